"""GPU parity of the scalar ends-free aligners against the oracle, byte for
byte: align_endsfree_with_buf (nwalign.rs:336-417) and
align_endsfree_homo_with_buf (nwalign.rs:435-528), and the compare routed
through them (homopolymer gap penalty, raw_align_dp nwalign.rs:1400-1402).

Cases: the reference's rolling-buffer fuzz grid shapes (nwalign.rs:2873-2952)
within the kernel's envelope, the reference's homopolymer pair
(nwalign.rs:3132-3173), unbanded short pairs, and long reads at the
long-read workload's shape (1,420-1,480 nt, band 32, homo_gap -1).
"""
import numpy as np
import pytest

from test_gpu_align import fuzz_grid

pytestmark = pytest.mark.gpu


def _check_pairs(ctx, oracle, pairs, band, homo_gap=None):
    s1s = [a for a, _ in pairs]
    s2s = [b for _, b in pairs]
    got = ctx.align_endsfree(s1s, s2s, band=band, homo_gap_p=homo_gap)
    params = oracle.make_params(band=band, homo_gap_p=homo_gap)
    for (a, b), g in zip(pairs, got):
        if homo_gap is None:
            want = oracle.align_endsfree(a, b, band=band)
        else:
            want = oracle.align_endsfree_homo(a, b, params)
        assert g == want, (band, homo_gap, len(a), len(b))
    return len(pairs)


def test_band_fuzz_grid(ctx, oracle, pkg):
    groups = {}
    for s1, s2, band, eg in fuzz_grid():
        if eg != 0 or not pkg.band_supported(len(s1), len(s2), band, True):
            continue
        groups.setdefault(band, []).append((s1, s2))
    n = 0
    for band, pairs in groups.items():
        n += _check_pairs(ctx, oracle, pairs, band)
        n += _check_pairs(ctx, oracle, pairs, band, homo_gap=-1)
    assert n > 1500, n


def test_band_homopolymer_pair(ctx, oracle):
    enc = oracle.encode
    pairs = [(enc("AGAAAAGGGGTTTAAAAAATTTTTCCCC"), enc("AAAAAGGGGTTTAAAAAATTTTTCCCC")),
             (enc("ACGTACGTACGT"), enc("ACGTTCGTACGT")), (enc("ACGTACGTACGT"), enc("ACGACGTACGT")),
             (enc("ACGTACGTACGT"), enc("ACGTACGTACGTAC")), (enc("GCGG"), enc("CG"))]
    for band in (16, 8, 32, -1):
        _check_pairs(ctx, oracle, pairs, band)
        _check_pairs(ctx, oracle, pairs, band, homo_gap=-1)
        _check_pairs(ctx, oracle, pairs, band, homo_gap=-3)


def test_band_unbanded_and_ascii(ctx, oracle):
    rng = np.random.default_rng(5)
    pairs = []
    for _ in range(60):
        la, lb = int(rng.integers(1, 50)), int(rng.integers(1, 50))
        pairs.append((rng.integers(1, 5, la).astype(np.uint8).tobytes(),
                      rng.integers(1, 5, lb).astype(np.uint8).tobytes()))
    _check_pairs(ctx, oracle, pairs, -1)
    _check_pairs(ctx, oracle, pairs, -1, homo_gap=-2)
    ascii_pairs = [(b"ACGTTTTACGGGA", b"ACGTTTACGGGGA"), (b"TTTTTTTT", b"TTTT"),
                   (b"GATTACA" * 5, b"GATTTACA" * 5)]
    _check_pairs(ctx, oracle, ascii_pairs, 16, homo_gap=-1)


def test_band_long_reads(ctx, oracle, pkg):
    """Pairs at the long-read shape: 1,420-1,480 nt, band 32, homo_gap -1."""
    from dada2_rs_amd import synth
    w = synth.workload_longread(n_variants=3, n_uniques=24, seed=synth.SEED_BASE + 40)
    pairs = [(w.seqs[int(c)].tobytes(), w.seqs[i].tobytes())
             for c in w.centers for i in range(0, w.n, 3)]
    pairs = [p for p in pairs if pkg.band_supported(len(p[0]), len(p[1]), 32, True)]
    assert len(pairs) >= 16
    _check_pairs(ctx, oracle, pairs, 32, homo_gap=-1)
    _check_pairs(ctx, oracle, pairs[:8], 32)


def test_compare_homopolymer_long_reads(ctx, oracle, pkg):
    """The compare with homo_gap_p = -1 routes every DP pair to the
    homopolymer aligner (nwalign.rs:1400-1402): lambda bits, hamming and
    skipped flags against the oracle's compare."""
    from dada2_rs_amd import synth
    w = synth.workload_longread(n_variants=4, n_uniques=160, seed=synth.SEED_BASE + 41)
    rs = oracle.RawSet(w.seqs, w.quals, w.reads, k=w.kmer_size)
    ctx.load_uniques(w.seqs, w.quals, w.reads, k=w.kmer_size)
    params = pkg.align_params(band=32, homo_gap_p=-1)
    oparams = oracle.make_params(band=32, homo_gap_p=-1)
    got = ctx.compare_batch(w.centers, w.err, params)
    assert got.stats["aligned_dp"] > 0
    for r, c in enumerate(int(x) for x in w.centers):
        lam, ham, skp = oracle.compare_center(rs, c, w.err, w.err.shape[1], oparams, nthreads=8)[:3]
        assert np.array_equal(got.lam[r].view(np.uint64), lam.view(np.uint64)), c
        assert np.array_equal(got.hamming[r], ham), c
        assert np.array_equal(got.skipped[r], skp), c


def test_band_wide_geometries(ctx, oracle, pkg):
    """Bands wider than one wavefront of lanes: 2 and 4 cells per lane."""
    rng = np.random.default_rng(17)
    cases = []
    for la, lb, band in [(200, 300, 32), (300, 200, 32), (240, 330, 16), (150, 150, -1),
                         (180, 120, -1), (1450, 1380, 32), (1380, 1460, 32)]:
        for _ in range(3):
            a = rng.integers(1, 5, la).astype(np.uint8)
            b = np.concatenate([a[: min(la, lb)], rng.integers(1, 5, max(0, lb - la))])[:lb]
            b = b.astype(np.uint8).copy()
            hit = rng.random(lb) < 0.05
            b[hit] = rng.integers(1, 5, int(hit.sum()))
            cases.append((a.tobytes(), b.tobytes(), band))
    for band in sorted({c[2] for c in cases}):
        pairs = [(a, b) for a, b, bd in cases if bd == band]
        for a, b in pairs:
            assert pkg.band_supported(len(a), len(b), band, True)
        _check_pairs(ctx, oracle, pairs, band)
        _check_pairs(ctx, oracle, pairs, band, homo_gap=-1)
    assert not pkg.band_supported(600, 600, -1, True)  # 1203 diagonals: beyond 4 x 128


@pytest.mark.parametrize("tile", [1, 2, 4])
def test_compare_long_reads_every_dp_tile(pkg, oracle, tile):
    """The row-band kernel at 1, 2 and 4 cells per lane (the DP tiles the
    autotuner chooses among): the compare's results do not change."""
    from dada2_rs_amd import synth
    w = synth.workload_longread(n_variants=4, n_uniques=120, seed=synth.SEED_BASE + 42)
    rs = oracle.RawSet(w.seqs, w.quals, w.reads, k=w.kmer_size)
    oparams = oracle.make_params(band=32, homo_gap_p=-1)
    c = pkg.Context(0)
    try:
        c.load_uniques(w.seqs, w.quals, w.reads, k=w.kmer_size)
        c.set_dp_tile(tile)
        got = c.compare_batch(w.centers, w.err, pkg.align_params(band=32, homo_gap_p=-1))
    finally:
        c.close()
    assert got.stats["aligned_dp"] > 0
    for r, cc in enumerate(int(x) for x in w.centers):
        lam, ham, skp = oracle.compare_center(rs, cc, w.err, w.err.shape[1], oparams, nthreads=8)[:3]
        assert np.array_equal(got.lam[r].view(np.uint64), lam.view(np.uint64)), cc
        assert np.array_equal(got.hamming[r], ham), cc


@pytest.mark.parametrize("sub", [0.02, 0.03])
def test_compare_long_reads_dense_substitutions(ctx, oracle, pkg, sub):
    """Long-read pairs 2-6% apart (35 to about 105 substitutions): more
    substitution records than the pair has lanes (34 + |len1 - len2| / 2 at
    band 32) and more than one 64-record register of the one-pair
    traceback.  Lambda bits, hamming and skipped flags against the oracle's
    compare, then the alignment strings of the same pairs."""
    from dada2_rs_amd import synth
    w = synth.workload_longread(n_variants=2, n_uniques=30, sub=sub, seed=synth.SEED_BASE + 43)
    rs = oracle.RawSet(w.seqs, w.quals, w.reads, k=w.kmer_size)
    ctx.load_uniques(w.seqs, w.quals, w.reads, k=w.kmer_size)
    params = pkg.align_params(band=32, homo_gap_p=-1)
    oparams = oracle.make_params(band=32, homo_gap_p=-1)
    got = ctx.compare_batch(w.centers, w.err, params)
    assert got.stats["aligned_dp"] > 0
    hams = []
    for r, c in enumerate(int(x) for x in w.centers):
        lam, ham, skp = oracle.compare_center(rs, c, w.err, w.err.shape[1], oparams, nthreads=8)[:3]
        assert np.array_equal(got.lam[r].view(np.uint64), lam.view(np.uint64)), c
        assert np.array_equal(got.hamming[r], ham), c
        assert np.array_equal(got.skipped[r], skp), c
        hams += [int(h) for h, s in zip(ham, skp) if not s and h != 0xFFFFFFFF]
    assert max(hams) > 64 and sum(h > 34 for h in hams) > 10, hams
    pairs = [(w.seqs[int(c)].tobytes(), w.seqs[i].tobytes())
             for c in w.centers for i in range(0, w.n, 2)]
    pairs = [p for p in pairs if pkg.band_supported(len(p[0]), len(p[1]), 32, True)]
    _check_pairs(ctx, oracle, pairs, 32, homo_gap=-1)
    _check_pairs(ctx, oracle, pairs, 32)
