"""raw_align_dp's routing per pair on the device (nwalign.rs:1315-1414):
a unique set on both sides of the vectorized aligner's 3,500-nt limit
(pairs with a longer side above it take the scalar ends-free aligner,
nwalign.rs:1379-1385), and pairs whose compressed band is wider than one
wavefront (length differences above ~90 at band 16).  Compare results,
the partition loop and the final-alignment counts equal the oracle's."""
import numpy as np
import pytest

from dada2_rs_amd import synth

pytestmark = pytest.mark.gpu


def _rows(oracle, seqs, quals, reads, centers, err, params, greedy=False):
    rs = oracle.RawSet(seqs, quals, reads, k=5)
    rows = [oracle.compare_center(rs, int(c), err, err.shape[1], params, greedy=greedy,
                                  nthreads=8)
            for c in centers]
    return (np.stack([r[0] for r in rows]), np.stack([r[1] for r in rows]),
            np.stack([r[2] for r in rows]))


def _same(got, lam, ham, skp):
    assert np.array_equal(got.lam.view(np.uint64), lam.view(np.uint64))
    assert np.array_equal(got.hamming, ham)
    assert np.array_equal(got.skipped, skp)


def _straddling(seed=23, n=36):
    """Uniques of 3,488-3,512 nt from two 3,512-nt variants (prefixes, 0.3%
    substitutions), so the set straddles 3,500 nt."""
    rng = np.random.default_rng(seed)
    variants = [rng.integers(1, 5, size=3512).astype(np.uint8) for _ in range(2)]
    seqs, seen = [], set()
    while len(seqs) < n:
        v = variants[len(seqs) % 2]
        L = int(rng.integers(3488, 3513))
        s = v[:L].copy()
        flip = rng.random(L) < 0.003
        s[flip] = rng.integers(1, 5, size=int(flip.sum()))
        if s.tobytes() not in seen:
            seen.add(s.tobytes())
            seqs.append(s)
    quals = [np.clip(rng.normal(36, 3, size=len(s)).round(), 0, 40).astype(np.uint8) for s in seqs]
    reads = np.maximum(1, rng.lognormal(2, 1.5, size=n).round()).astype(np.uint32)
    return seqs, quals, reads


def test_compare_straddling_the_vectorized_limit(ctx, oracle, pkg):
    seqs, quals, reads = _straddling()
    lens = np.array([len(s) for s in seqs])
    assert (lens <= 3500).any() and (lens > 3500).any()
    ctx.load_uniques(seqs, quals, reads, k=5)
    short = [i for i in range(len(seqs)) if lens[i] <= 3500][:3]
    long_ = [i for i in range(len(seqs)) if lens[i] > 3500][:3]
    centers = np.array(short + long_, np.uint32)
    err = synth.fixed_err(41)
    got = ctx.compare_batch(centers, err, pkg.align_params(band=16), greedy=False)
    _same(got, *_rows(oracle, seqs, quals, reads, centers, err,
                      oracle.make_params(band=16)))
    assert got.stats["aligned_dp"] + got.stats["gapless"] > len(centers) * len(seqs) // 2
    # the same through the single-center entry and the device store
    one = ctx.compare_center(int(short[0]), err, pkg.align_params(band=16), greedy=False)
    assert np.array_equal(one.lam.view(np.uint64), got.lam[0].view(np.uint64))


def test_run_dada_straddling_the_vectorized_limit(pkg, oracle):
    from oracle import partition as P
    seqs, quals, reads = _straddling(seed=24, n=30)
    err = synth.fixed_err(41)
    rs = oracle.RawSet(seqs, quals, reads, k=5)
    want = P.run_dada(rs, err, oracle.make_params(band=16))
    ctx = pkg.Context(0)
    try:
        ctx.load_uniques(seqs, quals, reads, k=5)
        got = ctx.run_dada(err, pkg.align_params(band=16))
        center_of = np.array([got.clusters[k]["center"] for k in got.cluster_of], np.int64)
        weight = np.where(got.map >= 0, reads.astype(np.int64), 0)
        tc = ctx.transition_counts(center_of, weight, pkg.align_params(band=16), ncol=41)
    finally:
        ctx.close()
    assert [c["center"] for c in got.clusters] == [c["center"] for c in want.clusters]
    assert np.array_equal(got.p, want.p) and np.array_equal(got.map, want.map)
    assert np.array_equal(got.comp_lambda, want.comp_lambda)
    from oracle import learn
    wtc = learn.transition_counts(rs, center_of, weight, oracle.make_params(band=16), True, 41)
    assert np.array_equal(tc, wtc)


def test_compare_with_wide_length_differences(ctx, oracle, pkg):
    """Uniques that are prefixes of shared variants with lengths 130-300:
    prefixes pass the k-mer screen (their k-mers are the longer one's), and
    pairs more than ~90 nt apart exceed nw_vec's 64 compressed columns at
    band 16, so they run in the one-pair kernel."""
    rng = np.random.default_rng(31)
    variants = [rng.integers(1, 5, size=300).astype(np.uint8) for _ in range(3)]
    seqs, seen = [], set()
    while len(seqs) < 240:
        v = variants[len(seqs) % 3]
        L = int(rng.integers(130, 301))
        s = v[:L].copy()
        flip = rng.random(L) < 0.005
        s[flip] = rng.integers(1, 5, size=int(flip.sum()))
        if s.tobytes() not in seen:
            seen.add(s.tobytes())
            seqs.append(s)
    quals = [np.clip(rng.normal(35, 3, size=len(s)).round(), 0, 40).astype(np.uint8) for s in seqs]
    reads = np.maximum(1, rng.lognormal(2, 1.5, size=len(seqs)).round()).astype(np.uint32)
    lens = np.array([len(s) for s in seqs])
    centers = np.array([int(np.argmax(lens)), int(np.argmin(lens)), 5, 17], np.uint32)
    ctx.load_uniques(seqs, quals, reads, k=5)
    err = synth.fixed_err(41)
    got = ctx.compare_batch(centers, err, pkg.align_params(band=16), greedy=False)
    _same(got, *_rows(oracle, seqs, quals, reads, centers, err, oracle.make_params(band=16)))
    wide = sum(1 for c in centers for u in range(len(seqs))
               if abs(int(lens[c]) - int(lens[u])) > 100 and got.hamming[list(centers).index(c), u]
               != 0xFFFFFFFF)
    print(f"{wide} aligned pairs more than 100 nt apart", flush=True)
    assert wide > 50
