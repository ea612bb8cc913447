"""GPU parity of the compare map (b_compare_parallel, cluster.rs:279-330)
against the oracle: lambda bits, hamming and skipped flags must be identical
for every (center, unique) pair.
"""
import numpy as np
import pytest

from dada2_rs_amd import synth

pytestmark = pytest.mark.gpu


def _oracle_rows(oracle, w, centers, params=None, use_quals=True, greedy=True, lock=None):
    rs = oracle.RawSet(w.seqs, w.quals, w.reads, k=w.kmer_size, locks=lock)
    params = params or oracle.make_params(band=w.band)
    rows = [oracle.compare_center(rs, int(c), w.err, w.err.shape[1], params,
                                  use_quals=use_quals, greedy=greedy, nthreads=8)
            for c in centers]
    lam = np.stack([r[0] for r in rows])
    ham = np.stack([r[1] for r in rows])
    skp = np.stack([r[2] for r in rows])
    return lam, ham, skp


def _assert_same(got, lam, ham, skp):
    bad = np.nonzero(got.lam.view(np.uint64) != lam.view(np.uint64))
    assert bad[0].size == 0, f"lambda bits differ at {list(zip(*bad))[:5]}"
    assert np.array_equal(got.hamming, ham)
    assert np.array_equal(got.skipped, skp)


@pytest.fixture(scope="module")
def plain():
    return synth.workload_plain()


@pytest.fixture(scope="module")
def batched_small():
    return synth.subsample(synth.workload_batched(), 3000, n_centers=24)


def test_plain_center(ctx, oracle, pkg, plain):
    ctx.load_uniques(plain.seqs, plain.quals, plain.reads, k=5)
    c = int(plain.centers[0])
    got = ctx.compare_center(c, plain.err)
    _assert_same(got, *[x[0] for x in _oracle_rows(oracle, plain, [c])])
    assert got.stats["comparisons"] == plain.n


def test_plain_every_tenth_center_no_greedy(ctx, oracle, pkg, plain):
    ctx.load_uniques(plain.seqs, plain.quals, plain.reads, k=5)
    centers = np.arange(0, plain.n, 100)
    got = ctx.compare_batch(centers, plain.err, greedy=False)
    _assert_same(got, *_oracle_rows(oracle, plain, centers, greedy=False))


def test_batched_uneven(ctx, oracle, pkg, batched_small):
    w = batched_small
    ctx.load_uniques(w.seqs, w.quals, w.reads, k=5)
    got = ctx.compare_batch(w.centers, w.err)
    lam, ham, skp = _oracle_rows(oracle, w, w.centers)
    _assert_same(got, lam, ham, skp)
    assert got.stats["aligned_dp"] > 100


def test_batched_lock_and_no_quals(ctx, oracle, pkg, batched_small):
    w = batched_small
    rng = np.random.default_rng(3)
    lock = (rng.random(w.n) < 0.1).astype(np.uint8)
    ctx.load_uniques(w.seqs, w.quals, w.reads, k=5)
    cs = w.centers[:8]
    got = ctx.compare_batch(cs, w.err, use_quals=False, lock=lock)
    _assert_same(got, *_oracle_rows(oracle, w, cs, use_quals=False, lock=lock))


def test_cluster0_cutoff_one(ctx, oracle, pkg, plain):
    """The cluster-0 compare uses kdist_cutoff = 1.0 (dada.rs:835-838): nothing
    is shrouded, every pair aligns."""
    sub = synth.subsample(plain, 300)
    ctx.load_uniques(sub.seqs, sub.quals, sub.reads, k=5)
    p = pkg.align_params(kdist_cutoff=1.0)
    c = synth.plain_center(sub.reads)
    got = ctx.compare_center(c, sub.err, params=p)
    lam, ham, skp = _oracle_rows(oracle, sub, [c], params=oracle.make_params(kdist_cutoff=1.0))
    _assert_same(got, lam[0], ham[0], skp[0])


def test_results_are_deterministic(ctx, pkg, batched_small):
    w = batched_small
    ctx.load_uniques(w.seqs, w.quals, w.reads, k=5)
    a = ctx.compare_batch(w.centers[:6], w.err)
    b = ctx.compare_batch(w.centers[:6], w.err)
    assert np.array_equal(a.lam.view(np.uint64), b.lam.view(np.uint64))
    assert np.array_equal(a.hamming, b.hamming)


def test_graph_replays_match_eager(pkg, oracle, batched_small):
    """The captured launch sequence, replayed with new center lists of the
    same length and with eager calls of other shapes in between, gives the
    eager results every time."""
    w = batched_small
    g = pkg.Context(0)                 # graphs on (default), no timing
    e = pkg.Context(0, timing=True)    # timed calls always launch eagerly
    try:
        for c in (g, e):
            c.load_uniques(w.seqs, w.quals, w.reads, k=5)
        rng = np.random.default_rng(8)
        for rep in range(5):
            cs = np.sort(rng.choice(w.centers, size=8, replace=False))
            got = g.compare_batch(cs, w.err)
            want = e.compare_batch(cs, w.err)
            assert np.array_equal(got.lam.view(np.uint64), want.lam.view(np.uint64)), rep
            assert np.array_equal(got.hamming, want.hamming) and np.array_equal(got.skipped, want.skipped)
            # an eager call of another shape in between
            g.compare_center(int(cs[0]), w.err)
            e.compare_batch(cs[:3], w.err)
        lam, ham, skp = _oracle_rows(oracle, w, cs)
        _assert_same(got, lam, ham, skp)
    finally:
        g.close()
        e.close()


@pytest.mark.parametrize("qblock,split,per_lane", [(256, 1, 1), (256, 4, 1), (128, 2, 1), (64, 1, 1),
                                                   (64, 4, 1), (256, 2, 2), (128, 4, 2), (128, 1, 2)])
def test_screen_shapes_same_results(ctx, oracle, pkg, batched_small, qblock, split, per_lane):
    """Each launch shape of the screen (64, 128 or 256 uniques x 16 centers
    per workgroup, over 1, 2 or 4 wavefront groups, 1 or 2 uniques per lane)
    gives the oracle's results; 21 centers leave the last workgroup's groups
    partly or wholly without centers."""
    w = batched_small
    ctx.load_uniques(w.seqs, w.quals, w.reads, k=5)
    cs = w.centers[:21]
    ctx.set_screen_shape(qblock, split, per_lane)
    try:
        got = ctx.compare_batch(cs, w.err)
    finally:
        ctx.set_screen_shape()
    _assert_same(got, *_oracle_rows(oracle, w, cs))


def test_first_call_outgrows_substitution_pool(pkg, oracle):
    """A fresh context starts with a small substitution pool; a first
    compare with more DP pairs than it holds (cutoff 1.0: every pair aligns)
    is rerun after the pool grows.  Kernels of the overflowing pass must stay
    inside the pool, and the results must be the oracle's."""
    w = synth.subsample(synth.workload_batched(), 2000, n_centers=12)
    print("case: 2000 uniques x 12 centers, kdist_cutoff 1.0", flush=True)
    c = pkg.Context(0)
    try:
        c.load_uniques(w.seqs, w.quals, w.reads, k=5)
        p = pkg.align_params(kdist_cutoff=1.0)
        got = c.compare_batch(w.centers, w.err, params=p)
        assert got.stats["aligned_dp"] * 262 > (1 << 20)
        _assert_same(got, *_oracle_rows(oracle, w, w.centers, params=oracle.make_params(kdist_cutoff=1.0)))
    finally:
        c.close()


@pytest.mark.parametrize("cutoff", [0.05, 0.2, 0.6, 0.0])
def test_shroud_thresholds_at_other_cutoffs(ctx, oracle, pkg, batched_small, cutoff):
    """The screen decides kdist > cutoff from a tabulated dot threshold per
    overlap length; at other cutoffs (0.0: every pair with kdist > 0 is
    shrouded) the shroud and dispatch decisions stay the oracle's."""
    w = batched_small
    ctx.load_uniques(w.seqs, w.quals, w.reads, k=5)
    cs = w.centers[:8]
    got = ctx.compare_batch(cs, w.err, params=pkg.align_params(kdist_cutoff=cutoff))
    _assert_same(got, *_oracle_rows(oracle, w, cs, params=oracle.make_params(kdist_cutoff=cutoff)))


def test_autotuned_shape_same_results(pkg, oracle, batched_small):
    """With autotuning the first call times every screen shape on its own
    inputs; that call and the next (on the kept shape) give the oracle's
    results."""
    w = batched_small
    c = pkg.Context(0)
    try:
        c.set_autotune(True)
        c.load_uniques(w.seqs, w.quals, w.reads, k=5)
        cs = w.centers[:10]
        want = _oracle_rows(oracle, w, cs)
        _assert_same(c.compare_batch(cs, w.err), *want)
        _assert_same(c.compare_batch(cs, w.err), *want)
    finally:
        c.close()


@pytest.mark.parametrize("k", [3, 6, 7, 8])
def test_other_kmer_sizes(ctx, oracle, pkg, k):
    """k = 3, 6, 7 (the reference's PacBio setting) and 8 (where the
    reference switches to its sparse k-mer vectors, kmers.rs:29-45, whose
    distances equal the dense ones bit for bit): the device's dense
    screen against the oracle at each k."""
    w = synth.subsample(synth.workload_batched(), 600, n_centers=8)
    ctx.load_uniques(w.seqs, w.quals, w.reads, k=k)
    print(f"k={k}: {w.n} uniques x {len(w.centers)} centers", flush=True)
    for cutoff in (0.42, 0.1):
        params = pkg.align_params(band=w.band, kmer_size=k, kdist_cutoff=cutoff)
        got = ctx.compare_batch(w.centers, w.err, params)
        rs = oracle.RawSet(w.seqs, w.quals, w.reads, k=k)
        op = oracle.make_params(band=w.band, kmer_size=k, kdist_cutoff=cutoff)
        rows = [oracle.compare_center(rs, int(c), w.err, 41, op, nthreads=8) for c in w.centers]
        _assert_same(got, np.stack([r[0] for r in rows]), np.stack([r[1] for r in rows]),
                     np.stack([r[2] for r in rows]))
        assert got.stats["aligned_dp"] > 0
        if k >= 6:  # at k = 3 every profile is close: nothing is shrouded
            assert got.stats["shrouded"] > 0


@pytest.mark.parametrize("k", [3, 4, 5])
def test_one_center_screen_kmer_sizes(ctx, oracle, pkg, k):
    """One center per call takes the one-center screen (a wavefront per 16
    uniques, whole-vector row loads; one word per lane at k <= 4, four at
    k = 5): every center of a small batch alone, at two cutoffs, against the
    oracle."""
    w = synth.subsample(synth.workload_batched(), 600, n_centers=6)
    ctx.load_uniques(w.seqs, w.quals, w.reads, k=k)
    rs = oracle.RawSet(w.seqs, w.quals, w.reads, k=k)
    for cutoff in (0.42, 0.1):
        params = pkg.align_params(band=w.band, kmer_size=k, kdist_cutoff=cutoff)
        op = oracle.make_params(band=w.band, kmer_size=k, kdist_cutoff=cutoff)
        for c in w.centers:
            got = ctx.compare_center(int(c), w.err, params)
            lam, ham, skp, _, _ = oracle.compare_center(rs, int(c), w.err, 41, op, nthreads=8)
            assert np.array_equal(got.lam.view(np.uint64), lam.view(np.uint64)), (k, cutoff, c)
            assert np.array_equal(got.hamming, ham), (k, cutoff, c)
            assert np.array_equal(got.skipped, skp), (k, cutoff, c)


def _saturated_workload(seed=91):
    """Uniques of 300-330 nt led by a 280-nt A run, so the AAAAA bucket
    saturates kmer8 at 255 (kmers.rs:390-401: kmer_dist8 returns -1.0) and
    the compare takes the u16 kmer_dist fallback (nwalign.rs:1270-1278);
    mixed with ordinary random uniques of the same lengths."""
    rng = np.random.default_rng(seed)
    seqs = []
    base_tail = rng.integers(1, 5, size=40).astype(np.uint8)
    for i in range(60):
        run = 280 - int(rng.integers(0, 8))
        tail = base_tail[: 20 + int(rng.integers(0, 20))].copy()
        flip = rng.random(len(tail)) < 0.05
        tail[flip] = rng.integers(1, 5, size=int(flip.sum()))
        s = np.concatenate([np.ones(run, np.uint8), tail])
        if i % 3 == 0:  # a few substitutions inside the run keep its buckets >= 255
            pos = rng.integers(0, run, size=2)
            s[pos] = rng.integers(2, 5, size=2)
        seqs.append(s)
    for _ in range(140):
        seqs.append(rng.integers(1, 5, size=int(rng.integers(300, 331))).astype(np.uint8))
    uniq = {s.tobytes(): s for s in seqs}
    seqs = list(uniq.values())
    quals = [np.clip(rng.normal(34, 3, size=len(s)).round(), 0, 40).astype(np.uint8) for s in seqs]
    reads = np.maximum(1, rng.lognormal(2, 1.5, size=len(seqs)).round()).astype(np.uint32)
    return seqs, quals, reads


def test_u16_fallback_on_saturated_buckets(ctx, oracle, pkg):
    """Pairs where both k-mer vectors hold a 255 bucket: the screen's exact
    min-sum meets the saturation, and the u16 kmer_dist decides (the
    device's u16_min_dot branch); results equal the oracle's, and the
    oracle's own fallback is checked to have fired."""
    seqs, quals, reads = _saturated_workload()
    kv = [oracle.kmer8(s) for s in seqs[:60]]
    assert all(int(v.max()) == 255 for v in kv)
    assert oracle.kmer_dist8(kv[0], len(seqs[0]), kv[1], len(seqs[1])) == -1.0
    ctx.load_uniques(seqs, quals, reads, k=5)
    centers = np.arange(0, 60, 6)
    for cutoff in (0.42, 0.05):
        params = pkg.align_params(band=16, kdist_cutoff=cutoff)
        got = ctx.compare_batch(centers, synth.fixed_err(41), params)
        rs = oracle.RawSet(seqs, quals, reads, k=5)
        op = oracle.make_params(band=16, kdist_cutoff=cutoff)
        rows = [oracle.compare_center(rs, int(c), synth.fixed_err(41), 41, op, nthreads=8)
                for c in centers]
        _assert_same(got, np.stack([r[0] for r in rows]), np.stack([r[1] for r in rows]),
                     np.stack([r[2] for r in rows]))
    # the saturated pairs aligned (the fallback passed them), the random ones shrouded
    assert (got.hamming[:, :60] != 0xFFFFFFFF).sum() > 0
    # the same centers one at a time: the one-center screen's own fallback
    params = pkg.align_params(band=16, kdist_cutoff=0.42)
    op = oracle.make_params(band=16, kdist_cutoff=0.42)
    rs = oracle.RawSet(seqs, quals, reads, k=5)
    for c in centers[:4]:
        one = ctx.compare_center(int(c), synth.fixed_err(41), params)
        lam, ham, skp, _, _ = oracle.compare_center(rs, int(c), synth.fixed_err(41), 41, op,
                                                    nthreads=8)
        assert np.array_equal(one.lam.view(np.uint64), lam.view(np.uint64))
        assert np.array_equal(one.hamming, ham) and np.array_equal(one.skipped, skp)


@pytest.mark.parametrize("tile", [0, 2, 1])
def test_vectorized_dp_tiles_same_results(pkg, oracle, batched_small, tile):
    """nw_vec with every pair count per wavefront the autotuner tries."""
    w = batched_small
    c = pkg.Context(0)
    try:
        c.load_uniques(w.seqs, w.quals, w.reads, k=5)
        c.set_dp_tile(tile)
        cs = w.centers[:8]
        _assert_same(c.compare_batch(cs, w.err), *_oracle_rows(oracle, w, cs))
    finally:
        c.close()


def test_autotuned_long_read_dp_tile(pkg, oracle):
    """Autotuning on the long-read shape times the row-band tiles; the
    first and the next call match the oracle."""
    w = synth.workload_longread(n_variants=3, n_uniques=90, seed=synth.SEED_BASE + 43)
    params = pkg.align_params(band=32, homo_gap_p=-1)
    op = oracle.make_params(band=32, homo_gap_p=-1)
    want = _oracle_rows(oracle, w, w.centers, params=op)
    c = pkg.Context(0)
    try:
        c.set_autotune(True)
        c.load_uniques(w.seqs, w.quals, w.reads, k=w.kmer_size)
        _assert_same(c.compare_batch(w.centers, w.err, params), *want)
        _assert_same(c.compare_batch(w.centers, w.err, params), *want)
    finally:
        c.close()
