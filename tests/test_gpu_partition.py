"""d2g_run_dada (the partition loop over the device compare) against the
oracle's restatement of run_dada (oracle/partition.py): identical clusters,
births, membership, p-values, read map and counters."""
import numpy as np
import pytest

from dada2_rs_amd import synth

pytestmark = pytest.mark.gpu

CASES = [
    # (name, workload args, run options)
    ("default", dict(n_true=8, n_near=2, n_prior=1, L=150, n_uniques=300, seed_off=5), {}),
    ("larger", dict(n_true=20, n_near=6, n_prior=2, L=250, n_uniques=1500, seed_off=11), {}),
    ("not_greedy", dict(n_true=10, n_near=3, n_prior=0, L=200, n_uniques=600, seed_off=12),
     dict(greedy=False)),
    ("singletons", dict(n_true=6, n_near=2, n_prior=1, L=150, n_uniques=400, seed_off=13),
     dict(detect_singletons=True, omega_c=1e-20)),
    ("max_clust", dict(n_true=12, n_near=4, n_prior=0, L=180, n_uniques=500, seed_off=14),
     dict(max_clust=5, max_shuffle=2)),
    ("no_quals_fold", dict(n_true=9, n_near=3, n_prior=1, L=160, n_uniques=450, seed_off=15),
     dict(use_quals=False, min_fold=2.0, min_hamming=2, min_abund=2)),
]


def run_both(pkg, oracle, w, opts):
    from oracle import partition as P
    kw = dict(omega_a=1e-40, omega_p=1e-4, omega_c=1e-40, min_fold=1.0, min_hamming=1,
              min_abund=1, max_clust=0, max_shuffle=10, greedy=True, detect_singletons=False,
              use_quals=True)
    opts = dict(opts)
    dev = opts.pop("device_shuffle", False)
    kw.update(opts)
    prior = w.extra["prior"]
    rs = oracle.RawSet(w.seqs, w.quals, w.reads, k=5)
    want = P.run_dada(rs, w.err, oracle.make_params(band=16), prior=prior, **kw)
    ctx = pkg.Context(0)
    try:
        ctx.load_uniques(w.seqs, w.quals, w.reads, k=5)
        got = ctx.run_dada(w.err, pkg.align_params(band=16),
                           pkg.dada_params(device_shuffle=dev, **kw), prior=prior)
    finally:
        ctx.close()
    return got, want


@pytest.mark.parametrize("device_shuffle", [False, True], ids=["host_shuffle", "device_shuffle"])
@pytest.mark.parametrize("name,wargs,opts", CASES, ids=[c[0] for c in CASES])
def test_run_dada_matches_oracle(pkg, oracle, name, wargs, opts, device_shuffle):
    wargs = dict(wargs)
    seed = synth.SEED_BASE + wargs.pop("seed_off")
    w = synth.workload_partition(seed=seed, **wargs)
    print(f"case {name}, device_shuffle={device_shuffle}: {w.n} uniques", flush=True)
    got, want = run_both(pkg, oracle, w, dict(opts, device_shuffle=device_shuffle))
    assert len(got.clusters) == len(want.clusters)
    for k, (g, o) in enumerate(zip(got.clusters, want.clusters)):
        for key in o:
            assert g[key] == o[key] or (g[key] != g[key] and o[key] != o[key]), (k, key)
    assert np.array_equal(got.cluster_of, want.cluster_of)
    assert np.array_equal(got.p, want.p)  # bit-exact
    assert np.array_equal(got.map, want.map)
    assert np.array_equal(got.comp_lambda, want.comp_lambda)
    assert np.array_equal(got.comp_hamming, want.comp_hamming)
    assert got.stats["nalign"] == want.nalign
    assert got.stats["nshroud"] == want.nshroud
    assert got.stats["shuffle_passes"] == want.shuffle_passes
    assert got.stats["compares"] == len(want.clusters)


def _run_pair(pkg, oracle, seqs, quals, reads, align_kw, prior=None, **kw):
    from oracle import partition as P
    reads = np.asarray(reads, np.uint32)
    rs = oracle.RawSet(seqs, quals, reads, k=5)
    want = P.run_dada(rs, synth.fixed_err(41), oracle.make_params(**align_kw), prior=prior, **kw)
    ctx = pkg.Context(0)
    try:
        ctx.load_uniques(seqs, quals, reads, k=5)
        got = ctx.run_dada(synth.fixed_err(41), pkg.align_params(**align_kw),
                           pkg.dada_params(**kw), prior=prior)
    finally:
        ctx.close()
    return got, want


def _same(got, want):
    assert [c["center"] for c in got.clusters] == [c["center"] for c in want.clusters]
    for g, o in zip(got.clusters, want.clusters):
        assert g == o or all(g[k] == o[k] or (g[k] != g[k] and o[k] != o[k]) for k in o)
    assert np.array_equal(got.cluster_of, want.cluster_of)
    assert np.array_equal(got.p, want.p)
    assert np.array_equal(got.map, want.map)
    assert got.stats["nalign"] == want.nalign and got.stats["nshroud"] == want.nshroud


def test_run_dada_edge_cases(pkg, oracle):
    rng = np.random.default_rng(synth.SEED_BASE + 30)
    base = rng.integers(1, 5, size=120).astype(np.uint8)
    q = np.full(120, 35, np.uint8)
    # one unique: a single cluster, its center corrected with p = 1
    got, want = _run_pair(pkg, oracle, [base], [q], [7], dict(band=16))
    _same(got, want)
    assert len(got.clusters) == 1 and got.p[0] == 1.0 and got.map[0] == 0
    # singletons only: no abundance p-value applies, no bud
    seqs = [base]
    for k in range(1, 40):
        s = base.copy()
        s[k] = s[k] % 4 + 1
        seqs.append(s)
    got, want = _run_pair(pkg, oracle, seqs, [q] * 40, [1] * 40, dict(band=16))
    _same(got, want)
    assert len(got.clusters) == 1
    # equal reads everywhere: the center is the last unique; with singleton
    # detection the prior/abundance tests still see every raw
    got, want = _run_pair(pkg, oracle, seqs, [q] * 40, [5] * 40, dict(band=16),
                          detect_singletons=True)
    _same(got, want)
    assert got.clusters[0]["center"] == 39


def test_run_dada_long_reads_homopolymer(pkg, oracle):
    w = synth.workload_partition(n_true=3, n_near=1, n_prior=0, L=1450, n_uniques=50,
                                 seed=synth.SEED_BASE + 31)
    print(f"long-read partition: {w.n} uniques of {len(w.seqs[0])} nt", flush=True)
    got, want = _run_pair(pkg, oracle, w.seqs, w.quals, w.reads, dict(band=32, homo_gap_p=-1))
    _same(got, want)
    # the same run with the shuffle's scan on the host's threads
    ctx = pkg.Context(0)
    try:
        ctx.load_uniques(w.seqs, w.quals, w.reads, k=5)
        dev = ctx.run_dada(synth.fixed_err(41), pkg.align_params(band=32, homo_gap_p=-1),
                           pkg.dada_params(device_shuffle=False))
    finally:
        ctx.close()
    _same(dev, want)
    assert sorted(c["center"] for c in got.clusters) == sorted(int(x) for x in w.extra["planted"])


def test_run_dada_errors(pkg):
    w = synth.workload_partition(4, 1, 0, 120, 60, seed=synth.SEED_BASE + 16)
    ctx = pkg.Context(0)
    try:
        with pytest.raises(pkg.D2GError):  # nothing loaded
            ctx.run_dada(w.err)
        ctx.load_uniques(w.seqs, w.quals, w.reads, k=5)
        with pytest.raises(pkg.D2GError):
            ctx.run_dada(w.err, dada=pkg.dada_params(max_shuffle=0))
        with pytest.raises(ValueError):
            ctx.run_dada(w.err, prior=np.zeros(3, np.uint8))
        ok = ctx.run_dada(w.err)
        assert len(ok.clusters) >= 1
    finally:
        ctx.close()


def test_quality_above_the_model_extends_it(pkg, oracle):
    """Qualities at or above the model's columns (Q41-45 under a Q0-40
    model): run_dada extends the matrix as dada_uniques_cached does
    (dada.rs:337-384) instead of refusing, and the partition and the aux
    transitions (now 46 columns) equal the oracle's on the extended matrix."""
    import warnings
    from oracle import learn
    from oracle import partition as P
    w = synth.workload_partition(n_true=8, n_near=2, n_prior=0, L=150, n_uniques=300,
                                 seed=synth.SEED_BASE + 41)
    quals = [np.asarray(q, np.uint8).copy() for q in w.quals]
    for i in range(0, len(quals), 3):
        quals[i][::7] = np.minimum(45, quals[i][::7].astype(np.int32) + 12).astype(np.uint8)
    qmax = max(int(q.max()) for q in quals)
    assert qmax >= 41
    ext = P.extend_err(w.err, qmax)
    rs = oracle.RawSet(w.seqs, quals, w.reads, k=5)
    want = P.run_dada(rs, ext, oracle.make_params(band=16))
    ctx = pkg.Context(0)
    try:
        ctx.load_uniques(w.seqs, quals, w.reads, k=5)
        with warnings.catch_warnings(record=True):
            warnings.simplefilter("always")
            got = ctx.run_dada(w.err, pkg.align_params(band=16))
        aux = ctx.compute_aux(got, pkg.align_params(band=16))
    finally:
        ctx.close()
    assert got.err_ncol == qmax + 1
    assert [c["center"] for c in got.clusters] == [c["center"] for c in want.clusters]
    assert np.array_equal(got.p, want.p) and np.array_equal(got.map, want.map)
    assert np.array_equal(got.comp_lambda, want.comp_lambda)
    centers = [c["center"] for c in want.clusters]
    weight = np.where(want.map >= 0, np.asarray(w.reads, np.int64), 0)
    want_aux = learn.dada_aux(rs, centers, want.cluster_of.astype(np.int64), weight,
                              oracle.make_params(band=16), True, qmax + 1, 150)
    assert aux["transitions"].shape == (16, qmax + 1)
    assert np.array_equal(aux["transitions"], want_aux["transitions"])
