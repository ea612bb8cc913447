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
    kw.update(opts)
    prior = w.extra["prior"]
    rs = oracle.RawSet(w.seqs, w.quals, w.reads, k=5)
    want = P.run_dada(rs, w.err, oracle.make_params(band=16), prior=prior, **kw)
    ctx = pkg.Context(0)
    try:
        ctx.load_uniques(w.seqs, w.quals, w.reads, k=5)
        got = ctx.run_dada(w.err, pkg.align_params(band=16), pkg.dada_params(**kw), prior=prior)
    finally:
        ctx.close()
    return got, want


@pytest.mark.parametrize("name,wargs,opts", CASES, ids=[c[0] for c in CASES])
def test_run_dada_matches_oracle(pkg, oracle, name, wargs, opts):
    wargs = dict(wargs)
    seed = synth.SEED_BASE + wargs.pop("seed_off")
    w = synth.workload_partition(seed=seed, **wargs)
    print(f"case {name}: {w.n} uniques", flush=True)
    got, want = run_both(pkg, oracle, w, opts)
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
