"""d2g_transition_counts and d2g_build_trans_mat (the device alignment of
each raw with its center, counted into a 16 x ncol matrix) against the
oracle's restatements (oracle/learn.py), bit-exact."""
from pathlib import Path

import numpy as np
import pytest

from dada2_rs_amd import formats, synth

pytestmark = pytest.mark.gpu
GOLDEN = Path(__file__).resolve().parent / "golden"


def _with_indels(seqs, quals, rng, frac=0.3):
    """A deletion or an insertion of one base in a fraction of the uniques
    (their quality string edited the same way), so the final alignments have
    gaps and unequal lengths."""
    seqs = [s.copy() for s in seqs]
    quals = [q.copy() for q in quals]
    for i in range(len(seqs)):
        if rng.random() >= frac:
            continue
        p = int(rng.integers(5, len(seqs[i]) - 5))
        if rng.random() < 0.5:
            seqs[i] = np.delete(seqs[i], p)
            quals[i] = np.delete(quals[i], p)
        else:
            seqs[i] = np.insert(seqs[i], p, np.uint8(rng.integers(1, 5)))
            quals[i] = np.insert(quals[i], p, quals[i][p])
    return seqs, quals


def _partition(oracle, seqs, quals, reads, err, params):
    from oracle import learn
    from oracle import partition as P
    rs = oracle.RawSet(seqs, quals, reads, k=5)
    out = P.run_dada(rs, err, params)
    center_of, weight = learn.center_weight(out, reads)
    return rs, center_of, weight


CASES = [
    # (name, workload args, align args, ncol, big weights)
    ("vectorized", dict(n_true=8, n_near=2, L=150, n_uniques=400, seed_off=50), dict(band=16), 41,
     False),
    ("gapless_band0", dict(n_true=6, n_near=2, L=140, n_uniques=300, seed_off=51), dict(band=0), 41,
     False),
    ("scalar_endsfree", dict(n_true=6, n_near=2, L=150, n_uniques=300, seed_off=52),
     dict(band=16, vectorized=False), 41, False),
    ("narrow_ncol", dict(n_true=6, n_near=2, L=150, n_uniques=300, seed_off=53), dict(band=16), 20,
     False),
    ("big_weights", dict(n_true=5, n_near=1, L=150, n_uniques=250, seed_off=54), dict(band=16), 41,
     True),
    ("long_homopolymer", dict(n_true=3, n_near=1, L=1450, n_uniques=40, seed_off=55),
     dict(band=32, homo_gap_p=-1), 41, False),
]


@pytest.mark.parametrize("name,wargs,akw,ncol,big", CASES, ids=[c[0] for c in CASES])
def test_transitions_match_oracle(pkg, oracle, name, wargs, akw, ncol, big):
    from oracle import learn
    wargs = dict(wargs)
    seed = synth.SEED_BASE + wargs.pop("seed_off")
    w = synth.workload_partition(seed=seed, **wargs)
    rng = np.random.default_rng(seed)
    seqs, quals = _with_indels(w.seqs, w.quals, rng)
    params = oracle.make_params(**akw)
    rs, center_of, weight = _partition(oracle, seqs, quals, w.reads, w.err, params)
    if big:  # u32 sums that wrap (transition_counts) or saturate (build_trans_mat)
        weight = np.where(weight > 0, rng.integers(1 << 28, 1 << 32, size=len(weight)), 0)
        rs = oracle.RawSet(seqs, quals, np.maximum(weight, 1), k=5)
    print(f"case {name}: {len(seqs)} uniques, {len(set(center_of.tolist()))} centers, "
          f"{int((weight > 0).sum())} counted", flush=True)
    want_tc = learn.transition_counts(rs, center_of, weight, params, True, ncol)
    want_bt = learn.build_trans_mat(seqs, quals, center_of, weight, params, ncol)
    ctx = pkg.Context(0)
    try:
        ctx.load_uniques(seqs, quals, w.reads, k=5)
        ap = pkg.align_params(**akw)
        got_tc = ctx.transition_counts(center_of, weight, ap, ncol=ncol)
        got_bt = ctx.build_trans_mat(center_of, weight, ap, nq=ncol)
    finally:
        ctx.close()
    assert np.array_equal(got_tc, want_tc)
    assert np.array_equal(got_bt, want_bt)
    assert want_tc.sum() > 0
    if big:
        assert (want_bt == 0xFFFFFFFF).any()


def test_transitions_fixture_reads(pkg, oracle):
    """The reference's fixture reads: dereplicated, partitioned, counted."""
    from oracle import learn
    seqs, quals, reads = formats.raws_from_derep(formats.derep_fastq(GOLDEN / "sam1F.fastq.gz"))
    err = synth.fixed_err(41)
    params = oracle.make_params()
    rs, center_of, weight = _partition(oracle, seqs, quals, reads, err, params)
    want_tc = learn.transition_counts(rs, center_of, weight, params, True, 41)
    want_bt = learn.build_trans_mat(seqs, quals, center_of, weight, params, 41)
    ctx = pkg.Context(0)
    try:
        ctx.load_uniques(seqs, quals, reads, k=5)
        got_tc = ctx.transition_counts(center_of, weight, ncol=41)
        got_bt = ctx.build_trans_mat(center_of, weight, nq=41)
    finally:
        ctx.close()
    print(f"fixture: {len(seqs)} uniques, {int(want_tc.sum())} counted bases", flush=True)
    assert np.array_equal(got_tc, want_tc)
    assert np.array_equal(got_bt, want_bt)


def test_transitions_edges_and_errors(pkg, oracle):
    from oracle import learn
    w = synth.workload_partition(4, 1, 0, 120, 60, seed=synth.SEED_BASE + 56)
    n = w.n
    params = oracle.make_params()
    ctx = pkg.Context(0)
    try:
        with pytest.raises(pkg.D2GError):  # nothing loaded
            ctx.n = n
            ctx.transition_counts(np.zeros(n, np.int64), np.ones(n))
        # no qualities: one column, and build_trans_mat counts nothing
        ctx.load_uniques(w.seqs, None, w.reads, k=5)
        center_of = np.zeros(n, np.int64)
        weight = np.asarray(w.reads, np.int64)
        rs0 = oracle.RawSet(w.seqs, None, w.reads, k=5)
        got = ctx.transition_counts(center_of, weight)
        assert got.shape == (16, 1)
        assert np.array_equal(got, learn.transition_counts(rs0, center_of, weight, params, False, 41))
        assert not ctx.build_trans_mat(center_of, weight).any()
        ctx.load_uniques(w.seqs, w.quals, w.reads, k=5)
        # nothing counted: every unique outside a cluster, or weight 0
        assert not ctx.transition_counts(np.full(n, -1), weight).any()
        assert not ctx.build_trans_mat(center_of, np.zeros(n)).any()
        # use_quals off: one column even with qualities loaded
        one = ctx.transition_counts(center_of, weight, use_quals=False)
        assert one.shape == (16, 1) and one.sum() == ctx.transition_counts(center_of, weight).sum()
        bad = center_of.copy()
        bad[3] = n
        with pytest.raises(pkg.D2GError):
            ctx.transition_counts(bad, weight)
        with pytest.raises(ValueError):
            ctx.transition_counts(center_of[:-1], weight)
        # the plain compare still works after a transition pass
        lam = ctx.compare_center(0, w.err).lam
        lam2, _, _, _, _ = oracle.compare_center(oracle.RawSet(w.seqs, w.quals, w.reads, k=5), 0,
                                                 w.err, 41, params)
        assert np.array_equal(lam, lam2)
    finally:
        ctx.close()


AUX_CASES = [
    # (name, workload args, align args, use quals)
    ("vectorized", dict(n_true=8, n_near=3, L=150, n_uniques=400, seed_off=60), dict(band=16), True),
    ("gapless_band0", dict(n_true=6, n_near=2, L=140, n_uniques=300, seed_off=61), dict(band=0),
     True),
    ("no_quals", dict(n_true=6, n_near=2, L=150, n_uniques=300, seed_off=62), dict(band=16), False),
    ("long_homopolymer", dict(n_true=3, n_near=2, L=1450, n_uniques=40, seed_off=63),
     dict(band=32, homo_gap_p=-1), True),
]


def _aux_both(pkg, oracle, seqs, quals, reads, err, akw, use_quals):
    from oracle import learn
    from oracle import partition as P
    params = oracle.make_params(**akw)
    rs = oracle.RawSet(seqs, quals if use_quals else None, reads, k=5)
    out = P.run_dada(rs, err, params, use_quals=use_quals)
    centers = [c["center"] for c in out.clusters]
    bfrom = [c["birth_from"] for c in out.clusters]
    weight = np.where(out.map >= 0, np.asarray(reads, np.int64), 0)
    maxlen = max(len(s) for s in seqs)
    want = learn.dada_aux(rs, centers, out.cluster_of.astype(np.int64), weight, params, use_quals,
                          41, maxlen)
    want_b = learn.birth_subs(rs, centers, bfrom, params, use_quals)
    ctx = pkg.Context(0)
    try:
        ctx.load_uniques(seqs, quals if use_quals else None, reads, k=5)
        ap = pkg.align_params(**akw)
        res = ctx.run_dada(err, ap, pkg.dada_params(use_quals=use_quals))
        got = ctx.compute_aux(res, ap, use_quals=use_quals, ncol=41)
        got_b = ctx.birth_subs(res, ap, use_quals=use_quals)
    finally:
        ctx.close()
    return out, res, want, want_b, got, got_b


@pytest.mark.parametrize("name,wargs,akw,use_quals", AUX_CASES, ids=[c[0] for c in AUX_CASES])
def test_dada_aux_matches_oracle(pkg, oracle, name, wargs, akw, use_quals):
    wargs = dict(wargs)
    seed = synth.SEED_BASE + wargs.pop("seed_off")
    w = synth.workload_partition(seed=seed, **wargs)
    seqs, quals = _with_indels(w.seqs, w.quals, np.random.default_rng(seed))
    out, res, want, want_b, got, got_b = _aux_both(pkg, oracle, seqs, quals, w.reads, w.err, akw,
                                                   use_quals)
    print(f"case {name}: {len(seqs)} uniques, {len(out.clusters)} clusters, "
          f"{len(want_b[0])} birth subs", flush=True)
    assert [c["center"] for c in res.clusters] == [c["center"] for c in out.clusters]
    assert np.array_equal(got["transitions"], want["transitions"])
    assert np.array_equal(got["nsubs"], want["nsubs"])
    gq, wq = got["cluster_quality"], want["cluster_quality"]
    assert np.array_equal(np.isnan(gq), np.isnan(wq))
    assert np.array_equal(gq[~np.isnan(gq)], wq[~np.isnan(wq)])  # bit-exact
    recs, nsubs, qave = got_b
    assert recs == want_b[0]
    assert nsubs.tolist() == want_b[1]
    assert np.array_equal(np.isnan(qave), np.isnan(want_b[2]))
    assert np.array_equal(qave[~np.isnan(qave)], np.asarray(want_b[2])[~np.isnan(want_b[2])])
    # cluster_stats from the same pieces
    for k, st in enumerate(got["cluster_stats"]):
        assert st["pval"] == out.clusters[k]["posthoc_pval"]
        m = (out.map == k)
        assert st["nunq"] == int(m.sum())
        assert st["n0"] == int(np.asarray(w.reads)[m & (want["nsubs"] == 0)].sum())
    assert len(out.clusters) > 1


def test_dada_aux_fixture_reads(pkg, oracle):
    seqs, quals, reads = formats.raws_from_derep(formats.derep_fastq(GOLDEN / "sam1F.fastq.gz"))
    out, res, want, want_b, got, got_b = _aux_both(pkg, oracle, seqs, quals, reads,
                                                   synth.fixed_err(41), {}, True)
    print(f"fixture: {len(out.clusters)} clusters, {len(want_b[0])} birth subs", flush=True)
    assert np.array_equal(got["transitions"], want["transitions"])
    assert np.array_equal(got["nsubs"], want["nsubs"])
    gq, wq = got["cluster_quality"], want["cluster_quality"]
    assert np.array_equal(np.isnan(gq), np.isnan(wq))
    assert np.array_equal(gq[~np.isnan(gq)], wq[~np.isnan(wq)])
    assert got_b[0] == want_b[0] and got_b[1].tolist() == want_b[1]


def test_birth_subs_capacity(pkg):
    """More birth records than the capacity: D2G_ERR_ARG, the count set and
    only `capacity` records written; the full call then returns them all."""
    import ctypes as C
    w = synth.workload_partition(n_true=5, n_near=2, L=150, n_uniques=200,
                                 seed=synth.SEED_BASE + 64)
    ctx = pkg.Context(0)
    try:
        ctx.load_uniques(w.seqs, w.quals, w.reads, k=5)
        res = ctx.run_dada(w.err)
        recs, _, _ = ctx.birth_subs(res)
        assert len(recs) > 2
        ncl = len(res.clusters)
        centers = np.ascontiguousarray([c["center"] for c in res.clusters], np.uint32)
        bfrom = np.ascontiguousarray([c["birth_from"] for c in res.clusters], np.uint32)
        ns, qa = np.zeros(ncl, np.uint32), np.zeros(ncl, np.float64)
        cl, pos = np.zeros(2, np.uint32), np.zeros(2, np.uint32)
        nt0, nt1, q = np.zeros(2, np.uint8), np.zeros(2, np.uint8), np.zeros(2, np.uint8)
        nrec = C.c_uint32()
        u32p, u8p = C.POINTER(C.c_uint32), C.POINTER(C.c_uint8)
        f64p = C.POINTER(C.c_double)
        rc = pkg._lib().d2g_birth_subs(
            ctx.h, ncl, centers.ctypes.data_as(u32p), bfrom.ctypes.data_as(u32p),
            C.byref(pkg.align_params()), 1, 2, ns.ctypes.data_as(u32p), qa.ctypes.data_as(f64p),
            cl.ctypes.data_as(u32p), pos.ctypes.data_as(u32p), nt0.ctypes.data_as(u8p),
            nt1.ctypes.data_as(u8p), q.ctypes.data_as(u8p), C.byref(nrec))
        assert rc == -1 and nrec.value == len(recs)
        assert [(int(cl[i]), int(pos[i])) for i in range(2)] == [r[:2] for r in recs[:2]]
    finally:
        ctx.close()


def test_dada_aux_edges(pkg, oracle):
    """cluster_quality truncated at a maxlen below the centers' length (as
    the oracle's restatement, error.rs:268); build_trans_mat with the k-mer
    screen on and a negative cutoff counts nothing (nwalign.rs:1296 with
    kdist 0.0); argument errors."""
    from oracle import learn
    from oracle import partition as P
    w = synth.workload_partition(n_true=4, n_near=2, L=150, n_uniques=200,
                                 seed=synth.SEED_BASE + 65)
    params = oracle.make_params()
    rs = oracle.RawSet(w.seqs, w.quals, w.reads, k=5)
    out = P.run_dada(rs, w.err, params)
    centers = [c["center"] for c in out.clusters]
    weight = np.where(out.map >= 0, np.asarray(w.reads, np.int64), 0)
    want = learn.dada_aux(rs, centers, out.cluster_of.astype(np.int64), weight, params, True, 41,
                          100)
    ctx = pkg.Context(0)
    try:
        ctx.load_uniques(w.seqs, w.quals, w.reads, k=5)
        res = ctx.run_dada(w.err)
        got = ctx.dada_aux(res, maxlen=100)
        gq, wq = got["cluster_quality"], want["cluster_quality"]
        assert gq.shape == (len(centers), 100)
        assert np.array_equal(np.isnan(gq), np.isnan(wq))
        assert np.array_equal(gq[~np.isnan(gq)], wq[~np.isnan(wq)])
        assert np.array_equal(got["transitions"], want["transitions"])
        center_of, wt = learn.center_weight(out, w.reads)
        neg = pkg.align_params(kdist_cutoff=-0.5)
        assert not ctx.build_trans_mat(center_of, wt, neg).any()
        assert np.array_equal(learn.build_trans_mat(w.seqs, w.quals, center_of, wt,
                                                    oracle.make_params(kdist_cutoff=-0.5), 41),
                              np.zeros((16, 41), np.uint32))
        # a cluster index past the clusters: D2G_ERR_ARG
        bad = res.cluster_of.copy()
        bad[0] = len(res.clusters)
        import copy
        r2 = copy.copy(res)
        r2.cluster_of = bad
        with pytest.raises(pkg.D2GError):
            ctx.dada_aux(r2)
    finally:
        ctx.close()
