"""Pins of the oracle's transition counts (oracle/learn.py): a hand-worked
alignment counted by hand, the closed form of gapless pairs, the two
restatements agreeing where their rules coincide, and the wrap / saturate /
no-quality rules of error.rs:194-236 and learn_errors.rs:618-690."""
import numpy as np
import pytest

import oracle
from oracle import learn
from dada2_rs_amd import synth


def _hand_case():
    """Center of 40 bases; the query drops the first two (a leading overhang),
    substitutes center position 10, deletes position 20, inserts a base after
    position 30 and appends three bases (a trailing overhang).  The deleted
    and inserted bases differ from their neighbours, so each gap has one
    optimal place."""
    rng = np.random.default_rng(synth.SEED_BASE + 40)
    while True:
        c = rng.integers(1, 5, size=40).astype(np.uint8)
        if c[19] != c[20] and c[20] != c[21] and c[30] != c[31]:
            break
    ins = next(b for b in (1, 2, 3, 4) if b != c[30] and b != c[31])
    sub = c[10] % 4 + 1
    q = list(c[2:20]) + list(c[21:31]) + [ins] + list(c[31:40]) + [1, 2, 3]
    q[10 - 2] = sub
    q = np.asarray(q, np.uint8)
    # center position -> query position, by construction
    cmap = {}
    for p in range(40):
        if p < 2 or p == 20:
            continue
        cmap[p] = p - 2 if p < 20 else (p - 3 if p <= 30 else p - 2)
    return c, q, cmap


def test_hand_worked_alignment():
    c, q, cmap = _hand_case()
    params = oracle.make_params(band=16)
    rs = oracle.RawSet([c, q], None, [1, 1], k=5)
    ok, _, _, (al0, al1), _ = oracle.raw_align(rs, 0, 1, learn.final_align(params))
    assert ok
    # the alignment the construction fixes
    got = {}
    i0 = i1 = -1
    for a0, a1 in zip(al0, al1):
        i0 += a0 != ord("-")
        i1 += a1 != ord("-")
        if a0 != ord("-") and a1 != ord("-"):
            got[i0] = i1
    assert got == cmap
    rng = np.random.default_rng(synth.SEED_BASE + 41)
    qual = rng.integers(0, 41, size=len(q)).astype(np.uint8)
    w = 3
    want = np.zeros((16, 41), np.uint32)
    for p0, p1 in cmap.items():
        want[4 * (c[p0] - 1) + (q[p1] - 1), qual[p1]] += w
    # the substitution lands in its off-diagonal row
    assert want[4 * (c[10] - 1) + (q[8] - 1)].sum() == w and c[10] != q[8]
    rs = oracle.RawSet([c, q], [np.full(40, 30, np.uint8), qual], [5, w], k=5)
    tc = learn.transition_counts(rs, [0, 0], [0, w], params, True, 41)
    assert np.array_equal(tc, want)
    bt = learn.build_trans_mat([c, q], [np.full(40, 30, np.uint8), qual], [0, 0], [0, w], params, 41)
    assert np.array_equal(bt, want)


def test_gapless_closed_form():
    """Band 0 aligns every pair gaplessly: the counts are a histogram of
    (center base, query base, quality) over equal positions."""
    rng = np.random.default_rng(synth.SEED_BASE + 42)
    n, L, ncol = 30, 60, 25
    seqs = [rng.integers(1, 5, size=L).astype(np.uint8) for _ in range(n)]
    quals = [rng.integers(0, 40, size=L).astype(np.uint8) for _ in range(n)]
    center_of = rng.integers(0, 4, size=n)
    center_of[5] = -1
    weight = rng.integers(0, 6, size=n)
    params = oracle.make_params(band=0)
    skip = np.zeros((16, ncol), np.int64)
    clamp = np.zeros((16, ncol), np.int64)
    for i in range(n):
        if center_of[i] < 0 or weight[i] == 0:
            continue
        t = 4 * (seqs[center_of[i]].astype(np.int64) - 1) + (seqs[i].astype(np.int64) - 1)
        ql = quals[i].astype(np.int64)
        keep = ql < ncol
        np.add.at(skip, (t[keep], ql[keep]), weight[i])
        np.add.at(clamp, (t, np.minimum(ql, ncol - 1)), weight[i])
    rs = oracle.RawSet(seqs, quals, np.maximum(weight, 1), k=5)
    assert np.array_equal(learn.transition_counts(rs, center_of, weight, params, True, ncol), skip)
    assert np.array_equal(learn.build_trans_mat(seqs, quals, center_of, weight, params, ncol),
                          clamp)
    # without qualities: one column (transition_counts), nothing (build_trans_mat)
    rs0 = oracle.RawSet(seqs, None, np.maximum(weight, 1), k=5)
    one = learn.transition_counts(rs0, center_of, weight, params, False, ncol)
    assert one.shape == (16, 1) and np.array_equal(one[:, 0], clamp.sum(axis=1))
    assert not learn.build_trans_mat(seqs, None, center_of, weight, params, ncol).any()


def test_restatements_agree_on_a_partition():
    """transition_counts (Sub map) and build_trans_mat (alignment strings)
    count the same pairs the same way when every quality is below ncol and
    nothing overflows; the diagonal rows hold all matches."""
    from oracle import partition as P
    w = synth.workload_partition(n_true=5, n_near=2, n_prior=0, L=150, n_uniques=200,
                                 seed=synth.SEED_BASE + 43)
    params = oracle.make_params(band=16)
    rs = oracle.RawSet(w.seqs, w.quals, w.reads, k=5)
    out = P.run_dada(rs, w.err, params)
    center_of, weight = learn.center_weight(out, w.reads)
    assert (weight > 0).sum() > len(out.clusters)
    tc = learn.transition_counts(rs, center_of, weight, params, True, 41)
    bt = learn.build_trans_mat(w.seqs, w.quals, center_of, weight, params, 41)
    assert np.array_equal(tc, bt)
    diag = tc[[0, 5, 10, 15]].sum()
    assert diag > 0.9 * tc.sum()
    # every aligned center position counts once: at most weight x center length
    lens = np.asarray([len(w.seqs[c]) if c >= 0 else 0 for c in center_of])
    assert tc.sum() <= (weight * lens).sum()


@pytest.mark.parametrize("reads", [1 << 31, (1 << 32) - 1])
def test_wrap_and_saturate(reads):
    """u32 sums: transition_counts wraps (a release build's +=),
    build_trans_mat saturates (saturating_add)."""
    base = np.asarray([1, 2, 3, 4, 2, 1, 4, 3, 1, 3], np.uint8)
    seqs = [base, base.copy()]
    quals = [np.full(10, 7, np.uint8), np.full(10, 7, np.uint8)]
    params = oracle.make_params(band=0)
    rs = oracle.RawSet(seqs, quals, [reads, reads], k=5)
    tc = learn.transition_counts(rs, [0, 0], [reads, reads], params, True, 8)
    bt = learn.build_trans_mat(seqs, quals, [0, 0], [reads, reads], params, 8)
    for nt in (1, 2, 3, 4):
        cnt = int((base == nt).sum()) * 2  # two raws
        t = 5 * (nt - 1)
        assert tc[t, 7] == (cnt * reads) % (1 << 32)
        assert bt[t, 7] == min(cnt * reads, 0xFFFFFFFF)


def test_gapless_unequal_lengths():
    """Band 0 with unequal lengths: the shorter sequence is padded with gaps
    on the right (nwalign.rs:645-655), so positions below the shorter length
    count and the rest do not."""
    rng = np.random.default_rng(synth.SEED_BASE + 44)
    c = rng.integers(1, 5, size=50).astype(np.uint8)
    short = c[:43].copy()
    short[7] = short[7] % 4 + 1
    long = np.concatenate([c, rng.integers(1, 5, size=6).astype(np.uint8)])
    seqs = [c, short, long]
    quals = [rng.integers(0, 41, size=len(s)).astype(np.uint8) for s in seqs]
    params = oracle.make_params(band=0)
    rs = oracle.RawSet(seqs, quals, [9, 2, 3], k=5)
    tc = learn.transition_counts(rs, [0, 0, 0], [0, 2, 3], params, True, 41)
    want = np.zeros((16, 41), np.int64)
    for i, w in ((1, 2), (2, 3)):
        m = min(len(c), len(seqs[i]))
        t = 4 * (c[:m].astype(np.int64) - 1) + (seqs[i][:m].astype(np.int64) - 1)
        np.add.at(want, (t, quals[i][:m].astype(np.int64)), w)
    assert np.array_equal(tc, want)
    assert tc.sum() == 2 * 43 + 3 * 50


def test_dada_aux_hand_worked():
    """cluster_quality and Sub.nsubs by hand: a center (5 reads, quality 20)
    and a one-substitution member (2 reads, quality p at position p); the
    read-weighted mean at p is (20 * 5 + p * 2) / 7."""
    c = np.array([1, 2, 3, 4, 1, 2, 3, 4, 2, 2, 1, 3], np.uint8)
    q = c.copy()
    q[3] = 1
    quals = [np.full(12, 20, np.uint8), np.arange(12, dtype=np.uint8)]
    rs = oracle.RawSet([c, q], quals, [5, 2], k=5)
    out = learn.dada_aux(rs, [0], [0, 0], [5, 2], oracle.make_params(band=16), True, 41, 14)
    assert out["nsubs"].tolist() == [0, 1]
    want = np.full((1, 14), np.nan)
    want[0, :12] = (20 * 5 + np.arange(12) * 2) / 7
    assert np.allclose(out["cluster_quality"], want, equal_nan=True, rtol=0, atol=1e-12)
    # maxlen below the center length truncates the rows; a raw outside the
    # read map (weight 0) adds nothing
    out = learn.dada_aux(rs, [0], [0, 0], [5, 0], oracle.make_params(band=16), True, 41, 10)
    assert np.allclose(out["cluster_quality"], np.full((1, 10), 20.0))
    assert out["nsubs"].tolist() == [0, -1]
    # without qualities: all NaN
    rs0 = oracle.RawSet([c, q], None, [5, 2], k=5)
    out = learn.dada_aux(rs0, [0], [0, 0], [5, 2], oracle.make_params(band=16), False, 41, 12)
    assert np.isnan(out["cluster_quality"]).all() and out["transitions"].shape == (16, 1)


def test_birth_subs_hand_worked():
    """Cluster 1 = two substitutions of cluster 0's center (positions 10 and
    40), cluster 2 = cluster 0's center with position 30 deleted, born from
    cluster 1: its Sub against cluster 1's center has the same two
    substitutions, the second one's query quality shifted by the deletion."""
    rng = np.random.default_rng(synth.SEED_BASE + 45)
    while True:
        a = rng.integers(1, 5, size=60).astype(np.uint8)
        if a[29] != a[30] and a[30] != a[31]:
            break
    b = a.copy()
    b[10] = b[10] % 4 + 1
    b[40] = b[40] % 4 + 1
    c = np.delete(a, 30)
    quals = [np.full(60, 30, np.uint8), np.arange(60, dtype=np.uint8),
             np.arange(59, dtype=np.uint8)]
    rs = oracle.RawSet([a, b, c], quals, [100, 20, 10], k=5)
    recs, nsubs, qave = learn.birth_subs(rs, [0, 1, 2], [0, 0, 1], oracle.make_params(), True)
    dec = b"?ACGT"
    assert recs == [(1, 10, dec[a[10]], dec[b[10]], 10), (1, 40, dec[a[40]], dec[b[40]], 40),
                    (2, 10, dec[b[10]], dec[a[10]], 10), (2, 40, dec[b[40]], dec[a[40]], 39)]
    assert nsubs == [-1, 2, 2]
    assert np.isnan(qave[0]) and qave[1] == 25.0 and qave[2] == 24.5
    recs, _, qave = learn.birth_subs(rs, [0, 1, 2], [0, 0, 1], oracle.make_params(), False)
    assert all(r[4] is None for r in recs) and all(x != x for x in qave)


def test_post_hoc_pval_two_clusters():
    """post_hoc_pvals with two clusters: cluster 0's compare stores every
    raw, so cluster 1's expected reads are lambda(center 0 -> center 1) x
    cluster 0's reads, from the (pinned) compare_center and calc_pA."""
    from oracle import partition as P
    w = synth.workload_partition(n_true=1, n_near=1, n_prior=0, L=150, n_uniques=60,
                                 seed=synth.SEED_BASE + 46)
    params = oracle.make_params(band=16)
    rs = oracle.RawSet(w.seqs, w.quals, w.reads, k=5)
    out = P.run_dada(rs, w.err, params)
    assert len(out.clusters) == 2
    c0, c1 = out.clusters[0]["center"], out.clusters[1]["center"]
    lam0, _, _, _, _ = oracle.compare_center(rs, c0, w.err, 41, params)
    want = oracle.calc_pA(int(w.reads[c1]), float(lam0[c1]) * float(out.clusters[0]["reads"]),
                          True)
    assert out.clusters[1]["posthoc_pval"] == want
    assert 0.0 <= out.clusters[0]["posthoc_pval"] <= 1.0


@pytest.mark.parametrize("nthreads", [1, 4])
def test_build_trans_mat_c_matches_python(nthreads):
    """The threaded C restatement (the CPU baseline's body) against the
    Python one, pinned above: same counts, saturation included."""
    from oracle import partition as P
    w = synth.workload_partition(n_true=5, n_near=2, n_prior=0, L=150, n_uniques=200,
                                 seed=synth.SEED_BASE + 47)
    params = oracle.make_params(band=16)
    rs = oracle.RawSet(w.seqs, w.quals, w.reads, k=5)
    out = P.run_dada(rs, w.err, params)
    center_of, weight = learn.center_weight(out, w.reads)
    for wt, nq in ((weight, 41), (weight, 20), (np.where(weight > 0, (1 << 32) - 1, 0), 41)):
        a = learn.build_trans_mat(w.seqs, w.quals, center_of, wt, params, nq)
        b = learn.build_trans_mat_c(w.seqs, w.quals, center_of, wt, params, nq, nthreads)
        assert np.array_equal(a, b)
