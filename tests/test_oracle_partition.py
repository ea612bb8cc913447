"""The oracle's partition loop (oracle/partition.py, run_dada of dada.rs:749)
pinned by hand-worked cases of its steps and by properties the algorithm
fixes end to end.  CPU only."""
import numpy as np
import pytest

from dada2_rs_amd import synth


@pytest.fixture(scope="module")
def P():
    from oracle import partition
    return partition


def tiny_partition(oracle, P, reads):
    L = 20
    rng = np.random.default_rng(7)
    seqs = [rng.integers(1, 5, size=L).astype(np.uint8) for _ in reads]
    quals = [np.full(L, 30, np.uint8) for _ in reads]
    rs = oracle.RawSet(seqs, quals, np.asarray(reads, np.uint32), k=5)
    return P.Partition(rs, reads, None)


def test_init_and_swap_remove(oracle, P):
    B = tiny_partition(oracle, P, [5, 9, 9, 3, 1])
    # containers.rs:363-391: one cluster in index order; the last of the
    # maximal raws is the center (max_by_key keeps the last maximum)
    assert B.cl[0].raws == [0, 1, 2, 3, 4]
    assert B.cl[0].center == 2 and B.cl[0].reads == 27
    assert B.e_minmax == [-999.0] * 5
    # Vec::swap_remove: the last element fills the hole
    assert B.pop_raw(0, 1) == 1
    assert B.cl[0].raws == [0, 4, 2, 3] and B.cl[0].reads == 18
    assert B.pop_raw(0, 3) == 3
    assert B.cl[0].raws == [0, 4, 2]


def test_bud_picks_min_p_then_reads_and_skips_first_member(oracle, P):
    B = tiny_partition(oracle, P, [50, 7, 7, 9, 4])
    B.cl[0].raws = [0, 1, 2]
    B.cl[0].reads = 64
    c1 = P.Bi()
    c1.raws = [3, 4]
    c1.reads = 13
    c1.center = 3
    B.cl.append(c1)
    for r in range(5):
        B.comp[r] = P.Comp(0, r, 1e-3, 2)
    # raw 0 (cluster 0's center) seeds the scan with its p; raw 3 sits in
    # position 0 of cluster 1 and is never a candidate, even with the smallest p
    B.p = [1.0, 1e-50, 1e-50, 1e-90, 1e-49]
    ni = P._bud(B, omega_a=1e-40, omega_p=1e-4, min_fold=1.0, min_hamming=1, min_abund=1)
    # raws 1 and 2 tie on p and reads: the first one scanned (strict <) wins
    assert ni == 2
    new = B.cl[2]
    assert new.raws == [1] and new.center == 1 and new.birth_type == 1 and new.birth_from == 0
    assert new.birth_pval == 1e-50 * 5
    assert new.birth_e == 1e-3 * 64
    assert new.birth_fold == 7 / (1e-3 * 64)
    assert B.cl[0].raws == [0, 2] and B.cl[0].reads == 57
    # min_hamming / min_abund / min_fold exclude candidates
    B.p = [1.0, 1.0, 1e-60, 1e-90, 1e-49]
    assert P._bud(B, 1e-40, 1e-4, 1.0, min_hamming=3, min_abund=1) is None
    assert P._bud(B, 1e-40, 1e-4, 1.0, min_hamming=1, min_abund=8) is None
    assert P._bud(B, 1e-40, 1e-4, min_fold=1e6, min_hamming=1, min_abund=1) is None


def test_bud_reads_break_p_ties_and_prior_path(oracle, P):
    B = tiny_partition(oracle, P, [50, 7, 9, 3])
    for r in range(4):
        B.comp[r] = P.Comp(0, r, 1e-3, 2)
    B.p = [1.0, 1e-50, 1e-50, 1e-9]
    assert P._bud(B, 1e-40, 1e-4, 1.0, 1, 1) == 1
    assert B.cl[1].raws == [2]  # more reads wins the tie
    # no abundance bud: a prior raw with p < omega_p buds (birth type 2)
    B = tiny_partition(oracle, P, [50, 7, 9, 3])
    for r in range(4):
        B.comp[r] = P.Comp(0, r, 1e-3, 2)
    B.prior = [False, False, False, True]
    B.p = [1.0, 1e-30, 1e-30, 1e-9]
    ni = P._bud(B, 1e-40, 1e-4, 1.0, 1, 1)
    assert ni == 1 and B.cl[1].raws == [3] and B.cl[1].birth_type == 2
    assert B.cl[1].birth_pval == 1e-9 and B.cl[1].birth_from == 0


def test_shuffle_pass_moves_to_max_expected_and_keeps_centers(oracle, P):
    B = tiny_partition(oracle, P, [40, 30, 5, 5, 5])
    B.cl[0].raws = [0, 2, 3, 4]
    B.cl[0].reads = 55
    c1 = P.Bi()
    c1.raws = [1]
    c1.reads = 30
    c1.center = 1
    B.cl.append(c1)
    # expected reads lambda * cluster reads, exact in binary:
    #   raw 0: 0.25*55 < 1.0*30, but it is cluster 0's center: stays
    #   raw 1: cluster 1's center (no comp there; cluster 0 would win): stays
    #   raw 2: 0.25*55 = 13.75 < 0.5*30 = 15: moves
    #   raw 3: 0.375*55 = 20.625 == 0.6875*30: tie, strict > keeps cluster 0
    #   raw 4: 0.25*55 = 13.75 > 0.125*30: stays
    B.cl[0].comp = [P.Comp(0, r, lam, 1) for r, lam in
                    [(0, 0.25), (1, 1.0), (2, 0.25), (3, 0.375), (4, 0.25)]]
    B.cl[1].comp = [P.Comp(1, 0, 1.0, 3), P.Comp(1, 2, 0.5, 1), P.Comp(1, 3, 0.6875, 1),
                    P.Comp(1, 4, 0.125, 1)]
    assert P.shuffle_pass(B) == 1
    # the pass walks each cluster from its end; raw 2 (position 1) is
    # swap-removed, raw 4 takes its slot
    assert B.cl[0].raws == [0, 4, 3] and B.cl[0].reads == 50
    assert B.cl[1].raws == [1, 2] and B.cl[1].reads == 35
    assert B.comp[2].i == 1 and B.comp[2].lam == 0.5
    assert B.cl[0].update_e and B.cl[1].update_e


def test_p_update_sets_greedy_locks(oracle, P):
    B = tiny_partition(oracle, P, [100, 2, 50])
    for r in range(3):
        B.comp[r] = P.Comp(0, r, [1.0, 0.05, 1e-4][r], [0, 1, 1][r])
    P._p_update(B, greedy=True, detect_singletons=False)
    # center 0 (100 reads) explains raw 1 (100 * 0.05 = 5 > 2) but not raw 2
    assert B.lock == [True, True, False]
    assert B.cl[0].check_locks is False and B.cl[0].update_e is False
    assert B.p[0] == 1.0  # hamming 0
    assert B.p[1] == oracle.get_pA(2, 0.05, 1, 152)


@pytest.mark.parametrize("seed_off,n_true,n_near,n_prior,L,n", [
    (5, 8, 2, 1, 150, 300), (6, 12, 4, 2, 200, 500), (7, 5, 3, 0, 120, 250)])
def test_recovers_planted_variants(oracle, P, seed_off, n_true, n_near, n_prior, L, n):
    w = synth.workload_partition(n_true, n_near, n_prior, L, n, seed=synth.SEED_BASE + seed_off)
    rs = oracle.RawSet(w.seqs, w.quals, w.reads, k=5)
    out = P.run_dada(rs, w.err, oracle.make_params(band=16), prior=w.extra["prior"])
    centers = sorted(c["center"] for c in out.clusters)
    assert centers == sorted(int(x) for x in w.extra["planted"])
    center_of = np.array([out.clusters[k]["center"] for k in out.cluster_of])
    assert np.array_equal(center_of, w.extra["origin"])
    # corrected raws map to their cluster; a raw is uncorrected exactly when
    # its final p is below omega_c (1e-40)
    corrected = out.p >= 1e-40
    assert np.array_equal(out.map[corrected], out.cluster_of[corrected].astype(np.int32))
    assert np.all(out.map[~corrected] == -1)
    assert np.all(corrected[w.extra["planted"]])
    births = sorted(c["birth_type"] for c in out.clusters)
    assert births == [0] + [1] * (n_true + n_near - 1) + [2] * n_prior
    # the final partition is converged: one more move pass moves nothing
    assert P.shuffle_pass(out.partition) == 0
    assert out.nalign >= n and out.nshroud > 0


def test_limits(oracle, P):
    w = synth.workload_partition(6, 2, 0, 120, 200, seed=synth.SEED_BASE + 9)
    rs = oracle.RawSet(w.seqs, w.quals, w.reads, k=5)
    one = P.run_dada(rs, w.err, oracle.make_params(), omega_a=0.0, omega_p=0.0)
    assert len(one.clusters) == 1 and np.all(one.cluster_of == 0)
    rs = oracle.RawSet(w.seqs, w.quals, w.reads, k=5)
    three = P.run_dada(rs, w.err, oracle.make_params(), max_clust=3)
    assert len(three.clusters) == 3


def test_extend_err_repeats_the_last_column(oracle, pkg):
    """dada_uniques_cached's extension (dada.rs:337-384): a Q0-40 model met
    with Q45 grows to 46 columns, columns 41..45 copies of column 40; data
    within the model leaves it untouched.  The package's host mirror does
    the same."""
    import warnings
    from oracle import partition as P
    rng = np.random.default_rng(3)
    err = rng.random((16, 41))
    ext = P.extend_err(err, 45)
    assert ext.shape == (16, 46)
    assert np.array_equal(ext[:, :41], err)
    for q in range(41, 46):
        assert np.array_equal(ext[:, q], err[:, 40])
    assert P.extend_err(err, 40) is not None and P.extend_err(err, 40).shape == (16, 41)
    with warnings.catch_warnings(record=True) as w:
        warnings.simplefilter("always")
        mine = pkg.extend_err(err, 45)
        assert any("Q45" in str(x.message) for x in w)
    assert np.array_equal(mine, ext)
    assert np.array_equal(pkg.extend_err(err, 40), err)
