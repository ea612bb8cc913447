"""Pins of the oracle's compare map, lambda, p-values and bimera search."""
import math

import numpy as np
import pytest

from dada2_rs_amd import synth


def test_lambda_identical_is_product_of_diagonal(oracle):
    """compute_lambda for an identical pair equals get_self's product of
    err[diag] (pval.rs:361-366), in position order."""
    w = synth.subsample(synth.workload_plain(), 50)
    rs = oracle.RawSet(w.seqs, w.quals, w.reads)
    p = oracle.make_params()
    for i in range(5):
        s, buf = oracle.sub_new(rs, i, i, p)
        assert s.nsubs == 0
        lam = oracle.compute_lambda(rs, i, s, w.err, 41)
        want = 1.0
        for nt, q in zip(w.seqs[i], w.quals[i]):
            want *= w.err[5 * (nt - 1), q]
        assert np.float64(lam).view(np.uint64) == np.float64(want).view(np.uint64)


def test_lambda_substitution_overrides_query_position(oracle):
    rng = np.random.default_rng(1)
    a = rng.integers(1, 5, size=120).astype(np.uint8)
    b = a.copy()
    b[40] = (b[40] % 4) + 1
    q = rng.integers(0, 41, size=120).astype(np.uint8)
    rs = oracle.RawSet([a, b], [q, q])
    err = synth.seeded_err(np.random.default_rng(9))
    s, _ = oracle.sub_new(rs, 0, 1, oracle.make_params())
    assert s.nsubs == 1 and s.pos[0] == 40
    lam = oracle.compute_lambda(rs, 1, s, err, 41)
    want = 1.0
    for pos in range(120):
        t = 4 * (a[pos] - 1) + (b[pos] - 1)
        want *= err[t, q[pos]]
    assert lam == want
    assert oracle.compute_lambda(rs, 1, None, err, 41) == 0.0


def test_parallel_matches_serial(oracle):
    """parallel_matches_serial (cluster.rs:1896-1922): the threaded map and the
    serial one agree bit for bit; err = 0.001 everywhere, ncol 41."""
    w = synth.subsample(synth.workload_batched(), 1500, n_centers=4)
    rs = oracle.RawSet(w.seqs, w.quals, w.reads)
    err = np.full((16, 41), 0.001)
    for greedy in (True, False):
        for c in w.centers[:3]:
            a = oracle.compare_center(rs, int(c), err, 41, oracle.make_params(), greedy=greedy,
                                      nthreads=1)
            b = oracle.compare_center(rs, int(c), err, 41, oracle.make_params(), greedy=greedy,
                                      nthreads=8, grain=32)
            assert np.array_equal(a[0].view(np.uint64), b[0].view(np.uint64))
            assert np.array_equal(a[1], b[1]) and np.array_equal(a[2], b[2])


def test_greedy_skip_semantics(oracle):
    """cluster.rs:295-299: more abundant or locked raws are skipped with
    lambda 0 and no hamming."""
    w = synth.subsample(synth.workload_plain(), 200)
    locks = np.zeros(200, np.uint8)
    locks[150:160] = 1
    rs = oracle.RawSet(w.seqs, w.quals, w.reads, locks=locks)
    c = 20
    lam, ham, skp, _, _ = oracle.compare_center(rs, c, w.err, 41, oracle.make_params())
    want = (w.reads > w.reads[c]) | (locks == 1)
    assert np.array_equal(skp.astype(bool), want)
    assert np.all(lam[want] == 0.0) and np.all(ham[want] == 0xFFFFFFFF)
    assert skp[c] == 0 and ham[c] == 0


def test_poisson_tail_identities(oracle):
    for e in (0.3, 2.0, 17.5, 140.0):
        # P(X >= 1) = 1 - exp(-e) (prior: no conditioning)
        assert math.isclose(oracle.calc_pA(1, e, True), -math.expm1(-e), rel_tol=1e-12)
        # successive tails differ by the pmf
        for k in (2, 5, 20, 160):
            if oracle.calc_pA(k, e, True) > 0.5:
                continue  # difference of two tails near 1 cancels
            d = oracle.calc_pA(k, e, True) - oracle.calc_pA(k + 1, e, True)
            pmf = math.exp(-e + k * math.log(e) - math.lgamma(k + 1))
            assert math.isclose(d, pmf, rel_tol=1e-9, abs_tol=1e-300)
    # conditioning on presence (pval.rs:229-240)
    e = 3.0
    assert math.isclose(oracle.calc_pA(4, e, False), oracle.calc_pA(4, e, True) / -math.expm1(-e),
                        rel_tol=1e-14)


def test_small_lambda_matches_reference_series(oracle):
    """For tiny expectations the upper tail is the reference's own series
    (poisson_upper_tail_direct, pval.rs:329-353)."""
    for reads, e in [(2, 1e-30), (5, 1e-12), (3, 1e-200)]:
        a = oracle.calc_pA(reads, e, True)
        b = oracle.poisson_upper_tail_direct(reads, e)
        assert math.isclose(a, b, rel_tol=1e-10)


def test_get_pA_sentinels(oracle):
    L = oracle.lib()
    assert L.orc_get_pA(1, 0, 1e-3, 3, 100, 0) == 1.0   # singleton
    assert L.orc_get_pA(5, 0, 1e-3, 0, 100, 0) == 1.0   # hamming 0
    assert L.orc_get_pA(5, 0, 0.0, 3, 100, 0) == 0.0    # lambda 0


def _chim(o):
    L, M, R = b"ACGTACGTACGTACGTACGT", b"TTTTGGGGCCCCAAAATTTT", b"GATTACAGATTACAGATTAC"
    return L, M, R


def test_plain_bimera_is_flagged(oracle):
    """pooled_diagnostics_flags_plain_bimera (chimera.rs:772-789)."""
    L, _, R = _chim(oracle)
    q = L + R
    pa, pb = L + b"C" * 20, b"G" * 20 + R
    p = oracle.bimera_params()
    assert oracle.is_bimera(q, [pa, pb], p)
    al0, al1 = oracle.align_vectorized(q, pa, band=16)
    assert oracle.get_lr(al0, al1, False, 16)[0] >= 20


def test_trimera_left_right_twenty(oracle):
    """pooled_diagnostics_detects_trimera (chimera.rs:665-726): the left and
    right parents each explain 20 bases, so the query is not a bimera."""
    L, M, R = _chim(oracle)
    q = L + M + R
    pa = L + b"A" * 20 + b"C" * 20
    pb = b"GC" * 20 + R
    pc = b"CA" * 10 + M + b"TG" * 10
    p = oracle.bimera_params()
    assert not oracle.is_bimera(q, [pa, pb, pc], p)
    la = oracle.get_lr(*oracle.align_vectorized(q, pa, band=16), False, 16)
    lb = oracle.get_lr(*oracle.align_vectorized(q, pb, band=16), False, 16)
    assert max(la[0], lb[0]) == 20 and max(la[1], lb[1]) == 20


def test_snp_variant_distance_three(oracle):
    """pooled_diagnostics_rejects_snp_variant (chimera.rs:733-768)."""
    pa = bytearray(b"ACGTTGCAACGTTGCAACGTTGCAACGTTGCAACGTTGCAACGTTGCAACGTTGCAACGT")
    q = bytearray(pa)
    for pos in (20, 30, 40):
        q[pos] = ord("C") if q[pos] == ord("A") else ord("A")
    al0, al1 = oracle.align_vectorized(bytes(q), bytes(pa), band=16)
    assert oracle.get_ham_endsfree(al0, al1) == 3


def test_table_bimera2_single_sample(oracle):
    L, _, R = _chim(oracle)
    seqs = [L + R, L + b"C" * 20, b"G" * 20 + R, b"T" * 40]
    mat = np.array([[2, 100, 100, 1]], np.uint32)
    p = oracle.bimera_params()
    nflag, nsam = oracle.table_bimera2(mat, seqs, 1.5, 2, p, nthreads=2)
    assert list(nsam) == [1, 1, 1, 1]
    assert nflag[0] == 1 and nflag[1] == 0 and nflag[2] == 0
    # the same verdict as is_bimera on the qualifying parents
    assert oracle.is_bimera(seqs[0], [seqs[1], seqs[2]], p)
