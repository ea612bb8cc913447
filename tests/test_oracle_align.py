"""Pins of the oracle's aligners (nwalign.rs) that do not run the reference:
known-answer scores from the reference's tests and dev oracle, exhaustive
enumeration on tiny inputs, rolling == full-matrix on the reference fuzz grid
(with zero reads of unwritten cells), vectorized == ends-free score parity,
and the homopolymer dispatch divergence."""
import itertools

import numpy as np
import pytest

from test_gpu_align import fuzz_grid


def enc(o, s):
    return o.encode(s)


@pytest.mark.parametrize("s1,s2,score", [
    ("ACGTACGTACGT", "ACGTACGTACGT", 60),   # nwalign.rs:2345-2349
    ("ACGTACGTACGT", "ACGTTCGTACGT", 51),   # nwalign.rs:2351-2356
    ("ACGTACGTACGT", "ACGACGTACGT", 47),    # nwalign.rs:2358-2363
])
def test_endsfree_known_scores_band16(oracle, s1, s2, score):
    a = oracle.align_endsfree(enc(oracle, s1), enc(oracle, s2), band=16)
    assert oracle.score_alignment(*a) == score


@pytest.mark.parametrize("s1,s2,score", [
    # nwalign.rs:2244-2273 and dev/wfa-ffi-oracle/oracle.c:62-75 (optima per its README)
    ("AACAGCGCAAACCAACTCGCTAGCTAGCAAAATCTTGTGTTTCTGCCTAGCG",
     "ACAGCGCAAACCAACTCGCTAGCTAGCAAAATCTTGTGTTTCTGCCTAGCG", 255),
    ("GCGG", "CG", 10),
    ("AACTAGACATCAAAGCCACAGAATTTGGCAGGAGATACAAGAAGTACTCCAGCTC",
     "AAACTAGACATCAAAGCCACAGAATTTGGCAGGAGATACAAGAAGTACTCCAGCTC", 275),
    ("TACACGCTACCTTGGAGACCTCCTGTGCTTGCAGCCCATCAATCTTTTTACACAAGG",
     "TACACGCTACCTTGGAGACCTCCTGTGCTTGCAGCCCATCAATCTTTTTACACAGG", 272),
    ("TAACGACGAACGAGTTTCTAAACCAAAGAACGTAACAAGGAGTTGCCTCTTAGCCTTGGCCTACGCAGGAA",
     "TAACGACGAACGAGTTTCTAAACCAAAGAACGTAACAAGGAGTTGCCTCTTAGCCTTGGCCTACGCAGGGAA", 347),
])
def test_unbanded_known_optima(oracle, s1, s2, score):
    a, b = enc(oracle, s1), enc(oracle, s2)
    assert oracle.score_alignment(*oracle.align_endsfree(a, b, band=-1)) == score
    assert oracle.score_alignment(*oracle.align_vectorized(a, b, band=-1)) == score


def _best_endsfree(a, b, match=5, mm=-4, gap=-8):
    """Exhaustive maximum over all global alignments scored ends-free."""
    best = None

    def rec(i, j, al0, al1):
        nonlocal best
        if i == len(a) and j == len(b):
            # score with the reference's ends-free rule (nwalign.rs:1814-1855)
            n = len(al0)
            s = 0
            for k in range(n):
                g0, g1 = al0[k] is None, al1[k] is None
                if not g0 and not g1:
                    s += match if al0[k] == al1[k] else mm
                elif g0:
                    if any(x is not None for x in al0[:k]) and any(x is not None for x in al0[k + 1:]):
                        s += gap
                else:
                    if any(x is not None for x in al1[:k]) and any(x is not None for x in al1[k + 1:]):
                        s += gap
            best = s if best is None else max(best, s)
            return
        if i < len(a) and j < len(b):
            rec(i + 1, j + 1, al0 + [a[i]], al1 + [b[j]])
        if j < len(b):
            rec(i, j + 1, al0 + [None], al1 + [b[j]])
        if i < len(a):
            rec(i + 1, j, al0 + [a[i]], al1 + [None])

    rec(0, 0, [], [])
    return best


def test_exhaustive_small(oracle):
    """Unbanded align_endsfree is the ends-free optimum.  (The vectorized
    aligner is not on tiny inputs: its end-gap recalculation starts on the
    row after an edge is reached, so e.g. 1x1 mismatches keep the diagonal;
    the reference's own sweep starts at length 5, nwalign.rs:3080-3096.)"""
    rng = np.random.default_rng(5)
    for _ in range(60):
        l1, l2 = int(rng.integers(1, 6)), int(rng.integers(1, 6))
        a = bytes(rng.integers(1, 5, size=l1).astype(np.uint8))
        b = bytes(rng.integers(1, 5, size=l2).astype(np.uint8))
        want = _best_endsfree(a, b)
        assert oracle.score_alignment(*oracle.align_endsfree(a, b, band=-1)) == want
    v = oracle.align_vectorized(bytes([4]), bytes([3]), band=-1)
    assert oracle.score_alignment(*v) == -4 and _best_endsfree(bytes([4]), bytes([3])) == 0


def test_rolling_equals_full_matrix_on_reference_grid(oracle):
    """rolling_d16_matches_full_matrix (nwalign.rs:2873-2952), plus: the DP never
    reads a cell that was not written in the same call."""
    n = 0
    buf = oracle.Buf()
    for s1, s2, band, eg in fuzz_grid():
        a = oracle.align_vectorized(s1, s2, band=band, end_gap_p=eg, buf=buf)
        f0, f1, stale = oracle.align_vectorized(s1, s2, band=band, end_gap_p=eg, full=True)
        assert a == (f0, f1), (len(s1), len(s2), band, eg)
        assert stale == 0
        n += 1
    assert n > 5000


@pytest.mark.parametrize("band", [8, 16, 32])
def test_vectorized_matches_endsfree_score_sweep(oracle, band):
    """sweep_vectorized_parity (nwalign.rs:3069-3118): 0 score mismatches."""
    st = 0x9E3779B97F4A7C15

    def rng(m):
        nonlocal st
        st = (st * 6364136223846793005 + 1) & 0xFFFFFFFFFFFFFFFF
        return (st >> 33) % m

    fails = 0
    for _ in range(3000):
        L = 5 + rng(60)
        s1 = [1 + rng(4) for _ in range(L)]
        s2 = list(s1)
        for _ in range(rng(5)):
            if not s2:
                break
            p = rng(len(s2))
            op = rng(3)
            if op == 0:
                s2[p] = 1 + rng(4)
            elif op == 1:
                s2.insert(p, 1 + rng(4))
            else:
                del s2[p]
        if len(s2) < 3:
            continue
        ef = oracle.align_endsfree(bytes(s1), bytes(s2), band=band)
        ve = oracle.align_vectorized(bytes(s1), bytes(s2), band=band)
        fails += oracle.score_alignment(*ef) != oracle.score_alignment(*ve)
    assert fails == 0


def test_geometry_envelope_covers_configs(oracle):
    """Active columns stay inside [1, ncol-2] for every length pair of the
    benchmark workloads (240-261 nt at band 16, 1420-1480 nt at bands 16/32),
    with no reads of unwritten cells."""
    rng = np.random.default_rng(0)
    for lo, hi, band in [(239, 261, 16), (1420, 1480, 32), (1420, 1480, 16)]:
        for l1 in range(lo, hi + 1, 3):
            for l2 in range(lo, hi + 1, 5):
                ncol, mn, mx = oracle.vectorized_geometry(l1, l2, band)
                assert 1 <= mn and mx <= ncol - 2
        for _ in range(4):
            l1, l2 = int(rng.integers(lo, hi + 1)), int(rng.integers(lo, hi + 1))
            a = rng.integers(1, 5, size=l1).astype(np.uint8)
            b = np.concatenate([a[:l2], rng.integers(1, 5, size=max(0, l2 - l1)).astype(np.uint8)])
            f0, f1, stale = oracle.align_vectorized(a, b, band=band, full=True)
            assert stale == 0
            assert (f0, f1) == oracle.align_vectorized(a, b, band=band)


def test_homopolymer_penalty_routes_to_homo_aligner(oracle):
    """homo_gap_penalty_routes_to_homopolymer_aligner (nwalign.rs:3124-3188)."""
    s1 = enc(oracle, "AGAAAAGGGGTTTAAAAAATTTTTCCCC")
    s2 = enc(oracle, "AAAAAGGGGTTTAAAAAATTTTTCCCC")
    p = oracle.make_params(homo_gap_p=-1)
    want = oracle.align_endsfree_homo(s1, s2, p)
    assert oracle.align_vectorized(s1, s2, band=16) != want
    rs = oracle.RawSet([s1, s2], reads=[10, 5])
    ok, _, _, al, method = oracle.raw_align(rs, 0, 1, p)
    assert ok and al == want and method == 2


def test_gapless_dispatch_on_equal_kmer_and_kord(oracle):
    """Equal-length substitution-only pairs take the gapless path
    (nwalign.rs:1324) when kord_dist equals kmer_dist."""
    rng = np.random.default_rng(2)
    a = rng.integers(1, 5, size=250).astype(np.uint8)
    b = a.copy()
    b[125] = (b[125] % 4) + 1
    rs = oracle.RawSet([a, b])
    ok, kd, ko, al, method = oracle.raw_align(rs, 0, 1, oracle.make_params())
    assert ok
    if abs(ko - kd) < np.finfo(np.float64).eps:
        assert method == 0 and al == (a.tobytes(), b.tobytes())
    else:
        assert method == 1


def test_al2subs_counts_mismatches(oracle):
    rng = np.random.default_rng(4)
    for _ in range(50):
        a = rng.integers(1, 5, size=80).astype(np.uint8)
        b = list(a)
        for _ in range(5):
            p = int(rng.integers(0, len(b)))
            op = int(rng.integers(0, 3))
            if op == 0:
                b[p] = int(rng.integers(1, 5))
            elif op == 1:
                b.insert(p, int(rng.integers(1, 5)))
            else:
                del b[p]
        al0, al1 = oracle.align_vectorized(a, bytes(b), band=16)
        s = oracle.al2subs(al0, al1)
        mism = [i for i, (x, y) in enumerate(zip(al0, al1)) if x != 45 and y != 45 and x != y]
        assert len(s["pos"]) == len(mism)
        assert s["len0"] == len(a)
        # map: reference position -> query position or 9999 at a query gap
        q = -1
        r = -1
        for x, y in zip(al0, al1):
            if y != 45:
                q += 1
            if x != 45:
                r += 1
                assert s["map"][r] == (q if y != 45 else 9999)


def test_standard_nw_penalises_end_gaps(oracle):
    a, b = enc(oracle, "ACGTACGTACGT"), enc(oracle, "CGTACGTACGT")
    std = oracle.align_standard(a, b, band=-1)
    ef = oracle.align_endsfree(a, b, band=-1)
    assert oracle.score_alignment(*ef) == 55
    assert std[0][0] != 45 and std[1][0] == 45  # the leading gap is still placed


def _dec(b: bytes) -> str:
    return "".join("-" if x == 45 else "ACGTN"[x - 1] for x in b)


def _score(al0: str, al1: str) -> int:
    """Ends-free score with free leading and trailing gaps (5 / -4 / -8)."""
    lo = min(i for i in range(len(al0)) if al0[i] != "-" and al1[i] != "-")
    hi = max(i for i in range(len(al0)) if al0[i] != "-" and al1[i] != "-")
    s = 0
    for a, b in zip(al0[lo:hi + 1], al1[lo:hi + 1]):
        s += -8 if "-" in (a, b) else (5 if a == b else -4)
    return s


@pytest.mark.parametrize("case", __import__("tie_cases").TIE_CASES, ids=lambda c: f"{c[0]}-{c[1]}-{c[2]}")
def test_tie_cases_worked_by_hand(oracle, case):
    """Co-optimal placements fixed by the tie-break lines, each derived by
    hand on the full matrix (tests/tie_cases.py): the oracle reproduces
    them string for string."""
    kind, s1, s2, w0, w1, score, _ = case
    a, b = enc(oracle, s1), enc(oracle, s2)
    if kind == "vectorized":
        got = oracle.align_vectorized(a, b, band=16)
    elif kind == "endsfree":
        got = oracle.align_endsfree(a, b, band=16)
    else:
        got = oracle.align_endsfree_homo(a, b, oracle.make_params(band=16, homo_gap_p=-1))
    assert (_dec(got[0]), _dec(got[1])) == (w0, w1)
    if kind != "homo":
        assert _score(w0, w1) == score
