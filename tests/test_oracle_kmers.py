"""Pins of the oracle's k-mer functions (kmers.rs) by the reference's own
invariants (kmers.rs:347-401) and by brute force."""
from collections import Counter

import numpy as np


def lcg_seq(n, seed):
    """make_seq, kmers.rs:335-345."""
    s = seed
    out = []
    for _ in range(n):
        s = (s * 6364136223846793005 + 1442695040888963407) & 0xFFFFFFFFFFFFFFFF
        out.append(((s >> 33) % 4) + 1)
    return bytes(out)


def brute_counts(seq, k):
    c = Counter()
    for i in range(len(seq) - k + 1):
        w = seq[i:i + k]
        if all(1 <= b <= 4 for b in w):
            idx = 0
            for b in w:
                idx = idx * 4 + (b - 1)
            c[idx] += 1
    return c


def test_kmer8_matches_brute_force(oracle):
    for k in (3, 5, 7):
        for seed in range(4):
            seq = lcg_seq(180, seed)
            v = oracle.kmer8(seq, k)
            want = np.zeros(1 << (2 * k), np.int64)
            for idx, n in brute_counts(seq, k).items():
                want[idx] = min(n, 255)
            assert np.array_equal(v.astype(np.int64), want)


def test_kmer_skips_non_acgt_windows(oracle):
    seq = bytes([1, 2, 3, 5, 4, 1, 2, 3, 4])
    v = oracle.kmer16(seq, 3)
    assert int(v.sum()) == sum(brute_counts(seq, 3).values()) == 4
    ko = oracle.kord(seq, 3)
    assert list(ko[1:4]) == [0, 0, 0]  # windows through the N map to 0 (kmers.rs:132-134)


def test_dist_formula_and_self_distance_zero(oracle):
    for seed in range(8):
        a, b = lcg_seq(250, seed), lcg_seq(240, seed + 100)
        ka, kb = oracle.kmer8(a), oracle.kmer8(b)
        dot = int(np.minimum(ka, kb).astype(np.int64).sum())
        want = 1.0 - dot / float(min(len(a), len(b)) - 5 + 1)
        assert oracle.kmer_dist8(ka, len(a), kb, len(b)) == want
        assert oracle.kmer_dist8(ka, len(a), ka, len(a)) == 0.0  # kmers.rs:378-385


def test_sparse_matches_dense_bits(oracle):
    """sparse_dist_matches_dense, kmers.rs:367-387."""
    for k in (6, 7, 8):
        for seed in range(6):
            a, b = lcg_seq(250, seed), lcg_seq(250, seed + 100)
            ka, kb = oracle.kmer8(a, k), oracle.kmer8(b, k)
            d = oracle.kmer_dist8(ka, 250, kb, 250, k)
            s = oracle.kmer_dist8_sparse(ka, 250, kb, 250, k)
            assert np.float64(d).view(np.uint64) == np.float64(s).view(np.uint64)


def test_overflow_returns_minus_one(oracle):
    """sparse_dist_overflow_matches_dense, kmers.rs:390-401: a 300-A
    homopolymer saturates the all-A bucket."""
    seq = bytes([1] * 300)
    v = oracle.kmer8(seq, 6)
    assert oracle.kmer_dist8(v, 300, v, 300, 6) == -1.0
    assert oracle.kmer_dist8_sparse(v, 300, v, 300, 6) == -1.0
    u = oracle.kmer16(seq, 6)
    assert oracle.kmer_dist(u, 300, u, 300, 6) == 0.0


def test_kord_dist(oracle):
    a = lcg_seq(100, 1)
    b = bytearray(a)
    b[50] = (b[50] % 4) + 1
    oa, ob = oracle.kord(a), oracle.kord(bytes(b))
    eq = int((oa == ob).sum())
    assert eq == 96 - 5  # one substitution spoils the k windows covering it
    assert oracle.kord_dist(oa, 100, ob, 100) == 1.0 - eq / 96.0
    assert oracle.kord_dist(oa, 100, ob[:-1], 99) == -1.0  # unequal lengths
