"""GPU parity of the vectorized aligner (align_vectorized_with_buf,
nwalign.rs:761-1140) against the oracle, byte for byte.

The cases follow the reference's own tests: the rolling-buffer fuzz grid
(rolling_d16_matches_full_matrix, nwalign.rs:2873-2952), buffer reuse across
descending lengths (nwalign.rs:2959-2987), the fixed known-answer pairs
(nwalign.rs:2989-3062) and the chimera constructs on ASCII bytes.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

GRID_LENGTHS = [(1, 1), (1, 5), (5, 1), (2, 3), (10, 10), (10, 13), (13, 10), (17, 16), (16, 17),
                (33, 33), (40, 40), (40, 55), (55, 40), (64, 64), (100, 97), (97, 100),
                (128, 128), (240, 240), (160, 160), (240, 160), (160, 240), (300, 299)]
GRID_BANDS = [-1, 1, 2, 3, 4, 8, 15, 16, 17, 32, 64]


def _xorshift(state):
    """rng_next, nwalign.rs:2862-2869."""
    x = state
    x ^= (x << 13) & 0xFFFFFFFFFFFFFFFF
    x ^= x >> 7
    x ^= (x << 17) & 0xFFFFFFFFFFFFFFFF
    return x


def fuzz_grid():
    """The pairs of rolling_d16_matches_full_matrix (nwalign.rs:2873-2952)."""
    state = 0x2545F4914F6CDD1D
    cases = []
    for l1, l2 in GRID_LENGTHS:
        for band in GRID_BANDS:
            for trial in range(12):
                s1 = []
                for _ in range(l1):
                    state = _xorshift(state)
                    s1.append(1 + state % 4)
                if trial == 0 and l1 == l2:
                    s2 = list(s1)
                else:
                    v = list(s1[:min(l2, l1)]) if l2 <= l1 else list(s1)
                    v += [1] * (l2 - len(v))
                    for i in range(len(v)):
                        state = _xorshift(state)
                        if state % 16 < trial:
                            state = _xorshift(state)
                            v[i] = 1 + state % 4
                    s2 = v
                for eg in (0, -8):
                    cases.append((bytes(s1), bytes(s2), band, eg))
    return cases


def _run_grouped(ctx, oracle, cases):
    """Align cases on the GPU grouped by (band, end_gap) and compare."""
    groups = {}
    for s1, s2, band, eg in cases:
        groups.setdefault((band, eg), []).append((s1, s2))
    n = 0
    for (band, eg), pairs in groups.items():
        got = ctx.align_vectorized([p[0] for p in pairs], [p[1] for p in pairs], band=band,
                                   end_gap_p=eg)
        for (s1, s2), g in zip(pairs, got):
            want = oracle.align_vectorized(s1, s2, band=band, end_gap_p=eg)
            assert g == want, f"band={band} end_gap={eg} l1={len(s1)} l2={len(s2)}"
            n += 1
    return n


def test_fuzz_grid_supported(ctx, oracle, pkg):
    cases = [c for c in fuzz_grid() if pkg.vectorized_supported(len(c[0]), len(c[1]), c[2])]
    n = _run_grouped(ctx, oracle, cases)
    assert n > 3000, n


def test_wide_geometries_match_oracle(ctx, oracle, pkg):
    """Compressed bands wider than one wavefront (nw_wide.hip): unbanded
    pairs, wide bands, and length differences above ~90 at band 16, each in
    both orders (swap), with and without shared sequence."""
    rng = np.random.default_rng(17)
    s1s, s2s, bands = [], [], []
    for l1, l2, band in [(300, 299, -1), (250, 120, 16), (120, 250, 16), (400, 180, 32),
                         (200, 200, 128), (333, 310, 256), (90, 260, -1), (1000, 700, 16),
                         (61, 160, 16), (160, 61, 16)]:
        for related in (True, False):
            a = rng.integers(1, 5, size=max(l1, l2)).astype(np.uint8)
            x = a[:l1].copy()
            if related:
                y = a[:l2].copy()
                flip = rng.random(l2) < 0.02
                y[flip] = rng.integers(1, 5, size=int(flip.sum()))
            else:
                y = rng.integers(1, 5, size=l2).astype(np.uint8)
            assert pkg.vectorized_supported(l1, l2, band)
            s1s.append(bytes(x))
            s2s.append(bytes(y))
            bands.append(band)
    for band in sorted(set(bands)):
        idx = [i for i, b in enumerate(bands) if b == band]
        got = ctx.align_vectorized([s1s[i] for i in idx], [s2s[i] for i in idx], band=band)
        for i, g in zip(idx, got):
            assert g == oracle.align_vectorized(s1s[i], s2s[i], band=band), (len(s1s[i]),
                                                                           len(s2s[i]), band)


def test_buffer_reuse_descending_lengths(ctx, oracle):
    """rolling_d16_survives_buffer_reuse (nwalign.rs:2959-2987): successive calls
    on one context with varying lengths."""
    state = 0x9E3779B97F4A7C15
    for L in [400, 240, 239, 120, 33, 32, 7, 250, 3, 260]:
        s1 = []
        for _ in range(L):
            state = _xorshift(state)
            s1.append(1 + state % 4)
        s2 = list(s1)
        for k in range(0, L, 7):
            s2[k] = 1 + (s2[k] % 4)
        got = ctx.align_vectorized([bytes(s1)], [bytes(s2)], band=16)[0]
        assert got == oracle.align_vectorized(bytes(s1), bytes(s2), band=16), L


def test_known_pairs(ctx, oracle):
    enc = oracle.encode
    pairs = [("ACGTACGTACGT", "ACGTACGTACGT"), ("ACGTACGTACGT", "ACGTTCGTACGT"),
             ("ACGTACGTACGT", "ACGACGTACGT"), ("ACGTACGTACGT", "ACGTACGTACGTAC"),
             ("AGAAAAGGGGTTTAAAAAATTTTTCCCC", "AAAAAGGGGTTTAAAAAATTTTTCCCC")]
    s1 = [enc(a) for a, _ in pairs]
    s2 = [enc(b) for _, b in pairs]
    for band in (16, 8, 32):
        got = ctx.align_vectorized(s1, s2, band=band)
        for a, b, g in zip(s1, s2, got):
            assert g == oracle.align_vectorized(a, b, band=band)


def test_random_amplicon_pairs(ctx, oracle):
    """Indel/substitution variants at amplicon lengths, bands 8/16/32
    (sweep_vectorized_parity shapes, nwalign.rs:3069-3118, at 200-260 nt)."""
    rng = np.random.default_rng(11)
    s1s, s2s = [], []
    for _ in range(300):
        L = int(rng.integers(200, 261))
        a = rng.integers(1, 5, size=L).astype(np.uint8)
        b = list(a)
        for _ in range(int(rng.integers(0, 8))):
            p = int(rng.integers(0, len(b)))
            op = int(rng.integers(0, 3))
            if op == 0:
                b[p] = int(rng.integers(1, 5))
            elif op == 1:
                b.insert(p, int(rng.integers(1, 5)))
            elif len(b) > 3:
                del b[p]
        s1s.append(a.tobytes())
        s2s.append(bytes(b))
    for band in (8, 16, 32):
        got = ctx.align_vectorized(s1s, s2s, band=band)
        for a, b, g in zip(s1s, s2s, got):
            assert g == oracle.align_vectorized(a, b, band=band), (band, len(a), len(b))


def test_ascii_chimera_constructs(ctx, oracle):
    """The bimera search aligns ASCII bytes (remove_bimera.rs:70-75) with
    band = max_shift 16; constructs from chimera.rs:665-789."""
    L, M, R = b"ACGTACGTACGTACGTACGT", b"TTTTGGGGCCCCAAAATTTT", b"GATTACAGATTACAGATTAC"
    q = L + M + R
    parents = [L + b"A" * 20 + b"C" * 20, b"GC" * 20 + R, b"CA" * 10 + M + b"TG" * 10,
               L + b"C" * 20, b"G" * 20 + R]
    got = ctx.align_vectorized([q] * len(parents), parents, band=16)
    for p, g in zip(parents, got):
        assert g == oracle.align_vectorized(q, p, band=16)


def _dec(b: bytes) -> str:
    return "".join("-" if x == 45 else "ACGTN"[x - 1] for x in b)


def test_tie_cases_worked_by_hand_on_device(ctx, oracle):
    """The hand-derived co-optimal placements (tests/tie_cases.py) on the
    device aligners: vectorized with and without swap, scalar ends-free and
    homopolymer."""
    from tie_cases import TIE_CASES
    for kind, s1, s2, w0, w1, _, why in TIE_CASES:
        a, b = oracle.encode(s1), oracle.encode(s2)
        if kind == "vectorized":
            got = ctx.align_vectorized([a], [b], band=16)[0]
        elif kind == "endsfree":
            got = ctx.align_endsfree([a], [b], band=16)[0]
        else:
            got = ctx.align_endsfree([a], [b], homo_gap_p=-1, band=16)[0]
        assert (_dec(got[0]), _dec(got[1])) == (w0, w1), (kind, s1, s2, why)


def test_tie_rich_two_letter_pairs(ctx, oracle):
    """Every pair over {A, T} with lengths 1-5 and 1-7 (ties everywhere),
    both orders: the device equals the oracle, whose tie rules the
    hand-worked cases pin."""
    import itertools
    s1s, s2s = [], []
    for l1 in range(1, 6):
        for l2 in range(l1, l1 + 3):
            for x in itertools.product((1, 4), repeat=l1):
                for y in itertools.product((1, 4), repeat=l2):
                    s1s.append(bytes(x))
                    s2s.append(bytes(y))
    pairs = list(zip(s1s, s2s)) + list(zip(s2s, s1s))
    print(f"{len(pairs)} tie-rich pairs", flush=True)
    for band in (16, 2):
        got = ctx.align_vectorized([p[0] for p in pairs], [p[1] for p in pairs], band=band)
        for (a, b), g in zip(pairs, got):
            assert g == oracle.align_vectorized(a, b, band=band), (a, b, band)
    got = ctx.align_endsfree([p[0] for p in pairs], [p[1] for p in pairs], band=16)
    for (a, b), g in zip(pairs, got):
        assert g == oracle.align_endsfree(a, b, band=16), (a, b)
