"""GPU parity of the bimera parent search against the oracle: get_lr
(chimera.rs:99-175) and get_ham_endsfree (chimera.rs:54-86) per pair,
is_bimera_with_buf (chimera.rs:218-292) per query and table_bimera2
(chimera.rs:317-433) per column, including the reference's own constructs
(chimera.rs:665-789)."""
import numpy as np
import pytest

from dada2_rs_amd import synth

pytestmark = pytest.mark.gpu


def _chim():
    return b"ACGTACGTACGTACGTACGT", b"TTTTGGGGCCCCAAAATTTT", b"GATTACAGATTACAGATTAC"


@pytest.mark.parametrize("one_off", [False, True])
def test_pairs_match_oracle(ctx, oracle, pkg, one_off):
    bs = synth.workload_bimera(n_pool=30, n_queries=40, parents_per_query=6, lmin=240,
                               lmax=260, seed=synth.SEED_BASE + 50)
    qs = [q for q, ps in zip(bs.queries, bs.parent_lists) for _ in ps]
    ps = [p for pl in bs.parent_lists for p in pl]
    params = pkg.bimera_params(allow_one_off=one_off)
    lr, ham = ctx.bimera_pairs(qs, ps, params)
    for i, (q, p) in enumerate(zip(qs, ps)):
        al0, al1 = oracle.align_vectorized(q, p, band=16)
        assert tuple(lr[i]) == oracle.get_lr(al0, al1, one_off, 16), i
        assert ham[i] == oracle.get_ham_endsfree(al0, al1), i


def test_reference_constructs(ctx, oracle, pkg):
    L, M, R = _chim()
    p = pkg.bimera_params()
    op = oracle.bimera_params()
    cases = [(L + R, [L + b"C" * 20, b"G" * 20 + R]),                                  # bimera
             (L + M + R, [L + b"A" * 20 + b"C" * 20, b"GC" * 20 + R, b"CA" * 10 + M + b"TG" * 10])]
    got = ctx.is_bimera_batch([c[0] for c in cases], [c[1] for c in cases], p)
    want = [oracle.is_bimera(q, par, op) for q, par in cases]
    assert list(got) == want == [True, False]


@pytest.mark.parametrize("one_off", [False, True])
def test_is_bimera_batch(ctx, oracle, pkg, one_off):
    bs = synth.workload_bimera(n_pool=40, n_queries=80, parents_per_query=8, lmin=240, lmax=260,
                               div=0.02, seed=synth.SEED_BASE + 51)
    params = pkg.bimera_params(allow_one_off=one_off)
    oparams = oracle.bimera_params(allow_one_off=one_off)
    got = ctx.is_bimera_batch(bs.queries, bs.parent_lists, params)
    want = np.array([oracle.is_bimera(q, ps, oparams) for q, ps in zip(bs.queries, bs.parent_lists)])
    assert np.array_equal(got, want)
    assert want.any() and not want.all()


def test_is_bimera_pool_matches_batch(ctx, oracle, pkg):
    """Parents as indices into one pool: the same flags as per-parent bytes
    and as the oracle."""
    bs = synth.workload_bimera(n_pool=30, n_queries=60, parents_per_query=10, lmin=240,
                               lmax=260, div=0.02, seed=synth.SEED_BASE + 54)
    for one_off in (False, True):
        params = pkg.bimera_params(allow_one_off=one_off)
        pooled = ctx.is_bimera_pool(bs.queries, bs.pool, bs.parent_idx, params)
        listed = ctx.is_bimera_batch(bs.queries, bs.parent_lists, params)
        op = oracle.bimera_params(allow_one_off=one_off)
        want = np.array([oracle.is_bimera(q, ps, op) for q, ps in zip(bs.queries, bs.parent_lists)])
        assert np.array_equal(pooled, want) and np.array_equal(listed, want)
    assert want.any() and not want.all()
    # queries without parents, and an empty search
    assert not ctx.is_bimera_pool(bs.queries[:3], bs.pool, [[], [], []]).any()
    with pytest.raises(pkg.D2GError):  # a parent index past the pool
        ctx.is_bimera_pool(bs.queries[:1], bs.pool, [[len(bs.pool)]])


def test_is_bimera_long_reads(ctx, oracle, pkg):
    """The long-read shape: ~1,450 nt ASCII, band max_shift 16."""
    bs = synth.workload_bimera(n_pool=20, n_queries=12, parents_per_query=5,
                               seed=synth.SEED_BASE + 52)
    got = ctx.is_bimera_batch(bs.queries, bs.parent_lists)
    op = oracle.bimera_params()
    want = np.array([oracle.is_bimera(q, ps, op) for q, ps in zip(bs.queries, bs.parent_lists)])
    assert np.array_equal(got, want)


def test_table_bimera2(ctx, oracle, pkg):
    L, _, R = _chim()
    seqs = [L + R, L + b"C" * 20, b"G" * 20 + R, b"T" * 40]
    mat = np.array([[2, 100, 100, 1]], np.uint32)
    nflag, nsam = ctx.table_bimera2(mat, seqs, 1.5, 2)
    want = oracle.table_bimera2(mat, seqs, 1.5, 2, oracle.bimera_params())
    assert np.array_equal(nflag, want[0]) and np.array_equal(nsam, want[1])
    # a larger synthetic table: pool parents abundant, chimeras rare
    bs = synth.workload_bimera(n_pool=12, n_queries=30, parents_per_query=4, lmin=240, lmax=260,
                               seed=synth.SEED_BASE + 53)
    seqs = list(dict.fromkeys([p for pl in bs.parent_lists for p in pl] + bs.queries))
    rng = np.random.default_rng(9)
    mat = np.zeros((5, len(seqs)), np.uint32)
    for j, s in enumerate(seqs):
        base = 200 if s not in bs.queries else 3
        mat[:, j] = np.where(rng.random(5) < 0.7, rng.integers(base // 2, base * 2, 5), 0)
    for one_off in (False, True):
        p = pkg.bimera_params(allow_one_off=one_off)
        got = ctx.table_bimera2(mat, seqs, 1.5, 2, p)
        want = oracle.table_bimera2(mat, seqs, 1.5, 2, oracle.bimera_params(allow_one_off=one_off),
                                    nthreads=8)
        assert np.array_equal(got[0], want[0]) and np.array_equal(got[1], want[1])
        assert want[0].sum() > 0
