"""The store of b_compare_parallel on the device (d2g_compare_store) against
the oracle's serial store (oracle/partition.py `store`, cluster.rs:351-397)
over the oracle's compare of each center in turn: the same kept
comparisons (cluster, raw, lambda bits, hamming) in the same order, the
same self lambdas, with e_minmax carried from call to call and across
centers of one call."""
import numpy as np
import pytest

from dada2_rs_amd import synth

pytestmark = pytest.mark.gpu

HAMMING_NONE = 0xFFFFFFFF


@pytest.fixture(scope="module")
def work():
    return synth.subsample(synth.workload_batched(), 3000, n_centers=24)


def _oracle_store(oracle, w, center_lists, greedy=True, lock=None):
    from oracle import partition as P
    rs = oracle.RawSet(w.seqs, w.quals, w.reads, k=5)
    if lock is not None:
        for r in range(len(rs)):
            rs.arr[r].lock = int(lock[r])
    e = [-999.0] * w.n
    total = float(int(np.asarray(w.reads, np.uint64).sum()) & 0xFFFFFFFF)
    kept, selfs, i = [], [], 0
    for centers in center_lists:
        for c in centers:
            lam, ham, skp, _, _ = oracle.compare_center(rs, int(c), w.err, 41, oracle.make_params(),
                                                        greedy=greedy)
            selfs.append(float(lam[int(c)]))
            kept += P.store(e, lam, ham, total, int(w.reads[int(c)]), i)
            i += 1
    return kept, np.array(selfs)


def _as_tuples(arr):
    return [(int(x["cluster"]), int(x["index"]), float(x["lambda"]), int(x["hamming"]))
            for x in arr]


def _check(got, want):
    assert len(got) == len(want)
    for g, o in zip(got, want):
        assert g[0] == o.i and g[1] == o.index and g[3] == o.hamming
        assert np.float64(g[2]).view(np.uint64) == np.float64(o.lam).view(np.uint64)


def test_store_batched_centers_match_serial_store(ctx, oracle, work):
    w = work
    c1, c2 = w.centers[:16], w.centers[16:24]
    want, want_self = _oracle_store(oracle, w, [c1, c2])
    ctx.load_uniques(w.seqs, w.quals, w.reads, k=5)
    ctx.store_reset()
    a, s1, st1 = ctx.compare_store(c1, w.err, first_cluster=0)
    b, s2, st2 = ctx.compare_store(c2, w.err, first_cluster=len(c1))
    print(f"kept {len(a)} + {len(b)} of {len(w.centers) * w.n} comparisons", flush=True)
    assert len(a) >= w.n  # the first center keeps every raw (e_minmax = -999)
    _check(_as_tuples(a) + _as_tuples(b), want)
    assert np.array_equal(np.concatenate([s1, s2]).view(np.uint64), want_self.view(np.uint64))
    # the resident store holds both calls' records in order
    allrec = ctx.store_fetch()
    _check(_as_tuples(allrec), want)
    assert st1["comparisons"] == len(c1) * w.n


def test_store_one_center_per_call_equals_batch(ctx, oracle, work):
    w = work
    cs = w.centers[:12]
    ctx.load_uniques(w.seqs, w.quals, w.reads, k=5)
    ctx.store_reset()
    got = []
    for i, c in enumerate(cs):
        rec, _, _ = ctx.compare_store([c], w.err, first_cluster=i)
        got += _as_tuples(rec)
    ctx.store_reset()
    batch, _, _ = ctx.compare_store(cs, w.err)
    assert got == _as_tuples(batch)
    want, _ = _oracle_store(oracle, w, [cs])
    _check(got, want)


def test_store_without_greedy_and_with_locks(ctx, oracle, work):
    w = work
    rng = np.random.default_rng(5)
    lock = (rng.random(w.n) < 0.3).astype(np.uint8)
    cs = w.centers[:6]
    want, _ = _oracle_store(oracle, w, [cs], greedy=True, lock=lock)
    ctx.load_uniques(w.seqs, w.quals, w.reads, k=5)
    ctx.store_reset()
    rec, _, _ = ctx.compare_store(cs, w.err, lock=lock)
    _check(_as_tuples(rec), want)
    want, _ = _oracle_store(oracle, w, [cs], greedy=False)
    ctx.store_reset()
    rec, _, _ = ctx.compare_store(cs, w.err, greedy=False)
    _check(_as_tuples(rec), want)


def test_store_count_only_and_state(ctx, pkg, work):
    w = work
    ctx.load_uniques(w.seqs, w.quals, w.reads, k=5)
    with pytest.raises(pkg.D2GError):  # no store before d2g_store_reset
        ctx.compare_store(w.centers[:2], w.err)
    ctx.store_reset()
    k, _, _ = ctx.compare_store(w.centers[:4], w.err, fetch=False)
    assert k == len(ctx.store_fetch())


def test_store_rejects_calls_past_its_index_range(ctx, pkg, work):
    """The store's index-range guard (d2g_compare_store checks, in 64-bit
    arithmetic, that a call has at most 65,535 centers and that the store
    stays under 2^32 comparisons): a call with 65,536 centers is refused
    with D2G_ERR_UNSUPPORTED before any launch, and the store is left as it
    was."""
    w = work
    ctx.load_uniques(w.seqs, w.quals, w.reads, k=5)
    ctx.store_reset()
    kept, _, _ = ctx.compare_store(w.centers[:2], w.err)
    before = ctx.store_fetch()
    too_many = np.zeros(65536, np.uint32)
    with pytest.raises(pkg.D2GError) as exc:
        ctx.compare_store(too_many, w.err, first_cluster=2, fetch=False)
    assert exc.value.status == -5  # D2G_ERR_UNSUPPORTED
    after = ctx.store_fetch()
    assert np.array_equal(before, after) and len(after) == len(kept)
