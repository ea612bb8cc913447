"""The allocation-free entry point d2g_pairs_lambda_noalloc against the
existing d2g_pairs_load + d2g_pairs_lambda + d2g_pairs_fetch on the same
pairs: identical bits for log-lambda, lambda, p-values and recompute flags,
on two calls in a row (the second with fewer pairs), with and without
qualities; bad inputs become status codes from the device-side checks.
The caller's device memory comes from the HIP runtime the library itself
links (hipMalloc / hipMemcpy through ctypes): the test plays the caller,
which owns its inputs and outputs."""
import ctypes as C

import numpy as np
import pytest

from dada2_rs_amd import synth

pytestmark = pytest.mark.gpu


def _existing(ctx, ps, idx, use_quals, omega):
    seqs = [ps.seqs[i] for i in idx]
    quals = [ps.quals[i] for i in idx]
    subs = [ps.subs[i] for i in idx]
    ctx.pairs_load(seqs, quals, subs, ps.reads[idx], ps.center_reads[idx])
    ll, lam, pv, fl, st = ctx.pairs_lambda(ps.err, use_quals=use_quals, omega_a=omega)
    return ll, lam, pv, fl, st


class _Hip:
    """hipMalloc / hipMemcpy / hipFree of the runtime the library links."""

    def __init__(self):
        self.L = C.CDLL("libamdhip64.so.7")
        self.L.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        self.L.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        self.L.hipFree.argtypes = [C.c_void_p]
        self.live = []

    def put(self, a):
        a = np.ascontiguousarray(a)
        p = C.c_void_p()
        assert self.L.hipMalloc(C.byref(p), max(a.nbytes, 1)) == 0
        self.live.append(p.value)
        assert self.L.hipMemcpy(p.value, a.ctypes.data, a.nbytes, 1) == 0  # host to device
        return p.value

    def get(self, ptr, n, dtype):
        out = np.empty(n, dtype)
        assert self.L.hipMemcpy(out.ctypes.data, ptr, out.nbytes, 2) == 0  # device to host
        return out

    def free(self):
        for p in self.live:
            self.L.hipFree(p)
        self.live = []


@pytest.fixture()
def hip(ctx):
    h = _Hip()
    yield h
    h.free()


def _upload(pkg, ps, idx, hip, with_qual=True):
    lay = pkg.pairs_layout([ps.seqs[i] for i in idx], [ps.quals[i] for i in idx],
                           [ps.subs[i] for i in idx], ps.reads[idx], ps.center_reads[idx])
    dev = {}
    for k in ("seq", "qual", "off", "len", "subs", "sub_off", "reads", "center_reads"):
        if k == "qual" and not with_qual:
            continue
        dev[k] = hip.put(lay[k])
    return lay, dev


def _noalloc(ctx, lay, dev, n, err, hip, use_quals, omega):
    kinds = {"loglam32": (np.float32, 7.0), "lambda": (np.float64, 7.0),
             "pval": (np.float64, 7.0), "recomputed": (np.uint8, 9)}
    out = {k: hip.put(np.full(n, v, t)) for k, (t, v) in kinds.items()}
    st = ctx.pairs_lambda_noalloc(n, int(lay["seq"].size), lay["n_subs"], dev, err, out,
                                  use_quals=use_quals, omega_a=omega)
    return {k: hip.get(out[k], n, kinds[k][0]) for k in kinds}, st


def _same(a, b, what):
    assert a.dtype == b.dtype and a.shape == b.shape, what
    assert np.array_equal(a.view(np.uint8), b.view(np.uint8)), what


@pytest.mark.parametrize("use_quals", [True, False])
def test_noalloc_matches_existing_two_calls(ctx, pkg, hip, use_quals):
    ps = synth.workload_precision(n_pairs=6000, n_centers=50, seed=synth.SEED_BASE + 71)
    omega = 1e-10  # puts some pairs within a factor 10, so the recompute runs
    for n in (6000, 1700):  # the second call is smaller: no stale pair may leak in
        idx = np.arange(n)
        ll, lam, pv, fl, st = _existing(ctx, ps, idx, use_quals, omega)
        lay, dev = _upload(pkg, ps, idx, hip, with_qual=use_quals)
        got, st2 = _noalloc(ctx, lay, dev, n, ps.err, hip, use_quals, omega)
        _same(got["loglam32"], ll, "loglam32")
        _same(got["lambda"], lam, "lambda")
        _same(got["pval"], pv, "pval")
        _same(got["recomputed"], fl, "recomputed")
        assert st2["pairs"] == n and st2["recomputed"] == st["recomputed"]
        if n == 6000 and use_quals:
            assert st["recomputed"] > 0


def test_noalloc_rejects_bad_inputs(ctx, pkg, hip):
    ps = synth.workload_precision(n_pairs=64, n_centers=4, seed=synth.SEED_BASE + 72)
    idx = np.arange(64)
    lay, dev = _upload(pkg, ps, idx, hip)

    def call(d, lay_=lay, ncol=41):
        err = ps.err if ncol == 41 else np.ones((16, ncol))
        with pytest.raises(pkg.D2GError) as ei:
            _noalloc(ctx, lay_, d, 64, err, hip, True, 1e-40)
        return ei.value.status

    bad = dict(dev)
    s = lay["seq"].copy()
    s[int(lay["off"][5]) + 3] = 5  # an N inside pair 5's row
    bad["seq"] = hip.put(s)
    assert call(bad) == -3  # D2G_ERR_NONACGT
    bad = dict(dev)
    o = lay["off"].copy()
    o[7] += 4  # a row off its 16-byte grid
    bad["off"] = hip.put(o)
    assert call(bad) == -1  # D2G_ERR_ARG
    bad = dict(dev)
    r = lay["reads"].copy()
    r[3] = pkg.NOALLOC_MAX_READS + 1
    bad["reads"] = hip.put(r)
    assert call(bad) == -5  # D2G_ERR_UNSUPPORTED
    bad = dict(dev)
    q = lay["qual"].copy()
    q[int(lay["off"][9])] = 41  # a quality outside the 41 columns
    bad["qual"] = hip.put(q)
    assert call(bad) == -1
    # after the rejected calls the next good call is exact again
    ll, lam, pv, fl, _ = _existing(ctx, ps, idx, True, 1e-40)
    got, _ = _noalloc(ctx, lay, dev, 64, ps.err, hip, True, 1e-40)
    _same(got["pval"], pv, "pval after errors")
    _same(got["lambda"], lam, "lambda after errors")
