"""Mixed-precision lambda on the GPU: FP32 Kahan log-lambda within 1e-6
relative error of the FP64 oracle's log of the exact ordered product
(compute_lambda, pval.rs:253-320); FP64 p-values (get_pA/calc_pA,
pval.rs:208-241, 376-398) with FP64 lambda recompute within a factor 10 of
omega_a, and no p < omega_a call that differs from the oracle's on exact
lambda.  The oracle's incomplete gamma is not statrs (parity unpinned at the
bit level); the call identity is what is asserted."""
import math

import numpy as np
import pytest

from dada2_rs_amd import synth

pytestmark = pytest.mark.gpu

REL_TOL_LOGLAM = 1e-6


@pytest.fixture(scope="module")
def pairs():
    return synth.workload_precision(n_pairs=50000, seed=synth.SEED_BASE + 22)


def test_loglam32_relative_error(ctx, oracle, pairs):
    ps = pairs
    ctx.pairs_load(ps.seqs, ps.quals, ps.subs, ps.reads, ps.center_reads)
    ll, lam, pv, fl, st = ctx.pairs_lambda(ps.err, omega_a=1e-40)
    exact = oracle.pairs_lambda_fp64(ps.seqs, ps.quals, ps.subs, ps.err)
    ref = np.log(exact)
    rel = np.abs(ll.astype(np.float64) - ref) / np.abs(ref)
    assert np.isfinite(rel).all()
    assert rel.max() < REL_TOL_LOGLAM, (rel.max(), rel.mean())
    # recomputed pairs carry the exact FP64 product, bit for bit
    idx = np.nonzero(fl)[0]
    assert np.array_equal(lam[idx].view(np.uint64), exact[idx].view(np.uint64))
    assert st["pairs"] == ps.n and st["recomputed"] == len(idx)


@pytest.mark.parametrize("omega", [1e-40, 1e-10, 1e-3])
def test_no_call_flips(ctx, oracle, pairs, omega):
    ps = pairs
    ctx.pairs_load(ps.seqs, ps.quals, ps.subs, ps.reads, ps.center_reads)
    ll, lam, pv, fl, st = ctx.pairs_lambda(ps.err, omega_a=omega)
    exact = oracle.pairs_lambda_fp64(ps.seqs, ps.quals, ps.subs, ps.err)
    ham = np.array([len(s) for s in ps.subs])
    want = np.array([oracle.get_pA(int(r), float(l), int(h), int(c))
                     for r, l, h, c in zip(ps.reads, exact, ham, ps.center_reads)])
    flips = np.count_nonzero((pv < omega) != (want < omega))
    assert flips == 0, flips
    near = (want >= omega / 10) & (want <= omega * 10)
    assert np.all(fl[near] == 1)
    # where both are normal numbers the p-values agree closely
    ok = (want > 1e-300) & (pv > 1e-300)
    rel = np.abs(pv[ok] - want[ok]) / want[ok]
    assert np.median(rel) < 1e-4, np.median(rel)


@pytest.mark.parametrize("L", [600, 1450])
def test_loglam32_long_pairs_dense_substitutions(ctx, oracle, L):
    """Pairs longer than one 256-position segment (the kernel's substitution
    patches then span segments) and with up to about one substitution in
    four positions (records past the first 64 of a pair take the kernel's
    loop): FP32 log-lambda within 1e-6 relative of the FP64 oracle's
    ordered product, and the listed pairs' FP64 lambda bit for bit."""
    rng = np.random.default_rng(synth.SEED_BASE + 40 + L)
    n = 600
    err = synth.seeded_err(rng, 41)
    centers = rng.integers(1, 5, size=(n, L)).astype(np.uint8)
    q = centers.copy()
    rate = np.linspace(0.0, 0.25, n)[:, None]
    flip = rng.random((n, L)) < rate
    shift = rng.integers(1, 4, size=int(flip.sum())).astype(np.uint8)
    q[flip] = ((q[flip] - 1 + shift) % 4 + 1).astype(np.uint8)
    quals = rng.integers(0, 41, size=(n, L)).astype(np.uint8)
    rows, cols = np.nonzero(q != centers)
    recs = (cols.astype(np.uint32) << 16) | centers[rows, cols].astype(np.uint32)
    subs = np.split(recs, np.cumsum(np.bincount(rows, minlength=n))[:-1])
    assert max(len(s) for s in subs) > 64 and min(len(s) for s in subs) == 0
    reads = synth.abundance(rng, n)
    center_reads = np.maximum(1, np.rint(rng.lognormal(6.0, 1.5, size=n))).astype(np.uint32)
    ctx.pairs_load(list(q), list(quals), subs, reads, center_reads)
    ll, lam, pv, fl, st = ctx.pairs_lambda(err, omega_a=1e-10)
    exact = oracle.pairs_lambda_fp64(list(q), list(quals), subs, err)
    # the dense pairs' product underflows FP64 (lambda 0, as compute_lambda
    # returns it), so log lambda is checked against the FP64 sum of the log
    # table entries, the quantity the FP32 sum estimates; where the product
    # is normal its log agrees with that sum
    t = 5 * (q.astype(np.int64) - 1)
    for i, rec in enumerate(subs):
        pos, nt0 = rec.astype(np.int64) >> 16, rec.astype(np.int64) & 0xFFFF
        t[i, pos] = 4 * (nt0 - 1) + (q[i, pos].astype(np.int64) - 1)
    ref = np.log(err)[t, quals].sum(axis=1)
    normal = exact > 1e-300
    assert normal.sum() > n // 4 and (~normal).sum() > 0
    assert np.allclose(np.log(exact[normal]), ref[normal], rtol=1e-12, atol=0)
    rel = np.abs(ll.astype(np.float64) - ref) / np.abs(ref)
    assert rel.max() < REL_TOL_LOGLAM, (rel.max(), rel.mean())
    idx = np.nonzero(fl)[0]
    assert np.array_equal(lam[idx].view(np.uint64), exact[idx].view(np.uint64))
