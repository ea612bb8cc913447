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
