"""The device Poisson log-likelihood of the compare and its hand-written
gradient kernel: against an FP64 CPU reference built from the oracle's own
alignments (sub_new, nwalign.rs:1429-1455, and compute_lambda's transitions,
pval.rs:294-309), against central finite differences (h = 1e-3), and through
torch.autograd."""
import math

import numpy as np
import pytest

from dada2_rs_amd import synth

pytestmark = pytest.mark.gpu

H = 1e-3


def _workload(name):
    if name == "batched":
        return synth.subsample(synth.workload_batched(), 2000, n_centers=16)
    # the gradient workload: 500 variants at a fixed 250 nt (50,000 uniques
    # generated, the 3,000 most abundant and the 40 centers among them used)
    w = synth.workload_batched(n_variants=500, L=250, n_uniques=50000, lmin=250, lmax=250,
                               seed=synth.SEED_BASE + 3)
    return synth.subsample(w, 3000, n_centers=40)


@pytest.fixture(scope="module", params=["batched", "gradient"])
def setup(ctx, oracle, pkg, request):
    w = _workload(request.param)
    print(f"loglik workload {request.param}: {w.n} uniques x {len(w.centers)} centers", flush=True)
    ctx.load_uniques(w.seqs, w.quals, w.reads, k=5)
    got = ctx.compare_batch(w.centers, w.err)
    rng = np.random.default_rng(4)
    logE = (np.log(w.err) + rng.normal(0, 0.05, w.err.shape)).astype(np.float32)
    # FP64 reference: per aligned pair (not skipped, not shrouded) its counts
    rs = oracle.RawSet(w.seqs, w.quals, w.reads, k=5)
    params = oracle.make_params(band=w.band)
    pairs = []
    for r, c in enumerate(int(x) for x in w.centers):
        for u in np.nonzero(got.hamming[r] != 0xFFFFFFFF)[0]:
            sub, buf = oracle.sub_new(rs, c, int(u), params)
            q = np.asarray(w.seqs[u], np.int64)
            t = 5 * (q - 1)
            for s in range(sub.nsubs):
                p0 = sub.pos[s]
                p1 = sub.map[p0]
                t[p1] = 4 * (sub.nt0[s] - 1) + (sub.nt1[s] - 1)
            cnt = np.zeros(w.err.shape, np.float64)
            np.add.at(cnt, (t, np.asarray(w.quals[u], np.int64)), 1.0)
            pairs.append((cnt, float(w.reads[u]), float(w.reads[c])))
    return w, logE, pairs


def _ref(logE, pairs):
    le = logE.astype(np.float64)
    loss, grad = 0.0, np.zeros_like(le)
    for cnt, a, nc in pairs:
        E = nc * math.exp(float((cnt * le).sum()))
        loss += a * math.log(E) - E - math.lgamma(a + 1.0)
        grad += (a - E) * cnt
    return loss, grad


def test_loss_and_gradient_match_fp64_reference(ctx, setup):
    w, logE, pairs = setup
    assert len(pairs) > 100
    loss = ctx.loglik_forward(logE)
    g32, g64 = ctx.loglik_backward()
    rloss, rgrad = _ref(logE, pairs)
    assert abs(loss - rloss) <= 1e-10 * abs(rloss), (loss, rloss)
    scale = np.abs(rgrad).max()
    assert np.abs(g64 - rgrad).max() <= 1e-10 * scale
    assert np.array_equal(g32, g64.astype(np.float32))


def test_gradient_matches_central_differences(ctx, setup):
    """Central differences at h = 1e-3 on the entries with the largest
    gradient.  Tolerance: 1e-4 relative to the entry's gradient."""
    w, logE, pairs = setup
    ctx.loglik_forward(logE)
    _, g = ctx.loglik_backward()
    idx = np.argsort(-np.abs(g).ravel())[:24]
    worst = 0.0
    for k in idx:
        t, q = divmod(int(k), logE.shape[1])
        lp, lm = logE.copy(), logE.copy()
        lp[t, q] += np.float32(H)
        lm[t, q] -= np.float32(H)
        h_eff = (float(lp[t, q]) - float(lm[t, q])) / 2.0  # the FP32-representable step
        fd = (ctx.loglik_forward(lp) - ctx.loglik_forward(lm)) / (2.0 * h_eff)
        worst = max(worst, abs(fd - g[t, q]) / abs(g[t, q]))
    print("max relative FD error", worst)
    assert worst < 1e-4, worst


def test_autograd_function(ctx, setup):
    torch = pytest.importorskip("torch")
    from dada2_rs_amd.autograd import poisson_loglik
    w, logE, pairs = setup
    dev = "cuda" if torch.cuda.is_available() else "cpu"
    x = torch.tensor(logE, device=dev, requires_grad=True)
    loss = poisson_loglik(x, ctx)
    loss.backward()
    ctx.loglik_forward(logE)
    _, g = ctx.loglik_backward()
    assert np.allclose(x.grad.cpu().numpy(), g.astype(np.float32), rtol=0, atol=0)
    assert abs(loss.item() - ctx.loglik_forward(logE)) == 0.0


def test_forward_and_backward_bits_repeat(ctx, setup):
    """The gradient is a fixed-order reduction: integer counts per pair, one
    FP64 multiply-add per bin, wavefronts and workgroups summed in order, so
    repeated calls return the same bits."""
    w, logE, pairs = setup
    loss = [ctx.loglik_forward(logE) for _ in range(3)]
    assert loss[0] == loss[1] == loss[2]
    grads = []
    for _ in range(3):
        ctx.loglik_forward(logE)
        grads.append(ctx.loglik_backward()[1])
    assert all(np.array_equal(grads[0].view(np.uint64), g.view(np.uint64)) for g in grads[1:])
