"""The FP64 pair-lambda oracle (a vectorised restatement of compute_lambda's
ordered product, pval.rs:311-313) pinned against the C oracle's full chain
sub_new -> al2subs -> compute_lambda (nwalign.rs:1429-1455, pval.rs:253-320),
and the precision workload's substitution lists against its sequences.
No GPU."""
import numpy as np

from dada2_rs_amd import synth


def test_pair_lambda_matches_sub_new_chain(oracle):
    ps = synth.workload_precision(n_pairs=300, seed=synth.SEED_BASE + 20)
    centers = []
    for s, rec in zip(ps.seqs, ps.subs):
        c = np.array(s, np.uint8)
        rec = np.asarray(rec, np.int64)
        c[rec >> 16] = (rec & 0xFFFF).astype(np.uint8)
        centers.append(c)
    vec = oracle.pairs_lambda_fp64(ps.seqs, ps.quals, ps.subs, ps.err)
    # band 0: positional (gapless) pairing, as the pairs are substitution-only
    params = oracle.make_params(band=0, use_kmers=False)
    for i in range(ps.n):
        rs = oracle.RawSet([centers[i], ps.seqs[i]], [ps.quals[i], ps.quals[i]], [10, 1],
                           with_kmers=False)
        sub, buf = oracle.sub_new(rs, 0, 1, params)
        assert sub.nsubs == len(ps.subs[i])
        lam = oracle.compute_lambda(rs, 1, sub, ps.err, ps.err.shape[1])
        assert np.float64(lam).view(np.uint64) == vec[i].view(np.uint64), i


def test_identical_pairs_are_diagonal_products(oracle):
    rng = np.random.default_rng(3)
    err = synth.seeded_err(rng)
    seqs = [rng.integers(1, 5, 50).astype(np.uint8) for _ in range(5)]
    quals = [rng.integers(0, 41, 50).astype(np.uint8) for _ in range(5)]
    lam = oracle.pairs_lambda_fp64(seqs, quals, [np.zeros(0, np.uint32)] * 5, err)
    for s, q, l in zip(seqs, quals, lam):
        want = 1.0
        for nt, qq in zip(s, q):
            want *= err[5 * (int(nt) - 1), int(qq)]
        assert l == want


def test_precision_workload_subs_are_consistent():
    ps = synth.workload_precision(n_pairs=2000, seed=synth.SEED_BASE + 21)
    assert ps.err.shape == (16, 41)
    for r in range(4):  # each (ref, q) row normalised over the query bases
        assert np.allclose(ps.err[4 * r:4 * r + 4].sum(axis=0), 1.0)
    nsub = 0
    for s, rec in zip(ps.seqs, ps.subs):
        rec = np.asarray(rec, np.int64)
        pos, nt0 = rec >> 16, rec & 0xFFFF
        assert np.all(np.diff(pos) > 0)
        assert np.all(s[pos] != nt0) and np.all((nt0 >= 1) & (nt0 <= 4))
        nsub += len(rec)
    assert 2.0 < nsub / ps.n < 30.0
