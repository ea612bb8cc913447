"""Writes the golden fixtures of tests/golden/ from the oracle (the C
restatement in oracle/) on the synthetic workloads of dada2-rs_amd/synth.py.

Each fixture holds its inputs and the oracle's outputs; tests/test_golden.py
checks that the oracle still reproduces them and tests/test_gpu_golden.py
checks the device against them.  Run from the repository root:

    python tests/golden/make_golden.py
"""
from __future__ import annotations

import sys
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parents[2]
sys.path.insert(0, str(ROOT))

import __graft_entry__  # noqa: E402
import oracle  # noqa: E402

OUT = Path(__file__).resolve().parent


def concat(seqs) -> tuple[np.ndarray, np.ndarray]:
    arr = [np.frombuffer(bytes(s), np.uint8) if isinstance(s, (bytes, bytearray))
           else np.asarray(s, np.uint8) for s in seqs]
    off = np.zeros(len(arr) + 1, np.uint64)
    off[1:] = np.cumsum([len(a) for a in arr])
    return (np.concatenate(arr) if arr else np.zeros(0, np.uint8)), off


def vectorized(synth) -> None:
    """align_vectorized_with_buf at band 16, (5, -4, -8), end gap 0."""
    w = synth.subsample(synth.workload_batched(), 600, n_centers=8)
    rng = np.random.default_rng(11)
    qi = rng.choice(w.n, size=48, replace=False)
    s1 = [w.seqs[int(w.centers[i % len(w.centers)])] for i in range(48)]
    s2 = [w.seqs[int(q)] for q in qi]
    al0, al1 = zip(*[oracle.align_vectorized(a, b, band=16) for a, b in zip(s1, s2)])
    c1, o1 = concat(s1)
    c2, o2 = concat(s2)
    a0, ao = concat(al0)
    a1, _ = concat(al1)
    score = np.array([oracle.score_alignment(x, y) for x, y in zip(al0, al1)], np.int32)
    np.savez_compressed(OUT / "vectorized_band16.npz", s1=c1, s1_off=o1, s2=c2, s2_off=o2,
                        al0=a0, al1=a1, al_off=ao, score=score)


def homopolymer(synth) -> None:
    """align_endsfree_homo_with_buf on long reads: band 32, homopolymer gap -1."""
    w = synth.workload_longread(n_variants=8, n_uniques=64)
    params = oracle.make_params(homo_gap_p=-1, band=32)
    rng = np.random.default_rng(12)
    idx = rng.choice(w.n, size=6, replace=False)
    s1 = [w.seqs[int(w.centers[0])]] * 6
    s2 = [w.seqs[int(i)] for i in idx]
    al0, al1 = zip(*[oracle.align_endsfree_homo(a, b, params) for a, b in zip(s1, s2)])
    c1, o1 = concat(s1)
    c2, o2 = concat(s2)
    a0, ao = concat(al0)
    a1, _ = concat(al1)
    np.savez_compressed(OUT / "homopolymer_band32.npz", s1=c1, s1_off=o1, s2=c2, s2_off=o2,
                        al0=a0, al1=a1, al_off=ao)


def compare(synth) -> None:
    """The b_compare_parallel map: per (center, unique) lambda bits, hamming
    and skipped flag, k=5, band 16, greedy, FP64 lambda, at kdist cutoff 0.42
    and at 1.0 (the cluster-0 compare: nothing shrouded, every pair aligns)."""
    w = synth.subsample(synth.workload_batched(), 160, n_centers=6)
    rs = oracle.RawSet(w.seqs, w.quals, w.reads, k=w.kmer_size)
    out = {}
    for tag, cut in (("", 0.42), ("_c1", 1.0)):
        params = oracle.make_params(band=w.band, kdist_cutoff=cut)
        rows = [oracle.compare_center(rs, int(c), w.err, w.err.shape[1], params, nthreads=4)
                for c in w.centers]
        out["lam_bits" + tag] = np.stack([r[0] for r in rows]).view(np.uint64)
        out["hamming" + tag] = np.stack([r[1] for r in rows])
        out["skipped" + tag] = np.stack([r[2] for r in rows])
    seq, off = concat(w.seqs)
    qual, _ = concat(w.quals)
    np.savez_compressed(OUT / "compare_160x6.npz", seq=seq, off=off, qual=qual,
                        reads=np.asarray(w.reads, np.uint32), err=w.err,
                        centers=np.asarray(w.centers, np.uint32), **out)


def pairs(synth) -> None:
    """compute_lambda (ordered FP64 product) and get_pA of aligned pairs."""
    ps = synth.workload_precision(n_pairs=48, n_centers=8)
    lam = oracle.pairs_lambda_fp64(ps.seqs, ps.quals, ps.subs, ps.err)
    pval = np.array([oracle.get_pA(int(r), float(l), len(s), int(cr))
                     for r, l, s, cr in zip(ps.reads, lam, ps.subs, ps.center_reads)])
    seq, off = concat(ps.seqs)
    qual, _ = concat(ps.quals)
    subs = np.concatenate([np.asarray(s, np.uint32) for s in ps.subs])
    sub_off = np.zeros(len(ps.subs) + 1, np.uint64)
    sub_off[1:] = np.cumsum([len(s) for s in ps.subs])
    np.savez_compressed(OUT / "pairs_48.npz", seq=seq, off=off, qual=qual, subs=subs,
                        sub_off=sub_off, reads=np.asarray(ps.reads, np.uint32),
                        center_reads=np.asarray(ps.center_reads, np.uint32), err=ps.err,
                        lam_bits=lam.view(np.uint64), pval=pval)


def bimera(synth) -> None:
    """is_bimera_with_buf (ASCII, band = max_shift 16, (5, -4, -8))."""
    bs = synth.workload_bimera(n_pool=24, n_queries=6, parents_per_query=6)
    params = oracle.bimera_params()
    flags = np.array([oracle.is_bimera(q, ps, params) for q, ps in zip(bs.queries, bs.parent_lists)],
                     np.uint8)
    q, qo = concat(bs.queries)
    par = [p for ps in bs.parent_lists for p in ps]
    p, po = concat(par)
    start = np.zeros(len(bs.parent_lists) + 1, np.uint32)
    start[1:] = np.cumsum([len(ps) for ps in bs.parent_lists])
    np.savez_compressed(OUT / "bimera_6x6.npz", q=q, q_off=qo, p=p, p_off=po, par_start=start,
                        flags=flags)


def main() -> None:
    pkg = __graft_entry__.load_package()
    from dada2_rs_amd import synth  # noqa: F401  (registered by load_package)
    oracle.build()
    for f in (vectorized, homopolymer, compare, pairs, bimera):
        f(synth)
        print("wrote", f.__name__, flush=True)


if __name__ == "__main__":
    main()
