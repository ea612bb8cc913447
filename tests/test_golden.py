"""The oracle against the committed golden fixtures (tests/golden/, written
by tests/golden/make_golden.py): a change to the oracle that alters any
alignment, lambda bit, flag or p-value fails here."""
from pathlib import Path

import numpy as np
import pytest

G = Path(__file__).resolve().parent / "golden"


def split(cat, off):
    return [bytes(cat[int(off[i]):int(off[i + 1])]) for i in range(len(off) - 1)]


def test_vectorized_band16(oracle):
    d = np.load(G / "vectorized_band16.npz")
    s1, s2 = split(d["s1"], d["s1_off"]), split(d["s2"], d["s2_off"])
    a0, a1 = split(d["al0"], d["al_off"]), split(d["al1"], d["al_off"])
    for i, (x, y) in enumerate(zip(s1, s2)):
        got = oracle.align_vectorized(x, y, band=16)
        assert (bytes(got[0]), bytes(got[1])) == (a0[i], a1[i]), i
        assert oracle.score_alignment(a0[i], a1[i]) == int(d["score"][i])


def test_homopolymer_band32(oracle):
    d = np.load(G / "homopolymer_band32.npz")
    params = oracle.make_params(homo_gap_p=-1, band=32)
    s1, s2 = split(d["s1"], d["s1_off"]), split(d["s2"], d["s2_off"])
    a0, a1 = split(d["al0"], d["al_off"]), split(d["al1"], d["al_off"])
    for i, (x, y) in enumerate(zip(s1, s2)):
        got = oracle.align_endsfree_homo(x, y, params)
        assert (bytes(got[0]), bytes(got[1])) == (a0[i], a1[i]), i


@pytest.mark.parametrize("tag,cut", [("", 0.42), ("_c1", 1.0)])
def test_compare_160x6(oracle, tag, cut):
    d = np.load(G / "compare_160x6.npz")
    seqs = [np.frombuffer(s, np.uint8) for s in split(d["seq"], d["off"])]
    quals = [np.frombuffer(s, np.uint8) for s in split(d["qual"], d["off"])]
    rs = oracle.RawSet(seqs, quals, d["reads"], k=5)
    params = oracle.make_params(band=16, kdist_cutoff=cut)
    for r, c in enumerate(d["centers"]):
        lam, ham, skp = oracle.compare_center(rs, int(c), d["err"], d["err"].shape[1], params,
                                              nthreads=4)[:3]
        assert np.array_equal(lam.view(np.uint64), d["lam_bits" + tag][r])
        assert np.array_equal(ham, d["hamming" + tag][r])
        assert np.array_equal(skp, d["skipped" + tag][r])


def test_pairs_48(oracle):
    d = np.load(G / "pairs_48.npz")
    seqs = [np.frombuffer(s, np.uint8) for s in split(d["seq"], d["off"])]
    quals = [np.frombuffer(s, np.uint8) for s in split(d["qual"], d["off"])]
    subs = [d["subs"][int(d["sub_off"][i]):int(d["sub_off"][i + 1])] for i in range(len(seqs))]
    lam = oracle.pairs_lambda_fp64(seqs, quals, subs, d["err"])
    assert np.array_equal(lam.view(np.uint64), d["lam_bits"])
    pval = np.array([oracle.get_pA(int(r), float(l), len(s), int(cr))
                     for r, l, s, cr in zip(d["reads"], lam, subs, d["center_reads"])])
    assert np.array_equal(pval, d["pval"])


def test_bimera_6x6(oracle):
    d = np.load(G / "bimera_6x6.npz")
    qs = split(d["q"], d["q_off"])
    ps = split(d["p"], d["p_off"])
    st = d["par_start"]
    params = oracle.bimera_params()
    got = [oracle.is_bimera(q, ps[int(st[i]):int(st[i + 1])], params) for i, q in enumerate(qs)]
    assert np.array_equal(np.array(got, np.uint8), d["flags"])
