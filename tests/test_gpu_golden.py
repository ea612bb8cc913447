"""The device against the committed golden fixtures (tests/golden/): the
oracle's alignments, lambda bits, flags and p-values on each workload, read
from files (no oracle call), so the check runs wherever the fixtures go."""
from pathlib import Path

import numpy as np
import pytest

pytestmark = pytest.mark.gpu
G = Path(__file__).resolve().parent / "golden"


def split(cat, off):
    return [bytes(cat[int(off[i]):int(off[i + 1])]) for i in range(len(off) - 1)]


def test_vectorized_band16(ctx):
    d = np.load(G / "vectorized_band16.npz")
    s1, s2 = split(d["s1"], d["s1_off"]), split(d["s2"], d["s2_off"])
    got = ctx.align_vectorized(s1, s2, band=16)
    want = list(zip(split(d["al0"], d["al_off"]), split(d["al1"], d["al_off"])))
    assert [(bytes(a), bytes(b)) for a, b in got] == want


def test_homopolymer_band32(ctx):
    d = np.load(G / "homopolymer_band32.npz")
    s1, s2 = split(d["s1"], d["s1_off"]), split(d["s2"], d["s2_off"])
    got = ctx.align_endsfree(s1, s2, homo_gap_p=-1, band=32)
    want = list(zip(split(d["al0"], d["al_off"]), split(d["al1"], d["al_off"])))
    assert [(bytes(a), bytes(b)) for a, b in got] == want


@pytest.mark.parametrize("tag,cut", [("", 0.42), ("_c1", 1.0)])
def test_compare_160x6(ctx, pkg, tag, cut):
    d = np.load(G / "compare_160x6.npz")
    seqs = [np.frombuffer(s, np.uint8) for s in split(d["seq"], d["off"])]
    quals = [np.frombuffer(s, np.uint8) for s in split(d["qual"], d["off"])]
    ctx.load_uniques(seqs, quals, d["reads"], k=5)
    got = ctx.compare_batch(d["centers"], d["err"], params=pkg.align_params(kdist_cutoff=cut))
    assert np.array_equal(got.lam.view(np.uint64), d["lam_bits" + tag])
    assert np.array_equal(got.hamming, d["hamming" + tag])
    assert np.array_equal(got.skipped, d["skipped" + tag])


def test_pairs_48(ctx):
    """FP32 log-lambda within 1e-6 (relative) of the log of the FP64
    product; p-values decide the same ASV calls at omega_a = 1e-40 (p-value
    bits of the two incomplete gamma evaluations are not pinned to each
    other)."""
    d = np.load(G / "pairs_48.npz")
    seqs = [np.frombuffer(s, np.uint8) for s in split(d["seq"], d["off"])]
    quals = [np.frombuffer(s, np.uint8) for s in split(d["qual"], d["off"])]
    subs = [d["subs"][int(d["sub_off"][i]):int(d["sub_off"][i + 1])] for i in range(len(seqs))]
    ctx.pairs_load(seqs, quals, subs, d["reads"], d["center_reads"])
    ll, lam, pval, rec, _ = ctx.pairs_lambda(d["err"], omega_a=1e-40)
    want = np.log(d["lam_bits"].view(np.float64))
    assert np.all(np.abs(ll.astype(np.float64) - want) <= 1e-6 * np.abs(want))
    assert np.array_equal(pval < 1e-40, d["pval"] < 1e-40)
    # where lambda was recomputed in FP64 it is the oracle's product
    assert np.array_equal(lam.view(np.uint64)[rec == 1], d["lam_bits"][rec == 1])


def test_bimera_6x6(ctx, pkg):
    d = np.load(G / "bimera_6x6.npz")
    qs = split(d["q"], d["q_off"])
    ps = split(d["p"], d["p_off"])
    st = d["par_start"]
    lists = [ps[int(st[i]):int(st[i + 1])] for i in range(len(qs))]
    got = ctx.is_bimera_batch(qs, lists)
    assert np.array_equal(np.asarray(got, np.uint8), d["flags"])
