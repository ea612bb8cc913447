"""The reference's four fixture samples (tests/golden/sam{1,2}{F,R}.fastq.gz,
1,500 MiSeq reads each) through the whole device chain: dereplication,
run_dada, compute_aux, the dada JSON, the sequence table of the forward and
of the reverse samples, and remove_bimera_denovo by each method; every
stage identical to the oracle's."""
from pathlib import Path

import numpy as np
import pytest

from dada2_rs_amd import formats, synth
from dada2_rs_amd import remove_bimera as RB

pytestmark = pytest.mark.gpu
GOLDEN = Path(__file__).resolve().parent / "golden"
SAMPLES = ["sam1F", "sam1R", "sam2F", "sam2R"]


@pytest.fixture(scope="module")
def denoised(pkg, oracle):
    """Per sample: (device result, oracle result, encoded uniques, dada doc)."""
    from oracle import learn
    from oracle import partition as P
    out = {}
    err = synth.fixed_err(41)
    ctx = pkg.Context(0)
    try:
        for name in SAMPLES:
            seqs, quals, reads = formats.raws_from_derep(
                formats.derep_fastq(GOLDEN / f"{name}.fastq.gz"))
            ctx.load_uniques(seqs, quals, reads, k=5)
            got = ctx.run_dada(err)
            aux = ctx.compute_aux(got)
            rs = oracle.RawSet(seqs, quals, reads, k=5)
            want = P.run_dada(rs, err, oracle.make_params())
            centers = [c["center"] for c in want.clusters]
            weight = np.where(want.map >= 0, np.asarray(reads, np.int64), 0)
            want_aux = learn.dada_aux(rs, centers, want.cluster_of.astype(np.int64), weight,
                                      oracle.make_params(), True, 41, 250)
            rp = formats.run_params(pkg.align_params(), pkg.dada_params())
            doc = formats.tagged("dada", formats.dada_output(got, seqs, name, f"{name}.fastq.gz",
                                                             rp, aux))
            assert doc["params"]["kdist_cutoff"] == 0.42 and doc["num_asvs"] == len(got.clusters)
            print(f"{name}: {len(seqs)} uniques, {len(got.clusters)} ASVs", flush=True)
            out[name] = (got, want, aux, want_aux, doc)
    finally:
        ctx.close()
    return out


@pytest.mark.parametrize("name", SAMPLES)
def test_fixture_sample_denoise(denoised, name):
    got, want, aux, want_aux, _ = denoised[name]
    assert [c["center"] for c in got.clusters] == [c["center"] for c in want.clusters]
    for g, o in zip(got.clusters, want.clusters):
        assert all(g[k] == o[k] or (g[k] != g[k] and o[k] != o[k]) for k in o)
    assert np.array_equal(got.p, want.p) and np.array_equal(got.map, want.map)
    assert np.array_equal(aux["transitions"], want_aux["transitions"])
    assert np.array_equal(aux["nsubs"], want_aux["nsubs"])
    gq, wq = aux["cluster_quality"], want_aux["cluster_quality"]
    assert np.array_equal(np.isnan(gq), np.isnan(wq)) and np.array_equal(gq[~np.isnan(gq)],
                                                                         wq[~np.isnan(wq)])


@pytest.mark.parametrize("group", [["sam1F", "sam2F"], ["sam1R", "sam2R"]], ids=["F", "R"])
@pytest.mark.parametrize("method", RB.METHODS)
def test_fixture_remove_bimera(pkg, denoised, group, method):
    from oracle import remove_bimera as ORB
    table = formats.make_sequence_table([denoised[n][4] for n in group])
    ctx = pkg.Context(0)
    try:
        got = RB.remove_bimera_denovo(ctx, table, method)
        flags = RB.bimera_flags(ctx, table, method)
    finally:
        ctx.close()
    want = ORB.remove_bimera_denovo(table, method)
    print(f"{'+'.join(group)} {method}: {len(table['sequences'])} sequences, "
          f"{len(table['sequences']) - len(got['sequences'])} removed", flush=True)
    assert np.array_equal(flags, ORB.bimera_flags(table, method))
    assert got == want
    assert len(table["sequences"]) > 2


@pytest.mark.parametrize("method", RB.METHODS)
def test_constructed_chimera_removed(pkg, method):
    from test_seqtab_bimera import chimera_table
    from oracle import remove_bimera as ORB
    table, (a, b, c, d) = chimera_table()
    ctx = pkg.Context(0)
    try:
        got = RB.remove_bimera_denovo(ctx, table, method)
    finally:
        ctx.close()
    assert got == ORB.remove_bimera_denovo(table, method)
    assert c not in got["sequences"] and len(got["sequences"]) == 3
