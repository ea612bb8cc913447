"""The reference's own fixture reads (tests/golden/sam1F.fastq.gz, 1,500
MiSeq reads of 250 nt) dereplicated and run through the device path: the
plain compare bit-exact against the oracle, and run_dada identical to the
oracle's partition loop."""
from pathlib import Path

import numpy as np
import pytest

from dada2_rs_amd import formats, synth

pytestmark = pytest.mark.gpu
GOLDEN = Path(__file__).resolve().parent / "golden"


@pytest.fixture(scope="module")
def raws():
    return formats.raws_from_derep(formats.derep_fastq(GOLDEN / "sam1F.fastq.gz"))


def test_fixture_compare_center(pkg, oracle, raws):
    seqs, quals, reads = raws
    err = synth.fixed_err(41)
    rs = oracle.RawSet(seqs, quals, reads, k=5)
    ctx = pkg.Context(0)
    try:
        ctx.load_uniques(seqs, quals, reads, k=5)
        for c in (int(np.argmax(reads)), 1, 17):
            got = ctx.compare_center(c, err)
            lam, ham, skp, _, _ = oracle.compare_center(rs, c, err, 41, oracle.make_params())
            assert np.array_equal(got.lam, lam) and np.array_equal(got.hamming, ham)
            assert np.array_equal(got.skipped, skp)
    finally:
        ctx.close()


def test_fixture_run_dada(pkg, oracle, raws):
    from oracle import partition as P
    seqs, quals, reads = raws
    err = synth.fixed_err(41)
    rs = oracle.RawSet(seqs, quals, reads, k=5)
    want = P.run_dada(rs, err, oracle.make_params())
    ctx = pkg.Context(0)
    try:
        ctx.load_uniques(seqs, quals, reads, k=5)
        got = ctx.run_dada(err)
    finally:
        ctx.close()
    print(f"fixture: {len(seqs)} uniques, {len(got.clusters)} clusters", flush=True)
    assert [c["center"] for c in got.clusters] == [c["center"] for c in want.clusters]
    for g, o in zip(got.clusters, want.clusters):
        assert all(g[k] == o[k] or (g[k] != g[k] and o[k] != o[k]) for k in o)
    assert np.array_equal(got.cluster_of, want.cluster_of)
    assert np.array_equal(got.p, want.p)
    assert np.array_equal(got.map, want.map)
    assert got.stats["nalign"] == want.nalign and got.stats["nshroud"] == want.nshroud
    assert len(got.clusters) > 1
