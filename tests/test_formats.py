"""FASTQ dereplication, the Raw build and the error-model JSON
(dada2-rs_amd/formats.py) against hand-worked cases and the reference's own
fixture reads (tests/golden/sam1F.fastq.gz, copied from the reference's
tests/fixtures).  CPU only."""
import gzip
import json
from pathlib import Path

import numpy as np
import pytest

GOLDEN = Path(__file__).resolve().parent / "golden"


@pytest.fixture(scope="module")
def F():
    from dada2_rs_amd import formats
    return formats


def test_dereplicate_handmade(F):
    recs = [(b"ACGT", b"IIII"), (b"ACGA", b"5555"), (b"ACGT", b"!!!!"), (b"AAAA", b"IIII"),
            (b"ACGA", b"IIII"), (b"CCCC", b"++++")]
    d = F.dereplicate(recs)
    # count descending, ties by sequence bytes ascending (derep.rs:158-163)
    assert d.uniques == [(b"ACGA", 2), (b"ACGT", 2), (b"AAAA", 1), (b"CCCC", 1)]
    # Phred+33 sums: 'I' 40, '5' 20, '!' 0, '+' 10
    assert [list(q) for q in d.quals] == [[60] * 4, [40] * 4, [40] * 4, [10] * 4]
    assert list(d.map) == [1, 0, 1, 2, 0, 3]
    seqs, quals, reads = F.raws_from_derep(d, k=3)
    assert [list(s) for s in seqs] == [[1, 2, 3, 1], [1, 2, 3, 4], [1, 1, 1, 1], [2, 2, 2, 2]]
    assert [list(q) for q in quals] == [[30] * 4, [20] * 4, [40] * 4, [10] * 4]
    assert list(reads) == [2, 2, 1, 1]


def test_quality_rounds_half_away_from_zero(F):
    # Raw::from_qual_sums: (sum / reads).round(); half-to-even would give 2 0 2 4
    d = F.Derep([(b"ACGTACGT", 2)], [np.array([5, 1, 3, 7, 4, 0, 2, 6], np.uint32)],
                np.zeros(2, np.int64))
    _, quals, _ = F.raws_from_derep(d, k=3)
    assert list(quals[0]) == [3, 1, 2, 4, 2, 0, 1, 3]
    d = F.Derep([(b"ACGTACGT", 3)], [np.array([4, 5, 7, 8, 0, 1, 2, 3], np.uint32)],
                np.zeros(3, np.int64))
    assert list(F.raws_from_derep(d, k=3)[1][0]) == [1, 2, 2, 3, 0, 0, 1, 1]


def test_validation(F):
    with pytest.raises(ValueError):  # non-ACGT base
        F.raws_from_derep(F.dereplicate([(b"ACGNACGT", b"IIIIIIII")]), k=3)
    with pytest.raises(ValueError):  # not longer than k
        F.raws_from_derep(F.dereplicate([(b"ACGT", b"IIII")]), k=5)
    with pytest.raises(ValueError):  # zero inputs
        F.raws_from_derep(F.dereplicate([]), k=5)
    # U and lower case encode as T and upper case
    s, _, _ = F.raws_from_derep(F.dereplicate([(b"acguACGU", b"IIIIIIII")]), k=3)
    assert list(s[0]) == [1, 2, 3, 4, 1, 2, 3, 4]


def test_fastq_reader(F, tmp_path):
    p = tmp_path / "r.fastq"
    p.write_bytes(b"@a\nACGT\n+\nIIII\n@b\nAC\n+\nII\n\n")
    assert F.read_fastq(p) == [(b"ACGT", b"IIII"), (b"AC", b"II")]
    p.write_bytes(b"@a\nACGT\n+\nIII\n")
    with pytest.raises(ValueError):
        F.read_fastq(p)
    p.write_bytes(b"@a\nACGT\n+\n")
    with pytest.raises(ValueError):
        F.read_fastq(p)


def test_fixture_derep_invariants(F):
    recs = F.read_fastq(GOLDEN / "sam1F.fastq.gz")
    d = F.dereplicate(recs)
    assert len(recs) == 1500 and len(d.uniques) == len({s for s, _ in recs}) == 896
    counts = [c for _, c in d.uniques]
    assert sum(counts) == 1500
    keys = [(-c, s) for s, c in d.uniques]
    assert keys == sorted(keys)
    assert np.array_equal(np.bincount(d.map, minlength=896), counts)
    # every read lands on its sequence, and the sums add up
    sums = [np.zeros(len(s), np.int64) for s, _ in d.uniques]
    for r, (s, q) in enumerate(recs):
        assert d.uniques[d.map[r]][0] == s
        sums[d.map[r]] += np.frombuffer(q, np.uint8).astype(np.int64) - 33
    assert all(np.array_equal(a, b) for a, b in zip(sums, d.quals))


def test_error_model_json(F, tmp_path):
    err = np.arange(16 * 41, dtype=np.float64).reshape(16, 41) / 1000.0
    doc = {"dada2_rs_command": "learn-errors", "dada2_rs_version": "0.2.0", "nq": 41,
           "err_in": (err * 2).tolist(), "err_out": err.tolist()}
    p = tmp_path / "err.json.gz"
    with gzip.open(p, "wt") as f:
        json.dump(doc, f)
    assert np.array_equal(F.read_error_model(p), err)
    assert np.array_equal(F.read_error_model(p, use_err_in=True), err * 2)
    q = tmp_path / "bad.json"
    q.write_text(json.dumps(dict(doc, dada2_rs_command="dada")))
    with pytest.raises(ValueError):
        F.read_error_model(q)
    q.write_text(json.dumps(dict(doc, err_out=err[:15].tolist())))
    with pytest.raises(ValueError):
        F.read_error_model(q)


def test_positive_gap_penalties_are_normalised(pkg):
    """resolve_dada_params (main.rs:4403-4418): gap_p = 8 means -8, and
    homo_gap_p defaults to the normalised gap_p; a positive homo_gap_p is
    normalised as well."""
    a, b = pkg.align_params(gap_p=8), pkg.align_params(gap_p=-8)
    for f, _ in pkg.AlignParams._fields_:
        assert getattr(a, f) == getattr(b, f), f
    assert a.gap_p == -8 and a.homo_gap_p == -8
    c = pkg.align_params(gap_p=8, homo_gap_p=1)
    assert c.gap_p == -8 and c.homo_gap_p == -1
    assert pkg.align_params(homo_gap_p=-1).homo_gap_p == -1
