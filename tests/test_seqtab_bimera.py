"""The sequence table (formats.make_sequence_table, sequence_table.rs), the
dada JSON writer, and remove_bimera_denovo's decisions (remove_bimera.rs):
the reference's own parse case, hand-built tables, and the oracle's
restatement on a constructed chimera."""
import gzip
import json

import numpy as np
import pytest

import oracle
from dada2_rs_amd import formats, synth
from dada2_rs_amd import remove_bimera as RB


def test_parse_gzipped_dada_output(tmp_path):
    """sequence_table.rs:152-173 (parse_file_reads_gzipped_dada_output)."""
    path = tmp_path / "F3D0.json.gz"
    with gzip.open(path, "wb") as f:
        f.write(b'{"dada2_rs_command":"dada","sample":"F3D0",'
                b'"asvs":[{"sequence":"ACGTACGT","abundance":7}]}')
    assert formats.parse_counts(path) == [("F3D0", {"ACGTACGT": 7})]


def _dada(sample, asvs, cmd="dada"):
    return {"dada2_rs_command": cmd, "sample": sample,
            "asvs": [{"sequence": s, "abundance": a} for s, a in asvs]}


def test_sequence_table_orders_and_ids():
    docs = [_dada("a", [("AAAA", 5), ("CCCC", 1), ("GGGG", 2)]),
            _dada("b", [("CCCC", 3), ("TTTT", 4)], cmd="dada-pooled")]
    t = formats.make_sequence_table(docs)
    # totals: AAAA 5, CCCC 4, GGGG 2, TTTT 4 -> AAAA, then the 4s in first-seen order
    assert t["samples"] == ["a", "b"]
    assert t["sequences"] == ["AAAA", "CCCC", "TTTT", "GGGG"]
    assert t["counts"] == [[5, 1, 0, 2], [0, 3, 4, 0]]
    assert t["sequence_ids"][0] == "098890dde069e9abad63f19a0d9e1f32"  # md5("AAAA")
    t = formats.make_sequence_table(docs, order_by="nsamples")
    assert t["sequences"][0] == "CCCC"  # the only sequence in both samples
    t = formats.make_sequence_table(docs, order_by="none", hash_algo="sha1", sample_names=["x", "y"])
    assert t["sequences"] == ["AAAA", "CCCC", "GGGG", "TTTT"] and t["samples"] == ["x", "y"]
    assert t["sequence_ids"][0] == "e2512172abf8cc9f67fdd49eb6cacf2df71bbad3"  # sha1("AAAA")
    merge = {"dada2_rs_command": "merge-pairs", "samples": [
        {"sample": "m1", "merged": [{"sequence": "AAAA", "abundance": 3, "accept": True},
                                    {"sequence": "CCCC", "abundance": 9, "accept": False},
                                    {"sequence": "", "abundance": 9, "accept": True}]}]}
    t = formats.make_sequence_table([merge])
    assert t["sequences"] == ["AAAA"] and t["counts"] == [[3]]
    with pytest.raises(ValueError):
        formats.make_sequence_table([{"asvs": []}])
    with pytest.raises(ValueError):
        formats.make_sequence_table([{"dada2_rs_command": "learn-errors"}])


def test_dada_output_round_trip(tmp_path):
    class R:
        clusters = [{"center": 1, "reads": 30, "birth_type": 0, "birth_pval": 0.0,
                     "birth_fold": 1.0, "birth_e": 40.0},
                    {"center": 0, "reads": 10, "birth_type": 1, "birth_pval": 1e-50,
                     "birth_fold": 3.5, "birth_e": 2.0}]
        map = np.array([1, 0, -1])
        stats = {"nalign": 6, "nshroud": 1}
    seqs = [np.array([1, 2, 3, 4], np.uint8), np.array([4, 4, 1, 1], np.uint8),
            np.array([2, 2, 2, 2], np.uint8)]
    aux = {"cluster_stats": [{"center": 1, "abundance": 30}, {"center": 0, "abundance": 10}],
           "cluster_quality": np.array([[30.0, np.nan], [20.0, 21.0]]),
           "birth_subs": [(1, 0, 84, 65, 33)], "transitions": np.arange(16).reshape(16, 1),
           "transitions_ncol": 1}
    from dada2_rs_amd import align_params, dada_params
    rp = formats.run_params(align_params(), dada_params(), n_prior=2)
    assert list(rp)[:3] == ["omega_a", "omega_c", "omega_p"] and rp["band"] == 16
    assert rp["kdist_cutoff"] == 0.42 and rp["n_prior"] == 2 and rp["greedy"] is True
    doc = formats.dada_output(R, seqs, "s1", "s1.fastq.gz", rp, aux)
    path = tmp_path / "s1.json"
    formats.write_tagged_json(path, "dada", doc)
    back = json.loads(path.read_text())
    assert back["dada2_rs_command"] == "dada" and list(back)[:2] == ["dada2_rs_command",
                                                                     "dada2_rs_version"]
    assert [a["sequence"] for a in back["asvs"]] == ["TTAA", "ACGT"]
    assert [a["birth_type"] for a in back["asvs"]] == ["Initial", "Abundance"]
    assert back["total_reads"] == 40 and back["map"] == [1, 0, None]
    assert back["params"]["omega_a"] == 1e-40 and back["params"]["use_kmers"] is True
    assert back["aux"]["cluster_quality"] == [[30.0, None], [20.0, 21.0]]  # NaN -> null
    assert back["aux"]["birth_subs"][0]["nt0"] == "T"
    assert back["aux"]["cluster_stats"][0]["sequence"] == "TTAA"
    t = formats.make_sequence_table([path])
    assert t["samples"] == ["s1"] and t["counts"] == [[30, 10]]


def test_consensus_vote_by_hand():
    """remove_bimera.rs:151-161 with min_sample_fraction 0.9, ignore 1:
    (1 of 3) 1 >= 1.8 no; (0 of 0) 0 >= 0 yes; (2 of 3) 2 >= 1.8 yes;
    (0 of 2) no; (1 of 1) yes; (1 of 2) 1 >= 0.9 yes."""
    got = RB.consensus_vote([1, 0, 2, 0, 1, 1], [3, 0, 3, 2, 1, 2], 0.9, 1)
    assert got.tolist() == [False, True, True, False, True, True]
    # ignore_n_negatives beyond nsam saturates at zero
    assert RB.consensus_vote([1], [2], 0.9, 5).tolist() == [True]


def chimera_table():
    """Parents A and B (abundant in both samples), C = A's first half + B's
    second half (low abundance), D unrelated (abundant in sample 2 only)."""
    rng = np.random.default_rng(synth.SEED_BASE + 70)
    bases = np.frombuffer(b"ACGT", np.uint8)

    def rnd(n):
        return bases[rng.integers(0, 4, size=n)].tobytes().decode()
    a, b, d = rnd(200), rnd(200), rnd(200)
    c = a[:100] + b[100:]
    docs = [_dada("s1", [(a, 1000), (b, 900), (c, 20), (d, 15)]),
            _dada("s2", [(a, 800), (b, 700), (c, 10), (d, 300)])]
    return formats.make_sequence_table(docs), (a, b, c, d)


@pytest.mark.parametrize("method", RB.METHODS)
def test_oracle_remove_bimera_constructed(method):
    from oracle import remove_bimera as ORB
    t, (a, b, c, d) = chimera_table()
    out = ORB.remove_bimera_denovo(t, method)
    assert c not in out["sequences"]
    assert sorted(out["sequences"]) == sorted([a, b, d])
    j = t["sequences"].index(a)
    assert out["counts"][0][out["sequences"].index(a)] == t["counts"][0][j]
    flags = ORB.bimera_flags(t, method)
    if method == "per-sample":
        assert flags[:, t["sequences"].index(c)].all() and flags.sum() == 2
    else:
        assert flags.tolist() == [s == c for s in t["sequences"]]
