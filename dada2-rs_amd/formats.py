"""On-disk formats that feed the device path: FASTQ dereplication, the Raw
build of dada_uniques_cached, and the tagged error-model JSON.

Restated from dada2-rs v0.2.0:
  * dereplicate (derep.rs:41-200): reads grouped by exact sequence bytes;
    per-position Phred sums (quality byte - offset) kept as integers;
    uniques ordered by read count descending, ties by sequence bytes
    ascending; `map` gives each read's unique;
  * Raw::from_qual_sums (containers.rs:120-137): the stored quality is
    round(sum / reads) with Rust's round (half away from zero); for the
    integer sums here that is (2 sum + reads) // (2 reads), exactly;
  * dada_uniques_cached validation (dada.rs:243-317): lengths below SEQLEN
    (9999) and above k, only A C G T U (either case), qualities as long as
    their sequence;
  * nt_encode (misc.rs:346-356): A=1 C=2 G=3 T/U=4;
  * read_tagged_json (misc.rs:116-147) and the error-model fields nq,
    err_in, err_out (main.rs:4361-4374), gzip by the .gz extension;
  * the tagged writer (Tagged, misc.rs:116-135: the command and version
    first, then the payload's fields) and the dada output document
    (DadaOutput, main.rs:1139-1250; AsvEntry and birth_type_str,
    main.rs:4776-4817; non-finite floats as null, as serde_json writes them);
  * make_sequence_table (sequence_table.rs:69-330): dada / dada-pooled /
    dada-pseudo and merge-pairs documents to a samples x sequences count
    table, columns by decreasing total abundance (or samples present, or
    first-seen order), md5 or sha1 sequence ids.  The reference's first-seen
    order comes from HashMap iteration, which is unspecified; here it is
    the order sequences first appear in the documents, so ties in the sort
    keep that order.
"""
from __future__ import annotations

import gzip
import hashlib
import json
import math
from dataclasses import dataclass
from pathlib import Path

import numpy as np

SEQLEN = 9999  # dada.rs: maximum input length (exclusive)

_ENC = np.zeros(256, np.uint8)
for _c, _v in ((b"A", 1), (b"C", 2), (b"G", 3), (b"T", 4), (b"U", 4)):
    _ENC[_c[0]] = _v
    _ENC[_c.lower()[0]] = _v


def _open(path):
    path = Path(path)
    return gzip.open(path, "rb") if path.suffix == ".gz" else open(path, "rb")


def read_fastq(path):
    """(sequence bytes, quality bytes) per record of a 4-line FASTQ file."""
    out = []
    with _open(path) as f:
        lines = f.read().split(b"\n")
    i, n = 0, len(lines)
    while i < n:
        if not lines[i].strip():
            i += 1
            continue
        if i + 3 >= n or not lines[i].startswith(b"@") or not lines[i + 2].startswith(b"+"):
            raise ValueError(f"malformed or truncated FASTQ record at line {i + 1} of {path}")
        seq, qual = lines[i + 1].rstrip(b"\r"), lines[i + 3].rstrip(b"\r")
        if len(seq) != len(qual):
            raise ValueError(f"sequence and quality lengths differ at line {i + 1} of {path}")
        out.append((seq, qual))
        i += 4
    return out


@dataclass
class Derep:
    """derep.rs:41-60: uniques (sequence bytes, count) by count descending then
    sequence ascending; quals[i] the per-position Phred sums (u32); map[r]
    the unique of read r."""
    uniques: list
    quals: list
    map: np.ndarray


def dereplicate(records, phred_offset: int = 33) -> Derep:
    """dereplicate / PartialDerep (derep.rs:62-200) over (seq, qual) byte pairs."""
    index: dict[bytes, int] = {}
    seqs: list[bytes] = []
    counts: list[int] = []
    sums: list[np.ndarray] = []
    rmap = np.empty(len(records), np.int64)
    for r, (seq, qual) in enumerate(records):
        k = index.get(seq)
        if k is None:
            k = len(seqs)
            index[seq] = k
            seqs.append(seq)
            counts.append(0)
            sums.append(np.zeros(0, np.int64))
        counts[k] += 1
        rmap[r] = k
        q = np.frombuffer(qual, np.uint8).astype(np.int64) - phred_offset
        if len(q) > len(sums[k]):
            grown = np.zeros(len(q), np.int64)
            grown[:len(sums[k])] = sums[k]
            sums[k] = grown
        sums[k][:len(q)] += q
    order = sorted(range(len(seqs)), key=lambda i: (-counts[i], seqs[i]))
    new_of = np.empty(len(seqs), np.int64)
    new_of[order] = np.arange(len(seqs))
    quals = []
    for i in order:
        s = sums[i]
        if (s > 0xFFFFFFFF).any():  # checked_qual_sum, derep.rs:27-39
            raise OverflowError("per-position Phred sum overflows u32")
        quals.append(np.maximum(s, 0).astype(np.uint32))
    return Derep([(seqs[i], counts[i]) for i in order], quals, new_of[rmap])


def derep_fastq(path, phred_offset: int = 33) -> Derep:
    return dereplicate(read_fastq(path), phred_offset)


def nt_encode(seq: bytes) -> np.ndarray:
    """misc.rs:346-356 (A=1 C=2 G=3 T/U=4); other bytes are rejected by the
    caller's validation."""
    return _ENC[np.frombuffer(seq, np.uint8)]


def raws_from_derep(d: Derep, k: int = 5):
    """dada_uniques_cached's Raw build (dada.rs:243-317, containers.rs:120-137):
    (encoded sequences, rounded mean qualities, reads), with its validation."""
    if not d.uniques:
        raise ValueError("Zero input sequences.")
    lens = [len(s) for s, _ in d.uniques]
    if max(lens) >= SEQLEN:
        raise ValueError(f"Input sequences exceed the maximum allowed length ({SEQLEN}).")
    if min(lens) <= k:
        raise ValueError(f"All input sequences must be longer than the k-mer size ({k}).")
    seqs, quals, reads = [], [], []
    for i, ((s, c), qs) in enumerate(zip(d.uniques, d.quals)):
        enc = nt_encode(s)
        bad = np.nonzero(enc == 0)[0]
        if len(bad):
            raise ValueError(f"Sequence {i} contains a non-ACGT base {chr(s[bad[0]])!r} at "
                             f"position {int(bad[0])} (0-based).")
        if len(qs) != len(s):
            raise ValueError(f"Sequence {i}: quality length {len(qs)} does not match "
                             f"sequence length {len(s)}.")
        q = (2 * qs.astype(np.uint64) + np.uint64(c)) // np.uint64(2 * c)
        if (q > 255).any():
            raise ValueError(f"Sequence {i}: mean quality above 255")
        seqs.append(enc)
        quals.append(q.astype(np.uint8))
        reads.append(c)
    return seqs, quals, np.asarray(reads, np.uint32)


def read_tagged_json(path, expected):
    """read_tagged_json (misc.rs:116-147): the payload of a dada2-rs JSON
    output whose dada2_rs_command is one of `expected` (.gz by extension)."""
    with _open(path) as f:
        doc = json.loads(f.read().decode())
    cmd = doc.get("dada2_rs_command")
    if cmd not in expected:
        raise ValueError(f"{path}: dada2_rs_command is {cmd!r}, expected one of {expected}")
    return doc


def read_error_model(path, use_err_in: bool = False) -> np.ndarray:
    """The error matrix of a learn-errors / errors-from-sample JSON
    (main.rs:4361-4374): err_out (or err_in), 16 rows of nq, as a 16 x nq
    f64 array."""
    doc = read_tagged_json(path, ["learn-errors", "errors-from-sample"])
    nq = int(doc["nq"])
    rows = doc["err_in"] if use_err_in else doc["err_out"]
    if len(rows) != 16 or any(len(r) != nq for r in rows):
        raise ValueError(f"Error model matrix must be 16 x {nq}, got {len(rows)} rows")
    return np.asarray(rows, np.float64)


# ---- writers ---------------------------------------------------------------

FORMAT_VERSION = "0.2.0"  # the dada2-rs release whose formats these mirror
BIRTH_TYPES = ("Initial", "Abundance", "Prior", "Singleton")  # main.rs:4792-4799
_DEC = np.frombuffer(b"?ACGTN", np.uint8)  # nt_decode, misc.rs:360-370


def nt_decode(enc) -> str:
    return _DEC[np.asarray(enc, np.uint8)].tobytes().decode()


def _finite(x):
    """serde_json writes NaN and infinities as null."""
    if isinstance(x, float) and not math.isfinite(x):
        return None
    if isinstance(x, list):
        return [_finite(v) for v in x]
    return x


def tagged(command: str, doc: dict) -> dict:
    """Tagged (misc.rs:116-135): the command and version, then the payload."""
    out = {"dada2_rs_command": command, "dada2_rs_version": FORMAT_VERSION}
    out.update(doc)
    return out


def write_tagged_json(path, command: str, doc: dict, compact: bool = False):
    text = json.dumps(_finite_doc(tagged(command, doc)), indent=None if compact else 2,
                      allow_nan=False)
    path = Path(path)
    data = text.encode()
    if path.suffix == ".gz":
        with gzip.open(path, "wb") as f:
            f.write(data)
    else:
        path.write_bytes(data)


def _finite_doc(x):
    if isinstance(x, dict):
        return {k: _finite_doc(v) for k, v in x.items()}
    if isinstance(x, (list, tuple)):
        return [_finite_doc(v) for v in x]
    if isinstance(x, (np.floating, float)):
        v = float(x)
        return v if math.isfinite(v) else None
    if isinstance(x, np.integer):
        return int(x)
    if isinstance(x, np.bool_):
        return bool(x)
    return x


def run_params(align, dada, pool: bool = False, n_prior: int = 0) -> dict:
    """DadaRunParams (main.rs:142-172) of a run, in the reference's field
    order, from the package's AlignParams and DadaParams."""
    return {"omega_a": dada.omega_a, "omega_c": dada.omega_c, "omega_p": dada.omega_p,
            "min_fold": dada.min_fold, "min_hamming": int(dada.min_hamming),
            "min_abund": int(dada.min_abund), "detect_singletons": bool(dada.detect_singletons),
            "band": int(align.band), "homo_gap_p": int(align.homo_gap_p),
            "gap_p": int(align.gap_p), "match_score": int(align.match_score),
            "mismatch": int(align.mismatch), "max_clust": int(dada.max_clust),
            "greedy": bool(dada.greedy), "use_quals": bool(dada.use_quals),
            "kdist_cutoff": align.kdist_cutoff, "kmer_size": int(align.kmer_size),
            "use_kmers": bool(align.use_kmers), "pool": bool(pool), "n_prior": int(n_prior)}


def dada_output(result, seqs, sample: str, input_file: str, run_params: dict,
                aux: dict | None = None) -> dict:
    """The dada command's payload (DadaOutput, main.rs:1139-1250) for a
    run_dada result over encoded uniques `seqs`: ASVs in cluster order, the
    read map, the run's counters and parameters, and the aux outputs when
    given (Context.compute_aux's dict)."""
    asvs = [{"sequence": nt_decode(seqs[c["center"]]), "abundance": int(c["reads"]),
             "birth_type": BIRTH_TYPES[c["birth_type"]], "birth_pval": c["birth_pval"],
             "birth_fold": c["birth_fold"], "birth_e": c["birth_e"]} for c in result.clusters]
    total = sum(int(c["reads"]) for c in result.clusters) & 0xFFFFFFFF
    doc = {"sample": sample, "input_file": input_file, "num_asvs": len(asvs),
           "total_reads": total, "asvs": asvs,
           "stats": {"nalign": int(result.stats["nalign"]), "nshroud": int(result.stats["nshroud"])},
           "params": dict(run_params),
           "map": [None if m < 0 else int(m) for m in np.asarray(result.map)]}
    if aux is not None:
        doc["aux"] = {
            "cluster_stats": [dict({k: v for k, v in st.items() if k != "center"},
                                   sequence=nt_decode(seqs[st["center"]]))
                              for st in aux["cluster_stats"]],
            "cluster_quality": np.asarray(aux["cluster_quality"]).tolist(),
            "cluster_quality_maxlen": int(np.asarray(aux["cluster_quality"]).shape[1]),
            "birth_subs": [{"cluster": c, "pos": p, "nt0": chr(a), "nt1": chr(b), "qual": q}
                           for c, p, a, b, q in aux["birth_subs"]],
            "transitions": np.asarray(aux["transitions"]).reshape(-1).tolist(),
            "transitions_ncol": int(aux["transitions_ncol"]),
        }
    return doc


# ---- the sequence table ----------------------------------------------------

def _load(doc_or_path):
    if isinstance(doc_or_path, dict):
        return doc_or_path, None
    with _open(doc_or_path) as f:
        return json.loads(f.read().decode()), Path(doc_or_path)


def _file_stem(path: Path) -> str:
    return Path(path.name).stem if path is not None else "unknown"


def parse_counts(doc_or_path, sample_name: str | None = None):
    """parse_file (sequence_table.rs:69-134): (sample name, {sequence:
    count}) per sample of one dada-like or merge-pairs document."""
    doc, path = _load(doc_or_path)
    cmd = doc.get("dada2_rs_command")
    if cmd is None:
        raise ValueError(f"{path}: missing dada2_rs_command tag (expected 'dada', 'dada-pooled', "
                         "'dada-pseudo', or 'merge-pairs')")
    if cmd in ("dada", "dada-pooled", "dada-pseudo"):
        name = sample_name or doc.get("sample") or _file_stem(path)
        counts = {}
        for a in doc["asvs"]:  # a later duplicate replaces an earlier one
            counts[a["sequence"]] = int(a["abundance"])
        return [(name, counts)]
    if cmd == "merge-pairs":
        out = []
        for r in doc["samples"]:
            counts = {}
            for m in r["merged"]:
                if m["accept"] and m["sequence"]:
                    counts[m["sequence"]] = int(m["abundance"])
            out.append((r["sample"], counts))
        return out
    raise ValueError(f"{path}: dada2_rs_command={cmd!r}, expected 'dada', 'dada-pooled', or "
                     "'merge-pairs'")


def make_sequence_table(inputs, sample_names=None, order_by: str = "abundance",
                        hash_algo: str = "md5") -> dict:
    """make_sequence_table (sequence_table.rs:233-330) over documents or
    paths: {samples, sequences, sequence_ids, counts[sample][sequence]}."""
    samples = []
    for i, x in enumerate(inputs):
        samples.extend(parse_counts(x, sample_names[i] if sample_names else None))
    index: dict[str, int] = {}
    for _, counts in samples:
        for seq in counts:
            index.setdefault(seq, len(index))
    seqs = list(index)
    mat = np.zeros((len(samples), len(seqs)), np.uint64)
    for i, (_, counts) in enumerate(samples):
        for seq, c in counts.items():
            mat[i, index[seq]] += np.uint64(c)
    if order_by == "abundance":
        key = mat.sum(axis=0)
    elif order_by == "nsamples":
        key = (mat > 0).sum(axis=0)
    else:
        key = None
    order = (sorted(range(len(seqs)), key=lambda j: -int(key[j])) if key is not None
             else list(range(len(seqs))))
    seqs = [seqs[j] for j in order]
    digest = hashlib.sha1 if hash_algo == "sha1" else hashlib.md5
    return {"samples": [n for n, _ in samples], "sequences": seqs,
            "sequence_ids": [digest(s.encode()).hexdigest() for s in seqs],
            "counts": mat[:, order].tolist()}


def read_sequence_table(path) -> dict:
    """A make-sequence-table or remove-bimera-denovo document
    (main.rs:2752)."""
    doc = read_tagged_json(path, ["make-sequence-table", "remove-bimera-denovo"])
    return {k: doc[k] for k in ("samples", "sequences", "sequence_ids", "counts")}
