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
    err_in, err_out (main.rs:4361-4374), gzip by the .gz extension.
"""
from __future__ import annotations

import gzip
import json
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
