"""remove_bimera_denovo (remove_bimera.rs:49-120) over a sequence table, every
bimera test on the device:

  * consensus (remove_bimera.rs:124-163): table_bimera2 (d2g_table_bimera2),
    then a sequence is bimeric when flagged in every sample it is in, or
    when nflag > 0 and nflag >= (nsam - ignore_n_negatives) x
    min_sample_fraction (the vote of R's removeBimeraDenovo);
  * pooled (:165-211): abundances summed over samples (u32), each sequence
    tested against the sequences more abundant than min_fold x its own with
    at least min_abund reads (ascending column order, two parents at least);
  * per-sample (:213-263): the same test per (sample, sequence) cell on that
    sample's counts, flagged cells zeroed.

Columns flagged by consensus or pooled are dropped (drop_columns,
:265-298); per-sample drops the columns left all zero (zero_and_drop,
:300-349).  The counts go to the device as u32 (`as u32`,
remove_bimera.rs:66-71)."""
from __future__ import annotations

import numpy as np

METHODS = ("consensus", "pooled", "per-sample")


def consensus_vote(nflag, nsam, min_sample_fraction: float, ignore_n_negatives: int):
    """remove_bimera.rs:151-161 per sequence."""
    out = []
    for f, s in zip(np.asarray(nflag).tolist(), np.asarray(nsam).tolist()):
        out.append(f >= s or (f > 0 and float(f) >= float(max(s - ignore_n_negatives, 0))
                                                     * min_sample_fraction))
    return np.asarray(out, bool)


def _parents(abund, j, min_fold: float, min_abund: int):
    """remove_bimera.rs:192-199 / 244-251: the candidate parents of column j."""
    a = float(abund[j])
    return [k for k in range(len(abund))
            if k != j and float(abund[k]) > min_fold * a and int(abund[k]) >= min_abund]


def _cells_bimeric(ctx, seqs, rows, min_fold, min_abund, params):
    """Flags of the cells whose count vectors are `rows` (one u32 abundance
    vector per cell group); one device batch for all queries."""
    queries, parent_idx, where = [], [], []
    flags = [np.zeros(len(seqs), bool) for _ in rows]
    for r, abund in enumerate(rows):
        for j in range(len(seqs)):
            if int(abund[j]) == 0:
                continue
            ps = _parents(abund, j, min_fold, min_abund)
            if len(ps) < 2:
                continue
            queries.append(seqs[j])
            parent_idx.append(ps)
            where.append((r, j))
    if queries:
        got = ctx.is_bimera_pool(queries, seqs, parent_idx, params)
        for (r, j), g in zip(where, got):
            flags[r][j] = bool(g)
    return flags


def bimera_flags(ctx, table: dict, method: str = "consensus", min_fold: float = 1.5,
                 min_abund: int = 2, min_sample_fraction: float = 0.9,
                 ignore_n_negatives: int = 1, params=None):
    """Per-column flags (consensus, pooled) or per-cell flags (per-sample)."""
    from . import bimera_params
    params = params or bimera_params()
    if method not in METHODS:
        raise ValueError(f"method must be one of {METHODS}")
    seqs = [s.encode() for s in table["sequences"]]
    mat = (np.asarray(table["counts"], np.uint64).reshape(len(table["samples"]), len(seqs))
           & np.uint64(0xFFFFFFFF)).astype(np.uint32)
    if method == "consensus":
        nflag, nsam = ctx.table_bimera2(mat, seqs, min_fold, min_abund, params)
        return consensus_vote(nflag, nsam, min_sample_fraction, ignore_n_negatives)
    if method == "pooled":
        pooled = (mat.astype(np.uint64).sum(axis=0) & np.uint64(0xFFFFFFFF)).astype(np.uint32)
        return _cells_bimeric(ctx, seqs, [pooled], min_fold, min_abund, params)[0]
    return np.stack(_cells_bimeric(ctx, seqs, list(mat), min_fold, min_abund, params))


def remove_bimera_denovo(ctx, table: dict, method: str = "consensus", **kw) -> dict:
    """The filtered table (same keys as make_sequence_table's)."""
    nrow, ncol = len(table["samples"]), len(table["sequences"])
    if nrow == 0 or ncol == 0:
        return table
    flags = bimera_flags(ctx, table, method, **kw)
    counts = np.asarray(table["counts"], np.uint64).reshape(nrow, ncol)
    if method == "per-sample":
        counts = np.where(flags, np.uint64(0), counts)
        keep = (counts > 0).any(axis=0)
    else:
        keep = ~flags
    cols = np.nonzero(keep)[0]
    return {"samples": list(table["samples"]),
            "sequences": [table["sequences"][j] for j in cols],
            "sequence_ids": [table["sequence_ids"][j] for j in cols],
            "counts": counts[:, cols].tolist()}
