"""remove_bimera.py — restatement of remove_bimera_denovo
(src/remove_bimera.rs:49-349) over the oracle's C is_bimera and
table_bimera2.  TEST INFRASTRUCTURE ONLY: the checker of the package's
device-backed remove_bimera_denovo; only tests/ may import it.

Each step follows the reference:
  * the column-major u32 matrix (`counts[i][j] as u32`, :66-71);
  * consensus (:124-163): table_bimera2, then
    nflag >= nsam || (nflag > 0 && nflag >= (nsam -sat ignore) * frac);
  * pooled (:165-211): u32 sums over samples; for each sequence with
    reads, the parents k != j with pooled[k] > min_fold x pooled[j] (f64)
    and pooled[k] >= min_abund, ascending; fewer than two: not bimeric;
  * per-sample (:213-263): the same per cell on the sample's counts;
  * drop_columns (:265-298) and zero_and_drop (:300-349).

Parity: pinned by hand-worked tables (a two-parent chimera flagged by
every method, its parents and an unrelated sequence kept, the vote's
threshold cases counted by hand, the per-sample zeroing); is_bimera and
table_bimera2 themselves are pinned by the reference's chimera constructs.
"""
from __future__ import annotations

import numpy as np

from . import bimera_params, is_bimera, table_bimera2


def _u32_matrix(table):
    nrow, ncol = len(table["samples"]), len(table["sequences"])
    m = np.zeros((nrow, ncol), np.uint32)
    for i in range(nrow):
        for j in range(ncol):
            m[i, j] = int(table["counts"][i][j]) & 0xFFFFFFFF
    return m


def _is_bim(seqs, abund, j, min_fold, min_abund, params):
    a = int(abund[j])
    if a == 0:
        return False
    parents = []
    for k in range(len(seqs)):
        if k != j and float(int(abund[k])) > min_fold * float(a) and int(abund[k]) >= min_abund:
            parents.append(seqs[k])
    if len(parents) < 2:
        return False
    return is_bimera(seqs[j], parents, params)


def bimera_flags(table, method="consensus", min_fold=1.5, min_abund=2, min_sample_fraction=0.9,
                 ignore_n_negatives=1, params=None):
    params = params or bimera_params()
    seqs = [s.encode() for s in table["sequences"]]
    m = _u32_matrix(table)
    nrow, ncol = m.shape
    if method == "consensus":
        nflag, nsam = table_bimera2(m, seqs, min_fold, min_abund, params)
        out = []
        for j in range(ncol):
            f, s = int(nflag[j]), int(nsam[j])
            thr = float(s - ignore_n_negatives if s > ignore_n_negatives else 0)
            out.append(f >= s or (f > 0 and float(f) >= thr * min_sample_fraction))
        return np.asarray(out, bool)
    if method == "pooled":
        pooled = [sum(int(m[i, j]) for i in range(nrow)) & 0xFFFFFFFF for j in range(ncol)]
        return np.asarray([_is_bim(seqs, pooled, j, min_fold, min_abund, params)
                           for j in range(ncol)], bool)
    if method == "per-sample":
        return np.asarray([[_is_bim(seqs, m[i], j, min_fold, min_abund, params)
                            for j in range(ncol)] for i in range(nrow)], bool)
    raise ValueError(method)


def remove_bimera_denovo(table, method="consensus", **kw):
    nrow, ncol = len(table["samples"]), len(table["sequences"])
    if nrow == 0 or ncol == 0:
        return table
    flags = bimera_flags(table, method, **kw)
    counts = [list(map(int, row)) for row in table["counts"]]
    if method == "per-sample":
        for i in range(nrow):
            for j in range(ncol):
                if flags[i][j]:
                    counts[i][j] = 0
        keep = [any(counts[i][j] > 0 for i in range(nrow)) for j in range(ncol)]
    else:
        keep = [not flags[j] for j in range(ncol)]
    cols = [j for j in range(ncol) if keep[j]]
    return {"samples": list(table["samples"]),
            "sequences": [table["sequences"][j] for j in cols],
            "sequence_ids": [table["sequence_ids"][j] for j in cols],
            "counts": [[counts[i][j] for j in cols] for i in range(nrow)]}
