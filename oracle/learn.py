"""learn.py — restatement of the transition-by-quality counts of a finished
partition.  TEST INFRASTRUCTURE ONLY: the checker of the device library's
d2g_transition_counts and d2g_build_trans_mat; only tests/ may import it.

  * transition_counts (src/error.rs:194-236) over compute_final_subs
    (src/dada.rs:646-673), aligned with compute_aux's final parameters
    (src/dada.rs:612-616: use_kmers false, kdist_cutoff 1.0): each counted
    raw's Sub against its center (the oracle's C sub_new, nwalign.rs:1429),
    then for every center position whose map is not GAP_GLYPH,
    mat[4 (nt0 - 1) + (nt1 - 1)][qual] += reads, qualities >= ncol skipped,
    the u32 sums wrapping;
  * build_trans_mat (src/learn_errors.rs:618-690): the alignment strings of
    raw_align_with_buf (the oracle's C raw_align) over Raws without k-mer
    vectors, walked position by position; qualities clamped to nq - 1, the
    u32 sums saturating, inputs without qualities skipped.

Parity: pinned by a hand-worked case (an insertion, a deletion, a
substitution and an end overhang, counted by hand), by the closed form of
gapless pairs (band 0: a histogram of (center base, query base, quality) at
equal positions), by the two restatements agreeing wherever their quality
and overflow rules coincide, and by the row sums (every aligned center
position counts once).
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import AlignParams, Buf, RawSet, raw_align, sub_new

GAP_GLYPH = 9999  # nwalign.rs:27


def final_align(align: AlignParams) -> AlignParams:
    """compute_aux's final-subs parameters (dada.rs:612-616)."""
    p = AlignParams.from_buffer_copy(align)
    p.use_kmers = 0
    p.kdist_cutoff = 1.0
    return p


def transition_counts(rs: RawSet, center_of, weight, align: AlignParams, has_quals: bool,
                      ncol: int) -> np.ndarray:
    """transition_counts (error.rs:194-236) over compute_final_subs
    (dada.rs:646-673); center_of[i] < 0 for none, weight[i] = reads when
    Raw.correct, else 0.  16 x ncol (x 1 without qualities), u32."""
    ncol = ncol if has_quals else 1
    fp = final_align(align)
    buf = Buf()
    mat = [0] * (16 * ncol)
    for i in range(len(rs)):
        c, w = int(center_of[i]), int(weight[i])
        if c < 0 or w == 0:
            continue
        r = sub_new(rs, c, i, fp, buf)
        if r is None:  # no Sub (dada.rs:666-668): nothing counted
            continue
        s, _buf = r
        cseq, qseq = rs.seqs[c], rs.seqs[i]
        for pos0 in range(s.len0):
            pos1 = s.map[pos0]
            if pos1 == GAP_GLYPH:
                continue
            nti0 = max(int(cseq[pos0]) - 1, 0)
            nti1 = max(int(qseq[pos1]) - 1, 0)
            qual = int(rs.quals[i][pos1]) if (has_quals and rs.quals is not None) else 0
            t = 4 * nti0 + nti1
            if t < 16 and qual < ncol:
                mat[t * ncol + qual] += w
    return (np.asarray(mat, dtype=object) % (1 << 32)).astype(np.uint32).reshape(16, ncol)


def build_trans_mat(seqs, quals, center_of, weight, align: AlignParams, nq: int) -> np.ndarray:
    """build_trans_mat (learn_errors.rs:618-690): seqs / quals the inputs
    (rounded qualities, or None), center_of[i] the unique at the center of
    i's cluster (< 0 when the result maps no cluster), weight[i] the input's
    abundance.  16 x nq, u32, saturating."""
    mat = [0] * (16 * nq)
    if quals is None:
        return np.zeros((16, nq), np.uint32)
    rs = RawSet(seqs, quals, weight, with_kmers=False)  # no k-mer vectors here
    buf = Buf()
    for i in range(len(rs)):
        c, w = int(center_of[i]), int(weight[i])
        if c < 0 or w == 0:
            continue
        ok, _kd, _ko, al, _m = raw_align(rs, c, i, align, buf)
        if not ok:
            continue
        al0, al1 = al
        qpos = 0
        for a0, a1 in zip(al0, al1):
            if 1 <= a0 <= 4 and 1 <= a1 <= 4:
                q = min(int(rs.quals[i][qpos]), nq - 1)
                mat[((a0 - 1) * 4 + (a1 - 1)) * nq + q] += w
            if a1 != ord("-"):
                qpos += 1
    return np.minimum(np.asarray(mat, dtype=object), 0xFFFFFFFF).astype(np.uint32).reshape(16, nq)


def build_trans_mat_c(seqs, quals, center_of, weight, align: AlignParams, nq: int,
                      nthreads: int = 1) -> np.ndarray:
    """build_trans_mat in the oracle's threaded C (oracle/learn.c): the same
    counts as build_trans_mat above, computed as the reference's rayon fold."""
    from . import lib
    if quals is None:
        return np.zeros((16, nq), np.uint32)
    rs = RawSet(seqs, quals, weight, with_kmers=False)
    co = np.ascontiguousarray(np.asarray(center_of, np.int64))
    w = np.ascontiguousarray(np.asarray(weight, np.uint64) & np.uint64(0xFFFFFFFF), np.uint32)
    out = np.zeros(16 * nq, np.uint32)
    rc = lib().orc_build_trans_mat(rs.arr, len(rs), co.ctypes.data_as(C.POINTER(C.c_int64)),
                                   w.ctypes.data_as(C.POINTER(C.c_uint32)), C.byref(align), nq,
                                   nthreads, out.ctypes.data_as(C.POINTER(C.c_uint32)))
    if rc < 0:
        raise RuntimeError(f"oracle build_trans_mat failed with status {rc}")
    return out.reshape(16, nq)


def center_weight(out, reads):
    """(center_of, weight) of a run_dada result: center_of[i] the center of
    i's cluster, weight[i] its reads when the read map keeps it (Raw.correct,
    dada.rs:502-533), else 0."""
    centers = np.asarray([c["center"] for c in out.clusters], np.int64)
    center_of = centers[np.asarray(out.cluster_of, np.int64)]
    weight = np.where(np.asarray(out.map) >= 0, np.asarray(reads, np.int64), 0)
    return center_of, weight


def dada_aux(rs: RawSet, centers, cluster_of, weight, align: AlignParams, has_quals: bool,
             ncol: int, maxlen: int):
    """compute_aux's per-raw part (dada.rs:610-631) over a finished
    partition: transition_counts, each counted raw's Sub.nsubs against its
    center (-1 otherwise) and cluster_quality (error.rs:251-305: per cluster
    and center position below min(center length, maxlen), the read-weighted
    mean quality, NaN where no counted read covers it or without qualities;
    the u32 read sums wrap)."""
    ncl = len(centers)
    center_of = np.asarray([centers[k] if k >= 0 else -1 for k in cluster_of], np.int64)
    trans = transition_counts(rs, center_of, weight, align, has_quals, ncol)
    fp = final_align(align)
    buf = Buf()
    nsubs = np.full(len(rs), -1, np.int64)
    nreads = [[0] * maxlen for _ in range(ncl)]
    qsum = [[0.0] * maxlen for _ in range(ncl)]
    for i in range(len(rs)):
        k, w = int(cluster_of[i]), int(weight[i])
        if k < 0 or w == 0:
            continue
        c = int(centers[k])
        s, _buf = sub_new(rs, c, i, fp, buf)
        nsubs[i] = s.nsubs
        if not has_quals or rs.quals is None:
            continue
        clen = min(len(rs.seqs[c]), maxlen)
        for pos0 in range(clen):
            pos1 = s.map[pos0]
            if pos1 == GAP_GLYPH:
                continue
            nreads[k][pos0] += w
            qsum[k][pos0] += float(rs.quals[i][pos1]) * float(w)
    qual = np.full((ncl, maxlen), np.nan)
    if has_quals and rs.quals is not None:
        for k in range(ncl):
            for p in range(maxlen):
                r = nreads[k][p] % (1 << 32)
                if r > 0:
                    qual[k, p] = qsum[k][p] / float(r)
    return {"transitions": trans, "nsubs": nsubs, "cluster_quality": qual}


_DECODE = b"?ACGTN"  # nt_decode, misc.rs:360-370


def birth_subs(rs: RawSet, centers, birth_from, align: AlignParams, has_quals: bool):
    """compute_birth_subs (dada.rs:679-693) with birth_align (dada.rs:617-621:
    the k-mer settings as given, kdist_cutoff 1.0), then birth_sub_records
    (error.rs:328-352) and cluster_stats' birth_qave (error.rs:90-100).
    Returns (records [(cluster, pos, nt0, nt1, qual or None)], nsubs per
    cluster (-1 without a Sub), qave per cluster (NaN when None))."""
    bp = AlignParams.from_buffer_copy(align)
    bp.kdist_cutoff = 1.0
    buf = Buf()
    ncl = len(centers)
    recs, nsubs, qave = [], [-1] * ncl, [float("nan")] * ncl
    for ci in range(1, ncl):
        pc, cc = centers[birth_from[ci]], centers[ci]
        if pc is None or cc is None or pc < 0 or cc < 0:
            continue
        r = sub_new(rs, int(pc), int(cc), bp, buf)
        if r is None:
            continue
        s, _buf = r
        nsubs[ci] = s.nsubs
        q1 = [int(s.q1[k]) for k in range(s.nsubs)] if (has_quals and rs.quals is not None) else []
        for k in range(s.nsubs):
            recs.append((ci, int(s.pos[k]), _DECODE[s.nt0[k]], _DECODE[s.nt1[k]],
                         q1[k] if q1 else None))
        if q1:
            qave[ci] = sum(float(q) for q in q1) / s.nsubs
    return recs, nsubs, qave
