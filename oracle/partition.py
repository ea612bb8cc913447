"""partition.py — restatement of dada2-rs's partition loop: run_dada
(src/dada.rs:749-1150) and the final per-raw p-value pass and read map
(src/dada.rs:502-533).  TEST INFRASTRUCTURE ONLY: the checker of the device
library's d2g_run_dada; only tests/ may import it.

Each cluster's compare is the oracle's C restatement of the b_compare map
(compare_center); get_pA / calc_pA are the oracle's C restatements of
pval.rs.  The rest follows the reference step by step:

  * B::new / init (containers.rs:337-391), assign_center (:450-463, the last
    raw among those with the most reads);
  * b_compare's serial store (cluster.rs:130-160);
  * b_p_update with greedy locks (pval.rs:45-130);
  * b_bud (cluster.rs:1331-1480; b_bud_incremental, :1501, is its cached
    equivalent);
  * the shuffle: repeated b_shuffle2 passes (cluster.rs:615-700), which
    b_shuffle_converge (:908) reproduces byte for byte; with the carried map
    (the default, :79-82, 957) calls after the first spend their first
    MAX_SHUFFLE iteration on a reconcile (:1037, 1066-1073), so they make at
    most max_shuffle - 1 move passes.

Parity: the loop itself is pinned only by properties the algorithm fixes
(recovery of planted variants, the no-bud and max_clust limits, a converged
final partition, swap-remove membership order checked on a hand-worked
case); the reference's own partition tests compare two of its code paths
with each other and hold no expected outputs.
"""
from __future__ import annotations

import sys
from dataclasses import dataclass

import numpy as np

from . import RawSet, calc_pA, compare_center, get_pA

HAMMING_NONE = 0xFFFFFFFF


@dataclass
class Comp:
    """Comparison, containers.rs:15-20."""
    i: int = 0
    index: int = 0
    lam: float = 0.0
    hamming: int = 0


class Bi:
    """The fields of Bi (containers.rs:200-265) the loop uses."""

    def __init__(self):
        self.raws: list[int] = []
        self.reads = 0
        self.center = None
        self.update_e = True
        self.check_locks = True
        self.self_lam = 0.0
        self.birth_type = 0  # BirthType::Initial (containers.rs:59-70)
        self.birth_from = 0
        self.birth_pval = 0.0
        self.birth_fold = 1.0
        self.birth_e = 0.0
        self.birth_comp = Comp()
        self.comp: list[Comp] = []


class Partition:
    """B (containers.rs:268-337) over the oracle's RawSet."""

    def __init__(self, rs: RawSet, reads, prior):
        n = len(rs)
        self.rs = rs
        self.reads_of = [int(r) for r in reads]
        self.prior = [bool(x) for x in prior] if prior is not None else [False] * n
        self.lock = [False] * n
        self.p = [0.0] * n
        self.comp = [Comp() for _ in range(n)]
        self.e_minmax = [-999.0] * n
        self.total = sum(self.reads_of) & 0xFFFFFFFF  # B.reads is a u32 sum
        self.nalign = 0
        self.nshroud = 0
        c0 = Bi()
        c0.birth_e = float(self.total)
        c0.raws = list(range(n))
        c0.reads = sum(self.reads_of)
        self.cl = [c0]
        self.assign_center(0)

    def add_raw(self, ci, r):  # bi_add_raw, containers.rs:403-417
        c = self.cl[ci]
        c.raws.append(r)
        c.reads += self.reads_of[r]
        c.update_e = True

    def pop_raw(self, ci, pos):  # bi_pop_raw, containers.rs:428-439 (swap_remove)
        c = self.cl[ci]
        r = c.raws[pos]
        last = c.raws.pop()
        if pos < len(c.raws):
            c.raws[pos] = last
        c.reads -= self.reads_of[r]
        c.update_e = True
        return r

    def assign_center(self, ci):  # containers.rs:450-463
        best = None
        for r in self.cl[ci].raws:
            if best is None or self.reads_of[r] >= self.reads_of[best]:
                best = r
        if best is not None:
            self.cl[ci].center = best


def _compare(B: Partition, i: int, err, align, greedy, use_quals, nthreads=1):
    """b_compare, cluster.rs:106-164 (the map over nthreads, cluster.rs:279-330)."""
    rs = B.rs
    for r in range(len(rs)):
        rs.arr[r].lock = int(B.lock[r])
    params = type(align).from_buffer_copy(align)
    if i == 0:
        params.kdist_cutoff = 1.0  # dada.rs:835-838
    center = B.cl[i].center
    lam, ham, skp, _, _ = compare_center(rs, center, err, err.shape[1], params,
                                         use_quals=use_quals, greedy=greedy, nthreads=nthreads)
    for r in range(len(rs)):
        if not skp[r]:
            B.nalign += 1
            if int(ham[r]) == HAMMING_NONE:
                B.nshroud += 1
    B.cl[i].self_lam = float(lam[center])
    for c in store(B.e_minmax, lam, ham, float(B.total), B.reads_of[center], i):
        B.cl[i].comp.append(c)
        if i == 0 or c.index == center:
            B.comp[c.index] = c


def store(e_minmax, lam, ham, total: float, center_reads: int, i: int) -> list:
    """The serial store of b_compare_parallel (cluster.rs:351-397) for
    cluster i: raw r keeps its comparison when lam[r] * total > e_minmax[r]
    (strict), e_minmax[r] then rises to lam[r] * center_reads if larger; a
    shrouded comparison is kept with hamming 0.  Mutates e_minmax; returns
    the kept comparisons in raw order."""
    out = []
    for r in range(len(lam)):
        lmb = float(lam[r])
        if lmb * total > e_minmax[r]:
            e = lmb * float(center_reads)
            if e > e_minmax[r]:
                e_minmax[r] = e
            out.append(Comp(i, r, lmb, 0 if int(ham[r]) == HAMMING_NONE else int(ham[r])))
    return out


def _p_update(B: Partition, greedy, detect_singletons):
    """b_p_update, pval.rs:45-130."""
    for c in B.cl:
        if c.update_e:
            for r in c.raws:
                B.p[r] = get_pA(B.reads_of[r], B.comp[r].lam, B.comp[r].hamming, c.reads,
                                prior=B.prior[r], detect_singletons=detect_singletons)
            c.update_e = False
        if greedy and c.check_locks:
            cr = B.reads_of[c.center] if c.center is not None else 0
            for r in c.raws:
                if float(cr) * B.comp[r].lam > float(B.reads_of[r]):
                    B.lock[r] = True
                if r == c.center:
                    B.lock[r] = True
            c.check_locks = False


def _bud(B: Partition, omega_a, omega_p, min_fold, min_hamming, min_abund):
    """b_bud, cluster.rs:1331-1480: the new cluster's index, or None."""
    nraw = float(len(B.rs))
    c0 = B.cl[0].center
    mini = mini_prior = None
    min_p = min_p_prior = B.p[c0]
    min_reads = min_reads_prior = B.reads_of[c0]
    for ci, c in enumerate(B.cl):
        for pos in range(1, len(c.raws)):
            r = c.raws[pos]
            rd = B.reads_of[r]
            if rd < min_abund or B.comp[r].hamming < min_hamming:
                continue
            if not (min_fold <= 1.0 or float(rd) >= min_fold * B.comp[r].lam * float(c.reads)):
                continue
            p = B.p[r]
            if p < min_p or (p == min_p and rd > min_reads):
                mini, min_p, min_reads = (ci, pos), p, rd
            if B.prior[r] and (p < min_p_prior or (p == min_p_prior and rd > min_reads_prior)):
                mini_prior, min_p_prior, min_reads_prior = (ci, pos), p, rd
    p_a = min_p * nraw
    if p_a < omega_a and mini is not None:
        return _pop_into(B, mini, 1, p_a)
    if min_p_prior < omega_p and mini_prior is not None:
        return _pop_into(B, mini_prior, 2, min_p_prior)
    return None


def _pop_into(B: Partition, where, birth_type, pval):
    ci, pos = where
    r = B.cl[ci].raws[pos]
    expected = B.comp[r].lam * float(B.cl[ci].reads)
    nb = Bi()
    nb.birth_type = birth_type
    nb.birth_from = ci if birth_type == 1 else 0
    nb.birth_pval = pval
    nb.birth_fold = float(B.reads_of[r]) / max(expected, sys.float_info.min)  # inf allowed
    nb.birth_e = expected
    nb.birth_comp = B.comp[r]
    B.pop_raw(ci, pos)
    B.cl.append(nb)
    ni = len(B.cl) - 1
    B.add_raw(ni, r)
    B.assign_center(ni)
    return ni


def shuffle_pass(B: Partition) -> int:
    """b_shuffle2, cluster.rs:615-700: one move pass; returns the moves."""
    n = len(B.rs)
    emax = [float("-inf")] * n
    best = [Comp() for _ in range(n)]
    c0r = float(B.cl[0].reads)
    for c in B.cl[0].comp:
        emax[c.index] = c.lam * c0r
        best[c.index] = c
    for c in B.cl[1:]:
        cr = float(c.reads)
        for cp in c.comp:
            e = cp.lam * cr
            if e > emax[cp.index]:
                emax[cp.index] = e
                best[cp.index] = cp
    moves = 0
    for ci in range(len(B.cl)):
        pos = len(B.cl[ci].raws)
        while pos > 0:
            pos -= 1
            r = B.cl[ci].raws[pos]
            if best[r].i == ci or B.cl[ci].center == r:
                continue
            B.pop_raw(ci, pos)
            B.add_raw(best[r].i, r)
            B.comp[r] = best[r]
            moves += 1
    return moves


def post_hoc_pvals(B: Partition) -> list:
    """post_hoc_pvals (error.rs:145-180): for each center, calc_pA of its
    reads against the reads the other clusters' stored comparisons expect
    of it (sum over clusters in order, comparisons in stored order)."""
    cluster_at = {c.center: ci for ci, c in enumerate(B.cl) if c.center is not None}
    tot = [0.0] * len(B.cl)
    for ci, c in enumerate(B.cl):
        for cp in c.comp:
            j = cluster_at.get(cp.index)
            if j is not None and j != ci:
                tot[j] += cp.lam * float(c.reads)
    return [calc_pA(B.reads_of[c.center], tot[ci], True) if c.center is not None else 1.0
            for ci, c in enumerate(B.cl)]


@dataclass
class DadaOut:
    clusters: list
    cluster_of: np.ndarray
    p: np.ndarray
    map: np.ndarray
    comp_lambda: np.ndarray
    comp_hamming: np.ndarray
    nalign: int
    nshroud: int
    shuffle_passes: int
    partition: Partition


def extend_err(err: np.ndarray, qmax: int) -> np.ndarray:
    """dada_uniques_cached's error-model extension (dada.rs:337-384): when
    the highest rounded quality qmax reaches the matrix's column count, the
    matrix grows to qmax + 1 columns, each new column a copy of the last
    learned one; otherwise it is returned unchanged."""
    err = np.ascontiguousarray(err, np.float64)
    old = err.shape[1]
    if old == 0 or qmax < old:
        return err
    ext = np.zeros((16, qmax + 1), np.float64)
    for t in range(16):
        ext[t, :old] = err[t]
        for q in range(old, qmax + 1):
            ext[t, q] = err[t, old - 1]
    return ext


def run_dada(rs: RawSet, err: np.ndarray, align, prior=None, omega_a=1e-40, omega_p=1e-4,
             omega_c=1e-40, min_fold=1.0, min_hamming=1, min_abund=1, max_clust=0,
             max_shuffle=10, greedy=True, detect_singletons=False, use_quals=True,
             nthreads=1) -> DadaOut:
    """run_dada (dada.rs:749-1150) then the final p pass and map (dada.rs:502-533)."""
    err = np.ascontiguousarray(err, np.float64)
    reads = [rs.arr[i].reads for i in range(len(rs))]
    B = Partition(rs, reads, prior)
    _compare(B, 0, err, align, greedy, use_quals, nthreads)
    _p_update(B, greedy, detect_singletons)
    limit = len(rs) if max_clust == 0 else max_clust
    passes = 0
    first = True
    while len(B.cl) < limit:
        ni = _bud(B, omega_a, omega_p, min_fold, min_hamming, min_abund)
        if ni is None:
            break
        _compare(B, ni, err, align, greedy, use_quals, nthreads)
        cap = max_shuffle if first else max(max_shuffle - 1, 1)
        first = False
        for _ in range(cap):
            passes += 1
            if shuffle_pass(B) == 0:
                break
        _p_update(B, greedy, detect_singletons)
    n = len(rs)
    cluster_of = np.zeros(n, np.uint32)
    pv = np.zeros(n, np.float64)
    mp = np.full(n, -1, np.int32)
    for ci, c in enumerate(B.cl):
        for r in c.raws:
            cluster_of[r] = ci
            if r == c.center:
                p, ok = 1.0, True
            else:
                p = calc_pA(B.reads_of[r], B.comp[r].lam * float(c.reads), True)
                ok = p >= omega_c
            pv[r] = p
            if ok:
                mp[r] = ci
    posthoc = post_hoc_pvals(B)
    clusters = [{"center": c.center, "reads": c.reads, "nraw": len(c.raws),
                 "birth_type": c.birth_type, "birth_from": c.birth_from,
                 "birth_hamming": c.birth_comp.hamming, "birth_pval": c.birth_pval,
                 "birth_fold": c.birth_fold, "birth_e": c.birth_e,
                 "birth_lambda": c.birth_comp.lam, "self_lambda": c.self_lam,
                 "posthoc_pval": posthoc[ci]} for ci, c in enumerate(B.cl)]
    return DadaOut(clusters, cluster_of, pv, mp,
                   np.array([B.comp[r].lam for r in range(n)], np.float64),
                   np.array([B.comp[r].hamming for r in range(n)], np.uint32),
                   B.nalign, B.nshroud, passes, B)
