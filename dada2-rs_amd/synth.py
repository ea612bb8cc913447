"""Deterministic synthetic amplicon workloads for the compare path.

Each workload is seeded (seed = 0xDADA0000 + workload number) and built with
numpy's PCG64.  Conventions:
  * bases are encoded A=1 C=2 G=3 T=4 (misc.rs:346-356);
  * the error matrix is 16 x nq f64, row t = 4*ref + query, column = Phred q;
    the fixed matrix uses p(q) = min(0.75, 10^(-q/10)), off-diagonal p/3 and
    diagonal 1-p;
  * per-read quality q[pos] = clamp(round(N(36 - 8 pos/(L-1), 3)), 0, 40);
  * abundances max(1, round(lognormal(2, 1.5)));
  * uniques are ordered by abundance descending, ties by sequence ascending
    (derep.rs:180-185), so index 0 is the most abundant.
"""
from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np

SEED_BASE = 0xDADA0000


def fixed_err(nq: int = 41) -> np.ndarray:
    q = np.arange(nq, dtype=np.float64)
    p = np.minimum(0.75, 10.0 ** (-q / 10.0))
    err = np.empty((16, nq), np.float64)
    for r in range(4):
        for c in range(4):
            err[4 * r + c] = (1.0 - p) if r == c else p / 3.0
    return err


def seeded_err(rng: np.random.Generator, nq: int = 41) -> np.ndarray:
    """Off-diagonal rates log-uniform in [1e-5, 1e-1], each (ref, q) row
    normalised over the four query bases."""
    err = np.empty((16, nq), np.float64)
    for r in range(4):
        off = 10.0 ** rng.uniform(-5.0, -1.0, size=(3, nq))
        diag = 1.0 - off.sum(axis=0)
        k = 0
        for c in range(4):
            if c == r:
                err[4 * r + c] = diag
            else:
                err[4 * r + c] = off[k]
                k += 1
    return err


def quality_profile(rng: np.random.Generator, L: int, mean0=36.0, mean1=28.0, sd=3.0,
                    qmax=40) -> np.ndarray:
    pos = np.arange(L, dtype=np.float64)
    mu = mean0 - (mean0 - mean1) * pos / max(L - 1, 1)
    return np.clip(np.rint(rng.normal(mu, sd)), 0, qmax).astype(np.uint8)


def abundance(rng: np.random.Generator, n: int) -> np.ndarray:
    return np.maximum(1, np.rint(rng.lognormal(2.0, 1.5, size=n))).astype(np.uint32)


def _substitute(rng, seq: np.ndarray, rate: np.ndarray) -> np.ndarray:
    out = seq.copy()
    hit = rng.random(len(seq)) < rate
    if hit.any():
        shift = rng.integers(1, 4, size=int(hit.sum()))
        out[hit] = ((out[hit] - 1 + shift) % 4) + 1
    return out


def _derep_order(seqs: list[np.ndarray], reads: np.ndarray) -> np.ndarray:
    keys = [s.tobytes() for s in seqs]
    return np.array(sorted(range(len(seqs)), key=lambda i: (-int(reads[i]), keys[i])), np.int64)


@dataclass
class Workload:
    name: str
    seqs: list[np.ndarray]
    quals: list[np.ndarray]
    reads: np.ndarray
    err: np.ndarray
    centers: np.ndarray
    band: int = 16
    kmer_size: int = 5
    extra: dict = field(default_factory=dict)

    @property
    def n(self) -> int:
        return len(self.seqs)


def _unique_set(rng, make, n: int, seen: set) -> list[np.ndarray]:
    out = []
    while len(out) < n:
        s = make()
        key = s.tobytes()
        if key in seen:
            continue
        seen.add(key)
        out.append(s)
    return out


def _finish(name, seqs, quals, reads, err, centers_fn, band=16, k=5, extra=None) -> Workload:
    order = _derep_order(seqs, reads)
    seqs = [seqs[i] for i in order]
    quals = [quals[i] for i in order]
    reads = reads[order]
    inv = np.empty_like(order)
    inv[order] = np.arange(len(order))
    centers = centers_fn(inv, reads)
    return Workload(name, seqs, quals, reads, err, np.asarray(centers, np.uint32), band, k,
                    extra or {})


def plain_center(reads: np.ndarray) -> int:
    """assign_center on the initial cluster: the last raw of maximal reads
    (containers.rs:450-463, max_by_key keeps the last maximum)."""
    m = reads.max()
    return int(np.nonzero(reads == m)[0][-1])


def workload_plain(n_variants=20, L=250, n_uniques=2000, seed=SEED_BASE + 0) -> Workload:
    """Plain single-center compare: variants x L nt, substitution copies drawn
    from the fixed error matrix at each read's per-position quality."""
    rng = np.random.default_rng(seed)
    err = fixed_err(41)
    variants = [rng.integers(1, 5, size=L).astype(np.uint8) for _ in range(n_variants)]
    seen = {v.tobytes() for v in variants}
    seqs = list(variants)
    quals = [quality_profile(rng, L) for _ in variants]
    origin = list(range(n_variants))
    while len(seqs) < n_uniques:
        v = int(rng.integers(0, n_variants))
        q = quality_profile(rng, L)
        p = np.minimum(0.75, 10.0 ** (-q.astype(np.float64) / 10.0))
        s = _substitute(rng, variants[v], p)
        if s.tobytes() in seen:
            continue
        seen.add(s.tobytes())
        seqs.append(s)
        quals.append(q)
        origin.append(v)
    reads = abundance(rng, len(seqs))
    return _finish("plain", seqs, quals, reads, err,
                   lambda inv, r: [plain_center(r)], extra={"origin": origin})


def workload_batched(n_variants=200, L=260, n_uniques=20000, lmin=240, lmax=260, sub=0.005,
                     indel=0.0005, seed=SEED_BASE + 1) -> Workload:
    """Batched uneven compare: uniques are variant prefixes of length
    U[lmin, lmax] with substitutions (sub/base) and 1-nt indels (indel/base).
    Centers: the most abundant unique of each variant."""
    rng = np.random.default_rng(seed)
    err = fixed_err(41)
    variants = [rng.integers(1, 5, size=L).astype(np.uint8) for _ in range(n_variants)]
    seqs, quals, origin = [], [], []
    seen: set = set()
    while len(seqs) < n_uniques:
        v = int(rng.integers(0, n_variants))
        Lu = int(rng.integers(lmin, lmax + 1))
        s = _substitute(rng, variants[v][:Lu], np.full(Lu, sub))
        nind = rng.random(Lu) < indel
        if nind.any():
            s = list(s)
            for pos in sorted(np.nonzero(nind)[0], reverse=True):
                if rng.random() < 0.5 and len(s) > 1:
                    del s[pos]
                else:
                    s.insert(int(pos), int(rng.integers(1, 5)))
            s = np.asarray(s, np.uint8)
        if s.tobytes() in seen:
            continue
        seen.add(s.tobytes())
        seqs.append(s)
        quals.append(quality_profile(rng, len(s)))
        origin.append(v)
    reads = abundance(rng, len(seqs))
    origin = np.asarray(origin)

    def centers(inv, r):
        o = origin[np.argsort(inv)]  # origin in sorted order
        best = {}
        for i in range(len(o)):  # sorted by reads desc: first hit is the most abundant
            best.setdefault(int(o[i]), i)
        return sorted(best.values())

    return _finish("batched", seqs, quals, reads, err, centers, extra={"origin": origin})


def workload_partition(n_true=8, n_near=2, n_prior=0, L=150, n_uniques=300,
                       seed=SEED_BASE + 5) -> Workload:
    """Partition-loop workload.  Planted sequences:
      * n_true random variants with 300-3000 reads;
      * n_near variants two substitutions from a true one with 60-150 reads
        (their abundance p-value is far below omega_a: each should bud);
      * n_prior variants two substitutions from a true one with 3 reads
        (no abundance bud; flagged prior, they should bud by omega_p);
    and error copies of the true variants (1-3 random substitutions, 1-4
    reads) up to n_uniques.  extra: "origin" (planted index of each unique's
    source in sorted order, itself for planted ones), "planted" (sorted
    indices of the planted sequences), "prior" (flags, sorted order)."""
    rng = np.random.default_rng(seed)
    err = fixed_err(41)
    seen: set = set()
    truth = _unique_set(rng, lambda: rng.integers(1, 5, size=L).astype(np.uint8), n_true, seen)

    def mutant(src, k):
        def make():
            s = src.copy()
            pos = rng.choice(L, size=k, replace=False)
            s[pos] = ((s[pos] - 1 + rng.integers(1, 4, size=k)) % 4) + 1
            return s
        return make

    seqs, reads, origin, prior = [], [], [], []
    for t in truth:
        seqs.append(t)
        reads.append(int(rng.integers(300, 3001)))
        origin.append(len(seqs) - 1)
        prior.append(0)
    for k in range(n_near + n_prior):
        src = int(rng.integers(0, n_true))
        (s,) = _unique_set(rng, mutant(truth[src], 2), 1, seen)
        seqs.append(s)
        reads.append(int(rng.integers(60, 151)) if k < n_near else 3)
        origin.append(len(seqs) - 1)
        prior.append(0 if k < n_near else 1)
    while len(seqs) < n_uniques:
        src = int(rng.integers(0, n_true))
        (s,) = _unique_set(rng, mutant(truth[src], int(rng.integers(1, 4))), 1, seen)
        seqs.append(s)
        reads.append(int(rng.integers(1, 5)))
        origin.append(src)
        prior.append(0)
    quals = [quality_profile(rng, L) for _ in seqs]
    reads = np.asarray(reads, np.uint32)
    order = _derep_order(seqs, reads)
    inv = np.empty_like(order)
    inv[order] = np.arange(len(order))
    n_planted = n_true + n_near + n_prior
    origin_sorted = np.asarray([inv[origin[i]] for i in order], np.int64)
    extra = {"origin": origin_sorted,
             "planted": np.sort(inv[:n_planted]),
             "prior": np.asarray(prior, np.uint8)[order]}
    return _finish("partition", seqs, quals, reads, err,
                   lambda inv_, r: [plain_center(r)], extra=extra)


def subsample(w: Workload, n_uniques: int, n_centers: int | None = None) -> Workload:
    """The first n_uniques uniques (abundance order) and the centers among them."""
    keep = w.centers[w.centers < n_uniques]
    if n_centers is not None:
        keep = keep[:n_centers]
    return Workload(w.name + f"[:{n_uniques}]", w.seqs[:n_uniques], w.quals[:n_uniques],
                    w.reads[:n_uniques], w.err, keep, w.band, w.kmer_size, dict(w.extra))


def _homopolymer_indels(rng, s: np.ndarray, rate: float) -> np.ndarray:
    """At each position with probability `rate`, lengthen the homopolymer run
    there by one base, or shorten it when the run has length >= 2."""
    hits = np.nonzero(rng.random(len(s)) < rate)[0]
    if len(hits) == 0:
        return s
    out = list(s)
    for pos in sorted(hits.tolist(), reverse=True):
        if pos >= len(out):
            continue
        in_run = (pos > 0 and out[pos - 1] == out[pos]) or \
                 (pos + 1 < len(out) and out[pos + 1] == out[pos])
        if in_run and rng.random() < 0.5 and len(out) > 1:
            del out[pos]
        else:
            out.insert(pos, out[pos])
    return np.asarray(out, np.uint8)


def workload_longread(n_variants=100, n_uniques=5000, lmin=1420, lmax=1480, sub=0.001,
                      homo_indel=0.003, seed=SEED_BASE + 4) -> Workload:
    """Long-read (PacBio-shape) compare: uniques are variant prefixes of length
    U[lmin, lmax] with substitutions (sub/base) and homopolymer-length indels
    (homo_indel/base); Phred 0-93 with mean 60, sd 10, and the fixed 16 x 94
    error matrix.  Band 32; the compare runs with homo_gap_p = -1.  Centers:
    the most abundant unique of each variant."""
    rng = np.random.default_rng(seed)
    err = fixed_err(94)
    variants = [rng.integers(1, 5, size=lmax).astype(np.uint8) for _ in range(n_variants)]
    seqs, quals, origin = [], [], []
    seen: set = set()
    while len(seqs) < n_uniques:
        v = int(rng.integers(0, n_variants))
        Lu = int(rng.integers(lmin, lmax + 1))
        s = _substitute(rng, variants[v][:Lu], np.full(Lu, sub))
        s = _homopolymer_indels(rng, s, homo_indel)
        if s.tobytes() in seen:
            continue
        seen.add(s.tobytes())
        seqs.append(s)
        quals.append(np.clip(np.rint(rng.normal(60.0, 10.0, size=len(s))), 0, 93).astype(np.uint8))
        origin.append(v)
    reads = abundance(rng, len(seqs))
    origin = np.asarray(origin)

    def centers(inv, r):
        o = origin[np.argsort(inv)]
        best = {}
        for i in range(len(o)):
            best.setdefault(int(o[i]), i)
        return sorted(best.values())

    return _finish("longread", seqs, quals, reads, err, centers, band=32,
                   extra={"origin": origin, "homo_gap_p": -1})


@dataclass
class PairSet:
    """Aligned (query, center) pairs with their substitution lists."""
    seqs: list[np.ndarray]
    quals: list[np.ndarray]
    subs: list[np.ndarray]       # (query pos << 16 | center base), ascending
    reads: np.ndarray
    center_reads: np.ndarray
    err: np.ndarray

    @property
    def n(self) -> int:
        return len(self.seqs)


def workload_precision(n_pairs=1_000_000, L=250, n_centers=2000, seed=SEED_BASE + 2) -> PairSet:
    """Aligned pairs at L nt for the mixed-precision lambda: centers uniform
    over ACGT; each query is a center copy with substitutions drawn from the
    seeded error matrix (off-diagonal rates log-uniform in [1e-5, 1e-1], rows
    normalised) at the query's per-position quality (the plain workload's
    profile); abundances lognormal(2, 1.5) for queries, lognormal(6, 1.5)
    for centers."""
    rng = np.random.default_rng(seed)
    err = seeded_err(rng, 41)
    centers = rng.integers(1, 5, size=(n_centers, L)).astype(np.uint8)
    cid = rng.integers(0, n_centers, size=n_pairs)
    ref = centers[cid]                                            # (n, L)
    pos = np.arange(L, dtype=np.float64)
    mu = 36.0 - 8.0 * pos / max(L - 1, 1)
    qual = np.clip(np.rint(rng.normal(mu, 3.0, size=(n_pairs, L))), 0, 40).astype(np.uint8)
    # query base given the reference base and quality: rows 4*ref + query
    r0 = (ref.astype(np.int64) - 1) * 4
    cum = np.zeros((n_pairs, L), np.float64)
    u = rng.random((n_pairs, L))
    qry = np.zeros((n_pairs, L), np.uint8)
    for qb in range(4):
        cum += err[r0 + qb, qual]
        pick = (qry == 0) & (u < cum)
        qry[pick] = qb + 1
    qry[qry == 0] = ref[qry == 0]
    seqs = list(qry)
    quals = list(qual)
    rows, cols = np.nonzero(qry != ref)
    recs = (cols.astype(np.uint32) << 16) | ref[rows, cols].astype(np.uint32)
    subs = np.split(recs, np.cumsum(np.bincount(rows, minlength=n_pairs))[:-1])
    reads = abundance(rng, n_pairs)
    center_reads = np.maximum(1, np.rint(rng.lognormal(6.0, 1.5, size=n_pairs))).astype(np.uint32)
    return PairSet(seqs, quals, subs, reads, center_reads, err)


@dataclass
class BimeraSet:
    """ASCII queries, each with its list of more-abundant parents."""
    queries: list[bytes]
    parent_lists: list[list[bytes]]
    chimeric: np.ndarray  # construction truth (not the reference's verdict)
    pool: list[bytes] = field(default_factory=list)  # the parents, ASCII
    parent_idx: list[list[int]] = field(default_factory=list)  # pool index per parent


def workload_bimera(n_pool=200, n_queries=2000, parents_per_query=50, lmin=1420, lmax=1480,
                    div=0.01, sub=0.001, chimera_frac=0.4, seed=SEED_BASE + 5) -> BimeraSet:
    """Bimera parent search at the long-read shape: a pool of related parents
    (an ancestor trimmed to U[lmin, lmax] with `div` substitutions per base);
    queries are two-parent chimeras at a random breakpoint (chimera_frac) or
    pool parents with `sub` substitutions; each query gets parents_per_query
    distinct pool members, including its true parents.  ASCII bytes."""
    rng = np.random.default_rng(seed)
    ascii_ = np.frombuffer(b"ACGT", np.uint8)
    anc = rng.integers(1, 5, size=lmax).astype(np.uint8)
    pool = []
    for _ in range(n_pool):
        L = int(rng.integers(lmin, lmax + 1))
        pool.append(_substitute(rng, anc[:L], np.full(L, div)))
    queries, lists, truth, idx = [], [], [], []
    for _ in range(n_queries):
        chim = rng.random() < chimera_frac
        if chim:
            a, b = (int(x) for x in rng.choice(n_pool, size=2, replace=False))
            m = min(len(pool[a]), len(pool[b]))
            bp = int(rng.integers(int(0.2 * m), int(0.8 * m)))
            q = np.concatenate([pool[a][:bp], pool[b][bp:]])
            must = [a, b]
        else:
            a = int(rng.integers(0, n_pool))
            q = _substitute(rng, pool[a], np.full(len(pool[a]), sub))
            must = [a]
        others = [int(x) for x in rng.permutation(n_pool) if int(x) not in must]
        par = must + others[:parents_per_query - len(must)]
        par = [par[i] for i in rng.permutation(len(par))]
        queries.append(ascii_[q - 1].tobytes())
        lists.append([ascii_[pool[k] - 1].tobytes() for k in par])
        idx.append(par)
        truth.append(chim)
    return BimeraSet(queries, lists, np.asarray(truth), [ascii_[s - 1].tobytes() for s in pool],
                     idx)
