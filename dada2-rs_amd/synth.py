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


def subsample(w: Workload, n_uniques: int, n_centers: int | None = None) -> Workload:
    """The first n_uniques uniques (abundance order) and the centers among them."""
    keep = w.centers[w.centers < n_uniques]
    if n_centers is not None:
        keep = keep[:n_centers]
    return Workload(w.name + f"[:{n_uniques}]", w.seqs[:n_uniques], w.quals[:n_uniques],
                    w.reads[:n_uniques], w.err, keep, w.band, w.kmer_size, dict(w.extra))
