"""ctypes binding of the CPU oracle (oracle/liboracle.so).

TEST INFRASTRUCTURE ONLY: this restatement of the dada2-rs compare path is the
checker for the HIP implementation.  Only tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg may import it.  Each wrapped function cites the
reference file:line it restates in oracle/*.c.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
LIB_PATH = HERE / "liboracle.so"

u8p = C.POINTER(C.c_uint8)
u16p = C.POINTER(C.c_uint16)
u32p = C.POINTER(C.c_uint32)
u64p = C.POINTER(C.c_uint64)
f64p = C.POINTER(C.c_double)


class AlignParams(C.Structure):
    """AlignParams, nwalign.rs:63-90 (NW backend)."""
    _fields_ = [("match_score", C.c_int32), ("mismatch", C.c_int32), ("gap_p", C.c_int32),
                ("homo_gap_p", C.c_int32), ("use_kmers", C.c_int32), ("kmer_size", C.c_int32),
                ("kdist_cutoff", C.c_double), ("band", C.c_int32), ("vectorized", C.c_int32),
                ("gapless", C.c_int32), ("_pad", C.c_int32)]


class VecScores(C.Structure):
    """VectorizedAlignScores, nwalign.rs:96-103."""
    _fields_ = [("match_score", C.c_int16), ("mismatch", C.c_int16), ("gap_p", C.c_int16),
                ("end_gap_p", C.c_int16), ("band", C.c_int32)]


class Raw(C.Structure):
    """Raw, containers.rs:78-108 (fields the compare reads)."""
    _fields_ = [("seq", u8p), ("qual", u8p), ("kmer8", u8p), ("kord", u16p), ("len", C.c_uint32),
                ("reads", C.c_uint32), ("lock", C.c_uint8), ("prior", C.c_uint8),
                ("_pad", C.c_uint8 * 6)]


class Sub(C.Structure):
    """Sub, containers.rs:30-51."""
    _fields_ = [("len0", C.c_uint32), ("nsubs", C.c_uint32), ("map", u16p), ("pos", u16p),
                ("nt0", u8p), ("nt1", u8p), ("q0", u8p), ("q1", u8p)]


class BimeraParams(C.Structure):
    """BimeraAlignParams, chimera.rs:182-196 (NW backend)."""
    _fields_ = [("allow_one_off", C.c_int32), ("min_one_off_par_dist", C.c_int32),
                ("match_score", C.c_int16), ("mismatch", C.c_int16), ("gap_p", C.c_int16),
                ("_pad", C.c_int16), ("max_shift", C.c_int32)]


def build(force: bool = False) -> Path:
    if force or not LIB_PATH.exists():
        subprocess.run(["make", "-s", "-C", str(HERE)], check=True)
    return LIB_PATH


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            build()
        L = C.CDLL(str(LIB_PATH))
        sig = {
            "orc_buf_new": (C.c_void_p, []),
            "orc_buf_free": (None, [C.c_void_p]),
            "orc_buf_al0": (u8p, [C.c_void_p]),
            "orc_buf_al1": (u8p, [C.c_void_p]),
            "orc_buf_alen": (C.c_uint32, [C.c_void_p]),
            "orc_buf_last_method": (C.c_int, [C.c_void_p]),
            "orc_assign_kmer": (None, [u8p, C.c_uint32, C.c_int, u16p]),
            "orc_assign_kmer8": (None, [u8p, C.c_uint32, C.c_int, u8p]),
            "orc_assign_kmer_order": (None, [u8p, C.c_uint32, C.c_int, u16p]),
            "orc_kmer_dist": (C.c_double, [u16p, C.c_uint32, u16p, C.c_uint32, C.c_int]),
            "orc_kmer_dist8": (C.c_double, [u8p, C.c_uint32, u8p, C.c_uint32, C.c_int]),
            "orc_kmer_dist8_sparse": (C.c_double, [u16p, u8p, C.c_uint32, C.c_uint32, u16p, u8p,
                                                   C.c_uint32, C.c_uint32, C.c_int]),
            "orc_kord_dist": (C.c_double, [u16p, C.c_uint32, u16p, C.c_uint32, C.c_int]),
            "orc_align_endsfree": (C.c_int, [u8p, C.c_uint32, u8p, C.c_uint32, C.c_int32,
                                             C.c_int32, C.c_int32, C.c_int32, C.c_void_p]),
            "orc_align_endsfree_homo": (C.c_int, [u8p, C.c_uint32, u8p, C.c_uint32,
                                                  C.POINTER(AlignParams), C.c_void_p]),
            "orc_align_standard": (C.c_int, [u8p, C.c_uint32, u8p, C.c_uint32, C.c_int32,
                                             C.c_int32, C.c_int32, C.c_int32, C.c_void_p]),
            "orc_align_gapless": (C.c_int, [u8p, C.c_uint32, u8p, C.c_uint32, C.c_void_p]),
            "orc_align_vectorized": (C.c_int, [u8p, C.c_uint32, u8p, C.c_uint32,
                                               C.POINTER(VecScores), C.c_void_p]),
            "orc_align_vectorized_full": (C.c_int, [u8p, C.c_uint32, u8p, C.c_uint32,
                                                    C.POINTER(VecScores), C.c_void_p, u64p]),
            "orc_vectorized_cells": (C.c_uint64, [C.c_uint32, C.c_uint32, C.c_int32]),
            "orc_scalar_cells": (C.c_uint64, [C.c_uint32, C.c_uint32, C.c_int32]),
            "orc_vectorized_geometry": (None, [C.c_uint32, C.c_uint32, C.c_int32,
                                               C.POINTER(C.c_int64), C.POINTER(C.c_int64),
                                               C.POINTER(C.c_int64)]),
            "orc_score_alignment": (C.c_int32, [u8p, u8p, C.c_uint32, C.c_int32, C.c_int32,
                                                C.c_int32]),
            "orc_al2subs": (C.c_int, [u8p, u8p, C.c_uint32, C.c_void_p, C.POINTER(Sub)]),
            "orc_raw_align": (C.c_int, [C.POINTER(Raw), C.POINTER(Raw), C.POINTER(AlignParams),
                                        C.c_void_p, f64p, f64p]),
            "orc_sub_new": (C.c_int, [C.POINTER(Raw), C.POINTER(Raw), C.POINTER(AlignParams),
                                      C.c_void_p, C.POINTER(Sub)]),
            "orc_compute_lambda": (C.c_int, [C.POINTER(Raw), C.POINTER(Sub), f64p, C.c_uint32,
                                             C.c_int, f64p]),
            "orc_calc_pA": (C.c_double, [C.c_uint32, C.c_double, C.c_int]),
            "orc_poisson_upper_tail_direct": (C.c_double, [C.c_uint32, C.c_double]),
            "orc_get_pA": (C.c_double, [C.c_uint32, C.c_int, C.c_double, C.c_uint32, C.c_uint32,
                                        C.c_int]),
            "orc_gamma_q": (C.c_double, [C.c_double, C.c_double]),
            "orc_ln_gamma": (C.c_double, [C.c_double]),
            "orc_gamma_lr": (C.c_double, [C.c_double, C.c_double]),
            "orc_compare_center": (C.c_int, [C.POINTER(Raw), C.c_uint32, C.c_uint32, f64p,
                                             C.c_uint32, C.POINTER(AlignParams), C.c_int, C.c_int,
                                             f64p, u32p, u8p, C.c_int, C.c_int, u64p, u64p]),
            "orc_get_lr": (None, [u8p, u8p, C.c_uint32, C.c_int, C.c_uint32, u32p]),
            "orc_get_ham_endsfree": (C.c_uint32, [u8p, u8p, C.c_uint32]),
            "orc_is_bimera": (C.c_int, [u8p, C.c_uint32, C.POINTER(u8p), u32p, C.c_uint32,
                                        C.POINTER(BimeraParams)]),
            "orc_table_bimera2": (C.c_int, [u32p, C.c_uint32, C.c_uint32, u8p, u64p, C.c_double,
                                            C.c_uint32, C.POINTER(BimeraParams), u32p, u32p,
                                            C.c_int]),
            "orc_build_trans_mat": (C.c_int, [C.POINTER(Raw), C.c_uint32,
                                              C.POINTER(C.c_int64), u32p, C.POINTER(AlignParams),
                                              C.c_uint32, C.c_int, u32p]),
        }
        for name, (res, args) in sig.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def _p(a: np.ndarray, t):
    return a.ctypes.data_as(t)


def _u8(a) -> np.ndarray:
    if isinstance(a, (bytes, bytearray, memoryview)):
        return np.frombuffer(bytes(a), dtype=np.uint8).copy()
    return np.ascontiguousarray(np.asarray(a, dtype=np.uint8))


class Buf:
    """AlignBuffers counterpart (nwalign.rs:118-164)."""

    def __init__(self):
        self.h = lib().orc_buf_new()

    def __del__(self):
        if getattr(self, "h", None):
            lib().orc_buf_free(self.h)
            self.h = None

    def alignment(self) -> tuple[bytes, bytes]:
        L = lib()
        n = L.orc_buf_alen(self.h)
        a0 = C.string_at(L.orc_buf_al0(self.h), n)
        a1 = C.string_at(L.orc_buf_al1(self.h), n)
        return a0, a1

    def last_method(self) -> int:
        return lib().orc_buf_last_method(self.h)


def _check(rc: int, what: str):
    if rc < 0:
        raise RuntimeError(f"oracle {what} failed with status {rc}")


def align_vectorized(s1, s2, match=5, mismatch=-4, gap_p=-8, end_gap_p=0, band=16, buf=None,
                     full=False):
    """align_vectorized_with_buf (nwalign.rs:761) or its full-matrix twin."""
    buf = buf or Buf()
    a, b = _u8(s1), _u8(s2)
    sc = VecScores(match, mismatch, gap_p, end_gap_p, band)
    if full:
        stale = C.c_uint64(0)
        rc = lib().orc_align_vectorized_full(_p(a, u8p), len(a), _p(b, u8p), len(b),
                                             C.byref(sc), buf.h, C.byref(stale))
        _check(rc, "align_vectorized_full")
        return (*buf.alignment(), stale.value)
    rc = lib().orc_align_vectorized(_p(a, u8p), len(a), _p(b, u8p), len(b), C.byref(sc), buf.h)
    _check(rc, "align_vectorized")
    return buf.alignment()


def align_endsfree(s1, s2, match=5, mismatch=-4, gap_p=-8, band=16, buf=None):
    """align_endsfree_with_buf (nwalign.rs:336)."""
    buf = buf or Buf()
    a, b = _u8(s1), _u8(s2)
    rc = lib().orc_align_endsfree(_p(a, u8p), len(a), _p(b, u8p), len(b), match, mismatch,
                                  gap_p, band, buf.h)
    _check(rc, "align_endsfree")
    return buf.alignment()


def align_standard(s1, s2, match=5, mismatch=-4, gap_p=-8, band=16, buf=None):
    buf = buf or Buf()
    a, b = _u8(s1), _u8(s2)
    rc = lib().orc_align_standard(_p(a, u8p), len(a), _p(b, u8p), len(b), match, mismatch,
                                  gap_p, band, buf.h)
    _check(rc, "align_standard")
    return buf.alignment()


def make_params(match=5, mismatch=-4, gap_p=-8, homo_gap_p=None, use_kmers=True, kmer_size=5,
                kdist_cutoff=0.42, band=16, vectorized=True, gapless=True) -> AlignParams:
    """Defaults of resolve_dada_params (main.rs:4393-4433)."""
    return AlignParams(match, mismatch, gap_p, gap_p if homo_gap_p is None else homo_gap_p,
                       int(use_kmers), kmer_size, kdist_cutoff, band, int(vectorized),
                       int(gapless), 0)


def align_endsfree_homo(s1, s2, params: AlignParams, buf=None):
    """align_endsfree_homo_with_buf (nwalign.rs:435)."""
    buf = buf or Buf()
    a, b = _u8(s1), _u8(s2)
    rc = lib().orc_align_endsfree_homo(_p(a, u8p), len(a), _p(b, u8p), len(b),
                                       C.byref(params), buf.h)
    _check(rc, "align_endsfree_homo")
    return buf.alignment()


def score_alignment(al0: bytes, al1: bytes, match=5, mismatch=-4, gap_p=-8) -> int:
    a, b = _u8(bytearray(al0)), _u8(bytearray(al1))
    return lib().orc_score_alignment(_p(a, u8p), _p(b, u8p), len(a), match, mismatch, gap_p)


def vectorized_cells(l1: int, l2: int, band: int) -> int:
    return lib().orc_vectorized_cells(l1, l2, band)


def scalar_cells(l1: int, l2: int, band: int) -> int:
    return lib().orc_scalar_cells(l1, l2, band)


def vectorized_geometry(l1: int, l2: int, band: int) -> tuple[int, int, int]:
    a, b, c = C.c_int64(), C.c_int64(), C.c_int64()
    lib().orc_vectorized_geometry(l1, l2, band, C.byref(a), C.byref(b), C.byref(c))
    return a.value, b.value, c.value


def al2subs(al0: bytes, al1: bytes) -> dict:
    """al2subs (nwalign.rs:1169)."""
    buf = Buf()
    a, b = _u8(bytearray(al0)), _u8(bytearray(al1))
    s = Sub()
    _check(lib().orc_al2subs(_p(a, u8p), _p(b, u8p), len(a), buf.h, C.byref(s)), "al2subs")
    return {"len0": s.len0, "map": [s.map[i] for i in range(s.len0)],
            "pos": [s.pos[i] for i in range(s.nsubs)], "nt0": [s.nt0[i] for i in range(s.nsubs)],
            "nt1": [s.nt1[i] for i in range(s.nsubs)]}


def kmer8(seq, k=5) -> np.ndarray:
    a = _u8(seq)
    out = np.zeros(1 << (2 * k), np.uint8)
    lib().orc_assign_kmer8(_p(a, u8p), len(a), k, _p(out, u8p))
    return out


def kmer16(seq, k=5) -> np.ndarray:
    a = _u8(seq)
    out = np.zeros(1 << (2 * k), np.uint16)
    lib().orc_assign_kmer(_p(a, u8p), len(a), k, _p(out, u16p))
    return out


def kord(seq, k=5) -> np.ndarray:
    a = _u8(seq)
    out = np.zeros(max(len(a) - k + 1, 1), np.uint16)
    lib().orc_assign_kmer_order(_p(a, u8p), len(a), k, _p(out, u16p))
    return out


def kmer_dist8(kv1, len1, kv2, len2, k=5) -> float:
    a, b = _u8(kv1), _u8(kv2)
    return lib().orc_kmer_dist8(_p(a, u8p), len1, _p(b, u8p), len2, k)


def kmer_dist(kv1, len1, kv2, len2, k=5) -> float:
    a = np.ascontiguousarray(kv1, np.uint16)
    b = np.ascontiguousarray(kv2, np.uint16)
    return lib().orc_kmer_dist(_p(a, u16p), len1, _p(b, u16p), len2, k)


def kmer_dist8_sparse(kv1, len1, kv2, len2, k=5) -> float:
    a, b = _u8(kv1), _u8(kv2)
    i1 = np.nonzero(a)[0].astype(np.uint16)
    i2 = np.nonzero(b)[0].astype(np.uint16)
    c1 = np.ascontiguousarray(a[i1])
    c2 = np.ascontiguousarray(b[i2])
    return lib().orc_kmer_dist8_sparse(_p(i1, u16p), _p(c1, u8p), len(i1), len1, _p(i2, u16p),
                                       _p(c2, u8p), len(i2), len2, k)


def kord_dist(o1, len1, o2, len2, k=5) -> float:
    a = np.ascontiguousarray(o1, np.uint16)
    b = np.ascontiguousarray(o2, np.uint16)
    return lib().orc_kord_dist(_p(a, u16p), len1, _p(b, u16p), len2, k)


class RawSet:
    """Owns numpy storage for a list of Raws (containers.rs:78-108) with kmer8/kord built
    as raw_assign_kmers does (kmers.rs:213-221)."""

    def __init__(self, seqs, quals=None, reads=None, k=5, with_kmers=True, locks=None):
        n = len(seqs)
        self.seqs = [_u8(s) for s in seqs]
        self.quals = None if quals is None else [_u8(q) for q in quals]
        self.k = k
        self.kmer8s = [kmer8(s, k) for s in self.seqs] if with_kmers else None
        self.kords = [kord(s, k) for s in self.seqs] if with_kmers else None
        self.arr = (Raw * n)()
        for i in range(n):
            r = self.arr[i]
            r.seq = _p(self.seqs[i], u8p)
            r.len = len(self.seqs[i])
            r.qual = _p(self.quals[i], u8p) if self.quals is not None else u8p()
            r.kmer8 = _p(self.kmer8s[i], u8p) if with_kmers else u8p()
            r.kord = _p(self.kords[i], u16p) if with_kmers else u16p()
            r.reads = int(reads[i]) if reads is not None else 1
            r.lock = int(locks[i]) if locks is not None else 0
            r.prior = 0

    def __len__(self):
        return len(self.seqs)


def sub_new(rs: RawSet, i0: int, i1: int, params: AlignParams, buf=None):
    """sub_new_with_buf (nwalign.rs:1429). Returns None when shrouded."""
    buf = buf or Buf()
    s = Sub()
    rc = lib().orc_sub_new(C.byref(rs.arr[i0]), C.byref(rs.arr[i1]), C.byref(params), buf.h,
                           C.byref(s))
    _check(rc, "sub_new")
    if rc == 0:
        return None
    return s, buf


def raw_align(rs: RawSet, i1: int, i2: int, params: AlignParams, buf=None):
    """raw_align_with_buf (nwalign.rs:1244): (aligned?, kdist, kodist, alignment, method)."""
    buf = buf or Buf()
    kd, ko = C.c_double(), C.c_double()
    rc = lib().orc_raw_align(C.byref(rs.arr[i1]), C.byref(rs.arr[i2]), C.byref(params), buf.h,
                             C.byref(kd), C.byref(ko))
    _check(rc, "raw_align")
    if rc == 0:
        return False, kd.value, ko.value, None, None
    return True, kd.value, ko.value, buf.alignment(), buf.last_method()


def compute_lambda(rs: RawSet, idx: int, sub, err: np.ndarray, ncol: int,
                   use_quals=True) -> float:
    out = C.c_double()
    e = np.ascontiguousarray(err, np.float64)
    sp = C.byref(sub) if sub is not None else C.POINTER(Sub)()
    rc = lib().orc_compute_lambda(C.byref(rs.arr[idx]), sp, _p(e, f64p), ncol, int(use_quals),
                                  C.byref(out))
    _check(rc, "compute_lambda")
    return out.value


def compare_center(rs: RawSet, center: int, err: np.ndarray, ncol: int, params: AlignParams,
                   use_quals=True, greedy=True, nthreads=1, grain=32):
    """The b_compare_parallel map (cluster.rs:279-330): per-raw (lambda, hamming, skipped)."""
    n = len(rs)
    lam = np.zeros(n, np.float64)
    ham = np.zeros(n, np.uint32)
    skp = np.zeros(n, np.uint8)
    cells = C.c_uint64()
    aligned = C.c_uint64()
    e = np.ascontiguousarray(err, np.float64)
    rc = lib().orc_compare_center(rs.arr, n, center, _p(e, f64p), ncol, C.byref(params),
                                  int(use_quals), int(greedy), _p(lam, f64p), _p(ham, u32p),
                                  _p(skp, u8p), nthreads, grain, C.byref(cells), C.byref(aligned))
    _check(rc, "compare_center")
    return lam, ham, skp, cells.value, aligned.value


def calc_pA(reads: int, e_reads: float, prior: bool) -> float:
    return lib().orc_calc_pA(reads, e_reads, int(prior))


def get_pA(reads: int, lam: float, hamming: int, bi_reads: int, prior=False,
           detect_singletons=False) -> float:
    """get_pA, pval.rs:376-398."""
    return lib().orc_get_pA(reads, int(prior), lam, hamming, bi_reads, int(detect_singletons))


def pairs_lambda_fp64(seqs, quals, subs, err: np.ndarray, use_quals=True) -> np.ndarray:
    """compute_lambda (pval.rs:253-320) for equal-length aligned pairs, as the
    ordered FP64 left fold from 1.0 over query positions (pval.rs:311-313),
    vectorised over pairs: t = 5 (nt-1) on the diagonal, 4 (nt0-1) + (nt-1)
    at a substitution record (query pos << 16 | center base)."""
    q = np.stack([np.asarray(s, np.int64) for s in seqs])
    n, L = q.shape
    t = 5 * (q - 1)
    for i, rec in enumerate(subs):
        rec = np.asarray(rec, np.int64)
        if rec.size:
            pos, nt0 = rec >> 16, rec & 0xFFFF
            t[i, pos] = 4 * (nt0 - 1) + (q[i, pos] - 1)
    qi = np.stack([np.asarray(x, np.int64) for x in quals]) if (use_quals and quals is not None) \
        else np.zeros((n, L), np.int64)
    lam = np.ones(n, np.float64)
    for pos in range(L):
        lam *= err[t[:, pos], qi[:, pos]]
    return lam


def poisson_upper_tail_direct(reads: int, lam: float) -> float:
    return lib().orc_poisson_upper_tail_direct(reads, lam)


def gamma_q(a: float, x: float) -> float:
    return lib().orc_gamma_q(a, x)


def ln_gamma(x: float) -> float:
    """statrs 0.18 ln_gamma (Lanczos), restated."""
    return lib().orc_ln_gamma(x)


def gamma_lr(a: float, x: float) -> float:
    """statrs 0.18 gamma_lr (Cephes igam with the exp(ax) cut-off), restated."""
    return lib().orc_gamma_lr(a, x)


def bimera_params(allow_one_off=False, min_one_off_par_dist=4, match=5, mismatch=-4, gap_p=-8,
                  max_shift=16) -> BimeraParams:
    return BimeraParams(int(allow_one_off), min_one_off_par_dist, match, mismatch, gap_p, 0,
                        max_shift)


def get_lr(al0: bytes, al1: bytes, allow_one_off: bool, max_shift: int):
    a, b = _u8(bytearray(al0)), _u8(bytearray(al1))
    out = np.zeros(4, np.uint32)
    lib().orc_get_lr(_p(a, u8p), _p(b, u8p), len(a), int(allow_one_off), max_shift,
                     _p(out, u32p))
    return tuple(int(x) for x in out)


def get_ham_endsfree(al0: bytes, al1: bytes) -> int:
    a, b = _u8(bytearray(al0)), _u8(bytearray(al1))
    return lib().orc_get_ham_endsfree(_p(a, u8p), _p(b, u8p), len(a))


def is_bimera(sq: bytes, parents: list[bytes], params: BimeraParams) -> bool:
    q = _u8(bytearray(sq))
    ps = [_u8(bytearray(p)) for p in parents]
    arr = (u8p * max(len(ps), 1))(*[_p(p, u8p) for p in ps])
    lens = np.array([len(p) for p in ps], np.uint32)
    rc = lib().orc_is_bimera(_p(q, u8p), len(q), arr, _p(lens, u32p), len(ps), C.byref(params))
    _check(rc, "is_bimera")
    return bool(rc)


def table_bimera2(mat: np.ndarray, seqs: list[bytes], min_fold: float, min_abund: int,
                  params: BimeraParams, nthreads: int = 1):
    """table_bimera2 (chimera.rs:317). mat is (nrow samples, ncol sequences)."""
    nrow, ncol = mat.shape
    colmajor = np.ascontiguousarray(mat.T.reshape(-1), np.uint32)  # mat[i + j*nrow]
    cat = _u8(bytearray(b"".join(seqs)))
    off = np.zeros(ncol + 1, np.uint64)
    off[1:] = np.cumsum([len(s) for s in seqs])
    nflag = np.zeros(ncol, np.uint32)
    nsam = np.zeros(ncol, np.uint32)
    rc = lib().orc_table_bimera2(_p(colmajor, u32p), nrow, ncol, _p(cat, u8p), _p(off, u64p),
                                 min_fold, min_abund, C.byref(params), _p(nflag, u32p),
                                 _p(nsam, u32p), nthreads)
    _check(rc, "table_bimera2")
    return nflag, nsam


def encode(s: str) -> bytes:
    """The reference tests' encoder (nwalign.rs:1857-1868): ACGT -> 1..4, else 5."""
    m = {"A": 1, "C": 2, "G": 3, "T": 4, "a": 1, "c": 2, "g": 3, "t": 4}
    return bytes(m.get(ch, 5) for ch in s)


def cpu_count_allowed() -> int:
    try:
        return len(os.sched_getaffinity(0))
    except AttributeError:  # pragma: no cover
        return os.cpu_count() or 1
