"""dada2-rs compare hot path on MI355X (gfx950): ctypes binding of the C ABI.

The library (lib/libdada2_amd.so, built from csrc/*.hip) is the product; this
module only marshals numpy arrays through the C ABI declared in
include/dada2_amd.h.  There is no CPU fallback: every call runs the HIP
kernels, and loading fails loudly when the library is missing.

Names follow the reference (dada2-rs v0.2.0): AlignParams (nwalign.rs:63-90),
VectorizedAlignScores (nwalign.rs:96-103), the b_compare_parallel map
(cluster.rs:279-330) and align_vectorized_with_buf (nwalign.rs:761).
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass
from pathlib import Path

import numpy as np

PKG_DIR = Path(__file__).resolve().parent
LIB_PATH = PKG_DIR / "lib" / "libdada2_amd.so"

D2G_OK = 0
HAMMING_NONE = 0xFFFFFFFF
NO_CENTER = 0xFFFFFFFF  # d2g_transition_counts: a unique outside every cluster

u8p = C.POINTER(C.c_uint8)
u32p = C.POINTER(C.c_uint32)
u64p = C.POINTER(C.c_uint64)
f64p = C.POINTER(C.c_double)


class AlignParams(C.Structure):
    """AlignParams (nwalign.rs:63-90), NW backend."""
    _fields_ = [("match_score", C.c_int32), ("mismatch", C.c_int32), ("gap_p", C.c_int32),
                ("homo_gap_p", C.c_int32), ("use_kmers", C.c_int32), ("kmer_size", C.c_int32),
                ("kdist_cutoff", C.c_double), ("band", C.c_int32), ("vectorized", C.c_int32),
                ("gapless", C.c_int32), ("_reserved", C.c_int32)]


class VecScores(C.Structure):
    """VectorizedAlignScores (nwalign.rs:96-103)."""
    _fields_ = [("match_score", C.c_int16), ("mismatch", C.c_int16), ("gap_p", C.c_int16),
                ("end_gap_p", C.c_int16), ("band", C.c_int32)]


class Stats(C.Structure):
    _fields_ = [("comparisons", C.c_uint64), ("skipped", C.c_uint64), ("shrouded", C.c_uint64),
                ("gapless", C.c_uint64), ("aligned_dp", C.c_uint64), ("dp_cells", C.c_uint64),
                ("ms_screen", C.c_double), ("ms_plan", C.c_double), ("ms_dp", C.c_double),
                ("ms_lambda", C.c_double), ("ms_total", C.c_double)]

    def as_dict(self) -> dict:
        return {name: getattr(self, name) for name, _ in self._fields_}


def align_params(match=5, mismatch=-4, gap_p=-8, homo_gap_p=None, use_kmers=True, kmer_size=5,
                 kdist_cutoff=0.42, band=16, vectorized=True, gapless=True) -> AlignParams:
    """Defaults of resolve_dada_params (main.rs:4393-4433), which turns a
    positive gap_p negative before homo_gap_p defaults to it, and a positive
    homo_gap_p negative too (main.rs:4403-4418)."""
    gap_p = -gap_p if gap_p > 0 else gap_p
    homo = gap_p if homo_gap_p is None else homo_gap_p
    homo = -homo if homo > 0 else homo
    return AlignParams(match, mismatch, gap_p, homo, int(use_kmers), kmer_size, kdist_cutoff,
                       band, int(vectorized), int(gapless), 0)


def extend_err(err, qmax: int) -> np.ndarray:
    """The error-model extension of dada_uniques_cached (dada.rs:337-384):
    when the data's highest rounded quality qmax is at or above the model's
    column count, the 16 x ncol matrix grows to qmax + 1 columns, the new
    ones repeating the last learned column, with the reference's warning;
    otherwise it is returned as is."""
    e = np.ascontiguousarray(err, np.float64)
    old = e.shape[1]
    if old == 0 or qmax < old:
        return e
    ext = np.concatenate([e, np.repeat(e[:, old - 1:old], qmax + 1 - old, axis=1)], axis=1)
    import warnings
    warnings.warn(f"input has quality up to Q{qmax} but the error model covers Q0-Q{old - 1}. "
                  f"Extending the model by repeating its last column (Q{old - 1}) for "
                  f"Q{old}-Q{qmax}. Rates for the extended columns are extrapolated, not learned; "
                  "re-run learn-errors on this data to avoid it.", stacklevel=3)
    return np.ascontiguousarray(ext)


class D2GError(RuntimeError):
    def __init__(self, status: int, what: str):
        self.status = status
        msg = _lib().d2g_status_string(status).decode() if _LIB is not None else str(status)
        super().__init__(f"{what}: status {status} ({msg})")


_LIB = None


def _lib():
    global _LIB
    if _LIB is None:
        if not LIB_PATH.exists():
            raise FileNotFoundError(
                f"{LIB_PATH} is missing: build it first (__graft_entry__.build())")
        L = C.CDLL(str(LIB_PATH))
        vp = C.c_void_p
        sig = {
            "d2g_ctx_create": (C.c_int, [C.c_int, C.POINTER(vp)]),
            "d2g_ctx_destroy": (C.c_int, [vp]),
            "d2g_ctx_set_timing": (C.c_int, [vp, C.c_int]),
            "d2g_ctx_stream": (vp, [vp]),
            "d2g_device_info": (C.c_int, [vp, C.POINTER(C.c_int), C.POINTER(C.c_int),
                                          C.POINTER(C.c_int), C.POINTER(C.c_int)]),
            "d2g_status_string": (C.c_char_p, [C.c_int]),
            "d2g_load_uniques": (C.c_int, [vp, C.c_uint32, u8p, u64p, u8p, u32p, C.c_uint32]),
            "d2g_compare_center": (C.c_int, [vp, C.c_uint32, f64p, C.c_uint32,
                                             C.POINTER(AlignParams), C.c_int, C.c_int, u8p, f64p,
                                             u32p, u8p, C.POINTER(Stats)]),
            "d2g_compare_batch": (C.c_int, [vp, u32p, C.c_uint32, f64p, C.c_uint32,
                                            C.POINTER(AlignParams), C.c_int, C.c_int, u8p, f64p,
                                            u32p, u8p, C.POINTER(Stats)]),
            "d2g_compare_batch_device": (C.c_int, [vp, u32p, C.c_uint32, f64p, C.c_uint32,
                                                   C.POINTER(AlignParams), C.c_int, C.c_int, u8p,
                                                   C.POINTER(Stats)]),
            "d2g_fetch_results": (C.c_int, [vp, f64p, u32p, u8p]),
            "d2g_store_reset": (C.c_int, [vp]),
            "d2g_compare_store": (C.c_int, [vp, u32p, C.c_uint32, C.c_uint32, C.c_double, f64p,
                                            C.c_uint32, C.POINTER(AlignParams), C.c_int, C.c_int,
                                            u8p, C.c_uint64, C.POINTER(Comparison), u64p, f64p,
                                            C.POINTER(Stats)]),
            "d2g_store_size": (C.c_int, [vp, u64p]),
            "d2g_store_fetch": (C.c_int, [vp, C.c_uint64, C.c_uint64, C.POINTER(Comparison)]),
            "d2g_align_vectorized": (C.c_int, [vp, C.c_uint32, u8p, u64p, u8p, u64p,
                                               C.POINTER(VecScores), u8p, u8p, u64p, u32p]),
            "d2g_vectorized_cells": (C.c_uint64, [C.c_uint32, C.c_uint32, C.c_int32]),
            "d2g_vectorized_supported": (C.c_int, [C.c_uint32, C.c_uint32, C.c_int32]),
            "d2g_align_endsfree": (C.c_int, [vp, C.c_uint32, u8p, u64p, u8p, u64p, C.c_int32,
                                             C.c_int32, C.c_int32, C.c_int32, C.c_int32, u8p,
                                             u8p, u64p, u32p]),
            "d2g_band_supported": (C.c_int, [C.c_uint32, C.c_uint32, C.c_int32, C.c_int]),
            "d2g_pairs_load": (C.c_int, [vp, C.c_uint32, u8p, u64p, u8p, u32p, u64p, u32p, u32p]),
            "d2g_pairs_lambda": (C.c_int, [vp, f64p, C.c_uint32, C.c_int, C.c_double,
                                           C.POINTER(PairStats)]),
            "d2g_pairs_fetch": (C.c_int, [vp, C.POINTER(C.c_float), f64p, f64p, u8p]),
            "d2g_pairs_lambda_noalloc": (C.c_int, [vp, C.POINTER(PairsDev), f64p, C.c_uint32,
                                                   C.c_int, C.c_double, vp, vp, vp, vp,
                                                   C.POINTER(PairStats)]),
            "d2g_bimera_pairs": (C.c_int, [vp, C.c_uint32, u8p, u64p, u8p, u64p,
                                           C.POINTER(BimeraParams), u32p, u32p]),
            "d2g_is_bimera_batch": (C.c_int, [vp, C.c_uint32, u8p, u64p, u8p, u64p, u32p,
                                              C.POINTER(BimeraParams), u8p]),
            "d2g_is_bimera_pool": (C.c_int, [vp, C.c_uint32, u8p, u64p, C.c_uint32, u8p, u64p,
                                             u32p, u32p, C.POINTER(BimeraParams), u8p]),
            "d2g_table_bimera2": (C.c_int, [vp, u32p, C.c_uint32, C.c_uint32, u8p, u64p,
                                            C.c_double, C.c_uint32, C.POINTER(BimeraParams),
                                            u32p, u32p]),
            "d2g_loglik_forward": (C.c_int, [vp, C.POINTER(C.c_float), C.c_uint32, C.c_int,
                                             f64p, C.POINTER(C.c_float)]),
            "d2g_loglik_backward": (C.c_int, [vp, C.POINTER(C.c_float), f64p,
                                              C.POINTER(C.c_float)]),
            "d2g_ctx_set_graphs": (C.c_int, [vp, C.c_int]),
            "d2g_ctx_set_screen_shape": (C.c_int, [vp, C.c_int, C.c_int, C.c_int]),
            "d2g_ctx_set_autotune": (C.c_int, [vp, C.c_int]),
            "d2g_ctx_set_dp_tile": (C.c_int, [vp, C.c_int]),
            "d2g_run_dada": (C.c_int, [vp, f64p, C.c_uint32, C.POINTER(AlignParams),
                                       C.POINTER(DadaParams), u8p, C.c_uint32,
                                       C.POINTER(C.c_uint32), C.POINTER(ClusterInfo), u32p, f64p,
                                       C.POINTER(C.c_int32), f64p, u32p, C.POINTER(DadaStats)]),
            "d2g_transition_counts": (C.c_int, [vp, u32p, u32p, C.POINTER(AlignParams), C.c_int,
                                                C.c_uint32, u32p]),
            "d2g_build_trans_mat": (C.c_int, [vp, u32p, u32p, C.POINTER(AlignParams), C.c_uint32,
                                              u32p]),
            "d2g_transition_stats": (C.c_int, [vp, C.POINTER(Stats)]),
            "d2g_dada_aux": (C.c_int, [vp, C.c_uint32, u32p, u32p, u32p, C.POINTER(AlignParams),
                                       C.c_int, C.c_uint32, C.c_uint32, u32p, u32p, f64p]),
            "d2g_birth_subs": (C.c_int, [vp, C.c_uint32, u32p, u32p, C.POINTER(AlignParams),
                                         C.c_int, C.c_uint32, u32p, f64p, u32p, u32p, u8p, u8p,
                                         u8p, u32p]),
        }
        for name, (res, args) in sig.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _LIB = L
    return _LIB


def exported_symbols() -> list[str]:
    """Names of the C ABI entry points (include/dada2_amd.h)."""
    return ["d2g_ctx_create", "d2g_ctx_destroy", "d2g_ctx_set_timing", "d2g_ctx_stream",
            "d2g_device_info", "d2g_status_string", "d2g_load_uniques", "d2g_compare_center",
            "d2g_compare_batch", "d2g_compare_batch_device", "d2g_fetch_results",
            "d2g_align_vectorized", "d2g_vectorized_cells", "d2g_vectorized_supported",
            "d2g_align_endsfree", "d2g_band_supported", "d2g_pairs_load", "d2g_pairs_lambda",
            "d2g_pairs_fetch", "d2g_bimera_pairs", "d2g_is_bimera_batch", "d2g_is_bimera_pool", "d2g_table_bimera2",
            "d2g_loglik_forward", "d2g_loglik_backward", "d2g_ctx_set_graphs",
            "d2g_ctx_set_screen_shape", "d2g_ctx_set_autotune", "d2g_run_dada",
            "d2g_transition_counts", "d2g_build_trans_mat", "d2g_transition_stats",
            "d2g_dada_aux", "d2g_birth_subs", "d2g_pairs_lambda_noalloc", "d2g_store_reset",
            "d2g_compare_store", "d2g_store_size", "d2g_store_fetch", "d2g_ctx_set_dp_tile"]


def _p(a: np.ndarray, t):
    return a.ctypes.data_as(t)


def _check(rc: int, what: str):
    if rc != D2G_OK:
        raise D2GError(rc, what)


def vectorized_cells(l1: int, l2: int, band: int) -> int:
    return int(_lib().d2g_vectorized_cells(l1, l2, band))


def vectorized_supported(l1: int, l2: int, band: int) -> bool:
    return bool(_lib().d2g_vectorized_supported(l1, l2, band))


def band_supported(l1: int, l2: int, band: int, homo: bool = False) -> bool:
    return bool(_lib().d2g_band_supported(l1, l2, band, int(homo)))


def _concat(seqs) -> tuple[np.ndarray, np.ndarray]:
    arrs = [np.frombuffer(bytes(s), np.uint8) if isinstance(s, (bytes, bytearray))
            else np.asarray(s, np.uint8) for s in seqs]
    off = np.zeros(len(arrs) + 1, np.uint64)
    off[1:] = np.cumsum([len(a) for a in arrs])
    cat = np.concatenate(arrs) if arrs else np.zeros(0, np.uint8)
    return np.ascontiguousarray(cat, np.uint8), off


class BimeraParams(C.Structure):
    """BimeraAlignParams (chimera.rs:183-196), NW backend."""
    _fields_ = [("allow_one_off", C.c_int32), ("min_one_off_par_dist", C.c_int32),
                ("match_score", C.c_int16), ("mismatch", C.c_int16), ("gap_p", C.c_int16),
                ("_pad", C.c_int16), ("max_shift", C.c_int32)]


def bimera_params(allow_one_off=False, min_one_off_par_dist=4, match=5, mismatch=-4, gap_p=-8,
                  max_shift=16) -> BimeraParams:
    return BimeraParams(int(allow_one_off), min_one_off_par_dist, match, mismatch, gap_p, 0,
                        max_shift)


class PairStats(C.Structure):
    _fields_ = [("pairs", C.c_uint64), ("recomputed", C.c_uint64), ("ms_loglam", C.c_float),
                ("ms_pval", C.c_float), ("ms_total", C.c_float), ("_pad", C.c_float)]

    def as_dict(self) -> dict:
        return {name: getattr(self, name) for name, _ in self._fields_ if name != "_pad"}


class Comparison(C.Structure):
    """d2g_comparison: one kept comparison of the store (Comparison,
    containers.rs:30-51)."""
    _fields_ = [("index", C.c_uint32), ("cluster", C.c_uint32), ("hamming", C.c_uint32),
                ("_pad", C.c_uint32), ("lambda", C.c_double)]


COMPARISON_DTYPE = np.dtype([("index", "<u4"), ("cluster", "<u4"), ("hamming", "<u4"),
                             ("_pad", "<u4"), ("lambda", "<f8")])


class PairsDev(C.Structure):
    """d2g_pairs_dev: a pair set in device memory (d2g_pairs_load's layout)."""
    _fields_ = [("n", C.c_uint32), ("_pad", C.c_uint32), ("pool_bytes", C.c_uint64),
                ("n_subs", C.c_uint64), ("seq", C.c_void_p), ("qual", C.c_void_p),
                ("off", C.c_void_p), ("len", C.c_void_p), ("subs", C.c_void_p),
                ("sub_off", C.c_void_p), ("reads", C.c_void_p), ("center_reads", C.c_void_p)]


NOALLOC_MAX_PAIRS = 1 << 20
NOALLOC_MAX_READS = (1 << 20) - 1
NOALLOC_MAX_NCOL = 256


def pairs_layout(seqs, quals, subs, reads, center_reads) -> dict:
    """Host arrays of a pair set in d2g_pairs_load's resident layout (rows at
    16-byte offsets, padded), ready to copy to device memory for
    d2g_pairs_lambda_noalloc: seq, qual (or None), off (n + 1), len, subs,
    sub_off (n + 1), reads, center_reads."""
    n = len(seqs)
    lens = np.array([len(x) for x in seqs], np.uint32)
    off = np.zeros(n + 1, np.uint64)
    off[1:] = np.cumsum((lens.astype(np.uint64) + 15) // 16 * 16)
    total = int(off[-1]) + 16
    seq = np.ones(total, np.uint8)
    qual = None if quals is None else np.zeros(total, np.uint8)
    for i in range(n):
        o = int(off[i])
        seq[o:o + lens[i]] = np.asarray(seqs[i], np.uint8)
        if qual is not None:
            qual[o:o + lens[i]] = np.asarray(quals[i], np.uint8)
    so = np.zeros(n + 1, np.uint64)
    so[1:] = np.cumsum([len(x) for x in subs])
    sc = (np.concatenate([np.asarray(x, np.uint32) for x in subs]) if int(so[-1])
          else np.zeros(1, np.uint32))
    return {"seq": seq, "qual": qual, "off": off, "len": lens,
            "subs": np.ascontiguousarray(sc, np.uint32), "sub_off": so,
            "n_subs": int(so[-1]), "reads": np.ascontiguousarray(reads, np.uint32),
            "center_reads": np.ascontiguousarray(center_reads, np.uint32)}


class DadaParams(C.Structure):
    """DadaParams fields of the partition loop (dada.rs:54-90)."""
    _fields_ = [("omega_a", C.c_double), ("omega_p", C.c_double), ("omega_c", C.c_double),
                ("min_fold", C.c_double), ("min_hamming", C.c_uint32), ("min_abund", C.c_uint32),
                ("max_clust", C.c_uint32), ("max_shuffle", C.c_uint32), ("greedy", C.c_int32),
                ("detect_singletons", C.c_int32), ("use_quals", C.c_int32),
                ("device_shuffle", C.c_int32)]


def dada_params(omega_a=1e-40, omega_p=1e-4, omega_c=1e-40, min_fold=1.0, min_hamming=1,
                min_abund=1, max_clust=0, max_shuffle=10, greedy=True, detect_singletons=False,
                use_quals=True, device_shuffle=True) -> DadaParams:
    """Defaults of resolve_dada_params (main.rs:4393-4433); MAX_SHUFFLE = 10
    (dada.rs:39).  device_shuffle runs the shuffle's best-cluster scan on
    the device (the same moves in the same order; off: the host's
    threads)."""
    return DadaParams(omega_a, omega_p, omega_c, min_fold, min_hamming, min_abund, max_clust,
                      max_shuffle, int(greedy), int(detect_singletons), int(use_quals),
                      int(device_shuffle))


class ClusterInfo(C.Structure):
    _fields_ = [("center", C.c_uint32), ("reads", C.c_uint32), ("nraw", C.c_uint32),
                ("birth_type", C.c_int32), ("birth_from", C.c_uint32),
                ("birth_hamming", C.c_uint32), ("birth_pval", C.c_double),
                ("birth_fold", C.c_double), ("birth_e", C.c_double),
                ("birth_lambda", C.c_double), ("self_lambda", C.c_double),
                ("posthoc_pval", C.c_double)]


class DadaStats(C.Structure):
    _fields_ = [("nalign", C.c_uint64), ("nshroud", C.c_uint64), ("compares", C.c_uint32),
                ("shuffle_passes", C.c_uint32), ("ms_compare", C.c_double),
                ("ms_store", C.c_double), ("ms_shuffle", C.c_double), ("ms_bud", C.c_double),
                ("ms_pupdate", C.c_double), ("ms_total", C.c_double), ("dp_cells", C.c_uint64),
                ("ms_dev_screen", C.c_double), ("ms_dev_plan", C.c_double),
                ("ms_dev_dp", C.c_double), ("ms_dev_lambda", C.c_double)]

    def as_dict(self):
        return {name: getattr(self, name) for name, _ in self._fields_}


@dataclass
class DadaResult:
    """The partition after run_dada and the final p pass (dada.rs:502-533):
    per cluster (center, reads, nraw, birth_*, self_lambda) and per unique
    (cluster_of, p, map with -1 for uncorrected, comp_lambda, comp_hamming)."""
    clusters: list
    cluster_of: np.ndarray
    p: np.ndarray
    map: np.ndarray
    comp_lambda: np.ndarray
    comp_hamming: np.ndarray
    stats: dict
    err_ncol: int = 41  # columns of the (possibly extended) error matrix of the run


@dataclass
class CompareResult:
    """Per-(center, unique) output of the compare map (cluster.rs:279-330)."""
    lam: np.ndarray
    hamming: np.ndarray
    skipped: np.ndarray
    stats: dict


class Context:
    """One device context: a HIP stream plus the resident uniques."""

    def __init__(self, device: int = 0, timing: bool = False):
        L = _lib()
        h = C.c_void_p()
        _check(L.d2g_ctx_create(device, C.byref(h)), "d2g_ctx_create")
        self.h = h
        self.n = 0
        self.has_quals = False
        if timing:
            _check(L.d2g_ctx_set_timing(self.h, 1), "d2g_ctx_set_timing")

    def close(self):
        if getattr(self, "h", None):
            _lib().d2g_ctx_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def set_timing(self, on: bool):
        _check(_lib().d2g_ctx_set_timing(self.h, int(on)), "d2g_ctx_set_timing")

    def set_graphs(self, on: bool):
        """Replay the compare's launches as a captured HIP graph (default on)."""
        _check(_lib().d2g_ctx_set_graphs(self.h, int(on)), "d2g_ctx_set_graphs")

    def set_autotune(self, on: bool):
        """Time the screen's launch shapes on the first compare of each
        workload class and keep the fastest (default off)."""
        _check(_lib().d2g_ctx_set_autotune(self.h, int(on)), "d2g_ctx_set_autotune")

    def set_dp_tile(self, tile: int = -1):
        """Fix the DP tile (vectorized: pairs per wavefront cap, 0 = all that
        fit; row band: least cells per lane 1, 2, 4); -1 leaves it to the
        autotuner."""
        _check(_lib().d2g_ctx_set_dp_tile(self.h, int(tile)), "d2g_ctx_set_dp_tile")

    def set_screen_shape(self, qblock: int = 64, split: int = 4, per_lane: int = 1):
        """Screen launch shape: uniques per workgroup (64, 128, 256), the
        wavefront groups (1, 2, 4) sharing its 16 centers, and uniques per
        lane (1, 2)."""
        _check(_lib().d2g_ctx_set_screen_shape(self.h, int(qblock), int(split), int(per_lane)),
               "d2g_ctx_set_screen_shape")

    def device_info(self) -> dict:
        a, b, c, d = C.c_int(), C.c_int(), C.c_int(), C.c_int()
        _check(_lib().d2g_device_info(self.h, C.byref(a), C.byref(b), C.byref(c), C.byref(d)),
               "d2g_device_info")
        return {"cu_count": a.value, "clock_khz": b.value, "mem_clock_khz": c.value,
                "mem_bus_width": d.value}

    def load_uniques(self, seqs, quals=None, reads=None, k: int = 5):
        """Resident uniques (dada_uniques_cached's Raw build, dada.rs:243-402)."""
        cat, off = _concat(seqs)
        q = None
        if quals is not None:
            q, qoff = _concat(quals)
            if not np.array_equal(qoff, off):
                raise ValueError("quality lengths must match sequence lengths")
        r = np.ascontiguousarray(np.ones(len(seqs)) if reads is None else reads, np.uint32)
        _check(_lib().d2g_load_uniques(self.h, len(seqs), _p(cat, u8p), _p(off, u64p),
                                       _p(q, u8p) if q is not None else None, _p(r, u32p), k),
               "d2g_load_uniques")
        self.n = len(seqs)
        self.has_quals = q is not None
        self.max_qual = int(q.max()) if q is not None and q.size else 0
        self.reads_total = float(int(r.astype(np.uint64).sum()) & 0xFFFFFFFF)  # B.reads, a u32 sum
        self.reads = r
        self.maxlen = int(np.diff(off).max())

    def _common(self, err, params, lock):
        e = np.ascontiguousarray(err, np.float64)
        if e.ndim != 2 or e.shape[0] != 16:
            raise ValueError("err must be a 16 x ncol matrix")
        lk = None if lock is None else np.ascontiguousarray(lock, np.uint8)
        return e, lk

    def compare_batch(self, centers, err, params: AlignParams | None = None, use_quals=True,
                      greedy=True, lock=None, device_only=False) -> CompareResult | dict:
        """All centers x all uniques.  Rows follow `centers`."""
        params = params or align_params()
        e, lk = self._common(err, params, lock)
        cs = np.ascontiguousarray(np.atleast_1d(centers), np.uint32)
        st = Stats()
        L = _lib()
        if device_only:
            _check(L.d2g_compare_batch_device(self.h, _p(cs, u32p), len(cs), _p(e, f64p),
                                              e.shape[1], C.byref(params), int(use_quals),
                                              int(greedy), _p(lk, u8p) if lk is not None else None,
                                              C.byref(st)), "d2g_compare_batch_device")
            return st.as_dict()
        shape = (len(cs), self.n)
        lam = np.zeros(shape, np.float64)
        ham = np.zeros(shape, np.uint32)
        skp = np.zeros(shape, np.uint8)
        _check(L.d2g_compare_batch(self.h, _p(cs, u32p), len(cs), _p(e, f64p), e.shape[1],
                                   C.byref(params), int(use_quals), int(greedy),
                                   _p(lk, u8p) if lk is not None else None, _p(lam, f64p),
                                   _p(ham, u32p), _p(skp, u8p), C.byref(st)),
               "d2g_compare_batch")
        return CompareResult(lam, ham, skp, st.as_dict())

    def store_reset(self):
        """B::new's store: e_minmax = -999.0 per unique, nothing kept."""
        _check(_lib().d2g_store_reset(self.h), "d2g_store_reset")

    def compare_store(self, centers, err, first_cluster: int = 0,
                      params: AlignParams | None = None, use_quals=True, greedy=True, lock=None,
                      fetch=True, out: np.ndarray | None = None):
        """The compare of `centers` (successive clusters first_cluster, ...)
        and b_compare_parallel's store (cluster.rs:351-397) on the device.
        Returns (kept comparisons as a COMPARISON_DTYPE array in push order,
        or their count without fetch; each center's self lambda; stats).
        `out` (COMPARISON_DTYPE, at least len(centers) * n records) lets a
        caller reuse one landing array across calls; the result is a view of it."""
        params = params or align_params()
        e, lk = self._common(err, params, lock)
        cs = np.ascontiguousarray(np.atleast_1d(centers), np.uint32)
        cap = len(cs) * self.n if fetch else 0
        if out is None or not fetch:
            out = np.zeros(cap, COMPARISON_DTYPE)
        elif out.dtype != COMPARISON_DTYPE or out.ndim != 1 or len(out) < cap \
                or not out.flags.c_contiguous:
            raise ValueError("out must be a contiguous COMPARISON_DTYPE array of "
                             f"at least {cap} records")
        nst = np.zeros(1, np.uint64)
        self_l = np.zeros(len(cs), np.float64)
        st = Stats()
        _check(_lib().d2g_compare_store(
            self.h, _p(cs, u32p), len(cs), int(first_cluster), self.reads_total, _p(e, f64p),
            e.shape[1], C.byref(params), int(use_quals), int(greedy),
            _p(lk, u8p) if lk is not None else None, cap,
            out.ctypes.data_as(C.POINTER(Comparison)) if fetch else None, _p(nst, u64p),
            _p(self_l, f64p), C.byref(st)), "d2g_compare_store")
        k = int(nst[0])
        return (out[:k] if fetch else k), self_l, st.as_dict()

    def store_fetch(self, first: int = 0, count: int | None = None) -> np.ndarray:
        """Records [first, first + count) of the context's store."""
        n = np.zeros(1, np.uint64)
        _check(_lib().d2g_store_size(self.h, _p(n, u64p)), "d2g_store_size")
        count = int(n[0]) - first if count is None else count
        out = np.zeros(max(count, 0), COMPARISON_DTYPE)
        _check(_lib().d2g_store_fetch(self.h, first, count,
                                      out.ctypes.data_as(C.POINTER(Comparison))), "d2g_store_fetch")
        return out

    def fetch_results(self, n_centers: int) -> CompareResult:
        shape = (n_centers, self.n)
        lam = np.zeros(shape, np.float64)
        ham = np.zeros(shape, np.uint32)
        skp = np.zeros(shape, np.uint8)
        _check(_lib().d2g_fetch_results(self.h, _p(lam, f64p), _p(ham, u32p), _p(skp, u8p)),
               "d2g_fetch_results")
        return CompareResult(lam, ham, skp, {})

    def compare_center(self, center: int, err, params: AlignParams | None = None,
                       use_quals=True, greedy=True, lock=None) -> CompareResult:
        """The plain path: one center vs every unique (b_compare_parallel's map)."""
        params = params or align_params()
        e, lk = self._common(err, params, lock)
        lam = np.zeros(self.n, np.float64)
        ham = np.zeros(self.n, np.uint32)
        skp = np.zeros(self.n, np.uint8)
        st = Stats()
        _check(_lib().d2g_compare_center(self.h, center, _p(e, f64p), e.shape[1],
                                         C.byref(params), int(use_quals), int(greedy),
                                         _p(lk, u8p) if lk is not None else None, _p(lam, f64p),
                                         _p(ham, u32p), _p(skp, u8p), C.byref(st)),
               "d2g_compare_center")
        return CompareResult(lam, ham, skp, st.as_dict())

    def run_dada(self, err, params: AlignParams | None = None, dada: DadaParams | None = None,
                 prior=None) -> DadaResult:
        """run_dada (dada.rs:749) over the loaded uniques, then the final p
        pass and read map (dada.rs:502-533).  Every cluster's compare runs on
        the device; the partition state lives on the host."""
        params = params or align_params()
        dada = dada or dada_params()
        e = np.ascontiguousarray(err, np.float64)
        if self.has_quals:
            e = extend_err(e, self.max_qual)
        n = self.n
        pr = None if prior is None else np.ascontiguousarray(prior, np.uint8)
        if pr is not None and pr.shape != (n,):
            raise ValueError("prior needs one flag per unique")
        cap = max(n, 1)
        infos = (ClusterInfo * cap)()
        ncl = C.c_uint32()
        cof = np.zeros(n, np.uint32)
        pv = np.zeros(n, np.float64)
        mp = np.zeros(n, np.int32)
        cl = np.zeros(n, np.float64)
        ch = np.zeros(n, np.uint32)
        st = DadaStats()
        _check(_lib().d2g_run_dada(self.h, _p(e, f64p), e.shape[1], C.byref(params),
                                   C.byref(dada), _p(pr, u8p) if pr is not None else None, cap,
                                   C.byref(ncl), infos, _p(cof, u32p), _p(pv, f64p),
                                   mp.ctypes.data_as(C.POINTER(C.c_int32)), _p(cl, f64p),
                                   _p(ch, u32p), C.byref(st)), "d2g_run_dada")
        clusters = [{name: getattr(infos[i], name) for name, _ in ClusterInfo._fields_}
                    for i in range(ncl.value)]
        return DadaResult(clusters, cof, pv, mp, cl, ch, st.as_dict(), int(e.shape[1]))

    def transition_counts(self, center_of, weight, params: AlignParams | None = None,
                          use_quals: bool = True, ncol: int = 41) -> np.ndarray:
        """transition_counts (error.rs:194-236) over the final alignments
        (compute_final_subs, dada.rs:646-673) of the loaded uniques:
        center_of[i] is the center unique of i's cluster (-1 or NO_CENTER for
        none), weight[i] the reads it adds (0 skips it: Raw.correct false).
        Returns a 16 x ncol u32 matrix (16 x 1 without qualities)."""
        params = params or align_params()
        co, w = self._center_weight(center_of, weight)
        nout = ncol if (use_quals and self.has_quals) else 1
        out = np.zeros(16 * nout, np.uint32)
        _check(_lib().d2g_transition_counts(self.h, _p(co, u32p), _p(w, u32p), C.byref(params),
                                            int(use_quals), ncol, _p(out, u32p)),
               "d2g_transition_counts")
        return out.reshape(16, nout)

    def build_trans_mat(self, center_of, weight, params: AlignParams | None = None,
                        nq: int = 41) -> np.ndarray:
        """build_trans_mat (learn_errors.rs:618-690): center_of as in
        transition_counts, weight[i] the abundance of input i when the
        result maps it (DadaResult.map[i] >= 0), else 0.  Returns 16 x nq."""
        params = params or align_params()
        co, w = self._center_weight(center_of, weight)
        out = np.zeros(16 * nq, np.uint32)
        _check(_lib().d2g_build_trans_mat(self.h, _p(co, u32p), _p(w, u32p), C.byref(params),
                                          nq, _p(out, u32p)), "d2g_build_trans_mat")
        return out.reshape(16, nq)

    def dada_aux(self, result: "DadaResult", params: AlignParams | None = None,
                 use_quals: bool = True, ncol: int | None = None, maxlen: int | None = None) -> dict:
        """compute_aux's per-raw part (dada.rs:610-631) for a run_dada
        result over the loaded uniques: the transitions, each raw's Sub.nsubs
        against its center (-1 when not counted) and cluster_quality
        (clusters x maxlen, NaN where uncovered; maxlen defaults to the
        longest unique, as the reference's)."""
        params = params or align_params()
        ncol = result.err_ncol if ncol is None else ncol
        n = self.n
        ncl = len(result.clusters)
        centers = np.ascontiguousarray([c["center"] for c in result.clusters], np.uint32)
        cof = np.ascontiguousarray(result.cluster_of, np.uint32)
        reads = np.ascontiguousarray(self.reads, np.uint32)
        w = np.ascontiguousarray(np.where(np.asarray(result.map) >= 0, reads, 0), np.uint32)
        maxlen = int(self.maxlen if maxlen is None else maxlen)
        nout = ncol if (use_quals and self.has_quals) else 1
        trans = np.zeros(16 * nout, np.uint32)
        nsubs = np.zeros(n, np.uint32)
        qual = np.zeros(ncl * maxlen, np.float64)
        _check(_lib().d2g_dada_aux(self.h, ncl, _p(centers, u32p), _p(cof, u32p), _p(w, u32p),
                                   C.byref(params), int(use_quals), ncol, maxlen,
                                   _p(trans, u32p), _p(nsubs, u32p), _p(qual, f64p)),
               "d2g_dada_aux")
        return {"transitions": trans.reshape(16, nout),
                "nsubs": np.where(nsubs == NO_CENTER, -1, nsubs.astype(np.int64)),
                "cluster_quality": qual.reshape(ncl, maxlen)}

    def birth_subs(self, result: "DadaResult", params: AlignParams | None = None,
                   use_quals: bool = True):
        """compute_birth_subs (dada.rs:679-693) for a run_dada result:
        (records [(cluster, pos, nt0, nt1, qual or None)], nsubs per cluster
        (-1 without a Sub), birth_qave per cluster (NaN when None))."""
        params = params or align_params()
        ncl = len(result.clusters)
        centers = np.ascontiguousarray([c["center"] for c in result.clusters], np.uint32)
        bfrom = np.ascontiguousarray([c["birth_from"] for c in result.clusters], np.uint32)
        nsubs = np.zeros(ncl, np.uint32)
        qave = np.zeros(ncl, np.float64)
        nrec = C.c_uint32()
        # a substitution sits at a parent position: at most the longest
        # length per birth pair
        cap = max(1, (ncl - 1) * int(self.maxlen))
        cl, pos = np.zeros(cap, np.uint32), np.zeros(cap, np.uint32)
        nt0, nt1, q = np.zeros(cap, np.uint8), np.zeros(cap, np.uint8), np.zeros(cap, np.uint8)
        _check(_lib().d2g_birth_subs(self.h, ncl, _p(centers, u32p), _p(bfrom, u32p),
                                     C.byref(params), int(use_quals), cap, _p(nsubs, u32p),
                                     _p(qave, f64p), _p(cl, u32p), _p(pos, u32p), _p(nt0, u8p),
                                     _p(nt1, u8p), _p(q, u8p), C.byref(nrec)), "d2g_birth_subs")
        k = nrec.value
        quals = use_quals and self.has_quals
        qs = q[:k].tolist() if quals else [None] * k
        recs = list(zip(cl[:k].tolist(), pos[:k].tolist(), nt0[:k].tolist(), nt1[:k].tolist(), qs))
        return recs, np.where(nsubs == NO_CENTER, -1, nsubs.astype(np.int64)), qave

    def compute_aux(self, result: "DadaResult", params: AlignParams | None = None,
                    use_quals: bool = True, ncol: int | None = None) -> dict:
        """compute_aux (dada.rs:610-640) for a run_dada result: DadaAux's
        cluster_stats (error.rs:55-140), cluster_quality, birth_subs records
        and transitions.  ncol defaults to the run's error-matrix columns
        (after any extension, as compute_aux reads params.err_ncol)."""
        params = params or align_params()
        ncol = result.err_ncol if ncol is None else ncol
        a = self.dada_aux(result, params, use_quals, ncol)
        recs, _bn, qave = self.birth_subs(result, params, use_quals)
        reads = np.asarray(self.reads, np.int64)
        correct = np.asarray(result.map) >= 0
        cof = np.asarray(result.cluster_of, np.int64)
        stats = []
        for k, c in enumerate(result.clusters):
            m = correct & (cof == k)
            ns = a["nsubs"]
            stats.append({
                "center": int(c["center"]),
                "abundance": int(reads[m].sum()) & 0xFFFFFFFF,
                "n0": int(reads[m & (ns == 0)].sum()) & 0xFFFFFFFF,
                "n1": int(reads[m & (ns == 1)].sum()) & 0xFFFFFFFF,
                "nunq": int(m.sum()),
                "pval": c["posthoc_pval"],
                "birth_from": None if k == 0 else int(c["birth_from"]),
                "birth_pval": 0.0 if k == 0 else c["birth_pval"],
                "birth_fold": 0.0 if k == 0 else c["birth_fold"],
                "birth_ham": 0 if k == 0 else int(c["birth_hamming"]),
                "birth_e": 0.0 if k == 0 else c["birth_e"],
                "birth_qave": None if k == 0 or qave[k] != qave[k] else float(qave[k]),
            })
        return {"cluster_stats": stats, "cluster_quality": a["cluster_quality"],
                "birth_subs": recs, "transitions": a["transitions"],
                "transitions_ncol": a["transitions"].shape[1], "nsubs": a["nsubs"]}

    def transition_stats(self) -> dict:
        """Counters (and, with timing on, kernel times) of the last
        transition_counts / build_trans_mat call."""
        st = Stats()
        _check(_lib().d2g_transition_stats(self.h, C.byref(st)), "d2g_transition_stats")
        return st.as_dict()

    def _center_weight(self, center_of, weight):
        co = np.asarray(center_of)
        if co.shape != (self.n,) or np.shape(weight) != (self.n,):
            raise ValueError("center_of and weight need one entry per unique")
        co = np.ascontiguousarray(np.where(co < 0, NO_CENTER, co), np.uint32)
        return co, np.ascontiguousarray(weight, np.uint32)

    def pairs_load(self, seqs, quals, subs, reads, center_reads):
        """Resident aligned pairs: query i with its substitution list subs[i]
        (records (query position << 16 | center base), ascending)."""
        cat, off = _concat(seqs)
        q = None
        if quals is not None:
            q, _ = _concat(quals)
        sc = np.ascontiguousarray(np.concatenate([np.asarray(x, np.uint32) for x in subs])
                                  if len(subs) else np.zeros(0, np.uint32), np.uint32)
        so = np.zeros(len(subs) + 1, np.uint64)
        so[1:] = np.cumsum([len(x) for x in subs])
        if sc.size == 0:
            sc = np.zeros(1, np.uint32)
        r = np.ascontiguousarray(reads, np.uint32)
        cr = np.ascontiguousarray(center_reads, np.uint32)
        _check(_lib().d2g_pairs_load(self.h, len(seqs), _p(cat, u8p), _p(off, u64p),
                                     _p(q, u8p) if q is not None else None, _p(sc, u32p),
                                     _p(so, u64p), _p(r, u32p), _p(cr, u32p)), "d2g_pairs_load")
        self.n_pairs = len(seqs)

    def pairs_lambda(self, err, use_quals=True, omega_a=1e-40, fetch=True):
        """FP32 Kahan log-lambda + FP64 p-values (FP64 lambda recompute within
        10x of omega_a).  Returns (loglam32, lambda, pval, recomputed, stats)."""
        e = np.ascontiguousarray(err, np.float64)
        st = PairStats()
        _check(_lib().d2g_pairs_lambda(self.h, _p(e, f64p), e.shape[1], int(use_quals),
                                       float(omega_a), C.byref(st)), "d2g_pairs_lambda")
        if not fetch:
            return st.as_dict()
        n = self.n_pairs
        ll = np.zeros(n, np.float32)
        lam = np.zeros(n, np.float64)
        pv = np.zeros(n, np.float64)
        fl = np.zeros(n, np.uint8)
        _check(_lib().d2g_pairs_fetch(self.h, ll.ctypes.data_as(C.POINTER(C.c_float)),
                                      _p(lam, f64p), _p(pv, f64p), _p(fl, u8p)), "d2g_pairs_fetch")
        return ll, lam, pv, fl, st.as_dict()

    def pairs_lambda_noalloc(self, n: int, pool_bytes: int, n_subs: int, ptrs: dict, err,
                             out_ptrs: dict, use_quals=True, omega_a=1e-40) -> dict:
        """d2g_pairs_lambda_noalloc over a pair set in device memory.  ptrs
        maps seq, qual (0 for none), off, len, subs, sub_off, reads and
        center_reads to device addresses (pairs_layout's arrays after the
        copy); out_ptrs maps loglam32, lambda, pval and recomputed to device
        addresses of n entries.  Returns the call's PairStats."""
        e = np.ascontiguousarray(err, np.float64)
        pd = PairsDev(n, 0, pool_bytes, n_subs, ptrs["seq"], ptrs.get("qual") or None,
                      ptrs["off"], ptrs["len"], ptrs["subs"], ptrs["sub_off"], ptrs["reads"],
                      ptrs["center_reads"])
        st = PairStats()
        _check(_lib().d2g_pairs_lambda_noalloc(self.h, C.byref(pd), _p(e, f64p), e.shape[1],
                                               int(use_quals), float(omega_a),
                                               out_ptrs["loglam32"], out_ptrs["lambda"],
                                               out_ptrs["pval"], out_ptrs["recomputed"],
                                               C.byref(st)), "d2g_pairs_lambda_noalloc")
        return st.as_dict()

    def loglik_forward(self, logE, use_quals=True) -> float:
        """Poisson log-likelihood of the abundances over the last compare's
        aligned pairs for the 16 x ncol FP32 log error matrix logE."""
        e = np.ascontiguousarray(logE, np.float32)
        loss = C.c_double()
        ms = C.c_float()
        _check(_lib().d2g_loglik_forward(self.h, e.ctypes.data_as(C.POINTER(C.c_float)),
                                         e.shape[1], int(use_quals), C.byref(loss), C.byref(ms)),
               "d2g_loglik_forward")
        self.last_loglik_ms = ms.value
        self._ll_shape = e.shape
        return loss.value

    def loglik_backward(self):
        """Gradient of the last forward's loss in logE: (float32, float64)."""
        g32 = np.zeros(self._ll_shape, np.float32)
        g64 = np.zeros(self._ll_shape, np.float64)
        ms = C.c_float()
        _check(_lib().d2g_loglik_backward(self.h, g32.ctypes.data_as(C.POINTER(C.c_float)),
                                          _p(g64, f64p), C.byref(ms)), "d2g_loglik_backward")
        self.last_loglik_bwd_ms = ms.value
        return g32, g64

    def bimera_pairs(self, queries, parents, params: BimeraParams | None = None):
        """get_lr + get_ham_endsfree per (query, parent) pair: returns
        (lr (n x 4: left, right, left_oo, right_oo), ham (n))."""
        params = params or bimera_params()
        n = len(queries)
        qc, qo = _concat(queries)
        pc, po = _concat(parents)
        lr = np.zeros((n, 4), np.uint32)
        ham = np.zeros(n, np.uint32)
        _check(_lib().d2g_bimera_pairs(self.h, n, _p(qc, u8p), _p(qo, u64p), _p(pc, u8p),
                                       _p(po, u64p), C.byref(params), _p(lr, u32p),
                                       _p(ham, u32p)), "d2g_bimera_pairs")
        return lr, ham

    def is_bimera_batch(self, queries, parent_lists, params: BimeraParams | None = None):
        """is_bimera_with_buf (chimera.rs:218) for each query against its own
        parent list."""
        params = params or bimera_params()
        nq = len(queries)
        qc, qo = _concat(queries)
        flat = [p for ps in parent_lists for p in ps]
        pc, po = _concat(flat) if flat else (np.zeros(1, np.uint8), np.zeros(1, np.uint64))
        ps_ = np.zeros(nq + 1, np.uint32)
        ps_[1:] = np.cumsum([len(ps) for ps in parent_lists])
        out = np.zeros(nq, np.uint8)
        _check(_lib().d2g_is_bimera_batch(self.h, nq, _p(qc, u8p), _p(qo, u64p), _p(pc, u8p),
                                          _p(po, u64p), _p(ps_, u32p), C.byref(params),
                                          _p(out, u8p)), "d2g_is_bimera_batch")
        return out.astype(bool)

    def is_bimera_pool(self, queries, pool, parent_idx, params: BimeraParams | None = None):
        """is_bimera_with_buf (chimera.rs:218) for each query against the
        pool sequences parent_idx[i] (one index list per query)."""
        params = params or bimera_params()
        nq = len(queries)
        qc, qo = _concat(queries)
        pc, po = _concat(pool)
        idx = [int(x) for ps in parent_idx for x in ps]
        pi = np.asarray(idx if idx else [0], np.uint32)
        ps_ = np.zeros(nq + 1, np.uint32)
        ps_[1:] = np.cumsum([len(ps) for ps in parent_idx])
        out = np.zeros(nq, np.uint8)
        _check(_lib().d2g_is_bimera_pool(self.h, nq, _p(qc, u8p), _p(qo, u64p), len(pool),
                                         _p(pc, u8p), _p(po, u64p), _p(pi, u32p), _p(ps_, u32p),
                                         C.byref(params), _p(out, u8p)), "d2g_is_bimera_pool")
        return out.astype(bool)

    def table_bimera2(self, mat, seqs, min_fold=1.5, min_abund=2,
                      params: BimeraParams | None = None):
        """table_bimera2 (chimera.rs:317): mat is (samples, sequences)."""
        params = params or bimera_params()
        mat = np.asarray(mat)
        nrow, ncol = mat.shape
        colmajor = np.ascontiguousarray(mat.T.reshape(-1), np.uint32)
        cat, off = _concat(seqs)
        nflag = np.zeros(ncol, np.uint32)
        nsam = np.zeros(ncol, np.uint32)
        _check(_lib().d2g_table_bimera2(self.h, _p(colmajor, u32p), nrow, ncol, _p(cat, u8p),
                                        _p(off, u64p), float(min_fold), int(min_abund),
                                        C.byref(params), _p(nflag, u32p), _p(nsam, u32p)),
               "d2g_table_bimera2")
        return nflag, nsam

    @staticmethod
    def _pair_buffers(s1s, s2s):
        if len(s1s) != len(s2s):
            raise ValueError("s1s and s2s differ in length")
        n = len(s1s)
        c1, o1 = _concat(s1s)
        c2, o2 = _concat(s2s)
        l1 = np.diff(o1)
        l2 = np.diff(o2)
        aoff = np.zeros(n + 1, np.uint64)
        aoff[1:] = np.cumsum(l1 + l2)
        total = int(aoff[-1]) + 1
        return (n, c1, o1, c2, o2, aoff, np.zeros(total, np.uint8), np.zeros(total, np.uint8),
                np.zeros(n, np.uint32))

    @staticmethod
    def _pairs_out(n, aoff, a0, a1, alen):
        out = []
        for i in range(n):
            s, m = int(aoff[i]), int(alen[i])
            out.append((a0[s:s + m].tobytes(), a1[s:s + m].tobytes()))
        return out

    def align_vectorized(self, s1s, s2s, match=5, mismatch=-4, gap_p=-8, end_gap_p=0,
                         band=16) -> list[tuple[bytes, bytes]]:
        """align_vectorized_with_buf (nwalign.rs:761) for each (s1, s2) pair."""
        n, c1, o1, c2, o2, aoff, a0, a1, alen = self._pair_buffers(s1s, s2s)
        if n == 0:
            return []
        sc = VecScores(match, mismatch, gap_p, end_gap_p, band)
        _check(_lib().d2g_align_vectorized(self.h, n, _p(c1, u8p), _p(o1, u64p), _p(c2, u8p),
                                           _p(o2, u64p), C.byref(sc), _p(a0, u8p), _p(a1, u8p),
                                           _p(aoff, u64p), _p(alen, u32p)),
               "d2g_align_vectorized")
        return self._pairs_out(n, aoff, a0, a1, alen)

    def align_endsfree(self, s1s, s2s, match=5, mismatch=-4, gap_p=-8, homo_gap_p=None,
                       band=16) -> list[tuple[bytes, bytes]]:
        """align_endsfree_with_buf (nwalign.rs:336), or with homo_gap_p the
        homopolymer variant align_endsfree_homo_with_buf (nwalign.rs:435),
        for each (s1, s2) pair."""
        n, c1, o1, c2, o2, aoff, a0, a1, alen = self._pair_buffers(s1s, s2s)
        if n == 0:
            return []
        hg = gap_p if homo_gap_p is None else homo_gap_p
        _check(_lib().d2g_align_endsfree(self.h, n, _p(c1, u8p), _p(o1, u64p), _p(c2, u8p),
                                         _p(o2, u64p), match, mismatch, gap_p, hg, band,
                                         _p(a0, u8p), _p(a1, u8p), _p(aoff, u64p), _p(alen, u32p)),
               "d2g_align_endsfree")
        return self._pairs_out(n, aoff, a0, a1, alen)
