"""Benchmark of the dada2-rs compare hot path on one MI355X.

Workload (one step): the batched uneven compare — 200 centers x 20,000
uniques of 239-261 nt (variant prefixes of 240-260 nt with 0.5% substitutions
and 0.05% 1-nt indels), k=5 screen at cutoff 0.42, band 16, (5, -4, -8),
greedy, FP64 lambda: 4.0e6 (center, unique) comparisons through
screen -> plan -> vectorized DP -> lambda in one device pipeline, inputs
resident in HBM.  Prints one JSON line.

Also measured: the plain path (one center per call, results copied to the
host as the reference's b_compare_parallel returns them) on the same
uniques, and the CPU oracle (C restatement of the reference compare,
pthreads over chunks of 32 raws) on a bounded sample of centers.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

import __graft_entry__  # noqa: E402

HBM_PEAK_GBS = 8000.0


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--plain-centers", type=int, default=20,
                    help="centers timed one call each for the plain path")
    ap.add_argument("--cpu-seconds", type=float, default=10.0,
                    help="minimum CPU time of the baseline sample (whole passes)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--copy-calibration-mib", type=int, default=0,
                    help="after the run, one device-to-device copy of this size (a plain copy "
                         "kernel of known bytes that scales the HBM byte counters)")
    return ap.parse_args()


def cpu_baseline(w, min_seconds: float) -> dict:
    """The oracle's compare (C, threads over chunks of raws) on whole passes of
    the same workload: every center against every unique, repeated until at
    least min_seconds of wall time have passed."""
    import oracle as O

    threads = O.cpu_count_allowed()
    rs = O.RawSet(w.seqs, w.quals, w.reads, k=w.kmer_size)
    p = O.make_params(band=w.band)
    cs = [int(c) for c in w.centers]
    O.compare_center(rs, cs[0], w.err, w.err.shape[1], p, nthreads=threads)  # warm
    passes = 0
    t0 = time.perf_counter()
    while True:
        for c in cs:
            O.compare_center(rs, c, w.err, w.err.shape[1], p, nthreads=threads)
        passes += 1
        dt = time.perf_counter() - t0
        if dt >= min_seconds:
            break
    comps = passes * len(cs) * w.n
    return {"value": comps / dt, "unit": "comparisons/s", "cores": threads, "kind": "port",
            "sample": f"{passes} passes of {len(cs)} centers x {w.n} uniques of the same workload "
                      f"({comps} comparisons, {dt:.1f} s, C oracle with {threads} threads)"}


def copy_calibration(mib: int) -> None:
    """One hipMemcpy device-to-device of mib MiB: the runtime's copy kernel
    moves exactly that many bytes each way, which calibrates FETCH_SIZE and
    WRITE_SIZE in a profiled run."""
    import ctypes as C

    hip = C.CDLL("libamdhip64.so")
    n = mib << 20
    a, b = C.c_void_p(), C.c_void_p()
    assert hip.hipMalloc(C.byref(a), C.c_size_t(n)) == 0
    assert hip.hipMalloc(C.byref(b), C.c_size_t(n)) == 0
    assert hip.hipMemset(a, 1, C.c_size_t(n)) == 0
    assert hip.hipDeviceSynchronize() == 0
    assert hip.hipMemcpy(b, a, C.c_size_t(n), 3) == 0  # hipMemcpyDeviceToDevice
    assert hip.hipDeviceSynchronize() == 0
    hip.hipFree(a)
    hip.hipFree(b)


def traffic_from_profile(kernel: str):
    """HBM bytes per launch of `kernel` from the committed PMC summary
    (profiles/traffic_batched.txt: 'kernel fetch_kib write_kib' lines already
    scaled by the copy-kernel factor), or None."""
    path = ROOT / "profiles" / "traffic_batched.txt"
    if not path.exists():
        return None
    for line in path.read_text().splitlines():
        f = line.split()
        if len(f) == 3 and f[0] == kernel:
            return (float(f[1]) + float(f[2])) * 1024.0
    return None


def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py measures one GPU: every box here has one MI355X and the compare "
              f"path shards by uniques without cross-GPU traffic; --gpus {args.gpus} is not run.",
              file=sys.stderr)
        sys.exit(2)
    pkg = __graft_entry__.load_package()
    from dada2_rs_amd import synth

    w = synth.workload_batched()
    params = pkg.align_params(band=w.band)
    ctx = pkg.Context(0, timing=True)
    info = ctx.device_info()
    ctx.load_uniques(w.seqs, w.quals, w.reads, k=w.kmer_size)
    centers = np.asarray(w.centers, np.uint32)
    n_pairs = len(centers) * w.n

    for _ in range(args.warmup):
        ctx.compare_batch(centers, w.err, params, device_only=True)
    stats = []
    t0 = time.perf_counter()
    for _ in range(args.steps):
        stats.append(ctx.compare_batch(centers, w.err, params, device_only=True))
    wall = time.perf_counter() - t0
    ms_step = wall / args.steps * 1e3
    value = n_pairs * args.steps / wall

    def avg(key):
        return float(np.mean([s[key] for s in stats]))

    st = stats[-1]
    screened = st["comparisons"] - st["skipped"]
    ms_screen, ms_dp, ms_lam, ms_plan = avg("ms_screen"), avg("ms_dp"), avg("ms_lambda"), avg("ms_plan")
    # screen: 4^k bytes of query k-mer vector per screened comparison
    screen_gbs = 1024.0 * screened / (ms_screen * 1e-3) / 1e9
    cu, clk = info["cu_count"], info["clock_khz"] * 1e3
    valu_cells_peak = cu * 64 * clk / 8.0  # 8 integer ops per cell, int32 lanes
    gcups = st["dp_cells"] / (ms_dp * 1e-3) / 1e9
    kernels = {"screen": ms_screen, "plan+scatter": ms_plan, "nw_vec": ms_dp, "lambda": ms_lam}
    dominant = max(kernels, key=kernels.get)
    roof_screen = {"bound": "hbm", "achieved": screen_gbs, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                   "frac": screen_gbs / HBM_PEAK_GBS, "traffic": traffic_from_profile("screen_kernel"),
                   "kernel": "screen_kernel", "algorithmic_bytes_per_launch": 1024 * screened,
                   "avg_ms": ms_screen}
    roof_dp = {"bound": "valu", "achieved": gcups * 8 / 1e3, "peak": valu_cells_peak * 8 / 1e12,
               "unit": "TOP/s", "frac": gcups * 1e9 / valu_cells_peak,
               "traffic": traffic_from_profile("nw_vec_kernel"),
               "kernel": "nw_vec_kernel", "gcups": gcups, "cells_per_launch": st["dp_cells"],
               "avg_ms": ms_dp}
    roofline = roof_screen if dominant in ("screen", "plan+scatter", "lambda") else roof_dp

    # plain path: one center per call, results copied to the host
    plain = None
    if args.plain_centers > 0:
        pc = centers[:args.plain_centers]
        ctx.set_timing(False)
        ctx.compare_center(int(pc[0]), w.err, params)
        t1 = time.perf_counter()
        for c in pc:
            ctx.compare_center(int(c), w.err, params)
        plain_s = time.perf_counter() - t1
        plain_value = len(pc) * w.n / plain_s
        plain = {"value": plain_value, "unit": "comparisons/s", "centers": int(len(pc)),
                 "batched_over_plain": value / plain_value}

    # batched end to end, results copied to the host (PCIe inclusive)
    t2 = time.perf_counter()
    ctx.compare_batch(centers, w.err, params)
    e2e_value = n_pairs / (time.perf_counter() - t2)
    ctx.close()
    if args.copy_calibration_mib > 0:
        copy_calibration(args.copy_calibration_mib)

    cpu = None if args.no_cpu else cpu_baseline(w, args.cpu_seconds)
    line = {
        "metric": "center-unique comparisons/sec + GCUPS, 1 GPU, batched vs plain, bit-exact align",
        "value": value, "unit": "comparisons/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": ms_step, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "i16 DP / f64 lambda",
        "data": "synthetic",
        "config": {"workload": "batched uneven compare: 200 centers x 20000 uniques, 239-261 nt, "
                               "k=5 cutoff 0.42, band 16, (5,-4,-8), greedy, FP64 lambda",
                   "pairs_per_step": n_pairs},
        "roofline": roofline,
        "cpu_baseline": cpu,
        "gcups": gcups,
        "kernel_ms": kernels,
        "roofline_screen": roof_screen,
        "roofline_dp": roof_dp,
        "plain_path": plain,
        "batched_e2e_with_d2h": e2e_value,
        "counts": {k: st[k] for k in ("comparisons", "skipped", "shrouded", "gapless",
                                      "aligned_dp", "dp_cells")},
        "device": info,
    }
    print(json.dumps(line))


if __name__ == "__main__":
    main()
