// plan.hip — binning of DP pairs by geometry, on the device.
//
// Both DP kernels walk anti-diagonals with a geometry fixed by (len1, len2,
// band) (nwalign.rs:875-1088 for the vectorized aligner; the row band of
// nwalign.rs:276-308 for the scalar ones).  Pairs with identical
// (len_center, len_query) share a wavefront so that the walk is uniform
// (scalar registers) while lanes compute cells.  plan_kernel turns per-bin
// pair counts into bin offsets and wavefront offsets (G pairs per
// wavefront, bin_groups); scatter_kernel places each pair into its bin.
#include "d2g_device.h"
#include "d2g_launch.h"

namespace d2g {

__global__ void zero_kernel(uint32_t* p, size_t n) {
  size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  size_t stride = (size_t)gridDim.x * blockDim.x;
  for (; i < n; i += stride) p[i] = 0;
}

// Two buffers in one launch: words [0, na) of a, then [0, nb) of b.
__global__ void zero2_kernel(uint32_t* a, size_t na, uint32_t* b, size_t nb) {
  size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  const size_t stride = (size_t)gridDim.x * blockDim.x;
  for (; i < na + nb; i += stride) {
    if (i < na)
      a[i] = 0;
    else
      b[i - na] = 0;
  }
}

hipError_t launch_zero2(uint32_t* a, size_t na, uint32_t* b, size_t nb, hipStream_t s) {
  if (na + nb == 0) return hipSuccess;
  size_t blocks = (na + nb + 255) / 256;
  if (blocks > 4096) blocks = 4096;
  hipLaunchKernelGGL(zero2_kernel, dim3((unsigned)blocks), dim3(256), 0, s, a, na, b, nb);
  return hipGetLastError();
}

hipError_t launch_zero(uint32_t* p, size_t n_words, hipStream_t s) {
  if (n_words == 0) return hipSuccess;
  size_t blocks = (n_words + 255) / 256;
  if (blocks > 4096) blocks = 4096;
  hipLaunchKernelGGL(zero_kernel, dim3((unsigned)blocks), dim3(256), 0, s, p, n_words);
  return hipGetLastError();
}

constexpr int kPlanThreads = 1024;

// Single-workgroup exclusive scan of v over the 1024 threads (returns the
// exclusive prefix; *total gets the sum).  Each wavefront scans its lanes
// with shuffles, wavefront 0 scans the 16 wavefront totals, and three
// barriers order the LDS traffic.  Every thread must call it.
__device__ uint32_t block_scan(uint32_t v, uint32_t* sh, uint32_t* total) {
  constexpr int kWaves = kPlanThreads / 64;
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  uint32_t x = v;
#pragma unroll
  for (int off = 1; off < 64; off <<= 1) {
    const uint32_t y = (uint32_t)__shfl_up((int)x, off);
    x += lane >= off ? y : 0u;
  }
  if (lane == 63) sh[w] = x;
  __syncthreads();
  if (w == 0) {
    uint32_t t = lane < kWaves ? sh[lane] : 0u;
#pragma unroll
    for (int off = 1; off < kWaves; off <<= 1) {
      const uint32_t y = (uint32_t)__shfl_up((int)t, off);
      t += lane >= off ? y : 0u;
    }
    if (lane < kWaves) sh[kWaves + lane] = t;  // inclusive wavefront prefixes
  }
  __syncthreads();
  const uint32_t before = w ? sh[kWaves + w - 1] : 0u;
  *total = sh[2 * kWaves - 1];
  __syncthreads();  // sh is reused by the next scan
  return before + x - v;
}

// Exclusive scans of the pair counts and of the wavefront counts over bins;
// the call's DP cell total (per-bin cells from the schedule); and, for a
// compare, the screen's per-block counts (skipped, shrouded, gapless, DP):
// totals, and each block's offset into the compacted gapless list.  Totals
// are written here rather than by atomics in the screen: device-scope
// atomics on one word serialise across the chip.
__global__ __launch_bounds__(kPlanThreads) void plan_kernel(BinSpace bs,
                                                            const uint32_t* bin_count,
                                                            uint32_t* bin_start,
                                                            uint32_t* wave_start,
                                                            uint32_t* bin_fill, Counters* ctr,
                                                            const BinInfo* info,
                                                            const uint32_t* partials,
                                                            uint32_t n_partials,
                                                            uint32_t* gl_off, PlanCopy copy,
                                                            int do_copy) {
  __shared__ uint32_t sh[kPlanThreads];
  __shared__ unsigned long long s_cells[kPlanThreads / 64];
  const int tid = threadIdx.x;
  const uint32_t nb = (uint32_t)bs.nbins;
  const uint32_t chunk = (nb + kPlanThreads - 1) / kPlanThreads;
  const uint32_t b0 = min(nb, tid * chunk), b1 = min(nb, b0 + chunk);
  uint32_t sp = 0, sw = 0, bad = 0;
  unsigned long long cells = 0;
  for (uint32_t b = b0; b < b1; ++b) {
    uint32_t cnt = bin_count[b];
    int l1, l2;
    bin_lengths(bs, b, &l1, &l2);
    const int G = bin_groups(bs, l1, l2);
    // vectorized bins too wide for nw_vec take the one-pair kernel
    if (cnt && !G && !(bs.kind == kDpVectorized && wide_fits(vec_geom(l1, l2, bs.band)))) bad = 1;
    sp += cnt;
    sw += G ? (cnt + G - 1) / G : 0;
    if (cnt && info) cells += (unsigned long long)cnt * info[b].cells;
  }
  for (int off = 32; off > 0; off >>= 1) cells += __shfl_xor(cells, off);
  if ((tid & 63) == 0) s_cells[tid >> 6] = cells;
  uint32_t tot_pairs, tot_waves;
  uint32_t rp = block_scan(sp, sh, &tot_pairs);
  uint32_t rw = block_scan(sw, sh, &tot_waves);
  for (uint32_t b = b0; b < b1; ++b) {
    uint32_t cnt = bin_count[b];
    int l1, l2;
    bin_lengths(bs, b, &l1, &l2);
    const int G = bin_groups(bs, l1, l2);
    bin_start[b] = rp;
    wave_start[b] = rw;
    bin_fill[b] = 0;
    rp += cnt;
    rw += G ? (cnt + G - 1) / G : 0;
  }
  // screen blocks: skip / shroud totals, gapless offsets
  uint32_t skip = 0, shroud = 0, ngl = 0, ndp = 0;
  const uint32_t pchunk = (n_partials + kPlanThreads - 1) / kPlanThreads;
  const uint32_t p0 = min(n_partials, tid * pchunk), p1 = min(n_partials, p0 + pchunk);
  if (partials) {
    for (uint32_t i = p0; i < p1; ++i) {
      skip += partials[4 * i];
      shroud += partials[4 * i + 1];
      ngl += partials[4 * i + 2];
      ndp += partials[4 * i + 3];
    }
  }
  uint32_t tot_skip, tot_shroud, tot_gl;
  (void)block_scan(skip, sh, &tot_skip);
  (void)block_scan(shroud, sh, &tot_shroud);
  uint32_t go = block_scan(ngl, sh, &tot_gl);
  if (do_copy) {
    // regions in order hold the DP tasks in bin order: region i's tasks go
    // to the exclusive prefix of the regions' DP counts, which is where
    // bin_start and the in-region order put them
    uint32_t tot_dp;
    uint32_t dpo = block_scan(ndp, sh, &tot_dp);
    for (uint32_t i = p0; i < p1; ++i) {
      const uint32_t nd = partials[4 * i + 3], ng = partials[4 * i + 2];
      const DpTask* ds = copy.dp_unsorted + (size_t)i * copy.region;
      const GaplessTask* gs = copy.gl_unsorted + (size_t)i * copy.region;
      for (uint32_t j = 0; j < nd; ++j) copy.dp_sorted[dpo + j] = ds[j];
      for (uint32_t j = 0; j < ng; ++j) copy.gl[go + j] = gs[j];
      dpo += nd;
      go += ng;
    }
  } else if (partials) {
    for (uint32_t i = p0; i < p1; ++i) {
      gl_off[i] = go;
      go += partials[4 * i + 2];
    }
  }
  if (tid == kPlanThreads - 1) {
    bin_start[nb] = tot_pairs;
    wave_start[nb] = tot_waves;
    ctr->total_waves = tot_waves;
    ctr->wave_counter = 0;
    unsigned long long tc = 0;
    for (int w = 0; w < kPlanThreads / 64; ++w) tc += s_cells[w];
    ctr->dp_cells = tc;
    if (partials) {
      ctr->n_skipped = tot_skip;
      ctr->n_shrouded = tot_shroud;
      ctr->n_gapless = tot_gl;
      ctr->n_dp = tot_pairs;
    }
  }
  if (bad) atomicOr(&ctr->err_flags, kErrGeometry);
}

hipError_t launch_plan(const BinSpace& bs, const uint32_t* bin_count, uint32_t* bin_start,
                       uint32_t* wave_start, uint32_t* bin_fill, Counters* ctr,
                       const BinInfo* info, const uint32_t* partials, uint32_t n_partials,
                       uint32_t* gl_off, hipStream_t s, const PlanCopy* copy) {
  if (copy && !partials) return hipErrorInvalidValue;
  PlanCopy cp{};
  if (copy) cp = *copy;
  hipLaunchKernelGGL(plan_kernel, dim3(1), dim3(kPlanThreads), 0, s, bs, bin_count, bin_start,
                     wave_start, bin_fill, ctr, info, partials, n_partials, gl_off, cp,
                     copy ? 1 : 0);
  return hipGetLastError();
}

// One wavefront per task region (regions strided over a bounded grid; most
// regions hold a few tasks): DP tasks go to their bin (bin_start plus a
// per-bin fill counter); gapless tasks are compacted at gl_off[region].
// Region r holds partials[4r+3] DP tasks and partials[4r+2] gapless tasks;
// without partials, `count` DP tasks fill the regions in order.
constexpr uint32_t kScatterWaves = 4;
constexpr uint32_t kScatterMaxBlocks = 1024;

__global__ __launch_bounds__(64 * kScatterWaves) void scatter_kernel(
    const DpTask* unsorted, DpTask* sorted, const uint32_t* bin_start, uint32_t* bin_fill,
    const uint32_t* partials, uint32_t n_regions, uint32_t count, uint32_t region,
    const GaplessTask* gl_unsorted, GaplessTask* gl, const uint32_t* gl_off) {
  const uint32_t lane = __lane_id();
  const uint32_t stride = gridDim.x * kScatterWaves;
  for (uint32_t r = (uint32_t)uni((int)(blockIdx.x * kScatterWaves + threadIdx.x / 64u));
       r < n_regions; r += stride) {
    const uint32_t ndp = partials ? partials[4 * r + 3] : min(region, count - r * region);
    const DpTask* src = unsorted + (size_t)r * region;
    // (a wavefront-uniform trip count: every lane takes part in the keyed add)
    for (uint32_t i0 = 0; i0 < ndp; i0 += 64u) {
      const uint32_t i = i0 + lane;
      const bool ok = i < ndp;
      DpTask t;
      if (ok) t = src[i];
      const uint32_t rank = wave_add_by_key(bin_fill, ok ? t.bin : 0u, ok);
      if (ok) sorted[bin_start[t.bin] + rank] = t;
    }
    if (partials) {
      const uint32_t ngl = partials[4 * r + 2];
      const GaplessTask* gsrc = gl_unsorted + (size_t)r * region;
      GaplessTask* gdst = gl + gl_off[r];
      for (uint32_t i = lane; i < ngl; i += 64u) gdst[i] = gsrc[i];
    }
  }
}

hipError_t launch_scatter(const DpTask* unsorted, DpTask* sorted, const uint32_t* bin_start,
                          uint32_t* bin_fill, const uint32_t* partials, uint32_t n_regions,
                          uint32_t region, uint32_t count, const GaplessTask* gl_unsorted,
                          GaplessTask* gl, const uint32_t* gl_off, hipStream_t s) {
  if (n_regions == 0) return hipSuccess;
  uint32_t blocks = (n_regions + kScatterWaves - 1) / kScatterWaves;
  if (blocks > kScatterMaxBlocks) blocks = kScatterMaxBlocks;
  hipLaunchKernelGGL(scatter_kernel, dim3(blocks), dim3(64 * kScatterWaves), 0, s, unsorted,
                     sorted, bin_start, bin_fill, partials, n_regions, count, region,
                     gl_unsorted, gl, gl_off);
  return hipGetLastError();
}

}  // namespace d2g

namespace d2g {

// Results are produced in the length-sorted unique order; this restores the
// caller's order: dst[c * n + orig[s]] = src[c * n + s].  With ctr_dst, the
// first workgroup also copies ctr_words words of the compare's counters
// (final by now) next to the results, so one copy brings both to the host.
__global__ void unpermute_kernel(uint32_t nc, uint32_t n, const uint32_t* orig,
                                 const double* src_lam, const uint32_t* src_ham,
                                 const uint8_t* src_skp, double* dst_lam, uint32_t* dst_ham,
                                 uint8_t* dst_skp, const uint32_t* ctr_src, uint32_t* ctr_dst,
                                 uint32_t ctr_words) {
  if (ctr_dst && blockIdx.x == 0)
    for (uint32_t t = threadIdx.x; t < ctr_words; t += blockDim.x) ctr_dst[t] = ctr_src[t];
  const size_t total = (size_t)nc * n;
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < total;
       i += (size_t)gridDim.x * blockDim.x) {
    const size_t c = i / n, s = i - c * n;
    const size_t d = c * n + orig[s];
    dst_lam[d] = src_lam[i];
    dst_ham[d] = src_ham[i];
    dst_skp[d] = src_skp[i];
  }
}

hipError_t launch_unpermute(uint32_t nc, uint32_t n, const uint32_t* orig, const double* src_lam,
                            const uint32_t* src_ham, const uint8_t* src_skp, double* dst_lam,
                            uint32_t* dst_ham, uint8_t* dst_skp, hipStream_t s,
                            const uint32_t* ctr_src, uint32_t* ctr_dst, uint32_t ctr_words) {
  size_t total = (size_t)nc * n;
  size_t blocks = (total + 255) / 256;
  if (blocks > 8192) blocks = 8192;
  if (blocks == 0) blocks = 1;
  hipLaunchKernelGGL(unpermute_kernel, dim3((unsigned)blocks), dim3(256), 0, s, nc, n, orig,
                     src_lam, src_ham, src_skp, dst_lam, dst_ham, dst_skp, ctr_src, ctr_dst,
                     ctr_words);
  return hipGetLastError();
}

}  // namespace d2g
