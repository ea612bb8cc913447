// plan.hip — binning of DP pairs by geometry, on the device.
//
// The vectorized DP walks anti-diagonals with a row schedule fixed by
// (len1, len2, band) (nwalign.rs:875-1088).  Pairs with identical
// (len_center, len_query) share a wavefront so that schedule is uniform
// (scalar registers) while lanes compute cells.  plan_kernel turns per-bin
// pair counts into bin offsets and wavefront offsets (G = 64 / ncol pairs
// per wavefront); scatter_kernel places each pair into its bin.
#include "d2g_launch.h"

namespace d2g {

__global__ void zero_kernel(uint32_t* p, size_t n) {
  size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  size_t stride = (size_t)gridDim.x * blockDim.x;
  for (; i < n; i += stride) p[i] = 0;
}

hipError_t launch_zero(uint32_t* p, size_t n_words, hipStream_t s) {
  if (n_words == 0) return hipSuccess;
  size_t blocks = (n_words + 255) / 256;
  if (blocks > 4096) blocks = 4096;
  hipLaunchKernelGGL(zero_kernel, dim3((unsigned)blocks), dim3(256), 0, s, p, n_words);
  return hipGetLastError();
}

constexpr int kPlanThreads = 1024;

// Exclusive scans of the pair counts and of the wavefront counts over bins;
// also the call's DP cell total (per-bin cells from the schedule) and the
// screen's skip/shroud totals from its per-block partial counts (written
// without atomics: device-scope atomics on one word serialise across the
// chip).
__global__ __launch_bounds__(kPlanThreads) void plan_kernel(BinSpace bs,
                                                            const uint32_t* bin_count,
                                                            uint32_t* bin_start,
                                                            uint32_t* wave_start,
                                                            uint32_t* bin_fill, Counters* ctr,
                                                            const BinInfo* info,
                                                            const uint32_t* partials,
                                                            uint32_t n_partials) {
  __shared__ uint32_t s_pairs[kPlanThreads];
  __shared__ uint32_t s_waves[kPlanThreads];
  __shared__ unsigned long long s_cells[kPlanThreads / 64];
  __shared__ uint32_t s_skip[kPlanThreads / 64], s_shroud[kPlanThreads / 64];
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
    VecGeom g = vec_geom(l1, l2, bs.band);
    int G = nw_groups(g);
    if (cnt && !G) bad = 1;
    sp += cnt;
    sw += G ? (cnt + G - 1) / G : 0;
    if (cnt && info) cells += (unsigned long long)cnt * info[b].cells;
  }
  uint32_t skip = 0, shroud = 0;
  for (uint32_t i = tid; i < n_partials; i += kPlanThreads) {
    skip += partials[2 * i];
    shroud += partials[2 * i + 1];
  }
  for (int off = 32; off > 0; off >>= 1) {
    cells += __shfl_xor(cells, off);
    skip += __shfl_xor(skip, off);
    shroud += __shfl_xor(shroud, off);
  }
  if ((tid & 63) == 0) {
    s_cells[tid >> 6] = cells;
    s_skip[tid >> 6] = skip;
    s_shroud[tid >> 6] = shroud;
  }
  s_pairs[tid] = sp;
  s_waves[tid] = sw;
  __syncthreads();
  for (int off = 1; off < kPlanThreads; off <<= 1) {
    uint32_t vp = tid >= off ? s_pairs[tid - off] : 0;
    uint32_t vw = tid >= off ? s_waves[tid - off] : 0;
    __syncthreads();
    s_pairs[tid] += vp;
    s_waves[tid] += vw;
    __syncthreads();
  }
  uint32_t rp = s_pairs[tid] - sp, rw = s_waves[tid] - sw;  // exclusive prefix
  for (uint32_t b = b0; b < b1; ++b) {
    uint32_t cnt = bin_count[b];
    int l1, l2;
    bin_lengths(bs, b, &l1, &l2);
    VecGeom g = vec_geom(l1, l2, bs.band);
    int G = nw_groups(g);
    bin_start[b] = rp;
    wave_start[b] = rw;
    bin_fill[b] = 0;
    rp += cnt;
    rw += G ? (cnt + G - 1) / G : 0;
  }
  if (tid == kPlanThreads - 1) {
    bin_start[nb] = s_pairs[tid];
    wave_start[nb] = s_waves[tid];
    ctr->total_waves = s_waves[tid];
    ctr->wave_counter = 0;
    unsigned long long tc = 0;
    uint32_t ts = 0, th = 0;
    for (int w = 0; w < kPlanThreads / 64; ++w) {
      tc += s_cells[w];
      ts += s_skip[w];
      th += s_shroud[w];
    }
    ctr->dp_cells = tc;
    if (partials) {
      ctr->n_skipped = ts;
      ctr->n_shrouded = th;
    }
  }
  if (bad) atomicOr(&ctr->err_flags, kErrGeometry);
}

hipError_t launch_plan(const BinSpace& bs, const uint32_t* bin_count, uint32_t* bin_start,
                       uint32_t* wave_start, uint32_t* bin_fill, Counters* ctr,
                       const BinInfo* info, const uint32_t* partials, uint32_t n_partials,
                       hipStream_t s) {
  hipLaunchKernelGGL(plan_kernel, dim3(1), dim3(kPlanThreads), 0, s, bs, bin_count, bin_start,
                     wave_start, bin_fill, ctr, info, partials, n_partials);
  return hipGetLastError();
}

__global__ void scatter_kernel(const DpTask* unsorted, DpTask* sorted, const uint32_t* bin_start,
                               uint32_t* bin_fill, const Counters* ctr, uint32_t capacity) {
  const uint32_t n = min(ctr->n_dp, capacity);
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    DpTask t = unsorted[i];
    uint32_t pos = bin_start[t.bin] + atomicAdd(&bin_fill[t.bin], 1u);
    sorted[pos] = t;
  }
}

hipError_t launch_scatter(const DpTask* unsorted, DpTask* sorted, const uint32_t* bin_start,
                          uint32_t* bin_fill, const Counters* ctr, uint32_t capacity,
                          hipStream_t s) {
  uint32_t blocks = (capacity + 255) / 256;
  if (blocks > 2048) blocks = 2048;
  if (blocks == 0) blocks = 1;
  hipLaunchKernelGGL(scatter_kernel, dim3(blocks), dim3(256), 0, s, unsorted, sorted, bin_start,
                     bin_fill, ctr, capacity);
  return hipGetLastError();
}

}  // namespace d2g

namespace d2g {

// Results are produced in the length-sorted unique order; this restores the
// caller's order: dst[c * n + orig[s]] = src[c * n + s].
__global__ void unpermute_kernel(uint32_t nc, uint32_t n, const uint32_t* orig,
                                 const double* src_lam, const uint32_t* src_ham,
                                 const uint8_t* src_skp, double* dst_lam, uint32_t* dst_ham,
                                 uint8_t* dst_skp) {
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
                            uint32_t* dst_ham, uint8_t* dst_skp, hipStream_t s) {
  size_t total = (size_t)nc * n;
  size_t blocks = (total + 255) / 256;
  if (blocks > 8192) blocks = 8192;
  if (blocks == 0) blocks = 1;
  hipLaunchKernelGGL(unpermute_kernel, dim3((unsigned)blocks), dim3(256), 0, s, nc, n, orig,
                     src_lam, src_ham, src_skp, dst_lam, dst_ham, dst_skp);
  return hipGetLastError();
}

}  // namespace d2g
