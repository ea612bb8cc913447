// store.hip — the store of b_compare_parallel (cluster.rs:351-397) on the
// device.
//
// After a compare, raw `index` (caller order) keeps the comparison with
// center c when lambda * total_reads > e_minmax[index] (strict), and then
// e_minmax[index] rises to lambda * center_reads if that is larger; a
// shrouded comparison is kept with hamming 0 (cluster.rs:389).  e_minmax
// stays resident across calls (in the length-sorted unique order, like the
// compare's results).  With several centers in one call they act as
// successive clusters: each raw walks them in order, so the flags are
// exactly those of one compare per center.  The kept comparisons are
// compacted in (center, raw index) order, the order in which the reference
// pushes them onto each cluster's comp vector:
//   1. store_flag_kernel: one thread per unique (sorted order, coalesced
//      result reads), the centers in order with the next lambdas loaded
//      ahead: the flags, e_minmax and each center's own lambda (Bi.self_);
//   2. store_count_kernel: per center and 256-raw block (caller order), the
//      count of kept comparisons;
//   3. store_scan_kernel: one workgroup, the exclusive scan of those counts
//      (center-major) by per-thread segments, the total after them;
//   4. store_write_kernel: per (block, center), the kept comparisons at their
//      scanned slots, appended to the context's store.
#include "d2g_device.h"
#include "d2g_launch.h"

namespace d2g {

namespace {

constexpr uint32_t kStoreBlock = 256;
constexpr uint32_t kScanThreads = 1024;

__device__ __forceinline__ uint32_t block_count(bool pred, uint32_t* wsum) {
  const uint32_t lane = __lane_id(), wave = threadIdx.x / 64u;
  const uint64_t m = __builtin_amdgcn_ballot_w64(pred);
  if (lane == 0) wsum[wave] = (uint32_t)__builtin_popcountll(m);
  __syncthreads();
  uint32_t t = 0;
  for (uint32_t w = 0; w < kStoreBlock / 64u; ++w) t += wsum[w];
  __syncthreads();
  return t;
}

constexpr uint32_t kFlagAhead = 8;  // lambdas loaded ahead of the e_minmax chain

__global__ __launch_bounds__(kStoreBlock) void store_flag_kernel(StoreArgs a) {
  const uint32_t s = blockIdx.x * kStoreBlock + threadIdx.x;
  if (s >= a.n) return;
  double e = a.e_minmax[s];
  for (uint32_t c0 = 0; c0 < a.nc; c0 += kFlagAhead) {
    double lam[kFlagAhead];
#pragma unroll
    for (uint32_t j = 0; j < kFlagAhead; ++j)
      lam[j] = c0 + j < a.nc ? a.lam[(size_t)(c0 + j) * a.n + s] : 0.0;
#pragma unroll
    for (uint32_t j = 0; j < kFlagAhead; ++j) {
      const uint32_t c = c0 + j;
      if (c >= a.nc) break;
      const uint32_t cs = a.centers[c];
      if (s == cs) a.self_out[c] = lam[j];
      const bool keep = lam[j] * a.total > e;
      if (keep) {
        const double ne = lam[j] * (double)a.reads[cs];
        if (ne > e) e = ne;
      }
      a.flag[(size_t)c * a.n + s] = keep ? 1 : 0;
    }
  }
  a.e_minmax[s] = e;
}

// Kept comparisons per (center, 256-raw block of the caller's order).
__global__ __launch_bounds__(kStoreBlock) void store_count_kernel(StoreArgs a) {
  __shared__ uint32_t wsum[kStoreBlock / 64u];
  const uint32_t c = blockIdx.y, rb = blockIdx.x;
  const uint32_t r = rb * kStoreBlock + threadIdx.x;
  const bool keep = r < a.n && a.flag[(size_t)c * a.n + a.inv[r]] != 0;
  const uint32_t t = block_count(keep, wsum);
  if (threadIdx.x == 0) a.bcount[(size_t)c * gridDim.x + rb] = t;
}

// Exclusive scan of bcount[0 .. m) in place; bcount[m] receives the total.
// Thread t owns the segment [t seg, (t + 1) seg): its sum, a workgroup scan
// of the sums, then the segment's offsets.
__global__ __launch_bounds__(kScanThreads) void store_scan_kernel(uint32_t* bcount, uint32_t m) {
  __shared__ uint32_t sh[kScanThreads];
  const uint32_t tid = threadIdx.x;
  const uint32_t seg = (m + kScanThreads - 1) / kScanThreads;
  const uint32_t b = min(tid * seg, m), e = min(b + seg, m);
  uint32_t sum = 0;
  for (uint32_t i = b; i < e; ++i) sum += bcount[i];
  sh[tid] = sum;
  __syncthreads();
  for (uint32_t off = 1; off < kScanThreads; off <<= 1) {
    const uint32_t x = tid >= off ? sh[tid - off] : 0u;
    __syncthreads();
    sh[tid] += x;
    __syncthreads();
  }
  uint32_t run = sh[tid] - sum;  // exclusive
  for (uint32_t i = b; i < e; ++i) {
    const uint32_t v = bcount[i];
    bcount[i] = run;
    run += v;
  }
  if (tid == kScanThreads - 1) bcount[m] = sh[tid];
}

__global__ __launch_bounds__(kStoreBlock) void store_write_kernel(StoreArgs a) {
  __shared__ uint32_t wsum[kStoreBlock / 64u];
  const uint32_t lane = __lane_id(), wave = threadIdx.x / 64u;
  const uint32_t c = blockIdx.y, rb = blockIdx.x;
  const uint32_t r = rb * kStoreBlock + threadIdx.x;
  const uint32_t s = r < a.n ? a.inv[r] : 0u;
  const bool keep = r < a.n && a.flag[(size_t)c * a.n + s] != 0;
  const uint64_t m = __builtin_amdgcn_ballot_w64(keep);
  if (lane == 0) wsum[wave] = (uint32_t)__builtin_popcountll(m);
  __syncthreads();
  uint32_t base = a.bcount[(size_t)c * gridDim.x + rb];
  for (uint32_t w = 0; w < wave; ++w) base += wsum[w];
  if (keep) {
    const uint32_t slot = base + (uint32_t)__builtin_popcountll(m & ((1ull << lane) - 1ull));
    const uint32_t h = a.ham[(size_t)c * a.n + s];
    ShuffleComp cp;
    cp.raw = r;
    cp.cl = a.first_cluster + c;
    cp.ham = h == kHammingNone ? 0u : h;
    cp._pad = 0u;
    cp.lambda = a.lam[(size_t)c * a.n + s];
    a.store[slot] = cp;
  }
}

__global__ void store_reset_kernel(double* e_minmax, uint32_t n) {
  for (uint32_t r = blockIdx.x * blockDim.x + threadIdx.x; r < n; r += gridDim.x * blockDim.x)
    e_minmax[r] = -999.0;  // containers.rs:342
}

// lock[idx[j]] = 1 for j < k (sorted unique order).
__global__ void lock_set_kernel(uint8_t* lock, const uint32_t* idx, uint32_t k) {
  for (uint32_t j = blockIdx.x * blockDim.x + threadIdx.x; j < k; j += gridDim.x * blockDim.x)
    lock[idx[j]] = 1;
}

}  // namespace

hipError_t launch_lock_set(uint8_t* lock, const uint32_t* idx, uint32_t k, hipStream_t s) {
  if (k == 0) return hipSuccess;
  const uint32_t g = std::max(1u, std::min((k + 255u) / 256u, 1024u));
  hipLaunchKernelGGL(lock_set_kernel, dim3(g), dim3(256), 0, s, lock, idx, k);
  return hipGetLastError();
}

uint32_t store_blocks(uint32_t n) { return (n + kStoreBlock - 1) / kStoreBlock; }

hipError_t launch_store_reset(double* e_minmax, uint32_t n, hipStream_t s) {
  const uint32_t g = std::max(1u, std::min((n + 255u) / 256u, 4096u));
  hipLaunchKernelGGL(store_reset_kernel, dim3(g), dim3(256), 0, s, e_minmax, n);
  return hipGetLastError();
}

hipError_t launch_store(const StoreArgs& a, hipStream_t s) {
  const uint32_t nrb = store_blocks(a.n);
  if (a.n == 0 || a.nc == 0) return hipSuccess;
  hipLaunchKernelGGL(store_flag_kernel, dim3(nrb), dim3(kStoreBlock), 0, s, a);
  hipLaunchKernelGGL(store_count_kernel, dim3(nrb, a.nc), dim3(kStoreBlock), 0, s, a);
  hipLaunchKernelGGL(store_scan_kernel, dim3(1), dim3(kScanThreads), 0, s, a.bcount, nrb * a.nc);
  hipLaunchKernelGGL(store_write_kernel, dim3(nrb, a.nc), dim3(kStoreBlock), 0, s, a);
  return hipGetLastError();
}

}  // namespace d2g
