// shuffle.hip — the best-cluster scan of a shuffle pass (b_shuffle2,
// cluster.rs:615-700) on the device.
//
// Each raw moves to the cluster whose stored comparison gives it the most
// expected reads, lambda x cluster reads, the first such cluster in
// ascending order on ties (the scan's strict `>` from -inf).  The stored
// comparisons live on the device, appended after each compare's store.  A
// pass is three kernels:
//   1. each raw's largest lambda x reads, as an atomic max of the f64 bits
//      (non-negative doubles order like their bit patterns); each center's
//      cluster is set here (a bud moves the center into its new cluster);
//   2. the smallest (cluster, comparison index) reaching it (atomic min);
//   3. over raws: the movers (best cluster not their own, and not a
//      center), compacted into a list with the winning comparison's lambda
//      and hamming; the raw's cluster is updated here and the per-raw
//      scratch reset for the next pass.
// The host applies the moves in the loop's order (cluster, then descending
// position), so membership order matches the host scan's exactly.
#include <algorithm>

#include "d2g_device.h"
#include "d2g_launch.h"

namespace d2g {

namespace {

__device__ __forceinline__ unsigned long long e_bits(double lambda, double reads) {
  return (unsigned long long)__double_as_longlong(lambda * reads);
}

__global__ __launch_bounds__(256) void shuffle_max_kernel(ShuffleArgs a) {
  const uint32_t tid = blockIdx.x * blockDim.x + threadIdx.x;
  if (tid == 0) a.moves[0].r = 0;  // the move count
  for (uint32_t c = tid; c < a.n_cl; c += gridDim.x * blockDim.x)
    if (a.ccenter[c] < a.n_raw) a.where[a.ccenter[c]] = c;
  for (uint32_t k = tid; k < a.n_comp; k += gridDim.x * blockDim.x) {
    const ShuffleComp cp = a.comps[k];
    atomicMax(&a.best_bits[cp.raw], e_bits(cp.lambda, a.creads[cp.cl]));
  }
}

__global__ __launch_bounds__(256) void shuffle_first_kernel(ShuffleArgs a) {
  for (uint32_t k = blockIdx.x * blockDim.x + threadIdx.x; k < a.n_comp;
       k += gridDim.x * blockDim.x) {
    const ShuffleComp cp = a.comps[k];
    if (e_bits(cp.lambda, a.creads[cp.cl]) == a.best_bits[cp.raw])
      atomicMin(&a.best_key[cp.raw], (unsigned long long)cp.cl << 32 | k);
  }
}

__global__ __launch_bounds__(256) void shuffle_moves_kernel(ShuffleArgs a) {
  const uint32_t lane = __lane_id();
  // a block-uniform trip count: every lane takes part in the append
  for (uint32_t r0 = blockIdx.x * blockDim.x; r0 < a.n_raw; r0 += gridDim.x * blockDim.x) {
    const uint32_t r = r0 + threadIdx.x;
    bool move = false;
    unsigned long long key = ~0ull;
    if (r < a.n_raw) {
      key = a.best_key[r];
      const uint32_t w = a.where[r];
      move = key != ~0ull && (uint32_t)(key >> 32) != w && a.ccenter[w] != r;
    }
    const uint64_t mask = __builtin_amdgcn_ballot_w64(move);
    uint32_t base = 0;
    if (mask) {
      const int leader = __builtin_ffsll((long long)mask) - 1;
      if ((int)lane == leader)
        base = atomicAdd(&a.moves[0].r, (uint32_t)__builtin_popcountll(mask));
      base = (uint32_t)__shfl((int)base, leader);
    }
    if (move) {
      const uint32_t slot = 1u + base + (uint32_t)__builtin_popcountll(mask & ((1ull << lane) - 1ull));
      const ShuffleComp cp = a.comps[(uint32_t)key];
      ShuffleMove m;
      m.r = r;
      m.to = cp.cl;
      m.ham = cp.ham;
      m._pad = 0;
      m.lambda = cp.lambda;
      a.moves[slot] = m;
      a.where[r] = cp.cl;
    }
    if (r < a.n_raw) {
      a.best_bits[r] = 0ull;
      a.best_key[r] = ~0ull;
    }
  }
}

__global__ __launch_bounds__(256) void shuffle_init_kernel(ShuffleArgs a) {
  for (uint32_t r = blockIdx.x * blockDim.x + threadIdx.x; r < a.n_raw;
       r += gridDim.x * blockDim.x) {
    a.best_bits[r] = 0ull;
    a.best_key[r] = ~0ull;
    a.where[r] = 0;
  }
}

uint32_t grid_for(uint32_t n) { return std::max(1u, std::min((n + 255u) / 256u, 4096u)); }

}  // namespace

hipError_t launch_shuffle_init(const ShuffleArgs& a, hipStream_t s) {
  hipLaunchKernelGGL(shuffle_init_kernel, dim3(grid_for(a.n_raw)), dim3(256), 0, s, a);
  return hipGetLastError();
}

hipError_t launch_shuffle_pass(const ShuffleArgs& a, hipStream_t s) {
  const dim3 gc(grid_for(a.n_comp)), gr(grid_for(a.n_raw)), b(256);
  hipLaunchKernelGGL(shuffle_max_kernel, dim3(std::max(gc.x, grid_for(a.n_cl))), b, 0, s, a);
  hipLaunchKernelGGL(shuffle_first_kernel, gc, b, 0, s, a);
  hipLaunchKernelGGL(shuffle_moves_kernel, gr, b, 0, s, a);
  return hipGetLastError();
}

}  // namespace d2g
