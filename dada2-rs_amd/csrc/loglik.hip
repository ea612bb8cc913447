// loglik.hip — the Poisson log-likelihood of the abundances over the aligned
// pairs of the last compare, and its gradient with respect to the 16 x ncol
// log error matrix.
//
// For an aligned pair j = (center c, unique u):
//   log lambda_j = sum over query positions of logE[t(pos)][q(pos)], with the
//                  transitions of compute_lambda (pval.rs:294-309): t = 5 (nt-1)
//                  on the diagonal, 4 (nt0-1) + (nt-1) at a substitution;
//   E_j          = reads[c] * exp(log lambda_j);
//   ll_j         = a_j log E_j - E_j - lgamma(a_j + 1), a_j = reads[u].
// loss = sum_j ll_j and d loss / d logE[t][q] = sum_j (a_j - E_j) count_j[t][q].
// The alignment is fixed (the compare's gapless pairs pair positions; DP
// pairs carry their al2subs list); count_j is recomputed by walking the pair
// again in the backward kernel, never stored.  The parameter is FP32; sums
// are FP64.
// The pairs are staged into LDS rows as d2g_stage.h describes (four
// wavefronts per 64 pairs; each wavefront stages its 16 pairs segment by
// segment), and then every staged pair is taken by all 64 lanes, four
// positions per lane:
//   forward:  each lane sums its four table entries of every staged pair in
//             FP64; one transposing reduction over the 16 pairs (pairwise
//             joins over lane bits 2..5, then 0 and 1) leaves pair j's
//             segment total on lanes 4j .. 4j + 3, the same tree for every
//             pair, and the pair's running sum takes it, so log lambda_j does
//             not depend on scheduling;
//   backward: per staged pair the lanes count the segment's table indices
//             into an integer histogram (u32 LDS atomics: order-free); the
//             lane whose increment found a bin at 0 owns that bin and, once
//             every lane has counted, applies its one FP64 multiply-add,
//             w_j x count, to the wavefront's gradient row and clears it.
//             Pairs and segments are visited in a fixed order and the
//             workgroup adds its wavefronts' rows in wavefront order, so the
//             gradient's bits repeat run to run.
// Positions past a pair's end index an extra table entry: 0.0 in the
// forward's log table and a discarded count in the backward.
#include "d2g_stage.h"

namespace d2g {

struct LoglikDev {
  TaskSet ts;
  uint32_t ncol;
  int use_quals;
  const Counters* ctr;
  double* w;            // per task: a_j - E_j
  double* block_loss;   // per block
  double* block_grad;   // per block: 16 * ncol
};

__global__ __launch_bounds__(64 * kStWaves) void loglik_fwd_kernel(LoglikDev a, const float* logE) {
  extern __shared__ double le[];  // 16 ncol entries and 0.0 after them, then the rows
  __shared__ double red[kStWaves];
  __shared__ double llp[kStWaves][kStPairsW];  // each staged pair's running sum
  const uint32_t lane = __lane_id(), wave = threadIdx.x / 64u;
  const uint32_t ncol = a.ncol, ne = 16u * ncol;
  for (uint32_t i = threadIdx.x; i <= ne; i += blockDim.x) le[i] = i < ne ? (double)logE[i] : 0.0;
  uint8_t* rows = reinterpret_cast<uint8_t*>(le + ne + 1) + wave * kStPairsW * kStRow;
  uint8_t* myb = rows + min(lane, (uint32_t)kStPairsW - 1u) * kStRow;
  const uint32_t n_g = a.ctr->n_gapless, n = n_g + a.ctr->n_dp;
  const bool quals = a.use_quals && a.ts.u.qual != nullptr;
  __syncthreads();
  double acc = 0.0;  // lanes 0 .. 15: their pairs, in order
  for (uint32_t base = blockIdx.x * (uint32_t)kStTile; base < n; base += gridDim.x * (uint32_t)kStTile) {
    const TaskRow me = task_row(a.ts, n_g, n, base + wave * kStPairsW + lane,
                                lane < (uint32_t)kStPairsW);
    const bool any_gl = __builtin_amdgcn_ballot_w64(me.gl) != 0;
    if (lane < (uint32_t)kStPairsW) llp[wave][lane] = 0.0;
    uint32_t ham = 0;
    SubCursor cur = sub_cursor(me);
    for (uint32_t s0 = 0; __builtin_amdgcn_ballot_w64(s0 < me.lq); s0 += kStSeg) {
      stage_segment(a.ts, me, s0, lane, quals, any_gl, rows, myb, ham, cur);
      wave_lds_sync();
      double part[kStPairsW];
#pragma unroll
      for (int pr = 0; pr < kStPairsW; ++pr) {
        // eight pairs' table reads in flight at a time
        if (pr == kStPairsW / 2) __builtin_amdgcn_sched_barrier(0);
        // positions past the pair's end carry the 0.0 entry
        const uint32_t w2 = *reinterpret_cast<const uint32_t*>(rows + pr * kStRow + 4u * lane);
        const uint32_t w3 = *reinterpret_cast<const uint32_t*>(rows + pr * kStRow + 4u * lane + 256u);
        part[pr] = le[code_index(w2 & 0xFFFFu, ncol)] + le[code_index(w2 >> 16, ncol)];
        part[pr] += le[code_index(w3 & 0xFFFFu, ncol)] + le[code_index(w3 >> 16, ncol)];
      }
      // reduce-scatter: at lane bit b, pairs (2i, 2i + 1) become slot i,
      // lanes with bit b clear holding pair 2i, the others 2i + 1
      static_assert(kStPairsW == 16, "16 staged pairs map onto lane bits 2..5");
#pragma unroll
      for (int step = 0; step < 4; ++step) {
        const int bit = 4 << step;
        const bool hi = (lane & (uint32_t)bit) != 0;
#pragma unroll
        for (int i = 0; i < kStPairsW / 2; ++i) {
          if (i >= (kStPairsW >> (step + 1))) break;
          const double give = hi ? part[2 * i] : part[2 * i + 1];
          const double keep = hi ? part[2 * i + 1] : part[2 * i];
          part[i] = keep + __shfl_xor(give, bit);
        }
      }
      double segt = part[0];
      segt += __shfl_xor(segt, 1);
      segt += __shfl_xor(segt, 2);
      if ((lane & 3u) == 0) llp[wave][lane >> 2] += segt;
      wave_lds_sync();
    }
    if (me.live) {
      const double av = (double)a.ts.u.reads[me.u_idx];
      const double E = (double)a.ts.u.reads[me.c_idx] * exp(llp[wave][lane]);
      a.w[me.t] = av - E;
      acc += av * log(E) - E - lgamma(av + 1.0);
    }
    wave_lds_sync();
  }
  for (int off = 32; off > 0; off >>= 1) acc += __shfl_xor(acc, off);
  if (lane == 0) red[wave] = acc;
  __syncthreads();
  if (threadIdx.x == 0) {
    double s = 0.0;
    for (int i = 0; i < kStWaves; ++i) s += red[i];
    a.block_loss[blockIdx.x] = s;
  }
}

__global__ __launch_bounds__(64 * kStWaves) void loglik_bwd_kernel(LoglikDev a) {
  // per wavefront: 16 ncol gradient sums; then per wavefront 16 ncol + 1
  // counts (the last one discarded); then the rows
  extern __shared__ double grad[];
  const uint32_t lane = __lane_id(), wave = threadIdx.x / 64u;
  const uint32_t ncol = a.ncol, ne = 16u * ncol;
  double* g = grad + (size_t)wave * ne;
  uint32_t* cnt0 = reinterpret_cast<uint32_t*>(grad + (size_t)kStWaves * ne);
  uint32_t* cnt = cnt0 + (size_t)wave * (ne + 1);
  uint8_t* rows = reinterpret_cast<uint8_t*>(cnt0 + (size_t)kStWaves * (ne + 1) + 1) + wave * kStPairsW * kStRow;
  rows = reinterpret_cast<uint8_t*>((reinterpret_cast<uintptr_t>(rows) + 7) & ~(uintptr_t)7);
  uint8_t* myb = rows + min(lane, (uint32_t)kStPairsW - 1u) * kStRow;
  for (uint32_t i = lane; i <= ne; i += 64u) {
    if (i < ne) g[i] = 0.0;
    cnt[i] = 0u;
  }
  const uint32_t n_g = a.ctr->n_gapless, n = n_g + a.ctr->n_dp;
  const bool quals = a.use_quals && a.ts.u.qual != nullptr;
  wave_lds_sync();
  for (uint32_t base = blockIdx.x * (uint32_t)kStTile; base < n; base += gridDim.x * (uint32_t)kStTile) {
    const TaskRow me = task_row(a.ts, n_g, n, base + wave * kStPairsW + lane,
                                lane < (uint32_t)kStPairsW);
    const bool any_gl = __builtin_amdgcn_ballot_w64(me.gl) != 0;
    const double wt = me.live ? a.w[me.t] : 0.0;
    uint32_t ham = 0;
    SubCursor cur = sub_cursor(me);
    for (uint32_t s0 = 0; __builtin_amdgcn_ballot_w64(s0 < me.lq); s0 += kStSeg) {
      stage_segment(a.ts, me, s0, lane, quals, any_gl, rows, myb, ham, cur);
      wave_lds_sync();
      for (uint32_t pr = 0; pr < (uint32_t)kStPairsW; ++pr) {
        const double wp = __shfl(wt, (int)pr);
        if (!__builtin_amdgcn_readfirstlane(__shfl((int)me.live, (int)pr))) continue;  // an empty slot
        const uint32_t w2 = *reinterpret_cast<const uint32_t*>(rows + pr * kStRow + 4u * lane);
        const uint32_t w3 = *reinterpret_cast<const uint32_t*>(rows + pr * kStRow + 4u * lane + 256u);
        // past the pair's end: index 16 ncol, the discarded count
        const uint32_t ix[4] = {code_index(w2 & 0xFFFFu, ncol), code_index(w2 >> 16, ncol),
                                code_index(w3 & 0xFFFFu, ncol), code_index(w3 >> 16, ncol)};
        uint32_t old[4];
#pragma unroll
        for (int e = 0; e < 4; ++e) old[e] = atomicAdd(&cnt[ix[e]], 1u);
        wave_lds_sync();
        // the bin's owner (its first increment) applies the pair's count
#pragma unroll
        for (int e = 0; e < 4; ++e) {
          if (old[e] == 0u && ix[e] < ne) {
            g[ix[e]] += wp * (double)cnt[ix[e]];
            cnt[ix[e]] = 0u;
          }
        }
        wave_lds_sync();
      }
    }
  }
  __syncthreads();
  for (uint32_t i = threadIdx.x; i < ne; i += blockDim.x) {  // bin-major partials
    double s = grad[i];
    for (int w = 1; w < kStWaves; ++w) s += grad[(size_t)w * ne + i];
    a.block_grad[(size_t)i * gridDim.x + blockIdx.x] = s;
  }
}

// out[i] = sum over blocks b of part[i * nblocks + b], one workgroup per i:
// thread t adds blocks t, t + 256, ... in order, then a fixed tree over the
// threads, so the sum's bits do not depend on scheduling.
__global__ __launch_bounds__(256) void reduce_blocks_kernel(const double* part, uint32_t nblocks,
                                                            double* out) {
  __shared__ double sh[256];
  const uint32_t i = blockIdx.x, t = threadIdx.x;
  const double* row = part + (size_t)i * nblocks;
  double s = 0.0;
  for (uint32_t b = t; b < nblocks; b += 256u) s += row[b];
  sh[t] = s;
  __syncthreads();
  for (uint32_t h = 128; h > 0; h >>= 1) {
    if (t < h) sh[t] += sh[t + h];
    __syncthreads();
  }
  if (t == 0) out[i] = sh[0];
}

static LoglikDev make_dev(const LoglikArgs& s) {
  LoglikDev d;
  d.ts.u = *s.u;
  d.ts.gl = s.gapless_tasks;
  d.ts.dp = s.dp_tasks;
  d.ts.sub_pool = s.sub_pool;
  d.ts.sub_stride = s.sub_stride;
  d.ts.sub_cnt = s.sub_cnt;
  d.ncol = s.ncol;
  d.use_quals = s.use_quals;
  d.ctr = s.ctr;
  d.w = s.w;
  d.block_loss = s.block_part;
  d.block_grad = s.block_part;
  return d;
}

static hipError_t lds_attr(const void* f, size_t lds) {
  if (lds <= 48 * 1024) return hipSuccess;
  return hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
}

hipError_t launch_loglik_fwd(const LoglikArgs& s, const float* logE, double* loss,
                             hipStream_t st) {
  LoglikDev d = make_dev(s);
  const size_t lds = ((size_t)16 * s.ncol + 1) * sizeof(double) + (size_t)kStTile * kStRow;
  hipError_t e = lds_attr((const void*)loglik_fwd_kernel, lds);
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(loglik_fwd_kernel, dim3(s.grid), dim3(64 * kStWaves), lds, st, d, logE);
  e = hipGetLastError();
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(reduce_blocks_kernel, dim3(1), dim3(256), 0, st, (const double*)s.block_part,
                     s.grid, loss);
  return hipGetLastError();
}

hipError_t launch_loglik_bwd(const LoglikArgs& s, double* grad, hipStream_t st) {
  LoglikDev d = make_dev(s);
  const uint32_t m = 16u * s.ncol;
  const size_t lds = (size_t)kStWaves * m * sizeof(double) +
                     ((size_t)kStWaves * (m + 1) + 1) * sizeof(uint32_t) + 8 +
                     (size_t)kStTile * kStRow;
  hipError_t e = lds_attr((const void*)loglik_bwd_kernel, lds);
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(loglik_bwd_kernel, dim3(s.grid), dim3(64 * kStWaves), lds, st, d);
  e = hipGetLastError();
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(reduce_blocks_kernel, dim3(m), dim3(256), 0, st,
                     (const double*)s.block_part, s.grid, grad);
  return hipGetLastError();
}

}  // namespace d2g
