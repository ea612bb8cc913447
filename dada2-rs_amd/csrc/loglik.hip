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
// wavefronts per 64 pairs, lanes 0..15 own them), so the per-pair walks read
// LDS only.  Positions past a pair's end index an extra table entry: 0.0 in
// the forward's log table (adding +0.0 leaves the f64 sum unchanged) and a
// discarded bin in the backward's histogram.
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
  const uint32_t lane = __lane_id(), wave = threadIdx.x / 64u;
  const uint32_t ncol = a.ncol, ne = 16u * ncol;
  for (uint32_t i = threadIdx.x; i <= ne; i += blockDim.x) le[i] = i < ne ? (double)logE[i] : 0.0;
  uint8_t* rows = reinterpret_cast<uint8_t*>(le + ne + 1) + wave * kStPairsW * kStRow;
  uint8_t* myb = rows + min(lane, (uint32_t)kStPairsW - 1u) * kStRow;
  const uint32_t n_g = a.ctr->n_gapless, n = n_g + a.ctr->n_dp;
  const bool quals = a.use_quals && a.ts.u.qual != nullptr;
  __syncthreads();
  double acc = 0.0;
  for (uint32_t base = blockIdx.x * (uint32_t)kStTile; base < n; base += gridDim.x * (uint32_t)kStTile) {
    const TaskRow me = task_row(a.ts, n_g, n, base + wave * kStPairsW + lane,
                                lane < (uint32_t)kStPairsW);
    const bool any_gl = __builtin_amdgcn_ballot_w64(me.gl) != 0;
    double ll_lam = 0.0;
    uint32_t ham = 0;
    SubCursor cur = sub_cursor(me);
    for (uint32_t s0 = 0; __builtin_amdgcn_ballot_w64(s0 < me.lq); s0 += kStSeg) {
      stage_segment(a.ts, me, s0, lane, quals, any_gl, rows, myb, ham, cur);
      wave_lds_sync();
      const uint32_t seg_n = me.lq > s0 ? min(me.lq - s0, (uint32_t)kStSeg) : 0u;
      if (lane < (uint32_t)kStPairsW) {
        const uint32_t* my = reinterpret_cast<const uint32_t*>(myb);
        for (uint32_t i = 0; __builtin_amdgcn_ballot_w64(i < seg_n); i += 32) {
          uint32_t w16[16];
#pragma unroll
          for (int h = 0; h < 16; ++h) w16[h] = my[(i >> 1) + (uint32_t)h];
          double f[32];
#pragma unroll
          for (int e = 0; e < 32; ++e) f[e] = le[code_index((w16[e >> 1] >> (16 * (e & 1))) & 0xFFFFu, ncol)];
#pragma unroll
          for (int e = 0; e < 32; ++e) ll_lam += f[e];
        }
      }
      wave_lds_sync();
    }
    if (me.live) {
      const double av = (double)a.ts.u.reads[me.u_idx];
      const double E = (double)a.ts.u.reads[me.c_idx] * exp(ll_lam);
      a.w[me.t] = av - E;
      acc += av * log(E) - E - lgamma(av + 1.0);
    }
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
  extern __shared__ double hist[];  // 16 ncol bins and a discarded one, then the rows
  const uint32_t lane = __lane_id(), wave = threadIdx.x / 64u;
  const uint32_t ncol = a.ncol, ne = 16u * ncol;
  for (uint32_t i = threadIdx.x; i <= ne; i += blockDim.x) hist[i] = 0.0;
  uint8_t* rows = reinterpret_cast<uint8_t*>(hist + ne + 1) + wave * kStPairsW * kStRow;
  uint8_t* myb = rows + min(lane, (uint32_t)kStPairsW - 1u) * kStRow;
  const uint32_t n_g = a.ctr->n_gapless, n = n_g + a.ctr->n_dp;
  const bool quals = a.use_quals && a.ts.u.qual != nullptr;
  __syncthreads();
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
      const uint32_t seg_n = me.lq > s0 ? min(me.lq - s0, (uint32_t)kStSeg) : 0u;
      if (lane < (uint32_t)kStPairsW && me.live) {
        const uint32_t* my = reinterpret_cast<const uint32_t*>(myb);
        for (uint32_t i = 0; i < seg_n; i += 2) {
          const uint32_t w2 = my[i >> 1];
          atomicAdd(&hist[code_index(w2 & 0xFFFFu, ncol)], wt);
          atomicAdd(&hist[code_index(w2 >> 16, ncol)], wt);  // past the end: the discarded bin
        }
      }
      wave_lds_sync();
    }
  }
  __syncthreads();
  double* out = a.block_grad + (size_t)blockIdx.x * ne;
  for (uint32_t i = threadIdx.x; i < ne; i += blockDim.x) out[i] = hist[i];
}

// Sums per-block partials in block order: out[i] = sum_b part[b * m + i].
__global__ void reduce_blocks_kernel(const double* part, uint32_t nblocks, uint32_t m, double* out) {
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < m; i += gridDim.x * blockDim.x) {
    double s = 0.0;
    for (uint32_t b = 0; b < nblocks; ++b) s += part[(size_t)b * m + i];
    out[i] = s;
  }
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
  hipLaunchKernelGGL(reduce_blocks_kernel, dim3(1), dim3(64), 0, st, (const double*)s.block_part,
                     s.grid, 1u, loss);
  return hipGetLastError();
}

hipError_t launch_loglik_bwd(const LoglikArgs& s, double* grad, hipStream_t st) {
  LoglikDev d = make_dev(s);
  const uint32_t m = 16u * s.ncol;
  const size_t lds = ((size_t)m + 1) * sizeof(double) + (size_t)kStTile * kStRow;
  hipError_t e = lds_attr((const void*)loglik_bwd_kernel, lds);
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(loglik_bwd_kernel, dim3(s.grid), dim3(64 * kStWaves), lds, st, d);
  e = hipGetLastError();
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(reduce_blocks_kernel, dim3((m + 255) / 256), dim3(256), 0, st,
                     (const double*)s.block_part, s.grid, m, grad);
  return hipGetLastError();
}

}  // namespace d2g
