// lambda.hip — compute_lambda (pval.rs:253-320) for every aligned pair.
//
// lambda = prod over query positions pos = 0..len-1, left to right from 1.0,
// of err[t(pos) * ncol + q(pos)], where t(pos) = 5*(nt-1) (the query base on
// the diagonal) unless an al2subs substitution maps the query position to a
// center base (pval.rs:294-309).  The f64 multiplications keep that exact
// order, so results are bit-identical to the reference.
//
// Gapless pairs (align_gapless_with_buf, nwalign.rs:645-655) pair positions
// directly: substitutions are the positions < min(len) where bases differ.
// DP pairs read the substitution list the traceback wrote: (query position,
// center position) records.
//
// One lane per pair (the product is one ordered chain), fed from LDS rows
// staged as d2g_stage.h describes, so that no global load sits on the
// chain.  Positions past a pair's end index an extra 1.0 entry after the
// error matrix (x * 1.0 = x for every f64), so the multiplications run
// without branches.
#include "d2g_stage.h"

namespace d2g {

struct LambdaDev {
  TaskSet ts;
  const double* err;
  uint32_t ncol;
  int use_quals;
  Counters* ctr;
  double* out_lambda;
  uint32_t* out_hamming;
  uint8_t* out_skipped;
};

__global__ __launch_bounds__(64 * kStWaves) void lambda_kernel(LambdaDev a) {
  extern __shared__ double err_lds[];
  const uint32_t lane = __lane_id();
  const uint32_t wave = threadIdx.x / 64u;
  const uint32_t ncol = a.ncol;
  const uint32_t ne = 16u * ncol;
  const uint32_t n_g = a.ctr->n_gapless;
  const uint32_t n = n_g + a.ctr->n_dp;
  // the error matrix and, after it, the 1.0 entry
  for (uint32_t i0 = 0; i0 <= ne; i0 += 4 * 64 * kStWaves) {
    double v[4];
#pragma unroll
    for (int k = 0; k < 4; ++k) v[k] = a.err[min(i0 + (uint32_t)k * 64u * kStWaves + threadIdx.x, ne - 1u)];
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      const uint32_t i = i0 + (uint32_t)k * 64u * kStWaves + threadIdx.x;
      if (i < ne) err_lds[i] = v[k];
      if (i == ne) err_lds[i] = 1.0;
    }
  }
  uint8_t* rows = reinterpret_cast<uint8_t*>(err_lds + ne + 1) + wave * kStPairsW * kStRow;
  uint8_t* myb = rows + min(lane, (uint32_t)kStPairsW - 1u) * kStRow;
  const bool quals = a.use_quals && a.ts.u.qual != nullptr;
  __syncthreads();
  for (uint32_t base = blockIdx.x * (uint32_t)kStTile; base < n; base += gridDim.x * (uint32_t)kStTile) {
    const TaskRow me = task_row(a.ts, n_g, n, base + wave * kStPairsW + lane,
                                lane < (uint32_t)kStPairsW);
    if (me.bad) atomicOr(&a.ctr->err_flags, kErrTraceback);
    const bool any_gl = __builtin_amdgcn_ballot_w64(me.gl) != 0;
    double lam = 1.0;
    uint32_t ham = 0;
    SubCursor cur = sub_cursor(me);
    for (uint32_t s0 = 0; __builtin_amdgcn_ballot_w64(s0 < me.lq); s0 += kStSeg) {
      stage_segment(a.ts, me, s0, lane, quals, any_gl, rows, myb, ham, cur);
      // the wavefront's staged rows are read by other lanes of the same
      // wavefront (segment counts differ between wavefronts, so no
      // workgroup barrier here)
      wave_lds_sync();
      // ---- the ordered product over the segment (owning lanes) ----
      const uint32_t seg_n = me.lq > s0 ? min(me.lq - s0, (uint32_t)kStSeg) : 0u;
      if (lane < (uint32_t)kStPairsW) {
        const uint32_t* my = reinterpret_cast<const uint32_t*>(myb);
        for (uint32_t i = 0; __builtin_amdgcn_ballot_w64(i < seg_n); i += 32) {
          uint32_t w16[16];
#pragma unroll
          for (int h = 0; h < 16; ++h) w16[h] = my[(i >> 1) + (uint32_t)h];
          double f[32];
#pragma unroll
          for (int e = 0; e < 32; ++e)
            f[e] = err_lds[code_index((w16[e >> 1] >> (16 * (e & 1))) & 0xFFFFu, ncol)];
#pragma unroll
          for (int e = 0; e < 32; ++e) lam *= f[e];
        }
      }
      wave_lds_sync();  // rows are restaged for the next segment
    }
    if (me.live) {
      a.out_lambda[me.out] = lam;
      a.out_hamming[me.out] = me.gl ? ham : me.cnt;
      a.out_skipped[me.out] = 0;
    }
  }
}

hipError_t launch_lambda(const LambdaArgs& s, hipStream_t st) {
  LambdaDev d;
  d.ts.u = *s.u;
  d.ts.gl = s.gapless_tasks;
  d.ts.dp = s.dp_tasks;
  d.ts.sub_pool = s.sub_pool;
  d.ts.sub_stride = s.sub_stride;
  d.ts.sub_cnt = s.sub_cnt;
  d.err = s.err;
  d.ncol = s.ncol;
  d.use_quals = s.use_quals;
  d.ctr = s.ctr;
  d.out_lambda = s.out_lambda;
  d.out_hamming = s.out_hamming;
  d.out_skipped = s.out_skipped;
  const size_t lds = ((size_t)16 * s.ncol + 1) * sizeof(double) + (size_t)kStTile * kStRow;
  if (lds > 48 * 1024) {
    hipError_t e = hipFuncSetAttribute((const void*)lambda_kernel,
                                       hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    if (e != hipSuccess) return e;
  }
  hipLaunchKernelGGL(lambda_kernel, dim3(s.grid), dim3(64 * kStWaves), lds, st, d);
  return hipGetLastError();
}

}  // namespace d2g
