// lambda.hip — compute_lambda (pval.rs:253-320) for every aligned pair.
//
// lambda = prod over query positions pos = 0..len-1, left to right from 1.0,
// of err[t(pos) * ncol + q(pos)], where t(pos) = 5*(nt-1) (the query base on
// the diagonal) unless an al2subs substitution maps the query position to a
// center base (pval.rs:294-309).  The f64 multiplications keep that exact
// order, so results are bit-identical to the reference.  One lane per pair;
// the error matrix sits in LDS.
//
// Gapless pairs (align_gapless_with_buf, nwalign.rs:645-655) pair positions
// directly: substitutions are the positions < min(len) where bases differ.
// DP pairs read the substitution list the traceback wrote: (query position,
// center position) records in descending query position.
#include "d2g_launch.h"

namespace d2g {

struct LambdaDev {
  UniqueDev u;
  const GaplessTask* gl;
  const DpTask* dp;
  const uint32_t* sub_pool;
  uint32_t sub_stride;
  const uint32_t* sub_cnt;
  const double* err;
  uint32_t ncol;
  int use_quals;
  Counters* ctr;
  double* out_lambda;
  uint32_t* out_hamming;
  uint8_t* out_skipped;
};

__global__ __launch_bounds__(256) void lambda_kernel(LambdaDev a) {
  extern __shared__ double err_lds[];
  const uint32_t ne = 16u * a.ncol;
  for (uint32_t i = threadIdx.x; i < ne; i += blockDim.x) err_lds[i] = a.err[i];
  __syncthreads();
  const uint32_t n_g = a.ctr->n_gapless;
  const uint32_t n_d = a.ctr->n_dp;
  const uint32_t ncol = a.ncol;
  const bool quals = a.use_quals && a.u.qual != nullptr;
  for (uint32_t t = blockIdx.x * blockDim.x + threadIdx.x; t < n_g + n_d;
       t += gridDim.x * blockDim.x) {
    double lam = 1.0;
    uint32_t ham = 0, out;
    if (t < n_g) {
      const GaplessTask task = a.gl[t];
      out = task.out_idx;
      const uint4* q16 = reinterpret_cast<const uint4*>(a.u.seq + a.u.off[task.u_idx]);
      const uint4* qq16 = quals ? reinterpret_cast<const uint4*>(a.u.qual + a.u.off[task.u_idx]) : nullptr;
      const uint4* c16 = reinterpret_cast<const uint4*>(a.u.seq + a.u.off[task.c_idx]);
      const uint32_t lq = a.u.len[task.u_idx], lc = a.u.len[task.c_idx];
      for (uint32_t base = 0; base < lq; base += 16) {
        const uint4 qv = q16[base >> 4];
        // rows are padded to 16 bytes; a longer query reads no center bytes
        const uint4 cv = base < lc ? c16[base >> 4] : make_uint4(0, 0, 0, 0);
        const uint4 qu = qq16 ? qq16[base >> 4] : make_uint4(0, 0, 0, 0);
        const uint32_t qw[4] = {qv.x, qv.y, qv.z, qv.w};
        const uint32_t cw[4] = {cv.x, cv.y, cv.z, cv.w};
        const uint32_t uw[4] = {qu.x, qu.y, qu.z, qu.w};
#pragma unroll
        for (int b = 0; b < 16; ++b) {
          const uint32_t pos = base + (uint32_t)b;
          if (pos < lq) {
            const uint32_t qn = (qw[b >> 2] >> (8 * (b & 3))) & 0xFFu;
            const uint32_t cn = (cw[b >> 2] >> (8 * (b & 3))) & 0xFFu;
            const uint32_t qi = (uw[b >> 2] >> (8 * (b & 3))) & 0xFFu;
            uint32_t tr = 5u * (qn - 1u);
            if (pos < lc && cn != qn) {  // gapless pairing, nwalign.rs:645-655
              tr = 4u * (cn - 1u) + (qn - 1u);
              ++ham;
            }
            lam *= err_lds[tr * ncol + qi];
          }
        }
      }
    } else {
      const uint32_t di = t - n_g;
      const DpTask task = a.dp[di];
      out = task.out_idx;
      const uint4* q16 = reinterpret_cast<const uint4*>(a.u.seq + a.u.off[task.u_idx]);
      const uint4* qq16 = quals ? reinterpret_cast<const uint4*>(a.u.qual + a.u.off[task.u_idx]) : nullptr;
      const uint32_t lq = a.u.len[task.u_idx];
      const uint32_t cnt_raw = a.sub_cnt[di];
      if (cnt_raw & 0x80000000u) atomicOr(&a.ctr->err_flags, kErrTraceback);
      const uint32_t cnt = cnt_raw & 0x7FFFFFFFu;
      const uint32_t* subs = a.sub_pool + (size_t)di * a.sub_stride;
      const uint8_t* cseq = a.u.seq + a.u.off[task.c_idx];
      // records are (query position << 16 | center position), descending
      int idx = (int)cnt - 1;
      uint32_t next = idx >= 0 ? subs[idx] : 0xFFFFFFFFu;
      for (uint32_t base = 0; base < lq; base += 16) {
        const uint4 qv = q16[base >> 4];
        const uint4 qu = qq16 ? qq16[base >> 4] : make_uint4(0, 0, 0, 0);
        const uint32_t qw[4] = {qv.x, qv.y, qv.z, qv.w};
        const uint32_t uw[4] = {qu.x, qu.y, qu.z, qu.w};
#pragma unroll
        for (int b = 0; b < 16; ++b) {
          const uint32_t pos = base + (uint32_t)b;
          if (pos < lq) {
            const uint32_t qn = (qw[b >> 2] >> (8 * (b & 3))) & 0xFFu;
            const uint32_t qi = (uw[b >> 2] >> (8 * (b & 3))) & 0xFFu;
            uint32_t tr = 5u * (qn - 1u);
            if ((next >> 16) == pos) {
              const uint32_t cn = cseq[next & 0xFFFFu];
              tr = 4u * (cn - 1u) + (qn - 1u);
              --idx;
              next = idx >= 0 ? subs[idx] : 0xFFFFFFFFu;
            }
            lam *= err_lds[tr * ncol + qi];
          }
        }
      }
      ham = cnt;
    }
    a.out_lambda[out] = lam;
    a.out_hamming[out] = ham;
    a.out_skipped[out] = 0;
  }
}

hipError_t launch_lambda(const LambdaArgs& s, hipStream_t st) {
  LambdaDev d;
  d.u = *s.u;
  d.gl = s.gapless_tasks;
  d.dp = s.dp_tasks;
  d.sub_pool = s.sub_pool;
  d.sub_stride = s.sub_stride;
  d.sub_cnt = s.sub_cnt;
  d.err = s.err;
  d.ncol = s.ncol;
  d.use_quals = s.use_quals;
  d.ctr = s.ctr;
  d.out_lambda = s.out_lambda;
  d.out_hamming = s.out_hamming;
  d.out_skipped = s.out_skipped;
  size_t lds = (size_t)16 * s.ncol * sizeof(double);
  if (lds > 48 * 1024) {
    hipError_t e = hipFuncSetAttribute((const void*)lambda_kernel,
                                       hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    if (e != hipSuccess) return e;
  }
  hipLaunchKernelGGL(lambda_kernel, dim3(s.grid), dim3(256), lds, st, d);
  return hipGetLastError();
}

}  // namespace d2g
