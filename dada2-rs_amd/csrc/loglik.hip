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
#include "d2g_launch.h"

namespace d2g {

struct LoglikDev {
  UniqueDev u;
  const GaplessTask* gl;
  const DpTask* dp;
  const uint32_t* sub_pool;
  uint32_t sub_stride;
  const uint32_t* sub_cnt;
  uint32_t ncol;
  int use_quals;
  const Counters* ctr;
  double* w;            // per task: a_j - E_j
  double* block_loss;   // per block
  double* block_grad;   // per block: 16 * ncol
};

// Walks the query positions of task t, calling f(transition row, quality).
// Gapless tasks: positional pairing (nwalign.rs:645-655); DP tasks: the
// substitution records (query pos << 16 | center pos), descending.
template <class F>
__device__ __forceinline__ void walk_task(const LoglikDev& a, uint32_t t, uint32_t n_g, bool quals,
                                          uint32_t* u_idx, uint32_t* c_idx, F&& f) {
  if (t < n_g) {
    const GaplessTask task = a.gl[t];
    *u_idx = task.u_idx;
    *c_idx = task.c_idx;
    const uint8_t* q = a.u.seq + a.u.off[task.u_idx];
    const uint8_t* qq = quals ? a.u.qual + a.u.off[task.u_idx] : nullptr;
    const uint8_t* cs = a.u.seq + a.u.off[task.c_idx];
    const uint32_t lq = a.u.len[task.u_idx], lc = a.u.len[task.c_idx];
    for (uint32_t pos = 0; pos < lq; ++pos) {
      const uint32_t qn = q[pos];
      uint32_t tr = 5u * (qn - 1u);
      if (pos < lc) {
        const uint32_t cn = cs[pos];
        if (cn != qn) tr = 4u * (cn - 1u) + (qn - 1u);
      }
      f(tr, qq ? (uint32_t)qq[pos] : 0u);
    }
  } else {
    const uint32_t di = t - n_g;
    const DpTask task = a.dp[di];
    *u_idx = task.u_idx;
    *c_idx = task.c_idx;
    const uint8_t* q = a.u.seq + a.u.off[task.u_idx];
    const uint8_t* qq = quals ? a.u.qual + a.u.off[task.u_idx] : nullptr;
    const uint8_t* cs = a.u.seq + a.u.off[task.c_idx];
    const uint32_t lq = a.u.len[task.u_idx];
    const uint32_t cnt = a.sub_cnt[di] & 0x7FFFFFFFu;
    const uint32_t* subs = a.sub_pool + (size_t)di * a.sub_stride;
    int idx = (int)cnt - 1;
    uint32_t next = idx >= 0 ? subs[idx] : 0xFFFFFFFFu;
    for (uint32_t pos = 0; pos < lq; ++pos) {
      const uint32_t qn = q[pos];
      uint32_t tr = 5u * (qn - 1u);
      if ((next >> 16) == pos) {
        tr = 4u * ((uint32_t)cs[next & 0xFFFFu] - 1u) + (qn - 1u);
        --idx;
        next = idx >= 0 ? subs[idx] : 0xFFFFFFFFu;
      }
      f(tr, qq ? (uint32_t)qq[pos] : 0u);
    }
  }
}

__global__ __launch_bounds__(256) void loglik_fwd_kernel(LoglikDev a, const float* logE) {
  extern __shared__ double sh[];
  double* le = sh;  // 16 * ncol
  __shared__ double red[256 / 64];
  const uint32_t ne = 16u * a.ncol;
  for (uint32_t i = threadIdx.x; i < ne; i += blockDim.x) le[i] = (double)logE[i];
  __syncthreads();
  const uint32_t n_g = a.ctr->n_gapless, n = n_g + a.ctr->n_dp;
  const bool quals = a.use_quals && a.u.qual != nullptr;
  const uint32_t ncol = a.ncol;
  double acc = 0.0;
  for (uint32_t t = blockIdx.x * blockDim.x + threadIdx.x; t < n; t += gridDim.x * blockDim.x) {
    double ll_lam = 0.0;
    uint32_t ui = 0, ci = 0;
    walk_task(a, t, n_g, quals, &ui, &ci, [&](uint32_t tr, uint32_t qi) { ll_lam += le[tr * ncol + qi]; });
    const double av = (double)a.u.reads[ui];
    const double E = (double)a.u.reads[ci] * exp(ll_lam);
    a.w[t] = av - E;
    acc += av * log(E) - E - lgamma(av + 1.0);
  }
  for (int off = 32; off > 0; off >>= 1) acc += __shfl_xor(acc, off);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = acc;
  __syncthreads();
  if (threadIdx.x == 0) {
    double s = 0.0;
    for (int i = 0; i < 256 / 64; ++i) s += red[i];
    a.block_loss[blockIdx.x] = s;
  }
}

__global__ __launch_bounds__(256) void loglik_bwd_kernel(LoglikDev a) {
  extern __shared__ double hist[];  // 16 * ncol
  const uint32_t ne = 16u * a.ncol;
  for (uint32_t i = threadIdx.x; i < ne; i += blockDim.x) hist[i] = 0.0;
  __syncthreads();
  const uint32_t n_g = a.ctr->n_gapless, n = n_g + a.ctr->n_dp;
  const bool quals = a.use_quals && a.u.qual != nullptr;
  const uint32_t ncol = a.ncol;
  for (uint32_t t = blockIdx.x * blockDim.x + threadIdx.x; t < n; t += gridDim.x * blockDim.x) {
    const double wt = a.w[t];
    uint32_t ui = 0, ci = 0;
    walk_task(a, t, n_g, quals, &ui, &ci,
              [&](uint32_t tr, uint32_t qi) { atomicAdd(&hist[tr * ncol + qi], wt); });
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
  d.u = *s.u;
  d.gl = s.gapless_tasks;
  d.dp = s.dp_tasks;
  d.sub_pool = s.sub_pool;
  d.sub_stride = s.sub_stride;
  d.sub_cnt = s.sub_cnt;
  d.ncol = s.ncol;
  d.use_quals = s.use_quals;
  d.ctr = s.ctr;
  d.w = s.w;
  d.block_loss = s.block_part;
  d.block_grad = s.block_part;
  return d;
}

hipError_t launch_loglik_fwd(const LoglikArgs& s, const float* logE, double* loss,
                             hipStream_t st) {
  LoglikDev d = make_dev(s);
  const size_t lds = (size_t)16 * s.ncol * sizeof(double);
  hipLaunchKernelGGL(loglik_fwd_kernel, dim3(s.grid), dim3(256), lds, st, d, logE);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(reduce_blocks_kernel, dim3(1), dim3(64), 0, st, (const double*)s.block_part,
                     s.grid, 1u, loss);
  return hipGetLastError();
}

hipError_t launch_loglik_bwd(const LoglikArgs& s, double* grad, hipStream_t st) {
  LoglikDev d = make_dev(s);
  const uint32_t m = 16u * s.ncol;
  const size_t lds = (size_t)m * sizeof(double);
  hipLaunchKernelGGL(loglik_bwd_kernel, dim3(s.grid), dim3(256), lds, st, d);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(reduce_blocks_kernel, dim3((m + 255) / 256), dim3(256), 0, st,
                     (const double*)s.block_part, s.grid, m, grad);
  return hipGetLastError();
}

}  // namespace d2g
