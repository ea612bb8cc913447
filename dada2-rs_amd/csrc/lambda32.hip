// lambda32.hip — mixed-precision lambda and the FP64 abundance p-values over
// a resident set of aligned pairs.
//
// A pair is a query (bases 1..4, Phred qualities) with its al2subs list
// against the center: records (query position << 16 | center base), in
// ascending query position.  compute_lambda (pval.rs:253-320) is the ordered
// product over query positions of err[t][q], t = 5 (nt-1) on the diagonal and
// 4 (nt0-1) + (nt-1) at a substitution.
//
// loglam32_kernel: log lambda as an FP32 Kahan (compensated) sum, in the same
// position order, of a table log_err[t][q] = (float)log(err[t][q]) computed in
// FP64.  Each table entry carries one FP32 rounding (<= 2^-24 relative) and
// every term has the same sign, so the sum's relative error stays near
// 2^-24 plus the compensated-summation bound.
//
// pval_kernel: get_pA / calc_pA (pval.rs:376-398, 208-241) in FP64 from
// lambda = exp(log lambda):  sentinels (singleton, hamming 0, lambda 0), the
// Poisson upper tail P(X >= reads) = P(reads, E) (regularised lower
// incomplete gamma, E = lambda * center reads), the reference's direct
// series when that underflows (pval.rs:329-353), and the 1 - e^{-E}
// conditioning with its Taylor replacement below 1e-7.  Pairs whose p-value
// lands within a factor 10 of omega_a get lambda recomputed as the exact
// ordered FP64 product and their p-value recomputed, so the p < omega_a call
// is taken on FP64 lambda wherever it is close.
//
// The incomplete gamma is evaluated with the series (x < a + 1) or the
// modified-Lentz continued fraction (Numerical Recipes, 3rd ed., 6.2), using
// the device lgamma/exp/log.
#include <float.h>

#include "d2g_launch.h"

namespace d2g {

struct PairDev {
  uint32_t n;
  const uint8_t* seq;       // rows padded to 16 bytes
  const uint8_t* qual;      // nullable, same layout
  const uint64_t* off;      // n + 1
  const uint32_t* len;      // n
  const uint32_t* subs;     // (query pos << 16 | center base)
  const uint64_t* sub_off;  // n + 1
  const uint32_t* reads;
  const uint32_t* center_reads;
};

__global__ __launch_bounds__(256) void loglam32_kernel(PairDev a, const double* err, uint32_t ncol,
                                                       int use_quals, float* out_loglam) {
  extern __shared__ float logerr[];
  const uint32_t ne = 16u * ncol;
  for (uint32_t i = threadIdx.x; i < ne; i += blockDim.x) logerr[i] = (float)log(err[i]);
  __syncthreads();
  const bool quals = use_quals && a.qual != nullptr;
  for (uint32_t t = blockIdx.x * blockDim.x + threadIdx.x; t < a.n; t += gridDim.x * blockDim.x) {
    const uint4* q16 = reinterpret_cast<const uint4*>(a.seq + a.off[t]);
    const uint4* u16 = quals ? reinterpret_cast<const uint4*>(a.qual + a.off[t]) : nullptr;
    const uint32_t lq = a.len[t];
    const uint32_t* subs = a.subs + a.sub_off[t];
    const uint32_t ns = (uint32_t)(a.sub_off[t + 1] - a.sub_off[t]);
    uint32_t si = 0;
    uint32_t next = ns ? subs[0] : 0xFFFFFFFFu;
    float sum = 0.0f, comp = 0.0f;
    bool zero = false;
    for (uint32_t base = 0; base < lq; base += 16) {
      const uint4 qv = q16[base >> 4];
      const uint4 uv = u16 ? u16[base >> 4] : make_uint4(0, 0, 0, 0);
      const uint32_t qw[4] = {qv.x, qv.y, qv.z, qv.w};
      const uint32_t uw[4] = {uv.x, uv.y, uv.z, uv.w};
#pragma unroll
      for (int b = 0; b < 16; ++b) {
        const uint32_t pos = base + (uint32_t)b;
        if (pos < lq) {
          const uint32_t nt = (qw[b >> 2] >> (8 * (b & 3))) & 0xFFu;
          const uint32_t qi = (uw[b >> 2] >> (8 * (b & 3))) & 0xFFu;
          uint32_t tr = 5u * (nt - 1u);
          if ((next >> 16) == pos) {
            tr = 4u * ((next & 0xFFFFu) - 1u) + (nt - 1u);
            ++si;
            next = si < ns ? subs[si] : 0xFFFFFFFFu;
          }
          const float term = logerr[tr * ncol + qi];
          zero = zero || term == -INFINITY;
          // Kahan: compensated sum in FP32 (no contraction: no products)
          const float y = term - comp;
          const float s = sum + y;
          comp = (s - sum) - y;
          sum = s;
        }
      }
    }
    out_loglam[t] = zero ? -INFINITY : sum;
  }
}

// Regularised lower incomplete gamma P(a, x).
__device__ double gamma_p(double a, double x) {
  if (x <= 0.0) return 0.0;
  const double gln = lgamma(a);
  if (x < a + 1.0) {
    double ap = a, del = 1.0 / a, sum = del;
    for (int n = 0; n < 100000; ++n) {
      ap += 1.0;
      del *= x / ap;
      sum += del;
      if (fabs(del) < fabs(sum) * DBL_EPSILON) break;
    }
    return sum * exp(-x + a * log(x) - gln);
  }
  const double fpmin = DBL_MIN / DBL_EPSILON;
  double bb = x + 1.0 - a, c = 1.0 / fpmin, d = 1.0 / bb, h = d;
  for (int i = 1; i < 100000; ++i) {
    const double an = -i * (i - a);
    bb += 2.0;
    d = an * d + bb;
    if (fabs(d) < fpmin) d = fpmin;
    c = bb + an / c;
    if (fabs(c) < fpmin) c = fpmin;
    d = 1.0 / d;
    const double del = d * c;
    h *= del;
    if (fabs(del - 1.0) <= DBL_EPSILON) break;
  }
  return 1.0 - exp(-x + a * log(x) - gln) * h;
}

// poisson_upper_tail_direct, pval.rs:329-353
__device__ double poisson_tail_direct(uint32_t reads, double lambda) {
  const double k = (double)reads;
  const double log_lambda = log(lambda);
  double log_kfact = 0.0;
  for (uint32_t i = 1; i <= reads; ++i) log_kfact += log((double)i);
  const double log_leading = -lambda + k * log_lambda - log_kfact;
  double series = 1.0, term = 1.0;
  for (uint32_t j = 1; j < 10000u; ++j) {
    term *= lambda / (k + (double)j);
    if (!isfinite(term) || term <= 0.0) break;
    const double ns = series + term;
    if (ns == series) break;
    series = ns;
  }
  return exp(log_leading + log(series));
}

// get_pA with prior = detect_singletons = false (b_p_update's per-raw call)
__device__ double get_pA_dev(uint32_t reads, double lambda, uint32_t hamming, uint32_t bi_reads) {
  if (reads == 1) return 1.0;
  if (hamming == 0) return 1.0;
  if (lambda == 0.0) return 0.0;
  const double e = lambda * (double)bi_reads;
  if (!(e > 0.0)) return 1.0;
  double p = gamma_p((double)reads, e);
  if (p == 0.0) p = poisson_tail_direct(reads, e);
  double norm = 1.0 - exp(-e);
  if (norm < 1e-7) norm = e - 0.5 * e * e;
  return p / norm;
}

// The exact ordered FP64 product of compute_lambda (pval.rs:311-313).
__device__ double lambda_fp64(const PairDev& a, uint32_t t, const double* err, uint32_t ncol,
                              bool quals) {
  const uint8_t* q = a.seq + a.off[t];
  const uint8_t* u = quals ? a.qual + a.off[t] : nullptr;
  const uint32_t lq = a.len[t];
  const uint32_t* subs = a.subs + a.sub_off[t];
  const uint32_t ns = (uint32_t)(a.sub_off[t + 1] - a.sub_off[t]);
  uint32_t si = 0;
  uint32_t next = ns ? subs[0] : 0xFFFFFFFFu;
  double lam = 1.0;
  for (uint32_t pos = 0; pos < lq; ++pos) {
    const uint32_t nt = q[pos];
    uint32_t tr = 5u * (nt - 1u);
    if ((next >> 16) == pos) {
      tr = 4u * ((next & 0xFFFFu) - 1u) + (nt - 1u);
      ++si;
      next = si < ns ? subs[si] : 0xFFFFFFFFu;
    }
    lam *= err[tr * ncol + (u ? u[pos] : 0u)];
  }
  return lam;
}

__global__ __launch_bounds__(256) void pval_kernel(PairDev a, const double* err, uint32_t ncol,
                                                   int use_quals, const float* loglam,
                                                   double omega, double* out_lambda,
                                                   double* out_p, uint8_t* out_flag,
                                                   uint32_t* n_recomputed) {
  const bool quals = use_quals && a.qual != nullptr;
  uint32_t recomputed = 0;
  for (uint32_t t = blockIdx.x * blockDim.x + threadIdx.x; t < a.n; t += gridDim.x * blockDim.x) {
    const uint32_t ham = (uint32_t)(a.sub_off[t + 1] - a.sub_off[t]);
    const float ll = loglam[t];
    double lam = ll == -INFINITY ? 0.0 : exp((double)ll);
    double p = get_pA_dev(a.reads[t], lam, ham, a.center_reads[t]);
    uint8_t flag = 0;
    if (p >= omega * 0.1 && p <= omega * 10.0) {
      lam = lambda_fp64(a, t, err, ncol, quals);
      p = get_pA_dev(a.reads[t], lam, ham, a.center_reads[t]);
      flag = 1;
      ++recomputed;
    }
    out_lambda[t] = lam;
    out_p[t] = p;
    out_flag[t] = flag;
  }
  for (int off = 32; off > 0; off >>= 1) recomputed += __shfl_xor(recomputed, off);
  if ((threadIdx.x & 63) == 0 && recomputed) atomicAdd(n_recomputed, recomputed);
}

hipError_t launch_pairs_lambda(const PairArgs& s, hipStream_t st) {
  PairDev d;
  d.n = s.n;
  d.seq = s.seq;
  d.qual = s.qual;
  d.off = s.off;
  d.len = s.len;
  d.subs = s.subs;
  d.sub_off = s.sub_off;
  d.reads = s.reads;
  d.center_reads = s.center_reads;
  if (s.n == 0) return hipSuccess;
  const size_t lds = (size_t)16 * s.ncol * sizeof(float);
  uint32_t blocks = (s.n + 255) / 256;
  if (blocks > s.max_blocks) blocks = s.max_blocks;
  hipLaunchKernelGGL(loglam32_kernel, dim3(blocks), dim3(256), lds, st, d, s.err, s.ncol,
                     s.use_quals, s.out_loglam);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  if (s.ev_mid) {
    e = hipEventRecord(s.ev_mid, st);
    if (e != hipSuccess) return e;
  }
  hipLaunchKernelGGL(pval_kernel, dim3(blocks), dim3(256), 0, st, d, s.err, s.ncol, s.use_quals,
                     s.out_loglam, s.omega, s.out_lambda, s.out_p, s.out_flag, s.n_recomputed);
  return hipGetLastError();
}

}  // namespace d2g
