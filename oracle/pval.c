/*
 * pval.c — restatement of dada2-rs src/pval.rs (TEST INFRASTRUCTURE ONLY).
 *
 * compute_lambda is a left-to-right f64 product over query positions starting
 * from 1.0 (pval.rs:311-313).  calc_pA calls Poisson::sf of statrs 0.18
 * (Cargo.toml:40, Cargo.lock:715-718), which is not vendored with the
 * reference.  statrs defines sf(x) = gamma_lr(x + 1, lambda) (pval.rs:211-218),
 * and its published gamma_lr is the Cephes incomplete gamma (igam: a power
 * series for x <= 1 or x <= a, else a continued fraction for the complement,
 * both stopped at a relative change of 1e-15), scaled by exp(ax) with
 * ax = a ln x - x - ln_gamma(a), which returns exactly 0 (a >= x) or 1 when
 * ax < -709.78271289338399 and 0 when x is within 1.1102230246251565e-15 of
 * 0; ln_gamma is a Lanczos approximation (g = 10.900511, 11 terms).  Both are
 * restated below in statrs's operation order.  The reference's own fallback
 * poisson_upper_tail_direct runs exactly where that gamma_lr returns 0.
 * PARITY UNPINNED at the bit level against statrs itself (not importable
 * here): the restatement is pinned by lgamma and the regularised incomplete
 * gamma of the system library / scipy to ~1e-14, by Poisson identities, and
 * by the cut-off band's switch to the direct series.
 */
#include <float.h>
#include <math.h>

#include "oracle.h"

/* compute_lambda, pval.rs:253-320 */
int orc_compute_lambda(const orc_raw* raw, const orc_sub* sub, const double* err, uint32_t ncol,
                       int use_quals, double* out) {
  if (!sub) {
    *out = 0.0;
    return ORC_OK;
  }
  uint32_t len = raw->len;
  /* tvec/qind over query positions (pval.rs:268-291) */
  uint8_t tvec_small[4096];
  uint8_t* tvec = tvec_small;
  if (len > sizeof(tvec_small)) return ORC_ERR_ARG;
  for (uint32_t pos = 0; pos < len; ++pos) {
    uint8_t nt = raw->seq[pos];
    if (nt < 1 || nt > 4) return ORC_ERR_NONACGT;
    uint32_t nti = (uint32_t)nt - 1;
    tvec[pos] = (uint8_t)(nti * 4 + nti);
  }
  /* substitution overrides (pval.rs:294-309) */
  for (uint32_t s = 0; s < sub->nsubs; ++s) {
    uint32_t pos0 = sub->pos[s];
    uint32_t pos1 = sub->map[pos0];
    if (pos1 >= len) return ORC_ERR_ARG;
    uint32_t nti0 = sub->nt0[s] ? (uint32_t)sub->nt0[s] - 1 : 0; /* saturating_sub(1) */
    uint32_t nti1 = sub->nt1[s] ? (uint32_t)sub->nt1[s] - 1 : 0;
    tvec[pos1] = (uint8_t)(nti0 * 4 + nti1);
  }
  double lambda = 1.0;
  for (uint32_t pos = 0; pos < len; ++pos) {
    uint32_t q = (use_quals && raw->qual) ? raw->qual[pos] : 0;
    if (q >= ncol) return ORC_ERR_ARG;
    lambda *= err[(size_t)tvec[pos] * ncol + q];
  }
  *out = lambda;
  return ORC_OK;
}

/* statrs 0.18 ln_gamma: Lanczos approximation, g = GAMMA_R, coefficients
 * GAMMA_DK, summed left to right from GAMMA_DK[0]. */
static const double GAMMA_R = 10.900511;
static const double GAMMA_DK[11] = {
    2.48574089138753565546e-5,  1.05142378581721974210,    -3.45687097222016235469,
    4.51227709466894823700,     -2.98285225323576655721,   1.05639711577126713077,
    -1.95428773191645869583e-1, 1.70970543404441224307e-2, -5.71926117404305781283e-4,
    4.63399473359905636708e-6,  -2.71994908488607703910e-9};
static const double LN_PI = 1.1447298858494001741434273513530587116472948129153;
static const double LN_2_SQRT_E_OVER_PI = 0.6207822376352452223455184457816472122518527279025978;
static const double E_CONST = 2.71828182845904523536028747135266250;
static const double PI_CONST = 3.14159265358979323846264338327950288;

double orc_ln_gamma(double x) {
  double s = GAMMA_DK[0];
  if (x < 0.5) {
    for (int i = 1; i < 11; ++i) s = s + GAMMA_DK[i] / ((double)i - x);
    return LN_PI - log(sin(PI_CONST * x)) - log(s) - LN_2_SQRT_E_OVER_PI -
           (0.5 - x) * log((0.5 - x + GAMMA_R) / E_CONST);
  }
  for (int i = 1; i < 11; ++i) s = s + GAMMA_DK[i] / (x + (double)i - 1.0);
  return log(s) + LN_2_SQRT_E_OVER_PI + (x - 0.5) * log((x - 0.5 + GAMMA_R) / E_CONST);
}

/* statrs 0.18 gamma_lr(a, x) = P(a, x) for a > 0, x > 0 (the callers' range). */
double orc_gamma_lr(double a, double x) {
  if (isnan(a) || isnan(x)) return NAN;
  const double eps = 0.000000000000001;
  const double big = 4503599627370496.0;
  const double big_inv = 2.22044604925031308085e-16;
  const double acc = 0.0000000000000011102230246251565; /* prec::DEFAULT_F64_ACC */
  if (fabs(a - 0.0) < acc) return 1.0;
  if (fabs(x - 0.0) < acc) return 0.0;
  const double ax = a * log(x) - x - orc_ln_gamma(a);
  if (ax < -709.78271289338399) return a < x ? 1.0 : 0.0;
  if (x <= 1.0 || x <= a) {
    double r2 = a, c2 = 1.0, ans2 = 1.0;
    for (;;) {
      r2 += 1.0;
      c2 *= x / r2;
      ans2 += c2;
      if (c2 / ans2 <= eps) break;
    }
    return exp(ax) * ans2 / a;
  }
  double y = 1.0 - a, z = x + y + 1.0;
  int c = 0;
  double p3 = 1.0, q3 = x, p2 = x + 1.0, q2 = z * x;
  double ans = p2 / q2;
  for (;;) {
    y += 1.0;
    z += 2.0;
    c += 1;
    const double yc = y * (double)c;
    const double p = p2 * z - p3 * yc;
    const double q = q2 * z - q3 * yc;
    p3 = p2;
    p2 = p;
    q3 = q2;
    q2 = q;
    if (fabs(p) > big) {
      p3 *= big_inv;
      p2 *= big_inv;
      q3 *= big_inv;
      q2 *= big_inv;
    }
    if (q != 0.0) {
      const double nextans = p / q;
      const double error = fabs((ans - nextans) / nextans);
      ans = nextans;
      if (error <= eps) break;
    }
  }
  return 1.0 - exp(ax) * ans;
}

static double gamma_p_impl(double a, double x) { return orc_gamma_lr(a, x); }

double orc_gamma_q(double a, double x) { return 1.0 - gamma_p_impl(a, x); }

/* poisson_upper_tail_direct, pval.rs:329-353 */
double orc_poisson_upper_tail_direct(uint32_t reads, double lambda) {
  double k = (double)reads;
  double log_lambda = log(lambda);
  double log_kfact = 0.0;
  for (uint32_t i = 1; i <= reads; ++i) log_kfact += log((double)i);
  double log_leading = -lambda + k * log_lambda - log_kfact;
  double series_sum = 1.0, term = 1.0;
  for (uint32_t j = 1; j < 10000u; ++j) {
    term *= lambda / (k + (double)j);
    if (!isfinite(term) || term <= 0.0) break;
    double ns = series_sum + term;
    if (ns == series_sum) break;
    series_sum = ns;
  }
  return exp(log_leading + log(series_sum));
}

/* calc_pA, pval.rs:208-241 */
double orc_calc_pA(uint32_t reads, double e_reads, int prior) {
  if (!(e_reads > 0.0) || isnan(e_reads)) return 1.0; /* Poisson::new error */
  double pval = gamma_p_impl((double)reads, e_reads); /* sf(reads-1) */
  if (pval == 0.0 && e_reads > 0.0) pval = orc_poisson_upper_tail_direct(reads, e_reads);
  if (prior) return pval;
  double norm = 1.0 - exp(-e_reads);
  if (norm < 1e-7) norm = e_reads - 0.5 * e_reads * e_reads; /* TAIL_APPROX_CUTOFF, pval.rs:18 */
  return pval / norm;
}

/* get_pA, pval.rs:376-398 */
double orc_get_pA(uint32_t reads, int prior, double lambda, uint32_t hamming, uint32_t bi_reads,
                  int detect_singletons) {
  if (reads == 1 && !prior && !detect_singletons) return 1.0;
  if (hamming == 0) return 1.0;
  if (lambda == 0.0) return 0.0;
  double e_reads = lambda * (double)bi_reads;
  return orc_calc_pA(reads, e_reads, prior || detect_singletons);
}
