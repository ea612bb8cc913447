/*
 * pval.c — restatement of dada2-rs src/pval.rs (TEST INFRASTRUCTURE ONLY).
 *
 * compute_lambda is a left-to-right f64 product over query positions starting
 * from 1.0 (pval.rs:311-313).  calc_pA calls statrs 0.18 Poisson::sf, which
 * is not vendored with the reference: here the Poisson upper tail is the
 * regularised lower incomplete gamma P(reads, e) (sf(x) = gamma_lr(x+1, e),
 * pval.rs:211-218), evaluated with the series / Lentz continued fraction of
 * Numerical Recipes (Press et al., 3rd ed., section 6.2).  PARITY UNPINNED
 * against statrs bit patterns; pinned only by Poisson identities and the
 * reference's own small-lambda series (poisson_upper_tail_direct).
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

/* Regularised lower incomplete gamma P(a, x); NR 3rd ed. 6.2. */
static double gamma_p_impl(double a, double x) {
  if (x <= 0.0) return 0.0;
  double gln = lgamma(a);
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
  /* Lentz continued fraction for Q(a, x) */
  const double FPMIN = DBL_MIN / DBL_EPSILON;
  double bb = x + 1.0 - a, c = 1.0 / FPMIN, d = 1.0 / bb, h = d;
  for (int i = 1; i < 100000; ++i) {
    double an = -i * (i - a);
    bb += 2.0;
    d = an * d + bb;
    if (fabs(d) < FPMIN) d = FPMIN;
    c = bb + an / c;
    if (fabs(c) < FPMIN) c = FPMIN;
    d = 1.0 / d;
    double del = d * c;
    h *= del;
    if (fabs(del - 1.0) <= DBL_EPSILON) break;
  }
  double q = exp(-x + a * log(x) - gln) * h;
  return 1.0 - q;
}

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
