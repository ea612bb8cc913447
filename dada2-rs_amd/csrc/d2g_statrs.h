// d2g_statrs.h — the Poisson upper tail behind calc_pA (pval.rs:208-241),
// host and device.
//
// The reference calls Poisson::sf of statrs 0.18 (Cargo.toml:40), defined as
// sf(x) = gamma_lr(x + 1, lambda) (pval.rs:211-218).  statrs's published
// gamma_lr is the Cephes incomplete gamma: ax = a ln x - x - ln_gamma(a);
// exactly 0 (a >= x) or 1 when ax < -709.78271289338399, and 0 when x is
// within 1.1102230246251565e-15 of 0 (prec::DEFAULT_F64_ACC); otherwise the
// power series exp(ax) (1 + x/(a+1) + ...) / a for x <= 1 or x <= a, else
// 1 - exp(ax) times a continued fraction with Cephes's rescaling, both
// stopped at a relative change of 1e-15.  ln_gamma is statrs's Lanczos
// approximation (g = 10.900511, 11 coefficients, summed left to right).
// These are restated in that operation order; where the reference's fallback
// (poisson_upper_tail_direct, pval.rs:329-353) takes over is therefore the
// same band of (reads, E).  The loops carry an iteration cap that the
// converging cases never reach, so device code cannot spin on a NaN.
#pragma once

#include <math.h>

#ifndef D2G_HD
#define D2G_HD __host__ __device__ __forceinline__
#endif

namespace d2g {

D2G_HD double statrs_ln_gamma(double x) {
  const double r = 10.900511;
  const double dk[11] = {2.48574089138753565546e-5,  1.05142378581721974210,
                         -3.45687097222016235469,    4.51227709466894823700,
                         -2.98285225323576655721,    1.05639711577126713077,
                         -1.95428773191645869583e-1, 1.70970543404441224307e-2,
                         -5.71926117404305781283e-4, 4.63399473359905636708e-6,
                         -2.71994908488607703910e-9};
  const double ln_pi = 1.1447298858494001741434273513530587116472948129153;
  const double ln_2_sqrt_e_over_pi = 0.6207822376352452223455184457816472122518527279025978;
  const double e = 2.71828182845904523536028747135266250;
  const double pi = 3.14159265358979323846264338327950288;
  double s = dk[0];
  if (x < 0.5) {
    for (int i = 1; i < 11; ++i) s = s + dk[i] / ((double)i - x);
    return ln_pi - log(sin(pi * x)) - log(s) - ln_2_sqrt_e_over_pi -
           (0.5 - x) * log((0.5 - x + r) / e);
  }
  for (int i = 1; i < 11; ++i) s = s + dk[i] / (x + (double)i - 1.0);
  return log(s) + ln_2_sqrt_e_over_pi + (x - 0.5) * log((x - 0.5 + r) / e);
}

// gamma_lr(a, x) = P(a, x) for a > 0 and x > 0 (calc_pA's range).
D2G_HD double statrs_gamma_lr(double a, double x) {
  const double eps = 0.000000000000001;
  const double big = 4503599627370496.0;
  const double big_inv = 2.22044604925031308085e-16;
  const double acc = 0.0000000000000011102230246251565;
  if (fabs(a - 0.0) < acc) return 1.0;
  if (fabs(x - 0.0) < acc) return 0.0;
  const double ax = a * log(x) - x - statrs_ln_gamma(a);
  if (ax < -709.78271289338399) return a < x ? 1.0 : 0.0;
  if (x <= 1.0 || x <= a) {
    double r2 = a, c2 = 1.0, ans2 = 1.0;
    for (int it = 0; it < 1000000; ++it) {
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
  for (int it = 0; it < 1000000; ++it) {
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

}  // namespace d2g
