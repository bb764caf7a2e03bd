// crmath.h -- correctly rounded (with overwhelming probability) FP64 exp,
// log and pow, one source for the host (gcc, the oracle) and the device
// (hipcc, the FP64 kernels), so both evaluate them bit for bit alike.
//
// Method: Tang's table-driven algorithms (P. T. P. Tang, "Table-driven
// implementation of the exponential function", ACM TOMS 15(2), 1989, and
// "... of the logarithm function", ACM TOMS 16(4), 1990) carried in
// double-double arithmetic (Dekker's exact product by FMA, Knuth's
// two-sum), so that the value before the final rounding is within about
// 2^-70 of the true result; its one rounding to double is then the correct
// rounding except where the true value lies within that distance of a
// rounding boundary (a fraction of about 2^-17 of arguments).
//
//   log(x)   = k ln2 - log(invc_i) + log1p(r),  x = 2^k m, m in [1, 2),
//              i = the nearest of 1 + i/256 to m, invc_i = 1/(1 + i/256)
//              rounded, r = m invc_i - 1 formed exactly; log1p(r) by its
//              Taylor series to r^9 (|r| <= 2^-9)
//   exp(x)   = 2^e 2^(j/256) exp(r),  x = (256 e + j) ln2/256 + r,
//              |r| <= ln2/512; exp(r) by its Taylor series to r^7
//   pow(x,y) = exp(y log(x)) with log(x) and y log(x) as double-doubles
//
// Arguments outside the ranges the model produces (x <= 0, non-finite
// values, |result exponent| near the overflow range) are passed to the
// platform's own function.  The tables come from gen_crmath_tables.py.
#pragma once
#include <math.h>
#include <stdint.h>

#if defined(__HIPCC__)
#define CRM_FN __host__ __device__ static inline
#define CRM_TABLE static constexpr
#else
#define CRM_FN static inline
#define CRM_TABLE static const
#endif

#include "crmath_tables.h"

CRM_FN uint64_t crm_bits(double x) {
  uint64_t u;
  __builtin_memcpy(&u, &x, 8);
  return u;
}
CRM_FN double crm_double(uint64_t u) {
  double x;
  __builtin_memcpy(&x, &u, 8);
  return x;
}
// s + e = a + b exactly (Knuth)
CRM_FN void crm_two_sum(double a, double b, double *s, double *e) {
  const double t = a + b, bb = t - a;
  *s = t;
  *e = (a - (t - bb)) + (b - bb);
}
// s + e = a + b exactly, |a| >= |b| (Dekker)
CRM_FN void crm_fast_two_sum(double a, double b, double *s, double *e) {
  const double t = a + b;
  *s = t;
  *e = b - (t - a);
}
// p + e = a b exactly
CRM_FN void crm_two_prod(double a, double b, double *p, double *e) {
  const double t = a * b;
  *p = t;
  *e = __builtin_fma(a, b, -t);
}

// log(x) as a double-double hi + lo, x positive and finite
CRM_FN void crm_log_dd(double x, double *hi, double *lo) {
  uint64_t ix = crm_bits(x);
  int k = (int)(ix >> 52) - 1023;
  if ((ix >> 52) == 0) {  // subnormal
    ix = crm_bits(x * 0x1p54);
    k = (int)(ix >> 52) - 1023 - 54;
  }
  double m = crm_double((ix & 0x000fffffffffffffULL) | 0x3ff0000000000000ULL);  // [1, 2)
  // nearest table point 1 + i/256: the top nine mantissa bits, rounded
  int i = (int)(((ix >> 43) & 0x1ff) + 1) >> 1;
  double rh, rl;
  if (i == CRM_N) {  // m near 2: log(x) = (k+1) ln2 + log1p(m/2 - 1)
    k += 1;
    m *= 0.5;
    i = 0;
  }
  if (i == 0) {
    rh = m - 1.0;  // exact (Sterbenz)
    rl = 0.0;
  } else {
    double p, e;
    crm_two_prod(m, CRM_LOG_INVC[i], &p, &e);
    crm_two_sum(p - 1.0, e, &rh, &rl);  // p - 1 is exact (Sterbenz)
  }
  // log1p(rh + rl) = rh - rh^2/2 + rh^3/3 - ... + rl (1 - rh + rh^2) + ...
  double sh, sl;
  crm_two_prod(rh, rh, &sh, &sl);
  const double poly =
      rh * sh *
      (1. / 3. + rh * (-1. / 4. + rh * (1. / 5. + rh * (-1. / 6. + rh * (1. / 7. + rh * (-1. / 8. + rh * (1. / 9.)))))));
  double a, ea;
  crm_two_sum(rh, -0.5 * sh, &a, &ea);
  const double l1 = ea + (rl + ((-0.5 * sl) + (-(rh * rl) + (sh * rl + poly))));
  // k ln2 + T_i + a + l1
  const double kd = (double)k;
  double b, eb, c, ec;
  crm_two_sum(kd * CRM_LN2_HI, CRM_LOG_T[2 * i], &b, &eb);
  crm_two_sum(b, a, &c, &ec);
  const double low = eb + ec + (kd * CRM_LN2_LO + CRM_LOG_T[2 * i + 1] + l1);
  crm_fast_two_sum(c, low, hi, lo);
}

CRM_FN double crm_log(double x) {
  if (!(x > 0.) || x == __builtin_inf()) return log(x);
  double hi, lo;
  crm_log_dd(x, &hi, &lo);
  return hi + lo;
}

// exp(xh + xl), |xl| <= ulp(xh), |xh| < 700
CRM_FN double crm_exp_dd(double xh, double xl) {
  const double kd = __builtin_rint(xh * CRM_INV_LN2N);
  const int k = (int)kd;
  // r = xh + xl - k ln2/256, ln2/256 = C1 + C2 + C3
  double rh, rl;
  crm_two_sum(xh - kd * CRM_LN2N_C1, -(kd * CRM_LN2N_C2), &rh, &rl);  // both products exact
  rl += xl - kd * CRM_LN2N_C3;
  crm_fast_two_sum(rh, rl, &rh, &rl);
  // expm1(rh + rl) = rh + rh^2/2 + rh^3/6 + ... + rl (1 + rh)
  double sh, sl;
  crm_two_prod(rh, rh, &sh, &sl);
  const double poly = rh * sh * (1. / 6. + rh * (1. / 24. + rh * (1. / 120. + rh * (1. / 720. + rh * (1. / 5040.)))));
  double a, ea;
  crm_two_sum(rh, 0.5 * sh, &a, &ea);
  double ph, pl;
  crm_fast_two_sum(a, ea + (rl + (0.5 * sl + (rh * rl + poly))), &ph, &pl);
  // 2^(j/256) (1 + p)
  const int j = k & (CRM_N - 1), e = (k - j) / CRM_N;
  const double th = CRM_EXP_T[2 * j], tl = CRM_EXP_T[2 * j + 1];
  double q1, q2, s, es;
  crm_two_prod(th, ph, &q1, &q2);
  crm_fast_two_sum(th, q1, &s, &es);
  const double r = s + (es + (q2 + (th * pl + (tl + tl * ph))));
  // times 2^e: exact for normal results (|e| <= 1010 here)
  return r * crm_double((uint64_t)(1023 + e) << 52);
}

CRM_FN double crm_exp(double x) {
  if (!(__builtin_fabs(x) < 700.)) return exp(x);
  return crm_exp_dd(x, 0.);
}

CRM_FN double crm_pow(double x, double y) {
  if (x == 1. || y == 0.) return 1.;
  if (!(x > 0.) || x == __builtin_inf() || !(__builtin_fabs(y) < __builtin_inf())) return pow(x, y);
  double lh, ll;
  crm_log_dd(x, &lh, &ll);
  double zh, zl;
  crm_two_prod(y, lh, &zh, &zl);
  zl += y * ll;
  crm_fast_two_sum(zh, zl, &zh, &zl);
  if (!(__builtin_fabs(zh) < 700.)) return pow(x, y);
  return crm_exp_dd(zh, zl);
}
