// dev_common.h -- device-side data model of the MI355X EPIC engine.
//
// Every kernel is launched over a bucket of members that share (ni,nj,nk):
// blockIdx.z selects the member, the remaining grid covers that member's
// cells.  All per-member pointers and constants live in a MemberDesc held in
// device memory; the per-member time-plane rotation of the Adams-Bashforth
// tendencies lives in device memory too (MemberDyn), so a captured step
// replays without host involvement.
//
// Layout of every 3-D field is the reference's padded (K,J,I):
//   index = I + (J+1)*Iadim + K*Nelem2d,  K = 0..nk+1, J = -1..nj+1, I = 0..ni+1
// (src/shared/epic_funcs_diag.c:654-664, include/epic.h:407-523).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

// FP64 exp, log and pow of every kernel: correctly rounded, from the same
// source the oracle's attribution build uses (crmath.h)
#include "crmath.h"

namespace epic {

constexpr int KMAX = 66;  // nk <= 64

struct MemberDyn {
  int it0, itm1, itm2;  // IT_ZERO, IT_MINUS1, IT_MINUS2
  int itime;            // value of grid.itime during the next step
  int err;              // sticky fatal-condition bits
  int pad[3];
};

struct MemberDesc {
  int ni, nj, nk, Ia, N2, N3;
  int ks, k_sponge;  // k_sigma, k_sponge
  int is_hs, is_gas, newt_cool_on, newt_cool_adjust, drag_v, drag_zonal_avg, jtp;
  int nrow_h, nrow_v, nlogn;  // filtered rows (h- and v-grid), log2(ni)
  int mixed;                  // precision mode
  double dt, kappa, cp, cpr, rgas, press0, zeta0, zeta1, alpha, sigma_sigma;
  double topo_scale, prandtl, tau_drag;
  double coef2, coef4, nu4, m0z, nudiv, eps13, mu3, mu2, g_coeff;
  // del^6 and del^8 hyperviscosity: factored coefficients and nu
  double coef6, coef8, nu6, nu8;
  // day-night relaxation of the hot-Jupiter case (is_hj)
  int is_hj;
  double pbot, teq_contrast, teq_p_contrast, tau_rad, ref_theta_b, ref_pbk, ref_a_over_kappa;
  double hj_lnref;  // log(pbot / teq_p_contrast)
  // sum of the cell areas 1/mn over J = 0..nj, I = 1..ni in that order (the
  // layer-mean removal of the Newtonian relaxation)
  double heat_area;
  // Newtonian cooling toward the sounding (is_tab): PCHIP tables against
  // -log p of the ironed sounding temperature and of the cooling time
  int is_tab, n_snd, n_cool;
  // the traditional pressure gradient in place of the finite-volume one
  int pgrad_trad;
  // advection scheme per advected field (unity, theta, hdry): 1 Akima, 0
  // the predictor-corrector of Hsu, Konor and Arakawa
  int adv_akima[3];
  // PV averaging of uv_core and gz_from_u: 1 Arakawa-Lamb (1981), 0 Sadourny
  int pv_al81;
  const double *snd_x, *snd_y, *snd_d, *cool_x, *cool_y, *cool_d;
  // host-libm tables: cos(lat), cos(lat)^4 per h-row J, cos(lon) per column I
  const double *coslat_h, *coslat4_h, *coslon_h;
  // reciprocals formed once on the host (the same IEEE quotients the
  // kernels would form per cell): 1/m on v-rows J = 0..nj+1 (0 at the
  // poles, as the meridional advection takes it) and 1/n on h-rows 0..nj
  const double *minv_v, *ninv_h;
  // per layer K >= k_sigma: THETA2's interval in the sigmatheta table and its width
  const int *t2_kay;
  const double *t2_h;
  // 1-D metrics and tables
  const double *lat, *lon, *m, *n, *mn, *f, *g, *rln, *rlt, *sgth, *dsgth, *dsgth_inv, *h_min, *t_sponge;
  const double *mn_inv, *hl_h, *hl_pv, *hcx, *hcy, *hcz, *vcx, *vcy, *vcz;
  const double *tw_f, *tw_i, *tr_f, *tr_i;  // four1 / realft twiddles, forward and inverse
  const float *tw_f32, *tw_i32, *tr_f32, *tr_i32;
  const int *rows_h, *rows_v;
  const int *rflag_h, *rflag_v;  // 1 where row J is in the filter band
  // state and diagnostics
  double *U, *V, *DUDT, *DVDT, *HDRY, *THETA, *P3, *P2, *THETA2, *H3, *T2, *T3, *RHO2, *RHO3;
  double *EXNER3, *GZ3, *MONT3, *HEAT3, *PV3, *DIV, *W3, *DZDT3, *USPIN, *GZS;
  double *RI2;  // formed on demand (epic_engine_read, the drop-in's write-back)
  // mixed mode: DUDT, DVDT are one FP64 plane (the current step's sum) and
  // DUDTf, DVDTf the three FP32 Adams-Bashforth history planes
  float *DUDTf, *DVDTf;
  // sA..sD: scratch, three 3-D copies each (the batched advection gives each
  // advected field its own copy; every other user takes the first)
  double *gz, *sA, *sB, *sC, *sD, *unity, *old, *spl;
  double *rowbuf;  // 8*(nj+3) doubles: zonal means and meridional profiles
  MemberDyn *dyn;
};

__device__ __forceinline__ size_t ix3(const MemberDesc &d, int K, int J, int I) {
  return (size_t)I + (size_t)(J + 1) * d.Ia + (size_t)K * d.N2;
}
__device__ __forceinline__ size_t ix2(const MemberDesc &d, int J, int I) {
  return (size_t)I + (size_t)(J + 1) * d.Ia;
}

__device__ __forceinline__ double dmax(double x, double y) { return x > y ? x : y; }

// Hyperviscosity orders (src/shared/epic_subgrid.c:590-1006): pass p forms
// the order 2(p+1) from the order before it with the factored coefficient
// coef[2(p+1)], and is applied when nu[2(p+1)] > 0; the passes run up to
// the highest order with nu > 0 (0 passes: no hyperviscosity).
__host__ __device__ __forceinline__ double hv_coef(const MemberDesc &d, int pass) {
  return pass == 0 ? d.coef2 : pass == 1 ? d.coef4 : pass == 2 ? d.coef6 : d.coef8;
}
__host__ __device__ __forceinline__ double hv_nu(const MemberDesc &d, int pass) {
  return pass == 1 ? d.nu4 : pass == 2 ? d.nu6 : pass == 3 ? d.nu8 : 0.;
}
// PV weights of uv_core's pv*uh (v grid) and pv*vh (u grid) terms
// (include/epic_pv_schemes.h:34-62): Sadourny (1975), three-point sums over
// 12 with the extra terms zero; Arakawa and Lamb (1981), weighted four-point
// sums over 24 with the extra terms PH1, PH2 (v grid) or EP1, EP2 (u grid);
// every sum in the header's order.  P(j, i) reads PV in the type T.
template <class T>
struct PvW {
  T al, be, ga, de, e1, e2, coef;
};
template <class T, class F>
__device__ __forceinline__ PvW<T> pv_weights_v(bool al81, const F &P, int J, int I) {
  PvW<T> w;
  if (al81) {
    w.ga = T(2.) * P(J + 1, I + 1) + P(J + 1, I) + T(2.) * P(J, I) + P(J, I + 1);
    w.de = P(J + 1, I + 1) + T(2.) * P(J + 1, I) + P(J, I) + T(2.) * P(J, I + 1);
    w.al = T(2.) * P(J, I + 1) + P(J, I) + T(2.) * P(J - 1, I) + P(J - 1, I + 1);
    w.be = P(J, I + 1) + T(2.) * P(J, I) + P(J - 1, I) + T(2.) * P(J - 1, I + 1);
    w.e1 = -P(J + 1, I + 1) + P(J + 1, I) + P(J, I) - P(J, I + 1);
    w.e2 = -P(J, I + 1) + P(J, I) + P(J - 1, I) - P(J - 1, I + 1);
    w.coef = T(1.) / T(24.);
  } else {
    w.ga = (P(J, I) + P(J, I + 1) + P(J + 1, I + 1));
    w.de = (P(J, I) + P(J + 1, I) + P(J, I + 1));
    w.al = (P(J - 1, I) + P(J, I) + P(J, I + 1));
    w.be = (P(J, I) + P(J, I + 1) + P(J - 1, I + 1));
    w.e1 = w.e2 = T(0.);
    w.coef = T(1.) / T(12.);
  }
  return w;
}
template <class T, class F>
__device__ __forceinline__ PvW<T> pv_weights_u(bool al81, const F &P, int J, int I) {
  PvW<T> w;
  if (al81) {
    w.al = T(2.) * P(J + 1, I + 1) + P(J + 1, I) + T(2.) * P(J, I) + P(J, I + 1);
    w.be = P(J + 1, I) + T(2.) * P(J + 1, I - 1) + P(J, I - 1) + T(2.) * P(J, I);
    w.ga = T(2.) * P(J + 1, I) + P(J + 1, I - 1) + T(2.) * P(J, I - 1) + P(J, I);
    w.de = P(J + 1, I + 1) + T(2.) * P(J + 1, I) + P(J, I) + T(2.) * P(J, I + 1);
    w.e1 = P(J + 1, I + 1) + P(J + 1, I) - P(J, I) - P(J, I + 1);
    w.e2 = P(J + 1, I) + P(J + 1, I - 1) - P(J, I - 1) - P(J, I);
    w.coef = T(1.) / T(24.);
  } else {
    w.al = (P(J, I) + P(J + 1, I) + P(J + 1, I + 1));
    w.be = (P(J + 1, I - 1) + P(J + 1, I) + P(J, I));
    w.ga = (P(J, I - 1) + P(J, I) + P(J + 1, I));
    w.de = (P(J, I) + P(J + 1, I) + P(J, I + 1));
    w.e1 = w.e2 = T(0.);
    w.coef = T(1.) / T(12.);
  }
  return w;
}

__host__ __device__ __forceinline__ int hv_passes(const MemberDesc &d) {
  return d.nu8 > 0. ? 4 : d.nu6 > 0. ? 3 : d.nu4 > 0. ? 2 : 0;
}
__device__ __forceinline__ double dmin(double x, double y) { return x < y ? x : y; }

// fcmp (src/shared/epic_funcs_util.c:4275-4346): compare to within one ulp of
// the larger magnitude's binade.
__device__ __forceinline__ int fcmp_d(double x1, double x2) {
  int e;
  frexp(fabs(x1) > fabs(x2) ? x1 : x2, &e);
  double delta = ldexp(2.220446049250313e-16, e), d = x1 - x2;
  if (d > delta) return 1;
  if (d < -delta) return -1;
  return 0;
}

// exp, log and pow of the FP64 hydrostatics: correctly rounded (crmath.h)
// in the FP64 path, whose results are the oracle's bit for bit; the device
// library's own (within an ulp or two) when F, the mixed mode, whose fields
// are held to FP32's accuracy and which gains the time of the double-double
// evaluation
template <bool F> __device__ __forceinline__ double m_exp(double x) {
  if constexpr (F) return exp(x);
  else return crm_exp(x);
}
template <bool F> __device__ __forceinline__ double m_log(double x) {
  if constexpr (F) return log(x);
  else return crm_log(x);
}
template <bool F> __device__ __forceinline__ double m_pow(double x, double y) {
  if constexpr (F) return pow(x, y);
  else return crm_pow(x, y);
}

// hybrid sigma-theta coordinate (src/shared/epic_funcs_diag.c:1238-1314,1751-1770)
template <bool F = false>
__device__ __forceinline__ double get_sigma(double pbot, double p, double ptop) {
  return m_log<F>(p / pbot) / m_log<F>(ptop / pbot);
}
template <bool F = false>
__device__ __forceinline__ double get_p_sigma(double pbot, double s, double ptop) {
  return pbot * m_exp<F>(s * m_log<F>(ptop / pbot));
}
template <bool F = false>
__device__ __forceinline__ double g_sig(const MemberDesc &d, double s) {
  if (s <= d.sigma_sigma) return 0.;
  if (s >= 1.) return 1.;
  return d.g_coeff * (1. - m_exp<F>(-d.alpha * (s - d.sigma_sigma)));
}
template <bool F = false>
__device__ __forceinline__ double f_sig(const MemberDesc &d, double s) {
  if (s < 0.) return d.zeta0;
  if (s <= d.sigma_sigma) return d.zeta0 + s * (d.zeta1 - d.zeta0);
  if (s < 1.) return (1. - g_sig<F>(d, s)) * (d.zeta0 + s * (d.zeta1 - d.zeta0));
  return 0.;
}
// (sigmatheta - f_sigma(s)) / g_sigma(s), the theta of a hybrid level at
// sigma s, with g_sigma formed once for both terms (the same operations as
// calling f_sig and g_sig, so the same result)
template <bool F = false>
__device__ __forceinline__ double theta_of_sigma(const MemberDesc &d, double sgth, double s) {
  const bool mid = (s > d.sigma_sigma && s < 1.);
  const double gm = mid ? d.g_coeff * (1. - m_exp<F>(-d.alpha * (s - d.sigma_sigma))) : 0.;
  const double lin = d.zeta0 + s * (d.zeta1 - d.zeta0);
  const double g = (s <= d.sigma_sigma) ? 0. : (s >= 1. ? 1. : gm);
  const double f = (s < 0.) ? d.zeta0 : (s <= d.sigma_sigma ? lin : (s < 1. ? (1. - gm) * lin : 0.));
  return (sgth - f) / g;
}
// return_sigmatheta; sets *bad when p is outside [ptop, pbot]
__device__ __forceinline__ double sgth_of(const MemberDesc &d, double theta, double p, double pbot, double ptop,
                                          int *bad) {
  if (p <= ptop) return theta;
  if (p <= pbot) {
    double s = get_sigma(pbot, p, ptop);
    return f_sig(d, s) + g_sig(d, s) * theta;
  }
  *bad = 1;
  return 0.;
}

// The same three functions with the column's log(ptop/pbot) formed once by
// the caller: identical operations on identical operands, so identical
// results, with one logarithm fewer per call.
template <bool F = false>
__device__ __forceinline__ double get_sigma_l(double pbot, double p, double lr) { return m_log<F>(p / pbot) / lr; }
template <bool F = false>
__device__ __forceinline__ double get_p_sigma_l(double pbot, double s, double lr) { return pbot * m_exp<F>(s * lr); }
template <bool F = false>
__device__ __forceinline__ double sgth_of_l(const MemberDesc &d, double theta, double p, double pbot, double ptop,
                                            double lr, int *bad) {
  if (p <= ptop) return theta;
  if (p <= pbot) {
    const double s = get_sigma_l<F>(pbot, p, lr);
    // f_sig(s) + g_sig(s) * theta with g_sigma's exponential formed once (the
    // same operations as the two calls, so the same result)
    const bool mid = (s > d.sigma_sigma && s < 1.);
    const double gm = mid ? d.g_coeff * (1. - m_exp<F>(-d.alpha * (s - d.sigma_sigma))) : 0.;
    const double g = (s <= d.sigma_sigma) ? 0. : (s >= 1. ? 1. : gm);
    const double f = (s < 0.) ? d.zeta0 : (s <= d.sigma_sigma ? d.zeta0 + s * (d.zeta1 - d.zeta0)
                                                              : (s < 1. ? (1. - gm) * (d.zeta0 + s * (d.zeta1 - d.zeta0)) : 0.));
    return f + g * theta;
  }
  *bad = 1;
  return 0.;
}

// The block's tile as a logical (x, y, z), renumbered for the XCDs: a
// launch deals its blocks to the 8 XCDs round-robin by linear block id, so
// neighbouring tiles (which share the rows of their stencils) land on
// different L2s.  Here XCD x takes the contiguous run of logical ids
// [x*per + min(x, rem), ...) instead, a bijection of the ids whatever the
// hardware does (only the speed depends on the dealing).  Launches of more
// than XCD_REMAP_MAX blocks keep the plain order: measured on MI355X, the
// runs speed the point-stencil kernels of a 256x128x20 grid (2,400 blocks)
// by up to a third and slow those of 1024x512x64 (130,000 blocks) by 3-7 %;
// the row kernels (one latitude row per lane group, FFT filter in LDS) run
// slower with them at any size and keep the plain order.
constexpr unsigned XCD_REMAP_MAX = 8192;
__device__ __forceinline__ uint3 xcd_block() {
  const unsigned nx = gridDim.x, ny = gridDim.y, n = nx * ny * gridDim.z;
  if (n > XCD_REMAP_MAX) return make_uint3(blockIdx.x, blockIdx.y, blockIdx.z);
  const unsigned l = blockIdx.x + nx * (blockIdx.y + ny * blockIdx.z);
  const unsigned per = n >> 3, rem = n & 7, x = l & 7, r = l >> 3;
  const unsigned lp = x * per + (x < rem ? x : rem) + r;
  return make_uint3(lp % nx, (lp / nx) % ny, lp / (nx * ny));
}

// bisection search of an ascending or descending table (find_place_in_table)
__device__ __forceinline__ int table_place(int n, const double *x, double xx, double *dx) {
  int il = -1, iu = n;
  int ascend = (x[n - 1] >= x[0]);
  int lo = fcmp_d(xx, x[0]), hi = fcmp_d(xx, x[n - 1]);
  if (ascend) {
    if (lo <= 0) { *dx = x[1] - x[0]; return 0; }
    if (hi >= 0) { *dx = x[n - 1] - x[n - 2]; return n - 2; }
  } else {
    if (lo >= 0) { *dx = x[1] - x[0]; return 0; }
    if (hi <= 0) { *dx = x[n - 1] - x[n - 2]; return n - 2; }
  }
  while (iu - il > 1) {
    int im = (iu + il) >> 1;
    if ((fcmp_d(xx, x[im]) >= 0) == ascend) il = im; else iu = im;
  }
  *dx = x[il + 1] - x[il];
  return il;
}

// the cubic of a PCHIP table on the interval k that table_place found, h its
// width (splint_pchip, src/shared/epic_funcs_util.c:926-999)
__device__ __forceinline__ double pchip_at(const double *x, const double *y, const double *dd, int k, double h,
                                           double xx) {
  const double delta = (y[k + 1] - y[k]) / h;
  const double e1 = (dd[k] - delta) / h, e2 = (dd[k + 1] - delta) / h;
  const double c3 = (e1 + e2) / h, c2 = -(e1 + e1 + e2);
  const double t = xx - x[k];
  return y[k] + t * (dd[k] + t * (c2 + t * c3));
}

// Newtonian cooling time at -log p = x from the table (t_rad,
// src/shared/epic_sensible_heating.c:294-311): the end values past the ends
__device__ __forceinline__ double tab_cool_time(const MemberDesc &d, double x) {
  const int n = d.n_cool;
  if (x >= d.cool_x[n - 1]) return d.cool_y[n - 1];
  if (x <= d.cool_x[0]) return d.cool_y[0];
  double h;
  const int k = table_place(n, d.cool_x, x, &h);
  return pchip_at(d.cool_x, d.cool_y, d.cool_d, k, h, x);
}
// the sounding's temperature at -log p = x (get_sounding,
// src/shared/epic_funcs_diag.c:3258-3290): its bottom value below the table
__device__ __forceinline__ double tab_sounding_t(const MemberDesc &d, double x) {
  if (x < d.snd_x[0]) return d.snd_y[0];
  double h;
  const int k = table_place(d.n_snd, d.snd_x, x, &h);
  return pchip_at(d.snd_x, d.snd_y, d.snd_d, k, h, x);
}
// the relaxation toward the sounding, -(cp / t_cool)(T - T_sounding), cp the
// centred derivative of the ideal-gas enthalpy cpr T rgas (return_cp,
// src/shared/epic_funcs_diag.c:3311-3337;
// src/shared/epic_sensible_heating.c:103-130)
__device__ __forceinline__ double tab_relax(const MemberDesc &d, double p, double t) {
  const double x = -crm_log(p), dT = t * 1.e-6;
  double h2 = d.cpr * (t + dT), h1 = d.cpr * (t - dT);
  h2 *= d.rgas;
  h1 *= d.rgas;
  const double cp_over_time = ((h2 - h1) / (2. * dT)) / tab_cool_time(d, x);
  return -cp_over_time * (t - tab_sounding_t(d, x));
}
// whether some Newtonian relaxation forms HEAT3
__device__ __forceinline__ bool newt_relax_on(const MemberDesc &d) {
  return d.newt_cool_on && (d.is_hs || d.is_hj || d.is_tab);
}

// table_place for a table whose neighbouring entries differ by more than two
// of its largest ulps, in one direction: every comparison the search makes
// (within one ulp of the larger magnitude, and the argument lies inside the
// table's range once the end checks pass) is then monotone along the table,
// so the binary search's interval is the unique i with x[i] "<=" xx and
// x[i+1] ">" xx.  It is found here by walking from the guess g.
__device__ __forceinline__ int table_place_near(int n, const double *x, double xx, int g, double *dx) {
  int ascend = (x[n - 1] >= x[0]);
  int lo = fcmp_d(xx, x[0]), hi = fcmp_d(xx, x[n - 1]);
  if (ascend) {
    if (lo <= 0) { *dx = x[1] - x[0]; return 0; }
    if (hi >= 0) { *dx = x[n - 1] - x[n - 2]; return n - 2; }
  } else {
    if (lo >= 0) { *dx = x[1] - x[0]; return 0; }
    if (hi <= 0) { *dx = x[n - 1] - x[n - 2]; return n - 2; }
  }
  g = g < 0 ? 0 : (g > n - 2 ? n - 2 : g);
  if ((fcmp_d(xx, x[g]) >= 0) == ascend) {
    while (g + 1 <= n - 2 && (fcmp_d(xx, x[g + 1]) >= 0) == ascend) g++;
  } else {
    while (g > 0 && !((fcmp_d(xx, x[g]) >= 0) == ascend)) g--;
  }
  *dx = x[g + 1] - x[g];
  return g;
}
// whether x[0..n) is such a table (then table_place_near may replace table_place)
__device__ __forceinline__ bool table_well_separated(int n, const double *x) {
  double mx = 0.;
  for (int e = 0; e < n; e++) mx = dmax(mx, fabs(x[e]));
  int ex;
  frexp(mx, &ex);
  const double gap = 2. * ldexp(2.220446049250313e-16, ex + 1);
  const bool ascend = (x[n - 1] >= x[0]);
  for (int e = 0; e + 1 < n; e++) {
    const double dd = ascend ? x[e + 1] - x[e] : x[e] - x[e + 1];
    if (!(dd > gap)) return false;
  }
  return n >= 2;
}

// strided variant: table entries at x[k*stride]
__device__ __forceinline__ int table_place_s(int n, const double *x, int st, double xx, double *dx) {
  int il = -1, iu = n;
  int ascend = (x[(n - 1) * st] >= x[0]);
  int lo = fcmp_d(xx, x[0]), hi = fcmp_d(xx, x[(n - 1) * st]);
  if (ascend) {
    if (lo <= 0) { *dx = x[st] - x[0]; return 0; }
    if (hi >= 0) { *dx = x[(n - 1) * st] - x[(n - 2) * st]; return n - 2; }
  } else {
    if (lo >= 0) { *dx = x[st] - x[0]; return 0; }
    if (hi <= 0) { *dx = x[(n - 1) * st] - x[(n - 2) * st]; return n - 2; }
  }
  while (iu - il > 1) {
    int im = (iu + il) >> 1;
    if ((fcmp_d(xx, x[im * st]) >= 0) == ascend) il = im; else iu = im;
  }
  *dx = x[(il + 1) * st] - x[il * st];
  return il;
}

// Barrier among the lanes of one wavefront (LDS traffic of a single wave is
// ordered; the fences keep the compiler from moving accesses across it).
__device__ __forceinline__ void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// s + a[0] + a[1] + ... + a[n-1] (or s - a[0] - a[1] - ...), accumulated left
// to right as the reference's loops do; the loads are issued eight at a time
// ahead of the additions so the chain waits on additions only.
template <bool SUB = false, class T = double>
__device__ __forceinline__ T seq_acc(const T *a, int n, T s = T(0)) {
  int i = 0;
  for (; i + 8 <= n; i += 8) {
    const T v0 = a[i], v1 = a[i + 1], v2 = a[i + 2], v3 = a[i + 3];
    const T v4 = a[i + 4], v5 = a[i + 5], v6 = a[i + 6], v7 = a[i + 7];
    if (SUB) { s -= v0; s -= v1; s -= v2; s -= v3; s -= v4; s -= v5; s -= v6; s -= v7; }
    else { s += v0; s += v1; s += v2; s += v3; s += v4; s += v5; s += v6; s += v7; }
  }
  for (; i < n; i++) {
    if (SUB) s -= a[i];
    else s += a[i];
  }
  return s;
}

// field selector: which 3-D array of the descriptor
enum Fld : int {
  F_U, F_V, F_HDRY, F_THETA, F_P3, F_P2, F_THETA2, F_H3, F_GZ3, F_HEAT3, F_PV3, F_DIV, F_W3, F_DZDT3,
  F_GZ, F_SA, F_SB, F_SC, F_SD, F_UNITY, F_OLD, F_DUDT_T, F_DVDT_T
};

__device__ __forceinline__ double *fld_ptr(const MemberDesc &d, int f) {
  switch (f) {
    case F_U: return d.U;
    case F_V: return d.V;
    case F_HDRY: return d.HDRY;
    case F_THETA: return d.THETA;
    case F_P3: return d.P3;
    case F_P2: return d.P2;
    case F_THETA2: return d.THETA2;
    case F_H3: return d.H3;
    case F_GZ3: return d.GZ3;
    case F_HEAT3: return d.HEAT3;
    case F_PV3: return d.PV3;
    case F_DIV: return d.DIV;
    case F_W3: return d.W3;
    case F_DZDT3: return d.DZDT3;
    case F_GZ: return d.gz;
    case F_SA: return d.sA;
    case F_SB: return d.sB;
    case F_SC: return d.sC;
    case F_SD: return d.sD;
    case F_UNITY: return d.unity;
    case F_OLD: return d.old;
    // the mixed mode keeps one FP64 plane for the current step's tendency
    case F_DUDT_T: return d.mixed ? d.DUDT : d.DUDT + (size_t)d.dyn->it0 * d.N3;
    case F_DVDT_T: return d.mixed ? d.DVDT : d.DVDT + (size_t)d.dyn->it0 * d.N3;
  }
  return nullptr;
}

// FP32 views used by the mixed mode: the tendency planes and the scratch
// arrays (an FP32 view of a scratch array uses the first half of its bytes)
template <class T> __device__ __forceinline__ T *fld_ptr_t(const MemberDesc &d, int f);
template <> __device__ __forceinline__ double *fld_ptr_t<double>(const MemberDesc &d, int f) { return fld_ptr(d, f); }
template <> __device__ __forceinline__ float *fld_ptr_t<float>(const MemberDesc &d, int f) {
  switch (f) {
    case F_DUDT_T: return d.DUDTf + (size_t)d.dyn->it0 * d.N3;
    case F_DVDT_T: return d.DVDTf + (size_t)d.dyn->it0 * d.N3;
    case F_SA: return reinterpret_cast<float *>(d.sA);
    case F_SB: return reinterpret_cast<float *>(d.sB);
    case F_SC: return reinterpret_cast<float *>(d.sC);
    case F_SD: return reinterpret_cast<float *>(d.sD);
  }
  return nullptr;
}

// launch geometry shared by host and device
constexpr int TPB = 256;

}  // namespace epic
