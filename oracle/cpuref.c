/*
 * cpuref.c -- CPU FP64 restatement of EPIC's dynamical-core timestep
 * (src/shared/epic_timestep.c:62-390 and its callees) for a dry globe with
 * the default schemes: Sadourny PV averaging, finite-volume pressure
 * gradient, Hsu-Konor-Arakawa advection, 3rd-order Adams-Bashforth,
 * del^4 hyperviscosity, polar zonal filter, Held-Suarez forcing.
 *
 * TEST INFRASTRUCTURE ONLY.  This is the oracle of the parity tests; the
 * product never links or calls it.
 *
 * Every phase states the arithmetic in the reference's operation order, so
 * that with the same compiler flags (no FMA contraction) and the same libm
 * it agrees with the reference build bit for bit; tests/test_oracle.py pins
 * it against oracle/_ref where that builds and against the committed golden
 * fixtures everywhere.  Where the reference reads state left over from a
 * previous call (static buffers), that state is kept here explicitly and
 * the place is named.  Layout is the reference's padded (K,J,I) with
 * Kshift = 0, Jshift = -1, Ishift = 0 (src/shared/epic_funcs_diag.c:654-664).
 */
#include "cpuref.h"

#include <float.h>
#include <math.h>
#ifdef CPUREF_LIBM_NUDGE
#include "libm_nudge.h"
#endif
#ifdef CPUREF_CRMATH
#include "libm_cr.h"
#endif
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#define DEGR (M_PI / 180.)
#define LAT0 45.
#define R_GAS 8.314472e+3
#define PV_COEF (1. / 12.)

enum { FK_U, FK_V, FK_H, FK_THETA, FK_DEF };

struct cpuref {
  epic_case c;
  int ni, nj, nk, Ia, N2, N3, k_sigma, k_sponge;
  double dt;
  int is_hs, is_gas, is_hj;
  /* Newtonian cooling toward the case's sounding: the sounding (sx = -log p,
   * sy = T, sz its PCHIP derivatives) and the cooling-time table (cx, cy, cz) */
  int is_tab, ns, nc;
  double *sx, *sy, *sz, *cx, *cy, *cz;
  /* fields */
  double *U, *V, *DUDT[3], *DVDT[3], *HDRY, *THETA, *P3, *P2, *THETA2, *H3;
  double *T2, *T3, *RHO2, *RHO3, *EXNER3, *GZ3, *MONT3, *HEAT3, *PV3, *DIV, *W3, *DZDT3, *RI2;
  double *USPIN, *GZS;
  /* scratch and persistent statics of the reference */
  double *gz, *sum_div, *unity, *old, *varz;
  double *b2[8];
  double *gzu2d, *gzv2d, *gz2d;
  double *w_wm, *w_wp, *w_a, *w_pred, *w_min, *w_ff, *w_d1, *w_d2, *w_aaa;
  /* tables */
  double *hl_h, *hl_pv, *mn_inv;
  double *hcx, *hcy, *hcz, *vcx, *vcy, *vcz;
  double coef[EPIC_MAX_NU_ORDER + 1], m0_zonal, nudiv, m0_div, eps13;
  double *fdata;
  double *tx, *ty, *tz;
  int it0, itm1, itm2;
  long itime;
  int err;
};

#define IX(K, J, I) ((size_t)(I) + (size_t)((J) + 1) * R->Ia + (size_t)(K) * R->N2)
#define A3(a, K, J, I) (a)[IX(K, J, I)]
#define A2(a, J, I) (a)[(size_t)(I) + (size_t)((J) + 1) * R->Ia]

static inline double mx(double x, double y) { return x > y ? x : y; }
static inline double mnv(double x, double y) { return x < y ? x : y; }

/* ------------------------------------------------------------------ util */

/* fcmp, src/shared/epic_funcs_util.c:4275-4346 */
int cpuref_fcmp(double x1, double x2) {
  int e;
  frexp(fabs(x1) > fabs(x2) ? x1 : x2, &e);
  double delta = ldexp(DBL_EPSILON, e), d = x1 - x2;
  if (d > delta) return 1;
  if (d < -delta) return -1;
  return 0;
}

/* bisection table search, src/shared/epic_funcs_util.c:2908-2990 */
static int table_place(int n, const double *x, double xx, double *dx) {
  int il = -1, iu = n;
  int ascend = (x[n - 1] >= x[0]);
  int lo = cpuref_fcmp(xx, x[0]), hi = cpuref_fcmp(xx, x[n - 1]);
  if (ascend) {
    if (lo <= 0) { *dx = x[1] - x[0]; return 0; }
    if (hi >= 0) { *dx = x[n - 1] - x[n - 2]; return n - 2; }
  } else {
    if (lo >= 0) { *dx = x[1] - x[0]; return 0; }
    if (hi <= 0) { *dx = x[n - 1] - x[n - 2]; return n - 2; }
  }
  while (iu - il > 1) {
    int im = (iu + il) >> 1;
    if ((cpuref_fcmp(xx, x[im]) >= 0) == ascend) il = im; else iu = im;
  }
  *dx = x[il + 1] - x[il];
  return il;
}

/* sign test of Fritsch-Carlson PCHIP, src/shared/epic_funcs_util.c:1000-1054 */
static double sgn_prod(double a, double b) {
  if (a == 0.) return 0.;
  if (a < 0.) return b < 0. ? 1. : (b == 0. ? 0. : -1.);
  return b < 0. ? -1. : (b == 0. ? 0. : 1.);
}

/* Monotone piecewise-cubic derivatives (Fritsch & Carlson 1980, Brodlie's
 * weights, three-point shape-preserving ends), following spline_pchip,
 * src/shared/epic_funcs_util.c:721-925. */
int cpuref_pchip(int n, const double *x, const double *y, double *d) {
  if (n < 2) return -1;
  for (int i = 1; i < n; i++)
    if (x[i] <= x[i - 1]) return -1;
  double h1 = x[1] - x[0], del1 = (y[1] - y[0]) / h1;
  if (n == 2) { d[0] = d[1] = del1; return 0; }
  double h2 = x[2] - x[1], del2 = (y[2] - y[1]) / h2;
  double hsum = h1 + h2, w1 = (h1 + hsum) / hsum, w2 = -h1 / hsum;
  d[0] = w1 * del1 + w2 * del2;
  if (sgn_prod(d[0], del1) <= 0.0) d[0] = 0.0;
  else if (sgn_prod(del1, del2) < 0.0) {
    double dmax = 3.0 * del1;
    if (fabs(dmax) < fabs(d[0])) d[0] = dmax;
  }
  for (int i = 2; i <= n - 1; i++) {
    if (2 < i) {
      h1 = h2;
      h2 = x[i] - x[i - 1];
      hsum = h1 + h2;
      del1 = del2;
      del2 = (y[i] - y[i - 1]) / h2;
    }
    d[i - 1] = 0.0;
    if (sgn_prod(del1, del2) > 0.0) {
      double hsumt3 = 3.0 * hsum;
      double a1 = (hsum + h1) / hsumt3, a2 = (hsum + h2) / hsumt3;
      double dmax = mx(fabs(del1), fabs(del2)), dmin = mnv(fabs(del1), fabs(del2));
      double r1 = del1 / dmax, r2 = del2 / dmax;
      d[i - 1] = dmin / (a1 * r1 + a2 * r2);
    }
  }
  w1 = -h2 / hsum;
  w2 = (h2 + hsum) / hsum;
  d[n - 1] = w1 * del1 + w2 * del2;
  if (sgn_prod(d[n - 1], del2) <= 0.0) d[n - 1] = 0.0;
  else if (sgn_prod(del1, del2) < 0.0) {
    double dmax = 3.0 * del2;
    if (fabs(dmax) < fabs(d[n - 1])) d[n - 1] = dmax;
  }
  return 0;
}

/* cubic Hermite evaluation on [x0,x1], src/shared/epic_funcs_util.c:926-999 */
static double hermite(const double *x, const double *y, const double *d, double h, double xx) {
  double delta = (y[1] - y[0]) / h;
  double e1 = (d[0] - delta) / h, e2 = (d[1] - delta) / h;
  double c3 = (e1 + e2) / h, c2 = -(e1 + e1 + e2);
  double t = xx - x[0];
  return y[0] + t * (d[0] + t * (c2 + t * c3));
}

double cpuref_pchip_eval(int n, const double *x, const double *y, const double *d, double xx) {
  double h;
  int k = table_place(n, x, xx, &h);
  return hermite(x + k, y + k, d + k, h, xx);
}

/* Radix-2 complex FFT of nn points, interleaved re/im, 1-based, in place,
 * with the trigonometric recurrence for the twiddles; the Danielson-Lanczos
 * scheme as Numerical Recipes states it (four1,
 * src/shared/epic_funcs_util.c:4023-4077). */
static void cfft(double *data, unsigned long nn, int isign) {
  unsigned long n = nn << 1, j = 1;
  for (unsigned long i = 1; i < n; i += 2) {
    if (j > i) {
      double t = data[j]; data[j] = data[i]; data[i] = t;
      t = data[j + 1]; data[j + 1] = data[i + 1]; data[i + 1] = t;
    }
    unsigned long m = n >> 1;
    while (m >= 2 && j > m) { j -= m; m >>= 1; }
    j += m;
  }
  unsigned long mmax = 2;
  while (n > mmax) {
    unsigned long istep = mmax << 1;
    double theta = isign * (2 * M_PI / mmax);
    double wtemp = sin(0.5 * theta);
    double wpr = -2.0 * wtemp * wtemp, wpi = sin(theta), wr = 1.0, wi = 0.0;
    for (unsigned long m = 1; m < mmax; m += 2) {
      for (unsigned long i = m; i <= n; i += istep) {
        j = i + mmax;
        double tr = wr * data[j] - wi * data[j + 1];
        double ti = wr * data[j + 1] + wi * data[j];
        data[j] = data[i] - tr;
        data[j + 1] = data[i + 1] - ti;
        data[i] += tr;
        data[i + 1] += ti;
      }
      wr = (wtemp = wr) * wpr - wi * wpi + wr;
      wi = wi * wpr + wtemp * wpi + wi;
    }
    mmax = istep;
  }
}

/* Real FFT by the half-length complex transform and the split step,
 * NR packing (data[1] = DC, data[2] = Nyquist); the inverse needs a 2/n
 * scale (realft, src/shared/epic_funcs_util.c:4084-4135). */
void cpuref_realft(double *data, unsigned long n, int isign) {
  double c1 = 0.5, c2, theta = M_PI / (double)(n >> 1);
  if (isign == 1) { c2 = -0.5; cfft(data, n >> 1, 1); }
  else { c2 = 0.5; theta = -theta; }
  double wtemp = sin(0.5 * theta);
  double wpr = -2.0 * wtemp * wtemp, wpi = sin(theta);
  double wr = 1.0 + wpr, wi = wpi;
  unsigned long np3 = n + 3;
  double h1r;
  for (unsigned long i = 2; i <= (n >> 2); i++) {
    unsigned long i1 = i + i - 1, i2 = 1 + i1, i3 = np3 - i2, i4 = 1 + i3;
    h1r = c1 * (data[i1] + data[i3]);
    double h1i = c1 * (data[i2] - data[i4]);
    double h2r = -c2 * (data[i2] + data[i4]);
    double h2i = c2 * (data[i1] - data[i3]);
    data[i1] = h1r + wr * h2r - wi * h2i;
    data[i2] = h1i + wr * h2i + wi * h2r;
    data[i3] = h1r - wr * h2r + wi * h2i;
    data[i4] = -h1i + wr * h2i + wi * h2r;
    wr = (wtemp = wr) * wpr - wi * wpi + wr;
    wi = wi * wpr + wtemp * wpi + wi;
  }
  if (isign == 1) {
    data[1] = (h1r = data[1]) + data[2];
    data[2] = h1r - data[2];
  } else {
    data[1] = c1 * ((h1r = data[1]) + data[2]);
    data[2] = c1 * (h1r - data[2]);
    cfft(data, n >> 1, -1);
  }
}

/* --------------------------------------------------- hybrid coordinate */
/* src/shared/epic_funcs_diag.c:1238-1314, 1751-1770 */
static double get_sigma(double pbot, double p, double ptop) { return log(p / pbot) / log(ptop / pbot); }
static double get_p_sigma(double pbot, double s, double ptop) { return pbot * exp(s * log(ptop / pbot)); }
static double g_sig(const cpuref *R, double s) {
  /* (exp), (log), (pow) in parentheses: values the GPU side takes from
   * host tables, formed with the host libm in every build of the oracle */
  const double coeff = 1. / (1 - (exp)(-R->c.hybrid_alpha * (1. - R->c.sigma_sigma)));
  if (s <= R->c.sigma_sigma) return 0.;
  if (s >= 1.) return 1.;
  return coeff * (1. - exp(-R->c.hybrid_alpha * (s - R->c.sigma_sigma)));
}
static double f_sig(const cpuref *R, double s) {
  if (s < 0.) return R->c.zeta0;
  if (s <= R->c.sigma_sigma) return R->c.zeta0 + s * (R->c.zeta1 - R->c.zeta0);
  if (s < 1.) return (1. - g_sig(R, s)) * (R->c.zeta0 + s * (R->c.zeta1 - R->c.zeta0));
  return 0.;
}
static double sgth_of(cpuref *R, double theta, double p, double pbot, double ptop) {
  if (p <= ptop) return theta;
  if (p <= pbot) {
    double s = get_sigma(pbot, p, ptop);
    return f_sig(R, s) + g_sig(R, s) * theta;
  }
  R->err |= 2;
  return NAN;
}

/* ----------------------------------------------------------- boundaries */
/* zonal periodicity, src/single/s_epic_functions.c:39-94 */
static void bc2(cpuref *R, double *a) {
  for (int J = 0; J <= R->nj + 1; J++) {
    A2(a, J, 0) = A2(a, J, R->ni);
    A2(a, J, R->ni + 1) = A2(a, J, 1);
  }
}
static void bc3(cpuref *R, double *a) {
  for (int K = 0; K <= R->nk + 1; K++) bc2(R, a + (size_t)K * R->N2);
}

/* --------------------------------------------------------- zonal filter */
/* src/shared/epic_adjust.c:102-242 (initialisation) */
static void filter_init(cpuref *R) {
  epic_case *c = &R->c;
  double lat0 = LAT0 * DEGR;
  double rln = c->re / sqrt(1. + (pow)(c->rp / c->re * tan(lat0), 2.));
  double rlt = rln / (cos(lat0) * ((pow)(sin(lat0), 2.) + (pow)(c->re / c->rp * cos(lat0), 2.)));
  double m0 = 1. / (rln * c->dln * DEGR), n0 = 1. / (rlt * c->dlt * DEGR);
  for (int pass = 0; pass < 2; pass++) {
    double *hl = pass ? R->hl_pv : R->hl_h;
    for (int J = pass ? 1 : 0; J <= R->nj; J++) {
      int jj = pass ? 2 * J : 2 * J + 1;
      A2(hl, J, 1) = 1.;
      double r = c->mn[jj] / (c->m[jj] * c->n[jj]);
      double t0 = (c->n[jj] / n0) / (c->m[jj] / m0);
      t0 = (pow)(t0, 2.00);
      for (int I = 2; I <= R->ni / 2 + 1; I++) {
        double t1 = sin((I - 1) * (c->dln) * DEGR / 2.);
        t1 = (pow)(t1, -2.00);
        t1 *= r * t0;
        t1 = (t1 < 1.) ? t1 : 1.;
        A2(hl, J, I) = t1;
      }
    }
  }
}

/* FFT part over one plane: rows with |lat| >= LAT0 */
static void filter_plane(cpuref *R, double *p, int vgrid) {
  int ni = R->ni;
  double *data = R->fdata;
  const double *hl = vgrid ? R->hl_pv : R->hl_h;
  for (int J = vgrid ? 1 : 0; J <= R->nj; J++) {
    double lat = vgrid ? R->c.lat[2 * J] : R->c.lat[2 * J + 1];
    if (fabs(lat) < LAT0) continue;
    for (int I = 1; I <= ni; I++) data[I] = A2(p, J, I);
    cpuref_realft(data, ni, 1);
    data[1] *= A2(hl, J, 1);
    for (int I = 2; I <= ni / 2; I++) {
      data[2 * I - 1] *= A2(hl, J, I);
      data[2 * I] *= A2(hl, J, I);
    }
    data[2] *= A2(hl, J, ni / 2 + 1);
    cpuref_realft(data, ni, -1);
    for (int I = 1; I <= ni; I++) A2(p, J, I) = data[I] * (2. / (ni));
  }
}

/* zonal_filter(index, plane, NULL, TWODIM), src/shared/epic_adjust.c:251-257,398-478 */
static void filter2(cpuref *R, double *plane, int kind) {
  filter_plane(R, plane, kind == FK_V);
  bc2(R, plane);
}

/* zonal_filter(index, cube, gz, THREEDIM): remove a PCHIP profile of the
 * variable against the filtered geopotential, taken at the column of lowest
 * bottom GZ of each row, filter, add it back (src/shared/epic_adjust.c:258-478). */
static void filter3(cpuref *R, double *b, int kind) {
  int ni = R->ni, nk = R->nk;
  int jstart = (kind == FK_V) ? 1 : 0;
  int kstart = (kind == FK_THETA) ? R->k_sigma - 1 : (kind == FK_DEF ? 0 : 1), kend = nk;
  double *gz = R->gz, *varz = R->varz;
  int n = kend - kstart + 1;
  for (int J = jstart; J <= R->nj; J++) {
    int Imin = 1;
    double gzmin = DBL_MAX;
    for (int I = 1; I <= ni; I++)
      if (A3(gz, nk, J, I) < gzmin) { Imin = I; gzmin = A3(gz, nk, J, I); }
    for (int pass = 0; pass <= ni; pass++) {
      /* pass 0 builds the table at Imin, passes 1..ni evaluate at I = pass */
      int I = pass ? pass : Imin;
      for (int K = kend; K >= kstart; K--) {
        double geezee;
        if (kind == FK_U) geezee = .5 * (A3(gz, K, J, I) + A3(gz, K, J, I - 1));
        else if (kind == FK_V) geezee = .5 * (A3(gz, K, J, I) + A3(gz, K, J - 1, I));
        else if (kind == FK_H) geezee = .5 * (A3(gz, K, J, I) + A3(gz, K - 1, J, I));
        else geezee = A3(gz, K, J, I);
        if (pass == 0) {
          R->tx[kend - K] = geezee;
          R->ty[kend - K] = (kind == FK_H) ? log(A3(b, K, J, I)) : A3(b, K, J, I);
        } else {
          double h;
          int kay = table_place(n, R->tx, geezee, &h);
          double v = hermite(R->tx + kay, R->ty + kay, R->tz + kay, h, geezee);
          if (kind == FK_H) v = exp(v);
          A3(varz, K, J, I) = v;
          A3(b, K, J, I) -= v;
        }
      }
      if (pass == 0 && cpuref_pchip(n, R->tx, R->ty, R->tz)) R->err |= 4;
    }
  }
  for (int K = kstart; K <= kend; K++) filter_plane(R, b + (size_t)K * R->N2, kind == FK_V);
  for (int K = kstart; K <= kend; K++)
    for (int J = jstart; J <= R->nj; J++)
      for (int I = 1; I <= ni; I++) A3(b, K, J, I) += A3(varz, K, J, I);
  bc3(R, b);
}

/* ------------------------------------------------------- diagnostics */
/* get_h layer value, src/shared/epic_funcs_diag.c:1780-1853 */
static double h_layer(cpuref *R, int K, int J, int I) {
  return ((double)A3(R->P3, K, J, I) - (double)A3(R->P3, K - 1, J, I)) /
         ((double)R->c.g[2 * J + 1] * (double)R->c.dsgth[2 * K]);
}

/* set_p2etc, src/shared/epic_funcs_diag.c:3352-3569 (dry: H2 == HDRY, P3 == PDRY3) */
static void set_p2etc(cpuref *R, int update_theta) {
  epic_case *c = &R->c;
  int nk = R->nk, ks = R->k_sigma;
  for (int J = 0; J <= R->nj; J++) {
    double g = c->g[2 * J + 1];
    for (int I = 0; I <= R->ni + 1; I++) {
      for (int K = 1; K <= nk; K++)
        A3(R->P3, K, J, I) = A3(R->P3, K - 1, J, I) + c->dsgth[2 * K] * g * A3(R->HDRY, K, J, I);
      double ptop = A3(R->P3, 0, J, I), pbot = A3(R->P3, nk, J, I);
      for (int K = ks - 1; K < nk; K++) {
        double sigma = (c->sigmatheta[2 * K + 1] - c->zeta0) / (c->zeta1 - c->zeta0);
        A3(R->P3, K, J, I) = get_p_sigma(pbot, sigma, ptop);
      }
      for (int K = 1; K <= nk; K++) A3(R->HDRY, K, J, I) = mx(c->h_min[K], h_layer(R, K, J, I));
    }
  }
  if (update_theta) {
    for (int J = 0; J <= R->nj; J++)
      for (int I = 0; I <= R->ni + 1; I++) {
        double ptop = A3(R->P3, 0, J, I), pbot = A3(R->P3, nk, J, I);
        for (int K = 0; K < ks - 1; K++) {
          double sigma = get_sigma(pbot, A3(R->P3, K, J, I), ptop);
          A3(R->THETA, K, J, I) = ((double)c->sigmatheta[2 * K + 1] - f_sig(R, sigma)) / g_sig(R, sigma);
        }
      }
  }
  double kp1 = c->kappa + 1.;
  for (int J = 0; J <= R->nj; J++)
    for (int I = 0; I <= R->ni + 1; I++) {
      for (int K = 1; K <= nk; K++) {
        /* onto_kk(P2_INDEX) (src/shared/epic_funcs_diag.c:2600-2616) */
        double top = A3(R->P3, K - 1, J, I), bot = A3(R->P3, K, J, I), p2;
        if (cpuref_fcmp(bot, top) == 0) p2 = .5 * (bot + top);
        else if (top < 0. || bot < 0.) { R->err |= 2; p2 = NAN; }
        else p2 = pow((pow(bot, kp1) - pow(top, kp1)) / (kp1 * (bot - top)), 1. / c->kappa);
        A3(R->P2, K, J, I) = p2;
      }
      A3(R->P2, 0, J, I) = A3(R->P3, 0, J, I) * A3(R->P3, 0, J, I) / A3(R->P2, 1, J, I);
      for (int K = 1; K <= nk; K++) {
        double h;
        if (K < nk) {
          int kk = 2 * K + 1;
          double topwt = c->dsgth[kk - 1], botwt = c->dsgth[kk + 1];
          h = (topwt * h_layer(R, K, J, I) + botwt * h_layer(R, K + 1, J, I)) / (topwt + botwt);
        } else {
          h = h_layer(R, K, J, I);
        }
        A3(R->H3, K, J, I) = h;
      }
      A3(R->H3, 0, J, I) = A3(R->H3, 1, J, I);
    }
  if (update_theta) {
    int nkay = nk - (ks - 1) + 1;
    for (int K = nk; K >= ks - 1; K--) R->tx[nk - K] = c->sigmatheta[2 * K + 1];
    for (int J = 0; J <= R->nj; J++)
      for (int I = 0; I <= R->ni + 1; I++) {
        double ptop = A3(R->P3, 0, J, I), pbot = A3(R->P3, nk, J, I);
        for (int K = 1; K < ks - 1; K++) {
          double sigma = get_sigma(pbot, A3(R->P2, K, J, I), ptop);
          A3(R->THETA2, K, J, I) = ((double)c->sigmatheta[2 * K] - f_sig(R, sigma)) / g_sig(R, sigma);
        }
        A3(R->THETA2, ks - 1, J, I) = .5 * (A3(R->THETA, ks - 1, J, I) + A3(R->THETA, ks - 2, J, I));
        for (int K = nk; K >= ks - 1; K--) R->ty[nk - K] = A3(R->THETA, K, J, I);
        for (int K = ks; K <= nk; K++) {
          double sg = c->sigmatheta[2 * K], h;
          int kay = table_place(nkay, R->tx, sg, &h);
          /* linint, src/shared/epic_funcs_util.c:636-647 */
          A3(R->THETA2, K, J, I) = R->ty[kay] + (R->ty[kay + 1] - R->ty[kay]) * (sg - R->tx[kay]) / h;
        }
      }
  }
}

/* get_kin, src/shared/epic_funcs_diag.c:2694-2847: Sadourny with the
 * Suarez-Takacs correction off the poles; Arakawa-Lamb the centred value */
static double kin_at(cpuref *R, const double *u, const double *v, int J, int I) {
  const double *mn = R->c.mn, *mni = R->mn_inv;
  int jj = 2 * J;
  double u2 = A2(u, J, I), u4 = A2(u, J, I + 1), v2 = A2(v, J, I), v4 = A2(v, J + 1, I);
  double kc = (u2 * u2 + u4 * u4 + mn[jj + 1] * (v2 * v2 * mni[jj] + v4 * v4 * mni[jj + 2])) * .25;
  if (R->c.pv_scheme || J == 0 || J == R->nj) return kc;
  u2 = .5 * (A2(u, J + 1, I) + A2(u, J - 1, I));
  u4 = .5 * (A2(u, J + 1, I + 1) + A2(u, J - 1, I + 1));
  v2 = .5 * (A2(v, J, I + 1) + A2(v, J, I - 1));
  v4 = .5 * (A2(v, J + 1, I + 1) + A2(v, J + 1, I - 1));
  double ks = (u2 * u2 + u4 * u4 + mn[jj + 1] * (v2 * v2 * mni[jj] + v4 * v4 * mni[jj + 2])) * .25;
  return (5. / 6.) * kc + (1. / 6.) * ks;
}

/* vorticity (POTENTIAL or ABSOLUTE) on one plane,
 * src/shared/epic_funcs_diag.c:4131-4445 (globe with both poles) */
static void vorticity(cpuref *R, int potential, const double *uu, const double *vv, const double *hh,
                      double *pv) {
  epic_case *c = &R->c;
  int ni = R->ni, nj = R->nj;
  for (int J = 1; J <= nj; J++) {
    double f2j = c->f[2 * J], m2j = c->m[2 * J], n2j = c->n[2 * J];
    double mm1i = 1. / c->m[2 * J - 1], mp1i = 1. / c->m[2 * J + 1];
    double mnm1i = 1. / c->mn[2 * J - 1], mnp1i = 1. / c->mn[2 * J + 1];
    double mn_pv = .5 / (mnm1i + mnp1i);
    for (int I = 1; I <= ni; I++) {
      double ze = m2j * ((A2(vv, J, I) - A2(vv, J, I - 1)) +
                         n2j * (A2(uu, J - 1, I) * mm1i - A2(uu, J, I) * mp1i));
      A2(pv, J, I) = (ze + f2j);
    }
    if (potential)
      for (int I = 1; I <= ni; I++) {
        double h_pv = ((A2(hh, J, I) + A2(hh, J, I - 1)) * mnp1i + (A2(hh, J - 1, I) + A2(hh, J - 1, I - 1)) * mnm1i) * mn_pv;
        A2(pv, J, I) /= h_pv;
      }
  }
  /* south pole: circulation of the adjacent u-row */
  double ze = 0.;
  for (int I = 1; I <= ni; I++) ze -= A2(uu, 0, I);
  ze *= c->mn[0] / ((c->m[1]) * (double)(ni));
  double pvp;
  if (potential) {
    double h = 0.;
    for (int I = 1; I <= ni; I++) h += A2(hh, 0, I);
    h /= (double)(ni);
    pvp = (ze + c->f[0]) / h;
  } else {
    pvp = ze + c->f[0];
  }
  for (int I = 1; I <= ni; I++) A2(pv, 0, I) = pvp;
  /* north pole */
  ze = 0.;
  for (int I = 1; I <= ni; I++) ze += A2(uu, nj, I);
  ze *= c->mn[2 * (nj + 1)] / (c->m[2 * nj + 1] * (double)(ni));
  if (potential) {
    double h = 0.;
    for (int I = 1; I <= ni; I++) h += A2(hh, nj, I);
    h /= (double)(ni);
    pvp = (ze + c->f[2 * (nj + 1)]) / h;
  } else {
    pvp = ze + c->f[2 * (nj + 1)];
  }
  for (int I = 1; I <= ni; I++) A2(pv, nj + 1, I) = pvp;
  bc2(R, pv);
}

/* divergence, src/shared/epic_funcs_diag.c:4064-4117 */
static void divergence(cpuref *R, const double *uu, const double *vv, double *di) {
  epic_case *c = &R->c;
  for (int J = 0; J <= R->nj; J++) {
    int jj = 2 * J + 1;
    double n_inv = 1. / c->n[jj], mn = c->mn[jj];
    double m2j_inv = (J == 0) ? 0. : 1. / c->m[jj - 1];
    double m2jp2_inv = (J == R->nj) ? 0. : 1. / c->m[jj + 1];
    for (int I = 1; I <= R->ni; I++)
      A2(di, J, I) = mn * ((A2(uu, J, I + 1) * n_inv - A2(uu, J, I) * n_inv) +
                           (A2(vv, J + 1, I) * m2jp2_inv - A2(vv, J, I) * m2j_inv));
  }
  bc2(R, di);
}

/* PV weights of uv_core's pv*uh (v grid) and pv*vh (u grid) terms,
 * include/epic_pv_schemes.h:34-62: Sadourny (1975), three-point sums over
 * 12, the extra terms zero; Arakawa and Lamb (1981), weighted four-point
 * sums over 24 with the extra terms PH1, PH2 (v grid) and EP1, EP2 (u
 * grid); each sum in the header's order */
typedef struct { double al, be, ga, de, e1, e2, coef; } pvw;
static pvw pv_v(const cpuref *R, const double *pv, int J, int I) {
#define P(J_, I_) A2(pv, J_, I_)
  pvw w;
  if (R->c.pv_scheme) {
    w.ga = 2. * P(J + 1, I + 1) + P(J + 1, I) + 2. * P(J, I) + P(J, I + 1);
    w.de = P(J + 1, I + 1) + 2. * P(J + 1, I) + P(J, I) + 2. * P(J, I + 1);
    w.al = 2. * P(J, I + 1) + P(J, I) + 2. * P(J - 1, I) + P(J - 1, I + 1);
    w.be = P(J, I + 1) + 2. * P(J, I) + P(J - 1, I) + 2. * P(J - 1, I + 1);
    w.e1 = -P(J + 1, I + 1) + P(J + 1, I) + P(J, I) - P(J, I + 1);
    w.e2 = -P(J, I + 1) + P(J, I) + P(J - 1, I) - P(J - 1, I + 1);
    w.coef = 1. / 24.;
  } else {
    w.ga = (P(J, I) + P(J, I + 1) + P(J + 1, I + 1));
    w.de = (P(J, I) + P(J + 1, I) + P(J, I + 1));
    w.al = (P(J - 1, I) + P(J, I) + P(J, I + 1));
    w.be = (P(J, I) + P(J, I + 1) + P(J - 1, I + 1));
    w.e1 = w.e2 = 0.;
    w.coef = PV_COEF;
  }
  return w;
}
static pvw pv_u(const cpuref *R, const double *pv, int J, int I) {
  pvw w;
  if (R->c.pv_scheme) {
    w.al = 2. * P(J + 1, I + 1) + P(J + 1, I) + 2. * P(J, I) + P(J, I + 1);
    w.be = P(J + 1, I) + 2. * P(J + 1, I - 1) + P(J, I - 1) + 2. * P(J, I);
    w.ga = 2. * P(J + 1, I) + P(J + 1, I - 1) + 2. * P(J, I - 1) + P(J, I);
    w.de = P(J + 1, I + 1) + 2. * P(J + 1, I) + P(J, I) + 2. * P(J, I + 1);
    w.e1 = P(J + 1, I + 1) + P(J + 1, I) - P(J, I) - P(J, I + 1);
    w.e2 = P(J + 1, I) + P(J + 1, I - 1) - P(J, I - 1) - P(J, I);
    w.coef = 1. / 24.;
  } else {
    w.al = (P(J, I) + P(J + 1, I) + P(J + 1, I + 1));
    w.be = (P(J + 1, I - 1) + P(J + 1, I) + P(J, I));
    w.ga = (P(J, I - 1) + P(J, I) + P(J + 1, I));
    w.de = (P(J, I) + P(J + 1, I) + P(J, I + 1));
    w.e1 = w.e2 = 0.;
    w.coef = PV_COEF;
  }
  return w;
#undef P
}

/* gz_from_u for gas giants, src/shared/epic_funcs_diag.c:4460-4610 */
static void gz_from_u(cpuref *R, const double *u1d, const double *p1d, const double *r1d, double *gz1d) {
  epic_case *c = &R->c;
  int nj = R->nj, nk = R->nk, j0 = c->jtp;
  double *u2d = R->gzu2d, *v2d = R->gzv2d;
  for (int J = 0; J <= nj; J++)
    for (int I = 0; I <= R->ni + 1; I++) A2(u2d, J, I) = u1d[J];
  double *pv = R->PV3 + (size_t)(nk + 1) * R->N2;
  vorticity(R, 0, u2d, v2d, NULL, pv);
  double udy[4096], pvh[4096], bern[4096];
  for (int J = 0; J <= nj; J++) udy[J] = A2(u2d, J, 1) / c->n[2 * J + 1];
  memset(pvh, 0, sizeof(double) * (nj + 2));
  const int I = 1;
  for (int J = 1; J <= nj; J++) {
    const pvw w = pv_v(R, pv, J, I);
    pvh[J] = (w.ga * udy[J] + w.de * udy[J] + w.al * udy[J - 1] + w.be * udy[J - 1]) * w.coef;
    pvh[J] += (p1d[J] - p1d[J - 1]) * 2. / (r1d[J] + r1d[J - 1]);
  }
  double kin0 = 0.;
  if (j0 >= 0 && j0 <= nj) kin0 = kin_at(R, u2d, v2d, j0, 1);
  bern[j0] = 0. + kin0;
  for (int J = j0 + 1; J <= nj; J++) bern[J] = bern[J - 1] - pvh[J];
  for (int J = j0 - 1; J >= 0; J--) bern[J] = bern[J + 1] + pvh[J + 1];
  for (int J = 0; J <= nj; J++) gz1d[J] = bern[J] - kin_at(R, u2d, v2d, J, 1);
}

/* store_pgrad_vars, src/shared/epic_funcs_diag.c:3595-3950 (ideal gas, x_h2 = 0) */
static void store_pgrad_vars(cpuref *R) {
  epic_case *c = &R->c;
  int nk = R->nk, nj = R->nj, ni = R->ni;
  if (R->itime > 1) {
    for (int K = 1; K <= nk; K++)
      for (int J = 0; J <= nj; J++)
        for (int I = 1; I <= ni; I++) A3(R->DZDT3, K, J, I) = A3(R->GZ3, K, J, I);
  } else {
    memset(R->DZDT3, 0, sizeof(double) * R->N3);
  }
  double mu3 = 1. / (c->rgas / R_GAS);
  double mu2 = .5 * (mu3 + mu3);
  for (int K = 0; K <= nk; K++)
    for (int J = 0; J <= nj; J++)
      for (int I = 0; I <= ni + 1; I++) {
        double th = A3(R->THETA, K, J, I), p = A3(R->P3, K, J, I);
        double t = th * pow(p / c->press0, c->kappa);
        A3(R->T3, K, J, I) = t;
        A3(R->EXNER3, K, J, I) = c->cp * t / th;
        A3(R->RHO3, K, J, I) = p * mu3 / (R_GAS * t);
      }
  for (int K = 1; K <= nk; K++)
    for (int J = 0; J <= nj; J++)
      for (int I = 0; I <= ni + 1; I++) {
        double p = A3(R->P2, K, J, I), th = A3(R->THETA2, K, J, I);
        double t = th * pow(p / c->press0, c->kappa);
        A3(R->T2, K, J, I) = t;
        A3(R->RHO2, K, J, I) = p * mu2 / (R_GAS * t);
      }
  for (int J = 0; J <= nj; J++)
    for (int I = 0; I <= ni + 1; I++) {
      A3(R->RHO2, 0, J, I) = A3(R->RHO3, 0, J, I) * A3(R->RHO3, 0, J, I) / A3(R->RHO2, 1, J, I);
      A3(R->THETA2, 0, J, I) = c->sigmatheta[0];
      A3(R->T2, 0, J, I) = A3(R->T3, 0, J, I);
    }
  if (R->is_gas) {
    double u1d[4096], p1d[4096], r1d[4096], gz1d[4096];
    for (int J = 0; J <= nj; J++) {
      u1d[J] = p1d[J] = r1d[J] = 0.;
      for (int I = 1; I <= ni; I++) {
        u1d[J] += A3(R->U, nk, J, I);
        p1d[J] += A3(R->P3, nk, J, I);
        r1d[J] += A3(R->RHO3, nk, J, I);
      }
      u1d[J] /= ni;
      p1d[J] /= ni;
      r1d[J] /= ni;
    }
    gz_from_u(R, u1d, p1d, r1d, gz1d);
    for (int J = 0; J <= nj; J++)
      for (int I = 0; I <= ni + 1; I++) A2(R->GZS, J, I) = gz1d[J];
  }
  double cpr = c->cpr, rgas = c->rgas;
  for (int J = 0; J <= nj; J++)
    for (int I = 0; I <= ni + 1; I++) {
      A3(R->GZ3, nk, J, I) = c->topo_scale * A2(R->GZS, J, I);
      A3(R->MONT3, nk, J, I) = A3(R->GZ3, nk, J, I) + cpr * A3(R->T3, nk, J, I) * rgas;
    }
#define P(K) A3(R->P3, K, J, I)
#define RH(K) A3(R->RHO3, K, J, I)
  for (int J = 0; J <= nj; J++)
    for (int I = 0; I <= ni + 1; I++) {
      int K = nk;
      A3(R->GZ3, K - 1, J, I) = A3(R->GZ3, K, J, I) + ((P(K) + P(K - 1)) / (RH(K) + RH(K - 1))) * log(P(K) / P(K - 1));
      A3(R->MONT3, K - 1, J, I) = A3(R->GZ3, K - 1, J, I) + cpr * A3(R->T3, K - 1, J, I) * rgas;
      for (K = nk - 1; K >= 2; K--) {
        double dz10 = log(P(K + 1) / P(K)), dz21 = log(P(K) / P(K - 1)), dz32 = log(P(K - 1) / P(K - 2));
        double dz20 = dz21 + dz10, dz30 = dz32 + dz20, dz31 = dz32 + dz21;
        double df0 = P(K + 1) / RH(K + 1), df1 = P(K) / RH(K), df2 = P(K - 1) / RH(K - 1), df3 = P(K - 2) / RH(K - 2);
        double dgz = -(df0 * (dz21 * dz21 * dz21 * (dz31 + dz32)) / (dz10 * dz20 * dz30) +
                       df1 * (dz21 * (dz21 * (dz21 + 2. * (dz10 - dz31)) - 6. * dz10 * dz31)) / (dz10 * dz31) +
                       df2 * (dz21 * (dz21 * (dz21 + 2. * (dz32 - dz20)) - 6. * dz20 * dz32)) / (dz20 * dz32) +
                       df3 * (dz21 * dz21 * dz21 * (dz10 + dz20)) / (dz30 * dz31 * dz32)) / 12.;
        A3(R->GZ3, K - 1, J, I) = A3(R->GZ3, K, J, I) + dgz;
        A3(R->MONT3, K - 1, J, I) = A3(R->GZ3, K - 1, J, I) + cpr * A3(R->T3, K - 1, J, I) * rgas;
      }
      K = 1;
      A3(R->GZ3, K - 1, J, I) = A3(R->GZ3, K, J, I) + ((P(K) + P(K - 1)) / (RH(K) + RH(K - 1))) * log(P(K) / P(K - 1));
      A3(R->MONT3, K - 1, J, I) = A3(R->GZ3, K - 1, J, I) + cpr * A3(R->T3, K - 1, J, I) * rgas;
    }
#undef P
#undef RH
  for (int K = 1; K <= nk; K++)
    for (int J = 0; J <= nj; J++)
      for (int I = 1; I <= ni; I++)
        if (A3(R->GZ3, K, J, I) >= A3(R->GZ3, K - 1, J, I)) R->err |= 1;
  if (R->itime > 1)
    for (int K = 1; K <= nk; K++)
      for (int J = 0; J <= nj; J++) {
        double dtg_inv = -1. / (R->dt * c->g[2 * J + 1]);
        for (int I = 1; I <= ni; I++) {
          A3(R->DZDT3, K, J, I) -= A3(R->GZ3, K, J, I);
          A3(R->DZDT3, K, J, I) *= dtg_inv;
        }
      }
  for (int K = 1; K <= nk; K++)
    for (int J = 0; J <= nj; J++)
      for (int I = 1; I <= ni; I++)
        A3(R->DZDT3, K, J, I) += A3(R->W3, K, J, I) * A3(R->H3, K, J, I) / A3(R->RHO3, K, J, I);
  for (int K = 1; K <= nk; K++)
    for (int J = 0; J <= nj; J++) {
      int jj = 2 * J;
      double gx3 = .5 * c->m[jj + 1] / c->g[jj + 1];
      double gy2 = .5 * c->n[jj] / c->g[jj];
      double gy4 = .5 * c->n[jj + 2] / c->g[jj + 2];
      for (int I = 1; I <= ni; I++) {
        const double *g3 = R->GZ3;
        A3(R->DZDT3, K, J, I) += A3(R->V, K, J, I) * (A3(g3, K, J, I) - A3(g3, K, J - 1, I)) * gy2 +
                                 A3(R->V, K, J + 1, I) * (A3(g3, K, J + 1, I) - A3(g3, K, J, I)) * gy4 +
                                 A3(R->U, K, J, I) * (A3(g3, K, J, I) - A3(g3, K, J, I - 1)) * gx3 +
                                 A3(R->U, K, J, I + 1) * (A3(g3, K, J, I + 1) - A3(g3, K, J, I)) * gx3;
      }
    }
  bc3(R, R->DZDT3);
}

/* get_brunt2 in a layer (kk = 2K; np = 1, no smoothing),
 * src/shared/epic_funcs_diag.c:2885-3064, with return_cp :3311-3337 (the
 * enthalpy of x_h2 = 0 is cpr*T*rgas, :6652-6760), avg_molar_mass
 * :5360-5398 (nq = 0) and return_density :6353-6395 (ideal gas) */
static double brunt2_layer(cpuref *R, int K, int J, int I) {
  epic_case *c = &R->c;
  const double pressure = A3(R->P2, K, J, I), temperature = A3(R->T2, K, J, I), density = A3(R->RHO2, K, J, I);
  double cp;
  if (R->is_hs) {
    cp = 1004.;
  } else {
    double deltaT = temperature * 1.e-6, h2, h1;
    h2 = c->cpr * (temperature + deltaT);
    h2 *= c->rgas;
    h1 = c->cpr * (temperature - deltaT);
    h1 *= c->rgas;
    cp = (h2 - h1) / (2. * deltaT);
  }
  const double mu3 = 1. / (c->rgas / R_GAS), mu = .5 * (mu3 + mu3);
  const double deltap = 0.001 * pressure, deltaT = 0.001 * temperature;
  const double drho_dp_T = ((pressure + deltap) * mu / (R_GAS * temperature) -
                            (pressure - deltap) * mu / (R_GAS * temperature)) / (2. * deltap);
  const double drho_dT_p = (pressure * mu / (R_GAS * (temperature + deltaT)) -
                            pressure * mu / (R_GAS * (temperature - deltaT))) / (2. * deltaT);
  const double Drho_Dp = (A3(R->RHO3, K, J, I) - A3(R->RHO3, K - 1, J, I)) /
                         (A3(R->P3, K, J, I) - A3(R->P3, K - 1, J, I));
  const double g = c->g[2 * J + 1];
  return g * g * (Drho_Dp - drho_dp_T + temperature / (cp * density * density) * drho_dT_p * drho_dT_p);
}

/* get_richardson for a layer (kk = 2K), src/shared/epic_funcs_diag.c:3066-3156:
 * the squared shears are plain differences (the reference forms dz_inv and
 * leaves it unused), and the v shear carries the reference's sign */
static double richardson_layer(cpuref *R, int K, int J, int I) {
  double dudz2 = .5 * (A3(R->U, K - 1, J, I) + A3(R->U, K - 1, J, I + 1) - A3(R->U, K, J, I) - A3(R->U, K, J, I + 1));
  double dvdz2 = .5 * (A3(R->V, K - 1, J, I) + A3(R->V, K - 1, J + 1, I) - A3(R->V, K, J, I) + A3(R->V, K, J + 1, I));
  dudz2 *= dudz2;
  dvdz2 *= dvdz2;
  if (cpuref_fcmp(dudz2, 0.) != 0) return brunt2_layer(R, K, J, I) / (dudz2 + dvdz2);
  return 1.e+6;
}

/* store_diag, src/shared/epic_funcs_diag.c:3963-4053 */
static void store_diag(cpuref *R) {
  double *buf = R->b2[7];
  for (int K = 0; K <= R->nk; K++) {
    memcpy(buf, R->V + (size_t)K * R->N2, sizeof(double) * R->N2);
    filter2(R, buf, FK_V);
    vorticity(R, 1, R->U + (size_t)K * R->N2, buf, R->H3 + (size_t)K * R->N2, R->PV3 + (size_t)K * R->N2);
  }
  for (int K = 0; K <= R->nk; K++) {
    memcpy(buf, R->U + (size_t)K * R->N2, sizeof(double) * R->N2);
    filter2(R, buf, FK_U);
    divergence(R, buf, R->V + (size_t)K * R->N2, R->DIV + (size_t)K * R->N2);
  }
  for (int J = 0; J <= R->nj; J++)
    for (int I = 1; I <= R->ni; I++)
      for (int K = 1; K <= R->nk; K++) A3(R->RI2, K, J, I) = richardson_layer(R, K, J, I);
  bc3(R, R->RI2);
}

/* calc_heating -> newtonian_cooling with the Held-Suarez T_eq and k_T,
 * src/shared/epic_funcs_diag.c:4793-4869, src/shared/epic_sensible_heating.c:51-180,217-314,325-463.
 * Layer by layer: the relaxation into a buffer, its area-weighted layer mean
 * removed for K <= k_sponge or with newt_cool_adjust (:139-167), then added
 * to the cleared HEAT3. */
/* t_rad's table branch (src/shared/epic_sensible_heating.c:291-311): the
 * cooling time at pressure p, held at the table's ends */
static double tab_t_rad(const cpuref *R, double p) {
  double neglogp = -log(p), dx;
  if (neglogp >= R->cx[R->nc - 1]) return R->cy[R->nc - 1];
  if (neglogp <= R->cx[0]) return R->cy[0];
  int ki = table_place(R->nc, R->cx, neglogp, &dx);
  return hermite(R->cx + ki, R->cy + ki, R->cz + ki, dx, neglogp);
}
/* get_sounding(..., "temperature") (src/shared/epic_funcs_diag.c:3259-3291):
 * held at the table's bottom below it */
static double tab_t_eq(const cpuref *R, double p) {
  double x = -log(p), xd;
  if (x < R->sx[0]) return R->sy[0];
  int ki = table_place(R->ns, R->sx, x, &xd);
  return hermite(R->sx + ki, R->sy + ki, R->sz + ki, xd, x);
}
/* return_cp for a dry ideal gas (x_h2 = 0, src/shared/epic_funcs_diag.c:
 * 3311-3337, 6652-6760): the centred difference of h = cpr T rgas */
static double cp_ideal(const epic_case *c, double t) {
  double deltaT = t * 1.e-6;
  double h2 = (c->cpr * (t + deltaT)) * c->rgas, h1 = (c->cpr * (t - deltaT)) * c->rgas;
  return (h2 - h1) / (2. * deltaT);
}

static void calc_heating(cpuref *R) {
  epic_case *c = &R->c;
  memset(R->HEAT3, 0, sizeof(double) * R->N3);
  if (!c->newt_cool_on || !(R->is_hs || R->is_hj || R->is_tab)) {
    bc3(R, R->HEAT3);
    return;
  }
  const double pbk = R->is_hj ? (pow)(c->ref_pb / c->press0, c->kappa) : 0.;
  double *buf = R->b2[6];
  for (int K = 1; K <= R->nk; K++) {
    for (int J = 0; J <= R->nj; J++) {
      double lat = c->lat[2 * J + 1];
      for (int I = 1; I <= R->ni; I++) {
        double p = A3(R->P3, K, J, I), t = A3(R->T3, K, J, I);
        if (R->is_tab) {
          /* toward the sounding with the tabulated cooling time
           * (src/shared/epic_sensible_heating.c:103-130) */
          const double cp_over_time = cp_ideal(c, t) / tab_t_rad(R, p);
          A2(buf, J, I) = -cp_over_time * (t - tab_t_eq(R, p));
        } else if (R->is_hs) {
          const double ka = 1. / (40. * 60. * 60. * 24.), ks = 1. / (4. * 60. * 60. * 24.);
          const double dt_y = 60., dth_z = 10., p0 = 1000. * 100.;
          double kt = ka, pbot = A3(R->P3, R->nk, J, I);
          double amp = (p / pbot - 0.7) / (1. - 0.7);
          if (amp > 0.) {
            double cl = cos(c->lat[2 * J + 1] * DEGR);
            kt += (ks - ka) * (pow)(cl, 4.) * amp;
          }
          double t_cool = 1. / kt;
          double cpt = 1004. / t_cool;
          double c2 = cos(lat * DEGR);
          c2 *= c2;
          double s2 = 1. - c2, pp0 = p / p0;
          double teq = (315. - dt_y * s2 - dth_z * log(pp0) * c2) * pow(pp0, c->kappa);
          teq = mx(200., teq);
          A2(buf, J, I) = -cpt * (t - teq);
        } else {
          /* Day-night Newtonian relaxation of the hot-Jupiter case (no
           * reference counterpart: its temp_eq has no such branch,
           * src/shared/epic_sensible_heating.c:394-460), on the interfaces
           * K = 1..nk like the Held-Suarez relaxation: cp (T_eq - T)/tau. */
          const double cl = cos(c->lat[2 * J + 1] * DEGR);
          const double pk = pow(p / c->press0, c->kappa);
          const double tref = (1. / (1. / c->ref_theta_b + c->ref_a_over_kappa * (pk - pbk))) * pk;
          double dT = c->teq_contrast;
          if (p > c->teq_p_contrast) dT *= mx(0., log(c->pbot / p) / (log)(c->pbot / c->teq_p_contrast));
          const double day = mx(0., cl * cos(c->lon[2 * I + 1] * DEGR));
          const double teq = tref + dT * (day - .25);
          A2(buf, J, I) = c->cp * (teq - t) / c->tau_rad;
        }
      }
    }
    if (K <= R->k_sponge || c->newt_cool_adjust) {
      double area = 0., ha = 0.;
      for (int J = 0; J <= R->nj; J++) {
        double da = 1. / c->mn[2 * J + 1];
        for (int I = 1; I <= R->ni; I++) {
          area += da;
          ha += A2(buf, J, I) * da;
        }
      }
      double havg = ha / area;
      for (int J = 0; J <= R->nj; J++)
        for (int I = 1; I <= R->ni; I++) A2(buf, J, I) -= havg;
    }
    for (int J = 0; J <= R->nj; J++)
      for (int I = 1; I <= R->ni; I++) A3(R->HEAT3, K, J, I) += A2(buf, J, I);
  }
  bc3(R, R->HEAT3);
}

/* calc_w, src/shared/epic_flux.c:123-335.  The reference's top- and
 * bottom-layer SUM_DIV statements use stale loop indices (:224-226,
 * :259-261) and land in pad cells that nothing reads, so SUM_DIV(1) and
 * SUM_DIV(nk) stay zero for every column; that is what is stated here. */
static void calc_w(cpuref *R) {
  epic_case *c = &R->c;
  int nk = R->nk, ks = R->k_sigma;
  double *W = R->W3, *P = R->P3;
  for (int J = 0; J <= R->nj; J++)
    for (int I = 0; I <= R->ni + 1; I++) {
      A3(W, 0, J, I) = 0.;
      A3(W, nk, J, I) = 0.;
    }
  for (int J = 0; J <= R->nj; J++)
    for (int I = 0; I <= R->ni + 1; I++) {
      double ptop = A3(P, 0, J, I), pbot = A3(P, nk, J, I);
      for (int K = 1; K < ks - 1; K++) {
        double sigma = get_sigma(pbot, A3(P, K, J, I), ptop);
        A3(W, K, J, I) = g_sig(R, sigma) * A3(R->HEAT3, K, J, I) / A3(R->EXNER3, K, J, I);
      }
      for (int K = ks - 1; K < nk; K++) A3(W, K, J, I) = 0.;
    }
  double *S = R->sum_div;
  memset(S, 0, sizeof(double) * R->N3);
  for (int K = 2; K < nk; K++)
    for (int J = 0; J <= R->nj; J++)
      for (int I = 0; I <= R->ni + 1; I++) {
#define PP(k) A3(P, k, J, I)
#define DV(k) A3(R->DIV, k, J, I)
        double dz10 = log(PP(K - 1) / PP(K - 2)), dz21 = log(PP(K) / PP(K - 1)), dz32 = log(PP(K + 1) / PP(K));
        double dz20 = dz21 + dz10, dz30 = dz32 + dz20, dz31 = dz32 + dz21;
        double df0 = PP(K - 2) * DV(K - 2), df1 = PP(K - 1) * DV(K - 1), df2 = PP(K) * DV(K), df3 = PP(K + 1) * DV(K + 1);
        double ds = -(df0 * (dz21 * dz21 * dz21 * (dz31 + dz32)) / (dz10 * dz20 * dz30) +
                      df1 * (dz21 * (dz21 * (dz21 + 2. * (dz10 - dz31)) - 6. * dz10 * dz31)) / (dz10 * dz31) +
                      df2 * (dz21 * (dz21 * (dz21 + 2. * (dz32 - dz20)) - 6. * dz20 * dz32)) / (dz20 * dz32) +
                      df3 * (dz21 * dz21 * dz21 * (dz10 + dz20)) / (dz30 * dz31 * dz32)) / 12.;
        A3(S, K, J, I) = A3(S, K - 1, J, I) + ds;
#undef PP
#undef DV
      }
  double eps = R->eps13;
  for (int J = 0; J <= R->nj; J++)
    for (int I = 0; I <= R->ni + 1; I++) {
      double ptop = A3(P, 0, J, I), psig = A3(P, ks - 1, J, I), pbot = A3(P, nk, J, I);
      for (int K = 1; K < nk; K++) {
        double th = A3(R->THETA, K, J, I), p3 = A3(P, K, J, I), pb3 = pbot;
        double dp = p3 * eps, dpb = pb3 * eps;
        double p2 = mx(p3 - dp * .5, ptop), p4 = mnv(p3 + dp * .5, pbot);
        if (p2 < psig && p4 >= psig) p2 = p3;
        double pb2 = mx(pb3 - dpb * .5, p3), pb4 = mx(pb3 + dpb * .5, p3);
        if (get_sigma(pb4, p3, ptop) > c->sigma_sigma && get_sigma(pb2, p3, ptop) <= c->sigma_sigma) pb4 = pb3;
        double dFdp = NAN, dFdpb = NAN;
        if (p2 < p4) dFdp = (sgth_of(R, th, p2, pb3, ptop) - sgth_of(R, th, p4, pb3, ptop)) / (p2 - p4);
        else R->err |= 2;
        if (pb2 < pb4) dFdpb = (sgth_of(R, th, p3, pb2, ptop) - sgth_of(R, th, p3, pb4, ptop)) / (pb2 - pb4);
        else R->err |= 2;
        A3(W, K, J, I) -= A3(S, K, J, I) * dFdp + A3(S, nk, J, I) * dFdpb;
      }
    }
}

/* source_sink: heating of the prognostic theta, src/shared/epic_funcs_diag.c:4768-4778 */
static void source_sink(cpuref *R) {
  for (int J = 0; J <= R->nj; J++)
    for (int I = 0; I <= R->ni + 1; I++)
      for (int K = R->k_sigma - 1; K <= R->nk; K++)
        A3(R->THETA, K, J, I) += R->dt * A3(R->HEAT3, K, J, I) / A3(R->EXNER3, K, J, I);
}

/* ------------------------------------------------------------ momentum */
/* uv_core, src/shared/epic_timestep.c:628-762 with the Sadourny weights of
 * include/epic_pv_schemes.h:48-61 */
static void uv_core(cpuref *R, double *dudt, double *dvdt) {
  epic_case *c = &R->c;
  int ni = R->ni, nj = R->nj;
  double *uh = R->b2[0], *vh = R->b2[1], *kin = R->b2[2];
  for (int K = 1; K < R->nk; K++) {
    memset(uh, 0, sizeof(double) * R->N2);
    memset(vh, 0, sizeof(double) * R->N2);
    memset(kin, 0, sizeof(double) * R->N2);
    const double *u = R->U + (size_t)K * R->N2, *v = R->V + (size_t)K * R->N2;
    const double *h3 = R->H3 + (size_t)K * R->N2, *pv = R->PV3 + (size_t)K * R->N2;
    for (int J = 0; J <= nj; J++) {
      double ni_ = 1. / c->n[2 * J + 1];
      for (int I = 1; I <= ni; I++) {
        double havg = .5 * (A2(h3, J, I) + A2(h3, J, I - 1));
        A2(uh, J, I) = A2(u, J, I) * havg * ni_;
      }
    }
    bc2(R, uh);
    for (int J = 1; J <= nj; J++) {
      double mi = 1. / c->m[2 * J];
      for (int I = 1; I <= ni; I++) {
        double havg = .5 * (A2(h3, J, I) + A2(h3, J - 1, I));
        A2(vh, J, I) = A2(v, J, I) * havg * mi;
      }
    }
    bc2(R, vh);
    for (int J = 0; J <= nj; J++)
      for (int I = 1; I <= ni; I++) A2(kin, J, I) = kin_at(R, u, v, J, I);
    bc2(R, kin);
    for (int J = 1; J <= nj; J++) {
      double n2j = c->n[2 * J];
      for (int I = 1; I <= ni; I++) {
        const pvw w = pv_v(R, pv, J, I);
        A3(dvdt, K, J, I) += ((-w.ga * A2(uh, J, I + 1) - w.de * A2(uh, J, I) - w.al * A2(uh, J - 1, I) -
                               w.be * A2(uh, J - 1, I + 1) + w.e2 * A2(vh, J - 1, I) - w.e1 * A2(vh, J + 1, I)) *
                                  w.coef +
                              A2(kin, J - 1, I) - A2(kin, J, I)) *
                             n2j;
      }
    }
    filter2(R, kin, FK_H);
    for (int J = 0; J <= nj; J++) {
      double m2 = c->m[2 * J + 1];
      for (int I = 1; I <= ni; I++) {
        const pvw w = pv_u(R, pv, J, I);
        A3(dudt, K, J, I) += ((w.al * A2(vh, J + 1, I) + w.be * A2(vh, J + 1, I - 1) + w.ga * A2(vh, J, I - 1) +
                               w.de * A2(vh, J, I) + w.e2 * A2(uh, J, I - 1) - w.e1 * A2(uh, J, I + 1)) *
                                  w.coef +
                              A2(kin, J, I - 1) - A2(kin, J, I)) *
                             m2;
      }
    }
  }
  for (int K = 1; K < R->nk; K++) {
    int kk = 2 * K + 1;
    double ds = c->sigmatheta[kk - 2] - c->sigmatheta[kk + 2];
    for (int J = 0; J <= nj; J++)
      for (int I = 1; I <= ni; I++)
        A3(dudt, K, J, I) -= .5 * (A3(R->W3, K, J, I) + A3(R->W3, K, J, I - 1)) *
                             (A3(R->U, K - 1, J, I) - A3(R->U, K + 1, J, I)) / ds;
    for (int J = 1; J <= nj; J++)
      for (int I = 1; I <= ni; I++)
        A3(dvdt, K, J, I) -= .5 * (A3(R->W3, K, J, I) + A3(R->W3, K, J - 1, I)) *
                             (A3(R->V, K - 1, J, I) - A3(R->V, K + 1, J, I)) / ds;
  }
}

/* uv_drag, src/shared/epic_timestep.c:1303-1509 */
/* t_rad, src/shared/epic_sensible_heating.c:217-314: the Held-Suarez
 * cooling time from k_a, k_s; the hot Jupiter's relaxation time is its
 * constant tau_rad */
static double t_rad(cpuref *R, int K, int J, int I) {
  if (R->is_hj) return R->c.tau_rad;
  if (R->is_tab) return tab_t_rad(R, A3(R->P3, K, J, I));
  const double ka = 1. / (40. * 60. * 60. * 24.), ks = 1. / (4. * 60. * 60. * 24.);
  double p = A3(R->P3, K, J, I), kt = ka, pbot = A3(R->P3, R->nk, J, I);
  double amp = (p / pbot - 0.7) / (1. - 0.7);
  if (amp > 0.) kt += (ks - ka) * (pow)(cos(R->c.lat[2 * J + 1] * DEGR), 4.) * amp;
  return 1. / kt;
}

static void uv_drag(cpuref *R, double *dudt, double *dvdt) {
  epic_case *c = &R->c;
  int ni = R->ni, nj = R->nj, nk = R->nk;
  for (int K = 0; K <= R->k_sponge; K++) {
    for (int J = 0; J <= nj; J++)
      for (int I = 1; I <= ni; I++)
        A3(dudt, K, J, I) -= (c->t_sponge_inv[K]) * (A3(R->U, K, J, I) - A3(R->USPIN, K, J, I));
    for (int J = 1; J <= nj; J++)
      for (int I = 1; I <= ni; I++) A3(dvdt, K, J, I) -= (c->t_sponge_inv[K]) * (A3(R->V, K, J, I) - 0.);
  }
  if ((c->newt_cool_on && c->prandtl > 0.) || (c->tau_drag > 0. && c->tau_drag < 1.e+20)) {
    for (int K = R->k_sponge + 1; K < nk; K++) {
      for (int J = 0; J <= nj; J++) {
        double uavg = 0.;
        for (int I = 1; I <= ni; I++) uavg += A3(R->U, K, J, I);
        uavg /= (ni);
        for (int I = 1; I <= ni; I++) {
          double uray = (c->drag_zonal_avg) * uavg + (1 - c->drag_zonal_avg) * A3(R->U, K, J, I);
          double nu0;
          if (c->newt_cool_on) nu0 = c->prandtl / (.5 * (t_rad(R, K, J, I) + t_rad(R, K, J, I - 1)));
          else nu0 = 1. / c->tau_drag;
          A3(dudt, K, J, I) -= nu0 * (uray - A3(R->USPIN, K, J, I));
        }
      }
      if (c->drag_v)
        for (int J = 1; J <= nj; J++) {
          double vavg = 0.;
          for (int I = 1; I <= ni; I++) vavg += A3(R->V, K, J, I);
          vavg /= ni;
          for (int I = 1; I <= ni; I++) {
            double vray = (c->drag_zonal_avg) * vavg + (1 - c->drag_zonal_avg) * A3(R->V, K, J, I);
            double nu0;
            if (c->newt_cool_on) nu0 = c->prandtl / (.5 * (t_rad(R, K, J, I) + t_rad(R, K, J - 1, I)));
            else nu0 = 1. / c->tau_drag;
            A3(dvdt, K, J, I) -= nu0 * vray;
          }
        }
    }
  }
  if (R->is_hs) {
    const double nu00 = (1. / (24. * 60. * 60.));
    for (int J = 0; J <= nj; J++)
      for (int I = 1; I <= ni; I++) {
        double pb = 0.5 * (A3(R->P3, nk, J, I) + A3(R->P3, nk, J, I - 1));
        for (int K = nk - 1; K >= 1; K--) {
          double p3 = .5 * (A3(R->P3, K, J, I) + A3(R->P3, K, J, I - 1));
          double amp = (p3 / pb - 0.7) / (1. - 0.7);
          if (amp < 0.) break;
          A3(dudt, K, J, I) -= nu00 * amp * A3(R->U, K, J, I);
        }
      }
    for (int J = 1; J <= nj; J++)
      for (int I = 1; I <= ni; I++) {
        double pb = 0.5 * (A3(R->P3, nk, J, I) + A3(R->P3, nk, J - 1, I));
        for (int K = nk - 1; K >= 1; K--) {
          double p3 = .5 * (A3(R->P3, K, J, I) + A3(R->P3, K, J - 1, I));
          double amp = (p3 / pb - 0.7) / (1. - 0.7);
          if (amp < 0.) break;
          A3(dvdt, K, J, I) -= nu00 * amp * A3(R->V, K, J, I);
        }
      }
  }
}

/* uv_pgrad_traditional, src/shared/epic_timestep.c:806-884 (the
 * reference's PGRAD_TRADITIONAL build; no fpara): the Montgomery-potential
 * form, grad(MONT3) - <EXNER3> grad(THETA), on K = 1..k_sigma-2, and on
 * K = k_sigma-1..nk-1 the h3-weighted GZ3 difference less the p-weighted
 * mean of the GZ3 differences of the interfaces above and below; u on
 * J = 0..nj, v on J = 1..nj, I = 1..ni (no pads) */
static void uv_pgrad_traditional(cpuref *R, double *dudt, double *dvdt) {
  epic_case *c = &R->c;
  const int ni = R->ni, nj = R->nj, nk = R->nk, ks = R->k_sigma;
  double *MO = R->MONT3, *EX = R->EXNER3, *TH = R->THETA, *G = R->GZ3, *H = R->H3, *P = R->P3;
  for (int K = 1; K <= ks - 2; K++) {
    for (int J = 0; J <= nj; J++) {
      const double m = c->m[2 * J + 1];
      for (int I = 1; I <= ni; I++)
        A3(dudt, K, J, I) -= m * (A3(MO, K, J, I) - A3(MO, K, J, I - 1) -
                                  .5 * (A3(EX, K, J, I) + A3(EX, K, J, I - 1)) * (A3(TH, K, J, I) - A3(TH, K, J, I - 1)));
    }
    for (int J = 1; J <= nj; J++) {
      const double n = c->n[2 * J];
      for (int I = 1; I <= ni; I++)
        A3(dvdt, K, J, I) -= n * (A3(MO, K, J, I) - A3(MO, K, J - 1, I) -
                                  .5 * (A3(EX, K, J, I) + A3(EX, K, J - 1, I)) * (A3(TH, K, J, I) - A3(TH, K, J - 1, I)));
    }
  }
  for (int K = ks - 1; K < nk; K++) {
    for (int J = 0; J <= nj; J++) {
      const double m = c->m[2 * J + 1];
      for (int I = 1; I <= ni; I++) {
        const double hw = (A3(G, K, J, I) * A3(H, K, J, I) - A3(G, K, J, I - 1) * A3(H, K, J, I - 1)) * 2. /
                          (A3(H, K, J, I) + A3(H, K, J, I - 1));
        const double up = (A3(G, K + 1, J, I) + A3(G, K + 1, J, I - 1)) * (A3(P, K + 1, J, I) - A3(P, K + 1, J, I - 1)) -
                          (A3(G, K - 1, J, I) + A3(G, K - 1, J, I - 1)) * (A3(P, K - 1, J, I) - A3(P, K - 1, J, I - 1));
        const double dn = A3(P, K + 1, J, I) + A3(P, K + 1, J, I - 1) - A3(P, K - 1, J, I) - A3(P, K - 1, J, I - 1);
        A3(dudt, K, J, I) -= m * (hw - up / dn);
      }
    }
    for (int J = 1; J <= nj; J++) {
      const double n = c->n[2 * J];
      for (int I = 1; I <= ni; I++) {
        const double hs = (A3(G, K, J, I) * A3(H, K, J, I) - A3(G, K, J - 1, I) * A3(H, K, J - 1, I)) * 2. /
                          (A3(H, K, J, I) + A3(H, K, J - 1, I));
        const double up = (A3(G, K + 1, J, I) + A3(G, K + 1, J - 1, I)) * (A3(P, K + 1, J, I) - A3(P, K + 1, J - 1, I)) -
                          (A3(G, K - 1, J, I) + A3(G, K - 1, J - 1, I)) * (A3(P, K - 1, J, I) - A3(P, K - 1, J - 1, I));
        const double dn = A3(P, K + 1, J, I) + A3(P, K + 1, J - 1, I) - A3(P, K - 1, J, I) - A3(P, K - 1, J - 1, I);
        A3(dvdt, K, J, I) -= n * (hs - up / dn);
      }
    }
  }
}

/* the traditional pressure gradient alone on the current state, into a
 * cleared plane IT_ZERO of the tendencies (the function-level check against
 * the reference's own uv_pgrad_traditional) */
void cpuref_pgrad_traditional(cpuref *R) {
  memset(R->DUDT[R->it0], 0, sizeof(double) * R->N3);
  memset(R->DVDT[R->it0], 0, sizeof(double) * R->N3);
  uv_pgrad_traditional(R, R->DUDT[R->it0], R->DVDT[R->it0]);
}

/* uv_pgrad_finite_volume, src/shared/epic_timestep.c:909-1273 */
static void uv_pgrad(cpuref *R, double *dudt, double *dvdt) {
  epic_case *c = &R->c;
  int ni = R->ni, nj = R->nj, nk = R->nk;
  double *G = R->GZ3, *P2 = R->P2;
  for (int K = 0; K <= nk; K++) {
    double t = 0.;
    for (int I = 1; I <= ni; I++) t += A3(G, K, 0, I);
    t /= ni;
    for (int I = 0; I <= ni + 1; I++) A3(G, K, -1, I) = t;
  }
  for (int K = 0; K <= nk; K++) {
    double t = 0.;
    for (int I = 1; I <= ni; I++) t += A3(G, K, nj, I);
    t /= ni;
    for (int I = 0; I <= ni + 1; I++) A3(G, K, nj + 1, I) = t;
  }
#define Z(k, j, i) A3(G, k, j, i)
#define PZ(k, j, i) A3(P2, k, j, i)
  const double TOL = 1.e-2;
  for (int K = 1; K < nk; K++)
    for (int J = 0; J <= nj; J++) {
      double rlns = c->rln[2 * J], rlnn = c->rln[2 * J + 2];
      int I = 0;
      double lne = c->lon[2 * I + 1] * DEGR;
      double zbse = .25 * (Z(K, J, I) + Z(K, J - 1, I) + Z(K + 1, J, I) + Z(K + 1, J - 1, I));
      double zbne = .25 * (Z(K, J, I) + Z(K, J + 1, I) + Z(K + 1, J, I) + Z(K + 1, J + 1, I));
      double ztse = .25 * (Z(K, J, I) + Z(K, J - 1, I) + Z(K - 1, J, I) + Z(K - 1, J - 1, I));
      double ztne = .25 * (Z(K, J, I) + Z(K, J + 1, I) + Z(K - 1, J, I) + Z(K - 1, J + 1, I));
      double zbe = .5 * (Z(K, J, I) + Z(K + 1, J, I)), zte = .5 * (Z(K, J, I) + Z(K - 1, J, I));
      double pbe = PZ(K + 1, J, I), pte = PZ(K, J, I);
      double a1 = ztne + ztse, b1 = ztne * ztse, a2 = zbne + zbse, b2 = zbne * zbse;
      double t1 = a1 * a1 - b1 - (a2 * a2 - b2), t2 = 3. * (a1 - a2);
      double east = (-t1 * (pbe - pte) + t2 * (zte * pbe - zbe * pte)) / (zte - zbe);
      for (I = 1; I <= ni; I++) {
        double rho = .5 * (A3(R->RHO3, K, J, I) + A3(R->RHO3, K, J, I - 1));
        double lnw = lne, zbsw = zbse, zbnw = zbne, zbw = zbe, ztsw = ztse, ztnw = ztne, ztw = zte;
        double pbw = pbe, ptw = pte, west = east;
        lne = c->lon[2 * I + 1] * DEGR;
        zbse = .25 * (Z(K, J, I) + Z(K, J - 1, I) + Z(K + 1, J, I) + Z(K + 1, J - 1, I));
        zbne = .25 * (Z(K, J, I) + Z(K, J + 1, I) + Z(K + 1, J, I) + Z(K + 1, J + 1, I));
        ztse = .25 * (Z(K, J, I) + Z(K, J - 1, I) + Z(K - 1, J, I) + Z(K - 1, J - 1, I));
        ztne = .25 * (Z(K, J, I) + Z(K, J + 1, I) + Z(K - 1, J, I) + Z(K - 1, J + 1, I));
        zbe = .5 * (Z(K, J, I) + Z(K + 1, J, I));
        zte = .5 * (Z(K, J, I) + Z(K - 1, J, I));
        pbe = PZ(K + 1, J, I);
        pte = PZ(K, J, I);
        a1 = ztne + ztse; b1 = ztne * ztse; a2 = zbne + zbse; b2 = zbne * zbse;
        t1 = a1 * a1 - b1 - (a2 * a2 - b2);
        t2 = 3. * (a1 - a2);
        east = (-t1 * (pbe - pte) + t2 * (zte * pbe - zbe * pte)) / (zte - zbe);
        double top, bottom;
        a1 = ztne + ztse; a2 = ztnw + ztsw;
        t2 = 3. * (a1 - a2);
        if (fabs(2. * (zte - ztw) / (zte + ztw - zbe - zbw)) > TOL) {
          b1 = ztne * ztse; b2 = ztnw * ztsw;
          t1 = a1 * a1 - b1 - (a2 * a2 - b2);
          top = (-t1 * (ptw - pte) + t2 * (zte * ptw - ztw * pte)) / (zte - ztw);
        } else {
          top = .5 * (ptw + pte) * t2;
        }
        a1 = zbne + zbse; a2 = zbnw + zbsw;
        t2 = 3. * (a1 - a2);
        if (fabs(2. * (zbe - zbw) / (zte + ztw - zbe - zbw)) > TOL) {
          b1 = zbne * zbse; b2 = zbnw * zbsw;
          t1 = a1 * a1 - b1 - (a2 * a2 - b2);
          bottom = (-t1 * (pbw - pbe) + t2 * (zbe * pbw - zbw * pbe)) / (zbe - zbw);
        } else {
          bottom = .5 * (pbw + pbe) * t2;
        }
        double north = 0., south = 0.;
        double pgf = top - bottom + south - north + west - east;
        double dzn = ztne - zbne + ztnw - zbnw, dzs = ztse - zbse + ztsw - zbsw;
        pgf /= (lne - lnw) * rho * (rlns * (dzs + .5 * dzn) + rlnn * (dzn + .5 * dzs));
        A3(dudt, K, J, I) += pgf;
      }
    }
  for (int K = 1; K < nk; K++)
    for (int I = 1; I <= ni; I++) {
      int J = 0;
      double rlnn = c->rln[2 * J + 1], ltn = c->lat[2 * J + 1] * DEGR;
      double zbnw = .25 * (Z(K, J, I) + Z(K, J, I - 1) + Z(K + 1, J, I) + Z(K + 1, J, I - 1));
      double zbne = .25 * (Z(K, J, I) + Z(K, J, I + 1) + Z(K + 1, J, I) + Z(K + 1, J, I + 1));
      double ztnw = .25 * (Z(K, J, I) + Z(K, J, I - 1) + Z(K - 1, J, I) + Z(K - 1, J, I - 1));
      double ztne = .25 * (Z(K, J, I) + Z(K, J, I + 1) + Z(K - 1, J, I) + Z(K - 1, J, I + 1));
      double zbn = .5 * (Z(K, J, I) + Z(K + 1, J, I)), ztn = .5 * (Z(K, J, I) + Z(K - 1, J, I));
      double pbn = PZ(K + 1, J, I), ptn = PZ(K, J, I);
      double a1 = ztne + ztnw, b1 = ztne * ztnw, a2 = zbne + zbnw, b2 = zbne * zbnw, b3;
      double t1 = a1 * a1 - b1 - (a2 * a2 - b2), t2 = 3. * (a1 - a2);
      double north = rlnn * (-t1 * (pbn - ptn) + t2 * (ztn * pbn - zbn * ptn)) / (ztn - zbn);
      for (J = 1; J <= nj; J++) {
        double rho = .5 * (A3(R->RHO3, K, J, I) + A3(R->RHO3, K, J - 1, I));
        double rlt = c->rlt[2 * J];
        double rlns = rlnn, lts = ltn;
        double zbsw = zbnw, zbse = zbne, ztsw = ztnw, ztse = ztne;
        double south = north;
        double zbs = zbn, zts = ztn, pbs = pbn, pts = ptn;
        rlnn = c->rln[2 * J + 1];
        ltn = c->lat[2 * J + 1] * DEGR;
        zbnw = .25 * (Z(K, J, I) + Z(K, J, I - 1) + Z(K + 1, J, I) + Z(K + 1, J, I - 1));
        zbne = .25 * (Z(K, J, I) + Z(K, J, I + 1) + Z(K + 1, J, I) + Z(K + 1, J, I + 1));
        ztnw = .25 * (Z(K, J, I) + Z(K, J, I - 1) + Z(K - 1, J, I) + Z(K - 1, J, I - 1));
        ztne = .25 * (Z(K, J, I) + Z(K, J, I + 1) + Z(K - 1, J, I) + Z(K - 1, J, I + 1));
        zbn = .5 * (Z(K, J, I) + Z(K + 1, J, I));
        ztn = .5 * (Z(K, J, I) + Z(K - 1, J, I));
        pbn = PZ(K + 1, J, I);
        ptn = PZ(K, J, I);
        a1 = ztne + ztnw; b1 = ztne * ztnw; a2 = zbne + zbnw; b2 = zbne * zbnw;
        t1 = a1 * a1 - b1 - (a2 * a2 - b2);
        t2 = 3. * (a1 - a2);
        north = rlnn * (-t1 * (pbn - ptn) + t2 * (ztn * pbn - zbn * ptn)) / (ztn - zbn);
        a1 = ztse + ztsw; b1 = ztse * ztsw; a2 = zbse + zbsw; b2 = zbse * zbsw;
        t1 = a1 * a1 - b1 - (a2 * a2 - b2);
        t2 = 3. * (a1 - a2);
        south = rlns * (-t1 * (pbs - pts) + t2 * (zts * pbs - zbs * pts)) / (zts - zbs);
        double zbe = .125 * (Z(K, J, I) + Z(K, J, I + 1) + Z(K, J - 1, I) + Z(K, J - 1, I + 1) + Z(K + 1, J, I) +
                             Z(K + 1, J, I + 1) + Z(K + 1, J - 1, I) + Z(K + 1, J - 1, I + 1));
        double zte = .125 * (Z(K - 1, J, I) + Z(K - 1, J, I + 1) + Z(K - 1, J - 1, I) + Z(K - 1, J - 1, I + 1) +
                             Z(K, J, I) + Z(K, J, I + 1) + Z(K, J - 1, I) + Z(K, J - 1, I + 1));
        double pbe = .25 * (PZ(K + 1, J, I) + PZ(K + 1, J, I + 1) + PZ(K + 1, J - 1, I) + PZ(K + 1, J - 1, I + 1));
        double pte = .25 * (PZ(K, J, I) + PZ(K, J, I + 1) + PZ(K, J - 1, I) + PZ(K, J - 1, I + 1));
        a1 = ztne + ztse; b1 = ztne * ztse; a2 = zbne + zbse; b2 = zbne * zbse;
        t1 = a1 * a1 - b1 - (a2 * a2 - b2);
        t2 = 3. * (a1 - a2);
        double east = .5 * (rlns - rlnn) * (-t1 * (pbe - pte) + t2 * (zte * pbe - zbe * pte)) / (zte - zbe);
        double zbw = .125 * (Z(K, J, I) + Z(K, J, I - 1) + Z(K, J - 1, I) + Z(K, J - 1, I - 1) + Z(K + 1, J, I) +
                             Z(K + 1, J, I - 1) + Z(K + 1, J - 1, I) + Z(K + 1, J - 1, I - 1));
        double ztw = .125 * (Z(K - 1, J, I) + Z(K - 1, J, I - 1) + Z(K - 1, J - 1, I) + Z(K - 1, J - 1, I - 1) +
                             Z(K, J, I) + Z(K, J, I - 1) + Z(K, J - 1, I) + Z(K, J - 1, I - 1));
        double pbw = .25 * (PZ(K + 1, J, I) + PZ(K + 1, J, I - 1) + PZ(K + 1, J - 1, I) + PZ(K + 1, J - 1, I - 1));
        double ptw = .25 * (PZ(K, J, I) + PZ(K, J, I - 1) + PZ(K, J - 1, I) + PZ(K, J - 1, I - 1));
        a1 = ztnw + ztsw; b1 = ztnw * ztsw; a2 = zbnw + zbsw; b2 = zbnw * zbsw;
        t1 = a1 * a1 - b1 - (a2 * a2 - b2);
        t2 = 3. * (a1 - a2);
        double west = .5 * (rlns - rlnn) * (-t1 * (pbw - ptw) + t2 * (ztw * pbw - zbw * ptw)) / (ztw - zbw);
        double top, bottom;
        a1 = ztne + ztnw; a2 = ztse + ztsw;
        t2 = 3. * (a1 - a2);
        if (fabs(2. * (ztn - zts) / (ztn + zts - zbn - zbs)) > TOL) {
          b1 = ztne * ztnw; b2 = ztse * ztsw; b3 = ztne * ztse + ztnw * ztsw;
          t1 = a1 * a1 - b1 - (a2 * a2 - b2);
          top = .5 * ((ptn - pts) * ((rlnn + rlns) * t1 + (rlnn - rlns) * (b1 + b2 - b3)) +
                      (rlnn + rlns) * t2 * (ztn * pts - zts * ptn)) / (ztn - zts);
        } else {
          top = .25 * (rlnn + rlns) * (pts + ptn) * t2;
        }
        a1 = zbne + zbnw; a2 = zbse + zbsw;
        t2 = 3. * (a1 - a2);
        if (fabs(2. * (zbn - zbs) / (ztn + zts - zbn - zbs)) > TOL) {
          b1 = zbne * zbnw; b2 = zbse * zbsw; b3 = zbne * zbse + zbnw * zbsw;
          t1 = a1 * a1 - b1 - (a2 * a2 - b2);
          bottom = .5 * ((pbn - pbs) * ((rlnn + rlns) * t1 + (rlnn - rlns) * (b1 + b2 - b3)) +
                         (rlnn + rlns) * t2 * (zbn * pbs - zbs * pbn)) / (zbn - zbs);
        } else {
          bottom = .25 * (rlnn + rlns) * (pbs + pbn) * t2;
        }
        double pgf = top - bottom + south - (north + west + east);
        double dzn = ztne - zbne + ztnw - zbnw, dzs = ztse - zbse + ztsw - zbsw;
        pgf /= (ltn - lts) * rho * rlt * (rlns * (dzs + .5 * dzn) + rlnn * (dzn + .5 * dzs));
        A3(dvdt, K, J, I) += pgf;
      }
    }
#undef Z
#undef PZ
}

/* divergence_damping, src/shared/epic_subgrid.c:1792-1885 (the taper tests
 * lat > LAT0 without fabs, so only northern rows use m0, :1860) */
static void divergence_damping(cpuref *R, double *dudt, double *dvdt) {
  epic_case *c = &R->c;
  double *buf = R->b2[7], *div = R->b2[6];
  for (int K = 1; K < R->nk; K++) {
    memcpy(buf, R->U + (size_t)K * R->N2, sizeof(double) * R->N2);
    filter2(R, buf, FK_U);
    divergence(R, buf, R->V + (size_t)K * R->N2, div);
    for (int J = 0; J <= R->nj; J++) {
      double coef = (c->lat[2 * J + 1] > LAT0) ? R->nudiv * R->m0_div : R->nudiv * c->m[2 * J + 1];
      for (int I = 1; I <= R->ni; I++) A3(dudt, K, J, I) += coef * (A2(div, J, I) - A2(div, J, I - 1));
    }
    for (int J = 1; J <= R->nj; J++) {
      double coef = R->nudiv * c->n[2 * J];
      for (int I = 1; I <= R->ni; I++) A3(dvdt, K, J, I) += coef * (A2(div, J, I) - A2(div, J - 1, I));
    }
  }
}

/* ------------------------------------------------------ hyperviscosity */
static double nu_tail(const cpuref *R, int from) {
  double t = 0.;
  for (int q = from; q <= EPIC_MAX_NU_ORDER; q += 2) t += R->c.nu[q];
  return t;
}

/* d2dy2: 3-point Legendre-fit meridional Laplacian, src/shared/epic_subgrid.c:1023-1232 */
static void d2dy2(cpuref *R, int vgrid, const double *hh, double dc, double *out) {
  int nj = R->nj;
  for (int J = (vgrid ? 0 : 0); J <= (vgrid ? nj + 1 : nj); J++) {
    const double *cx = vgrid ? R->vcx : R->hcx, *cy = vgrid ? R->vcy : R->hcy, *cz = vgrid ? R->vcz : R->hcz;
    int j0, j1, j2;
    if (J == 0) { j0 = 0; j1 = 1; j2 = 2; }
    else if (J == (vgrid ? nj + 1 : nj)) { j0 = J - 2; j1 = J - 1; j2 = J; }
    else { j0 = J - 1; j1 = J; j2 = J + 1; }
    for (int I = 1; I <= R->ni; I++) {
      double a0 = dc * cx[J], a1 = dc * cy[J], a2 = dc * cz[J];
      A2(out, J, I) = a0 * A2(hh, j0, I) + a1 * A2(hh, j1, I) + a2 * A2(hh, j2, I);
    }
  }
  bc2(R, out);
}

static void d2dy2_init(cpuref *R) {
  epic_case *c = &R->c;
  int nj = R->nj;
  for (int pass = 0; pass < 2; pass++) {
    int jmax = pass ? nj + 1 : nj;
    double *cx = pass ? R->vcx : R->hcx, *cy = pass ? R->vcy : R->hcy, *cz = pass ? R->vcz : R->hcz;
    for (int J = 0; J <= jmax; J++) {
      int jj = pass ? 2 * J : 2 * J + 1;
      double r2 = (double)c->rlt[jj];
      r2 = 1. / (r2 * r2);
      double l0, l1, l2;
      int kind;
      if (J == 0) { l0 = c->lat[jj] * DEGR; l1 = c->lat[jj + 2] * DEGR; l2 = c->lat[jj + 4] * DEGR; kind = 0; }
      else if (J == jmax) { l0 = c->lat[jj - 4] * DEGR; l1 = c->lat[jj - 2] * DEGR; l2 = c->lat[jj] * DEGR; kind = 2; }
      else { l0 = c->lat[jj - 2] * DEGR; l1 = c->lat[jj] * DEGR; l2 = c->lat[jj + 2] * DEGR; kind = 1; }
      double s0 = sin(l0), s1 = sin(l1), s2 = sin(l2);
      if (kind == 0) {
        cx[J] = r2 * (1. + s0 * (-3. * s0 + s1 + s2)) / ((s1 - s0) * .5 * (s2 - s0));
        cy[J] = -r2 * 2. * (cos(2. * l0) + s0 * s2) / ((s1 - s0) * (s2 - s1));
        cz[J] = r2 * (cos(2. * l0) + s0 * s1) / (.5 * (s2 - s0) * (s2 - s1));
      } else if (kind == 1) {
        cx[J] = r2 * (cos(2. * l1) + s1 * s2) / ((s1 - s0) * .5 * (s2 - s0));
        cy[J] = -r2 * 2. * (1. + s1 * (s0 - 3. * s1 + s2)) / ((s1 - s0) * (s2 - s1));
        cz[J] = r2 * (cos(2. * l1) + s0 * s1) / (.5 * (s2 - s0) * (s2 - s1));
      } else {
        cx[J] = r2 * (cos(2. * l2) + s1 * s2) / ((s1 - s0) * .5 * (s2 - s0));
        cy[J] = -r2 * 2. * (cos(2. * l2) + s0 * s2) / ((s1 - s0) * (s2 - s1));
        cz[J] = r2 * (1. + s2 * (s0 + s1 - 3. * s2)) / (.5 * (s2 - s0) * (s2 - s1));
      }
    }
  }
}

/* meridional_hyperviscosity, src/shared/epic_subgrid.c:829-1006 */
static void merid_hv(cpuref *R, int kind, double *field) {
  if (nu_tail(R, 4) == 0.) return;
  int vg = (kind == FK_V);
  int ks = (kind == FK_THETA) ? R->k_sigma - 1 : 1;
  int ke = (kind == FK_U || kind == FK_V) ? R->nk - 1 : R->nk;
  double *hh0 = R->b2[0], *dd0 = R->b2[1];
  for (int K = ks; K <= ke; K++) {
    double *a = field + (size_t)K * R->N2, *hh = hh0, *dd = dd0;
    for (int J = 0; J <= R->nj + 1; J++)
      for (int I = 0; I <= R->ni + 1; I++) A2(hh, J, I) = A2(a, J, I);
    int sign = 1, it = 2;
    double tmp = 1.;
    while (tmp > 0.) {
      d2dy2(R, vg, hh, R->coef[it], dd);
      if (R->c.nu[it] > 0.) {
        filter2(R, dd, vg ? FK_V : FK_H);
        for (int J = vg ? 1 : 0; J <= R->nj; J++)
          for (int I = 1; I <= R->ni; I++) A2(a, J, I) += R->dt * (double)sign * A2(dd, J, I);
        bc2(R, a);
      }
      sign *= -1;
      it += 2;
      double *t = hh; hh = dd; dd = t;
      tmp = nu_tail(R, it);
    }
  }
}

/* zonal_hyperviscosity, src/shared/epic_subgrid.c:590-818 */
static void zonal_hv(cpuref *R, int kind, double *field) {
  if (nu_tail(R, 4) == 0.) return;
  int jbot = (kind == FK_V) ? 1 : 0, jh = (kind == FK_V) ? 0 : 1;
  int ks = (kind == FK_THETA) ? R->k_sigma - 1 : 1;
  int ke = (kind == FK_U || kind == FK_V) ? R->nk - 1 : R->nk;
  int ni = R->ni;
  double *hh0 = R->b2[0], *dd0 = R->b2[1];
  for (int K = ks; K <= ke; K++) {
    double *a = field + (size_t)K * R->N2, *hh = hh0, *dd = dd0;
    for (int J = 0; J <= R->nj + 1; J++)
      for (int I = 0; I <= ni + 1; I++) A2(hh, J, I) = A2(a, J, I);
    int sign = 1, it = 2;
    double tmp = 1.;
    while (tmp > 0.) {
      for (int J = jbot; J <= R->nj; J++) {
        double m = R->c.m[2 * J + jh], dc;
        if (fabs(R->c.lat[2 * J + jh]) > LAT0) dc = R->coef[it] * R->m0_zonal * R->m0_zonal;
        else dc = R->coef[it] * m * m;
        for (int I = 1; I <= ni; I++)
          A2(dd, J, I) = dc * (A2(hh, J, I - 1) - 2. * A2(hh, J, I) + A2(hh, J, I + 1));
        A2(dd, J, ni + 1) = A2(dd, J, 1);
        A2(dd, J, 0) = A2(dd, J, ni);
        double fm = 0.;
        for (int I = 1; I <= ni; I++) fm += A2(dd, J, I);
        fm /= ni;
        for (int I = 0; I <= ni + 1; I++) A2(dd, J, I) -= fm;
        if (R->c.nu[it] > 0.)
          for (int I = 1; I <= ni; I++) A2(a, J, I) += R->dt * (double)sign * A2(dd, J, I);
      }
      if (R->c.nu[it] > 0.) bc2(R, a);
      sign *= -1;
      it += 2;
      double *t = hh; hh = dd; dd = t;
      tmp = nu_tail(R, it);
    }
  }
}

/* restore_mass, src/shared/epic_adjust.c:45-96 */
static void restore_mass(cpuref *R) {
  for (int K = 1; K <= R->nk; K++)
    for (int J = 0; J <= R->nj; J++)
      for (int I = 0; I <= R->ni + 1; I++)
        A3(R->HDRY, K, J, I) = mx(R->c.h_min[K], A3(R->HDRY, K, J, I));
}

/* ------------------------------------------------------------ advection */
enum { AV_P3, AV_THETA, AV_HDRY };

/* hsu_advection, src/shared/epic_flux.c:652-1488: vertical column pass,
 * then meridional and zonal passes per layer.  The column work arrays are
 * the reference's statics: an interface variable starting above K = 1
 * (theta) reads wm[k_first-1] as left by the previous call's last column. */
static void hsu(cpuref *R, int which, double *b) {
  epic_case *c = &R->c;
  int ni = R->ni, nj = R->nj, nk = R->nk;
  const double ep = 1.e-10, DT = R->dt;
  const int layer = (which == AV_HDRY);
  int kf = (which == AV_THETA) ? R->k_sigma - 1 : 1;
  double *wm = R->w_wm, *wp = R->w_wp, *wa = R->w_a, *pr = R->w_pred, *amin = R->w_min;
  double *ff = R->w_ff, *d1 = R->w_d1, *d2 = R->w_d2, *aaa = R->w_aaa;
  const double *dsi = c->dsgth_inv;
  if (layer) for (int K = kf; K <= nk; K++) amin[K] = c->h_min[K];
  else for (int K = kf - 1; K <= nk; K++) amin[K] = 0.;
  /* ---- vertical ---- */
  for (int J = 0; J <= nj; J++)
    for (int I = 0; I <= ni + 1; I++) {
      for (int K = layer ? kf : kf - 1; K <= nk; K++) wa[K] = A3(b, K, J, I);
      if (layer) {
        for (int K = kf - 1; K <= nk; K++) {
          double w = A3(R->W3, K, J, I);
          wm[K] = mnv(0., w);
          wp[K] = mx(0., w);
        }
      } else {
        for (int K = kf; K <= nk; K++) {
          double w = (K == 1 || K == nk) ? 0. : .5 * (A3(R->W3, K, J, I) + A3(R->W3, K - 1, J, I));
          wm[K] = mnv(0., w);
          wp[K] = mx(0., w);
        }
      }
      if (layer) {
        ff[kf - 1] = 0.;
        ff[nk] = 0.;
        for (int K = kf; K < nk; K++) ff[K] = wp[K] * wa[K + 1] + wm[K] * wa[K];
        for (int K = kf; K <= nk; K++) pr[K] = wa[K] + (ff[K] - ff[K - 1]) * dsi[2 * K] * DT;
        for (int K = kf; K < nk; K++) ff[K] = .5 * (wp[K] * (pr[K] + wa[K + 1]) + wm[K] * (pr[K + 1] + wa[K]));
        for (int K = kf; K <= nk; K++) {
          double t;
          if (K == 1) { d1[K] = pr[K] - wa[K + 1]; d2[K] = 0.; t = fabs(wa[K + 1] - wa[K]) + ep; }
          else if (K == nk) { d1[K] = 0.; d2[K] = pr[K] - wa[K - 1]; t = fabs(-wa[K] + wa[K - 1]) + ep; }
          else { d1[K] = pr[K] - wa[K + 1]; d2[K] = pr[K] - wa[K - 1]; t = fabs(wa[K + 1] - 2. * wa[K] + wa[K - 1]) + ep; }
          aaa[K] = t * t;
        }
        for (int K = kf; K < nk; K++) {
          double whp = sqrt(wp[K] * wp[K + 1]), whm = -sqrt(wm[K] * wm[K - 1]);
          double mu = (wp[K] - wm[K]) * dsi[2 * K + 1] * DT, al = (1. + mu) / 6.;
          double gp = aaa[K + 1] / (aaa[K + 1] + (wa[K + 1] - amin[K + 1]) * (wa[K] - amin[K]));
          gp = gp * gp;
          double gm = aaa[K] / (aaa[K] + (wa[K + 1] - amin[K + 1]) * (wa[K] - amin[K]));
          gm = gm * gm;
          double bp = 1. + (1. / (2. * al) - 1.) * gp, bm = 1. + (1. / (2. * al) - 1.) * gm;
          double bhp = 1. - gp, bhm = 1. - gm;
          ff[K] += -al * (wp[K] * bp * d1[K] - whp * bhp * d1[K + 1] + wm[K] * bm * d2[K + 1] - whm * bhm * d2[K]);
        }
        for (int K = kf; K <= nk; K++) A3(b, K, J, I) += (ff[K] - ff[K - 1]) * dsi[2 * K] * DT;
      } else {
        for (int K = kf; K <= nk; K++) ff[K] = wp[K] * wa[K] + wm[K] * wa[K - 1];
        ff[nk + 1] = 0.;
        pr[kf - 1] = wa[kf - 1];
        for (int K = kf; K <= nk; K++) pr[K] = wa[K] + (ff[K + 1] - ff[K]) * dsi[2 * K + 1] * DT;
        for (int K = kf; K <= nk; K++) ff[K] = .5 * (wp[K] * (pr[K - 1] + wa[K]) + wm[K] * (pr[K] + wa[K - 1]));
        for (int K = kf - 1; K <= nk; K++) {
          double t;
          if (K == 0) t = fabs(wa[K + 1] - wa[K]) + ep;
          else if (K == nk) t = fabs(-wa[K] + wa[K - 1]) + ep;
          else t = fabs(wa[K + 1] - 2. * wa[K] + wa[K - 1]) + ep;
          aaa[K] = t * t;
        }
        d2[kf - 1] = 0.;
        for (int K = kf; K <= nk; K++) {
          d1[K] = pr[K - 1] - wa[K];
          d2[K] = pr[K] - wa[K - 1];
        }
        for (int K = kf; K <= nk; K++) {
          double whm = (K == 1) ? 0. : -sqrt(wm[K] * wm[K - 1]);
          double whp = (K == nk) ? 0. : sqrt(wp[K] * wp[K + 1]);
          double mu = (wp[K] - wm[K]) * dsi[2 * K] * DT, al = (1. + mu) / 6.;
          double gp = aaa[K] / (aaa[K] + (wa[K] - amin[K]) * (wa[K - 1] - amin[K - 1]));
          gp = gp * gp;
          double gm = aaa[K - 1] / (aaa[K - 1] + (wa[K] - amin[K]) * (wa[K - 1] - amin[K - 1]));
          gm = gm * gm;
          double bp = 1. + (1. / (2. * al) - 1.) * gp, bm = 1. + (1. / (2. * al) - 1.) * gm;
          double bhp = 1. - gp, bhm = 1. - gm, g;
          if (K < nk) g = -al * (wp[K] * bp * d1[K] - whp * bhp * d1[K + 1] + wm[K] * bm * d2[K] - whm * bhm * d2[K - 1]);
          else g = -al * (wp[K] * bp * d1[K] + wm[K] * bm * d2[K] - whm * bhm * d2[K - 1]);
          ff[K] += g;
        }
        ff[nk + 1] = 0.;
        for (int K = kf; K <= nk; K++) A3(b, K, J, I) += (ff[K + 1] - ff[K]) * dsi[2 * K + 1] * DT;
      }
    }
  /* ---- meridional, per layer or interface K ---- */
  double *vm = R->b2[0], *vp = R->b2[1], *v2 = R->b2[2], *df1 = R->b2[2], *df2 = R->b2[3];
  double *fl = R->b2[4], *ap = R->b2[5], *aq = R->b2[5];
  for (int K = 1; K <= nk; K++) {
    if (which == AV_THETA && K <= R->k_sigma - 2) continue;
    double a_min = layer ? c->h_min[K] : 0.;
    double *a = b + (size_t)K * R->N2;
    memset(vm, 0, sizeof(double) * R->N2);
    memset(vp, 0, sizeof(double) * R->N2);
    memset(v2, 0, sizeof(double) * R->N2);
    if (layer) {
      for (int J = 1; J <= nj; J++)
        for (int I = 1; I <= ni; I++) A2(v2, J, I) = .5 * (A3(R->V, K, J, I) + A3(R->V, K - 1, J, I));
      bc2(R, v2);
    } else {
      for (int J = 0; J <= nj + 1; J++)
        for (int I = 0; I <= ni + 1; I++) A2(v2, J, I) = A3(R->V, K, J, I);
    }
    for (int J = 0; J <= nj + 1; J++) {
      double mi = (fabs(c->lat[2 * J]) == 90.) ? 0. : 1. / c->m[2 * J];
      for (int I = 1; I <= ni; I++) {
        A2(vm, J, I) = mnv(0., A2(v2, J, I) * mi);
        A2(vp, J, I) = mx(0., A2(v2, J, I) * mi);
      }
    }
    memset(fl, 0, sizeof(double) * R->N2);
    for (int J = 1; J <= nj; J++)
      for (int I = 1; I <= ni; I++) A2(fl, J, I) = A2(vp, J, I) * A2(a, J - 1, I) + A2(vm, J, I) * A2(a, J, I);
    bc2(R, fl);
    memset(ap, 0, sizeof(double) * R->N2);
    for (int J = 0; J <= nj; J++) {
      double mn = c->mn[2 * J + 1];
      for (int I = 1; I <= ni; I++) A2(ap, J, I) = A2(a, J, I) + (A2(fl, J, I) - A2(fl, J + 1, I)) * mn * DT;
    }
    bc2(R, ap);
    for (int J = 1; J <= nj; J++)
      for (int I = 1; I <= ni; I++)
        A2(fl, J, I) = .5 * (A2(vp, J, I) * (A2(ap, J, I) + A2(a, J - 1, I)) + A2(vm, J, I) * (A2(ap, J - 1, I) + A2(a, J, I)));
    memset(df1, 0, sizeof(double) * R->N2);
    memset(df2, 0, sizeof(double) * R->N2);
    for (int J = 1; J < nj; J++)
      for (int I = 1; I <= ni; I++) {
        A2(df1, J, I) = A2(ap, J, I) - A2(a, J - 1, I);
        A2(df2, J, I) = A2(ap, J, I) - A2(a, J + 1, I);
      }
    for (int I = 1; I <= ni; I++) {
      A2(df1, 0, I) = 0.;
      A2(df2, 0, I) = A2(ap, 0, I) - A2(a, 1, I);
      A2(df1, nj, I) = A2(ap, nj, I) - A2(a, nj - 1, I);
      A2(df2, nj, I) = 0.;
    }
    bc2(R, df1);
    bc2(R, df2);
    memset(aq, 0, sizeof(double) * R->N2);
    for (int J = 1; J < nj; J++)
      for (int I = 1; I <= ni; I++) {
        double t = fabs(A2(a, J - 1, I) - 2. * A2(a, J, I) + A2(a, J + 1, I)) + ep;
        A2(aq, J, I) = t * t;
      }
    for (int I = 1; I <= ni; I++) {
      double t = fabs(-A2(a, 0, I) + A2(a, 1, I)) + ep;
      A2(aq, 0, I) = t * t;
      t = fabs(A2(a, nj - 1, I) - A2(a, nj, I)) + ep;
      A2(aq, nj, I) = t * t;
    }
    bc2(R, aq);
    for (int J = 1; J <= nj; J++) {
      double mn = c->mn[2 * J];
      for (int I = 1; I <= ni; I++) {
        double vhp = sqrt(A2(vp, J, I) * A2(vp, J - 1, I)), vhm = -sqrt(A2(vm, J, I) * A2(vm, J + 1, I));
        double mu = (A2(vp, J, I) - A2(vm, J, I)) * mn * DT, al = (1. + mu) / 6.;
        double gp = A2(aq, J - 1, I) / (A2(aq, J - 1, I) + (A2(a, J - 1, I) - a_min) * (A2(a, J, I) - a_min));
        gp = gp * gp;
        double gm = A2(aq, J, I) / (A2(aq, J, I) + (A2(a, J - 1, I) - a_min) * (A2(a, J, I) - a_min));
        gm = gm * gm;
        double bp = 1. + (1. / (2. * al) - 1.) * gp, bm = 1. + (1. / (2. * al) - 1.) * gm;
        double bhp = 1. - gp, bhm = 1. - gm;
        A2(fl, J, I) += -al * (A2(vp, J, I) * bp * A2(df1, J, I) - vhp * bhp * A2(df1, J - 1, I) +
                               A2(vm, J, I) * bm * A2(df2, J - 1, I) - vhm * bhm * A2(df2, J, I));
      }
    }
    bc2(R, fl);
    for (int J = 0; J <= nj; J++) {
      double mn = c->mn[2 * J + 1];
      for (int I = 1; I <= ni; I++) A2(a, J, I) += (A2(fl, J, I) - A2(fl, J + 1, I)) * mn * DT;
    }
    bc2(R, a);
  }
  /* ---- zonal, per K, on a filtered copy ---- */
  double *um = R->b2[0], *up = R->b2[1], *u2 = R->b2[2], *af = R->b2[6];
  for (int K = 1; K <= nk; K++) {
    if (which == AV_THETA && K <= R->k_sigma - 2) continue;
    double a_min = layer ? c->h_min[K] : 0.;
    memcpy(af, b + (size_t)K * R->N2, sizeof(double) * R->N2);
    filter2(R, af, FK_H);
    double *a = af;
    memset(um, 0, sizeof(double) * R->N2);
    memset(up, 0, sizeof(double) * R->N2);
    memset(u2, 0, sizeof(double) * R->N2);
    if (layer) {
      for (int J = 0; J <= nj; J++)
        for (int I = 1; I <= ni; I++) A2(u2, J, I) = .5 * (A3(R->U, K, J, I) + A3(R->U, K - 1, J, I));
      bc2(R, u2);
    } else {
      for (int J = 0; J <= nj; J++)
        for (int I = 0; I <= ni + 1; I++) A2(u2, J, I) = A3(R->U, K, J, I);
    }
    for (int J = 0; J <= nj; J++) {
      double n_inv = 1. / c->n[2 * J + 1];
      for (int I = 0; I <= ni + 1; I++) {
        A2(um, J, I) = mnv(0., A2(u2, J, I) * n_inv);
        A2(up, J, I) = mx(0., A2(u2, J, I) * n_inv);
      }
    }
    memset(fl, 0, sizeof(double) * R->N2);
    for (int J = 0; J <= nj; J++)
      for (int I = 1; I <= ni; I++) A2(fl, J, I) = A2(up, J, I) * A2(a, J, I - 1) + A2(um, J, I) * A2(a, J, I);
    bc2(R, fl);
    memset(ap, 0, sizeof(double) * R->N2);
    for (int J = 0; J <= nj; J++) {
      double mn = c->mn[2 * J + 1];
      for (int I = 1; I <= ni; I++) A2(ap, J, I) = A2(a, J, I) + (A2(fl, J, I) - A2(fl, J, I + 1)) * mn * DT;
    }
    bc2(R, ap);
    memset(df1, 0, sizeof(double) * R->N2);
    memset(df2, 0, sizeof(double) * R->N2);
    for (int J = 0; J <= nj; J++)
      for (int I = 1; I <= ni; I++) {
        A2(fl, J, I) = .5 * (A2(up, J, I) * (A2(ap, J, I) + A2(a, J, I - 1)) + A2(um, J, I) * (A2(ap, J, I - 1) + A2(a, J, I)));
        A2(df1, J, I) = A2(ap, J, I) - A2(a, J, I - 1);
        A2(df2, J, I) = A2(ap, J, I) - A2(a, J, I + 1);
      }
    bc2(R, df1);
    bc2(R, df2);
    memset(aq, 0, sizeof(double) * R->N2);
    for (int J = 0; J <= nj; J++)
      for (int I = 1; I <= ni; I++) {
        double t = fabs(A2(a, J, I - 1) - 2. * A2(a, J, I) + A2(a, J, I + 1)) + ep;
        A2(aq, J, I) = t * t;
      }
    bc2(R, aq);
    for (int J = 0; J <= nj; J++) {
      double mn = c->mn[2 * J + 1];
      for (int I = 1; I <= ni; I++) {
        double uhp = sqrt(A2(up, J, I) * A2(up, J, I - 1)), uhm = -sqrt(A2(um, J, I) * A2(um, J, I + 1));
        double mu = (A2(up, J, I) - A2(um, J, I)) * mn * DT, al = (1. + mu) / 6.;
        double gp = A2(aq, J, I - 1) / (A2(aq, J, I - 1) + (A2(a, J, I - 1) - a_min) * (A2(a, J, I) - a_min));
        gp = gp * gp;
        double gm = A2(aq, J, I) / (A2(aq, J, I) + (A2(a, J, I - 1) - a_min) * (A2(a, J, I) - a_min));
        gm = gm * gm;
        double bp = 1. + (1. / (2. * al) - 1.) * gp, bm = 1. + (1. / (2. * al) - 1.) * gm;
        double bhp = 1. - gp, bhm = 1. - gm;
        A2(fl, J, I) += -al * (A2(up, J, I) * bp * A2(df1, J, I) - uhp * bhp * A2(df1, J, I - 1) +
                               A2(um, J, I) * bm * A2(df2, J, I - 1) - uhm * bhm * A2(df2, J, I));
      }
    }
    bc2(R, fl);
    for (int J = 0; J <= nj; J++) {
      double mn = c->mn[2 * J + 1];
      for (int I = 1; I <= ni; I++) A3(b, K, J, I) += (A2(fl, J, I) - A2(fl, J, I + 1)) * mn * DT;
    }
  }
  bc3(R, b);
}

/* akima_advection, src/shared/epic_flux.c:1502-1924: the monotonized
 * fourth-order centred scheme, directions split as in hsu: vertical column
 * pass, then meridional and zonal passes per layer or interface (the zonal
 * one on a filtered copy).  Differences dA of neighbouring values, their
 * harmonic-mean slopes dela = 2 dA- dA+ / (dA- + dA+) where the two agree
 * in sign (else 0), the edge value .5 (a- + a+) - (dela+ - dela-)/6 times the
 * velocity as the flux.  The vertical pass runs over the rows J = 0..nj
 * (the reference also runs it on the pad rows J = -1 and nj+1, where W3 is
 * zero and it changes nothing). */
static double akima_dela(double d0, double d1) {
  double tmp = d0 * d1;
  return (tmp > 0.) ? 2. * tmp / (d0 + d1) : 0.;
}
static void akima(cpuref *R, int which, double *b) {
  epic_case *c = &R->c;
  int ni = R->ni, nj = R->nj, nk = R->nk;
  const double DT = R->dt;
  const int layer = (which == AV_HDRY);
  const int kf = (which == AV_THETA) ? R->k_sigma - 1 : 1;
  double *wa = R->w_a, *w = R->w_wm, *da = R->w_d1, *dl = R->w_d2, *ff = R->w_ff;
  const double *dsi = c->dsgth_inv;
  /* ---- vertical ---- */
  for (int J = 0; J <= nj; J++)
    for (int I = 0; I <= ni + 1; I++) {
      if (layer) {
        for (int K = kf; K <= nk; K++) wa[K] = A3(b, K, J, I);
        for (int K = kf - 1; K <= nk; K++) w[K] = A3(R->W3, K, J, I);
        da[kf - 1] = 0.;
        for (int K = kf; K < nk; K++) da[K] = wa[K] - wa[K + 1];
        da[nk] = 0.;
        for (int K = kf; K <= nk; K++) dl[K] = akima_dela(da[K], da[K - 1]);
        ff[kf - 1] = 0.; /* mass_flux_top */
        for (int K = kf; K < nk; K++) ff[K] = w[K] * (.5 * (wa[K] + wa[K + 1]) - (dl[K] - dl[K + 1]) / 6.);
        ff[nk] = 0.; /* mass_flux_bot */
        for (int K = kf; K <= nk; K++) A3(b, K, J, I) += (ff[K] - ff[K - 1]) * dsi[2 * K] * DT;
      } else {
        for (int K = kf - 1; K <= nk; K++) wa[K] = A3(b, K, J, I);
        for (int K = kf; K <= nk; K++)
          w[K] = (K == 1 || K == nk) ? 0. : .5 * (A3(R->W3, K, J, I) + A3(R->W3, K - 1, J, I));
        for (int K = kf; K <= nk; K++) da[K] = wa[K - 1] - wa[K];
        dl[kf - 1] = 0.;
        for (int K = kf; K < nk; K++) dl[K] = akima_dela(da[K], da[K + 1]);
        dl[nk] = 0.;
        for (int K = kf; K <= nk; K++) ff[K] = w[K] * (.5 * (wa[K - 1] + wa[K]) - (dl[K - 1] - dl[K]) / 6.);
        ff[nk + 1] = 0.;
        for (int K = kf; K <= nk; K++) A3(b, K, J, I) += (ff[K + 1] - ff[K]) * dsi[2 * K + 1] * DT;
      }
    }
  /* ---- meridional, per layer or interface K ---- */
  double *v2 = R->b2[0], *dA = R->b2[1], *dL = R->b2[2], *fl = R->b2[3];
  for (int K = 1; K <= nk; K++) {
    if (which == AV_THETA && K <= R->k_sigma - 2) continue;
    double *a = b + (size_t)K * R->N2;
    memset(v2, 0, sizeof(double) * R->N2);
    memset(dA, 0, sizeof(double) * R->N2);
    memset(dL, 0, sizeof(double) * R->N2);
    memset(fl, 0, sizeof(double) * R->N2);
    if (layer) {
      for (int J = 1; J <= nj; J++)
        for (int I = 1; I <= ni; I++) A2(v2, J, I) = .5 * (A3(R->V, K, J, I) + A3(R->V, K - 1, J, I));
      bc2(R, v2);
    } else {
      for (int J = -1; J <= nj + 1; J++)
        for (int I = 0; I <= ni + 1; I++) A2(v2, J, I) = A3(R->V, K, J, I);
    }
    for (int J = 1; J <= nj; J++)
      for (int I = 1; I <= ni; I++) A2(dA, J, I) = A2(a, J, I) - A2(a, J - 1, I);
    bc2(R, dA);
    for (int J = 0; J <= nj; J++)
      for (int I = 1; I <= ni; I++) A2(dL, J, I) = akima_dela(A2(dA, J, I), A2(dA, J + 1, I));
    bc2(R, dL);
    for (int J = 1; J <= nj; J++) {
      double mi = (fabs(c->lat[2 * J]) == 90.) ? 0. : 1. / c->m[2 * J];
      for (int I = 1; I <= ni; I++)
        A2(fl, J, I) = A2(v2, J, I) * (.5 * (A2(a, J, I) + A2(a, J - 1, I)) - (A2(dL, J, I) - A2(dL, J - 1, I)) / 6.) * mi;
    }
    bc2(R, fl);
    for (int J = 0; J <= nj; J++) {
      double mn = c->mn[2 * J + 1];
      for (int I = 1; I <= ni; I++) A2(a, J, I) += (A2(fl, J, I) - A2(fl, J + 1, I)) * mn * DT;
    }
    bc2(R, a);
  }
  /* ---- zonal, per K, on a filtered copy ---- */
  double *u2 = R->b2[0], *af = R->b2[6];
  for (int K = 1; K <= nk; K++) {
    if (which == AV_THETA && K <= R->k_sigma - 2) continue;
    memcpy(af, b + (size_t)K * R->N2, sizeof(double) * R->N2);
    filter2(R, af, FK_H);
    double *a = af;
    memset(u2, 0, sizeof(double) * R->N2);
    memset(dA, 0, sizeof(double) * R->N2);
    memset(dL, 0, sizeof(double) * R->N2);
    memset(fl, 0, sizeof(double) * R->N2);
    if (layer) {
      for (int J = 0; J <= nj; J++)
        for (int I = 1; I <= ni; I++) A2(u2, J, I) = .5 * (A3(R->U, K, J, I) + A3(R->U, K - 1, J, I));
      bc2(R, u2);
    } else {
      for (int J = -1; J <= nj + 1; J++)
        for (int I = 0; I <= ni + 1; I++) A2(u2, J, I) = A3(R->U, K, J, I);
    }
    for (int J = 0; J <= nj; J++)
      for (int I = 1; I <= ni; I++) A2(dA, J, I) = A2(a, J, I) - A2(a, J, I - 1);
    bc2(R, dA);
    for (int J = 0; J <= nj; J++)
      for (int I = 1; I <= ni; I++) A2(dL, J, I) = akima_dela(A2(dA, J, I), A2(dA, J, I + 1));
    bc2(R, dL);
    for (int J = 0; J <= nj; J++) {
      double n_inv = 1. / c->n[2 * J + 1];
      for (int I = 1; I <= ni; I++)
        A2(fl, J, I) = A2(u2, J, I) * (.5 * (A2(a, J, I) + A2(a, J, I - 1)) - (A2(dL, J, I) - A2(dL, J, I - 1)) / 6.) * n_inv;
    }
    bc2(R, fl);
    for (int J = 0; J <= nj; J++) {
      double mn = c->mn[2 * J + 1];
      for (int I = 1; I <= ni; I++) A3(b, K, J, I) += (A2(fl, J, I) - A2(fl, J, I + 1)) * mn * DT;
    }
  }
  bc3(R, b);
}

/* advection, src/shared/epic_flux.c:346-637 (dry: unity, theta, hdry); the
 * unity field and theta take theta's scheme, hdry its own */
static void advection(cpuref *R) {
  int nk = R->nk, nj = R->nj, ni = R->ni;
  void (*nonmass)(cpuref *, int, double *) = R->c.adv_nonmass ? akima : hsu;
  for (int K = 0; K <= nk + 1; K++)
    for (int J = 0; J <= nj; J++)
      for (int I = 0; I <= ni + 1; I++) A3(R->unity, K, J, I) = 1.;
  nonmass(R, AV_P3, R->unity);
  for (int K = R->k_sigma - 1; K <= nk + 1; K++)
    for (int J = 0; J <= nj; J++)
      for (int I = 0; I <= ni + 1; I++) A3(R->old, K, J, I) = A3(R->THETA, K, J, I);
  nonmass(R, AV_THETA, R->THETA);
  for (int K = R->k_sigma - 1; K <= nk + 1; K++)
    for (int J = 0; J <= nj; J++)
      for (int I = 0; I <= ni + 1; I++)
        A3(R->THETA, K, J, I) -= A3(R->old, K, J, I) * (A3(R->unity, K, J, I) - 1.);
  if (R->c.adv_mass) akima(R, AV_HDRY, R->HDRY);
  else hsu(R, AV_HDRY, R->HDRY);
  restore_mass(R);
  set_p2etc(R, 0);
}

/* Adams-Bashforth 1/2/3, src/shared/epic_timestep.c:406-498 */
static void ab_step(cpuref *R, double *val, double **tend, int jlo) {
  double ab0, ab1, ab2;
  if (A3(tend[R->itm2], 1, jlo, 1) == DBL_MAX) {
    if (A3(tend[R->itm1], 1, jlo, 1) == DBL_MAX) { ab0 = 1.; ab1 = 0.; ab2 = 0.; }
    else { ab0 = 3. / 2.; ab1 = -1. / 2.; ab2 = 0.; }
  } else {
    ab0 = 23. / 12.; ab1 = -16. / 12.; ab2 = 5. / 12.;
  }
  for (int K = 1; K < R->nk; K++)
    for (int J = jlo; J <= R->nj; J++)
      for (int I = 1; I <= R->ni; I++)
        A3(val, K, J, I) += R->dt * (ab0 * A3(tend[R->it0], K, J, I) + ab1 * A3(tend[R->itm1], K, J, I) +
                                     ab2 * A3(tend[R->itm2], K, J, I));
  bc3(R, val);
}

/* --------------------------------------------------------------- step */
int cpuref_sync(cpuref *R) {
  set_p2etc(R, 1);
  store_pgrad_vars(R);
  store_diag(R);
  calc_heating(R);
  calc_w(R);
  return R->err;
}

/* timestep(STEP_PROGS_AND_SYNC_DIAGS), src/shared/epic_timestep.c:104-327 */
int cpuref_step(cpuref *R) {
  int nk = R->nk, nj = R->nj, ni = R->ni;
  double *dudt = R->DUDT[R->it0], *dvdt = R->DVDT[R->it0];
  memset(dudt, 0, sizeof(double) * R->N3);
  memset(dvdt, 0, sizeof(double) * R->N3);
  source_sink(R);
  calc_w(R);
  uv_core(R, dudt, dvdt);
  uv_drag(R, dudt, dvdt);
  merid_hv(R, FK_H, R->HDRY);
  zonal_hv(R, FK_H, R->HDRY);
  restore_mass(R);
  merid_hv(R, FK_THETA, R->THETA);
  zonal_hv(R, FK_THETA, R->THETA);
  advection(R);
  store_pgrad_vars(R);
  /* filtered, monotone copy of GZ3 (src/shared/epic_timestep.c:176-196) */
  memcpy(R->gz, R->GZ3, sizeof(double) * R->N3);
  for (int K = nk - 1; K >= 0; K--) {
    filter2(R, R->gz + (size_t)K * R->N2, FK_DEF);
    for (int J = 0; J <= nj; J++)
      for (int I = 0; I <= ni + 1; I++) {
        double t = A3(R->gz, K + 1, J, I);
        if (cpuref_fcmp(t, 0.) == 0) A3(R->gz, K, J, I) = 1.e-3;
        else A3(R->gz, K, J, I) = mx(A3(R->gz, K, J, I), A3(R->gz, K + 1, J, I) + fabs(A3(R->gz, K + 1, J, I)) * 1.e-5);
      }
  }
  filter3(R, dudt, FK_U);
  filter3(R, dvdt, FK_V);
  filter3(R, R->HDRY, FK_H);
  restore_mass(R);
  filter3(R, R->THETA, FK_THETA);
  set_p2etc(R, 1);
  store_pgrad_vars(R);
  if (R->c.pgrad_traditional) uv_pgrad_traditional(R, dudt, dvdt);
  else uv_pgrad(R, dudt, dvdt);
  divergence_damping(R, dudt, dvdt);
  ab_step(R, R->U, R->DUDT, 0);
  ab_step(R, R->V, R->DVDT, 1);
  int t = R->itm2;
  R->itm2 = R->itm1;
  R->itm1 = R->it0;
  R->it0 = t;
  merid_hv(R, FK_U, R->U);
  merid_hv(R, FK_V, R->V);
  zonal_hv(R, FK_U, R->U);
  zonal_hv(R, FK_V, R->V);
  store_diag(R);
  calc_heating(R);
  R->itime++;
  return R->err;
}

/* ------------------------------------------------------------ lifetime */
static cpuref *finish_create(cpuref *R);

cpuref *cpuref_create(const epic_case_spec *spec) {
  cpuref *R = (cpuref *)calloc(1, sizeof(cpuref));
  if (epic_case_build(spec, &R->c)) {
    free(R);
    return NULL;
  }
  return finish_create(R);
}

/* the state of an epic.nc file, its Adams-Bashforth history included, as
 * the reference's epic_main takes it up (var_read, then HDRY from P3) */
cpuref *cpuref_create_nc(const char *path) {
  cpuref *R = (cpuref *)calloc(1, sizeof(cpuref));
  if (epic_nc_read(path, &R->c)) {
    free(R);
    return NULL;
  }
  return finish_create(R);
}

static cpuref *finish_create(cpuref *R) {
  epic_case *c = &R->c;
  R->ni = c->ni; R->nj = c->nj; R->nk = c->nk;
  R->Ia = c->Iadim; R->N2 = c->Nelem2d; R->N3 = c->Nelem3d;
  R->k_sigma = c->k_sigma; R->k_sponge = c->k_sponge;
  R->dt = (double)c->dt;
  R->is_hs = (strcmp(c->planet_name, "held_suarez") == 0);
  R->is_hj = (strcmp(c->planet_name, "hot_jupiter") == 0) && c->ntp == 0;
  R->is_gas = (strcmp(c->planet_type, "gas-giant") == 0);
  R->is_tab = c->newt_cool_on && !R->is_hs && c->ntp > 0;
  if (R->is_tab) {
    /* get_sounding's first call (src/shared/epic_funcs_diag.c:3200-3256):
     * the temperature table in -log p and its PCHIP derivatives, then the
     * negative-slope regions of theta ironed out from the top down, each
     * ironed level's temperature formed again from its theta (the table's
     * derivatives stay those of the original temperatures, as there); and
     * t_rad's first call (src/shared/epic_sensible_heating.c:264-271) */
    int n = R->ns = c->ntp;
    R->sx = (double *)calloc(n, sizeof(double));
    R->sy = (double *)calloc(n, sizeof(double));
    R->sz = (double *)calloc(n, sizeof(double));
    double *th = (double *)calloc(n, sizeof(double));
    for (int k = 0; k < n; k++) {
      R->sx[k] = -(log)(c->pdat[k]);
      R->sy[k] = c->tdat[k];
      th[k] = c->tdat[k] * (pow)(c->press0 / c->pdat[k], c->kappa);
    }
    if (cpuref_pchip(n, R->sx, R->sy, R->sz)) R->err |= 4;
    for (int k = n - 2; k >= 0; k--) {
      double tmp = th[k + 1];
      if (th[k] > tmp) {
        th[k] = tmp;
        R->sy[k] = th[k] * (pow)(c->pdat[k] / c->press0, c->kappa);
      }
    }
    free(th);
    n = R->nc = c->n_t_cool;
    R->cx = (double *)calloc(n, sizeof(double));
    R->cy = (double *)calloc(n, sizeof(double));
    R->cz = (double *)calloc(n, sizeof(double));
    for (int k = 0; k < n; k++) {
      R->cx[k] = c->t_cool_x[k];
      R->cy[k] = c->t_cool_y[k];
    }
    if (cpuref_pchip(n, R->cx, R->cy, R->cz)) R->err |= 4;
  }
  size_t N3 = R->N3, N2 = R->N2;
#define ALLOC3(p) p = (double *)calloc(N3, sizeof(double))
  ALLOC3(R->U); ALLOC3(R->V); ALLOC3(R->HDRY); ALLOC3(R->THETA); ALLOC3(R->P3); ALLOC3(R->P2);
  ALLOC3(R->THETA2); ALLOC3(R->H3); ALLOC3(R->T2); ALLOC3(R->T3); ALLOC3(R->RHO2); ALLOC3(R->RHO3);
  ALLOC3(R->EXNER3); ALLOC3(R->GZ3); ALLOC3(R->MONT3); ALLOC3(R->HEAT3); ALLOC3(R->PV3); ALLOC3(R->DIV);
  ALLOC3(R->W3); ALLOC3(R->DZDT3); ALLOC3(R->RI2); ALLOC3(R->USPIN); ALLOC3(R->gz); ALLOC3(R->sum_div);
  ALLOC3(R->unity); ALLOC3(R->old); ALLOC3(R->varz);
  for (int p = 0; p < 3; p++) { ALLOC3(R->DUDT[p]); ALLOC3(R->DVDT[p]); }
#undef ALLOC3
  for (int q = 0; q < 8; q++) R->b2[q] = (double *)calloc(N2, sizeof(double));
  R->GZS = (double *)calloc(N2, sizeof(double));
  R->gzu2d = (double *)calloc(N2, sizeof(double));
  R->gzv2d = (double *)calloc(N2, sizeof(double));
  R->gz2d = (double *)calloc(N2, sizeof(double));
  R->hl_h = (double *)calloc(N2, sizeof(double));
  R->hl_pv = (double *)calloc(N2, sizeof(double));
  int nkk = R->nk + 3;
  double **cols[] = {&R->w_wm, &R->w_wp, &R->w_a, &R->w_pred, &R->w_min, &R->w_ff, &R->w_d1, &R->w_d2, &R->w_aaa,
                     &R->tx, &R->ty, &R->tz};
  for (size_t q = 0; q < sizeof(cols) / sizeof(cols[0]); q++) *cols[q] = (double *)calloc(nkk, sizeof(double));
  R->mn_inv = (double *)calloc(2 * (R->nj + 1) + 1, sizeof(double));
  for (int jj = 0; jj <= 2 * (R->nj + 1); jj++) R->mn_inv[jj] = 1. / c->mn[jj];
  R->hcx = (double *)calloc(R->nj + 2, sizeof(double)); R->hcy = (double *)calloc(R->nj + 2, sizeof(double));
  R->hcz = (double *)calloc(R->nj + 2, sizeof(double)); R->vcx = (double *)calloc(R->nj + 2, sizeof(double));
  R->vcy = (double *)calloc(R->nj + 2, sizeof(double)); R->vcz = (double *)calloc(R->nj + 2, sizeof(double));
  R->fdata = (double *)calloc(R->ni + 1, sizeof(double));
  filter_init(R);
  d2dy2_init(R);
  /* hyperviscosity coefficients, factored (src/shared/epic_subgrid.c:82-158, 627-650) */
  {
    double dt = (double)c->dt, dy2 = c->dy0 * c->dy0, mxn[EPIC_MAX_NU_ORDER + 1];
    mxn[0] = (1. / 3.) / dt;
    mxn[2] = (dy2 / 30.) / dt;
    mxn[4] = (dy2 / 240.) * (dy2 / dt);
    mxn[6] = (dy2 / 800.) * (dy2 / dt) * dy2;
    mxn[8] = (dy2 / 2400.) * (dy2 / dt) * dy2 * dy2;
    c->nu[2] = 0.;
    for (int i2 = 2; i2 <= EPIC_MAX_NU_ORDER; i2 += 2) {
      R->coef[i2] = (c->nu[i2] > 0.) ? c->nu[i2] : 0.01 * mxn[i2];
      for (int i = 2; i < i2; i += 2) R->coef[i2] /= R->coef[i];
    }
    double rln = c->re / sqrt(1. + (pow)(c->rp / c->re * tan(LAT0 * DEGR), 2.));
    R->m0_zonal = 1. / (rln * c->dln * DEGR);
    R->m0_div = R->m0_zonal;
    R->nudiv = c->nudiv_nondim * mxn[2];
  }
  {
    double e = 1.;
    while (1. + e != 1.) e *= .5;
    e *= 2.;
    R->eps13 = (pow)(e, 1. / 3.);
  }
  /* state */
  memcpy(R->U, c->u, sizeof(double) * N3);
  memcpy(R->V, c->v, sizeof(double) * N3);
  memcpy(R->THETA, c->theta, sizeof(double) * N3);
  memcpy(R->P3, c->p3, sizeof(double) * N3);
  memcpy(R->USPIN, c->u_spinup, sizeof(double) * N3);
  memcpy(R->GZS, c->gz_surface, sizeof(double) * N2);
  /* HDRY from P3 as epic_main reconstitutes it (src/shared/epic_main.c:290-297) */
  for (int K = 1; K <= R->nk; K++)
    for (int J = 0; J <= R->nj; J++)
      for (int I = 1; I <= R->ni; I++) A3(R->HDRY, K, J, I) = h_layer(R, K, J, I);
  bc3(R, R->HDRY);
  R->it0 = 0; R->itm1 = 1; R->itm2 = 2;
  if (c->dudt_hist && c->dvdt_hist) {
    /* IT_MINUS1 and IT_MINUS2 from the file (var_read) */
    for (int p = 1; p <= 2; p++) {
      memcpy(R->DUDT[p], c->dudt_hist + (size_t)(p - 1) * N3, sizeof(double) * N3);
      memcpy(R->DVDT[p], c->dvdt_hist + (size_t)(p - 1) * N3, sizeof(double) * N3);
    }
  } else {
    for (int K = 1; K <= R->nk; K++)
      for (int J = 0; J <= R->nj; J++)
        for (int I = 1; I <= R->ni; I++) {
          A3(R->DUDT[2], K, J, I) = A3(R->DUDT[1], K, J, I) = DBL_MAX;
          A3(R->DVDT[2], K, J, I) = A3(R->DVDT[1], K, J, I) = DBL_MAX;
        }
  }
  R->itime = 0;
  return R;
}

void cpuref_destroy(cpuref *R) {
  if (!R) return;
  double *all[] = {R->U, R->V, R->HDRY, R->THETA, R->P3, R->P2, R->THETA2, R->H3, R->T2, R->T3, R->RHO2,
                   R->RHO3, R->EXNER3, R->GZ3, R->MONT3, R->HEAT3, R->PV3, R->DIV, R->W3, R->DZDT3, R->RI2, R->USPIN,
                   R->gz, R->sum_div, R->unity, R->old, R->varz, R->GZS, R->gzu2d, R->gzv2d, R->gz2d, R->hl_h,
                   R->hl_pv, R->w_wm, R->w_wp, R->w_a, R->w_pred, R->w_min, R->w_ff, R->w_d1, R->w_d2, R->w_aaa,
                   R->tx, R->ty, R->tz, R->mn_inv, R->hcx, R->hcy, R->hcz, R->vcx, R->vcy, R->vcz, R->fdata,
                   R->sx, R->sy, R->sz, R->cx, R->cy, R->cz};
  for (size_t q = 0; q < sizeof(all) / sizeof(all[0]); q++) free(all[q]);
  for (int p = 0; p < 3; p++) { free(R->DUDT[p]); free(R->DVDT[p]); }
  for (int q = 0; q < 8; q++) free(R->b2[q]);
  epic_case_free(&R->c);
  free(R);
}

int cpuref_read(cpuref *R, int field, double *out) {
  double *src = NULL;
  size_t n = R->N3;
  switch (field) {
    case EPIC_F_U: src = R->U; break;
    case EPIC_F_V: src = R->V; break;
    case EPIC_F_HDRY: src = R->HDRY; break;
    case EPIC_F_THETA: src = R->THETA; break;
    case EPIC_F_P3: src = R->P3; break;
    case EPIC_F_P2: src = R->P2; break;
    case EPIC_F_H3: src = R->H3; break;
    case EPIC_F_THETA2: src = R->THETA2; break;
    case EPIC_F_T2: src = R->T2; break;
    case EPIC_F_T3: src = R->T3; break;
    case EPIC_F_RHO2: src = R->RHO2; break;
    case EPIC_F_RHO3: src = R->RHO3; break;
    case EPIC_F_EXNER3: src = R->EXNER3; break;
    case EPIC_F_GZ3: src = R->GZ3; break;
    case EPIC_F_MONT3: src = R->MONT3; break;
    case EPIC_F_PV3: src = R->PV3; break;
    case EPIC_F_DIV_UV3: src = R->DIV; break;
    case EPIC_F_W3: src = R->W3; break;
    case EPIC_F_HEAT3: src = R->HEAT3; break;
    case EPIC_F_DZDT3: src = R->DZDT3; break;
    case EPIC_F_DUDT0: case EPIC_F_DUDT1: case EPIC_F_DUDT2: src = R->DUDT[field - EPIC_F_DUDT0]; break;
    case EPIC_F_DVDT0: case EPIC_F_DVDT1: case EPIC_F_DVDT2: src = R->DVDT[field - EPIC_F_DVDT0]; break;
    case EPIC_F_GZ_SURFACE: src = R->GZS; n = R->N2; break;
    case EPIC_F_RI2: src = R->RI2; break;
    default: return -1;
  }
  memcpy(out, src, sizeof(double) * n);
  return 0;
}

int cpuref_write(cpuref *R, int field, const double *in) {
  if (field == EPIC_F_GZ_SURFACE || (field >= EPIC_F_DUDT0 && field <= EPIC_F_DVDT2)) return -1;
  double *tmp = (double *)malloc(sizeof(double) * R->N3);
  if (!tmp || cpuref_read(R, field, tmp) != 0) {
    free(tmp);
    return -1;
  }
  free(tmp);
  double *dst = NULL;
  switch (field) {
    case EPIC_F_U: dst = R->U; break;
    case EPIC_F_V: dst = R->V; break;
    case EPIC_F_HDRY: dst = R->HDRY; break;
    case EPIC_F_THETA: dst = R->THETA; break;
    case EPIC_F_P3: dst = R->P3; break;
    case EPIC_F_P2: dst = R->P2; break;
    case EPIC_F_H3: dst = R->H3; break;
    case EPIC_F_THETA2: dst = R->THETA2; break;
    case EPIC_F_T2: dst = R->T2; break;
    case EPIC_F_T3: dst = R->T3; break;
    case EPIC_F_RHO2: dst = R->RHO2; break;
    case EPIC_F_RHO3: dst = R->RHO3; break;
    case EPIC_F_EXNER3: dst = R->EXNER3; break;
    case EPIC_F_GZ3: dst = R->GZ3; break;
    case EPIC_F_MONT3: dst = R->MONT3; break;
    case EPIC_F_PV3: dst = R->PV3; break;
    case EPIC_F_DIV_UV3: dst = R->DIV; break;
    case EPIC_F_W3: dst = R->W3; break;
    case EPIC_F_HEAT3: dst = R->HEAT3; break;
    case EPIC_F_DZDT3: dst = R->DZDT3; break;
    case EPIC_F_RI2: dst = R->RI2; break;
    default: return -1;
  }
  memcpy(dst, in, sizeof(double) * R->N3);
  return 0;
}

void cpuref_richardson(cpuref *R) {
  memset(R->RI2, 0, sizeof(double) * R->N3);
  for (int J = 0; J <= R->nj; J++)
    for (int I = 1; I <= R->ni; I++)
      for (int K = 1; K <= R->nk; K++) A3(R->RI2, K, J, I) = richardson_layer(R, K, J, I);
  bc3(R, R->RI2);
}

void cpuref_time_planes(cpuref *R, int *it3, long *steps) {
  it3[0] = R->it0; it3[1] = R->itm1; it3[2] = R->itm2;
  *steps = R->itime;
}
int cpuref_error(cpuref *R) { return R->err; }
void cpuref_dims(cpuref *R, int *d) {
  d[0] = R->ni; d[1] = R->nj; d[2] = R->nk; d[3] = R->N3; d[4] = R->N2;
}

/* exp (fn 0), log (1) or pow (2) as this build of the oracle evaluates
 * them: the host libm, or crmath.h in the CPUREF_CRMATH build */
double cpuref_math(int fn, double x, double y) {
  if (fn == 0) return exp(x);
  if (fn == 1) return log(x);
  return pow(x, y);
}
/* calls of exp, log, pow so far and how many of them glibc rounded
 * differently from crmath.h (CPUREF_CRMATH build; zeros otherwise) */
/* CPUREF_CRMATH build: which of exp (bit 0), log (1), pow (2) take
 * crmath.h's result (default all); the others take glibc's */
void cpuref_cr_select(int mask) {
#ifdef CPUREF_CRMATH
  cpuref_cr_mask = mask;
#else
  (void)mask;
#endif
}
void cpuref_libm_stats(long *calls3, long *diffs3, int reset) {
  for (int q = 0; q < 3; q++) {
#ifdef CPUREF_CRMATH
    calls3[q] = cpuref_cr_calls[q];
    diffs3[q] = cpuref_cr_diffs[q];
    if (reset) cpuref_cr_calls[q] = cpuref_cr_diffs[q] = 0;
#else
    (void)reset;
    calls3[q] = diffs3[q] = 0;
#endif
  }
}
