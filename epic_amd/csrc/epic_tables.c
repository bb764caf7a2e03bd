/*
 * epic_tables.c -- the run-once tables of the engine: the callee-internal
 * statics of the reference, formed on the host with the host libm and the
 * same expressions as their producers:
 *   zonal_filter gains       src/shared/epic_adjust.c:148-242
 *   four1/realft twiddles    src/shared/epic_funcs_util.c:4023-4135 (the
 *                            trigonometric recurrences, run once here)
 *   d2dy2 weights            src/shared/epic_subgrid.c:1049-1140
 *   set_max_nu, coef[]       src/shared/epic_subgrid.c:82-158, 627-650
 *   divergence damping       src/shared/epic_subgrid.c:1814-1847
 *   get_kin mn_inv           src/shared/epic_funcs_diag.c:2718-2732
 *   g_sigma coefficient      src/shared/epic_funcs_diag.c:1297-1299
 *   machine_epsilon          src/shared/epic_funcs_util.c:1247-1264
 *   sounding and cooling-time PCHIP tables of the Newtonian cooling
 *                            src/shared/epic_funcs_diag.c:3200-3256,
 *                            src/shared/epic_sensible_heating.c:264-271
 */
#include "tables.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

#define DEGR (M_PI / 180.)
#define LAT0 45.

/* sign of a*b without forming it (the Fritsch-Carlson sign test,
 * src/shared/epic_funcs_util.c:1000-1054) */
static double sign_of_product(double a, double b) {
  if (a == 0. || b == 0.) return 0.;
  return ((a < 0.) == (b < 0.)) ? 1. : -1.;
}

/* Monotone piecewise-cubic (PCHIP) derivatives at the points of an ascending
 * table, each formed from its own three neighbours (Fritsch & Carlson with
 * Brodlie's weights inside, the three-point shape-preserving formula at the
 * ends), as spline_pchip (src/shared/epic_funcs_util.c:721-925) forms them
 * along its sweep.  Returns -1 unless x rises strictly and n >= 2. */
static int pchip_slopes(int n, const double *x, const double *y, double *d) {
  if (n < 2) return -1;
  for (int i = 1; i < n; i++)
    if (!(x[i] > x[i - 1])) return -1;
  if (n == 2) {
    d[0] = d[1] = (y[1] - y[0]) / (x[1] - x[0]);
    return 0;
  }
  for (int i = 1; i + 1 < n; i++) {
    const double h1 = x[i] - x[i - 1], h2 = x[i + 1] - x[i], hsum = h1 + h2;
    const double del1 = (y[i] - y[i - 1]) / h1, del2 = (y[i + 1] - y[i]) / h2;
    d[i] = 0.;
    if (sign_of_product(del1, del2) > 0.) {
      const double h3 = 3. * hsum, a1 = (hsum + h1) / h3, a2 = (hsum + h2) / h3;
      const double big = fmax(fabs(del1), fabs(del2)), small = fmin(fabs(del1), fabs(del2));
      d[i] = small / (a1 * (del1 / big) + a2 * (del2 / big));
    }
  }
  for (int end = 0; end < 2; end++) {
    /* the end point e, its neighbour and the next: the slope of the nearer
     * interval weighted against the farther one */
    const int e = end ? n - 1 : 0;
    const double hn = end ? x[e] - x[e - 1] : x[1] - x[0];
    const double hf = end ? x[e - 1] - x[e - 2] : x[2] - x[1];
    const double dn = end ? (y[e] - y[e - 1]) / hn : (y[1] - y[0]) / hn;
    const double df = end ? (y[e - 1] - y[e - 2]) / hf : (y[2] - y[1]) / hf;
    const double hsum = hn + hf;
    const double v = ((hn + hsum) / hsum) * dn + (-hn / hsum) * df;
    double r = v;
    if (sign_of_product(v, dn) <= 0.) {
      r = 0.;
    } else if (sign_of_product(df, dn) < 0.) {
      const double lim = 3. * dn;
      if (fabs(lim) < fabs(v)) r = lim;
    }
    d[e] = r;
  }
  return 0;
}

int epic_case_table_cooling(const epic_case *c) {
  return c->newt_cool_on && strcmp(c->planet_name, "held_suarez") != 0 && c->ntp > 0;
}

/* The sounding table of get_sounding's first call: T against -log p, its
 * PCHIP slopes from those temperatures, then theta = T (press0/p)^kappa
 * ironed from the top down (a level warmer in theta than the one above it
 * takes that one's theta) and each ironed level's T formed again from its
 * theta; the slopes are not formed again.  And the cooling-time table with
 * its slopes (t_rad's first call). */
static int cooling_tables(const epic_case *c, epic_tables *T) {
  const int n = c->ntp, m = c->n_t_cool;
  if (n < 2 || m < 2) return -1;
  for (int q = 0; q < 3; q++) {
    T->snd[q] = (double *)calloc((size_t)n, sizeof(double));
    T->cool[q] = (double *)calloc((size_t)m, sizeof(double));
  }
  double *th = (double *)calloc((size_t)n, sizeof(double));
  for (int k = 0; k < n; k++) {
    if (!(c->pdat[k] > 0.)) { free(th); return -1; }
    T->snd[0][k] = -log(c->pdat[k]);
    T->snd[1][k] = c->tdat[k];
    th[k] = c->tdat[k] * pow(c->press0 / c->pdat[k], c->kappa);
  }
  int rc = pchip_slopes(n, T->snd[0], T->snd[1], T->snd[2]);
  for (int k = n - 2; k >= 0; k--)
    if (th[k] > th[k + 1]) {
      th[k] = th[k + 1];
      T->snd[1][k] = th[k] * pow(c->pdat[k] / c->press0, c->kappa);
    }
  free(th);
  for (int k = 0; k < m; k++) {
    T->cool[0][k] = c->t_cool_x[k];
    T->cool[1][k] = c->t_cool_y[k];
  }
  if (pchip_slopes(m, T->cool[0], T->cool[1], T->cool[2])) rc = -1;
  T->n_snd = n;
  T->n_cool = m;
  return rc;
}

static int four1_tw(int nn, int isign, double *tw) {
  int n = nn << 1, mmax = 2, k = 0;
  while (n > mmax) {
    double theta = isign * (2 * M_PI / mmax);
    double wtemp = sin(0.5 * theta);
    double wpr = -2.0 * wtemp * wtemp, wpi = sin(theta), wr = 1.0, wi = 0.0;
    for (int m = 1; m < mmax; m += 2) {
      tw[k++] = wr;
      tw[k++] = wi;
      wr = (wtemp = wr) * wpr - wi * wpi + wr;
      wi = wi * wpr + wtemp * wpi + wi;
    }
    mmax <<= 1;
  }
  return k;
}

static int realft_tw(int n, int isign, double *tw) {
  double theta = M_PI / (double)(n >> 1);
  if (isign != 1) theta = -theta;
  double wtemp = sin(0.5 * theta);
  double wpr = -2.0 * wtemp * wtemp, wpi = sin(theta);
  double wr = 1.0 + wpr, wi = wpi;
  int k = 0;
  for (int i = 2; i <= (n >> 2); i++) {
    tw[k++] = wr;
    tw[k++] = wi;
    wr = (wtemp = wr) * wpr - wi * wpi + wr;
    wi = wi * wpr + wtemp * wpi + wi;
  }
  return k;
}

void epic_tables_free(epic_tables *T) {
  free(T->mn_inv); free(T->hl_h); free(T->hl_pv);
  for (int q = 0; q < 3; q++) { free(T->hc[q]); free(T->vc[q]); free(T->snd[q]); free(T->cool[q]); }
  free(T->twf); free(T->twi); free(T->trf); free(T->tri);
  free(T->rows_h); free(T->rows_v);
  memset(T, 0, sizeof(*T));
}

int epic_tables_build(const epic_case *c, epic_tables *T) {
  memset(T, 0, sizeof(*T));
  int ni = c->ni, nj = c->nj, nJ = 2 * (nj + 1) + 1;
  T->n_mn_inv = nJ;
  T->mn_inv = (double *)calloc(nJ, sizeof(double));
  for (int jj = 0; jj < nJ; jj++) T->mn_inv[jj] = 1. / c->mn[jj];
  /* filter gains, layout [J*(ni/2+2) + I], J = 0..nj+1 */
  double lat0 = LAT0 * DEGR;
  double rln = c->re / sqrt(1. + pow(c->rp / c->re * tan(lat0), 2.));
  double rlt = rln / (cos(lat0) * (pow(sin(lat0), 2.) + pow(c->re / c->rp * cos(lat0), 2.)));
  double m0 = 1. / (rln * c->dln * DEGR), n0 = 1. / (rlt * c->dlt * DEGR);
  int W = ni / 2 + 2;
  T->n_hl = (nj + 2) * W;
  T->hl_h = (double *)malloc(sizeof(double) * T->n_hl);
  T->hl_pv = (double *)malloc(sizeof(double) * T->n_hl);
  for (int q = 0; q < T->n_hl; q++) T->hl_h[q] = T->hl_pv[q] = 1.;
  for (int pass = 0; pass < 2; pass++) {
    double *hl = pass ? T->hl_pv : T->hl_h;
    for (int J = pass ? 1 : 0; J <= nj; J++) {
      int jj = pass ? 2 * J : 2 * J + 1;
      hl[J * W + 1] = 1.;
      double r = c->mn[jj] / (c->m[jj] * c->n[jj]);
      double t0 = (c->n[jj] / n0) / (c->m[jj] / m0);
      t0 = pow(t0, 2.00);
      for (int I = 2; I <= ni / 2 + 1; I++) {
        double t1 = sin((I - 1) * (c->dln) * DEGR / 2.);
        t1 = pow(t1, -2.00);
        t1 *= r * t0;
        t1 = (t1 < 1.) ? t1 : 1.;
        hl[J * W + I] = t1;
      }
    }
  }
  T->rows_h = (int *)calloc(nj + 2, sizeof(int));
  T->rows_v = (int *)calloc(nj + 2, sizeof(int));
  for (int J = 0; J <= nj; J++)
    if (!(fabs(c->lat[2 * J + 1]) < LAT0)) T->rows_h[T->nrow_h++] = J;
  for (int J = 1; J <= nj; J++)
    if (!(fabs(c->lat[2 * J]) < LAT0)) T->rows_v[T->nrow_v++] = J;
  /* d2dy2 weights for the h/u rows (J = 0..nj) and the v/pv rows (J = 0..nj+1) */
  for (int pass = 0; pass < 2; pass++) {
    int jmax = pass ? nj + 1 : nj;
    double **cc = pass ? T->vc : T->hc;
    for (int q = 0; q < 3; q++) cc[q] = (double *)calloc(nj + 2, sizeof(double));
    for (int J = 0; J <= jmax; J++) {
      int jj = pass ? 2 * J : 2 * J + 1;
      double r2 = (double)c->rlt[jj];
      r2 = 1. / (r2 * r2);
      double l0, l1, l2;
      int kind;
      if (J == 0) { l0 = c->lat[jj] * DEGR; l1 = c->lat[jj + 2] * DEGR; l2 = c->lat[jj + 4] * DEGR; kind = 0; }
      else if (J == jmax) { l0 = c->lat[jj - 4] * DEGR; l1 = c->lat[jj - 2] * DEGR; l2 = c->lat[jj] * DEGR; kind = 2; }
      else { l0 = c->lat[jj - 2] * DEGR; l1 = c->lat[jj] * DEGR; l2 = c->lat[jj + 2] * DEGR; kind = 1; }
      double s0 = sin(l0), s1 = sin(l1), s2 = sin(l2), x, y, z;
      if (kind == 0) {
        x = r2 * (1. + s0 * (-3. * s0 + s1 + s2)) / ((s1 - s0) * .5 * (s2 - s0));
        y = -r2 * 2. * (cos(2. * l0) + s0 * s2) / ((s1 - s0) * (s2 - s1));
        z = r2 * (cos(2. * l0) + s0 * s1) / (.5 * (s2 - s0) * (s2 - s1));
      } else if (kind == 1) {
        x = r2 * (cos(2. * l1) + s1 * s2) / ((s1 - s0) * .5 * (s2 - s0));
        y = -r2 * 2. * (1. + s1 * (s0 - 3. * s1 + s2)) / ((s1 - s0) * (s2 - s1));
        z = r2 * (cos(2. * l1) + s0 * s1) / (.5 * (s2 - s0) * (s2 - s1));
      } else {
        x = r2 * (cos(2. * l2) + s1 * s2) / ((s1 - s0) * .5 * (s2 - s0));
        y = -r2 * 2. * (cos(2. * l2) + s0 * s2) / ((s1 - s0) * (s2 - s1));
        z = r2 * (1. + s2 * (s0 + s1 - 3. * s2)) / (.5 * (s2 - s0) * (s2 - s1));
      }
      cc[0][J] = x;
      cc[1][J] = y;
      cc[2][J] = z;
    }
  }
  /* FFT twiddles: forward and inverse, four1 on ni/2 points and the split step */
  T->n_tw = ni;  /* 2*(ni/2 - 1) doubles used */
  T->n_tr = ni / 2;
  T->twf = (double *)calloc(T->n_tw, sizeof(double));
  T->twi = (double *)calloc(T->n_tw, sizeof(double));
  T->trf = (double *)calloc(T->n_tr, sizeof(double));
  T->tri = (double *)calloc(T->n_tr, sizeof(double));
  four1_tw(ni / 2, 1, T->twf);
  four1_tw(ni / 2, -1, T->twi);
  realft_tw(ni, 1, T->trf);
  realft_tw(ni, -1, T->tri);
  /* viscosity coefficients (nu[2] is forced to zero by the initialisers) */
  double dt = (double)c->dt, dy2 = c->dy0 * c->dy0;
  double mx2 = (dy2 / 30.) / dt, mx4 = (dy2 / 240.) * (dy2 / dt);
  double mx6 = (dy2 / 800.) * (dy2 / dt) * dy2, mx8 = (dy2 / 2400.) * (dy2 / dt) * dy2 * dy2;
  T->coef2 = 0.01 * mx2;
  T->coef4 = (c->nu[4] > 0.) ? c->nu[4] : 0.01 * mx4;
  T->coef4 /= T->coef2;
  T->coef6 = (c->nu[6] > 0.) ? c->nu[6] : 0.01 * mx6;
  T->coef6 /= T->coef2;
  T->coef6 /= T->coef4;
  T->coef8 = (c->nu[8] > 0.) ? c->nu[8] : 0.01 * mx8;
  T->coef8 /= T->coef2;
  T->coef8 /= T->coef4;
  T->coef8 /= T->coef6;
  double rlnz = c->re / sqrt(1. + pow(c->rp / c->re * tan(LAT0 * DEGR), 2.));
  T->m0z = 1. / (rlnz * c->dln * DEGR);
  T->nudiv = c->nudiv_nondim * mx2;
  double e = 1.;
  while (1. + e != 1.) e *= .5;
  e *= 2.;
  T->eps13 = pow(e, 1. / 3.);
  T->gcoef = 1. / (1 - exp(-c->hybrid_alpha * (1. - c->sigma_sigma)));
  T->mu3 = 1. / (c->rgas / 8.314472e+3);
  T->mu2 = .5 * (T->mu3 + T->mu3);
  if (epic_case_table_cooling(c) && cooling_tables(c, T)) return -1;
  return 0;
}
