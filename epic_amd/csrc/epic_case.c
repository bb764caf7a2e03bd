/*
 * epic_case.c -- synthetic cases for the EPIC dynamical core: planet
 * constants, the run-once metric arrays, the hybrid sigma-theta vertical
 * levels and a seeded initial state, all laid out as the reference lays
 * out its globals.
 *
 * Metric arrays follow the reference's producers term by term so that
 * they agree bit for bit when built with the same libm:
 *   set_lonlat   src/shared/epic_funcs_diag.c:1324-1415
 *   set_fmn      src/shared/epic_funcs_diag.c:1424-1593
 *   set_gravity  src/shared/epic_funcs_diag.c:1607-1633
 *   set_dsgth    src/shared/epic_funcs_diag.c:1639-1690
 * The vertical coordinate follows the construction of the `initial`
 * program (src/shared/epic_initial.c:1562-2006): log-p reference levels,
 * k_sigma, sigma_sigma, sigmatheta[kk], h_min[K].  The hybrid coordinate
 * functions are those of src/shared/epic_funcs_diag.c:1238-1314,1751-1770.
 *
 * The reference has no Gaussian generator; noise comes from a counter-based
 * generator (splitmix64 of seed, field and cell index) and Box-Muller, so a
 * case is a pure function of its spec.
 */
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "../../include/epic_amd.h"
#include "case_internal.h"

#define DEGR (M_PI / 180.)
#define LAT0_DEG 45.

/* ---------------------------------------------------------------- rng */
static uint64_t mix64(uint64_t z) {
  z += 0x9e3779b97f4a7c15ULL;
  z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
  z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
  return z ^ (z >> 31);
}
/* uniform in (0,1), from (seed, stream, index) */
static double uniform01(uint64_t seed, uint64_t stream, uint64_t idx) {
  uint64_t h = mix64(seed ^ mix64(stream * 0x100000001b3ULL ^ mix64(idx)));
  return ((double)(h >> 11) + 0.5) * (1.0 / 9007199254740992.0);
}
static double gauss(uint64_t seed, uint64_t stream, uint64_t idx) {
  double u1 = uniform01(seed, stream, 2 * idx), u2 = uniform01(seed, stream, 2 * idx + 1);
  return sqrt(-2. * log(u1)) * cos(2. * M_PI * u2);
}

/* ------------------------------------------------- hybrid coordinate */
typedef struct { double zeta0, zeta1, alpha, sigma_sigma; } hyb;

static double hyb_get_sigma(double pbot, double p, double ptop) {
  return log(p / pbot) / log(ptop / pbot);
}
static double hyb_p_sigma(double pbot, double sigma, double ptop) {
  return pbot * exp(sigma * log(ptop / pbot));
}
static double hyb_g(const hyb *h, double sigma) {
  const double coeff = 1. / (1 - exp(-h->alpha * (1. - h->sigma_sigma)));
  if (sigma <= h->sigma_sigma) return 0.;
  if (sigma >= 1.) return 1.;
  return coeff * (1. - exp(-h->alpha * (sigma - h->sigma_sigma)));
}
static double hyb_f(const hyb *h, double sigma) {
  if (sigma < 0.) return h->zeta0;
  if (sigma <= h->sigma_sigma) return h->zeta0 + sigma * (h->zeta1 - h->zeta0);
  if (sigma < 1.) return (1. - hyb_g(h, sigma)) * (h->zeta0 + sigma * (h->zeta1 - h->zeta0));
  return 0.;
}
static double hyb_sgth(const hyb *h, double theta, double p, double pbot, double ptop) {
  if (p <= ptop) return theta;
  double s = hyb_get_sigma(pbot, p, ptop);
  return hyb_f(h, s) + hyb_g(h, s) * theta;
}

/* ------------------------------------------------- planet profiles */
typedef struct {
  int kind;
  double p0, kappa;
  /* Held-Suarez equilibrium */
  /* constant-N^2 profile: theta(p) = 1/(1/theta_b + (a/kappa)((p/p0)^k - (pb/p0)^k)) */
  double theta_b, pb, a_over_kappa;
} profile;

static double theta_profile(const profile *pr, double latdeg, double p) {
  if (pr->kind == EPIC_PLANET_HELD_SUAREZ) {
    /* T_eq of Held & Suarez (1994), as temp_eq() writes it
     * (src/shared/epic_sensible_heating.c:404-412), as potential temperature. */
    double c2 = cos(latdeg * DEGR);
    c2 *= c2;
    double s2 = 1. - c2, pp0 = p / pr->p0;
    double t = (315. - 60. * s2 - 10. * log(pp0) * c2) * pow(pp0, pr->kappa);
    if (t < 200.) t = 200.;
    return t * pow(pr->p0 / p, pr->kappa);
  }
  double y = 1. / pr->theta_b +
             pr->a_over_kappa * (pow(p / pr->p0, pr->kappa) - pow(pr->pb / pr->p0, pr->kappa));
  return 1. / y;
}

/* ------------------------------------------------------------ build */
static double *zalloc(size_t n) { return (double *)calloc(n ? n : 1, sizeof(double)); }

size_t epic_member_footprint(int ni, int nj, int nk, int precision_mode) {
  size_t n3 = (size_t)(ni + 2) * (nj + 3) * (nk + 2);
  /* 37 FP64 3-D arrays of state, diagnostics and scratch, six tendency
   * planes (FP32 in the mixed mode, which adds two FP64 current planes)
   * and small tables. */
  size_t b = n3 * 8 * 37 + (precision_mode == EPIC_MIXED ? n3 * 8 * 2 + n3 * 4 * 6 : n3 * 8 * 6);
  return b + (size_t)(nj + 3) * (ni + 2) * 8 * 6 + 65536;
}

void epic_case_free(epic_case *c) {
  double **arrs[] = {&c->lon, &c->lat, &c->rln, &c->rlt, &c->f, &c->m, &c->n, &c->mn, &c->g,
                     &c->sigmatheta, &c->dsgth, &c->dsgth_inv, &c->p_ref, &c->theta_ref,
                     &c->h_min, &c->t_sponge_inv, &c->u, &c->v, &c->hdry, &c->theta,
                     &c->p3, &c->u_spinup, &c->gz_surface, &c->pdat, &c->tdat, &c->t_cool_x,
                     &c->t_cool_y, &c->dudt_hist, &c->dvdt_hist};
  for (size_t i = 0; i < sizeof(arrs) / sizeof(arrs[0]); i++) {
    free(*arrs[i]);
    *arrs[i] = NULL;
  }
  epic_nc_header_free(c->nc_header);
  c->nc_header = NULL;
}

/* set_lonlat (src/shared/epic_funcs_diag.c:1324-1415), globe branch */
static void build_lonlat(epic_case *c) {
  int nj = c->nj, ni = c->ni;
  c->lon[0] = c->globe_lonbot - c->dln;
  for (int ii = 1; ii <= 2 * (ni + 1); ii++) c->lon[ii] = c->lon[ii - 1] + c->dln * .5;
  c->lat[0] = -90. + c->dlt * sqrt(.5);
  for (int jj = 1; jj <= 2 * (nj + 1); jj++) c->lat[jj] = c->lat[jj - 1] + c->dlt * .5;
  c->lat[0] = -90.;
  c->lat[2 * (nj + 1)] = 90.;
  /* mirror symmetry across the equator, same sweep as the reference */
  int jjn = 2 * (nj + 1), jjs = 0;
  while (c->lat[jjs] < 0. && c->lat[jjn] > 0.) {
    for (; jjs < jjn; jjs++) {
      if (fabs(c->lat[jjs] + c->lat[jjn]) < 1.e-3) {
        c->lat[jjn--] = -c->lat[jjs++];
        break;
      }
    }
  }
}

/* set_fmn (src/shared/epic_funcs_diag.c:1424-1593), globe with both poles */
static void build_fmn(epic_case *c) {
  int nj = c->nj;
  double omega = c->omega, re = c->re, rp = c->rp;
  double dlnr = c->dln * DEGR, dltr = c->dlt * DEGR;
  for (int jj = 1; jj < 2 * (nj + 1); jj++) {
    double lat = DEGR * c->lat[jj];
    double rln = re / sqrt(1. + pow(rp / re * tan(lat), 2.));
    double rlt = rln / (cos(lat) * (pow(sin(lat), 2.) + pow(re / rp * cos(lat), 2.)));
    c->rln[jj] = rln;
    c->rlt[jj] = rlt;
    c->f[jj] = 2. * omega * sin(lat);
    c->m[jj] = 1. / (rln * dlnr);
    c->n[jj] = 1. / (rlt * dltr);
    c->mn[jj] = c->m[jj] * c->n[jj];
  }
  /* south pole */
  double lat = DEGR * c->lat[0];
  c->rln[0] = 0.;
  c->rlt[0] = re * re / rp;
  c->f[0] = 2. * omega * sin(lat);
  c->n[1] /= 1. + sqrt(.5);
  c->mn[1] = c->n[1] * c->n[1] / (.5 * dlnr);
  c->mn[0] = 2. * c->mn[1];
  /* north pole */
  int jp = 2 * (nj + 1);
  lat = DEGR * c->lat[jp];
  c->rln[jp] = 0.;
  c->rlt[jp] = re * re / rp;
  c->f[jp] = 2. * omega * sin(lat);
  c->n[jp - 1] /= 1. + sqrt(.5);
  c->mn[jp - 1] = c->n[jp - 1] * c->n[jp - 1] / (.5 * dlnr);
  c->mn[jp] = 2. * c->mn[jp - 1];
  /* dy at LAT0 */
  double lat0 = LAT0_DEG * DEGR;
  double rln = re / sqrt(1. + pow(rp / re * tan(lat0), 2.));
  double rlt = rln / (cos(lat0) * (pow(sin(lat0), 2.) + pow(re / rp * cos(lat0), 2.)));
  c->dy0 = rlt * c->dlt * DEGR;
}

/* set_gravity (src/shared/epic_funcs_diag.c:1607-1633) */
static void build_gravity(epic_case *c) {
  double a = c->re, b = c->rp;
  double spin = (c->omega * a) * (c->omega * a) * (a / c->GM);
  double ge = (1. + 1.5 * c->J2 - spin) * c->GM / (a * a);
  double gp = (1. - 3. * (a / b) * (a / b) * c->J2) * c->GM / (b * b);
  for (int jj = 0; jj <= 2 * (c->nj + 1); jj++) {
    double s2 = sin(c->lat[jj] * DEGR);
    s2 *= s2;
    double c2 = cos(c->lat[jj] * DEGR);
    c2 *= c2;
    c->g[jj] = (a * ge * c2 + b * gp * s2) / sqrt(a * a * c2 + b * b * s2);
  }
}

/* set_dsgth (src/shared/epic_funcs_diag.c:1639-1690) */
static int build_dsgth(epic_case *c) {
  int nk = c->nk;
  for (int K = 0; K <= nk; K++) {
    int kk = 2 * K;
    if (K == 0)
      c->dsgth[kk + 1] = c->sigmatheta[kk + 1] - c->sigmatheta[kk + 2];
    else if (K == nk)
      c->dsgth[kk + 1] = c->sigmatheta[kk] - c->sigmatheta[kk + 1];
    else
      c->dsgth[kk + 1] = c->sigmatheta[kk] - c->sigmatheta[kk + 2];
    if (!(c->dsgth[kk + 1] > 0.)) return -3;
    c->dsgth_inv[kk + 1] = 1. / c->dsgth[kk + 1];
  }
  for (int K = 1; K <= nk; K++) {
    int kk = 2 * K;
    c->dsgth[kk] = c->sigmatheta[kk - 1] - c->sigmatheta[kk + 1];
    if (!(c->dsgth[kk] > 0.)) return -3;
    c->dsgth_inv[kk] = 1. / c->dsgth[kk];
  }
  return 0;
}

/* Hsu-Arakawa layer pressure, onto_kk(P2_INDEX) (src/shared/epic_funcs_diag.c:2600-2616) */
static double layer_p(double kappa, double ptop_i, double pbot_i) {
  double kp1 = kappa + 1.;
  return pow((pow(pbot_i, kp1) - pow(ptop_i, kp1)) / (kp1 * (pbot_i - ptop_i)), 1. / kappa);
}

/* Root of sgth(theta(lat,p), p) = target for p in [plo, phi] (bisection in ln p). */
static double solve_level(const hyb *h, const profile *pr, double latdeg, double target,
                          double pbot, double ptop, double plo, double phi) {
  double a = log(plo), b = log(phi);
  double fa = hyb_sgth(h, theta_profile(pr, latdeg, plo), plo, pbot, ptop) - target;
  for (int it = 0; it < 200; it++) {
    double m = .5 * (a + b);
    if (m <= a || m >= b) break;
    double pm = exp(m);
    double fm = hyb_sgth(h, theta_profile(pr, latdeg, pm), pm, pbot, ptop) - target;
    if ((fm > 0.) == (fa > 0.)) {
      a = m;
      fa = fm;
    } else {
      b = m;
    }
  }
  return exp(.5 * (a + b));
}

/* The arrays of a case whose extents, k_sponge and dt are set (zeroed), and
 * the sponge rates of set_sponge (src/shared/epic_funcs_diag.c:1696-1731):
 * relax u to its initial value and v to zero, strongest (1/(30 dt)) in the
 * top layer. */
void epic_case_alloc(epic_case *c) {
  const int ni = c->ni, nj = c->nj, nk = c->nk;
  const int nJ = 2 * (nj + 1) + 1, nK = 2 * (nk + 1) + 1;
  c->Iadim = ni + 2;
  c->Jadim = nj + 3;
  c->Kadim = nk + 2;
  c->Nelem2d = c->Iadim * c->Jadim;
  c->Nelem3d = c->Nelem2d * c->Kadim;
  c->lon = zalloc(2 * (ni + 1) + 1);
  c->lat = zalloc(nJ);
  c->rln = zalloc(nJ);
  c->rlt = zalloc(nJ);
  c->f = zalloc(nJ);
  c->m = zalloc(nJ);
  c->n = zalloc(nJ);
  c->mn = zalloc(nJ);
  c->g = zalloc(nJ);
  c->sigmatheta = zalloc(nK);
  c->dsgth = zalloc(nK);
  c->dsgth_inv = zalloc(nK);
  c->p_ref = zalloc(nK);
  c->theta_ref = zalloc(nK);
  c->h_min = zalloc(nk + 2);
  c->t_sponge_inv = zalloc(c->k_sponge + 1);
  if (c->k_sponge > 0) {
    const double t0_inv = 1. / (30. * (double)c->dt);
    for (int K = 0; K <= c->k_sponge; K++) {
      const double tmp = (double)(c->k_sponge - K) / (c->k_sponge);
      c->t_sponge_inv[K] = t0_inv * .5 * (1. - cos(M_PI * tmp));
    }
  }
  const size_t N3 = (size_t)c->Nelem3d;
  c->u = zalloc(N3);
  c->v = zalloc(N3);
  c->hdry = zalloc(N3);
  c->theta = zalloc(N3);
  c->p3 = zalloc(N3);
  c->u_spinup = zalloc(N3);
  c->gz_surface = zalloc(c->Nelem2d);
}

/* lon/lat, the map factors and Coriolis parameter, gravity (set_lonlat,
 * set_fmn, set_gravity) from the grid and planet scalars */
void epic_case_metrics(epic_case *c) {
  build_lonlat(c);
  build_fmn(c);
  build_gravity(c);
}

/* dsgth and its inverse from sigmatheta (set_dsgth); 0 or -3 */
int epic_case_dsgth(epic_case *c) { return build_dsgth(c) ? -3 : 0; }

int epic_case_build(const epic_case_spec *s, epic_case *c) {
  memset(c, 0, sizeof(*c));
  if (s->ni < 8 || (s->ni & (s->ni - 1)) || s->nj < 4 || s->nk < 4 || s->dt <= 0) return -1;
  if (s->adv_mass < 0 || s->adv_mass > 1 || s->adv_nonmass < 0 || s->adv_nonmass > 1) return -1;
  if (s->pv_scheme < 0 || s->pv_scheme > 1) return -1;
  int ni = s->ni, nj = s->nj, nk = s->nk;
  c->ni = ni;
  c->nj = nj;
  c->nk = nk;
  c->dt = s->dt;

  /* ---- planet constants ---- */
  double gsurf, ptop, pbot, n2 = 0., theta_b = 0.;
  int gas = 0;
  switch (s->planet) {
    case EPIC_PLANET_HELD_SUAREZ:
      strcpy(c->planet_name, "held_suarez");
      c->planet_index = EPIC_HELD_SUAREZ_INDEX;
      c->re = c->rp = 6.371e6;
      c->omega = 7.292e-5;
      c->cp = 1004.;
      c->rgas = 287.;
      gsurf = 9.81;
      ptop = 2000.;
      pbot = 1.e5;
      c->newt_cool_on = 1;
      c->kinvisc = 14.60e-6;
      break;
    case EPIC_PLANET_JUPITER_LIKE:
      strcpy(c->planet_name, "jupiter_like");
      c->planet_index = 1;
      c->re = c->rp = 7.1492e7;
      c->omega = 1.7585e-4;
      c->cp = 12360.;
      c->rgas = 3614.;
      gsurf = 23.1;
      ptop = 1000.;
      pbot = 3.e5;
      n2 = 5.e-5;
      theta_b = 160.;
      gas = 1;
      c->kinvisc = 101.00e-6;
      break;
    case EPIC_PLANET_SATURN_LIKE:
      strcpy(c->planet_name, "saturn_like");
      c->planet_index = 2;
      c->re = c->rp = 6.0268e7;
      c->omega = 1.638e-4;
      c->cp = 11500.;
      c->rgas = 3500.;
      gsurf = 10.4;
      ptop = 1000.;
      pbot = 3.e5;
      n2 = 2.e-5;
      theta_b = 135.;
      gas = 1;
      c->kinvisc = 101.00e-6;
      break;
    case EPIC_PLANET_VENUS_LIKE:
      strcpy(c->planet_name, "venus_like");
      c->planet_index = 6;
      c->re = c->rp = 6.0518e6;
      c->omega = 2.99e-7;
      c->cp = 1000.;
      c->rgas = 191.4;
      gsurf = 8.87;
      ptop = 5000.;
      pbot = 1.e6;
      n2 = 1.e-4;
      theta_b = 400.;
      c->kinvisc = 14.60e-6;
      break;
    case EPIC_PLANET_HOT_JUPITER:
      strcpy(c->planet_name, "hot_jupiter");
      c->planet_index = 18; /* HOT_JUPITER_INDEX, include/epic_datatypes.h:89 */
      c->re = c->rp = 9.44e7;
      c->omega = 2.06e-5;
      c->cp = 14300.;
      c->rgas = 3523.;
      gsurf = 9.42;
      ptop = 100.;
      pbot = 1.e6;
      n2 = 1.5e-6;
      theta_b = 1500.;
      gas = 1;
      c->newt_cool_on = 1;
      c->kinvisc = 101.00e-6;
      break;
    default:
      return -1;
  }
  strcpy(c->planet_type, gas ? "gas-giant" : "terrestrial");
  c->cpr = c->cp / c->rgas;
  c->kappa = 1. / c->cpr;
  c->x_h2 = 0.;
  /* Uniform surface gravity from GM and J2 through set_gravity()'s formula:
   * with a = b, ge == gp requires J2 = s/4.5 (s = Omega^2 a^3/GM), and then
   * (1-3 J2) GM/a^2 = g gives GM = g a^2 + (2/3) Omega^2 a^3. */
  c->GM = gsurf * c->re * c->re + (2. / 3.) * c->omega * c->omega * c->re * c->re * c->re;
  c->J2 = c->omega * c->omega * c->re * c->re * c->re / (4.5 * c->GM);

  /* ---- grid scalars ---- */
  c->globe_lonbot = -180.;
  c->globe_lontop = 180.;
  c->globe_latbot = -90.;
  c->globe_lattop = 90.;
  c->dln = (c->globe_lontop - c->globe_lonbot) / ni;
  c->dlt = 180. / ((nj + 1) + sqrt(.5) + sqrt(.5));
  c->jtp = nj / 2;
  c->press0 = 1.e5;
  c->pbot = pbot;
  c->ptop = ptop;
  c->topo_scale = 1.;
  c->prandtl = s->prandtl;
  c->tau_drag = (s->tau_drag > 0.) ? s->tau_drag : 1.e20;
  c->drag_v = s->drag_v ? 1 : 0;
  c->drag_zonal_avg = s->drag_zonal_avg ? 1 : 0;
  c->newt_cool_adjust = s->newt_cool_adjust ? 1 : 0;
  c->pgrad_traditional = s->pgrad_traditional ? 1 : 0;
  c->adv_mass = s->adv_mass;
  c->adv_nonmass = s->adv_nonmass;
  c->pv_scheme = s->pv_scheme;
  c->nudiv_nondim = s->nudiv_nondim;
  /* Newtonian cooling switched on for a planet with no relaxation of its
   * own relaxes toward a sounding with a cooling-time table, as the
   * reference's other planets do (formed below, once the profile is known) */
  int want_sounding = 0;
  if (s->newt_cool > 0) {
    if (!c->newt_cool_on) want_sounding = 1;
    c->newt_cool_on = 1;
  } else if (s->newt_cool < 0) {
    c->newt_cool_on = 0;
  }
  /* the top sponge: the `initial` program's choice
   * (src/shared/epic_initial.c:1249-1266) or the spec's */
  if (s->k_sponge == EPIC_K_SPONGE_DEFAULT) {
    if (s->planet == EPIC_PLANET_HELD_SUAREZ || nk < 5) {
      c->k_sponge = 0;
    } else {
      const double x = (double)nk * .2;
      int k = x > 0. ? (int)(x + .5) : (int)(x - .5);
      c->k_sponge = k < 1 ? 1 : (k > 4 ? 4 : k);
    }
  } else if (s->k_sponge >= 0 && s->k_sponge < nk) {
    c->k_sponge = s->k_sponge;
  } else {
    return -1;
  }

  /* ---- arrays ---- */
  epic_case_alloc(c);
  epic_case_metrics(c);

  /* ---- reference profile ---- */
  profile pr;
  memset(&pr, 0, sizeof(pr));
  pr.kind = s->planet;
  pr.p0 = c->press0;
  pr.kappa = c->kappa;
  if (s->planet != EPIC_PLANET_HELD_SUAREZ) {
    pr.pb = pbot;
    pr.theta_b = theta_b;
    pr.a_over_kappa = n2 * c->rgas / (gsurf * gsurf) / c->kappa;
  }
  if (s->planet == EPIC_PLANET_HOT_JUPITER) {
    c->teq_contrast = 1000.;
    c->teq_p_contrast = 1.e5;
    c->tau_rad = 1.e5;
    c->ref_theta_b = pr.theta_b;
    c->ref_pb = pr.pb;
    c->ref_a_over_kappa = pr.a_over_kappa;
  }
  if (want_sounding) {
    /* The sounding: 48 pressures log-spaced from 2 pbot up to ptop/2 (so
     * pdat falls and -log(p) rises along the table), at the temperature of
     * the initial profile (theta_profile at the equator times
     * (p/press0)^kappa) shifted linearly in ln p from +2 K at the bottom to
     * -2 K at the top: the relaxation warms the deep layers and cools the
     * upper ones.  The cooling time: 16 pressures over the same range,
     * 2e6 s (p/pbot)^(1/2), i.e. 2e6 s at pbot and faster aloft. */
    const double p_hi = 2. * pbot, p_lo = .5 * ptop, lhi = log(p_hi), llo = log(p_lo);
    c->ntp = 48;
    c->pdat = zalloc(c->ntp);
    c->tdat = zalloc(c->ntp);
    for (int i = 0; i < c->ntp; i++) {
      const double lp = lhi + (double)i / (double)(c->ntp - 1) * (llo - lhi), p = exp(lp);
      c->pdat[i] = p;
      c->tdat[i] = theta_profile(&pr, 0., p) * pow(p / pr.p0, pr.kappa) + (4. * (lp - llo) / (lhi - llo) - 2.);
    }
    c->n_t_cool = 16;
    c->t_cool_x = zalloc(c->n_t_cool);
    c->t_cool_y = zalloc(c->n_t_cool);
    for (int i = 0; i < c->n_t_cool; i++) {
      const double lp = lhi + (double)i / (double)(c->n_t_cool - 1) * (llo - lhi), p = exp(lp);
      c->t_cool_x[i] = -log(p);
      c->t_cool_y[i] = 2.e6 * sqrt(p / pbot);
    }
  }
  /* theta_ref(p): cos-weighted meridional mean of the profile */
  double *pref = c->p_ref, *tref = c->theta_ref;
  pref[1] = ptop;
  pref[2 * nk + 1] = pbot;
  for (int K = 1; K < nk; K++) {
    double neglogp = -log(ptop) + (double)K / (double)nk * (-log(pbot) + log(ptop));
    pref[2 * K + 1] = exp(-neglogp);
  }
  for (int K = 1; K <= nk; K++) pref[2 * K] = layer_p(c->kappa, pref[2 * K - 1], pref[2 * K + 1]);
  for (int kk = 1; kk <= 2 * nk + 1; kk++) {
    double sw = 0., st = 0.;
    for (int q = 0; q < 90; q++) {
      double latq = -89. + 2. * q;
      double w = cos(latq * DEGR);
      sw += w;
      st += w * theta_profile(&pr, latq, pref[kk]);
    }
    tref[kk] = st / sw;
  }

  /* ---- hybrid coordinate: k_sigma, zeta0/zeta1 (tangent line through the
   * knee at the top of layer k_sigma, as the generic branch of `initial`
   * does, src/shared/epic_initial.c:1695-1740), sigmatheta[kk] ---- */
  int k_sigma = nk - nk / 3;
  /* The hot Jupiter's day-night heating would move whole isentropic layers
   * across one another within ~100 steps at nk = 64; its levels are sigma
   * levels below the top layer, where heating changes theta in place. */
  if (s->planet == EPIC_PLANET_HOT_JUPITER) k_sigma = 2;
  if (k_sigma < 2) k_sigma = 2;
  c->k_sigma = k_sigma;
  hyb h;
  double tmin = tref[1];
  for (int kk = 1; kk <= 2 * nk + 1; kk++)
    if (tref[kk] < tmin) tmin = tref[kk];
  h.zeta0 = tmin * 0.7;
  double p_knee = pref[2 * (k_sigma - 1) + 1];
  double s_knee = hyb_get_sigma(pbot, p_knee, ptop);
  h.zeta1 = (tref[2 * (k_sigma - 1) + 1] - h.zeta0 * (1. - s_knee)) / s_knee;
  h.alpha = 50.;
  h.sigma_sigma = s_knee;
  c->sigmatheta[0] = c->sigmatheta[1] = tref[1];
  c->sigmatheta[2 * nk + 1] = hyb_sgth(&h, tref[2 * nk + 1], pbot, pbot, ptop);
  if (gas) c->sigmatheta[2 * (nk + 1)] = c->sigmatheta[2 * nk + 1];
  for (int kk = 2; kk < 2 * nk + 1; kk++) c->sigmatheta[kk] = hyb_sgth(&h, tref[kk], pref[kk], pbot, ptop);
  h.sigma_sigma = (c->sigmatheta[2 * k_sigma - 1] - h.zeta0) / (h.zeta1 - h.zeta0);
  c->zeta0 = h.zeta0;
  c->zeta1 = h.zeta1;
  c->hybrid_alpha = h.alpha;
  c->sigma_sigma = h.sigma_sigma;
  if (build_dsgth(c)) return -3;
  {
    double tmp = 0.02 / c->g[2 * c->jtp + 1];
    for (int K = 1; K <= nk; K++)
      c->h_min[K] = tmp * (pref[2 * K + 1] - pref[2 * (K - 1) + 1]) * c->dsgth_inv[2 * K];
    c->h_min[0] = c->h_min[1];
  }

  /* ---- viscosity: set_max_nu (src/shared/epic_subgrid.c:82-158, AB3 branch) ---- */
  {
    double dt = (double)c->dt, dy2 = c->dy0 * c->dy0;
    double max_nu4 = (dy2 / 240.) * (dy2 / dt);
    double max_nu6 = (dy2 / 800.) * (dy2 / dt) * dy2;
    double max_nu8 = (dy2 / 2400.) * (dy2 / dt) * dy2 * dy2;
    for (int q = 0; q <= EPIC_MAX_NU_ORDER; q++) c->nu[q] = c->nu_nondim[q] = 0.;
    c->nu_nondim[4] = s->nu4_nondim;
    c->nu[4] = s->nu4_nondim * max_nu4;
    c->nu_nondim[6] = s->nu6_nondim;
    c->nu[6] = s->nu6_nondim * max_nu6;
    c->nu_nondim[8] = s->nu8_nondim;
    c->nu[8] = s->nu8_nondim * max_nu8;
    c->nu[2] = 0.; /* the hyperviscosity initialisers force nu[2] = 0 */
  }

  /* ---- initial state ---- */
  const int Ia = c->Iadim, N2 = c->Nelem2d;
#define IX(K, J, I) ((size_t)(I) + (size_t)((J) + 1) * Ia + (size_t)(K) * N2)
  double *col = zalloc(nk + 1);
  for (int J = 0; J <= nj; J++) {
    double latd = c->lat[2 * J + 1];
    double gJ = c->g[2 * J + 1];
    /* interface pressures of this latitude */
    col[0] = ptop;
    col[nk] = pbot;
    double psig = hyb_p_sigma(pbot, h.sigma_sigma, ptop);
    for (int K = k_sigma - 1; K < nk; K++) {
      double sigma = (c->sigmatheta[2 * K + 1] - h.zeta0) / (h.zeta1 - h.zeta0);
      col[K] = hyb_p_sigma(pbot, sigma, ptop);
    }
    for (int K = 1; K < k_sigma - 1; K++)
      col[K] = solve_level(&h, &pr, latd, c->sigmatheta[2 * K + 1], pbot, ptop, ptop, psig);
    for (int K = 1; K <= nk; K++)
      if (!(col[K] > col[K - 1])) {
        free(col);
        return -4;
      }
    for (int I = 0; I <= ni + 1; I++) {
      c->p3[IX(0, J, I)] = ptop;
      for (int K = 1; K <= nk; K++) {
        c->p3[IX(K, J, I)] = col[K];
        c->hdry[IX(K, J, I)] = (col[K] - col[K - 1]) / (gJ * c->dsgth[2 * K]);
      }
      for (int K = 0; K <= nk; K++) c->theta[IX(K, J, I)] = theta_profile(&pr, latd, col[K]);
    }
  }
  free(col);
  /* zonal jets of the gas giants: u = sum_n A_n cos(n pi lat / 90), A_n ~ U(-60,60) */
  if (gas && s->planet != EPIC_PLANET_HOT_JUPITER) {
    double A[9];
    for (int q = 1; q <= 8; q++) A[q] = -60. + 120. * uniform01(s->seed, 7, (uint64_t)q);
    for (int J = 0; J <= nj; J++) {
      double latd = c->lat[2 * J + 1], uu = 0.;
      for (int q = 1; q <= 8; q++) uu += A[q] * cos(q * M_PI * latd / 90.);
      for (int K = 0; K <= nk; K++)
        for (int I = 0; I <= ni + 1; I++) c->u[IX(K, J, I)] = uu;
    }
  }
  /* seeded noise on the prognostic interior, then zonal periodicity */
  for (int K = 0; K <= nk; K++)
    for (int J = 0; J <= nj; J++)
      for (int I = 1; I <= ni; I++) {
        uint64_t idx = IX(K, J, I);
        if (K >= k_sigma - 1 && s->noise_theta > 0.) c->theta[idx] += s->noise_theta * gauss(s->seed, 1, idx);
        if (K >= 1 && K < nk && s->noise_uv > 0.) {
          c->u[idx] += s->noise_uv * gauss(s->seed, 2, idx);
          if (J >= 1) c->v[idx] += s->noise_uv * gauss(s->seed, 3, idx);
        }
      }
  double *per[] = {c->u, c->v, c->theta, c->hdry, c->p3};
  for (int q = 0; q < 5; q++)
    for (int K = 0; K <= nk + 1; K++)
      for (int J = -1; J <= nj + 1; J++) {
        per[q][IX(K, J, 0)] = per[q][IX(K, J, ni)];
        per[q][IX(K, J, ni + 1)] = per[q][IX(K, J, 1)];
      }
#undef IX
  return 0;
}
