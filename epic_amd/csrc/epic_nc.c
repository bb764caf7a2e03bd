/*
 * epic_nc.c -- the model state of one member as the reference's epic.nc
 * file: the netCDF-3 layout that var_write writes and var_read reads for
 * ALL_DATA (src/shared/epic_funcs_io.c:730-1368 var_read, :1369-2257
 * var_write, :2643-3300 define_netcdf; variable names, units and standard
 * names from set_var_props, src/shared/epic_funcs_diag.c:109-275):
 *
 *   global attributes  the planet, grid and run parameters, one attribute
 *                      per value (the planet_*, grid_*, var_* names of the
 *                      READ and WRITE macros, include/epic_io_macros.h);
 *                      floating values as double, times as int
 *   dimensions         time (record); lon_G (ni), lat_G (nj+1) and
 *                      sigmatheta_G (nk+1) for each grid G of pv3, u, v,
 *                      hdry, p3
 *   coordinates        time [days since the start time], L_s, and lon_G,
 *                      lat_G, sigmatheta_G (the staggered C-grid positions)
 *   record variables   u (u-grid), v (v-grid), p3, theta (p3-grid),
 *                      u_spinup (u-grid): (time, K = 0..nk, J = 0..nj,
 *                      I = 1..ni)
 *   fixed variables    du_IT_MINUS1dt, du_IT_MINUS2dt (u-grid),
 *                      dv_IT_MINUS1dt, dv_IT_MINUS2dt (v-grid): the two
 *                      older Adams-Bashforth tendency planes, (K, J, I);
 *                      gz_surface (J, I; terrestrial planets only)
 *
 * Reading does what var_read and epic_main do with the file: the metric
 * arrays formed again from the scalars (set_lonlat, set_fmn, set_gravity,
 * set_dsgth, set_sponge), the sounding pressures from hPa, the cooling
 * table's -log p, u_spinup's top plane copied from the one below, zonal
 * periodicity applied (bc_lateral); HDRY is formed from P3 when the case is
 * added to an engine, as epic_main forms it.
 */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

#include "../../include/epic_amd.h"
#include "case_internal.h"
#include "ncio.h"

#define N_STR 256
#define GEOM_STR 16
#define TOPDIM 3
#define N_ON_LIST 54 /* LAST_INDEX - FIRST_INDEX + 1 */
#define DATA_VERSION 4.30
#define DBL_FLAG 1.7976931348623157e308

static char g_err[256];
const char *epic_nc_error(void) { return g_err; }
static int fail(int rc, const char *fmt, const char *what) {
  snprintf(g_err, sizeof(g_err), fmt, what);
  return rc;
}

void epic_nc_header_free(void *h) {
  if (!h) return;
  ncio_free((ncio_file *)h);
  free(h);
}

/* ------------------------------------------------------------- writing */
static void w_text(ncio_file *f, const char *name, const char *s, size_t width) {
  // fixed-width character arrays (WRITEC's num) padded with NULs
  char *b = (char *)calloc(width + 1, 1);
  strncpy(b, s, width);
  ncio_put_text(f, -1, name, b, width);
  free(b);
}
static void w_str(ncio_file *f, int varid, const char *name, const char *s) {
  ncio_put_text(f, varid, name, s, strlen(s) + 1);  // strlen + 1, as the reference writes
}
static void w_d(ncio_file *f, const char *name, size_t n, const double *v) {
  ncio_put_att(f, -1, name, NCIO_DOUBLE, n, v);
}
static void w_d1(ncio_file *f, const char *name, double v) { w_d(f, name, 1, &v); }
static void w_i(ncio_file *f, const char *name, size_t n, const int *v) { ncio_put_att(f, -1, name, NCIO_INT, n, v); }
static void w_i1(ncio_file *f, const char *name, int v) { w_i(f, name, 1, &v); }

/* a global of the stash (the case's file of origin) or a default */
static double stash_d(const epic_case *c, const char *name, double dflt) {
  const ncio_file *h = (const ncio_file *)c->nc_header;
  const ncio_att *a = h ? ncio_att_of(h, -1, name) : NULL;
  if (a && a->n >= 1 && a->type == NCIO_DOUBLE) return ((const double *)a->val)[0];
  if (a && a->n >= 1 && a->type == NCIO_FLOAT) return ((const float *)a->val)[0];
  return dflt;
}
static int stash_i(const epic_case *c, const char *name, int dflt) {
  const ncio_file *h = (const ncio_file *)c->nc_header;
  const ncio_att *a = h ? ncio_att_of(h, -1, name) : NULL;
  return (a && a->n >= 1 && a->type == NCIO_INT) ? ((const int *)a->val)[0] : dflt;
}

/* K = 0..nk, J = 0..nj, I = 1..ni of a padded array, K outermost */
static void gather3(const epic_case *c, const double *a, double *out) {
  const int ni = c->ni, nj = c->nj, nk = c->nk;
  size_t o = 0;
  for (int K = 0; K <= nk; K++)
    for (int J = 0; J <= nj; J++)
      for (int I = 1; I <= ni; I++)
        out[o++] = a[(size_t)I + (size_t)(J + 1) * c->Iadim + (size_t)K * c->Nelem2d];
}

int epic_nc_write(const char *path, const epic_case *c, const epic_nc_state *st) {
  const int ni = c->ni, nj = c->nj, nk = c->nk;
  const size_t n3 = (size_t)(nk + 1) * (nj + 1) * ni;
  // classic (32-bit offsets) where it fits, as the reference creates it;
  // 64-bit offsets beyond
  const size_t approx = n3 * 8 * 10 + 65536;
  ncio_file f;
  ncio_init(&f, approx > 0x7fff0000u ? 2 : 1);
  f.numrecs = 1;
  const int terrestrial = strcmp(c->planet_type, "terrestrial") == 0;
  // global attributes of the file of origin first: those without a value
  // in the case pass through
  const ncio_file *h = (const ncio_file *)c->nc_header;
  if (h)
    for (int i = 0; i < h->ngatts; i++) ncio_put_att(&f, -1, h->gatts[i].name, h->gatts[i].type, h->gatts[i].n, h->gatts[i].val);

  // time and L_s
  const int d_time = ncio_def_dim(&f, "time", 0);
  const int v_time = ncio_def_var(&f, "time", NCIO_DOUBLE, 1, &d_time);
  {
    char units[96];
    time_t t0 = (time_t)c->start_time;
    struct tm g;
    gmtime_r(&t0, &g);
    strftime(units, sizeof(units), "days since %Y-%m-%d %H:%M:%S 0", &g);
    w_str(&f, v_time, "units", units);
    w_str(&f, v_time, "standard_name", "time");
    w_str(&f, v_time, "long_name", "elapsed time");
    w_str(&f, v_time, "calendar", "julian");
  }
  const int v_ls = ncio_def_var(&f, "L_s", NCIO_DOUBLE, 1, &d_time);
  w_str(&f, v_ls, "long_name", "solar longitude (planetocentric)");
  w_str(&f, v_ls, "units", "degrees");

  // the staggered grids: pv3, u, v, hdry, p3
  const char *gname[5] = {"pv3", "u", "v", "hdry", "p3"};
  int gdim[5][3], gvar[5][3];
  for (int g = 0; g < 5; g++) {
    char nm[64];
    snprintf(nm, sizeof(nm), "lon_%s", gname[g]);
    gdim[g][0] = ncio_def_dim(&f, nm, (size_t)ni);
    gvar[g][0] = ncio_def_var(&f, nm, NCIO_DOUBLE, 1, &gdim[g][0]);
    w_str(&f, gvar[g][0], "standard_name", "longitude");
    w_str(&f, gvar[g][0], "long_name", "longitude");
    w_str(&f, gvar[g][0], "units", "degrees_east");
    snprintf(nm, sizeof(nm), "lat_%s", gname[g]);
    gdim[g][1] = ncio_def_dim(&f, nm, (size_t)nj + 1);
    gvar[g][1] = ncio_def_var(&f, nm, NCIO_DOUBLE, 1, &gdim[g][1]);
    w_str(&f, gvar[g][1], "standard_name", "latitude");
    w_str(&f, gvar[g][1], "long_name", "latitude (planetographic)");
    w_str(&f, gvar[g][1], "units", "degrees_north");
    snprintf(nm, sizeof(nm), "sigmatheta_%s", gname[g]);
    gdim[g][2] = ncio_def_dim(&f, nm, (size_t)nk + 1);
    gvar[g][2] = ncio_def_var(&f, nm, NCIO_DOUBLE, 1, &gdim[g][2]);
    w_str(&f, gvar[g][2], "long_name", "hybrid sigma-theta vertical coordinate");
    w_str(&f, gvar[g][2], "positive", "up");
    w_str(&f, gvar[g][2], "units", "K");
  }
  // record variable dims (time, sigmatheta, lat, lon) of grid g; the
  // tendencies take the last three
  int rd[5][4];
  for (int g = 0; g < 5; g++) {
    rd[g][0] = d_time;
    rd[g][1] = gdim[g][2];
    rd[g][2] = gdim[g][1];
    rd[g][3] = gdim[g][0];
  }
  struct {
    const char *name, *std, *lng, *units;
    int grid;
  } prog[2] = {{"u", "eastward_wind", "zonal wind", "m/s", 1}, {"v", "northward_wind", "meridional wind", "m/s", 2}};
  int v_prog[2], v_tend[2][2];
  for (int q = 0; q < 2; q++) {
    v_prog[q] = ncio_def_var(&f, prog[q].name, NCIO_DOUBLE, 4, rd[prog[q].grid]);
    w_str(&f, v_prog[q], "standard_name", prog[q].std);
    w_str(&f, v_prog[q], "long_name", prog[q].lng);
    w_str(&f, v_prog[q], "units", prog[q].units);
    for (int p = 0; p < 2; p++) {
      char nm[64], s1[96], s2[96];
      snprintf(nm, sizeof(nm), "d%s_IT_MINUS%ddt", prog[q].name, p + 1);
      v_tend[q][p] = ncio_def_var(&f, nm, NCIO_DOUBLE, 3, rd[prog[q].grid] + 1);
      w_str(&f, v_tend[q][p], "units", "m/s/s");
      snprintf(s1, sizeof(s1), "tendency_of_%s", prog[q].std);
      snprintf(s2, sizeof(s2), "tendency of %s", prog[q].lng);
      w_str(&f, v_tend[q][p], "standard_name", s1);
      w_str(&f, v_tend[q][p], "long_name", s2);
    }
  }
  const int v_p3 = ncio_def_var(&f, "p3", NCIO_DOUBLE, 4, rd[4]);
  w_str(&f, v_p3, "standard_name", "air_pressure");
  w_str(&f, v_p3, "long_name", "total air pressure (interface)");
  w_str(&f, v_p3, "units", "Pa");
  const int v_th = ncio_def_var(&f, "theta", NCIO_DOUBLE, 4, rd[4]);
  w_str(&f, v_th, "standard_name", "air_potential_temperature");
  w_str(&f, v_th, "long_name", "potential temperature");
  w_str(&f, v_th, "units", "K");
  const int v_us = ncio_def_var(&f, "u_spinup", NCIO_DOUBLE, 4, rd[1]);
  w_str(&f, v_us, "long_name", "Rayleigh-drag zonal wind profile");
  w_str(&f, v_us, "units", "m/s");
  int v_gzs = -1;
  if (terrestrial) {
    int d2[2] = {gdim[4][1], gdim[4][0]};
    v_gzs = ncio_def_var(&f, "gz_surface", NCIO_DOUBLE, 2, d2);
    w_str(&f, v_gzs, "standard_name", "surface_geopotential");
    w_str(&f, v_gzs, "long_name", "surface geopotential");
    w_str(&f, v_gzs, "units", "m^2/s^2");
  }

  // ---- global attributes (var_write's order) ----
  char msg[N_STR];
  w_str(&f, -1, "Conventions", "CF-1.0");
  snprintf(msg, sizeof(msg),
           "Explicit Planetary Isentropic-Coordinate (EPIC) Atmospheric Model, Version %4.2f: runable data file",
           DATA_VERSION);
  w_str(&f, -1, "title", msg);
  snprintf(msg, sizeof(msg), "EPIC Model Version %4.2f: model state (prognostic variables and tendencies)",
           DATA_VERSION);
  w_str(&f, -1, "source", msg);
  w_str(&f, -1, "institution", "Developed at the Comparative Planetology Laboratory (CPL), Louisville, KY");
  w_str(&f, -1, "history", "This file produced by the EPIC Atmospheric model");
  w_str(&f, -1, "references",
        "The EPIC Model is downloadable as open source from the NASA Planetary Data System (PDS) Atmospheres Node\n");
  w_d1(&f, "grid_data_version", stash_d(c, "grid_data_version", DATA_VERSION));
  w_i1(&f, "var_start_time", (int)c->start_time);
  w_i1(&f, "planet_index", c->planet_index);
  w_text(&f, "planet_name", c->planet_name, 32);
  w_text(&f, "planet_type", c->planet_type, 16);
  {
    const ncio_att *a = h ? ncio_att_of(h, -1, "planet_orbital_epoch") : NULL;
    if (!a) w_text(&f, "planet_orbital_epoch", "J2000", 8);
  }
  w_d1(&f, "planet_re", c->re);
  w_d1(&f, "planet_rp", c->rp);
  w_d1(&f, "planet_omega_sidereal", c->omega);
  w_d1(&f, "planet_omega_synodic", stash_d(c, "planet_omega_synodic", c->omega));
  w_d1(&f, "planet_cp", c->cp);
  w_d1(&f, "planet_rgas", c->rgas);
  w_d1(&f, "planet_GM", c->GM);
  w_d1(&f, "planet_J2", c->J2);
  w_d1(&f, "planet_x_he", stash_d(c, "planet_x_he", 0.));
  w_d1(&f, "planet_x_h2", c->x_h2);
  w_d1(&f, "planet_x_3", stash_d(c, "planet_x_3", 1.));
  // orbital elements: the file's, else a circular orbit at 1 AU
  w_d1(&f, "planet_a", stash_d(c, "planet_a", 1.));
  w_d1(&f, "planet_e", stash_d(c, "planet_e", 0.));
  w_d1(&f, "planet_i", stash_d(c, "planet_i", 0.));
  w_d1(&f, "planet_lon_ascending_node", stash_d(c, "planet_lon_ascending_node", 0.));
  w_d1(&f, "planet_lon_perihelion", stash_d(c, "planet_lon_perihelion", 0.));
  w_d1(&f, "planet_mean_lon", stash_d(c, "planet_mean_lon", 0.));
  w_d1(&f, "planet_orbit_period", stash_d(c, "planet_orbit_period", 1.));
  w_d1(&f, "planet_vernal_equinox_anomaly", stash_d(c, "planet_vernal_equinox_anomaly", 0.));
  w_d1(&f, "planet_kinvisc", c->kinvisc);
  w_d1(&f, "planet_dynvisc", stash_d(c, "planet_dynvisc", 0.));
  w_d1(&f, "planet_k_a", stash_d(c, "planet_k_a", 0.));
  w_i1(&f, "grid_nk", nk);
  w_i1(&f, "grid_nj", nj);
  w_i1(&f, "grid_ni", ni);
  w_i1(&f, "grid_jtp", c->jtp);
  w_text(&f, "grid_geometry", "globe", GEOM_STR);
  w_text(&f, "grid_uv_timestep_scheme", "3rd-order Adams-Bashforth", N_STR);
  w_text(&f, "grid_turbulence_scheme", "off", N_STR);
  {
    const ncio_att *a = h ? ncio_att_of(h, -1, "grid_stability_factor") : NULL;
    if (!a) w_text(&f, "grid_stability_factor", "off", N_STR);
  }
  {
    // mass scheme: hdry and the species; non-mass: theta, fpara, nu_turb
    // (as the reference's initial program assigns them)
    const char *mass = c->adv_mass ? EPIC_ADV_AKIMA : EPIC_ADV_HSU, *nonmass = c->adv_nonmass ? EPIC_ADV_AKIMA : EPIC_ADV_HSU;
    w_text(&f, "var_hdry_advection_scheme", mass, N_STR);
    w_text(&f, "var_theta_advection_scheme", nonmass, N_STR);
    const char *other[3] = {"var_fpara_advection_scheme", "var_nu_turb_advection_scheme",
                            "var_species_advection_scheme"};
    for (int q = 0; q < 3; q++)
      if (!(h && ncio_att_of(h, -1, other[q]))) w_text(&f, other[q], q == 2 ? mass : nonmass, N_STR);
  }
  w_d1(&f, "grid_globe_lonbot", c->globe_lonbot);
  w_d1(&f, "grid_globe_lontop", c->globe_lontop);
  w_d1(&f, "grid_globe_latbot", c->globe_latbot);
  w_d1(&f, "grid_globe_lattop", c->globe_lattop);
  if (!(h && ncio_att_of(h, -1, "grid_f_plane_map"))) w_text(&f, "grid_f_plane_map", "none", GEOM_STR);
  w_d1(&f, "grid_f_plane_lat0", stash_d(c, "grid_f_plane_lat0", 0.));
  w_d1(&f, "grid_f_plane_half_width", stash_d(c, "grid_f_plane_half_width", 0.));
  {
    const int wrap[TOPDIM] = {1, 0, 0}, pad[TOPDIM] = {1, 1, 1};
    w_i(&f, "grid_wrap", TOPDIM, wrap);
    w_i(&f, "grid_pad", TOPDIM, pad);
  }
  w_i1(&f, "grid_jlo", 0);
  w_i1(&f, "grid_jfirst", 1);
  w_i1(&f, "grid_jlast", nj - 1);
  w_i1(&f, "grid_k_sponge", c->k_sponge);
  w_i1(&f, "grid_k_sigma", c->k_sigma);
  w_d1(&f, "grid_du_vert", stash_d(c, "grid_du_vert", 0.));
  w_i1(&f, "grid_newt_cool_on", c->newt_cool_on);
  w_i1(&f, "grid_microphysics_on", 0);
  w_i1(&f, "grid_nmt_physics_on", 0);
  {
    // u, v, hdry and theta on; nothing else carried
    int on[N_ON_LIST], ex[N_ON_LIST];
    for (int q = 0; q < N_ON_LIST; q++) on[q] = ex[q] = 0;
    on[0] = on[1] = on[2] = on[3] = 1;
    w_i(&f, "var_on_list", N_ON_LIST, on);
    w_i1(&f, "var_extract_on", 0);
    w_i(&f, "var_extract_on_list", N_ON_LIST, ex);
  }
  w_i1(&f, "var_ntp", c->ntp);
  w_i1(&f, "var_n_t_cool", c->n_t_cool);
  w_d1(&f, "grid_dlt", c->dlt);
  w_d1(&f, "grid_dln", c->dln);
  w_i1(&f, "grid_dt", c->dt);
  w_d1(&f, "var_time_fp_bar", stash_d(c, "var_time_fp_bar", 3.e8 * 1000. * 100.));
  if (c->ntp > 0) {
    double *b = (double *)malloc(sizeof(double) * c->ntp);
    for (int i = 0; i < c->ntp; i++) b[i] = c->pdat[i] / 100.;  // Pa to hPa
    w_d(&f, "var_pdat", (size_t)c->ntp, b);
    w_str(&f, -1, "var_pdat_units", "hPa");
    w_d(&f, "var_tdat", (size_t)c->ntp, c->tdat);
    w_str(&f, -1, "var_tdat_units", "K");
    // dtdat: the file's, else zeros (nothing in the dry step reads it)
    const ncio_att *a = h ? ncio_att_of(h, -1, "var_dtdat") : NULL;
    if (!(a && a->n == (size_t)c->ntp)) {
      for (int i = 0; i < c->ntp; i++) b[i] = 0.;
      w_d(&f, "var_dtdat", (size_t)c->ntp, b);
    }
    w_str(&f, -1, "var_dtdat_units", "K");
    free(b);
  }
  if (c->n_t_cool > 0) {
    double *b = (double *)malloc(sizeof(double) * c->n_t_cool);
    for (int i = 0; i < c->n_t_cool; i++) b[i] = exp(-c->t_cool_x[i]) / 100.;
    w_d(&f, "var_t_cool_p", (size_t)c->n_t_cool, b);
    w_str(&f, -1, "var_t_cool_p_units", "hPa");
    w_d(&f, "var_t_cool_tau", (size_t)c->n_t_cool, c->t_cool_y);
    w_str(&f, -1, "var_t_cool_tau_units", "s");
    free(b);
  }
  const int nK = 2 * (nk + 1) + 1;
  w_d(&f, "grid_sigmatheta", (size_t)nK, c->sigmatheta);
  w_d(&f, "grid_p_ref", (size_t)nK, c->p_ref);
  w_d(&f, "grid_theta_ref", (size_t)nK, c->theta_ref);
  w_d(&f, "grid_h_min", (size_t)nk + 1, c->h_min);
  w_d1(&f, "grid_sgth_bot", stash_d(c, "grid_sgth_bot", c->sigmatheta[2 * nk + 1]));
  w_d1(&f, "grid_sgth_top", stash_d(c, "grid_sgth_top", c->sigmatheta[1]));
  w_d1(&f, "grid_zeta0", c->zeta0);
  w_d1(&f, "grid_zeta1", c->zeta1);
  w_d1(&f, "grid_hybrid_alpha", c->hybrid_alpha);
  w_d1(&f, "grid_sigma_sigma", c->sigma_sigma);
  w_i1(&f, "grid_newt_cool_adjust", c->newt_cool_adjust);
  w_i1(&f, "grid_diffusion_direction", stash_i(c, "grid_diffusion_direction", 0));
  w_text(&f, "grid_eos", "ideal", 8);
  w_d1(&f, "grid_prandtl", c->prandtl);
  w_d1(&f, "grid_tau_drag", c->tau_drag);
  w_i1(&f, "grid_drag_v", c->drag_v);
  w_i1(&f, "grid_drag_zonal_avg", c->drag_zonal_avg);
  w_d1(&f, "grid_pbot", c->pbot);
  w_d1(&f, "grid_press0", c->press0);
  w_d1(&f, "grid_mont0", stash_d(c, "grid_mont0", 0.));
  w_d1(&f, "grid_topo_scale", c->topo_scale);
  w_d1(&f, "grid_nudiv_nondim", c->nudiv_nondim);
  w_d(&f, "grid_nu_nondim", EPIC_MAX_NU_ORDER + 1, c->nu_nondim);
  w_d(&f, "grid_nu", EPIC_MAX_NU_ORDER + 1, c->nu);
  w_i1(&f, "var_model_time", (int)st->model_time);
  w_i1(&f, "grid_cfl_dt", stash_i(c, "grid_cfl_dt", c->dt));
  w_i1(&f, "grid_aux_a", stash_i(c, "grid_aux_a", 0));
  w_i1(&f, "grid_aux_b", stash_i(c, "grid_aux_b", 0));
  w_i1(&f, "grid_aux_c", stash_i(c, "grid_aux_c", 0));
  w_d1(&f, "grid_aux_fa", stash_d(c, "grid_aux_fa", 0.));
  w_d1(&f, "grid_aux_fb", stash_d(c, "grid_aux_fb", 0.));
  w_d1(&f, "grid_aux_fc", stash_d(c, "grid_aux_fc", 0.));

  // ---- values ----
  double *b = (double *)malloc(sizeof(double) * (n3 + 2 * (size_t)(ni + nj + nk) + 8));
  {
    const double tday = (st->model_time - c->start_time) / 86400.;
    ncio_put_var(&f, v_time, &tday);
    ncio_put_var(&f, v_ls, &st->l_s_deg);
  }
  for (int g = 0; g < 5; g++) {
    const int pvlike = (g == 0 || g == 1);  // u and pv3 on lon[2I]; v, hdry, p3 on lon[2I+1]
    for (int I = 1; I <= ni; I++) b[I - 1] = c->lon[2 * I + (pvlike ? 0 : 1)];
    ncio_put_var(&f, gvar[g][0], b);
    const int vlat = (g == 0 || g == 2);  // pv3 and v on lat[2J]
    for (int J = 0; J <= nj; J++) b[J] = c->lat[2 * J + (vlat ? 0 : 1)];
    ncio_put_var(&f, gvar[g][1], b);
    for (int K = 0; K <= nk; K++) b[K] = c->sigmatheta[2 * K + (g == 3 ? 0 : 1)];
    ncio_put_var(&f, gvar[g][2], b);
  }
  const double *pv[5] = {st->u, st->v, st->p3, st->theta, st->u_spinup};
  const int vid[5] = {v_prog[0], v_prog[1], v_p3, v_th, v_us};
  for (int q = 0; q < 5; q++) {
    gather3(c, pv[q], b);
    ncio_put_var(&f, vid[q], b);
  }
  const double *td[2][2] = {{st->dudt_m1, st->dudt_m2}, {st->dvdt_m1, st->dvdt_m2}};
  for (int q = 0; q < 2; q++)
    for (int p = 0; p < 2; p++) {
      if (td[q][p]) {
        gather3(c, td[q][p], b);
      } else {
        for (size_t i = 0; i < n3; i++) b[i] = DBL_FLAG;
      }
      ncio_put_var(&f, v_tend[q][p], b);
    }
  if (v_gzs >= 0) {
    size_t o = 0;
    for (int J = 0; J <= nj; J++)
      for (int I = 1; I <= ni; I++) b[o++] = st->gz_surface[(size_t)I + (size_t)(J + 1) * c->Iadim];
    ncio_put_var(&f, v_gzs, b);
  }
  free(b);
  int rc = ncio_write(&f, path);
  ncio_free(&f);
  return rc ? fail(-1, "cannot write %s", path) : 0;
}

/* ------------------------------------------------------------- reading */
typedef struct {
  const ncio_file *f;
  int bad;
  char what[96];
} rd;

static const ncio_att *r_att(rd *r, const char *name, size_t n) {
  const ncio_att *a = ncio_att_of(r->f, -1, name);
  if (!a || (n && a->n < n)) {
    if (!r->bad) snprintf(r->what, sizeof(r->what), "global attribute %s", name);
    r->bad = 1;
    return NULL;
  }
  return a;
}
/* floating values: double, or float from a single-precision build */
static void r_dn(rd *r, const char *name, size_t n, double *out) {
  const ncio_att *a = r_att(r, name, n);
  if (!a) return;
  for (size_t i = 0; i < n; i++) {
    if (a->type == NCIO_DOUBLE) out[i] = ((const double *)a->val)[i];
    else if (a->type == NCIO_FLOAT) out[i] = ((const float *)a->val)[i];
    else {
      if (!r->bad) snprintf(r->what, sizeof(r->what), "type of global attribute %s", name);
      r->bad = 1;
      return;
    }
  }
}
static double r_d(rd *r, const char *name) {
  double v = 0.;
  r_dn(r, name, 1, &v);
  return v;
}
static int r_i(rd *r, const char *name) {
  const ncio_att *a = r_att(r, name, 1);
  if (!a) return 0;
  if (a->type != NCIO_INT) {
    if (!r->bad) snprintf(r->what, sizeof(r->what), "type of global attribute %s", name);
    r->bad = 1;
    return 0;
  }
  return ((const int *)a->val)[0];
}
static void r_text(rd *r, const char *name, char *out, size_t cap) {
  const ncio_att *a = r_att(r, name, 0);
  memset(out, 0, cap);
  if (!a) return;
  if (a->type != NCIO_CHAR) {
    if (!r->bad) snprintf(r->what, sizeof(r->what), "type of global attribute %s", name);
    r->bad = 1;
    return;
  }
  size_t n = a->n < cap - 1 ? a->n : cap - 1;
  memcpy(out, a->val, n);
}

/* a variable's values as doubles, record 0 of a record variable; n values */
static double *r_var(rd *r, const char *name, size_t n) {
  const int id = ncio_var_id(r->f, name);
  if (id < 0) {
    if (!r->bad) snprintf(r->what, sizeof(r->what), "variable %s", name);
    r->bad = 1;
    return NULL;
  }
  int rec;
  const size_t cnt = ncio_var_count(r->f, id, &rec);
  const ncio_var *v = &r->f->vars[id];
  if (cnt != n || (rec && r->f->numrecs < 1) || (v->type != NCIO_DOUBLE && v->type != NCIO_FLOAT)) {
    if (!r->bad) snprintf(r->what, sizeof(r->what), "shape or type of variable %s", name);
    r->bad = 1;
    return NULL;
  }
  double *out = (double *)malloc(sizeof(double) * (n ? n : 1));
  for (size_t i = 0; i < n; i++)
    out[i] = (v->type == NCIO_DOUBLE) ? ((const double *)v->data)[i] : ((const float *)v->data)[i];
  return out;
}

/* the file's K = 0..nk, J = 0..nj, I = 1..ni values into a padded array, and
 * zonal periodicity over every plane and row (bc_lateral) */
static void scatter3(const epic_case *c, const double *in, double *a) {
  const int ni = c->ni, nj = c->nj, nk = c->nk;
  size_t o = 0;
  for (int K = 0; K <= nk; K++)
    for (int J = 0; J <= nj; J++)
      for (int I = 1; I <= ni; I++) a[(size_t)I + (size_t)(J + 1) * c->Iadim + (size_t)K * c->Nelem2d] = in[o++];
  for (int K = 0; K <= nk + 1; K++)
    for (int J = -1; J <= nj + 1; J++) {
      const size_t row = (size_t)(J + 1) * c->Iadim + (size_t)K * c->Nelem2d;
      a[row] = a[row + ni];
      a[row + ni + 1] = a[row + 1];
    }
}

static int text_is(const char *s, const char *want) { return strcmp(s, want) == 0; }

int epic_nc_read(const char *path, epic_case *c) {
  memset(c, 0, sizeof(*c));
  g_err[0] = 0;
  ncio_file *f = (ncio_file *)calloc(1, sizeof(ncio_file));
  int rc = ncio_read(f, path);
  if (rc) {
    free(f);
    return fail(rc == -1 ? -1 : -2, rc == -1 ? "cannot read %s" : "%s is not a netCDF classic file", path);
  }
  rd r = {f, 0, ""};
  char s[N_STR + 1];
  // the configurations the GPU step implements
  const char *need_text[][2] = {{"grid_geometry", "globe"},
                                {"grid_uv_timestep_scheme", "3rd-order Adams-Bashforth"},
                                {"grid_turbulence_scheme", "off"},
                                {"grid_eos", "ideal"}};
  for (size_t q = 0; q < sizeof(need_text) / sizeof(need_text[0]); q++) {
    r_text(&r, need_text[q][0], s, sizeof(s));
    if (r.bad) break;
    if (!text_is(s, need_text[q][1])) {
      snprintf(g_err, sizeof(g_err), "%.40s = \"%.60s\" is not implemented (only \"%.60s\")", need_text[q][0], s,
               need_text[q][1]);
      epic_nc_header_free(f);
      return -4;
    }
  }
  // the two advection schemes the GPU step implements
  const char *adv_att[2] = {"var_hdry_advection_scheme", "var_theta_advection_scheme"};
  int adv[2] = {0, 0};
  for (int q = 0; q < 2 && !r.bad; q++) {
    r_text(&r, adv_att[q], s, sizeof(s));
    if (r.bad) break;
    if (text_is(s, EPIC_ADV_AKIMA)) adv[q] = 1;
    else if (!text_is(s, EPIC_ADV_HSU)) {
      snprintf(g_err, sizeof(g_err), "%.40s = \"%.60s\" is not implemented", adv_att[q], s);
      epic_nc_header_free(f);
      return -4;
    }
  }
  c->adv_mass = adv[0];
  c->adv_nonmass = adv[1];
  r_text(&r, "planet_name", c->planet_name, sizeof(c->planet_name));
  r_text(&r, "planet_type", c->planet_type, sizeof(c->planet_type));
  c->planet_index = r_i(&r, "planet_index");
  c->re = r_d(&r, "planet_re");
  c->rp = r_d(&r, "planet_rp");
  c->omega = r_d(&r, "planet_omega_sidereal");
  c->cp = r_d(&r, "planet_cp");
  c->rgas = r_d(&r, "planet_rgas");
  c->cpr = c->cp / c->rgas;  // as var_read forms them
  c->kappa = 1. / c->cpr;
  c->GM = r_d(&r, "planet_GM");
  c->J2 = r_d(&r, "planet_J2");
  c->x_h2 = r_d(&r, "planet_x_h2");
  c->kinvisc = r_d(&r, "planet_kinvisc");
  c->nk = r_i(&r, "grid_nk");
  c->nj = r_i(&r, "grid_nj");
  c->ni = r_i(&r, "grid_ni");
  c->jtp = r_i(&r, "grid_jtp");
  c->globe_lonbot = r_d(&r, "grid_globe_lonbot");
  c->globe_lontop = r_d(&r, "grid_globe_lontop");
  c->globe_latbot = r_d(&r, "grid_globe_latbot");
  c->globe_lattop = r_d(&r, "grid_globe_lattop");
  c->k_sponge = r_i(&r, "grid_k_sponge");
  c->k_sigma = r_i(&r, "grid_k_sigma");
  c->newt_cool_on = r_i(&r, "grid_newt_cool_on");
  const int micro = r_i(&r, "grid_microphysics_on"), nmt = r_i(&r, "grid_nmt_physics_on");
  const ncio_att *onl = r_att(&r, "var_on_list", N_ON_LIST);
  c->ntp = r_i(&r, "var_ntp");
  c->n_t_cool = r_i(&r, "var_n_t_cool");
  c->dlt = r_d(&r, "grid_dlt");
  c->dln = r_d(&r, "grid_dln");
  c->dt = r_i(&r, "grid_dt");
  c->newt_cool_adjust = r_i(&r, "grid_newt_cool_adjust");
  c->prandtl = r_d(&r, "grid_prandtl");
  c->tau_drag = r_d(&r, "grid_tau_drag");
  c->drag_v = r_i(&r, "grid_drag_v");
  c->drag_zonal_avg = r_i(&r, "grid_drag_zonal_avg");
  c->pbot = r_d(&r, "grid_pbot");
  c->press0 = r_d(&r, "grid_press0");
  c->topo_scale = r_d(&r, "grid_topo_scale");
  c->nudiv_nondim = r_d(&r, "grid_nudiv_nondim");
  r_dn(&r, "grid_nu_nondim", EPIC_MAX_NU_ORDER + 1, c->nu_nondim);
  r_dn(&r, "grid_nu", EPIC_MAX_NU_ORDER + 1, c->nu);
  c->zeta0 = r_d(&r, "grid_zeta0");
  c->zeta1 = r_d(&r, "grid_zeta1");
  c->hybrid_alpha = r_d(&r, "grid_hybrid_alpha");
  c->sigma_sigma = r_d(&r, "grid_sigma_sigma");
  c->model_time = r_i(&r, "var_model_time");
  c->start_time = r_i(&r, "var_start_time");
  if (r.bad) {
    snprintf(g_err, sizeof(g_err), "%s: missing or malformed %s", path, r.what);
    epic_nc_header_free(f);
    return -3;
  }
  if (onl && onl->type == NCIO_INT) {
    const int *on = (const int *)onl->val;
    int ok = on[0] == 1 && on[1] == 1 && on[2] == 1 && on[3] == 1;
    for (int q = 4; q < 16; q++) ok = ok && on[q] != 1;  // nu_turb, fpara, species
    if (!ok) {
      snprintf(g_err, sizeof(g_err), "var_on_list: only u, v, hdry and theta are implemented");
      epic_nc_header_free(f);
      return -4;
    }
  }
  if (micro || nmt || c->x_h2 != 0.) {
    snprintf(g_err, sizeof(g_err), "microphysics, NMT physics and x_h2 > 0 are not implemented");
    epic_nc_header_free(f);
    return -4;
  }
  if (c->ni < 1 || c->nj < 1 || c->nk < 1 || c->k_sponge < 0 || c->k_sponge > c->nk || c->dt <= 0 ||
      c->ntp < 0 || c->n_t_cool < 0) {
    snprintf(g_err, sizeof(g_err), "%s: grid sizes out of range", path);
    epic_nc_header_free(f);
    return -3;
  }

  epic_case_alloc(c);
  epic_case_metrics(c);
  const int ni = c->ni, nj = c->nj, nk = c->nk, nK = 2 * (nk + 1) + 1;
  r_dn(&r, "grid_sigmatheta", (size_t)nK, c->sigmatheta);
  r_dn(&r, "grid_p_ref", (size_t)nK, c->p_ref);
  r_dn(&r, "grid_theta_ref", (size_t)nK, c->theta_ref);
  r_dn(&r, "grid_h_min", (size_t)nk + 1, c->h_min);
  c->ptop = c->p_ref[1];
  if (c->ntp > 0) {
    c->pdat = (double *)calloc((size_t)c->ntp, sizeof(double));
    c->tdat = (double *)calloc((size_t)c->ntp, sizeof(double));
    r_dn(&r, "var_pdat", (size_t)c->ntp, c->pdat);
    r_dn(&r, "var_tdat", (size_t)c->ntp, c->tdat);
    for (int i = 0; i < c->ntp; i++) c->pdat[i] *= 100.;  // hPa to Pa
  }
  if (c->n_t_cool > 0) {
    c->t_cool_x = (double *)calloc((size_t)c->n_t_cool, sizeof(double));
    c->t_cool_y = (double *)calloc((size_t)c->n_t_cool, sizeof(double));
    r_dn(&r, "var_t_cool_p", (size_t)c->n_t_cool, c->t_cool_x);
    r_dn(&r, "var_t_cool_tau", (size_t)c->n_t_cool, c->t_cool_y);
    for (int i = 0; i < c->n_t_cool; i++) c->t_cool_x[i] = -log(c->t_cool_x[i] * 100.);
  }
  if (!r.bad && epic_case_dsgth(c)) {
    snprintf(g_err, sizeof(g_err), "%s: sigmatheta is not monotone", path);
    epic_case_free(c);
    epic_nc_header_free(f);
    return -3;
  }
  const size_t n3 = (size_t)(nk + 1) * (nj + 1) * ni, N3 = (size_t)c->Nelem3d;
  struct {
    const char *name;
    double *dst;
  } st[5] = {{"u", c->u}, {"v", c->v}, {"p3", c->p3}, {"theta", c->theta}, {"u_spinup", c->u_spinup}};
  for (int q = 0; q < 5 && !r.bad; q++) {
    double *v = r_var(&r, st[q].name, n3);
    if (!v) break;
    if (q == 4)  // u_spinup above the model top as in the top layer (var_read)
      for (size_t i = 0; i < (size_t)(nj + 1) * ni; i++) v[i] = v[(size_t)(nj + 1) * ni + i];
    scatter3(c, v, st[q].dst);
    free(v);
  }
  if (!r.bad && strcmp(c->planet_type, "terrestrial") == 0) {
    double *v = r_var(&r, "gz_surface", (size_t)(nj + 1) * ni);
    if (v) {
      size_t o = 0;
      for (int J = 0; J <= nj; J++)
        for (int I = 1; I <= ni; I++) c->gz_surface[(size_t)I + (size_t)(J + 1) * c->Iadim] = v[o++];
      for (int J = -1; J <= nj + 1; J++) {
        const size_t row = (size_t)(J + 1) * c->Iadim;
        c->gz_surface[row] = c->gz_surface[row + ni];
        c->gz_surface[row + ni + 1] = c->gz_surface[row + 1];
      }
      free(v);
    }
  }
  if (!r.bad) {
    c->dudt_hist = (double *)calloc(2 * N3, sizeof(double));
    c->dvdt_hist = (double *)calloc(2 * N3, sizeof(double));
    const char *tn[2][2] = {{"du_IT_MINUS1dt", "du_IT_MINUS2dt"}, {"dv_IT_MINUS1dt", "dv_IT_MINUS2dt"}};
    for (int q = 0; q < 2 && !r.bad; q++)
      for (int p = 0; p < 2 && !r.bad; p++) {
        double *v = r_var(&r, tn[q][p], n3);
        if (v) {
          scatter3(c, v, (q ? c->dvdt_hist : c->dudt_hist) + (size_t)p * N3);
          free(v);
        }
      }
  }
  if (r.bad) {
    snprintf(g_err, sizeof(g_err), "%s: missing or malformed %s", path, r.what);
    epic_case_free(c);
    epic_nc_header_free(f);
    return -3;
  }
  // keep the global attributes only
  for (int i = 0; i < f->nvars; i++) {
    free(f->vars[i].data);
    f->vars[i].data = NULL;
    f->vars[i].nval = 0;
  }
  c->nc_header = f;
  return 0;
}
