/* dropin_host.c -- a minimal EPIC-shaped host program for the drop-in
 * timestep() of libepic_amd.so: it defines the globals an EPIC executable
 * defines (grid, var, IT_ZERO, IT_MINUS1, IT_MINUS2; struct layouts from
 * include/epic_amd.h), fills them from a seeded case the way the
 * reference's start-up does (make_arrays, get_h, the Adams-Bashforth
 * FLOAT_MAX flags), then calls timestep(SYNC_DIAGS_ONLY) once and
 * timestep(STEP_PROGS_AND_SYNC_DIAGS) NSTEPS times, advancing grid.itime as
 * epic_main does.
 *
 * Usage: dropin_host [--OPTION=VALUE ...] PLANET NI NJ NK DT SEED NU4 NUDIV NOISE_TH NOISE_UV NSTEPS OUT LAZY [NAME]
 * (OPTION: an epic_case_spec switch, as for the oracle's reference harness)
 * OUT holds the dump format of the oracle's reference harness: an int32
 * header {magic, ni, nj, nk, 1, nfields, nelem3d, nelem2d, 0} and one dump
 * of the final state.  LAZY=1 selects epic_amd_set_sync_mode(1) and a
 * single epic_amd_flush() at the end. */
#include <float.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

#include "../../include/epic_amd.h"

struct epic_grid_abi grid;
struct epic_var_abi var;
int IT_ZERO, IT_MINUS1, IT_MINUS2;

static double *arr(size_t n) {
  double *p = (double *)calloc(n, sizeof(double));
  if (!p) exit(3);
  return p;
}

/* --name=value switches of epic_case_spec (the options of the oracle's
 * reference harness); returns 0 or -1 */
static int spec_option(epic_case_spec *sp, const char *arg) {
  const char *eq = strchr(arg, '=');
  if (strncmp(arg, "--", 2) != 0 || !eq) return -1;
  size_t n = (size_t)(eq - arg - 2);
  const char *v = eq + 1, *k = arg + 2;
#define OPT_I(name) if (n == strlen(#name) && !strncmp(k, #name, n)) { sp->name = atoi(v); return 0; }
#define OPT_D(name) if (n == strlen(#name) && !strncmp(k, #name, n)) { sp->name = atof(v); return 0; }
  OPT_I(k_sponge) OPT_I(newt_cool) OPT_I(newt_cool_adjust) OPT_I(drag_v) OPT_I(drag_zonal_avg)
  OPT_D(tau_drag) OPT_D(prandtl) OPT_D(nu6_nondim) OPT_D(nu8_nondim) OPT_I(adv_mass) OPT_I(adv_nonmass) OPT_I(pv_scheme)
#undef OPT_I
#undef OPT_D
  return -1;
}

int main(int argc, char **argv) {
  epic_case_spec sp;
  memset(&sp, 0, sizeof(sp));
  while (argc > 1 && strncmp(argv[1], "--", 2) == 0) {
    if (spec_option(&sp, argv[1])) {
      fprintf(stderr, "unknown option %s\n", argv[1]);
      return 2;
    }
    argv++;
    argc--;
  }
  if (argc < 14) {
    fprintf(stderr, "usage: %s [--OPTION=VALUE ...] PLANET NI NJ NK DT SEED NU4 NUDIV NOISE_TH NOISE_UV NSTEPS OUT LAZY [NAME]\n",
            argv[0]);
    return 2;
  }
  sp.planet = atoi(argv[1]);
  sp.ni = atoi(argv[2]);
  sp.nj = atoi(argv[3]);
  sp.nk = atoi(argv[4]);
  sp.dt = atoi(argv[5]);
  sp.seed = strtoull(argv[6], NULL, 10);
  sp.nu4_nondim = atof(argv[7]);
  sp.nudiv_nondim = atof(argv[8]);
  sp.noise_theta = atof(argv[9]);
  sp.noise_uv = atof(argv[10]);
  int nsteps = atoi(argv[11]);
  const char *out = argv[12];
  int lazy = atoi(argv[13]);

  epic_case c;
  if (epic_case_build(&sp, &c)) return 3;
  const size_t N3 = (size_t)c.Nelem3d, N2 = (size_t)c.Nelem2d;

  static struct epic_planet_abi pl;
  pl.index = c.planet_index;
  /* NAME renames the planet while keeping the case's settings: a caller's
   * configuration that timestep() must refuse before it touches the GPU */
  snprintf(pl.name, sizeof(pl.name), "%s", argc > 14 ? argv[14] : c.planet_name);
  /* the reference's index of a renamed planet (include/epic_datatypes.h) */
  if (!strcmp(pl.name, "titan")) pl.index = 10;
  if (!strcmp(pl.name, "hot_jupiter")) pl.index = 18;
  if (!strcmp(pl.name, "venus_llr05")) pl.index = 20;
  snprintf(pl.type, sizeof(pl.type), "%s", c.planet_type);
  pl.re = c.re; pl.rp = c.rp; pl.omega_sidereal = pl.omega_synodic = c.omega;
  pl.cp = c.cp; pl.cpr = c.cpr; pl.rgas = c.rgas; pl.kappa = c.kappa;
  pl.GM = c.GM; pl.J2 = c.J2; pl.kinvisc = c.kinvisc;

  strcpy(grid.geometry, "globe");
  strcpy(grid.uv_timestep_scheme, "3rd-order Adams-Bashforth");
  strcpy(grid.turbulence_scheme, "off");
  strcpy(grid.eos, "ideal");
  grid.dt = c.dt; grid.nk = c.nk; grid.nj = c.nj; grid.ni = c.ni; grid.jtp = c.jtp;
  grid.dln = c.dln; grid.dlt = c.dlt; grid.dy0 = c.dy0;
  grid.globe_lonbot = c.globe_lonbot; grid.globe_lontop = c.globe_lontop;
  grid.globe_latbot = c.globe_latbot; grid.globe_lattop = c.globe_lattop;
  grid.pbot = c.pbot; grid.press0 = c.press0;
  grid.zeta0 = c.zeta0; grid.zeta1 = c.zeta1; grid.hybrid_alpha = c.hybrid_alpha; grid.sigma_sigma = c.sigma_sigma;
  grid.k_sponge = c.k_sponge; grid.k_sigma = c.k_sigma;
  grid.newt_cool_on = c.newt_cool_on; grid.newt_cool_adjust = c.newt_cool_adjust;
  grid.drag_v = c.drag_v; grid.drag_zonal_avg = c.drag_zonal_avg;
  grid.topo_scale = c.topo_scale; grid.prandtl = c.prandtl; grid.tau_drag = c.tau_drag;
  grid.nudiv_nondim = c.nudiv_nondim;
  for (int q = 0; q <= EPIC_MAX_NU_ORDER; q++) {
    grid.nu[q] = c.nu[q];
    grid.nu_nondim[q] = c.nu_nondim[q];
  }
  grid.lon = c.lon; grid.lat = c.lat; grid.rln = c.rln; grid.rlt = c.rlt; grid.f = c.f;
  grid.m = c.m; grid.n = c.n; grid.mn = c.mn; grid.g = c.g;
  grid.sigmatheta = c.sigmatheta; grid.dsgth = c.dsgth; grid.dsgth_inv = c.dsgth_inv;
  grid.p_ref = c.p_ref; grid.theta_ref = c.theta_ref; grid.h_min = c.h_min; grid.t_sponge_inv = c.t_sponge_inv;

  var.u.on = var.v.on = var.hdry.on = var.theta.on = 1;
  {
    const char *adv[2] = {"Predictor-corrector (Hsu, Konor, and Arakawa)", "Monotonized, 4th-order centered (Akima)"};
    strcpy(var.hdry.advection_scheme, adv[c.adv_mass]);
    strcpy(var.theta.advection_scheme, adv[c.adv_nonmass]);
  }
  var.u.value = arr(N3); var.v.value = arr(N3);
  var.u.tendency = arr(3 * N3); var.v.tendency = arr(3 * N3);
  var.hdry.value = arr(N3); var.theta.value = arr(N3);
  var.pdry3.value = arr(N3); var.p3.value = var.pdry3.value;
  var.hdry3.value = arr(N3); var.h3.value = var.hdry3.value; var.h2.value = var.hdry.value;
  struct epic_diag_abi *diags[] = {&var.p2, &var.theta2, &var.t2, &var.t3, &var.rho2, &var.rho3, &var.exner3,
                                   &var.gz3, &var.mont3, &var.heat3, &var.pv3, &var.div_uv3, &var.w3,
                                   &var.dzdt3, &var.u_spinup, &var.ri2};
  for (unsigned q = 0; q < sizeof(diags) / sizeof(diags[0]); q++) diags[q]->value = arr(N3);
  var.gz_surface.value = arr(N2);
  /* the sounding and the cooling-time table ({x, y, z} triplets) */
  if (c.ntp > 0) {
    var.ntp = c.ntp;
    var.pdat = c.pdat;
    var.tdat = c.tdat;
    var.n_t_cool = c.n_t_cool;
    double *trip = arr(3 * (size_t)c.n_t_cool);
    for (int q = 0; q < c.n_t_cool; q++) {
      trip[3 * q] = c.t_cool_x[q];
      trip[3 * q + 1] = c.t_cool_y[q];
    }
    var.t_cool_table = trip;
  }
  memcpy(var.u.value, c.u, sizeof(double) * N3);
  memcpy(var.v.value, c.v, sizeof(double) * N3);
  memcpy(var.theta.value, c.theta, sizeof(double) * N3);
  memcpy(var.p3.value, c.p3, sizeof(double) * N3);
  memcpy(var.u_spinup.value, c.u_spinup, sizeof(double) * N3);
  memcpy(var.gz_surface.value, c.gz_surface, sizeof(double) * N2);

  /* h from the layer's pressure thickness (get_h, dry), periodic in I */
  const int Ia = c.Iadim;
#define IX(K, J, I) ((size_t)(I) + (size_t)((J) + 1) * Ia + (size_t)(K) * N2)
  for (int K = 1; K <= c.nk; K++)
    for (int J = 0; J <= c.nj; J++)
      for (int I = 1; I <= c.ni; I++)
        var.hdry.value[IX(K, J, I)] = (c.p3[IX(K, J, I)] - c.p3[IX(K - 1, J, I)]) / (c.g[2 * J + 1] * c.dsgth[2 * K]);
  for (int K = 0; K <= c.nk + 1; K++)
    for (int J = 0; J <= c.nj + 1; J++) {
      var.hdry.value[IX(K, J, 0)] = var.hdry.value[IX(K, J, c.ni)];
      var.hdry.value[IX(K, J, c.ni + 1)] = var.hdry.value[IX(K, J, 1)];
    }
  IT_ZERO = 0; IT_MINUS1 = 1; IT_MINUS2 = 2;
  grid.it_uv_tend = IT_ZERO;
  for (int p = 1; p <= 2; p++)
    for (int K = 1; K <= c.nk; K++)
      for (int J = 0; J <= c.nj; J++)
        for (int I = 1; I <= c.ni; I++) {
          var.u.tendency[p * N3 + IX(K, J, I)] = DBL_MAX;
          var.v.tendency[p * N3 + IX(K, J, I)] = DBL_MAX;
        }

  if (lazy) epic_amd_set_sync_mode(1);
  /* a caller built with the Arakawa-Lamb PV_SCHEME says so */
  epic_amd_set_pv_scheme(sp.pv_scheme);
  grid.itime = 0;
  timestep(&pl, EPIC_SYNC_DIAGS_ONLY, NULL);
  struct timespec t0, t1, t2, t3;
  clock_gettime(CLOCK_MONOTONIC, &t0);
  for (int s = 1; s <= nsteps; s++) {
    if (s == 2) clock_gettime(CLOCK_MONOTONIC, &t2);
    grid.itime = (unsigned long)(s - 1);
    timestep(&pl, EPIC_STEP_PROGS_AND_SYNC_DIAGS, NULL);
  }
  clock_gettime(CLOCK_MONOTONIC, &t3);
  if (lazy) epic_amd_flush();
  clock_gettime(CLOCK_MONOTONIC, &t1);
  const double sec = (double)(t1.tv_sec - t0.tv_sec) + 1e-9 * (double)(t1.tv_nsec - t0.tv_nsec);
  /* wall time of the stepping calls (and the final flush in lazy mode),
   * host-to-device and back included */
  fprintf(stderr, "timestep calls %d mean_ms %.4f pt_steps_per_s %.4e\n", nsteps, 1e3 * sec / nsteps,
          (double)c.ni * (c.nj + 1) * c.nk * nsteps / sec);
  /* the same without the first stepping call (which records the step's
   * graph) and the final flush: the per-call cost of a long run */
  if (nsteps > 1) {
    const double ss = (double)(t3.tv_sec - t2.tv_sec) + 1e-9 * (double)(t3.tv_nsec - t2.tv_nsec);
    fprintf(stderr, "steady calls %d mean_ms %.4f pt_steps_per_s %.4e\n", nsteps - 1, 1e3 * ss / (nsteps - 1),
            (double)c.ni * (c.nj + 1) * c.nk * (nsteps - 1) / ss);
  }

  FILE *fo = fopen(out, "wb");
  if (!fo) return 4;
  int hdr[9] = {0x45504943, c.ni, c.nj, c.nk, 1, EPIC_F_COUNT, (int)N3, (int)N2, 0};
  fwrite(hdr, sizeof(int), 9, fo);
  int h4[4] = {nsteps, IT_ZERO, IT_MINUS1, IT_MINUS2};
  fwrite(h4, sizeof(int), 4, fo);
  const double *f3[] = {var.u.value, var.v.value, var.hdry.value, var.theta.value, var.p3.value,
                        var.p2.value, var.h3.value, var.theta2.value, var.t2.value, var.t3.value,
                        var.rho2.value, var.rho3.value, var.exner3.value, var.gz3.value,
                        var.mont3.value, var.pv3.value, var.div_uv3.value, var.w3.value,
                        var.heat3.value, var.dzdt3.value};
  for (int q = 0; q < 20; q++) fwrite(f3[q], sizeof(double), N3, fo);
  fwrite(var.u.tendency, sizeof(double), 3 * N3, fo);
  fwrite(var.v.tendency, sizeof(double), 3 * N3, fo);
  fwrite(var.gz_surface.value, sizeof(double), N2, fo);
  fwrite(var.ri2.value, sizeof(double), N3, fo);
  fclose(fo);
  printf("model_time %ld it_uv_tend %d\n", (long)var.model_time, grid.it_uv_tend);
  return 0;
}
