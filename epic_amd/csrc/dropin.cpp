// dropin.cpp -- the reference's timestep(planet, action, Buff2D) over the
// MI355X engine (include/epic.h:654-656, src/shared/epic_timestep.c:62-390).
//
// The caller's globals `grid`, `var`, IT_ZERO, IT_MINUS1, IT_MINUS2
// (src/shared/epic_globals.c, include/epic.h:303-304) are bound as weak
// undefined symbols: inside an EPIC executable they resolve to the
// executable's own definitions; in any other process they are null and the
// call stops with an error.  The first call builds one engine member from the
// caller's grid and planet.  Every call then
//   1. uploads the caller's state (prognostics, tendency planes, the
//      diagnostics the step reads) unless the lazy mode holds it resident,
//   2. runs SYNC_DIAGS_ONLY or STEP_PROGS_AND_SYNC_DIAGS on the GPU,
//   3. writes every array the reference's timestep() writes back into the
//      caller's var.* (unless lazy), rotates IT_* and sets grid.it_uv*,
//      as the reference's timeplane_bookkeeping() does
//      (src/shared/epic_funcs_diag.c:7086-7100),
//   4. advances var.model_time by grid.dt and updates L_s with the caller's
//      own solar_longitude() (src/shared/epic_timestep.c:299-305).
// Configurations the GPU path does not implement stop with an error message
// rather than running a different algorithm.
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

#include <vector>

#include "case_internal.h"
#include "engine_internal.h"

extern "C" {
extern struct epic_grid_abi grid __attribute__((weak));
extern struct epic_var_abi var __attribute__((weak));
extern int IT_ZERO __attribute__((weak));
extern int IT_MINUS1 __attribute__((weak));
extern int IT_MINUS2 __attribute__((weak));
double solar_longitude(struct epic_planet_abi *planet, time_t date) __attribute__((weak));
void epic_error(char *calling_function, char *Message) __attribute__((weak));
}

using epic::MemberDesc;

namespace {

struct DropIn {
  epic_engine *E = nullptr;
  int slot = -1;
  int lazy = 0;
  int pgrad_traditional = 0;
  int pv_al81 = 0;
  bool resident = false;  // the device copy is newer than or equal to var.*
  bool dirty_host = false;  // var.* lags the device copy (lazy mode)
  int ni = 0, nj = 0, nk = 0;
  // the member's time planes and itime as the device holds them (after a
  // step the device rotates them as the host does and counts itime on), so a
  // call whose IT_* and itime match sends nothing
  bool planes_known = false;
  int planes[4] = {0, 0, 0, 0};
};
DropIn g_di;

[[noreturn]] void fail(const char *msg) {
  static char fn[] = "timestep";
  char buf[512];
  snprintf(buf, sizeof(buf), "%s", msg);
  if (epic_error) epic_error(fn, buf);
  fprintf(stderr, "\n\nEPIC (MI355X) Error: %s(): %s\n\n", fn, buf);
  fflush(stderr);
  exit(1);
}

void need(bool ok, const char *what) {
  if (!ok) fail(what);
}

struct Xfer {
  double *host;
  double *dev;
  size_t n;
  bool down;     // written by the step: copied back to the caller
  bool up = true;  // read by the step: copied to the device
};

// Every array of var.* that the step reads or writes, paired with the
// member's device copy.  The dry-air aliases of the reference
// (src/shared/epic_funcs_diag.c:879-926: p3 == pdry3, h3 == hdry3) mean
// one device array serves both names.
std::vector<Xfer> transfers(const MemberDesc &d) {
  const size_t N3 = (size_t)d.N3, N2 = (size_t)d.N2;
  std::vector<Xfer> x = {
      {var.u.value, d.U, N3, true},           {var.v.value, d.V, N3, true},
      {var.u.tendency, d.DUDT, 3 * N3, true}, {var.v.tendency, d.DVDT, 3 * N3, true},
      {var.hdry.value, d.HDRY, N3, true},     {var.theta.value, d.THETA, N3, true},
      {var.p3.value, d.P3, N3, true},         {var.p2.value, d.P2, N3, true},
      {var.theta2.value, d.THETA2, N3, true}, {var.h3.value, d.H3, N3, true},
      {var.t2.value, d.T2, N3, true},         {var.t3.value, d.T3, N3, true},
      {var.rho2.value, d.RHO2, N3, true},     {var.rho3.value, d.RHO3, N3, true},
      {var.exner3.value, d.EXNER3, N3, true}, {var.gz3.value, d.GZ3, N3, true},
      {var.mont3.value, d.MONT3, N3, true},   {var.heat3.value, d.HEAT3, N3, true},
      {var.pv3.value, d.PV3, N3, true},       {var.div_uv3.value, d.DIV, N3, true},
      {var.w3.value, d.W3, N3, true},         {var.dzdt3.value, d.DZDT3, N3, true},
      {var.u_spinup.value, d.USPIN, N3, false}, {var.gz_surface.value, d.GZS, N2, true},
      {var.ri2.value, d.RI2, N3, true, false},
  };
  return x;
}

void upload(const MemberDesc &d) {
  hipStream_t s = engine_stream(g_di.E);
  for (auto &t : transfers(d)) {
    if (!t.host || !t.up) continue;
    need(hipMemcpyAsync(t.dev, t.host, sizeof(double) * t.n, hipMemcpyHostToDevice, s) == hipSuccess,
         "host-to-device copy failed");
  }
  need(hipStreamSynchronize(s) == hipSuccess, "upload failed");
}

void download(const MemberDesc &d) {
  hipStream_t s = engine_stream(g_di.E);
  // RI2 is formed from the synced state when it is copied back
  if (var.ri2.value) need(engine_richardson(g_di.E, g_di.slot) == 0, "RI2 launch failed");
  for (auto &t : transfers(d)) {
    if (!t.host || !t.down) continue;
    need(hipMemcpyAsync(t.host, t.dev, sizeof(double) * t.n, hipMemcpyDeviceToHost, s) == hipSuccess,
         "device-to-host copy failed");
  }
  need(hipStreamSynchronize(s) == hipSuccess, "download failed");
}

// index of an advection scheme's name (epic_case_spec.adv_mass), -1 if none
static int adv_scheme(const char *name) {
  if (strcmp(name, EPIC_ADV_HSU) == 0) return 0;
  if (strcmp(name, EPIC_ADV_AKIMA) == 0) return 1;
  return -1;
}

// The configurations the GPU step implements (the reference's defaults for
// a dry globe).
void check_config(const struct epic_planet_abi *planet) {
  need(&grid != nullptr && &var != nullptr && &IT_ZERO != nullptr && &IT_MINUS1 != nullptr &&
           &IT_MINUS2 != nullptr,
       "the caller defines no EPIC globals (grid, var, IT_ZERO, IT_MINUS1, IT_MINUS2)");
  need(strcmp(grid.geometry, "globe") == 0, "only grid.geometry \"globe\" is implemented");
  need(strcmp(grid.uv_timestep_scheme, "3rd-order Adams-Bashforth") == 0,
       "only the 3rd-order Adams-Bashforth uv scheme is implemented");
  need(strcmp(grid.turbulence_scheme, "off") == 0, "grid.turbulence_scheme must be \"off\"");
  need(strcmp(grid.eos, "ideal") == 0, "only the ideal equation of state is implemented");
  need(grid.nq == 0 && grid.microphysics_on == 0 && grid.nmt_physics_on == 0,
       "species, microphysics and NMT physics are not implemented");
  need(!var.fpara.on && !var.nu_turb.on, "fpara and nu_turb are not implemented");
  need(var.hdry.on && var.theta.on, "hdry and theta must be on");
  need(adv_scheme(var.theta.advection_scheme) >= 0 && adv_scheme(var.hdry.advection_scheme) >= 0,
       "only the Hsu-Konor-Arakawa and Akima advection schemes are implemented");
  need(var.p3.value == var.pdry3.value && var.h3.value == var.hdry3.value,
       "dry runs alias p3/pdry3 and h3/hdry3");
  need(strcmp(planet->type, "gas-giant") == 0 || strcmp(planet->type, "terrestrial") == 0,
       "planet->type must be gas-giant or terrestrial");
  need(var.u.value && var.v.value && var.u.tendency && var.v.tendency && var.hdry.value && var.theta.value,
       "prognostic arrays are not allocated");
  need(planet->x_h2 == 0., "the ortho/para hydrogen thermodynamics (planet->x_h2 > 0) are not implemented");
  // calc_heating: Held-Suarez's analytic relaxation, or, on every other
  // planet (a hot Jupiter included), the relaxation toward the t_vs_p
  // sounding with t_cool_table's times (src/shared/epic_sensible_heating.c:
  // 103-130, 294-311); Venus (LLR05) and Titan take their temp_eq instead,
  // which is not built here.  The engine's own day-night relaxation is not
  // the reference's and is not offered here.
  if (grid.newt_cool_on && strcmp(planet->name, "held_suarez") != 0) {
    need(planet->index != 20 && planet->index != 10,
         "grid.newt_cool_on: the temp_eq relaxations of venus_llr05 and titan are not implemented");
    need(var.ntp >= 2 && var.pdat && var.tdat && var.n_t_cool >= 2 && var.t_cool_table,
         "grid.newt_cool_on needs the t_vs_p sounding (var.ntp, var.pdat, var.tdat) and var.t_cool_table");
  }
}

void create_member(const struct epic_planet_abi *planet) {
  epic_case c;
  memset(&c, 0, sizeof(c));
  snprintf(c.planet_name, sizeof(c.planet_name), "%s", planet->name);
  snprintf(c.planet_type, sizeof(c.planet_type), "%s", planet->type);
  c.planet_index = planet->index;
  c.re = planet->re; c.rp = planet->rp; c.omega = planet->omega_sidereal;
  c.cp = planet->cp; c.rgas = planet->rgas; c.cpr = planet->cpr; c.kappa = planet->kappa;
  c.GM = planet->GM; c.J2 = planet->J2; c.x_h2 = planet->x_h2; c.kinvisc = planet->kinvisc;
  c.ni = grid.ni; c.nj = grid.nj; c.nk = grid.nk; c.dt = grid.dt; c.jtp = grid.jtp;
  c.k_sigma = grid.k_sigma; c.k_sponge = grid.k_sponge;
  c.newt_cool_on = grid.newt_cool_on; c.newt_cool_adjust = grid.newt_cool_adjust;
  c.drag_v = grid.drag_v; c.drag_zonal_avg = grid.drag_zonal_avg;
  c.pgrad_traditional = g_di.pgrad_traditional;
  c.pv_scheme = g_di.pv_al81;
  c.adv_mass = adv_scheme(var.hdry.advection_scheme);
  c.adv_nonmass = adv_scheme(var.theta.advection_scheme);
  c.dln = grid.dln; c.dlt = grid.dlt; c.dy0 = grid.dy0;
  c.globe_lonbot = grid.globe_lonbot; c.globe_lontop = grid.globe_lontop;
  c.globe_latbot = grid.globe_latbot; c.globe_lattop = grid.globe_lattop;
  c.press0 = grid.press0; c.pbot = grid.pbot;
  c.zeta0 = grid.zeta0; c.zeta1 = grid.zeta1; c.hybrid_alpha = grid.hybrid_alpha; c.sigma_sigma = grid.sigma_sigma;
  c.topo_scale = grid.topo_scale; c.prandtl = grid.prandtl; c.tau_drag = grid.tau_drag;
  c.nudiv_nondim = grid.nudiv_nondim;
  for (int i = 0; i <= EPIC_MAX_NU_ORDER; i++) {
    c.nu[i] = grid.nu[i];
    c.nu_nondim[i] = grid.nu_nondim[i];
  }
  c.Iadim = grid.ni + 2; c.Jadim = grid.nj + 3; c.Kadim = grid.nk + 2;
  c.Nelem2d = c.Iadim * c.Jadim; c.Nelem3d = c.Nelem2d * c.Kadim;
  c.lon = grid.lon; c.lat = grid.lat; c.rln = grid.rln; c.rlt = grid.rlt; c.f = grid.f;
  c.m = grid.m; c.n = grid.n; c.mn = grid.mn; c.g = grid.g;
  c.sigmatheta = grid.sigmatheta; c.dsgth = grid.dsgth; c.dsgth_inv = grid.dsgth_inv;
  c.p_ref = grid.p_ref; c.theta_ref = grid.theta_ref; c.h_min = grid.h_min; c.t_sponge_inv = grid.t_sponge_inv;
  c.u = var.u.value; c.v = var.v.value; c.hdry = var.hdry.value; c.theta = var.theta.value;
  c.p3 = var.p3.value; c.u_spinup = var.u_spinup.value; c.gz_surface = var.gz_surface.value;
  need(c.u_spinup && c.gz_surface && c.p3, "u_spinup, gz_surface and p3 must be allocated");
  // the sounding and the cooling-time table ({x, y, z} triplets, x = -log p)
  std::vector<double> tcx, tcy;
  if (c.newt_cool_on && strcmp(planet->name, "held_suarez") != 0) {
    c.ntp = var.ntp; c.pdat = var.pdat; c.tdat = var.tdat;
    const double *trip = (const double *)var.t_cool_table;
    for (int q = 0; q < var.n_t_cool; q++) {
      tcx.push_back(trip[3 * q]);
      tcy.push_back(trip[3 * q + 1]);
    }
    c.n_t_cool = var.n_t_cool; c.t_cool_x = tcx.data(); c.t_cool_y = tcy.data();
  }

  if (!g_di.E) {
    g_di.E = epic_engine_create(EPIC_FP64, 1, 0);
    need(g_di.E != nullptr, "could not create the GPU engine");
  }
  int slot = engine_add_member(g_di.E, &c, var.hdry.value, false);
  if (slot < 0) fail(engine_errbuf(g_di.E));
  g_di.slot = slot;
  g_di.planes_known = false;
  g_di.ni = grid.ni; g_di.nj = grid.nj; g_di.nk = grid.nk;
  g_di.resident = false;
}

}  // namespace

extern "C" void epic_amd_set_sync_mode(int lazy) { g_di.lazy = lazy ? 1 : 0; }

extern "C" void epic_amd_set_pgrad_traditional(int on) { g_di.pgrad_traditional = on ? 1 : 0; }
extern "C" void epic_amd_set_pv_scheme(int al81) { g_di.pv_al81 = al81 ? 1 : 0; }

extern "C" void epic_amd_flush(void) {
  if (!g_di.E || g_di.slot < 0 || !g_di.dirty_host) return;
  download(*engine_member_desc(g_di.E, g_di.slot));
  g_di.dirty_host = false;
}

extern "C" void epic_amd_invalidate(void) {
  if (g_di.E && g_di.slot >= 0) epic_engine_remove(g_di.E, g_di.slot);
  g_di.slot = -1;
  g_di.planes_known = false;
  g_di.resident = false;
  g_di.dirty_host = false;
}

extern "C" void timestep(struct epic_planet_abi *planet, int action, double **Buff2D) {
  (void)Buff2D;
  if (action != EPIC_STEP_PROGS_AND_SYNC_DIAGS && action != EPIC_SYNC_DIAGS_ONLY) {
    char msg[64];
    snprintf(msg, sizeof(msg), "unrecognized action=%d", action);
    fail(msg);
  }
  if (g_di.slot < 0 || g_di.ni != grid.ni || g_di.nj != grid.nj || g_di.nk != grid.nk) {
    check_config(planet);
    if (g_di.slot >= 0) epic_amd_invalidate();
    create_member(planet);
  }
  const MemberDesc &d = *engine_member_desc(g_di.E, g_di.slot);
  if (!g_di.lazy || !g_di.resident) {
    upload(d);
    g_di.resident = true;
  }
  int planes[4] = {IT_ZERO, IT_MINUS1, IT_MINUS2, (int)grid.itime};
  if (!g_di.planes_known || memcmp(planes, g_di.planes, sizeof(planes)) != 0) {
    if (epic_engine_set_time_planes(g_di.E, g_di.slot, planes) != 0) fail(engine_errbuf(g_di.E));
  }

  if (action == EPIC_SYNC_DIAGS_ONLY) {
    if (engine_sync_member(g_di.E, g_di.slot) != 0) fail(engine_errbuf(g_di.E));
    memcpy(g_di.planes, planes, sizeof(planes));
  } else {
    if (epic_engine_step(g_di.E, 1) != 0) fail(engine_errbuf(g_di.E));
    // the step's closing rotation: (it0, itm1, itm2) <- (itm2, it0, itm1), itime + 1
    const int rot[4] = {planes[2], planes[0], planes[1], planes[3] + 1};
    memcpy(g_di.planes, rot, sizeof(rot));
  }
  g_di.planes_known = true;
  if (epic_engine_sync(g_di.E) != 0) fail(engine_errbuf(g_di.E));

  if (action == EPIC_STEP_PROGS_AND_SYNC_DIAGS) {
    // Adams-Bashforth plane rotation and bookkeeping
    int itmp = IT_MINUS2;
    IT_MINUS2 = IT_MINUS1;
    IT_MINUS1 = IT_ZERO;
    IT_ZERO = itmp;
    grid.it_uv = 0;
    grid.it_uv_dis = 0;
    grid.it_uv_tend = IT_ZERO;
    var.model_time += (time_t)grid.dt;
    if (solar_longitude) var.l_s_value = solar_longitude(planet, var.model_time);
  }
  if (g_di.lazy) {
    g_di.dirty_host = true;
  } else {
    download(d);
  }
}
