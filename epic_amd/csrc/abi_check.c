/* abi_check.c -- compile-time pins of the restated reference struct layouts
 * of include/epic_amd.h (x86-64 SysV, EPIC_PRECISION 8).  Compiled with
 * -fsyntax-only by the CPU tests; any drift fails the build.  The same
 * offsets are asserted against the reference's own headers in
 * oracle/ref_harness.c where the reference tree is present. */
#include <stddef.h>
#include "../../include/epic_amd.h"

#define AT(S, m, off) _Static_assert(offsetof(struct S, m) == (off), #S "." #m)
#define SZ(S, n) _Static_assert(sizeof(struct S) == (n), "sizeof " #S)

SZ(epic_planet_abi, 2584);
AT(epic_planet_abi, re, 64);
AT(epic_planet_abi, omega_sidereal, 80);
AT(epic_planet_abi, cp, 96);
AT(epic_planet_abi, cpr, 104);
AT(epic_planet_abi, rgas, 112);
AT(epic_planet_abi, kappa, 120);
AT(epic_planet_abi, GM, 128);
AT(epic_planet_abi, J2, 136);
AT(epic_planet_abi, orbit, 168);
AT(epic_planet_abi, kinvisc, 232);
AT(epic_planet_abi, u, 256);

SZ(epic_grid_abi, 2024);
AT(epic_grid_abi, globe_lonbot, 1304);
AT(epic_grid_abi, dt, 1368);
AT(epic_grid_abi, nk, 1376);
AT(epic_grid_abi, nj, 1380);
AT(epic_grid_abi, ni, 1384);
AT(epic_grid_abi, jtp, 1388);
AT(epic_grid_abi, jlo, 1416);
AT(epic_grid_abi, dln, 1432);
AT(epic_grid_abi, pbot, 1472);
AT(epic_grid_abi, press0, 1480);
AT(epic_grid_abi, zeta0, 1496);
AT(epic_grid_abi, sigma_sigma, 1520);
AT(epic_grid_abi, k_sponge, 1528);
AT(epic_grid_abi, k_sigma, 1532);
AT(epic_grid_abi, newt_cool_on, 1536);
AT(epic_grid_abi, du_vert, 1560);
AT(epic_grid_abi, nq, 1616);
AT(epic_grid_abi, sigmatheta, 1640);
AT(epic_grid_abi, h_min, 1664);
AT(epic_grid_abi, rln, 1672);
AT(epic_grid_abi, dsgth_inv, 1752);
AT(epic_grid_abi, drag_v, 1768);
AT(epic_grid_abi, t_sponge_inv, 1776);
AT(epic_grid_abi, topo_scale, 1784);
AT(epic_grid_abi, nudiv_nondim, 1808);
AT(epic_grid_abi, ab, 1816);
AT(epic_grid_abi, nu_nondim, 1840);
AT(epic_grid_abi, nu, 1912);
AT(epic_grid_abi, it_uv, 2000);
AT(epic_grid_abi, itime, 2016);

SZ(epic_wind_abi, 344);
AT(epic_wind_abi, value, 8);
AT(epic_wind_abi, tendency, 16);
SZ(epic_thermo_abi, 352);
AT(epic_thermo_abi, value, 8);
SZ(epic_diag_abi, 96);
AT(epic_diag_abi, value, 8);

SZ(epic_var_abi, 20496);
AT(epic_var_abi, ntp, 436);
AT(epic_var_abi, u, 464);
AT(epic_var_abi, v, 808);
AT(epic_var_abi, hdry, 1152);
AT(epic_var_abi, theta, 1504);
AT(epic_var_abi, nu_turb, 2208);
AT(epic_var_abi, species, 2560);
AT(epic_var_abi, pdat, 16640);
AT(epic_var_abi, hdry3, 16664);
AT(epic_var_abi, gz_surface, 19256);
AT(epic_var_abi, nmt_diag, 19352);
AT(epic_var_abi, t_cool_table, 20312);
AT(epic_var_abi, l_s_value, 20320);
AT(epic_var_abi, time_fp_bar, 20488);
