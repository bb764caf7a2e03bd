/*
 * epic_amd.h -- C ABI of the MI355X (gfx950) implementation of EPIC's
 * dynamical-core timestep.
 *
 * Two surfaces are exported by libepic_amd.so:
 *
 *  1. The drop-in.  `timestep(planet, action, Buff2D)` with the exact
 *     signature of the reference (include/epic.h:654-656,
 *     src/shared/epic_timestep.c:62-390).  It reads and writes the caller's
 *     global `grid` and `var` structures (src/shared/epic_globals.c), whose
 *     memory layout is restated below (struct epic_*_abi) from the measured
 *     offsets of the reference headers (include/epic_datatypes.h:238-627,
 *     EPIC_PRECISION=8, x86-64 SysV).  Only the members the step touches are
 *     named; the rest of each struct is opaque padding of the right size.
 *
 *  2. The native batched engine (epic_engine_*): many independent
 *     integrations ("members") of possibly different grid sizes held
 *     resident in HBM and advanced together, in FP64 or in the mixed
 *     FP32/FP64 precision mode.  It has no counterpart in the reference;
 *     `timestep()` is its single-member special case.
 *
 * All entry points take plain pointers and sizes.  Errors are reported as
 * negative return codes (engine) or, for the drop-in, as an EPIC-style
 * message on stderr followed by exit(1), as the reference's epic_error()
 * does (src/shared/epic_funcs_io.c:4877-4889).
 */
#ifndef EPIC_AMD_H
#define EPIC_AMD_H

#include <stddef.h>
#include <stdint.h>
#include <time.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ------------------------------------------------------------------------
 * Constants shared with the reference (include/epic.h:294-295,
 * include/epic_datatypes.h:72-188).
 * ---------------------------------------------------------------------- */
#define EPIC_STEP_PROGS_AND_SYNC_DIAGS 0
#define EPIC_SYNC_DIAGS_ONLY           1

#define EPIC_HELD_SUAREZ_INDEX 19
#define EPIC_MAX_NU_ORDER       8

/* ------------------------------------------------------------------------
 * Reference struct layouts (EPIC_PRECISION == 8).  Offsets were measured
 * with gcc 11.4 on the reference headers; each is pinned by a static
 * assertion in epic_amd/csrc/abi_check.c.
 * ---------------------------------------------------------------------- */

/* planetspec, include/epic_datatypes.h:238-274 (2584 bytes) */
struct epic_planet_abi {
  int32_t index;                  /*    0 */
  char    name[32];               /*    4 */
  char    type[16];               /*   36 */
  char    orbital_epoch[8];       /*   52 */
  double  re, rp;                 /*   64, 72 */
  double  omega_sidereal;         /*   80 */
  double  omega_synodic;          /*   88 */
  double  cp, cpr, rgas, kappa;   /*   96 .. 120 */
  double  GM, J2;                 /*  128, 136 */
  double  x_h2, x_he, x_3;        /*  144 .. 160 */
  double  orbit[8];               /*  168: a,e,i,lon_asc,lon_peri,mean_lon,period,vernal_eq_anom */
  double  kinvisc, dynvisc, k_a;  /*  232 .. 248 */
  void   *u;                      /*  256: zonal-wind profile function */
  unsigned char cloud[2320];      /*  264: microphysics_spec[10], unused by the step */
};

/* gridspec, include/epic_datatypes.h:279-394 (2024 bytes) */
struct epic_grid_abi {
  char    geometry[16];           /*    0 */
  char    advection_scheme[256];  /*   16 */
  char    uv_timestep_scheme[256];/*  272 */
  char    turbulence_scheme[256]; /*  528 */
  char    stability_factor[256];  /*  784 */
  char    extract_append[256];    /* 1040 */
  double  data_version;           /* 1296 */
  double  globe_lonbot, globe_lontop, globe_latbot, globe_lattop; /* 1304 .. 1328 */
  char    f_plane_map[16];        /* 1336 */
  double  f_plane_lat0, f_plane_half_width; /* 1352, 1360 */
  int32_t dt, cfl_dt, nk, nj, ni, jtp;      /* 1368 .. 1388 */
  int32_t wrap[3], pad[3];        /* 1392, 1404 */
  int32_t jlo, jfirst, jlast, we_num_nodes; /* 1416 .. 1428 */
  double  dln, dlt, dy0;          /* 1432 .. 1448 */
  double  sgth_bot, sgth_top;     /* 1456, 1464 */
  double  pbot, press0, mont0;    /* 1472 .. 1488 */
  double  zeta0, zeta1, hybrid_alpha, sigma_sigma; /* 1496 .. 1520 */
  int32_t k_sponge, k_sigma;      /* 1528, 1532 */
  int32_t newt_cool_on, newt_cool_adjust; /* 1536, 1540 */
  int32_t microphysics_on, nmt_physics_on, diffusion_direction; /* 1544 .. 1552 */
  int32_t _pad0;
  double  du_vert;                /* 1560 */
  char    eos[8];                 /* 1568 */
  int32_t aux_a, aux_b, aux_c;    /* 1576 */
  int32_t _pad1;
  double  aux_fa, aux_fb, aux_fc; /* 1592 */
  int32_t nq;                     /* 1616 */
  int32_t _pad2;
  int32_t *is, *ip;               /* 1624, 1632 */
  double  *sigmatheta;            /* 1640 */
  double  *p_ref, *theta_ref, *h_min; /* 1648 .. 1664 */
  double  *rln, *rlt, *f, *m, *n, *mn, *g, *lat, *lon, *dsgth, *dsgth_inv; /* 1672 .. 1752 */
  int32_t is_spole, is_npole, drag_v, drag_zonal_avg; /* 1760 .. 1772 */
  double  *t_sponge_inv;          /* 1776 */
  double  topo_scale, prandtl, tau_drag, nudiv_nondim; /* 1784 .. 1808 */
  double  ab[3];                  /* 1816 */
  double  nu_nondim[9];           /* 1840 */
  double  nu[9];                  /* 1912 */
  int32_t itback, itsave, itextract, itrun; /* 1984 .. 1996 */
  int32_t it_uv, it_uv_dis, it_uv_tend, it_h; /* 2000 .. 2012 */
  unsigned long itime;            /* 2016 */
};

/* wind_variable, include/epic_datatypes.h:536-545 (344 bytes) */
struct epic_wind_abi {
  int32_t on, extract_on;
  double *value;                  /*   8 */
  double *tendency;               /*  16 */
  unsigned char info[320];        /*  24: id_information[4] */
};

/* thermo_variable, include/epic_datatypes.h:547-555 (352 bytes) */
struct epic_thermo_abi {
  int32_t on, extract_on;
  double *value;                  /*   8 */
  char    advection_scheme[256];  /*  16 */
  unsigned char info[80];         /* 272 */
};

/* diagnostic_variable, include/epic_datatypes.h:557-563 (96 bytes) */
struct epic_diag_abi {
  int32_t on, extract_on;
  double *value;                  /*   8 */
  unsigned char info[80];         /*  16 */
};

/* variablespec, include/epic_datatypes.h:566-627 (20496 bytes) */
struct epic_var_abi {
  int32_t on_list[54];            /*     0 */
  int32_t extract_on;             /*   216 */
  int32_t extract_on_list[54];    /*   220 */
  int32_t ntp, n_t_cool;          /*   436, 440 */
  int32_t _pad0;
  time_t  start_time, model_time; /*   448, 456 */
  struct epic_wind_abi u, v;      /*   464, 808 */
  struct epic_thermo_abi hdry, theta, fpara, nu_turb; /* 1152 .. 2208 */
  unsigned char species[14080];   /*  2560: species_variable[16], dry runs only */
  double *pdat, *tdat, *dtdat;    /* 16640 */
  struct epic_diag_abi hdry3, pdry3, p2, p3, theta2, h2, h3, t2, t3, rho2, rho3,
                       exner3, fgibb3, gz3, mont3, heat3, pv3, eddy_pv3, ri2,
                       vort3, div_uv3, w3, dzdt3, u_spinup, diffusion_coef_uv,
                       diffusion_coef_theta, diffusion_coef_mass, gz_surface;
                                  /* 16664 .. 19256 */
  unsigned char nmt_diag[960];    /* 19352: ten nmt diagnostic_variables */
  void   *t_cool_table;           /* 20312 */
  double  l_s_value;              /* 20320: time_variable l_s (value first) */
  unsigned char l_s_info[80];     /* 20328 */
  unsigned char info[80];         /* 20408 */
  double  time_fp_bar;            /* 20488 */
};

/*
 * Drop-in replacement for the reference's timestep()
 * (include/epic.h:654-656; body src/shared/epic_timestep.c:62-390).
 * Reads the caller's globals `grid`, `var`, IT_ZERO, IT_MINUS1, IT_MINUS2
 * (include/epic.h:303-304), keeps the state resident on the GPU between
 * calls and writes the post-conditions of the reference back into the
 * caller's arrays (see epic_amd_set_sync_mode()).
 */
#ifndef EPIC_AMD_NO_DROPIN_DECL
void timestep(struct epic_planet_abi *planet, int action, double **Buff2D);
#endif

/* Host write-back policy of the drop-in: 0 = after every call (default),
 * 1 = only when epic_amd_flush() is called (e.g. from a var_write wrapper). */
void epic_amd_set_sync_mode(int lazy);
/* Copy the device-resident state and diagnostics into the caller's var.*. */
void epic_amd_flush(void);
/* Forget the device copy; the next timestep() re-uploads var.* and grid.*. */
void epic_amd_invalidate(void);
/* The pressure gradient of a caller built with PGRAD_TYPE =
 * PGRAD_TRADITIONAL (a compile-time switch of the reference's
 * src/shared/epic_timestep.c:772-774, invisible to the library): 1 selects
 * uv_pgrad_traditional, 0 the default finite-volume form.  Call before the
 * first timestep() (or invalidate after). */
void epic_amd_set_pgrad_traditional(int on);
/* The PV averaging of a caller built with PV_SCHEME = ARAKAWA_LAMB_1981
 * (the other compile-time value of the reference's include/epic.h:79-86,
 * invisible to the library): 1 selects Arakawa and Lamb (1981), 0 the
 * default Sadourny (1975).  Call before the first timestep() (or
 * invalidate after). */
void epic_amd_set_pv_scheme(int al81);

/* ------------------------------------------------------------------------
 * Synthetic cases: planet + grid + seeded initial state, laid out exactly
 * as the reference lays out its globals (padded (K,J,I), Kshift=0,
 * Jshift=-1, Ishift=0; src/shared/epic_funcs_diag.c:654-664).
 * ---------------------------------------------------------------------- */
enum epic_planet_kind {
  EPIC_PLANET_HELD_SUAREZ = 0, /* dry Earth, Held & Suarez (1994) forcing   */
  EPIC_PLANET_JUPITER_LIKE = 1,/* ideal-gas gas giant, jets, no forcing      */
  EPIC_PLANET_SATURN_LIKE = 2, /* ideal-gas gas giant, jets, no forcing      */
  EPIC_PLANET_VENUS_LIKE = 3,  /* slow rotator, Held-Suarez-type relaxation  */
  EPIC_PLANET_HOT_JUPITER = 4  /* tidally locked gas giant: Newtonian
                                  relaxation to a day-night T_eq            */
};

/* k_sponge value that asks for the reference `initial` program's choice:
 * 0 for Held-Suarez, NINT(0.2 nk) clamped to 1..4 for any other planet with
 * nk >= 5, 0 for nk <= 4 (src/shared/epic_initial.c:1249-1266). */
#define EPIC_K_SPONGE_DEFAULT (-1)

typedef struct {
  int    planet;          /* enum epic_planet_kind                         */
  int    ni, nj, nk;      /* ni a power of two, nj >= 4, nk >= 4           */
  int    dt;              /* seconds                                       */
  uint64_t seed;          /* noise and jet coefficients                    */
  double nu4_nondim;      /* del^4 hyperviscosity, fraction of maximum     */
  double nudiv_nondim;    /* divergence damping, fraction of maximum       */
  double noise_theta;     /* K, standard deviation of theta noise          */
  double noise_uv;        /* m/s, standard deviation of u,v noise          */
  /* The reference's drag and forcing switches (gridspec members of the same
   * names, src/shared/epic_timestep.c:1303-1509,
   * src/shared/epic_sensible_heating.c:51-180).  A zeroed tail gives no
   * sponge, no Rayleigh drag and the planet's own Newtonian-cooling
   * default (on for Held-Suarez and the hot Jupiter). */
  int    k_sponge;        /* top sponge layers (t_sponge_inv as set_sponge
                             forms it), or EPIC_K_SPONGE_DEFAULT            */
  int    newt_cool;       /* 0: planet default, 1: on, -1: off; on for a
                             planet without its own relaxation it relaxes
                             toward a sounding with a cooling-time table
                             the case builder forms (epic_case: ntp ...)    */
  int    newt_cool_adjust;/* remove the layer mean of the heating           */
  int    drag_v;          /* Rayleigh drag on v as well as u                */
  int    drag_zonal_avg;  /* Rayleigh drag acts on the zonal mean           */
  double tau_drag;        /* s, Rayleigh drag time; 0 or >= 1e20: none      */
  double prandtl;         /* Rayleigh drag rate prandtl/t_rad (with
                             Newtonian cooling on); 0: none                 */
  double nu6_nondim;      /* del^6 hyperviscosity, fraction of maximum     */
  double nu8_nondim;      /* del^8 hyperviscosity, fraction of maximum     */
  int    pgrad_traditional; /* 1: the reference's PGRAD_TRADITIONAL
                             pressure gradient (uv_pgrad_traditional,
                             src/shared/epic_timestep.c:806-884) in place of
                             the finite-volume one of its default build    */
  int    adv_mass;        /* advection scheme of hdry, and adv_nonmass of
                             theta (and of the unity field its correction
                             uses): the reference's list
                             (src/shared/epic_initial.c:146-149), 0 its
                             default predictor-corrector (Hsu, Konor and
                             Arakawa), 1 the monotonized fourth-order
                             centred scheme (Akima)                         */
  int    adv_nonmass;
  int    pv_scheme;       /* how PV is averaged onto the u and v grids: 0
                             Sadourny (1975), the reference's default
                             PV_SCHEME; 1 Arakawa and Lamb (1981), the other
                             value include/epic.h:79-86 offers
                             (include/epic_pv_schemes.h:34-47)             */
} epic_case_spec;

typedef struct {
  /* planet */
  char   planet_name[32];
  char   planet_type[16];
  int    planet_index;
  double re, rp, omega, cp, rgas, cpr, kappa, GM, J2, x_h2, kinvisc;
  /* grid scalars */
  int    ni, nj, nk, dt, jtp, k_sigma, k_sponge;
  int    newt_cool_on, newt_cool_adjust, drag_v, drag_zonal_avg;
  double dln, dlt, dy0, globe_lonbot, globe_lontop, globe_latbot, globe_lattop;
  double press0, pbot, ptop, zeta0, zeta1, hybrid_alpha, sigma_sigma;
  double topo_scale, prandtl, tau_drag, nudiv_nondim;
  double nu[EPIC_MAX_NU_ORDER + 1], nu_nondim[EPIC_MAX_NU_ORDER + 1];
  /* day-night relaxation (EPIC_PLANET_HOT_JUPITER; unused otherwise):
   * HEAT3 = cp (T_eq - T)/tau_rad with T_eq = T_ref(p) + dT(p) (max(0,
   * cos(lat) cos(lon)) - 1/4), T_ref the initial profile's temperature
   * 1/theta = 1/ref_theta_b + ref_a_over_kappa ((p/press0)^kappa -
   * (ref_pb/press0)^kappa) times (p/press0)^kappa, dT = teq_contrast for
   * p <= teq_p_contrast, falling linearly in ln p to 0 at pbot. */
  double teq_contrast, teq_p_contrast, tau_rad, ref_theta_b, ref_pb, ref_a_over_kappa;
  /* padded extents */
  int    Iadim, Jadim, Kadim, Nelem2d, Nelem3d;
  /* 1-D arrays (length 2*(ni+1)+1 for lon, 2*(nj+1)+1 for the J-arrays,
   * 2*(nk+1)+1 for the kk-arrays, nk+2 for h_min, k_sponge+1 for the sponge) */
  double *lon, *lat, *rln, *rlt, *f, *m, *n, *mn, *g;
  double *sigmatheta, *dsgth, *dsgth_inv, *p_ref, *theta_ref, *h_min, *t_sponge_inv;
  /* initial state, Nelem3d each (gz_surface: Nelem2d) */
  double *u, *v, *hdry, *theta, *p3, *u_spinup, *gz_surface;
  /* Newtonian cooling toward a sounding, for newt_cool_on on any planet but
   * Held-Suarez (a hot Jupiter without a sounding takes the day-night
   * relaxation above instead): the
   * reference's var.ntp, var.pdat [Pa] and var.tdat [K] (get_sounding,
   * src/shared/epic_funcs_diag.c:3160-3294) and var.n_t_cool,
   * var.t_cool_table (x = -log(p [Pa]) ascending, y = cooling time [s];
   * t_rad, src/shared/epic_sensible_heating.c:217-314).  0 and null when
   * the case has no such cooling. */
  int    ntp, n_t_cool;
  double *pdat, *tdat, *t_cool_x, *t_cool_y;
  /* restart state (a case read from an epic.nc file; null and zero for a
   * built case): the u and v Adams-Bashforth tendency planes IT_MINUS1 and
   * IT_MINUS2 (two Nelem3d planes each, in that order; null: the start-up
   * FLOAT_MAX flags), the model and start times [s] (var.model_time,
   * var.start_time), and the file's global attributes, which a case
   * written back carries unchanged where it has no value of its own */
  double *dudt_hist, *dvdt_hist;
  double model_time, start_time;
  void   *nc_header;
  /* 1: the traditional pressure gradient (epic_case_spec.pgrad_traditional) */
  int    pgrad_traditional;
  /* advection schemes (epic_case_spec.adv_mass, adv_nonmass) */
  int    adv_mass, adv_nonmass;
  /* PV averaging (epic_case_spec.pv_scheme) */
  int    pv_scheme;
} epic_case;

/* Build a case from a spec: metric arrays (set_lonlat/set_fmn/set_gravity/
 * set_dsgth restated), hybrid sigma-theta levels and a seeded initial
 * state.  Returns 0 or a negative error code. */
int  epic_case_build(const epic_case_spec *spec, epic_case *out);
void epic_case_free(epic_case *c);
/* Bytes of HBM one member of this size holds in the given precision mode. */
size_t epic_member_footprint(int ni, int nj, int nk, int precision_mode);

/* ------------------------------------------------------------------------
 * Native engine.
 * ---------------------------------------------------------------------- */
enum epic_precision { EPIC_FP64 = 0, EPIC_MIXED = 1 };

enum epic_field {
  EPIC_F_U = 0, EPIC_F_V, EPIC_F_HDRY, EPIC_F_THETA, EPIC_F_P3, EPIC_F_P2,
  EPIC_F_H3, EPIC_F_THETA2, EPIC_F_T2, EPIC_F_T3, EPIC_F_RHO2, EPIC_F_RHO3,
  EPIC_F_EXNER3, EPIC_F_GZ3, EPIC_F_MONT3, EPIC_F_PV3, EPIC_F_DIV_UV3,
  EPIC_F_W3, EPIC_F_HEAT3, EPIC_F_DZDT3, EPIC_F_DUDT0, EPIC_F_DUDT1,
  EPIC_F_DUDT2, EPIC_F_DVDT0, EPIC_F_DVDT1, EPIC_F_DVDT2, EPIC_F_GZ_SURFACE,
  EPIC_F_RI2,
  EPIC_F_COUNT
};

typedef struct epic_engine epic_engine;

/* Create an engine that may hold up to `max_members` members using at most
 * `hbm_budget` bytes of device memory (0 = no limit). */
epic_engine *epic_engine_create(int precision_mode, int max_members, size_t hbm_budget);
void         epic_engine_destroy(epic_engine *e);
/* Upload a case; primes the diagnostics (the reference's SYNC_DIAGS_ONLY).
 * Returns the member slot (>= 0), or a negative code (-2: no memory). */
int  epic_engine_add(epic_engine *e, const epic_case *c);
/* epic_case_build() + epic_engine_add() + epic_case_free() in one call. */
int  epic_engine_add_spec(epic_engine *e, const epic_case_spec *spec);
int  epic_engine_remove(epic_engine *e, int slot);
/* Adams-Bashforth time planes of a member: {IT_ZERO, IT_MINUS1, IT_MINUS2,
 * itime of the next step} (include/epic.h:303-304, epic_timestep.c:329-345). */
int  epic_engine_time_planes(epic_engine *e, int slot, int *out4);
int  epic_engine_set_time_planes(epic_engine *e, int slot, const int *in4);
/* Advance every active member by nsteps (STEP_PROGS_AND_SYNC_DIAGS each). */
int  epic_engine_step(epic_engine *e, int nsteps);
/* Advance only the listed members. */
int  epic_engine_step_members(epic_engine *e, const int *slots, int nslots, int nsteps);
/* Advance every member of one grid shape (one bucket) by nsteps, from a
 * graph of that bucket alone: a scheduler can give each shape its own
 * number of steps (no counterpart in the reference, which steps one
 * domain).  -1 if no member has that shape. */
int  epic_engine_step_shape(epic_engine *e, int ni, int nj, int nk, int nsteps);
/* Copy one field of one member into host memory (Nelem3d doubles, or
 * Nelem2d for EPIC_F_GZ_SURFACE; tendencies are indexed by time plane). */
int  epic_engine_read(epic_engine *e, int slot, int field, double *out);
/* Overwrite one prognostic field (U, V, HDRY, THETA) of a member. */
int  epic_engine_write(epic_engine *e, int slot, int field, const double *in);
long epic_engine_steps_done(epic_engine *e, int slot);
size_t epic_engine_bytes_in_use(epic_engine *e);
/* Wait for all queued device work; -4 if any member has raised a fatal
 * condition (non-monotone GZ3 = 1, pressure out of range = 2, spline = 4). */
int  epic_engine_sync(epic_engine *e);
/* Wait for all queued device work without checking the members. */
int  epic_engine_wait(epic_engine *e);
/* The sticky fatal-condition bits of one member (0: healthy). */
int  epic_engine_member_error(epic_engine *e, int slot);
/* Record device timing of each kernel over the next `nsteps`; fills
 * names/ms arrays (up to cap entries).  Returns the entry count. */
int  epic_engine_profile(epic_engine *e, int nsteps, char (*names)[48], double *ms_total,
                         int *launches, int cap);
/* Exact restart (the engine's counterpart of the reference's restart file,
 * src/shared/epic_funcs_io.c:730-2257, in a binary layout of its own):
 * write every array of a member that its next step reads, with its
 * Adams-Bashforth history and time planes, to `path`; a member restored
 * from that file with the same case continues bit for bit as the original
 * would.  restore returns the new slot. */
int  epic_engine_checkpoint(epic_engine *e, int slot, const char *path);
int  epic_engine_restore(epic_engine *e, const epic_case *c, const char *path);
/* epic_case_build() + epic_engine_restore() + epic_case_free() in one call. */
int  epic_engine_restore_spec(epic_engine *e, const epic_case_spec *spec, const char *path);

/* ---- epic.nc: the reference's model-state file (netCDF-3) ----------------
 * The layout var_write writes and var_read reads for ALL_DATA
 * (src/shared/epic_funcs_io.c:730-2257): the planet and grid scalars as
 * global attributes; u, v, p3, theta, u_spinup at one time record; the
 * Adams-Bashforth tendency planes du_IT_MINUS1dt, du_IT_MINUS2dt,
 * dv_IT_MINUS1dt, dv_IT_MINUS2dt; gz_surface on terrestrial planets.
 *
 * epic_nc_read: a case from a file, its metric arrays formed from the
 * scalars as the reference's set_lonlat/set_fmn/set_gravity/set_dsgth/
 * set_sponge form them, the tendency planes in dudt_hist/dvdt_hist and
 * model_time/start_time set; epic_engine_add() then continues the run
 * (HDRY formed from P3 as epic_main forms it).  Returns 0, -1 unreadable,
 * -2 not a netCDF classic file, -3 an attribute or variable missing or
 * malformed, -4 a configuration the GPU step does not implement
 * (epic_nc_error() says which).  Free the case with epic_case_free().
 *
 * epic_nc_write: the state of one member (padded Nelem3d arrays; null
 * tendency planes are written as the start-up FLOAT_MAX flags) with the
 * header of its case; the global attributes of the file a case was read
 * from pass through where the case has no value of its own.
 * epic_engine_write_nc: the same for an engine member, its state read from
 * the device after a sync; model_time is the case's plus the member's steps
 * times dt. */
typedef struct {
  const double *u, *v, *p3, *theta, *u_spinup, *gz_surface;
  const double *dudt_m1, *dudt_m2, *dvdt_m1, *dvdt_m2;
  double model_time;  /* s, var.model_time */
  double l_s_deg;     /* solar longitude L_s, degrees */
} epic_nc_state;
int  epic_nc_read(const char *path, epic_case *out);
int  epic_nc_write(const char *path, const epic_case *c, const epic_nc_state *state);
int  epic_engine_write_nc(epic_engine *e, int slot, const epic_case *c, const char *path);
const char *epic_nc_error(void);
void epic_nc_header_free(void *header);
/* Autotuner (no counterpart in the reference, whose loops have no launch
 * shapes): for each bucket, time every candidate launch shape of the tunable
 * kernels (columns per column tile of set_p2etc, gz3, calc_w and the
 * vertical advection pass, and the tile's threads; rows per
 * zonal-hyperviscosity block, per meridional-advection tile and per column
 * tile of the del^4 meridional second differences) over
 * `steps` steps of the bucket's own members and keep the fastest.  The
 * members' state is copied aside and put back, so the call changes no
 * result and no state; it needs room for one copy of every member.  Writes
 * one line per choice into report (up to cap bytes).  0 or a negative code. */
int  epic_engine_autotune(epic_engine *e, int steps, char *report, int cap);
/* One launch shape set by hand for every bucket (knob: 0 columns per
 * set_p2etc tile, 1 gz3, 2 calc_w, 3 vertical advection, 4 rows per
 * zonal-hyperviscosity block, 5 rows per meridional-advection tile, 6 rows
 * per column tile of the del^4 meridional second differences (8, 16, 32); a
 * column-tile value is C, 1000 n + C for blocks of 256 n threads, n = 2
 * or 4, or 10000 + C for one thread per level of each of the C columns; 0
 * restores the built-in choice).  Values outside the candidates the
 * autotuner tries, or whose LDS would not fit, are refused (-1).  For tests
 * that every shape gives the same bits. */
int  epic_engine_set_shape(epic_engine *e, int knob, int value);
const char *epic_engine_last_error(epic_engine *e);

/* ------------------------------------------------------------------------
 * Allocation-free step.  No counterpart in the reference: it is for callers
 * that cannot allocate device memory (code under HIP graph capture, tight
 * serving loops, harnesses that forbid hipMalloc).  It does what
 * epic_engine_add (action EPIC_SYNC_DIAGS_ONLY: prime the diagnostics) or
 * epic_engine_step (action EPIC_STEP_PROGS_AND_SYNC_DIAGS: nsteps steps)
 * does for one FP64 member, with the same kernels and the same results, on
 * fields the caller holds in device memory: fields[f] is the device array of
 * epic_field f (Nelem3d doubles; Nelem2d for EPIC_F_GZ_SURFACE; the three
 * u and the three v tendency planes consecutive, fields[EPIC_F_DUDT0 + p] ==
 * fields[EPIC_F_DUDT0] + p * Nelem3d).  RI2 is formed at the end of the call.
 * The case supplies the metrics, tables and u_spinup (its state arrays are
 * not read).  it4: {IT_ZERO, IT_MINUS1, IT_MINUS2, itime} in and out.
 * The library's working memory is a fixed static device array of
 * EPIC_NOALLOC_SCRATCH_BYTES per GPU (room for Nelem3d up to about 780,000,
 * e.g. 256x128x20), cleared at the start of every call; larger members are
 * refused (-2).  The call queues its work on `stream` (a hipStream_t; NULL:
 * the default stream) and waits for it.  Calls that overlap on one GPU share
 * the array and race.  Returns 0, or a negative code with the reason in
 * err (-1 bad arguments or configuration, -2 too large, -4 a fatal
 * condition of the step, -5 a HIP error).
 */
#define EPIC_NOALLOC_SCRATCH_BYTES ((size_t)96 << 20)
int  epic_step_noalloc(const epic_case *c, double *const *fields, int *it4, int action, int nsteps,
                       void *stream, char *err, size_t errcap);

/* Measurement helper with no counterpart in the reference: time a plain
 * device-to-device streaming copy of nbytes (reads nbytes, writes nbytes),
 * the measured copy peak beside the 8 TB/s specification. */
int  epic_copy_probe(size_t nbytes, int reps, double *seconds_per_copy);

#ifdef __cplusplus
}
#endif
#endif /* EPIC_AMD_H */
