// engine.cpp -- host side of the MI355X EPIC engine: member memory, the
// run-once tables, shape buckets, the step and diagnostic-sync sequences,
// hipGraph capture of the step, and per-kernel event timing.
//
// Tables (filter gains, FFT twiddles, d2dy2 weights, hyperviscosity
// coefficients) are the callee-internal statics of the reference and are
// formed here on the host with the host libm, with the same expressions as
// their producers:
//   zonal_filter init       src/shared/epic_adjust.c:148-242
//   four1/realft twiddles   src/shared/epic_funcs_util.c:4023-4135
//   d2dy2 init              src/shared/epic_subgrid.c:1049-1140
//   set_max_nu / coef[]     src/shared/epic_subgrid.c:82-158, 627-650
//   divergence damping      src/shared/epic_subgrid.c:1814-1847
//   get_kin mn_inv          src/shared/epic_funcs_diag.c:2718-2732
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <algorithm>
#include <array>
#include <map>
#include <string>
#include <vector>

#include "../../include/epic_amd.h"
#include "engine_internal.h"
#include "kernels.h"
#include "tables.h"

#define HIPCHK(x)                                                                      \
  do {                                                                                 \
    hipError_t e_ = (x);                                                               \
    if (e_ != hipSuccess) {                                                            \
      snprintf(E->err, sizeof(E->err), "%s: %s", #x, hipGetErrorString(e_));           \
      return -5;                                                                       \
    }                                                                                  \
  } while (0)

namespace epic {

bool g_prof_on = false;
struct ProfRec {
  hipEvent_t a, b;
  const char *name;
};
static std::vector<ProfRec> g_prof;
static std::vector<hipEvent_t> g_evpool;
static size_t g_evnext = 0;
static hipEvent_t ev_get() {
  if (g_evnext < g_evpool.size()) return g_evpool[g_evnext++];
  hipEvent_t e;
  (void)hipEventCreate(&e);
  g_evpool.push_back(e);
  g_evnext++;
  return e;
}
void prof_begin(hipStream_t s, const char *name) {
  ProfRec r;
  r.a = ev_get();
  r.b = nullptr;
  r.name = name;
  (void)hipEventRecord(r.a, s);
  g_prof.push_back(r);
}
void prof_end(hipStream_t s) {
  g_prof.back().b = ev_get();
  (void)hipEventRecord(g_prof.back().b, s);
}

}  // namespace epic

using namespace epic;

struct Member {
  int active;
  MemberDesc h;         // host copy of the descriptor (device pointers)
  void *blob;           // one device allocation for everything
  size_t bytes;
  MemberDyn *dyn;       // device
  double nu4, nudiv;
  int adjust;
  long steps;
};

struct epic_engine {
  int precision, max_members;
  size_t budget, in_use;
  std::vector<Member> m;
  MemberDesc *dD;       // device descriptor array (bucket order)
  std::vector<int> order;  // bucket order -> slot
  std::vector<Bucket> buckets;
  // launch shapes chosen by epic_engine_autotune, by (ni, nj, nk, members)
  std::map<std::array<int, 4>, std::array<int, KN_COUNT>> tuned;
  bool dirty;
  hipStream_t stream;
  // a second stream (default priority) for the step's one independent
  // branch, forked from and joined back into `stream` by two events
  SideStream side;
  hipGraph_t graph;
  hipGraphExec_t gexec;
  bool have_graph;
  // one graph per bucket (epic_engine_step_shape), captured when first used
  std::vector<hipGraph_t> bgraph;
  std::vector<hipGraphExec_t> bgexec;
  char err[512];
};

static size_t align256(size_t x) { return (x + 255) & ~(size_t)255; }

// ------------------------------------------------------------------ engine
extern "C" epic_engine *epic_engine_create(int precision_mode, int max_members, size_t hbm_budget) {
  epic_engine *E = new epic_engine();
  E->precision = precision_mode;
  E->max_members = max_members > 0 ? max_members : 1;
  E->budget = hbm_budget;
  E->in_use = 0;
  E->m.resize(E->max_members);
  for (auto &x : E->m) x.active = 0;
  E->dD = nullptr;
  E->dirty = true;
  E->have_graph = false;
  E->err[0] = 0;
  if (hipStreamCreate(&E->stream) != hipSuccess) {
    delete E;
    return nullptr;
  }
  if (hipStreamCreate(&E->side.s2) != hipSuccess ||
      hipEventCreateWithFlags(&E->side.fork, hipEventDisableTiming) != hipSuccess ||
      hipEventCreateWithFlags(&E->side.join, hipEventDisableTiming) != hipSuccess) {
    (void)hipStreamDestroy(E->stream);
    delete E;
    return nullptr;
  }
  if (hipMalloc((void **)&E->dD, sizeof(MemberDesc) * E->max_members) != hipSuccess) {
    (void)hipStreamDestroy(E->stream);
    delete E;
    return nullptr;
  }
  return E;
}

extern "C" const char *epic_engine_last_error(epic_engine *E) { return E ? E->err : "no engine"; }

static void drop_graph(epic_engine *E) {
  if (E->have_graph) {
    (void)hipGraphExecDestroy(E->gexec);
    (void)hipGraphDestroy(E->graph);
    E->have_graph = false;
  }
  for (size_t i = 0; i < E->bgexec.size(); i++)
    if (E->bgexec[i]) {
      (void)hipGraphExecDestroy(E->bgexec[i]);
      (void)hipGraphDestroy(E->bgraph[i]);
    }
  E->bgexec.clear();
  E->bgraph.clear();
}

extern "C" void epic_engine_destroy(epic_engine *E) {
  if (!E) return;
  (void)hipStreamSynchronize(E->stream);
  drop_graph(E);
  for (auto &x : E->m)
    if (x.active) (void)hipFree(x.blob);
  (void)hipFree(E->dD);
  (void)hipStreamSynchronize(E->side.s2);
  (void)hipEventDestroy(E->side.fork);
  (void)hipEventDestroy(E->side.join);
  (void)hipStreamDestroy(E->side.s2);
  (void)hipStreamDestroy(E->stream);
  delete E;
}

// Members share a bucket when their grids are the same; in the mixed mode
// also only when both or neither apply a hyperviscosity order above del^4:
// a del^4 member's FP32 zonal means are formed by the fused del^4 row
// kernel, in its order, whatever the other members of its grid apply.
static int hv_class(const epic_engine *E, const MemberDesc &h) {
  return (E->precision == EPIC_MIXED && hv_passes(h) > 2) ? 1 : 0;
}

static int rebuild_buckets(epic_engine *E) {
  E->order.clear();
  for (int s = 0; s < E->max_members; s++)
    if (E->m[s].active) E->order.push_back(s);
  std::stable_sort(E->order.begin(), E->order.end(), [&](int a, int b) {
    const MemberDesc &x = E->m[a].h, &y = E->m[b].h;
    if (x.ni != y.ni) return x.ni < y.ni;
    if (x.nj != y.nj) return x.nj < y.nj;
    if (x.nk != y.nk) return x.nk < y.nk;
    return hv_class(E, x) < hv_class(E, y);
  });
  std::vector<MemberDesc> hd;
  for (int s : E->order) hd.push_back(E->m[s].h);
  if (!hd.empty())
    HIPCHK(hipMemcpyAsync(E->dD, hd.data(), sizeof(MemberDesc) * hd.size(), hipMemcpyHostToDevice, E->stream));
  E->buckets.clear();
  for (size_t i = 0; i < hd.size();) {
    size_t j = i;
    while (j < hd.size() && hd[j].ni == hd[i].ni && hd[j].nj == hd[i].nj && hd[j].nk == hd[i].nk &&
           hv_class(E, hd[j]) == hv_class(E, hd[i]))
      j++;
    Bucket b;
    b.D = E->dD;
    b.mb = (int)i;
    b.nm = (int)(j - i);
    b.ni = hd[i].ni;
    b.nj = hd[i].nj;
    b.nk = hd[i].nk;
    b.mixed = (E->precision == EPIC_MIXED);
    b.any_gas = b.any_hs = b.any_ray = 0;
    b.hv_passes = 0;
    b.adv_schemes = 0;
    for (size_t k = i; k < j; k++) {
      const MemberDesc &h = hd[k];
      b.hv_passes = std::max(b.hv_passes, hv_passes(h));
      b.any_gas |= (h.is_gas != 0);
      b.any_hs |= (h.is_hs != 0);
      b.any_ray |= ((h.newt_cool_on && h.prandtl > 0.) || (h.tau_drag > 0. && h.tau_drag < 1.e+20));
      b.adv_schemes |= adv_schemes(h);
    }
    auto t = E->tuned.find({b.ni, b.nj, b.nk, b.nm});
    for (int k = 0; k < KN_COUNT; k++) b.knob[k] = (t == E->tuned.end()) ? 0 : t->second[k];
    E->buckets.push_back(b);
    i = j;
  }
  E->dirty = false;
  drop_graph(E);
  return 0;
}

// store_pgrad_vars (src/shared/epic_funcs_diag.c:3595-3950)
// set_p2etc then store_pgrad_vars.  full = 1: everything.  full = 0 (the
// pair that follows the advection inside the step, with DONT_UPDATE_THETA):
// of what it writes, the step reads only P3, h, T3, RHO3, GZ_SURFACE and GZ3
// (through the filtered GZ3 copy of the 3-D filter) before the second pair
// writes every output again, so the first forms those only and leaves out
// P2, H3, EXNER3, T2, RHO2 and DZDT3 -- the step's results are the same.
// With `side` (full = 1, gas giants), gz_from_u runs on the second stream
// beside k_thermo: it reads U, P3 and theta of the bottom interface (forming
// RHO3 there itself) and writes its own row buffer, k_thermo writes the
// per-cell diagnostics; k_gz3 waits for both.
static bool fork_worth(const Bucket &b) { return (size_t)b.ni * (b.nj + 1) * b.nk * b.nm >= ((size_t)1 << 18); }
static void rec_p2etc_pgrad_vars(hipStream_t s, const Bucket &b, int full, const SideStream *side) {
  L_set_p2etc(s, b, full);  // full = 0: with store_pgrad_vars' T3, RHO3
  const bool fork = full && side && b.any_gas && fork_worth(b);
  if (fork) {
    (void)hipEventRecord(side->fork, s);
    (void)hipStreamWaitEvent(side->s2, side->fork, 0);
    L_gz_from_u(side->s2, b);
    (void)hipEventRecord(side->join, side->s2);
  }
  if (full) L_thermo(s, b);
  if (fork) (void)hipStreamWaitEvent(s, side->join, 0);
  else L_gz_from_u(s, b);
  L_gz3(s, b);
  if (full) L_dzdt3(s, b);
}

// store_diag (src/shared/epic_funcs_diag.c:3963-4053); the filtered copies
// of V and U are formed row by row inside the PV and divergence blocks.
// rotate: the step's closing rotation of the time planes rides along.
static void rec_store_diag(hipStream_t s, const Bucket &b, int rotate) { L_store_diag(s, b, rotate); }

static void rec_sync(hipStream_t s, const Bucket &b, bool adjust, const SideStream *side) {
  rec_p2etc_pgrad_vars(s, b, 1, side);
  rec_store_diag(s, b, 0);
  L_heating(s, b, adjust, false);
  L_calc_w(s, b, 0);
}

// timestep(STEP_PROGS_AND_SYNC_DIAGS) in the order of
// src/shared/epic_timestep.c:104-327.  With `side`, the hyperviscosity of
// hdry and theta runs on the second stream beside uv_core and uv_drag: the
// two branches read what calc_w leaves and write disjoint arrays (the
// tendencies and scratch copy 0; hdry, theta and scratch copies 1, 2), so
// the results are those of the reference's order.
static void rec_step(hipStream_t s, const Bucket &b, bool hv, bool divd, bool adjust, const SideStream *side) {
  L_calc_w(s, b, 1);  // source_sink + calc_w, and the step's tendency planes cleared
  // hdry and theta take their hyperviscosity together (neither reads the
  // other, and restore_mass touches hdry only)
  const FieldSet hth = {2, {F_HDRY, F_THETA}, {2, 3}, 1};
  // forked only for buckets of 2^18 grid points or more: at 256x128x20 the
  // branch saves 3% of the FP64 step, at 64x32x10 the cross-queue join
  // costs more than the branch (+9%)
  const bool fork = hv && side && fork_worth(b);
  // (a failure of the event calls shows in the caller's hipGetLastError,
  // as a failed launch does)
  if (fork) {
    (void)hipEventRecord(side->fork, s);
    (void)hipStreamWaitEvent(side->s2, side->fork, 0);
    L_hv(side->s2, b, hth);
    (void)hipEventRecord(side->join, side->s2);
  }
  L_uv_core(s, b);
  L_uv_drag(s, b);
  if (fork) {
    (void)hipStreamWaitEvent(s, side->join, 0);
  } else if (hv) {
    L_hv(s, b, hth);
  }
  L_advection(s, b);  // restore_mass first, inside
  rec_p2etc_pgrad_vars(s, b, 0, side);
  // filtered, monotone copy of GZ3
  L_gz_filtercopy(s, b);
  L_gz_mono(s, b);
  // the four 3-D filters in one pass (each reads only its own field and the
  // filtered GZ3 copy); restore_mass, which touches hdry only and follows
  // them, is applied to the filtered hdry as it is stored
  const FieldSet f4 = {4, {F_DUDT_T, F_DVDT_T, F_HDRY, F_THETA}, {0, 1, 2, 3}};
  L_filter3(s, b, f4);
  rec_p2etc_pgrad_vars(s, b, 1, side);
  L_pgrad(s, b);
  if (divd) L_divdamp(s, b);
  L_ab(s, b);
  if (hv) {
    const FieldSet uv = {2, {F_U, F_V}, {0, 1}};
    L_hv(s, b, uv);
  }
  // store_diag, then the time planes rotated (in the same launch: nothing
  // after the Adams-Bashforth update reads them)
  rec_store_diag(s, b, 1);
  L_heating(s, b, adjust, true);  // the relaxation itself was formed with T3 (thermo_cell)
}

void member_record(hipStream_t s, const Bucket &b, bool sync_only, bool hv, bool divd, bool adjust) {
  if (sync_only) rec_sync(s, b, adjust, nullptr);
  else rec_step(s, b, hv, divd, adjust, nullptr);
}

static void bucket_flags(epic_engine *E, const Bucket &b, bool &hv, bool &divd, bool &adj) {
  hv = divd = adj = false;
  for (int i = 0; i < b.nm; i++) {
    const Member &x = E->m[E->order[b.mb + i]];
    hv |= hv_passes(x.h) > 0;
    divd |= x.nudiv != 0.;
    adj |= (x.adjust != 0);
  }
}

static void record_all(epic_engine *E, bool sync_only, const std::vector<int> *only) {
  for (const Bucket &b : E->buckets) {
    bool hv, divd, adj;
    bucket_flags(E, b, hv, divd, adj);
    if (only) {
      // run member by member for a subset (no graph)
      for (int i = 0; i < b.nm; i++) {
        int slot = E->order[b.mb + i];
        if (std::find(only->begin(), only->end(), slot) == only->end()) continue;
        Bucket one = b;
        one.mb = b.mb + i;
        one.nm = 1;
        if (sync_only) rec_sync(E->stream, one, adj, &E->side);
        else rec_step(E->stream, one, hv, divd, adj, &E->side);
      }
    } else {
      if (sync_only) rec_sync(E->stream, b, adj, &E->side);
      else rec_step(E->stream, b, hv, divd, adj, &E->side);
    }
  }
}

extern "C" size_t epic_engine_bytes_in_use(epic_engine *E) { return E ? E->in_use : 0; }

// Upload a case into a free slot.  hdry_src: the prognostic h of the caller
// (nullptr: reconstitute it from p3 as epic_main does).  prime: run the
// SYNC_DIAGS_ONLY sequence for the new member.
// Host copies of the device's fcmp_d / table_place_s (dev_common.h), for
// searches whose table and argument are the same in every column: the
// index (and interval width) is formed once here, by the same operations.
static int fcmp_host(double x1, double x2) {
  int e;
  frexp(fabs(x1) > fabs(x2) ? x1 : x2, &e);
  double delta = ldexp(2.220446049250313e-16, e), d = x1 - x2;
  if (d > delta) return 1;
  if (d < -delta) return -1;
  return 0;
}
static int table_place_s_host(int n, const double *x, int st, double xx, double *dx) {
  int il = -1, iu = n;
  int ascend = (x[(n - 1) * st] >= x[0]);
  int lo = fcmp_host(xx, x[0]), hi = fcmp_host(xx, x[(n - 1) * st]);
  if (ascend) {
    if (lo <= 0) { *dx = x[st] - x[0]; return 0; }
    if (hi >= 0) { *dx = x[(n - 1) * st] - x[(n - 2) * st]; return n - 2; }
  } else {
    if (lo >= 0) { *dx = x[st] - x[0]; return 0; }
    if (hi <= 0) { *dx = x[(n - 1) * st] - x[(n - 2) * st]; return n - 2; }
  }
  while (iu - il > 1) {
    int im = (iu + il) >> 1;
    if ((fcmp_host(xx, x[im * st]) >= 0) == ascend) il = im; else iu = im;
  }
  *dx = x[(il + 1) * st] - x[il * st];
  return il;
}

// Whether the engine implements a case; writes the reason into err (cap
// bytes) and returns -1 when it does not.
int member_check(const epic_case *c, int precision, char *err, size_t cap) {
  if (c->ni > 1024 || c->ni < 8 || (c->ni & (c->ni - 1)) || c->nk > 64 || c->nk < 4) {
    snprintf(err, cap, "unsupported grid %dx%dx%d", c->ni, c->nj, c->nk);
    return -1;
  }
  if (strcmp(c->planet_type, "gas-giant") == 0 && (c->jtp < 0 || c->jtp > c->nj)) {
    snprintf(err, cap, "tie-point row jtp=%d outside 0..%d", c->jtp, c->nj);
    return -1;
  }
  // Newtonian cooling on any planet but Held-Suarez relaxes toward the
  // sounding with the tabulated cooling time; the hot Jupiter without a
  // sounding has its own day-night relaxation
  if (c->newt_cool_on && strcmp(c->planet_name, "held_suarez") != 0 &&
      (c->ntp > 0 || strcmp(c->planet_name, "hot_jupiter") != 0)) {
    bool ok = c->ntp >= 2 && c->n_t_cool >= 2 && c->pdat && c->tdat && c->t_cool_x && c->t_cool_y;
    for (int k = 0; ok && k < c->ntp; k++) ok = c->pdat[k] > 0. && (k == 0 || -log(c->pdat[k]) > -log(c->pdat[k - 1]));
    for (int k = 1; ok && k < c->n_t_cool; k++) ok = c->t_cool_x[k] > c->t_cool_x[k - 1];
    if (!ok) {
      snprintf(err, cap, "newt_cool_on (planet %s) needs the sounding (ntp >= 2 pressures pdat, falling, and "
               "temperatures tdat) and the cooling-time table (n_t_cool >= 2 points, t_cool_x = -log p rising)",
               c->planet_name);
      return -1;
    }
  }
  if (c->nu[4] < 0. || c->nu[6] < 0. || c->nu[8] < 0.) {
    snprintf(err, cap, "negative hyperviscosity coefficient");
    return -1;
  }
  if (c->k_sponge < 0 || c->k_sponge >= c->nk) {
    snprintf(err, cap, "k_sponge=%d outside 0..nk-1", c->k_sponge);
    return -1;
  }
  if (precision != EPIC_FP64 && precision != EPIC_MIXED) {
    snprintf(err, cap, "unknown precision mode %d", precision);
    return -1;
  }
  return 0;
}

// The arrays of one member's descriptor in their fixed order, with the
// scalars of the descriptor filled in.  Each part names the descriptor's
// pointer, its size, its initial contents (null: zero) and, for the arrays a
// caller of the engine reads and writes, the field id (EPIC_F_*; the tendency
// planes as EPIC_F_DUDT0 / EPIC_F_DVDT0, three planes each); working memory
// and tables have field -1.  The host arrays that parts point into live in
// `hold` until the copies are done.
void member_plan(const epic_case *c, bool mixed, const epic_tables &T, MemberDesc &d, std::vector<MemberPart> &parts,
                 MemberHold &hold) {
  memset(&d, 0, sizeof(d));
  parts.clear();
  const size_t N3 = (size_t)c->Nelem3d, N2 = (size_t)c->Nelem2d;
  const int nJ = 2 * (c->nj + 1) + 1, nK = 2 * (c->nk + 1) + 1;
  auto add = [&](void *pp, size_t bytes, const void *src, int field = -1) {
    parts.push_back(MemberPart{(void **)pp, bytes, src, field});
  };
  const double *nul = nullptr;
  add(&d.lat, sizeof(double) * nJ, c->lat);
  add(&d.lon, sizeof(double) * (2 * (c->ni + 1) + 1), c->lon);
  add(&d.m, sizeof(double) * nJ, c->m);
  add(&d.n, sizeof(double) * nJ, c->n);
  add(&d.mn, sizeof(double) * nJ, c->mn);
  add(&d.f, sizeof(double) * nJ, c->f);
  add(&d.g, sizeof(double) * nJ, c->g);
  add(&d.rln, sizeof(double) * nJ, c->rln);
  add(&d.rlt, sizeof(double) * nJ, c->rlt);
  add(&d.sgth, sizeof(double) * nK, c->sigmatheta);
  add(&d.dsgth, sizeof(double) * nK, c->dsgth);
  add(&d.dsgth_inv, sizeof(double) * nK, c->dsgth_inv);
  add(&d.h_min, sizeof(double) * (c->nk + 2), c->h_min);
  // cos(lat) and cos(lat)^4 per h-row and cos(lon) per column for the
  // Newtonian relaxations, with the host libm like the oracle's (every case
  // carries them; small).  Rows 0..nj and columns 0..ni are defined; the
  // last entries stay 0.
  hold.cl_h.assign(c->nj + 2, 0.);
  hold.cl4_h.assign(c->nj + 2, 0.);
  hold.clon_h.assign(c->ni + 2, 0.);
  for (int J = 0; J <= c->nj; J++) {
    hold.cl_h[J] = cos(c->lat[2 * J + 1] * (M_PI / 180.));
    hold.cl4_h[J] = pow(hold.cl_h[J], 4.);
  }
  for (int I = 0; I <= c->ni; I++) hold.clon_h[I] = cos(c->lon[2 * I + 1] * (M_PI / 180.));
  // THETA2 above the sigma region interpolates theta against sigmatheta:
  // the interval of layer K in the interface table (set_p2etc,
  // src/shared/epic_funcs_diag.c:3522-3562) is the same in every column
  hold.t2_kay.assign(c->nk + 2, 0);
  hold.t2_h.assign(c->nk + 2, 1.);
  {
    const int nk = c->nk, nkay = nk - (c->k_sigma - 1) + 1;
    const double *tx = c->sigmatheta + 2 * nk + 1;
    for (int K = c->k_sigma; K <= nk; K++)
      hold.t2_kay[K] = table_place_s_host(nkay, tx, -2, c->sigmatheta[2 * K], &hold.t2_h[K]);
  }
  add(&d.t2_kay, sizeof(int) * (c->nk + 2), hold.t2_kay.data());
  add(&d.t2_h, sizeof(double) * (c->nk + 2), hold.t2_h.data());
  hold.minv_v.assign(c->nj + 2, 0.);
  hold.ninv_h.assign(c->nj + 2, 0.);
  for (int J = 0; J <= c->nj + 1; J++) hold.minv_v[J] = (fabs(c->lat[2 * J]) == 90.) ? 0. : 1. / c->m[2 * J];
  for (int J = 0; J <= c->nj; J++) hold.ninv_h[J] = 1. / c->n[2 * J + 1];
  add(&d.minv_v, sizeof(double) * (c->nj + 2), hold.minv_v.data());
  add(&d.ninv_h, sizeof(double) * (c->nj + 2), hold.ninv_h.data());
  add(&d.coslat_h, sizeof(double) * (c->nj + 2), hold.cl_h.data());
  add(&d.coslat4_h, sizeof(double) * (c->nj + 2), hold.cl4_h.data());
  add(&d.coslon_h, sizeof(double) * (c->ni + 2), hold.clon_h.data());
  add(&d.t_sponge, sizeof(double) * (c->k_sponge + 1), c->t_sponge_inv);
  // the sounding and cooling-time tables (one entry when there are none)
  const size_t ns = T.n_snd ? T.n_snd : 1, nc = T.n_cool ? T.n_cool : 1;
  add(&d.snd_x, sizeof(double) * ns, T.snd[0]);
  add(&d.snd_y, sizeof(double) * ns, T.snd[1]);
  add(&d.snd_d, sizeof(double) * ns, T.snd[2]);
  add(&d.cool_x, sizeof(double) * nc, T.cool[0]);
  add(&d.cool_y, sizeof(double) * nc, T.cool[1]);
  add(&d.cool_d, sizeof(double) * nc, T.cool[2]);
  add(&d.mn_inv, sizeof(double) * nJ, T.mn_inv);
  add(&d.hl_h, sizeof(double) * T.n_hl, T.hl_h);
  add(&d.hl_pv, sizeof(double) * T.n_hl, T.hl_pv);
  add(&d.hcx, sizeof(double) * (c->nj + 2), T.hc[0]);
  add(&d.hcy, sizeof(double) * (c->nj + 2), T.hc[1]);
  add(&d.hcz, sizeof(double) * (c->nj + 2), T.hc[2]);
  add(&d.vcx, sizeof(double) * (c->nj + 2), T.vc[0]);
  add(&d.vcy, sizeof(double) * (c->nj + 2), T.vc[1]);
  add(&d.vcz, sizeof(double) * (c->nj + 2), T.vc[2]);
  add(&d.tw_f, sizeof(double) * T.n_tw, T.twf);
  add(&d.tw_i, sizeof(double) * T.n_tw, T.twi);
  add(&d.tr_f, sizeof(double) * T.n_tr, T.trf);
  add(&d.tr_i, sizeof(double) * T.n_tr, T.tri);
  // FP32 copies of the twiddles for the filters of FP32 arrays (mixed mode)
  hold.twf32.assign(T.twf, T.twf + T.n_tw);
  hold.twi32.assign(T.twi, T.twi + T.n_tw);
  hold.trf32.assign(T.trf, T.trf + T.n_tr);
  hold.tri32.assign(T.tri, T.tri + T.n_tr);
  add(&d.tw_f32, sizeof(float) * T.n_tw, hold.twf32.data());
  add(&d.tw_i32, sizeof(float) * T.n_tw, hold.twi32.data());
  add(&d.tr_f32, sizeof(float) * T.n_tr, hold.trf32.data());
  add(&d.tr_i32, sizeof(float) * T.n_tr, hold.tri32.data());
  add(&d.rows_h, sizeof(int) * (c->nj + 2), T.rows_h);
  add(&d.rows_v, sizeof(int) * (c->nj + 2), T.rows_v);
  hold.flag_h.assign(c->nj + 2, 0);
  hold.flag_v.assign(c->nj + 2, 0);
  for (int r = 0; r < T.nrow_h; r++) hold.flag_h[T.rows_h[r]] = 1;
  for (int r = 0; r < T.nrow_v; r++) hold.flag_v[T.rows_v[r]] = 1;
  add(&d.rflag_h, sizeof(int) * (c->nj + 2), hold.flag_h.data());
  add(&d.rflag_v, sizeof(int) * (c->nj + 2), hold.flag_v.data());
  add(&d.U, sizeof(double) * N3, c->u, EPIC_F_U);
  add(&d.V, sizeof(double) * N3, c->v, EPIC_F_V);
  add(&d.THETA, sizeof(double) * N3, c->theta, EPIC_F_THETA);
  add(&d.P3, sizeof(double) * N3, c->p3, EPIC_F_P3);
  add(&d.USPIN, sizeof(double) * N3, c->u_spinup);
  add(&d.GZS, sizeof(double) * N2, c->gz_surface, EPIC_F_GZ_SURFACE);
  struct { double **p; int f; } fields[] = {
      {&d.HDRY, EPIC_F_HDRY},     {&d.P2, EPIC_F_P2},       {&d.THETA2, EPIC_F_THETA2}, {&d.H3, EPIC_F_H3},
      {&d.T2, EPIC_F_T2},         {&d.T3, EPIC_F_T3},       {&d.RHO2, EPIC_F_RHO2},     {&d.RHO3, EPIC_F_RHO3},
      {&d.EXNER3, EPIC_F_EXNER3}, {&d.GZ3, EPIC_F_GZ3},     {&d.MONT3, EPIC_F_MONT3},   {&d.HEAT3, EPIC_F_HEAT3},
      {&d.PV3, EPIC_F_PV3},       {&d.DIV, EPIC_F_DIV_UV3}, {&d.W3, EPIC_F_W3},         {&d.DZDT3, EPIC_F_DZDT3},
      {&d.RI2, EPIC_F_RI2},       {&d.gz, -1},              {&d.unity, -1},             {&d.old, -1}};
  for (auto &p : fields) add(p.p, sizeof(double) * N3, nul, p.f);
  // scratch: three copies each, one per field of the batched advection
  for (auto p : {&d.sA, &d.sB, &d.sC, &d.sD}) add(p, sizeof(double) * 3 * N3, nul);
  if (mixed) {
    add(&d.DUDT, sizeof(double) * N3, nul);
    add(&d.DVDT, sizeof(double) * N3, nul);
    add(&d.DUDTf, sizeof(float) * 3 * N3, nul);
    add(&d.DVDTf, sizeof(float) * 3 * N3, nul);
  } else {
    add(&d.DUDT, sizeof(double) * 3 * N3, nul, EPIC_F_DUDT0);
    add(&d.DVDT, sizeof(double) * 3 * N3, nul, EPIC_F_DVDT0);
  }
  // PCHIP tables of the batched 3-D filter, one set per field
  add(&d.spl, sizeof(double) * 4 * (size_t)(c->nj + 2) * 3 * (c->nk + 2), nul);
  add(&d.rowbuf, sizeof(double) * 8 * (c->nj + 3), nul);
  add(&d.dyn, sizeof(MemberDyn), nul);

  // scalars
  d.ni = c->ni; d.nj = c->nj; d.nk = c->nk; d.Ia = c->Iadim; d.N2 = c->Nelem2d; d.N3 = c->Nelem3d;
  d.ks = c->k_sigma; d.k_sponge = c->k_sponge;
  d.is_hs = (strcmp(c->planet_name, "held_suarez") == 0);
  d.is_gas = (strcmp(c->planet_type, "gas-giant") == 0);
  d.is_tab = epic_case_table_cooling(c);
  d.is_hj = (strcmp(c->planet_name, "hot_jupiter") == 0) && !d.is_tab;
  d.pbot = c->pbot;
  d.teq_contrast = c->teq_contrast;
  d.teq_p_contrast = c->teq_p_contrast;
  d.tau_rad = c->tau_rad;
  d.ref_theta_b = c->ref_theta_b;
  d.ref_a_over_kappa = c->ref_a_over_kappa;
  // (ref_pb/press0)^kappa, with the host libm like the oracle's
  d.ref_pbk = d.is_hj ? pow(c->ref_pb / c->press0, c->kappa) : 0.;
  d.hj_lnref = d.is_hj ? log(c->pbot / c->teq_p_contrast) : 0.;
  d.n_snd = T.n_snd;
  d.n_cool = T.n_cool;
  d.pgrad_trad = c->pgrad_traditional ? 1 : 0;
  d.adv_akima[0] = d.adv_akima[1] = c->adv_nonmass ? 1 : 0;  // unity and theta
  d.adv_akima[2] = c->adv_mass ? 1 : 0;                      // hdry
  d.pv_al81 = c->pv_scheme ? 1 : 0;
  d.heat_area = 0.;
  for (int J = 0; J <= c->nj; J++) {
    const double da = 1. / c->mn[2 * J + 1];
    for (int I = 1; I <= c->ni; I++) d.heat_area += da;
  }
  d.newt_cool_on = c->newt_cool_on; d.newt_cool_adjust = c->newt_cool_adjust;
  d.drag_v = c->drag_v; d.drag_zonal_avg = c->drag_zonal_avg; d.jtp = c->jtp;
  d.nrow_h = T.nrow_h; d.nrow_v = T.nrow_v;
  int lg = 0;
  while ((1 << lg) < c->ni) lg++;
  d.nlogn = lg;
  d.mixed = mixed ? 1 : 0;
  d.dt = (double)c->dt; d.kappa = c->kappa; d.cp = c->cp; d.cpr = c->cpr; d.rgas = c->rgas;
  d.press0 = c->press0; d.zeta0 = c->zeta0; d.zeta1 = c->zeta1; d.alpha = c->hybrid_alpha;
  d.sigma_sigma = c->sigma_sigma; d.topo_scale = c->topo_scale; d.prandtl = c->prandtl; d.tau_drag = c->tau_drag;
  d.coef2 = T.coef2; d.coef4 = T.coef4; d.nu4 = c->nu[4]; d.m0z = T.m0z; d.nudiv = T.nudiv; d.eps13 = T.eps13;
  d.coef6 = T.coef6; d.coef8 = T.coef8; d.nu6 = c->nu[6]; d.nu8 = c->nu[8];
  d.mu3 = T.mu3; d.mu2 = T.mu2; d.g_coeff = T.gcoef;
}

int engine_add_member(epic_engine *E, const epic_case *c, const double *hdry_src, bool prime) {
  if (!E || !c) return -1;
  if (member_check(c, E->precision, E->err, sizeof(E->err))) return -1;
  const bool mixed = (E->precision == EPIC_MIXED);
  int slot = -1;
  for (int s = 0; s < E->max_members; s++)
    if (!E->m[s].active) { slot = s; break; }
  if (slot < 0) return -3;
  epic_tables T;
  epic_tables_build(c, &T);
  struct TablesGuard { epic_tables *t; ~TablesGuard() { epic_tables_free(t); } } tg{&T};
  const size_t N3 = (size_t)c->Nelem3d;
  Member &M = E->m[slot];
  MemberDesc &d = M.h;
  std::vector<MemberPart> parts;
  MemberHold hold;
  member_plan(c, mixed, T, d, parts, hold);
  size_t total = 0;
  for (auto &p : parts) total += align256(p.bytes);
  if (E->budget && E->in_use + total > E->budget) {
    snprintf(E->err, sizeof(E->err), "HBM budget exceeded (%zu + %zu > %zu)", E->in_use, total, E->budget);
    return -2;
  }
  void *blob = nullptr;
  if (hipMalloc(&blob, total) != hipSuccess) {
    snprintf(E->err, sizeof(E->err), "hipMalloc(%zu) failed", total);
    return -2;
  }
  HIPCHK(hipMemsetAsync(blob, 0, total, E->stream));
  size_t off = 0;
  for (auto &p : parts) {
    *p.dst = (char *)blob + off;
    if (p.src) HIPCHK(hipMemcpyAsync(*p.dst, p.src, p.bytes, hipMemcpyHostToDevice, E->stream));
    off += align256(p.bytes);
  }
  // HDRY reconstituted from P3 as epic_main does (src/shared/epic_main.c:290-297),
  // and the FLOAT_MAX Adams-Bashforth start-up flags in planes 1 and 2
  std::vector<double> h(N3, 0.), t1(N3, 0.);
  const int Ia = c->Iadim, N2i = c->Nelem2d;
  auto IX = [&](int K, int J, int I) { return (size_t)I + (size_t)(J + 1) * Ia + (size_t)K * N2i; };
  if (hdry_src) {
    memcpy(h.data(), hdry_src, sizeof(double) * N3);
  } else {
    for (int K = 1; K <= c->nk; K++)
      for (int J = 0; J <= c->nj; J++)
        for (int I = 1; I <= c->ni; I++)
          h[IX(K, J, I)] = ((double)c->p3[IX(K, J, I)] - (double)c->p3[IX(K - 1, J, I)]) /
                           ((double)c->g[2 * J + 1] * (double)c->dsgth[2 * K]);
    for (int K = 0; K <= c->nk + 1; K++)
      for (int J = 0; J <= c->nj + 1; J++) {
        h[IX(K, J, 0)] = h[IX(K, J, c->ni)];
        h[IX(K, J, c->ni + 1)] = h[IX(K, J, 1)];
      }
  }
  for (int K = 1; K <= c->nk; K++)
    for (int J = 0; J <= c->nj; J++)
      for (int I = 1; I <= c->ni; I++) t1[IX(K, J, I)] = 1.7976931348623157e308;
  HIPCHK(hipMemcpyAsync(d.HDRY, h.data(), sizeof(double) * N3, hipMemcpyHostToDevice, E->stream));
  // Adams-Bashforth planes 1 and 2 (IT_MINUS1, IT_MINUS2): a restart case's
  // history, else the FLOAT_MAX start-up flags; FP32 copies in the mixed
  // mode, the flag as FP32's largest value
  std::vector<float> hf[4];
  for (int p = 1; p <= 2; p++)
    for (int w = 0; w < 2; w++) {
      const double *hist = w ? c->dvdt_hist : c->dudt_hist;
      const double *src = hist ? hist + (size_t)(p - 1) * N3 : t1.data();
      double *dst = (w ? d.DVDT : d.DUDT) + p * N3;
      if (mixed) {
        std::vector<float> &b = hf[2 * (p - 1) + w];
        b.assign(N3, 0.f);
        for (size_t i = 0; i < N3; i++)
          b[i] = (src[i] == 1.7976931348623157e308) ? 3.40282347e+38f : (float)src[i];
        HIPCHK(hipMemcpyAsync((w ? d.DVDTf : d.DUDTf) + p * N3, b.data(), sizeof(float) * N3, hipMemcpyHostToDevice,
                              E->stream));
      } else {
        HIPCHK(hipMemcpyAsync(dst, src, sizeof(double) * N3, hipMemcpyHostToDevice, E->stream));
      }
    }
  MemberDyn dy;
  memset(&dy, 0, sizeof(dy));
  dy.it0 = 0; dy.itm1 = 1; dy.itm2 = 2; dy.itime = 0;
  HIPCHK(hipMemcpyAsync(d.dyn, &dy, sizeof(dy), hipMemcpyHostToDevice, E->stream));
  // the host arrays above are read by the copies: wait for them here
  HIPCHK(hipStreamSynchronize(E->stream));
  M.blob = blob;
  M.bytes = total;
  M.dyn = d.dyn;
  M.nu4 = c->nu[4];
  M.nudiv = T.nudiv;
  M.adjust = (c->newt_cool_adjust || c->k_sponge >= 1) && c->newt_cool_on;
  M.steps = 0;
  M.active = 1;
  E->in_use += total;
  int rc = rebuild_buckets(E);
  if (rc) return rc;
  if (!prime) {
    HIPCHK(hipStreamSynchronize(E->stream));
    return slot;
  }
  // prime the diagnostics of this member alone
  std::vector<int> only{slot};
  record_all(E, true, &only);
  HIPCHK(hipGetLastError());
  HIPCHK(hipStreamSynchronize(E->stream));
  MemberDyn chk;
  HIPCHK(hipMemcpy(&chk, d.dyn, sizeof(chk), hipMemcpyDeviceToHost));
  if (chk.err) {
    snprintf(E->err, sizeof(E->err), "fatal condition %d while priming member %d", chk.err, slot);
    return -4;
  }
  return slot;
}

extern "C" int epic_engine_add(epic_engine *E, const epic_case *c) { return engine_add_member(E, c, nullptr, true); }

// SYNC_DIAGS_ONLY for one member (the drop-in's priming call)
int engine_sync_member(epic_engine *E, int slot) {
  if (!E || slot < 0 || slot >= E->max_members || !E->m[slot].active) return -1;
  if (E->dirty && rebuild_buckets(E)) return -5;
  std::vector<int> only{slot};
  record_all(E, true, &only);
  HIPCHK(hipGetLastError());
  return 0;
}

// RI2 of one member, queued on the engine's stream.  store_diag forms it
// from u, v, P2, T2, RHO2, P3 and RHO3 as the step leaves them, and nothing
// after store_diag in the step changes those, so forming it when it is read
// gives store_diag's values (unless u or v were written since the step)
// and keeps it out of the step itself.
int engine_richardson(epic_engine *E, int slot) {
  if (!E || slot < 0 || slot >= E->max_members || !E->m[slot].active) return -1;
  if (E->dirty && rebuild_buckets(E)) return -5;
  for (const Bucket &b : E->buckets)
    for (int i = 0; i < b.nm; i++)
      if (E->order[b.mb + i] == slot) {
        Bucket one = b;
        one.mb = b.mb + i;
        one.nm = 1;
        L_richardson(E->stream, one);
        HIPCHK(hipGetLastError());
        return 0;
      }
  return -1;
}

const MemberDesc *engine_member_desc(epic_engine *E, int slot) {
  if (!E || slot < 0 || slot >= E->max_members || !E->m[slot].active) return nullptr;
  return &E->m[slot].h;
}

hipStream_t engine_stream(epic_engine *E) { return E->stream; }

char *engine_errbuf(epic_engine *E) { return E->err; }

extern "C" int epic_engine_remove(epic_engine *E, int slot) {
  if (!E || slot < 0 || slot >= E->max_members || !E->m[slot].active) return -1;
  HIPCHK(hipStreamSynchronize(E->stream));
  HIPCHK(hipFree(E->m[slot].blob));
  E->in_use -= E->m[slot].bytes;
  E->m[slot].active = 0;
  E->dirty = true;
  return 0;
}

int engine_check_errors(epic_engine *E);
static int check_errors(epic_engine *E) { return engine_check_errors(E); }
int engine_check_errors(epic_engine *E) {
  for (int s : E->order) {
    MemberDyn chk;
    HIPCHK(hipMemcpy(&chk, E->m[s].dyn, sizeof(chk), hipMemcpyDeviceToHost));
    if (chk.err) {
      snprintf(E->err, sizeof(E->err), "fatal condition %d in member %d (non-monotone GZ3=1, p range=2, spline=4)",
               chk.err, s);
      return -4;
    }
  }
  return 0;
}

extern "C" int epic_engine_step(epic_engine *E, int nsteps) {
  if (!E) return -1;
  if (E->dirty && rebuild_buckets(E)) return -5;
  if (E->order.empty()) return 0;
  const bool use_graph = !g_prof_on && getenv("EPIC_AMD_NO_GRAPH") == nullptr;
  if (use_graph && !E->have_graph) {
    HIPCHK(hipStreamSynchronize(E->stream));
    HIPCHK(hipStreamBeginCapture(E->stream, hipStreamCaptureModeThreadLocal));
    record_all(E, false, nullptr);
    HIPCHK(hipStreamEndCapture(E->stream, &E->graph));
    HIPCHK(hipGraphInstantiate(&E->gexec, E->graph, nullptr, nullptr, 0));
    E->have_graph = true;
  }
  for (int n = 0; n < nsteps; n++) {
    if (use_graph) HIPCHK(hipGraphLaunch(E->gexec, E->stream));
    else record_all(E, false, nullptr);
  }
  HIPCHK(hipGetLastError());
  for (int s : E->order) E->m[s].steps += nsteps;
  return 0;
}

extern "C" int epic_engine_step_shape(epic_engine *E, int ni, int nj, int nk, int nsteps) {
  if (!E) return -1;
  if (E->dirty && rebuild_buckets(E)) return -5;
  std::vector<size_t> sel;  // the buckets of the shape (two in the mixed mode when hv orders differ)
  for (size_t bi = 0; bi < E->buckets.size(); bi++)
    if (E->buckets[bi].ni == ni && E->buckets[bi].nj == nj && E->buckets[bi].nk == nk) sel.push_back(bi);
  if (sel.empty()) {
    snprintf(E->err, sizeof(E->err), "no member of shape %dx%dx%d", ni, nj, nk);
    return -1;
  }
  const bool use_graph = !g_prof_on && getenv("EPIC_AMD_NO_GRAPH") == nullptr;
  if (use_graph && E->bgexec.size() != E->buckets.size()) {
    E->bgexec.assign(E->buckets.size(), nullptr);
    E->bgraph.assign(E->buckets.size(), nullptr);
  }
  for (size_t bi : sel) {
    if (!use_graph || E->bgexec[bi]) continue;
    const Bucket &b = E->buckets[bi];
    bool hv, divd, adj;
    bucket_flags(E, b, hv, divd, adj);
    HIPCHK(hipStreamSynchronize(E->stream));
    HIPCHK(hipStreamBeginCapture(E->stream, hipStreamCaptureModeThreadLocal));
    rec_step(E->stream, b, hv, divd, adj, &E->side);
    HIPCHK(hipStreamEndCapture(E->stream, &E->bgraph[bi]));
    HIPCHK(hipGraphInstantiate(&E->bgexec[bi], E->bgraph[bi], nullptr, nullptr, 0));
  }
  for (int n = 0; n < nsteps; n++)
    for (size_t bi : sel) {
      if (use_graph) {
        HIPCHK(hipGraphLaunch(E->bgexec[bi], E->stream));
      } else {
        const Bucket &b = E->buckets[bi];
        bool hv, divd, adj;
        bucket_flags(E, b, hv, divd, adj);
        rec_step(E->stream, b, hv, divd, adj, &E->side);
      }
    }
  HIPCHK(hipGetLastError());
  for (size_t bi : sel)
    for (int i = 0; i < E->buckets[bi].nm; i++) E->m[E->order[E->buckets[bi].mb + i]].steps += nsteps;
  return 0;
}

extern "C" int epic_engine_step_members(epic_engine *E, const int *slots, int nslots, int nsteps) {
  if (!E) return -1;
  if (E->dirty && rebuild_buckets(E)) return -5;
  std::vector<int> only(slots, slots + nslots);
  for (int n = 0; n < nsteps; n++) record_all(E, false, &only);
  HIPCHK(hipGetLastError());
  for (int s : only) E->m[s].steps += nsteps;
  return 0;
}

extern "C" int epic_engine_sync(epic_engine *E) {
  if (!E) return -1;
  HIPCHK(hipStreamSynchronize(E->stream));
  return check_errors(E);
}

// Sticky fatal-condition bits of one member (0: healthy); waits for the
// queued work first
extern "C" int epic_engine_member_error(epic_engine *E, int slot) {
  if (!E || slot < 0 || slot >= E->max_members || !E->m[slot].active) return -1;
  MemberDyn chk;
  HIPCHK(hipStreamSynchronize(E->stream));
  HIPCHK(hipMemcpy(&chk, E->m[slot].dyn, sizeof(chk), hipMemcpyDeviceToHost));
  return chk.err;
}

// Wait for the queued work without checking the members' error bits
extern "C" int epic_engine_wait(epic_engine *E) {
  if (!E) return -1;
  HIPCHK(hipStreamSynchronize(E->stream));
  return 0;
}

extern "C" long epic_engine_steps_done(epic_engine *E, int slot) {
  if (!E || slot < 0 || slot >= E->max_members || !E->m[slot].active) return -1;
  return E->m[slot].steps;
}

static double *field_ptr(MemberDesc &d, int field, size_t &n, int itplanes[3]) {
  n = (size_t)d.N3;
  switch (field) {
    case EPIC_F_U: return d.U;
    case EPIC_F_V: return d.V;
    case EPIC_F_HDRY: return d.HDRY;
    case EPIC_F_THETA: return d.THETA;
    case EPIC_F_P3: return d.P3;
    case EPIC_F_P2: return d.P2;
    case EPIC_F_H3: return d.H3;
    case EPIC_F_THETA2: return d.THETA2;
    case EPIC_F_T2: return d.T2;
    case EPIC_F_T3: return d.T3;
    case EPIC_F_RHO2: return d.RHO2;
    case EPIC_F_RHO3: return d.RHO3;
    case EPIC_F_EXNER3: return d.EXNER3;
    case EPIC_F_GZ3: return d.GZ3;
    case EPIC_F_MONT3: return d.MONT3;
    case EPIC_F_PV3: return d.PV3;
    case EPIC_F_DIV_UV3: return d.DIV;
    case EPIC_F_W3: return d.W3;
    case EPIC_F_HEAT3: return d.HEAT3;
    case EPIC_F_DZDT3: return d.DZDT3;
    case EPIC_F_DUDT0: case EPIC_F_DUDT1: case EPIC_F_DUDT2:
      return d.DUDT + (size_t)(field - EPIC_F_DUDT0) * d.N3;
    case EPIC_F_DVDT0: case EPIC_F_DVDT1: case EPIC_F_DVDT2:
      return d.DVDT + (size_t)(field - EPIC_F_DVDT0) * d.N3;
    case EPIC_F_GZ_SURFACE: n = (size_t)d.N2; return d.GZS;
    case EPIC_F_RI2: return d.RI2;
  }
  (void)itplanes;
  return nullptr;
}

extern "C" int epic_engine_read(epic_engine *E, int slot, int field, double *out) {
  if (!E || slot < 0 || slot >= E->max_members || !E->m[slot].active || !out) return -1;
  const MemberDesc &hd = E->m[slot].h;
  if (hd.mixed && field >= EPIC_F_DUDT0 && field <= EPIC_F_DVDT2) {
    const bool isu = field <= EPIC_F_DUDT2;
    const int plane = isu ? field - EPIC_F_DUDT0 : field - EPIC_F_DVDT0;
    const float *src = (isu ? hd.DUDTf : hd.DVDTf) + (size_t)plane * hd.N3;
    std::vector<float> tmp(hd.N3);
    HIPCHK(hipStreamSynchronize(E->stream));
    HIPCHK(hipMemcpy(tmp.data(), src, sizeof(float) * hd.N3, hipMemcpyDeviceToHost));
    for (int i = 0; i < hd.N3; i++) out[i] = tmp[i];
    return 0;
  }
  size_t n;
  double *p = field_ptr(E->m[slot].h, field, n, nullptr);
  if (!p) return -1;
  if (field == EPIC_F_RI2 && engine_richardson(E, slot) != 0) return -1;
  HIPCHK(hipStreamSynchronize(E->stream));
  HIPCHK(hipMemcpy(out, p, sizeof(double) * n, hipMemcpyDeviceToHost));
  return 0;
}

extern "C" int epic_engine_write(epic_engine *E, int slot, int field, const double *in) {
  if (!E || slot < 0 || slot >= E->max_members || !E->m[slot].active || !in) return -1;
  if (field != EPIC_F_U && field != EPIC_F_V && field != EPIC_F_HDRY && field != EPIC_F_THETA) return -1;
  size_t n;
  double *p = field_ptr(E->m[slot].h, field, n, nullptr);
  HIPCHK(hipStreamSynchronize(E->stream));
  HIPCHK(hipMemcpy(p, in, sizeof(double) * n, hipMemcpyHostToDevice));
  return 0;
}

// time-plane state of a member: it0, itm1, itm2, itime
extern "C" int epic_engine_time_planes(epic_engine *E, int slot, int *out4) {
  if (!E || slot < 0 || slot >= E->max_members || !E->m[slot].active) return -1;
  MemberDyn chk;
  HIPCHK(hipStreamSynchronize(E->stream));
  HIPCHK(hipMemcpy(&chk, E->m[slot].dyn, sizeof(chk), hipMemcpyDeviceToHost));
  out4[0] = chk.it0; out4[1] = chk.itm1; out4[2] = chk.itm2; out4[3] = chk.itime;
  return 0;
}
extern "C" int epic_engine_set_time_planes(epic_engine *E, int slot, const int *in4) {
  if (!E || slot < 0 || slot >= E->max_members || !E->m[slot].active) return -1;
  MemberDyn chk;
  HIPCHK(hipStreamSynchronize(E->stream));
  HIPCHK(hipMemcpy(&chk, E->m[slot].dyn, sizeof(chk), hipMemcpyDeviceToHost));
  chk.it0 = in4[0]; chk.itm1 = in4[1]; chk.itm2 = in4[2]; chk.itime = in4[3];
  HIPCHK(hipMemcpy(E->m[slot].dyn, &chk, sizeof(chk), hipMemcpyHostToDevice));
  return 0;
}

extern "C" int epic_engine_profile(epic_engine *E, int nsteps, char (*names)[48], double *ms_total, int *launches,
                                   int cap) {
  if (!E) return -1;
  if (E->dirty && rebuild_buckets(E)) return -5;
  HIPCHK(hipStreamSynchronize(E->stream));
  g_prof.clear();
  g_evnext = 0;
  g_prof_on = true;
  for (int n = 0; n < nsteps; n++) record_all(E, false, nullptr);
  g_prof_on = false;
  HIPCHK(hipStreamSynchronize(E->stream));
  for (int s : E->order) E->m[s].steps += nsteps;
  std::map<std::string, std::pair<double, int>> agg;
  std::vector<std::string> ord;
  for (auto &r : g_prof) {
    float ms = 0.f;
    (void)hipEventElapsedTime(&ms, r.a, r.b);
    auto it = agg.find(r.name);
    if (it == agg.end()) {
      agg[r.name] = {ms, 1};
      ord.push_back(r.name);
    } else {
      it->second.first += ms;
      it->second.second += 1;
    }
  }
  int k = 0;
  for (auto &nm : ord) {
    if (k >= cap) break;
    snprintf(names[k], 48, "%s", nm.c_str());
    ms_total[k] = agg[nm].first;
    launches[k] = agg[nm].second;
    k++;
  }
  return k;
}

// ------------------------------------------------------ checkpoint/restart
// A member's state for an exact restart: every array the next step reads
// before writing it (the prognostic fields, the Adams-Bashforth history in
// its own precision, the diagnostics of the last step: pressure,
// thermodynamics, geopotential, PV, divergence, heating, W3, DZDT3,
// GZ_SURFACE), and the time-plane state.  Scratch, the filtered copies and
// RI2 are formed again by the step.  File layout (little-endian):
//   "EPICAMD1", int32 {version 1, ni, nj, nk, mixed, it0, itm1, itm2, itime,
//   err}, int64 steps, int32 array count, then per array int32 id, int64
//   bytes and the raw device bytes.
namespace {
struct Ckpt {
  int id;
  void *dev;
  size_t bytes;
};
std::vector<Ckpt> ckpt_arrays(const MemberDesc &d) {
  const size_t N3 = (size_t)d.N3 * sizeof(double), N2 = (size_t)d.N2 * sizeof(double);
  std::vector<Ckpt> a = {
      {EPIC_F_U, d.U, N3},           {EPIC_F_V, d.V, N3},         {EPIC_F_HDRY, d.HDRY, N3},
      {EPIC_F_THETA, d.THETA, N3},   {EPIC_F_P3, d.P3, N3},       {EPIC_F_P2, d.P2, N3},
      {EPIC_F_H3, d.H3, N3},         {EPIC_F_THETA2, d.THETA2, N3}, {EPIC_F_T2, d.T2, N3},
      {EPIC_F_T3, d.T3, N3},         {EPIC_F_RHO2, d.RHO2, N3},   {EPIC_F_RHO3, d.RHO3, N3},
      {EPIC_F_EXNER3, d.EXNER3, N3}, {EPIC_F_GZ3, d.GZ3, N3},     {EPIC_F_MONT3, d.MONT3, N3},
      {EPIC_F_PV3, d.PV3, N3},       {EPIC_F_DIV_UV3, d.DIV, N3}, {EPIC_F_W3, d.W3, N3},
      {EPIC_F_HEAT3, d.HEAT3, N3},   {EPIC_F_DZDT3, d.DZDT3, N3}, {EPIC_F_GZ_SURFACE, d.GZS, N2},
      {100, d.USPIN, N3},
  };
  if (d.mixed) {
    a.push_back({101, d.DUDT, N3});
    a.push_back({102, d.DVDT, N3});
    a.push_back({103, d.DUDTf, 3 * (size_t)d.N3 * sizeof(float)});
    a.push_back({104, d.DVDTf, 3 * (size_t)d.N3 * sizeof(float)});
  } else {
    a.push_back({101, d.DUDT, 3 * N3});
    a.push_back({102, d.DVDT, 3 * N3});
  }
  return a;
}
const char CKPT_MAGIC[8] = {'E', 'P', 'I', 'C', 'A', 'M', 'D', '1'};
}  // namespace

// epic.nc of one member: its state from the device, the header from its case
extern "C" int epic_engine_write_nc(epic_engine *E, int slot, const epic_case *c, const char *path) {
  if (!E || !c || !path || slot < 0 || slot >= E->max_members || !E->m[slot].active) return -1;
  const MemberDesc &hd = E->m[slot].h;
  if (hd.ni != c->ni || hd.nj != c->nj || hd.nk != c->nk) {
    snprintf(E->err, sizeof(E->err), "the case's grid is not the member's");
    return -1;
  }
  if (epic_engine_sync(E) != 0) return -1;
  const size_t N3 = (size_t)hd.N3;
  std::vector<double> u(N3), v(N3), p3(N3), th(N3), gzs((size_t)hd.N2), tend(4 * N3);
  int planes[4];
  if (epic_engine_read(E, slot, EPIC_F_U, u.data()) || epic_engine_read(E, slot, EPIC_F_V, v.data()) ||
      epic_engine_read(E, slot, EPIC_F_P3, p3.data()) || epic_engine_read(E, slot, EPIC_F_THETA, th.data()) ||
      epic_engine_read(E, slot, EPIC_F_GZ_SURFACE, gzs.data()) || epic_engine_time_planes(E, slot, planes))
    return -1;
  // the planes IT_MINUS1 and IT_MINUS2 of u and v; FP32's largest value
  // (the mixed mode's start-up flag) back to FLOAT_MAX
  for (int w = 0; w < 2; w++)
    for (int q = 0; q < 2; q++) {
      double *dst = tend.data() + (size_t)(2 * w + q) * N3;
      if (epic_engine_read(E, slot, (w ? EPIC_F_DVDT0 : EPIC_F_DUDT0) + planes[1 + q], dst)) return -1;
      if (hd.mixed)
        for (size_t i = 0; i < N3; i++)
          if (dst[i] == (double)3.40282347e+38f) dst[i] = 1.7976931348623157e308;
    }
  epic_nc_state st;
  memset(&st, 0, sizeof(st));
  st.u = u.data(); st.v = v.data(); st.p3 = p3.data(); st.theta = th.data();
  st.u_spinup = c->u_spinup; st.gz_surface = gzs.data();
  st.dudt_m1 = tend.data(); st.dudt_m2 = tend.data() + N3;
  st.dvdt_m1 = tend.data() + 2 * N3; st.dvdt_m2 = tend.data() + 3 * N3;
  st.model_time = c->model_time + (double)E->m[slot].steps * (double)c->dt;
  st.l_s_deg = 0.;
  if (epic_nc_write(path, c, &st)) {
    snprintf(E->err, sizeof(E->err), "%s", epic_nc_error());
    return -1;
  }
  return 0;
}

extern "C" int epic_engine_checkpoint(epic_engine *E, int slot, const char *path) {
  if (!E || slot < 0 || slot >= E->max_members || !E->m[slot].active || !path) return -1;
  const MemberDesc &d = E->m[slot].h;
  MemberDyn dy;
  HIPCHK(hipStreamSynchronize(E->stream));
  HIPCHK(hipMemcpy(&dy, d.dyn, sizeof(dy), hipMemcpyDeviceToHost));
  FILE *f = fopen(path, "wb");
  if (!f) {
    snprintf(E->err, sizeof(E->err), "checkpoint: cannot open %s", path);
    return -3;
  }
  auto arr = ckpt_arrays(d);
  int hdr[10] = {1, d.ni, d.nj, d.nk, d.mixed, dy.it0, dy.itm1, dy.itm2, dy.itime, dy.err};
  long long steps = E->m[slot].steps;
  int n = (int)arr.size();
  bool ok = fwrite(CKPT_MAGIC, 1, 8, f) == 8 && fwrite(hdr, sizeof(int), 10, f) == 10 &&
            fwrite(&steps, sizeof(steps), 1, f) == 1 && fwrite(&n, sizeof(int), 1, f) == 1;
  std::vector<char> buf;
  for (auto &a : arr) {
    if (!ok) break;
    buf.resize(a.bytes);
    if (hipMemcpy(buf.data(), a.dev, a.bytes, hipMemcpyDeviceToHost) != hipSuccess) {
      ok = false;
      break;
    }
    long long nb = (long long)a.bytes;
    ok = fwrite(&a.id, sizeof(int), 1, f) == 1 && fwrite(&nb, sizeof(nb), 1, f) == 1 &&
         fwrite(buf.data(), 1, a.bytes, f) == a.bytes;
  }
  if (fclose(f) != 0) ok = false;
  if (!ok) {
    snprintf(E->err, sizeof(E->err), "checkpoint: write of %s failed", path);
    return -3;
  }
  return 0;
}

// A new member from the case it was built from and a checkpoint of it:
// the case supplies the metrics and tables (built deterministically), the
// file the state and time planes.  Returns the slot.
extern "C" int epic_engine_restore(epic_engine *E, const epic_case *c, const char *path) {
  if (!E || !c || !path) return -1;
  FILE *f = fopen(path, "rb");
  if (!f) {
    snprintf(E->err, sizeof(E->err), "restore: cannot open %s", path);
    return -3;
  }
  char magic[8];
  int hdr[10], n = 0;
  long long steps = 0;
  bool ok = fread(magic, 1, 8, f) == 8 && memcmp(magic, CKPT_MAGIC, 8) == 0 && fread(hdr, sizeof(int), 10, f) == 10 &&
            fread(&steps, sizeof(steps), 1, f) == 1 && fread(&n, sizeof(int), 1, f) == 1 && hdr[0] == 1;
  const int mixed = (E->precision == EPIC_MIXED);
  if (!ok || hdr[1] != c->ni || hdr[2] != c->nj || hdr[3] != c->nk || hdr[4] != mixed) {
    fclose(f);
    snprintf(E->err, sizeof(E->err), "restore: %s is not a checkpoint of a %dx%dx%d member in this precision", path,
             c->ni, c->nj, c->nk);
    return -3;
  }
  int slot = engine_add_member(E, c, nullptr, false);
  if (slot < 0) {
    fclose(f);
    return slot;
  }
  const MemberDesc &d = E->m[slot].h;
  auto arr = ckpt_arrays(d);
  std::vector<char> buf;
  HIPCHK(hipStreamSynchronize(E->stream));  // the member's set-up uploads first
  for (int q = 0; q < n && ok; q++) {
    int id;
    long long nb;
    ok = fread(&id, sizeof(int), 1, f) == 1 && fread(&nb, sizeof(nb), 1, f) == 1;
    if (!ok) break;
    auto it = std::find_if(arr.begin(), arr.end(), [&](const Ckpt &a) { return a.id == id; });
    if (it == arr.end() || (size_t)nb != it->bytes) {
      ok = false;
      break;
    }
    buf.resize(it->bytes);
    ok = fread(buf.data(), 1, it->bytes, f) == it->bytes &&
         hipMemcpy(it->dev, buf.data(), it->bytes, hipMemcpyHostToDevice) == hipSuccess;
  }
  fclose(f);
  if (!ok || n != (int)arr.size()) {
    epic_engine_remove(E, slot);
    snprintf(E->err, sizeof(E->err), "restore: %s is truncated or holds other arrays", path);
    return -3;
  }
  MemberDyn dy;
  memset(&dy, 0, sizeof(dy));
  dy.it0 = hdr[5];
  dy.itm1 = hdr[6];
  dy.itm2 = hdr[7];
  dy.itime = hdr[8];
  dy.err = hdr[9];
  HIPCHK(hipMemcpy(d.dyn, &dy, sizeof(dy), hipMemcpyHostToDevice));
  E->m[slot].steps = (long)steps;
  return slot;
}

// --------------------------------------------------------------- autotune
// Picks, per bucket, the launch shape of each tunable kernel (columns per
// column tile, rows per zonal-hyperviscosity block or meridional-advection
// tile) by timing every
// candidate over a few steps of the bucket's own members.  The members'
// blobs are copied aside first and put back after every candidate and at
// the end, so tuning leaves the state (and the step counts) as it found
// them; the shapes never change a result.  The graph is re-captured with
// the chosen shapes at the next step.
static const char *const knob_name[KN_COUNT] = {"cols_p2etc", "cols_gz3", "cols_calcw",
                                                "cols_hsuv",  "rows_zhv", "rows_mer", "rows_d2"};

static bool knob_fits(const Bucket &b, int k, int v) {
  const size_t W = (size_t)b.nk + 1, L = (size_t)b.nk + 2, lim = 128 * 1024;
  if (k <= KN_COLS_HSUV) {
    // lanes per column beyond twice the column's length would idle
    // the kernel's levels per column: nk+2 for the vertical advection pass,
    // nk+1 for the others
    const int C = knob_cols(v), lev = (k == KN_COLS_HSUV) ? (int)L : (int)W, nt = knob_threads(v, lev);
    if (nt > 1024 || (size_t)(nt / C) > 2 * L) return false;
    v = C;
  }
  switch (k) {
    case KN_COLS_P2ETC: return sizeof(double) * (2 * v * W + 4 * v) <= lim;
    case KN_COLS_GZ3: return sizeof(double) * 3 * v * W <= lim;
    case KN_COLS_CALCW: return sizeof(double) * (4 * v * W + v) <= lim;
    case KN_COLS_HSUV: return sizeof(double) * 4 * v * L <= lim;
    case KN_ROWS_MER: return sizeof(double) * (size_t)(b.ni < 64 ? b.ni : 64) * (6 * v + 15) <= lim;
    case KN_ROWS_D2: return true;  // registers only
    default:  // KN_ROWS_ZHV: one row array per row, and at most 16 points per lane
      return sizeof(double) * ((size_t)v * ((b.ni + 2) | 1) + v) <= lim && v / 4 * (b.ni >= 64 ? b.ni / 64 : 1) <= 16;
  }
}

// the candidates: column tiles of 8-64 columns per block of 256, 512 or
// 1024 threads, or of 4-16 columns with one lane per level; 4-16 rows per zonal-hyperviscosity block; 2-8 rows per
// meridional-advection tile; 8-32 rows per column tile of the del^4
// meridional second differences
static std::vector<int> knob_candidates(int k) {
  static const int cols[16] = {8,    16,   32,   64,   2008,  2016,  2032,  2064,
                               4008, 4016, 4032, 4064, 10004, 10008, 10012, 10016};
  static const int zhv[3] = {4, 8, 16}, mer[3] = {2, 4, 8}, d2[3] = {8, 16, 32};
  if (k <= KN_COLS_HSUV) return std::vector<int>(cols, cols + 16);
  if (k == KN_ROWS_D2) return std::vector<int>(d2, d2 + 3);
  return k == KN_ROWS_ZHV ? std::vector<int>(zhv, zhv + 3) : std::vector<int>(mer, mer + 3);
}

static bool knob_times(const char *kernel, int k) {
  static const char *const pfx[KN_COUNT] = {"k_set_p2etc", "k_gz3", "k_calc_w", "k_hsu_vert", "k_zonal_hv",
                                            "k_hsu_mer", "k_d2dy2_cols"};
  return strncmp(kernel, pfx[k], strlen(pfx[k])) == 0;
}

extern "C" int epic_engine_autotune(epic_engine *E, int steps, char *report, int cap) {
  if (!E) return -1;
  if (E->dirty && rebuild_buckets(E)) return -5;
  if (steps < 1) steps = 1;
  if (report && cap > 0) report[0] = 0;
  HIPCHK(hipStreamSynchronize(E->stream));
  // put the state aside
  std::vector<void *> snap(E->order.size(), nullptr);
  std::vector<long> done(E->order.size());
  auto release = [&]() {
    for (void *p : snap)
      if (p) (void)hipFree(p);
  };
  for (size_t i = 0; i < E->order.size(); i++) {
    Member &M = E->m[E->order[i]];
    if (hipMalloc(&snap[i], M.bytes) != hipSuccess) {
      release();
      snprintf(E->err, sizeof(E->err), "autotune: no memory for the state copy (%zu B)", M.bytes);
      return -2;
    }
    HIPCHK(hipMemcpyAsync(snap[i], M.blob, M.bytes, hipMemcpyDeviceToDevice, E->stream));
    done[i] = M.steps;
  }
  auto restore = [&]() {
    for (size_t i = 0; i < E->order.size(); i++) {
      Member &M = E->m[E->order[i]];
      (void)hipMemcpyAsync(M.blob, snap[i], M.bytes, hipMemcpyDeviceToDevice, E->stream);
      M.steps = done[i];
    }
  };
  size_t used = 0;
  for (Bucket &b : E->buckets) {
    bool hv, divd, adj;
    bucket_flags(E, b, hv, divd, adj);
    for (int k = 0; k < KN_COUNT; k++) {
      double best = 1e300;
      int bestv = b.knob[k];
      for (int v : knob_candidates(k)) {
        if (!v || !knob_fits(b, k, v)) continue;
        b.knob[k] = v;
        restore();
        rec_step(E->stream, b, hv, divd, adj, &E->side);  // warm-up, not timed
        HIPCHK(hipStreamSynchronize(E->stream));
        g_prof.clear();
        g_evnext = 0;
        g_prof_on = true;
        for (int n = 0; n < steps; n++) rec_step(E->stream, b, hv, divd, adj, &E->side);
        g_prof_on = false;
        HIPCHK(hipStreamSynchronize(E->stream));
        double t = 0.;
        for (auto &r : g_prof) {
          if (!knob_times(r.name, k)) continue;
          float ms = 0.f;
          (void)hipEventElapsedTime(&ms, r.a, r.b);
          t += ms;
        }
        if (t < best) {
          best = t;
          bestv = v;
        }
      }
      b.knob[k] = bestv;
      if (report && used < (size_t)cap) {
        char shape[32];
        const int lev = (k == KN_COLS_HSUV) ? b.nk + 2 : b.nk + 1;
        if (k <= KN_COLS_HSUV && knob_threads(bestv, lev) != 256)
          snprintf(shape, sizeof(shape), "%d/%dthr", knob_cols(bestv), knob_threads(bestv, lev));
        else
          snprintf(shape, sizeof(shape), "%d", bestv);
        if (best < 1e300)
          used += snprintf(report + used, cap - used, "%dx%dx%d x%d %s=%s (%.1f us/step)\n", b.ni, b.nj, b.nk,
                           b.nm, knob_name[k], shape, best / steps * 1e3);
        else
          used += snprintf(report + used, cap - used, "%dx%dx%d x%d %s=built-in (no candidate fits)\n", b.ni,
                           b.nj, b.nk, b.nm, knob_name[k]);
      }
    }
    std::array<int, KN_COUNT> kv;
    for (int k = 0; k < KN_COUNT; k++) kv[k] = b.knob[k];
    E->tuned[{b.ni, b.nj, b.nk, b.nm}] = kv;
  }
  restore();
  HIPCHK(hipStreamSynchronize(E->stream));
  release();
  HIPCHK(hipGetLastError());
  drop_graph(E);
  return 0;
}

extern "C" int epic_engine_set_shape(epic_engine *E, int knob, int value) {
  if (!E || knob < 0 || knob >= KN_COUNT) return -1;
  if (E->dirty && rebuild_buckets(E)) return -5;
  if (value) {
    const std::vector<int> c = knob_candidates(knob);
    if (std::find(c.begin(), c.end(), value) == c.end()) {
      snprintf(E->err, sizeof(E->err), "set_shape: %d is not a candidate of %s", value, knob_name[knob]);
      return -1;
    }
    for (Bucket &b : E->buckets)
      if (!knob_fits(b, knob, value)) {
        snprintf(E->err, sizeof(E->err), "set_shape: %s=%d does not fit the %dx%dx%d bucket", knob_name[knob],
                 value, b.ni, b.nj, b.nk);
        return -1;
      }
  }
  HIPCHK(hipStreamSynchronize(E->stream));
  for (Bucket &b : E->buckets) {
    b.knob[knob] = value;
    auto t = E->tuned.find({b.ni, b.nj, b.nk, b.nm});
    if (t == E->tuned.end()) {
      std::array<int, KN_COUNT> kv{};
      kv[knob] = value;
      E->tuned[{b.ni, b.nj, b.nk, b.nm}] = kv;
    } else {
      t->second[knob] = value;
    }
  }
  drop_graph(E);
  return 0;
}

// --------------------------------------------------------- copy probe
// A plain streaming copy (16-byte vector loads and stores, grid-stride),
// the reference point for bandwidth figures and for scaling the
// FETCH_SIZE / WRITE_SIZE counters.
__global__ void __launch_bounds__(256) k_copy_probe(const double2 *__restrict__ src, double2 *__restrict__ dst, size_t n) {
  size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x, st = (size_t)gridDim.x * blockDim.x;
  for (; i < n; i += st) dst[i] = src[i];
}

extern "C" int epic_copy_probe(size_t nbytes, int reps, double *seconds_per_copy) {
  if (nbytes < 4096 || reps < 1 || !seconds_per_copy) return -1;
  const size_t n = nbytes / sizeof(double2);
  double2 *a = nullptr, *b = nullptr;
  if (hipMalloc((void **)&a, n * sizeof(double2)) != hipSuccess) return -2;
  if (hipMalloc((void **)&b, n * sizeof(double2)) != hipSuccess) {
    (void)hipFree(a);
    return -2;
  }
  int rc = 0;
  hipEvent_t e0, e1;
  (void)hipEventCreate(&e0);
  (void)hipEventCreate(&e1);
  (void)hipMemset(a, 0, n * sizeof(double2));
  const dim3 g(4096), blk(256);
  for (int r = 0; r < 3; r++) hipLaunchKernelGGL(k_copy_probe, g, blk, 0, 0, a, b, n);
  (void)hipEventRecord(e0, 0);
  for (int r = 0; r < reps; r++) hipLaunchKernelGGL(k_copy_probe, g, blk, 0, 0, a, b, n);
  (void)hipEventRecord(e1, 0);
  if (hipEventSynchronize(e1) != hipSuccess) rc = -5;
  float ms = 0.f;
  (void)hipEventElapsedTime(&ms, e0, e1);
  *seconds_per_copy = ms / 1e3 / reps;
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  (void)hipFree(a);
  (void)hipFree(b);
  return rc;
}

extern "C" int epic_engine_restore_spec(epic_engine *E, const epic_case_spec *sp, const char *path) {
  epic_case c;
  int rc = epic_case_build(sp, &c);
  if (rc) {
    if (E) snprintf(E->err, sizeof(E->err), "case build failed (%d)", rc);
    epic_case_free(&c);
    return -1;
  }
  rc = epic_engine_restore(E, &c, path);
  epic_case_free(&c);
  return rc;
}

extern "C" int epic_engine_add_spec(epic_engine *E, const epic_case_spec *sp) {
  epic_case c;
  int rc = epic_case_build(sp, &c);
  if (rc) {
    if (E) snprintf(E->err, sizeof(E->err), "case build failed (%d)", rc);
    epic_case_free(&c);
    return -1;
  }
  rc = epic_engine_add(E, &c);
  epic_case_free(&c);
  return rc;
}
