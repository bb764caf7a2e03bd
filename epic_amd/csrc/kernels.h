// kernels.h -- host-side launchers of the engine's HIP kernels.
// Every launcher covers the members [mb, mb+nm) of one bucket (same ni,nj,nk)
// on stream s.  Field selectors name a 3-D array of the MemberDesc.
#pragma once
#include <hip/hip_runtime.h>

#include "dev_common.h"

namespace epic {

// Optional per-launch timing: when enabled, every launch is bracketed by HIP
// events on its stream and the durations are aggregated per kernel name.
void prof_begin(hipStream_t s, const char *name);
void prof_end(hipStream_t s);
extern bool g_prof_on;
#define LAUNCH(k, g, b, shm, s, ...) LAUNCH_AS(#k, k, g, b, shm, s, __VA_ARGS__)
// the same under a given profile name (for a kernel whose template arguments
// hold commas, which is passed in parentheses)
#define LAUNCH_AS(name, k, g, b, shm, s, ...)            \
  do {                                                   \
    if (::epic::g_prof_on) ::epic::prof_begin(s, name);  \
    hipLaunchKernelGGL(k, g, b, shm, s, __VA_ARGS__);    \
    if (::epic::g_prof_on) ::epic::prof_end(s);          \
  } while (0)

// Up to four fields processed by one launch of a batched kernel, each with
// its kind (0 U, 1 V, 2 HDRY, 3 THETA); blockIdx.z = field index * members
// + member, and field i uses scratch copy copy0 + i (the hyperviscosity
// launches of hdry and theta take copies 1 and 2, so that they can run beside
// uv_core, which uses copy 0).
struct FieldSet {
  int n;
  int f[4], kind[4];
  int copy0;
};

// Launch shapes a bucket can be tuned over (epic_engine_autotune); 0 takes
// the built-in choice.  None of them changes a result: every cell is formed
// by the same operations whatever the tile.
enum Knob { KN_COLS_P2ETC, KN_COLS_GZ3, KN_COLS_CALCW, KN_COLS_HSUV, KN_ROWS_ZHV, KN_ROWS_MER, KN_ROWS_D2, KN_COUNT };
// A column-tile knob holds the columns per block C and, from 1000 up, the
// block's threads: v = C (256 threads), v = 1000 n + C (256 n threads,
// n = 2 or 4: more lanes per column, more waves in flight per CU for the
// same LDS), or v = 10000 + C (one lane per level: C times the kernel's
// levels per column, `levels`).
static inline int knob_cols(int v) { return v % 1000; }
static inline int knob_threads(int v, int levels) {
  return v >= 10000 ? (v % 1000) * levels : v >= 1000 ? (v / 1000) * 256 : 256;
}

struct Bucket {
  const MemberDesc *D;  // device array of descriptors
  int mb, nm;           // first member, member count
  int ni, nj, nk;
  int mixed;            // FP32 tendencies and horizontal arithmetic
  int knob[KN_COUNT];   // columns per column tile / rows per zonal-mean block
  // some member is a gas giant (gz_from_u), Held-Suarez (boundary-layer
  // drag), or has Rayleigh drag: launches none of the members needs are
  // left out (their kernels would return at once for every member)
  int any_gas, any_hs, any_ray;
  // hyperviscosity passes of the member that needs the most (hv_passes)
  int hv_passes;
  // advection schemes some member's field takes: bit 0 Hsu-Konor-Arakawa,
  // bit 1 Akima (the launches of a scheme no field takes are left out)
  int adv_schemes;
};



// the Bucket::adv_schemes bits of one member
inline int adv_schemes(const MemberDesc &d) {
  int s = 0;
  for (int f = 0; f < 3; f++) s |= d.adv_akima[f] ? 2 : 1;
  return s;
}

// ---- diagnostics (k_diag.hip)
void L_fill(hipStream_t s, const Bucket &b, int fld, double v);
void L_set_p2etc(hipStream_t s, const Bucket &b, int update_theta);
void L_richardson(hipStream_t s, const Bucket &b);
void L_gz_from_u(hipStream_t s, const Bucket &b);
void L_gz3(hipStream_t s, const Bucket &b);
void L_dzdt3(hipStream_t s, const Bucket &b);
void L_bc3(hipStream_t s, const Bucket &b, int fld);
void L_heating(hipStream_t s, const Bucket &b, int adjust, bool in_thermo);
void L_calc_w(hipStream_t s, const Bucket &b, int with_source);
void L_copy_planes(hipStream_t s, const Bucket &b, int src, int dst, int k0, int k1);
void L_vorticity(hipStream_t s, const Bucket &b, int uu, int vv, int k0, int k1);
void L_divergence(hipStream_t s, const Bucket &b, int uu, int vv, int dst, int k0, int k1);
// store_diag (PV3 and DIV_UV3 of U, V on K = 0..nk) in one launch; rotate:
// also rotate the time planes and advance itime (the step's last action)
void L_store_diag(hipStream_t s, const Bucket &b, int rotate);
// store_pgrad_vars' per-cell values after the FULL set_p2etc
void L_thermo(hipStream_t s, const Bucket &b);

// ---- zonal filter (k_filter.hip)
// 2-D filter of planes k0..k1 of fld; vgrid selects the v/pv rows; mode 0
// also refreshes the zonal pads of every row, mode 1 leaves pads alone;
// copies > 1 filters the same planes of that many consecutive 3-D copies.
void L_filter(hipStream_t s, const Bucket &b, int fld, int k0, int k1, int vgrid, int mode, int copies = 1);
void L_filter_f32(hipStream_t s, const Bucket &b, int fld, int k0, int k1, int vgrid, int mode, int copies = 1);
// the 3-D filter with the PCHIP profile removal, on every field of the set
void L_filter3(hipStream_t s, const Bucket &b, const FieldSet &fs);

// ---- momentum (k_dyn.hip)
void L_uv_core(hipStream_t s, const Bucket &b);
void L_uv_drag(hipStream_t s, const Bucket &b);
void L_pgrad(hipStream_t s, const Bucket &b);
void L_divdamp(hipStream_t s, const Bucket &b);
void L_ab(hipStream_t s, const Bucket &b);
void L_gz_mono(hipStream_t s, const Bucket &b);
void L_gz_filtercopy(hipStream_t s, const Bucket &b);
// hyperviscosity on every field of the set (at most two: the meridional
// part keeps its del^4 result in scratch copy i of sB)
void L_merid_hv(hipStream_t s, const Bucket &b, const FieldSet &fs);
void L_zonal_hv(hipStream_t s, const Bucket &b, const FieldSet &fs);
void L_hv(hipStream_t s, const Bucket &b, const FieldSet &fs);

// ---- mixed-precision momentum kernels (k_mixed.hip)
void LM_uv_core(hipStream_t s, const Bucket &b);
void LM_uv_drag(hipStream_t s, const Bucket &b);
void LM_divdamp_apply(hipStream_t s, const Bucket &b);
void LM_ab(hipStream_t s, const Bucket &b);

// ---- advection (k_adv.hip)
void L_advection(hipStream_t s, const Bucket &b);

}  // namespace epic
