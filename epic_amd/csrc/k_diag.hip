// k_diag.hip -- diagnostic kernels of the EPIC step on gfx950 (FP64).
//
//   set_p2etc         src/shared/epic_funcs_diag.c:3352-3569  (column scan)
//   store_pgrad_vars  src/shared/epic_funcs_diag.c:3595-3950  (column scans)
//   gz_from_u         src/shared/epic_funcs_diag.c:4460-4610
//   calc_heating      src/shared/epic_funcs_diag.c:4793-4869 +
//                     src/shared/epic_sensible_heating.c:51-463 (Held-Suarez,
//                     and the hot Jupiter's day-night relaxation; inside the
//                     step it is formed with T3 by thermo_cell)
//   calc_w            src/shared/epic_flux.c:123-335           (column scan)
//   vorticity         src/shared/epic_funcs_diag.c:4131-4445
//   divergence        src/shared/epic_funcs_diag.c:4064-4117
//
// Column work runs in LDS column tiles (per-level arithmetic on every lane,
// the K recurrences in one lane per column); zonal means stage the row in
// LDS and sum it in one lane in the reference's order.  The arithmetic of
// each expression follows the reference's operation order.
#include <algorithm>

#include "kernels.h"
#include "filter_row.h"

namespace epic {


#define MEMBER const MemberDesc &d = D[mb + blockIdx.z]
// flat (J,I) over nJ rows starting at j0 and nI columns starting at i0
#define FLAT2(j0, nJ, i0, nI)                          \
  int t_ = blockIdx.x * TPB + threadIdx.x;             \
  if (t_ >= (nJ) * (nI)) return;                       \
  const int J = (j0) + t_ / (nI);                      \
  const int I = (i0) + t_ % (nI);
// the same with the block renumbered for the XCDs (xcd_block): bx_.y and
// bx_.z stand for blockIdx.y and blockIdx.z
#define MEMBER_X                                       \
  const uint3 bx_ = xcd_block();                       \
  const MemberDesc &d = D[mb + bx_.z]
#define FLAT2_X(j0, nJ, i0, nI)                        \
  int t_ = bx_.x * TPB + threadIdx.x;                  \
  if (t_ >= (nJ) * (nI)) return;                       \
  const int J = (j0) + t_ / (nI);                      \
  const int I = (i0) + t_ % (nI);

static inline dim3 grid2(int nJ, int nI, int nK, int nm) {
  return dim3((unsigned)((nJ * nI + TPB - 1) / TPB), (unsigned)nK, (unsigned)nm);
}

// ------------------------------------------------------------ utilities
__global__ void __launch_bounds__(256) k_fill(const MemberDesc *D, int mb, int f, double v) {
  MEMBER;
  double *a = fld_ptr(d, f);
  size_t i = (size_t)blockIdx.x * TPB + threadIdx.x;
  size_t stride = (size_t)gridDim.x * TPB;
  for (; i < (size_t)d.N3; i += stride) a[i] = v;
}
void L_fill(hipStream_t s, const Bucket &b, int f, double v) {
  LAUNCH(k_fill, dim3(256, 1, b.nm), dim3(TPB), 0, s, b.D, b.mb, f, v);
}

// zonal periodicity of a whole 3-D array (bc_lateral THREEDIM,
// src/single/s_epic_functions.c:72-88): K = 0..nk+1, J = 0..nj+1
__global__ void __launch_bounds__(256) k_bc3(const MemberDesc *D, int mb, int f) {
  MEMBER;
  double *a = fld_ptr(d, f);
  int t = blockIdx.x * TPB + threadIdx.x;
  int nrow = (d.nj + 2) * (d.nk + 2);
  if (t >= nrow) return;
  int K = t / (d.nj + 2), J = t % (d.nj + 2);
  a[ix3(d, K, J, 0)] = a[ix3(d, K, J, d.ni)];
  a[ix3(d, K, J, d.ni + 1)] = a[ix3(d, K, J, 1)];
}
void L_bc3(hipStream_t s, const Bucket &b, int f) {
  int nrow = (b.nj + 2) * (b.nk + 2);
  LAUNCH(k_bc3, dim3((nrow + TPB - 1) / TPB, 1, b.nm), dim3(TPB), 0, s, b.D, b.mb, f);
}

// copy planes k0..k1 (whole planes) src -> dst
__global__ void __launch_bounds__(256) k_copy_planes(const MemberDesc *D, int mb, int src, int dst, int k0) {
  MEMBER;
  const double *a = fld_ptr(d, src);
  double *o = fld_ptr(d, dst);
  int t = blockIdx.x * TPB + threadIdx.x;
  if (t >= d.N2) return;
  size_t off = (size_t)(k0 + blockIdx.y) * d.N2 + t;
  o[off] = a[off];
}
void L_copy_planes(hipStream_t s, const Bucket &b, int src, int dst, int k0, int k1) {
  int N2 = (b.ni + 2) * (b.nj + 3);
  LAUNCH(k_copy_planes, dim3((N2 + TPB - 1) / TPB, k1 - k0 + 1, b.nm), dim3(TPB), 0, s, b.D, b.mb,
                     src, dst, k0);
}

// ------------------------------------------------------------ column tiles
// A block holds C consecutive (J,I) columns (J = 0..nj, I = 0..ni+1) and
// all their levels.  The per-level work (transcendentals, Leslie-Purser
// weights) runs on all lanes, (column, level) pairs strided over the block;
// the recurrences in K (pressure from the top, geopotential from the
// bottom, the running SUM_DIV) run in one lane per column over values staged
// in LDS, in the reference's order.  C is 64 where the block's LDS stays
// small; deep grids take 32 or 16 columns so that several blocks still fit
// in a CU's LDS.
struct Tile {
  int C, c, kk, KS, J, I;
  bool live;
};
__device__ __forceinline__ Tile tile_of(const MemberDesc &d, int C) {
  Tile t;
  t.C = C;
  t.c = threadIdx.x % C;
  t.kk = threadIdx.x / C;
  t.KS = blockDim.x / C;
  const int nI = d.ni + 2, col = blockIdx.x * C + t.c;
  t.live = col < (d.nj + 1) * nI;
  t.J = t.live ? col / nI : 0;
  t.I = t.live ? col % nI : 0;
  return t;
}
// columns per block for a kernel keeping `arrays` (nk+1)-long columns in LDS
static inline int tile_cols(const Bucket &b, int arrays) {
  int C = 64;
  while (C > 16 && (size_t)arrays * C * (b.nk + 1) * sizeof(double) > 40 * 1024) C /= 2;
  return C;
}
static inline dim3 tile_grid(const Bucket &b, int C) {
  return dim3((unsigned)(((b.nj + 1) * (b.ni + 2) + C - 1) / C), 1, (unsigned)b.nm);
}

// ------------------------------------------------------------ set_p2etc
// P3 from the top down (prefix of dsgth*g*h), the sigma-region pressures,
// h = max(h_min, dp/(g dsgth)), and the layer diagnostics P2, H3, THETA2
// (and the theta of the sigma region when update_theta)
// The FULL = 0 launch then forms store_pgrad_vars' T3 and RHO3 of its
// columns (thermo_cell).
// FULL = 1: set_p2etc(UPDATE_THETA) with P2, H3, THETA2.  FULL = 0: the
// call that follows the advection inside the step forms P3 and h (and the
// pressure-range check) only: the P2 and H3 it would form are written again
// by the step's second set_p2etc before anything reads them, so the values
// the step leaves are the same.
template <bool F>
__device__ __forceinline__ void thermo_cell(const MemberDesc &d, int K, int J, int I, bool full, double th, double p);
// F: the mixed mode's exp/log/pow (m_exp)
template <int FULL, bool F>
__global__ void __launch_bounds__(1024) k_set_p2etc(const MemberDesc *D, int mb, int C) {
  constexpr bool update_theta = FULL;
  MEMBER;
  extern __shared__ double lds[];
  const Tile T = tile_of(d, C);
  const int nk = d.nk, ks = d.ks, W = nk + 1, J = T.J, I = T.I;
  double *a = lds + T.c * W, *pw = lds + C * W + T.c * W;
  // theta(ks-2), H3(1), P2(1), log(ptop/pbot) of the column
  double *aux = lds + 2 * C * W + 4 * T.c;
  double *P3 = d.P3, *H = d.HDRY;
  const double g = d.g[2 * J + 1];
  if (T.live)
    for (int K = T.kk; K <= nk; K += T.KS)
      a[K] = (K == 0) ? P3[ix3(d, 0, J, I)] : d.dsgth[2 * K] * g * H[ix3(d, K, J, I)];
  __syncthreads();
  // running sum in a register (a[K-1] + a[K], in order), not through LDS,
  // its terms read eight at a time ahead of the chain
  if (T.live && T.kk == 0) {
    double acc = a[0];
    int K = 1;
    for (; K + 7 <= nk; K += 8) {
      double t8[8];
#pragma unroll
      for (int u = 0; u < 8; u++) t8[u] = a[K + u];
#pragma unroll
      for (int u = 0; u < 8; u++) {
        acc = acc + t8[u];
        a[K + u] = acc;
      }
    }
    for (; K <= nk; K++) {
      acc = acc + a[K];
      a[K] = acc;
    }
    // log(ptop/pbot) once per column (the sigma-region formulas read it)
    aux[3] = m_log<F>(a[0] / a[nk]);
  }
  __syncthreads();
  if (T.live) {
    const double pbot = a[nk], lr = aux[3];
    for (int K = ks - 1 + T.kk; K < nk; K += T.KS) {
      double sigma = (d.sgth[2 * K + 1] - d.zeta0) / (d.zeta1 - d.zeta0);
      a[K] = get_p_sigma_l<F>(pbot, sigma, lr);
    }
  }
  __syncthreads();
  // p^(kappa+1) once per interface (each is the top of one layer and the
  // bottom of the next)
  if (FULL && T.live)
    for (int K = T.kk; K <= nk; K += T.KS) pw[K] = m_pow<F>(a[K], d.kappa + 1.);
  __syncthreads();
  if (T.live) {
    const double pbot = a[nk], kp1 = d.kappa + 1., lr = aux[3];
    const double *tx = d.sgth + 2 * nk + 1;  // tx[kay*(-2)] = sgth[2*(nk-kay)+1]
    for (int K = T.kk; K <= nk; K += T.KS) {
      const size_t q = ix3(d, K, J, I);
      P3[q] = a[K];
      if (update_theta && K < ks - 1) {
        double sigma = get_sigma_l<F>(pbot, a[K], lr);
        double th = theta_of_sigma<F>(d, d.sgth[2 * K + 1], sigma);
        d.THETA[q] = th;
        if (K == ks - 2) aux[0] = th;
      }
      if (K == 0) continue;
      double hl = (a[K] - a[K - 1]) / (g * d.dsgth[2 * K]);
      H[q] = dmax(d.h_min[K], hl);
      if (!FULL) {
        if (fcmp_d(a[K], a[K - 1]) != 0 && (a[K - 1] < 0. || a[K] < 0.)) atomicOr(&d.dyn->err, 2);
        continue;
      }
      double top = a[K - 1], bot = a[K], pp;
      if (fcmp_d(bot, top) == 0) pp = .5 * (bot + top);
      else if (top < 0. || bot < 0.) { atomicOr(&d.dyn->err, 2); pp = 0.; }
      else pp = m_pow<F>((pw[K] - pw[K - 1]) / (kp1 * (bot - top)), 1. / d.kappa);
      if (K == 1) aux[2] = pp;
      d.P2[q] = pp;
      double h3;
      if (K < nk) {
        int kk = 2 * K + 1;
        double topwt = d.dsgth[kk - 1], botwt = d.dsgth[kk + 1];
        double hl1 = (a[K + 1] - a[K]) / (g * d.dsgth[2 * (K + 1)]);
        h3 = (topwt * hl + botwt * hl1) / (topwt + botwt);
      } else {
        h3 = hl;
      }
      d.H3[q] = h3;
      if (K == 1) aux[1] = h3;
      if (update_theta) {
        if (K < ks - 1) {
          double sigma = get_sigma_l<F>(pbot, pp, lr);
          d.THETA2[q] = theta_of_sigma<F>(d, d.sgth[2 * K], sigma);
        } else if (K >= ks) {
          // linear interpolation of theta against sigmatheta, table ordered
          // from the bottom interface up (src/shared/epic_funcs_diag.c:3522-3562);
          // the interval is the same in every column, found once on the host
          // by the same search
          const double sg = d.sgth[2 * K], h = d.t2_h[K];
          const int kay = d.t2_kay[K];
          double y0 = d.THETA[ix3(d, nk - kay, J, I)], y1 = d.THETA[ix3(d, nk - kay - 1, J, I)];
          d.THETA2[q] = y0 + (y1 - y0) * (sg - tx[-2 * kay]) / h;
        }
      }
    }
  }
  __syncthreads();
  if (FULL && T.live && T.kk == 0) {
    const size_t q0 = ix3(d, 0, J, I);
    d.P2[q0] = a[0] * a[0] / aux[2];
    d.H3[q0] = aux[1];
    if (update_theta) d.THETA2[ix3(d, ks - 1, J, I)] = .5 * (d.THETA[ix3(d, ks - 1, J, I)] + aux[0]);
  }
  // FULL = 0: store_pgrad_vars' per-cell part (T3, RHO3) on the column's
  // levels from the P3 in LDS.  FULL = 1 leaves its larger per-cell part to
  // k_thermo, whose one lane per cell hides the latency of its pow chains
  // better than the tile's few lanes per column.
  if (!FULL && T.live)
    for (int K = T.kk; K <= nk; K += T.KS) thermo_cell<F>(d, K, J, I, false, d.THETA[ix3(d, K, J, I)], a[K]);
}
void L_set_p2etc(hipStream_t s, const Bucket &b, int update_theta) {
  const int v = b.knob[KN_COLS_P2ETC], C = v ? knob_cols(v) : tile_cols(b, 2), nt = knob_threads(v, b.nk + 1);
  size_t shm = sizeof(double) * (2 * C * (b.nk + 1) + 4 * C);
  if (b.mixed) {
    if (update_theta) LAUNCH_AS("k_set_p2etc<1>", (k_set_p2etc<1, true>), tile_grid(b, C), dim3(nt), shm, s, b.D, b.mb, C);
    else LAUNCH_AS("k_set_p2etc<0>", (k_set_p2etc<0, true>), tile_grid(b, C), dim3(nt), shm, s, b.D, b.mb, C);
  } else {
    if (update_theta) LAUNCH_AS("k_set_p2etc<1>", (k_set_p2etc<1, false>), tile_grid(b, C), dim3(nt), shm, s, b.D, b.mb, C);
    else LAUNCH_AS("k_set_p2etc<0>", (k_set_p2etc<0, false>), tile_grid(b, C), dim3(nt), shm, s, b.D, b.mb, C);
  }
}

// ------------------------------------------------- store_pgrad_vars parts
// T3, EXNER3, RHO3, T2, RHO2 per cell and the top values; DZDT3 <- GZ3 when
// itime > 1, DZDT3 zeroed whole while itime <= 1
// (src/shared/epic_funcs_diag.c:3674-3690).  One lane per (K,J,I), K = 0..nk.
// full = 0 (the step's first store_pgrad_vars): T3 and RHO3 only, all that
// the geopotential integration and the filtered GZ3 copy read before the
// second call writes every output again.
template <bool F = false>
__device__ __forceinline__ double hs_relax(const MemberDesc &d, int J, double p, double t, double pbot, double pk0);
template <bool F = false>
__device__ __forceinline__ double hj_relax(const MemberDesc &d, int J, int I, double p, double t, double pk);
// full = 1 also forms the Newtonian relaxation HEAT3 of calc_heating on
// K = 1..nk, J = 0..nj and the zonal pads: it reads only P3 and T3, which
// nothing changes between here and the end of the step, where the reference
// forms it, and it takes the p^kappa formed here for T3 (a pad takes the
// value of the column it copies)
// T3 and RHO3 of an interface cell from its theta and p (return_temp and
// return_density with x_h2 = 0); pk = (p/press0)^kappa
template <bool F = false>
__device__ __forceinline__ double t3_of(const MemberDesc &d, double th, double p, double *pk) {
  *pk = m_pow<F>(p / d.press0, d.kappa);
  return th * *pk;
}
__device__ __forceinline__ double rho3_of(const MemberDesc &d, double p, double t) {
  const double R_GAS = 8.314472e+3;
  return p * d.mu3 / (R_GAS * t);
}
template <bool F>
__device__ __forceinline__ void thermo_cell(const MemberDesc &d, int K, int J, int I, bool full, double th, double p) {
  const size_t q = ix3(d, K, J, I);
  if (full && K >= 1 && d.dyn->itime > 1 && I >= 1 && I <= d.ni) d.DZDT3[q] = d.GZ3[q];
  const double R_GAS = 8.314472e+3;
  double pk;
  double t = t3_of<F>(d, th, p, &pk);
  double rho3 = rho3_of(d, p, t);
  d.T3[q] = t;
  d.RHO3[q] = rho3;
  if (!full) return;
  if (K >= 1 && newt_relax_on(d)) {
    double h;
    if (d.is_tab) {
      h = 0. + tab_relax(d, p, t);
    } else if (d.is_hj) {
      h = hj_relax<F>(d, J, (I == 0) ? d.ni : (I == d.ni + 1 ? 1 : I), p, t, pk);
    } else {
      const double pk0 = (d.press0 == 1000. * 100.) ? pk : m_pow<F>(p / (1000. * 100.), d.kappa);
      h = 0. + hs_relax<F>(d, J, p, t, d.P3[ix3(d, d.nk, J, I)], pk0);
    }
    d.HEAT3[q] = h;
  }
  d.EXNER3[q] = d.cp * t / th;
  if (K >= 1) {
    double p2 = d.P2[q], th2 = d.THETA2[q];
    double t2 = th2 * m_pow<F>(p2 / d.press0, d.kappa);
    d.T2[q] = t2;
    d.RHO2[q] = p2 * d.mu2 / (R_GAS * t2);
  } else {
    const size_t q1 = ix3(d, 1, J, I);
    double p21 = d.P2[q1], th21 = d.THETA2[q1];
    double t21 = th21 * m_pow<F>(p21 / d.press0, d.kappa);
    double rho21 = p21 * d.mu2 / (R_GAS * t21);
    d.RHO2[q] = rho3 * rho3 / rho21;
    d.THETA2[q] = d.sgth[0];
    d.T2[q] = t;
  }
}

// store_pgrad_vars' per-cell part with full = 1, one lane per (K,J,I),
// K = 0..nk, J = 0..nj, I = 0..ni+1 (after the FULL set_p2etc)
template <bool F>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(8))) k_thermo(const MemberDesc *D, int mb) {
  MEMBER;
  if (d.dyn->itime <= 1) {
    // no lane of this kernel writes DZDT3 while itime <= 1
    const size_t nthr = (size_t)gridDim.x * gridDim.y * blockDim.x;
    for (size_t i = ((size_t)blockIdx.y * gridDim.x + blockIdx.x) * blockDim.x + threadIdx.x; i < (size_t)d.N3;
         i += nthr)
      d.DZDT3[i] = 0.;
  }
  FLAT2(0, d.nj + 1, 0, d.ni + 2);
  const int K = blockIdx.y;
  const size_t q = ix3(d, K, J, I);
  thermo_cell<F>(d, K, J, I, true, d.THETA[q], d.P3[q]);
}
void L_thermo(hipStream_t s, const Bucket &b) {
  if (b.mixed) LAUNCH_AS("k_thermo", k_thermo<true>, grid2(b.nj + 1, b.ni + 2, b.nk + 1, b.nm), dim3(TPB), 0, s, b.D, b.mb);
  else LAUNCH_AS("k_thermo", k_thermo<false>, grid2(b.nj + 1, b.ni + 2, b.nk + 1, b.nm), dim3(TPB), 0, s, b.D, b.mb);
}

// ------------------------------------------------------------ RI2
// Richardson number in the layers, N^2 / ((du)^2 + (dv)^2) (get_richardson
// with kk = 2K, src/shared/epic_funcs_diag.c:3066-3156; the reference forms
// dz_inv there and leaves it unused, and its v difference carries a + on
// V(K,J+1)), with get_brunt2 of a layer (:2885-3064, no smoothing:
// Drho/Dp from the interfaces above and below, the partial derivatives of
// the ideal-gas density by centred differences of 0.1 %), return_cp
// (:3311-3337: 1004 for Held-Suarez, else the centred derivative of the
// enthalpy cpr*T*rgas of x_h2 = 0) and the layer molar mass mu2.
// Store_diag computes it for J = 0..nj, I = 1..ni, K = 1..nk and applies
// bc_lateral; every other cell stays zero.
__global__ void __launch_bounds__(256) k_richardson(const MemberDesc *D, int mb) {
  MEMBER;
  FLAT2(0, d.nj + 1, 1, d.ni);
  const int K = 1 + blockIdx.y;
  const double *U = d.U, *V = d.V;
  double dudz2 = .5 * (U[ix3(d, K - 1, J, I)] + U[ix3(d, K - 1, J, I + 1)] - U[ix3(d, K, J, I)] -
                       U[ix3(d, K, J, I + 1)]);
  double dvdz2 = .5 * (V[ix3(d, K - 1, J, I)] + V[ix3(d, K - 1, J + 1, I)] - V[ix3(d, K, J, I)] +
                       V[ix3(d, K, J + 1, I)]);
  dudz2 *= dudz2;
  dvdz2 *= dvdz2;
  double ri = 1.e+6;
  if (fcmp_d(dudz2, 0.) != 0) {
    const double R_GAS = 8.314472e+3;
    const size_t q = ix3(d, K, J, I);
    const double pressure = d.P2[q], temperature = d.T2[q], density = d.RHO2[q];
    double cp;
    if (d.is_hs) {
      cp = 1004.;
    } else {
      const double dT = temperature * 1.e-6;
      double h2 = d.cpr * (temperature + dT), h1 = d.cpr * (temperature - dT);
      h2 *= d.rgas;
      h1 *= d.rgas;
      cp = (h2 - h1) / (2. * dT);
    }
    const double mu = d.mu2, deltap = 0.001 * pressure, deltaT = 0.001 * temperature;
    const double drho_dp_T = ((pressure + deltap) * mu / (R_GAS * temperature) -
                              (pressure - deltap) * mu / (R_GAS * temperature)) / (2. * deltap);
    const double drho_dT_p = (pressure * mu / (R_GAS * (temperature + deltaT)) -
                              pressure * mu / (R_GAS * (temperature - deltaT))) / (2. * deltaT);
    const double Drho_Dp = (d.RHO3[q] - d.RHO3[ix3(d, K - 1, J, I)]) / (d.P3[q] - d.P3[ix3(d, K - 1, J, I)]);
    const double g = d.g[2 * J + 1];
    const double brunt2 =
        g * g * (Drho_Dp - drho_dp_T + temperature / (cp * density * density) * drho_dT_p * drho_dT_p);
    ri = brunt2 / (dudz2 + dvdz2);
  }
  d.RI2[ix3(d, K, J, I)] = ri;
  if (I == d.ni) d.RI2[ix3(d, K, J, 0)] = ri;
  if (I == 1) d.RI2[ix3(d, K, J, d.ni + 1)] = ri;
}
void L_richardson(hipStream_t s, const Bucket &b) {
  LAUNCH(k_richardson, grid2(b.nj + 1, b.ni, b.nk, b.nm), dim3(TPB), 0, s, b.D, b.mb);
}

// kinetic energy of a zonally uniform u row (get_kin with v = 0)
__device__ double kin_uniform(const double *mn, const double *mni, int nj, const double *u1, int J, bool al81) {
  int jj = 2 * J;
  double u2 = u1[J], u4 = u1[J], v2 = 0., v4 = 0.;
  double kc = (u2 * u2 + u4 * u4 + mn[jj + 1] * (v2 * v2 * mni[jj] + v4 * v4 * mni[jj + 2])) * .25;
  if (al81 || J == 0 || J == nj) return kc;
  u2 = .5 * (u1[J + 1] + u1[J - 1]);
  u4 = .5 * (u1[J + 1] + u1[J - 1]);
  v2 = .5 * (0. + 0.);
  v4 = .5 * (0. + 0.);
  double ks = (u2 * u2 + u4 * u4 + mn[jj + 1] * (v2 * v2 * mni[jj] + v4 * v4 * mni[jj + 2])) * .25;
  return (5. / 6.) * kc + (1. / 6.) * ks;
}

// Gas giants: GZ_SURFACE from gradient balance with the zonal-mean bottom
// wind, in two launches: the zonal means of u, p, rho on the bottom
// interface (one wavefront per row; lanes 0-2 sum in the reference's
// sequential order), the meridional integral (one block); the column
// kernel k_gz3 that must follow copies the results into GZ_SURFACE and
// PV3's plane nk+1 as it starts each column.
template <bool F>
__global__ void __launch_bounds__(256) k_gzu_means(const MemberDesc *D, int mb) {
  MEMBER;
  if (!d.is_gas) return;
  extern __shared__ double row[];
  const int J = blockIdx.x, ni = d.ni, nk = d.nk, R = d.nj + 3;
  // RHO3 of the bottom interface formed here from P3 and theta, as
  // store_pgrad_vars forms it (the same values), so that this launch need
  // not wait for the kernel that writes RHO3
  for (int i = threadIdx.x; i < ni; i += blockDim.x) {
    const size_t q = ix3(d, nk, J, i + 1);
    const double p = d.P3[q];
    double pk;
    row[i] = d.U[q];
    row[ni + i] = p;
    row[2 * ni + i] = rho3_of(d, p, t3_of<F>(d, d.THETA[q], p, &pk));
  }
  __syncthreads();
  if (threadIdx.x < 3) d.rowbuf[threadIdx.x * R + J] = seq_acc(row + threadIdx.x * ni, ni) / ni;
}
__global__ void __launch_bounds__(256) k_gzu_core(const MemberDesc *D, int mb) {
  MEMBER;
  if (!d.is_gas) return;
  extern __shared__ double sh[];
  const int nj = d.nj, ni = d.ni, R = nj + 3, NJ = 2 * (nj + 1) + 1;
  double *u1 = sh, *p1 = sh + R, *r1 = sh + 2 * R, *pvr = sh + 3 * R, *pvh = sh + 4 * R, *bern = sh + 5 * R;
  double *kn = sh + 6 * R;
  // the metric columns the serial integral reads, staged once
  double *f = sh + 7 * R, *m = f + NJ, *n = m + NJ, *mn = n + NJ, *mni = mn + NJ;
  for (int J = threadIdx.x; J <= nj; J += blockDim.x) {
    u1[J] = d.rowbuf[J];
    p1[J] = d.rowbuf[R + J];
    r1[J] = d.rowbuf[2 * R + J];
  }
  for (int jj = threadIdx.x; jj < NJ; jj += blockDim.x) {
    f[jj] = d.f[jj];
    m[jj] = d.m[jj];
    n[jj] = d.n[jj];
    mn[jj] = d.mn[jj];
    mni[jj] = d.mn_inv[jj];
  }
  __syncthreads();
  // absolute vorticity of the uniform row field
  for (int J = 1 + threadIdx.x; J <= nj; J += blockDim.x) {
    double f2j = f[2 * J], m2j = m[2 * J], n2j = n[2 * J];
    double mm1i = 1. / m[2 * J - 1], mp1i = 1. / m[2 * J + 1];
    double ze = m2j * ((0. - 0.) + n2j * (u1[J - 1] * mm1i - u1[J] * mp1i));
    pvr[J] = (ze + f2j);
  }
  if (threadIdx.x == 0) {
    double ze = 0.;
    for (int I = 1; I <= ni; I++) ze -= u1[0];
    ze *= mn[0] / ((m[1]) * (double)(ni));
    pvr[0] = ze + f[0];
  } else if (threadIdx.x == 64) {
    double ze = 0.;
    for (int I = 1; I <= ni; I++) ze += u1[nj];
    ze *= mn[2 * (nj + 1)] / (m[2 * nj + 1] * (double)(ni));
    pvr[nj + 1] = ze + f[2 * (nj + 1)];
  }
  __syncthreads();
  for (int J = threadIdx.x; J <= nj + 1; J += blockDim.x) {
    double v = 0.;
    if (J >= 1 && J <= nj) {
      double udyJ = u1[J] / n[2 * J + 1], udyM = u1[J - 1] / n[2 * J - 1];
      // the uniform rows' PV in the weights of the v grid
      auto P = [&](int j, int) { return pvr[j]; };
      const PvW<double> w = pv_weights_v<double>(d.pv_al81, P, J, 1);
      v = (w.ga * udyJ + w.de * udyJ + w.al * udyM + w.be * udyM) * w.coef;
      v += (p1[J] - p1[J - 1]) * 2. / (r1[J] + r1[J - 1]);
    }
    pvh[J] = v;
    if (J <= nj) kn[J] = kin_uniform(mn, mni, nj, u1, J, d.pv_al81);
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    // Bernoulli function integrated from the tie-point row both ways
    // (running values in a register, in the same order; the increments are
    // read eight at a time ahead of the chain)
    const int j0 = d.jtp;
    const double b0 = 0. + kn[j0];
    bern[j0] = b0;
    double b = b0;
    int J = j0 + 1;
    for (; J + 7 <= nj; J += 8) {
      double q[8];
#pragma unroll
      for (int u = 0; u < 8; u++) q[u] = pvh[J + u];
#pragma unroll
      for (int u = 0; u < 8; u++) {
        b = b - q[u];
        bern[J + u] = b;
      }
    }
    for (; J <= nj; J++) {
      b = b - pvh[J];
      bern[J] = b;
    }
    b = b0;
    J = j0 - 1;
    for (; J - 7 >= 0; J -= 8) {
      double q[8];
#pragma unroll
      for (int u = 0; u < 8; u++) q[u] = pvh[J - u + 1];
#pragma unroll
      for (int u = 0; u < 8; u++) {
        b = b + q[u];
        bern[J - u] = b;
      }
    }
    for (; J >= 0; J--) {
      b = b + pvh[J + 1];
      bern[J] = b;
    }
  }
  __syncthreads();
  for (int J = threadIdx.x; J <= nj + 1; J += blockDim.x) {
    d.rowbuf[3 * R + J] = pvr[J];
    if (J <= nj) d.rowbuf[4 * R + J] = bern[J] - kn[J];
  }
}
void L_gz_from_u(hipStream_t s, const Bucket &b) {
  if (!b.any_gas) return;
  if (b.mixed) LAUNCH_AS("k_gzu_means", k_gzu_means<true>, dim3(b.nj + 1, 1, b.nm), dim3(64), sizeof(double) * 3 * b.ni, s, b.D, b.mb);
  else LAUNCH_AS("k_gzu_means", k_gzu_means<false>, dim3(b.nj + 1, 1, b.nm), dim3(64), sizeof(double) * 3 * b.ni, s, b.D, b.mb);
  LAUNCH(k_gzu_core, dim3(1, 1, b.nm), dim3(256), sizeof(double) * (7 * (b.nj + 3) + 5 * (2 * b.nj + 3)), s, b.D,
         b.mb);
  // GZ_SURFACE and PV3's plane nk+1 are filled in by the k_gz3 that follows
}

// hydrostatic GZ3 and MONT3 from the bottom up (Leslie-Purser 4-point
// integral in ln p) and the monotonicity screen
template <bool F>
__global__ void __launch_bounds__(1024) __attribute__((amdgpu_waves_per_eu(6))) k_gz3(const MemberDesc *D, int mb, int C) {
  MEMBER;
  extern __shared__ double lds[];
  const Tile T = tile_of(d, C);
  const int nk = d.nk, W = nk + 1, J = T.J, I = T.I;
  double *inc = lds + T.c * W, *tt = lds + C * W + T.c * W, *lg = lds + 2 * C * W + T.c * W;
  const double *P = d.P3, *RH = d.RHO3;
#define PP(k) P[ix3(d, k, J, I)]
#define RR(k) RH[ix3(d, k, J, I)]
  // the column's surface geopotential, read ahead by the lane that runs the
  // chain; on a gas giant gz_from_u's results for this column: GZ_SURFACE,
  // and plane nk+1 of PV3 (the absolute vorticity of the uniform rows; row
  // nj+1 by the columns of row nj)
  double gs = 0.;
  if (T.live && T.kk == 0) {
    if (d.is_gas) {
      const int R = d.nj + 3;
      gs = d.rowbuf[4 * R + J];
      d.GZS[ix2(d, J, I)] = gs;
      double *pv = d.PV3 + (size_t)(d.nk + 1) * d.N2;
      pv[ix2(d, J, I)] = d.rowbuf[3 * R + J];
      if (J == d.nj) pv[ix2(d, J + 1, I)] = d.rowbuf[3 * R + J + 1];
    } else {
      gs = d.GZS[ix2(d, J, I)];
    }
  }
  // lg[k] = log(P3(k)/P3(k-1)): every ln-p interval the integral uses, once
  if (T.live)
    for (int K = T.kk; K <= nk; K += T.KS) {
      tt[K] = d.T3[ix3(d, K, J, I)];
      if (K >= 1) lg[K] = m_log<F>(PP(K) / PP(K - 1));
    }
  __syncthreads();
  if (T.live)
    for (int K = 1 + T.kk; K <= nk; K += T.KS) {
      double v;
      if (K == nk || K == 1) {
        v = ((PP(K) + PP(K - 1)) / (RR(K) + RR(K - 1))) * lg[K];
      } else {
        double dz10 = lg[K + 1], dz21 = lg[K], dz32 = lg[K - 1];
        double dz20 = dz21 + dz10, dz30 = dz32 + dz20, dz31 = dz32 + dz21;
        double df0 = PP(K + 1) / RR(K + 1), df1 = PP(K) / RR(K), df2 = PP(K - 1) / RR(K - 1),
               df3 = PP(K - 2) / RR(K - 2);
        v = -(df0 * (dz21 * dz21 * dz21 * (dz31 + dz32)) / (dz10 * dz20 * dz30) +
              df1 * (dz21 * (dz21 * (dz21 + 2. * (dz10 - dz31)) - 6. * dz10 * dz31)) / (dz10 * dz31) +
              df2 * (dz21 * (dz21 * (dz21 + 2. * (dz32 - dz20)) - 6. * dz20 * dz32)) / (dz20 * dz32) +
              df3 * (dz21 * dz21 * dz21 * (dz10 + dz20)) / (dz30 * dz31 * dz32)) /
            12.;
      }
      inc[K] = v;
    }
#undef PP
#undef RR
  __syncthreads();
  if (T.live && T.kk == 0) {
    const double cpr = d.cpr, rgas = d.rgas;
    double *G = d.GZ3, *M = d.MONT3;
    double gprev = d.topo_scale * gs;
    G[ix3(d, nk, J, I)] = gprev;
    M[ix3(d, nk, J, I)] = gprev + cpr * tt[nk] * rgas;
    const bool inner = (I >= 1 && I <= d.ni);
    int bad = 0;
    // the increments and temperatures are read eight levels at a time ahead
    // of the chain
    int K = nk;
    for (; K - 7 >= 1; K -= 8) {
      double i8[8], t8[8];
#pragma unroll
      for (int u = 0; u < 8; u++) {
        i8[u] = inc[K - u];
        t8[u] = tt[K - u - 1];
      }
#pragma unroll
      for (int u = 0; u < 8; u++) {
        double gn = gprev + i8[u];
        G[ix3(d, K - u - 1, J, I)] = gn;
        M[ix3(d, K - u - 1, J, I)] = gn + cpr * t8[u] * rgas;
        if (inner && gprev >= gn) bad = 1;
        gprev = gn;
      }
    }
    for (; K >= 1; K--) {
      double gn = gprev + inc[K];
      G[ix3(d, K - 1, J, I)] = gn;
      M[ix3(d, K - 1, J, I)] = gn + cpr * tt[K - 1] * rgas;
      if (inner && gprev >= gn) bad = 1;
      gprev = gn;
    }
    if (bad) atomicOr(&d.dyn->err, 1);
  }
}
void L_gz3(hipStream_t s, const Bucket &b) {
  const int v = b.knob[KN_COLS_GZ3], C = v ? knob_cols(v) : tile_cols(b, 3);
  const size_t shm = sizeof(double) * 3 * C * (b.nk + 1);
  if (b.mixed) LAUNCH_AS("k_gz3", k_gz3<true>, tile_grid(b, C), dim3(knob_threads(v, b.nk + 1)), shm, s, b.D, b.mb, C);
  else LAUNCH_AS("k_gz3", k_gz3<false>, tile_grid(b, C), dim3(knob_threads(v, b.nk + 1)), shm, s, b.D, b.mb, C);
}

// finish DZDT3 (diagnostic vertical velocity)
__global__ void __launch_bounds__(256) k_dzdt3(const MemberDesc *D, int mb) {
  MEMBER_X;
  FLAT2_X(0, d.nj + 1, 1, d.ni);
  const int K = 1 + bx_.y;
  size_t q = ix3(d, K, J, I);
  const double *G = d.GZ3;
  double z = d.DZDT3[q];
  if (d.dyn->itime > 1) {
    double dtg_inv = -1. / (d.dt * d.g[2 * J + 1]);
    z -= G[q];
    z *= dtg_inv;
  }
  z += d.W3[q] * d.H3[q] / d.RHO3[q];
  int jj = 2 * J;
  double gx3 = .5 * d.m[jj + 1] / d.g[jj + 1];
  double gy2 = .5 * d.n[jj] / d.g[jj];
  double gy4 = .5 * d.n[jj + 2] / d.g[jj + 2];
  z += d.V[q] * (G[q] - G[ix3(d, K, J - 1, I)]) * gy2 +
       d.V[ix3(d, K, J + 1, I)] * (G[ix3(d, K, J + 1, I)] - G[q]) * gy4 +
       d.U[q] * (G[q] - G[ix3(d, K, J, I - 1)]) * gx3 +
       d.U[ix3(d, K, J, I + 1)] * (G[ix3(d, K, J, I + 1)] - G[q]) * gx3;
  d.DZDT3[q] = z;
  if (I == d.ni) d.DZDT3[ix3(d, K, J, 0)] = z;
  if (I == 1) d.DZDT3[ix3(d, K, J, d.ni + 1)] = z;
}
void L_dzdt3(hipStream_t s, const Bucket &b) {
  LAUNCH(k_dzdt3, grid2(b.nj + 1, b.ni, b.nk, b.nm), dim3(TPB), 0, s, b.D, b.mb);
}

// ------------------------------------------------------------ heating
// the Held-Suarez relaxation of a cell from its p, T, the column's bottom
// pressure and (p/p0)^kappa, p0 = 1000 hPa
template <bool F>
__device__ __forceinline__ double hs_relax(const MemberDesc &d, int J, double p, double t, double pbot, double pk0) {
  const double ka = 1. / (40. * 60. * 60. * 24.), ks = 1. / (4. * 60. * 60. * 24.);
  const double dt_y = 60., dth_z = 10., p0 = 1000. * 100.;
  double kt = ka;
  double amp = (p / pbot - 0.7) / (1. - 0.7);
  // cos(lat) and its fourth power from the host tables (the host libm, as
  // the oracle forms them)
  if (amp > 0.) kt += (ks - ka) * d.coslat4_h[J] * amp;
  double t_cool = 1. / kt;
  double cpt = 1004. / t_cool;
  double c2 = d.coslat_h[J];
  c2 *= c2;
  double s2 = 1. - c2, pp0 = p / p0;
  double teq = (315. - dt_y * s2 - dth_z * m_log<F>(pp0) * c2) * pk0;
  teq = dmax(200., teq);
  return -cpt * (t - teq);
}
__device__ double hs_newton(const MemberDesc &d, int K, int J, int I) {
  const double p = d.P3[ix3(d, K, J, I)];
  return hs_relax(d, J, p, d.T3[ix3(d, K, J, I)], d.P3[ix3(d, d.nk, J, I)], crm_pow(p / (1000. * 100.), d.kappa));
}

// Day-night relaxation of the hot-Jupiter case (no reference counterpart):
// HEAT3 = cp (T_eq - T)/tau_rad, T_eq = T_ref(p) + dT(p) (max(0, cos(lat)
// cos(lon)) - 1/4), T_ref the initial profile's temperature, dT the full
// contrast above teq_p_contrast and linear in ln p to 0 at pbot.
// the day-night relaxation of a cell from its p, T, (p/press0)^kappa and the
// column I whose longitude it takes
template <bool F>
__device__ __forceinline__ double hj_relax(const MemberDesc &d, int J, int I, double p, double t, double pk) {
  const double tref = (1. / (1. / d.ref_theta_b + d.ref_a_over_kappa * (pk - d.ref_pbk))) * pk;
  double dT = d.teq_contrast;
  // the per-member log and the per-row/column cosines come from host tables
  // formed with the host libm, as the oracle forms them
  if (p > d.teq_p_contrast) dT *= dmax(0., m_log<F>(d.pbot / p) / d.hj_lnref);
  const double cl = d.coslat_h[J];
  const double day = dmax(0., cl * d.coslon_h[I]);
  const double teq = tref + dT * (day - .25);
  return d.cp * (teq - t) / d.tau_rad;
}
__device__ double hj_newton(const MemberDesc &d, int K, int J, int I) {
  const double p = d.P3[ix3(d, K, J, I)];
  return hj_relax(d, J, I, p, d.T3[ix3(d, K, J, I)], crm_pow(p / d.press0, d.kappa));
}

// HEAT3 = 0 everywhere, then the Held-Suarez, day-night or sounding
// relaxation on K = 1..nk with its zonal pads (the bc_lateral at the end of
// calc_heating)
__global__ void __launch_bounds__(256) k_heating(const MemberDesc *D, int mb) {
  MEMBER;
  FLAT2(0, d.nj + 1, 1, d.ni);
  const int K = 1 + blockIdx.y;
  if (!newt_relax_on(d)) return;
  double h;
  if (d.is_tab) {
    const size_t q = ix3(d, K, J, I);
    h = 0. + tab_relax(d, d.P3[q], d.T3[q]);
  } else {
    h = d.is_hj ? hj_newton(d, K, J, I) : 0. + hs_newton(d, K, J, I);
  }
  d.HEAT3[ix3(d, K, J, I)] = h;
  if (I == d.ni) d.HEAT3[ix3(d, K, J, 0)] = h;
  if (I == 1) d.HEAT3[ix3(d, K, J, d.ni + 1)] = h;
}
// Layer-mean removal of the relaxation (K <= k_sponge or newt_cool_adjust,
// src/shared/epic_sensible_heating.c:139-167), from the relaxation HEAT3
// holds: one block per layer stages rows of HEAT3 times the cell area
// 1/mn in LDS, and one lane adds them in the reference's order (J, then I);
// the total area, which depends on the metric only, is formed on the host
// by the same loop.  The layer mean goes to sD[K].
__global__ void __launch_bounds__(256) k_heat_layer_mean(const MemberDesc *D, int mb) {
  MEMBER;
  extern __shared__ double lds[];
  const int K = 1 + blockIdx.x;
  if (!newt_relax_on(d)) return;
  if (!(K <= d.k_sponge || d.newt_cool_adjust)) return;
  const int ni = d.ni, nrow = d.nj + 1;
  const int rpc = (2048 / ni) > 0 ? 2048 / ni : 1;  // rows per LDS chunk
  double ha = 0.;
  for (int j0 = 0; j0 < nrow; j0 += rpc) {
    const int nr = (nrow - j0) < rpc ? (nrow - j0) : rpc;
    for (int t = threadIdx.x; t < nr * ni; t += blockDim.x) {
      const int J = j0 + t / ni, I = 1 + t % ni;
      const double da = 1. / d.mn[2 * J + 1];
      lds[t] = d.HEAT3[ix3(d, K, J, I)] * da;
    }
    __syncthreads();
    if (threadIdx.x == 0) ha = seq_acc(lds, nr * ni, ha);
    __syncthreads();
  }
  if (threadIdx.x == 0) d.sD[K] = ha / d.heat_area;
}
__global__ void __launch_bounds__(256) k_heat_adjust_apply(const MemberDesc *D, int mb) {
  MEMBER;
  FLAT2(0, d.nj + 1, 0, d.ni + 2);
  const int K = 1 + blockIdx.y;
  if (!newt_relax_on(d)) return;
  if (!(K <= d.k_sponge || d.newt_cool_adjust)) return;
  const size_t q = ix3(d, K, J, I);
  d.HEAT3[q] = 0. + (d.HEAT3[q] - d.sD[K]);
}
// in_thermo: the relaxation was formed by this step's thermo_cell (full = 1),
// and every other cell of HEAT3 is zero (a member's blob is cleared when it
// is created, a drop-in upload brings the caller's array, which the
// reference zeroes there, and no kernel writes outside K = 1..nk, J = 0..nj),
// so only the layer-mean adjustment is left
void L_heating(hipStream_t s, const Bucket &b, int adjust, bool in_thermo) {
  if (!in_thermo) {
    L_fill(s, b, F_HEAT3, 0.);
    LAUNCH(k_heating, grid2(b.nj + 1, b.ni, b.nk, b.nm), dim3(TPB), 0, s, b.D, b.mb);
  }
  if (adjust) {
    LAUNCH(k_heat_layer_mean, dim3(b.nk, 1, b.nm), dim3(256), sizeof(double) * 2048, s, b.D, b.mb);
    LAUNCH(k_heat_adjust_apply, grid2(b.nj + 1, b.ni + 2, b.nk, b.nm), dim3(TPB), 0, s, b.D, b.mb);
  }
}

// ------------------------------------------------------------ calc_w
// Column tiles.  Per level: the theta heating of source_sink
// (src/shared/epic_funcs_diag.c:4768-4778, with_source), the heating part
// of W3, the Leslie-Purser increment of SUM_DIV and dF/dp; then the running
// SUM_DIV in one lane per column.  SUM_DIV(1) and SUM_DIV(nk) are zero (the
// reference's top/bottom statements write stale-index pad cells), so the
// dF/dpbot term vanishes and is not formed.
// With with_source (the step's call) the blocks past the ncb column tiles
// first clear the step's tendency planes of u and v (the FP64 plane IT_ZERO,
// or the mixed mode's one FP64 plane), which nothing in this launch reads.
template <bool F>
__global__ void __launch_bounds__(1024) k_calc_w(const MemberDesc *D, int mb, int with_source, int C, int ncb) {
  MEMBER;
  if ((int)blockIdx.x >= ncb) {
    const size_t off = d.mixed ? 0 : (size_t)d.dyn->it0 * d.N3;
    double *du = d.DUDT + off, *dv = d.DVDT + off;
    const size_t st = (size_t)(gridDim.x - ncb) * blockDim.x;
    for (size_t i = (size_t)(blockIdx.x - ncb) * blockDim.x + threadIdx.x; i < (size_t)d.N3; i += st) {
      du[i] = 0.;
      dv[i] = 0.;
    }
    return;  // a whole block: no barrier below waits for it
  }
  extern __shared__ double lds[];
  const Tile T = tile_of(d, C);
  const int nk = d.nk, ks = d.ks, W = nk + 1, J = T.J, I = T.I;
  double *dsv = lds + T.c * W, *dfv = lds + C * W + T.c * W, *wbv = lds + 2 * C * W + T.c * W;
  double *lg = lds + 3 * C * W + T.c * W, *lrv = lds + 4 * C * W + T.c;
  const double *P = d.P3;
  double *TH = d.THETA;
  int bad = 0;
#define PP(k) P[ix3(d, k, J, I)]
#define DV(k) d.DIV[ix3(d, k, J, I)]
  // lg[k] = log(P3(k)/P3(k-1)), the ln-p intervals of the Leslie-Purser weights
  // and log(ptop/pbot) once per column
  if (T.live) {
    for (int K = 1 + T.kk; K <= nk; K += T.KS) lg[K] = m_log<F>(PP(K) / PP(K - 1));
    if (T.kk == 0) lrv[0] = m_log<F>(PP(0) / PP(nk));
  }
  __syncthreads();
  if (T.live) {
    const double ptop = P[ix3(d, 0, J, I)], pbot = P[ix3(d, nk, J, I)], psig = P[ix3(d, ks - 1, J, I)];
    const double eps = d.eps13, lr = lrv[0];
    for (int K = 1 + T.kk; K <= nk; K += T.KS) {
      const size_t q = ix3(d, K, J, I);
      double th = TH[q];
      if (with_source && K >= ks - 1) {
        th += d.dt * d.HEAT3[q] / d.EXNER3[q];
        TH[q] = th;
      }
      if (K == nk) continue;
      double wb = 0.;
      if (K < ks - 1) {
        double sigma = get_sigma_l<F>(pbot, P[q], lr);
        wb = g_sig<F>(d, sigma) * d.HEAT3[q] / d.EXNER3[q];
      }
      double ds = 0.;
      if (K >= 2) {
        double dz10 = lg[K - 1], dz21 = lg[K], dz32 = lg[K + 1];
        double dz20 = dz21 + dz10, dz30 = dz32 + dz20, dz31 = dz32 + dz21;
        double df0 = PP(K - 2) * DV(K - 2), df1 = PP(K - 1) * DV(K - 1), df2 = PP(K) * DV(K),
               df3 = PP(K + 1) * DV(K + 1);
        ds = -(df0 * (dz21 * dz21 * dz21 * (dz31 + dz32)) / (dz10 * dz20 * dz30) +
               df1 * (dz21 * (dz21 * (dz21 + 2. * (dz10 - dz31)) - 6. * dz10 * dz31)) / (dz10 * dz31) +
               df2 * (dz21 * (dz21 * (dz21 + 2. * (dz32 - dz20)) - 6. * dz20 * dz32)) / (dz20 * dz32) +
               df3 * (dz21 * dz21 * dz21 * (dz10 + dz20)) / (dz30 * dz31 * dz32)) /
             12.;
      }
      double p3 = P[q], pb3 = pbot;
      double dp = p3 * eps;
      double p2 = dmax(p3 - dp * .5, ptop), p4 = dmin(p3 + dp * .5, pbot);
      if (p2 < psig && p4 >= psig) p2 = p3;
      double dFdp = 0.;
      if (p2 < p4)
        dFdp = (sgth_of_l<F>(d, th, p2, pb3, ptop, lr, &bad) - sgth_of_l<F>(d, th, p4, pb3, ptop, lr, &bad)) / (p2 - p4);
      else
        bad = 1;
      dsv[K] = ds;
      dfv[K] = dFdp;
      wbv[K] = wb;
    }
  }
#undef PP
#undef DV
  __syncthreads();
  if (T.live && T.kk == 0) {
    double *Wf = d.W3;
    Wf[ix3(d, 0, J, I)] = 0.;
    Wf[ix3(d, nk, J, I)] = 0.;
    // (the three columns read eight levels at a time ahead of the chain)
    double S = 0.;
    int K = 1;
    for (; K + 7 < nk; K += 8) {
      double s8[8], w8[8], f8[8];
#pragma unroll
      for (int u = 0; u < 8; u++) {
        s8[u] = dsv[K + u];
        w8[u] = wbv[K + u];
        f8[u] = dfv[K + u];
      }
#pragma unroll
      for (int u = 0; u < 8; u++) {
        if (K + u >= 2) S = S + s8[u];
        Wf[ix3(d, K + u, J, I)] = w8[u] - S * f8[u];
      }
    }
    for (; K < nk; K++) {
      if (K >= 2) S = S + dsv[K];
      Wf[ix3(d, K, J, I)] = wbv[K] - S * dfv[K];
    }
  }
  if (bad) atomicOr(&d.dyn->err, 2);
}
void L_calc_w(hipStream_t s, const Bucket &b, int with_source) {
  const int v = b.knob[KN_COLS_CALCW], C = v ? knob_cols(v) : tile_cols(b, 4), nt = knob_threads(v, b.nk + 1);
  dim3 g = tile_grid(b, C);
  const int ncb = (int)g.x;
  if (with_source) {
    // blocks that clear the tendency planes: about 16 cells per lane
    const size_t n3 = (size_t)(b.ni + 2) * (b.nj + 3) * (b.nk + 2);
    g.x += (unsigned)std::max<size_t>(1, std::min<size_t>(1024, n3 / ((size_t)nt * 16)));
  }
  const size_t shm = sizeof(double) * (4 * C * (b.nk + 1) + C);
  if (b.mixed) LAUNCH_AS("k_calc_w", k_calc_w<true>, g, dim3(nt), shm, s, b.D, b.mb, with_source, C, ncb);
  else LAUNCH_AS("k_calc_w", k_calc_w<false>, g, dim3(nt), shm, s, b.D, b.mb, with_source, C, ncb);
}

// ------------------------------------------------ vorticity / divergence
// PV3 planes k0..k1 from U (uu) and the zonally filtered V (vv), with H3.
// One block per (K,J) pv-row group: row J of vv is filtered in LDS (rows of
// the v-grid filter band), then one lane per I.  bx is the row group.
__device__ __forceinline__ void vort_interior_rows(const MemberDesc &d, double *lds, int bx, int K, int uf, int vf) {
  const RowMap m = row_map(d.ni);
  const int Jr = 1 + bx * m.rpb + m.g, ni = d.ni;
  const bool live = Jr <= d.nj;
  const int J = live ? Jr : 1;
  const double *uu = fld_ptr(d, uf) + (size_t)K * d.N2, *vv = fld_ptr(d, vf) + (size_t)K * d.N2;
  const double *hh = d.H3 + (size_t)K * d.N2;
  double *pv = d.PV3 + (size_t)K * d.N2;
  double *xv = lds + (size_t)m.g * 2 * ni, *work = xv + ni;  // xv[I-1] = filtered vv(J,I)
  if (live)
    for (int I = 1 + m.lane; I <= ni; I += m.stride) xv[I - 1] = vv[ix2(d, J, I)];
  __syncthreads();
  filter_row_lds(d, xv, work, J, 1, live && d.rflag_v[J] != 0, m.lane);
  __syncthreads();
  if (!live) return;
  double f2j = d.f[2 * J], m2j = d.m[2 * J], n2j = d.n[2 * J];
  double mm1i = 1. / d.m[2 * J - 1], mp1i = 1. / d.m[2 * J + 1];
  double mnm1i = 1. / d.mn[2 * J - 1], mnp1i = 1. / d.mn[2 * J + 1];
  double mn_pv = .5 / (mnm1i + mnp1i);
  for (int I = 1 + m.lane; I <= ni; I += m.stride) {
    const double vc = xv[I - 1], vw = xv[I == 1 ? ni - 1 : I - 2];
    double ze = m2j * ((vc - vw) + n2j * (uu[ix2(d, J - 1, I)] * mm1i - uu[ix2(d, J, I)] * mp1i));
    double v = (ze + f2j);
    double h_pv = ((hh[ix2(d, J, I)] + hh[ix2(d, J, I - 1)]) * mnp1i +
                   (hh[ix2(d, J - 1, I)] + hh[ix2(d, J - 1, I - 1)]) * mnm1i) * mn_pv;
    v /= h_pv;
    pv[ix2(d, J, I)] = v;
    if (I == ni) pv[ix2(d, J, 0)] = v;
    if (I == 1) pv[ix2(d, J, ni + 1)] = v;
  }
}
// pole PV from the circulation of the adjacent u-row (north = 1: J = nj+1
// from row nj, else J = 0 from row 0): the row is staged in LDS and lanes
// 0/1 form the two zonal sums in the reference's sequential order.  Every
// lane of the block reaches both barriers.
__device__ __forceinline__ void vort_pole(const MemberDesc &d, double *row, int K, int north, int uf) {
  const double *uu = fld_ptr(d, uf) + (size_t)K * d.N2;
  const double *hh = d.H3 + (size_t)K * d.N2;
  double *pv = d.PV3 + (size_t)K * d.N2;
  const int ni = d.ni, nj = d.nj, Jr = north ? nj : 0;
  for (int i = threadIdx.x; i < ni; i += blockDim.x) {
    row[i] = uu[ix2(d, Jr, i + 1)];
    row[ni + i] = hh[ix2(d, Jr, i + 1)];
  }
  __syncthreads();
  double *res = row + 2 * ni;
  if (threadIdx.x == 0) {
    double ze;
    if (!north) {
      ze = seq_acc<true>(row, ni);
      ze *= d.mn[0] / ((d.m[1]) * (double)(ni));
    } else {
      ze = seq_acc(row, ni);
      ze *= d.mn[2 * (nj + 1)] / (d.m[2 * nj + 1] * (double)(ni));
    }
    res[0] = ze;
  } else if (threadIdx.x == 1) {
    double h = seq_acc(row + ni, ni);
    h /= (double)(ni);
    res[1] = h;
  }
  __syncthreads();
  const double val = (res[0] + (north ? d.f[2 * (nj + 1)] : d.f[0])) / res[1];
  const int Jd = north ? nj + 1 : 0;
  for (int I = threadIdx.x; I <= ni + 1; I += blockDim.x) pv[ix2(d, Jd, I)] = val;
}
// DIV of (uu zonally filtered, vv) on plane K: one block per (K,J) h-row
// group, row J of uu filtered in LDS (rows of the h-grid filter band)
__device__ __forceinline__ void divergence_rows(const MemberDesc &d, double *lds, int bx, int K, int uf, int vf, int df) {
  const RowMap m = row_map(d.ni);
  const int Jr = bx * m.rpb + m.g, ni = d.ni;
  const bool live = Jr <= d.nj;
  const int J = live ? Jr : 0;
  const double *uu = fld_ptr(d, uf) + (size_t)K * d.N2, *vv = fld_ptr(d, vf) + (size_t)K * d.N2;
  double *di = fld_ptr(d, df) + (size_t)K * d.N2;
  double *xu = lds + (size_t)m.g * 2 * ni, *work = xu + ni;  // xu[I-1] = filtered uu(J,I)
  if (live)
    for (int I = 1 + m.lane; I <= ni; I += m.stride) xu[I - 1] = uu[ix2(d, J, I)];
  __syncthreads();
  filter_row_lds(d, xu, work, J, 0, live && d.rflag_h[J] != 0, m.lane);
  __syncthreads();
  if (!live) return;
  int jj = 2 * J + 1;
  double n_inv = 1. / d.n[jj], mn = d.mn[jj];
  double m2j_inv = (J == 0) ? 0. : 1. / d.m[jj - 1];
  double m2jp2_inv = (J == d.nj) ? 0. : 1. / d.m[jj + 1];
  for (int I = 1 + m.lane; I <= ni; I += m.stride) {
    const double ue = xu[I == ni ? 0 : I], uc = xu[I - 1];
    double v = mn * ((ue * n_inv - uc * n_inv) + (vv[ix2(d, J + 1, I)] * m2jp2_inv - vv[ix2(d, J, I)] * m2j_inv));
    di[ix2(d, J, I)] = v;
    if (I == ni) di[ix2(d, J, 0)] = v;
    if (I == 1) di[ix2(d, J, ni + 1)] = v;
  }
}

__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(8))) k_vort_interior(const MemberDesc *D, int mb, int uf, int vf, int k0) {
  MEMBER;
  extern __shared__ double lds[];
  vort_interior_rows(d, lds, blockIdx.x, k0 + blockIdx.y, uf, vf);
}
__global__ void __launch_bounds__(256) k_vort_poles(const MemberDesc *D, int mb, int uf, int k0) {
  MEMBER;
  extern __shared__ double row[];
  vort_pole(d, row, k0 + (blockIdx.x >> 1), blockIdx.x & 1, uf);
}
void L_vorticity(hipStream_t s, const Bucket &b, int uu, int vv, int k0, int k1) {
  int nK = k1 - k0 + 1;
  const int rpb = rows_per_block(b.ni);
  LAUNCH(k_vort_interior, dim3((b.nj + rpb - 1) / rpb, nK, b.nm), dim3(256), sizeof(double) * rpb * 2 * b.ni, s, b.D,
         b.mb, uu, vv, k0);
  LAUNCH(k_vort_poles, dim3(2 * nK, 1, b.nm), dim3(64), sizeof(double) * (2 * b.ni + 2), s, b.D, b.mb, uu, k0);
}

__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(8))) k_divergence(const MemberDesc *D, int mb, int uf, int vf, int df, int k0) {
  MEMBER;
  extern __shared__ double lds[];
  divergence_rows(d, lds, blockIdx.x, k0 + blockIdx.y, uf, vf, df);
}

// store_diag in one launch: on planes K = 0..nk, blocks [0, nv) form the PV
// rows, [nv, nv + nd) the divergence rows, and the last two the pole PV of
// the plane (the three are independent: each reads U, V and H3 only).  With
// rotate set, the first block of plane 0 of each member then rotates its
// time planes (IT_ZERO <- IT_MINUS2 <- IT_MINUS1 <- IT_ZERO) and advances
// itime: the step's last use of them was its Adams-Bashforth update, and no
// block of this launch reads them.
__global__ void __launch_bounds__(256) k_store_diag(const MemberDesc *D, int mb, int nv, int nd, int rotate) {
  MEMBER;
  extern __shared__ double lds[];
  const int bx = blockIdx.x, K = blockIdx.y;
  if (bx < nv) vort_interior_rows(d, lds, bx, K, F_U, F_V);
  else if (bx < nv + nd) divergence_rows(d, lds, bx - nv, K, F_U, F_V, F_DIV);
  else vort_pole(d, lds, K, bx - nv - nd, F_U);
  if (rotate && bx == 0 && K == 0 && threadIdx.x == 0) {
    MemberDyn *y = d.dyn;
    int t = y->itm2;
    y->itm2 = y->itm1;
    y->itm1 = y->it0;
    y->it0 = t;
    y->itime += 1;
  }
}
void L_store_diag(hipStream_t s, const Bucket &b, int rotate) {
  const int rpb = rows_per_block(b.ni);
  const int nv = (b.nj + rpb - 1) / rpb, nd = (b.nj + 1 + rpb - 1) / rpb;
  const size_t shm = sizeof(double) * std::max(rpb * 2 * b.ni, 2 * b.ni + 2);
  LAUNCH(k_store_diag, dim3(nv + nd + 2, b.nk + 1, b.nm), dim3(256), shm, s, b.D, b.mb, nv, nd, rotate);
}
void L_divergence(hipStream_t s, const Bucket &b, int uu, int vv, int dst, int k0, int k1) {
  const int rpb = rows_per_block(b.ni);
  LAUNCH(k_divergence, dim3((b.nj + 1 + rpb - 1) / rpb, k1 - k0 + 1, b.nm), dim3(256), sizeof(double) * rpb * 2 * b.ni,
         s, b.D, b.mb, uu, vv, dst, k0);
}

}  // namespace epic
