// k_dyn.hip -- momentum tendencies, time step and hyperviscosity (FP64).
//
//   uv_core                 src/shared/epic_timestep.c:628-762 with the
//                           Sadourny weights of include/epic_pv_schemes.h:48-61
//   uv_drag                 src/shared/epic_timestep.c:1303-1509
//   uv_pgrad_finite_volume  src/shared/epic_timestep.c:909-1273
//   divergence_damping      src/shared/epic_subgrid.c:1792-1885
//   adams_bashforth_step    src/shared/epic_timestep.c:406-498
//   zonal/meridional hyperviscosity, d2dy2  src/shared/epic_subgrid.c:590-1232
//
// (restore_mass, src/shared/epic_adjust.c:45-96, runs inside the launches
// next to each of its calls: k_adv_elem in k_adv.hip, k_filter3 in
// k_filter.hip.)
//
// The reference carries the finite-volume faces west->east and
// south->north; here every lane recomputes the face shared with its
// neighbour with the same expression, so the result is the same bits
// without the carried dependence.
#include "kernels.h"
#include "filter_row.h"

namespace epic {

#define MEMBER const MemberDesc &d = D[mb + blockIdx.z]
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

__device__ __forceinline__ double *tend(const MemberDesc &d, double *base, int plane) {
  return base + (size_t)plane * d.N3;
}
// the plane the current step's tendency accumulates in: the time plane
// IT_ZERO, or the mixed mode's single FP64 plane
__device__ __forceinline__ double *cur_tend(const MemberDesc &d, double *base) {
  return d.mixed ? base : base + (size_t)d.dyn->it0 * d.N3;
}

// ------------------------------------------------------------- uv_core
// UH (sA), VH (sB), KIN (sC) on planes K = 1..nk-1 with zonal pads; VH rows
// 0 and nj+1 are zero as in the reference's cleared buffers
__global__ void __launch_bounds__(256) k_uvh_kin(const MemberDesc *D, int mb) {
  MEMBER_X;
  FLAT2_X(0, d.nj + 2, 1, d.ni);
  const int K = 1 + bx_.y;
  const double *u = d.U + (size_t)K * d.N2, *v = d.V + (size_t)K * d.N2, *h3 = d.H3 + (size_t)K * d.N2;
  double *uh = d.sA + (size_t)K * d.N2, *vh = d.sB + (size_t)K * d.N2, *kin = d.sC + (size_t)K * d.N2;
  const int ni = d.ni, nj = d.nj;
  if (J <= nj) {
    double havg = .5 * (h3[ix2(d, J, I)] + h3[ix2(d, J, I - 1)]);
    double x = u[ix2(d, J, I)] * havg * d.ninv_h[J];
    uh[ix2(d, J, I)] = x;
    if (I == ni) uh[ix2(d, J, 0)] = x;
    if (I == 1) uh[ix2(d, J, ni + 1)] = x;
    // get_kin (src/shared/epic_funcs_diag.c:2694-2847)
    const double *mn = d.mn, *mni = d.mn_inv;
    int jj = 2 * J;
    double u2 = u[ix2(d, J, I)], u4 = u[ix2(d, J, I + 1)], v2 = v[ix2(d, J, I)], v4 = v[ix2(d, J + 1, I)];
    double kc = (u2 * u2 + u4 * u4 + mn[jj + 1] * (v2 * v2 * mni[jj] + v4 * v4 * mni[jj + 2])) * .25, k;
    if (d.pv_al81 || J == 0 || J == nj) {
      k = kc;
    } else {
      u2 = .5 * (u[ix2(d, J + 1, I)] + u[ix2(d, J - 1, I)]);
      u4 = .5 * (u[ix2(d, J + 1, I + 1)] + u[ix2(d, J - 1, I + 1)]);
      v2 = .5 * (v[ix2(d, J, I + 1)] + v[ix2(d, J, I - 1)]);
      v4 = .5 * (v[ix2(d, J + 1, I + 1)] + v[ix2(d, J + 1, I - 1)]);
      double ks = (u2 * u2 + u4 * u4 + mn[jj + 1] * (v2 * v2 * mni[jj] + v4 * v4 * mni[jj + 2])) * .25;
      k = (5. / 6.) * kc + (1. / 6.) * ks;
    }
    kin[ix2(d, J, I)] = k;
    if (I == ni) kin[ix2(d, J, 0)] = k;
    if (I == 1) kin[ix2(d, J, ni + 1)] = k;
  }
  double y = 0.;
  if (J >= 1 && J <= nj) {
    double havg = .5 * (h3[ix2(d, J, I)] + h3[ix2(d, J - 1, I)]);
    y = v[ix2(d, J, I)] * havg * d.minv_v[J];
  }
  vh[ix2(d, J, I)] = y;
  if (I == ni) vh[ix2(d, J, 0)] = y;
  if (I == 1) vh[ix2(d, J, ni + 1)] = y;
}

// DVDT = (DVDT + PV*UH terms + dKIN/dy) - vertical advection
__global__ void __launch_bounds__(256) k_dvdt_core(const MemberDesc *D, int mb) {
  MEMBER_X;
  FLAT2_X(1, d.nj, 1, d.ni);
  const int K = 1 + bx_.y;
  const size_t o = (size_t)K * d.N2;
  const double *uh = d.sA + o, *vh = d.sB + o, *kin = d.sC + o, *pv = d.PV3 + o;
  double *dv = cur_tend(d, d.DVDT) + o;
  auto PV = [&](int j, int i) { return pv[ix2(d, j, i)]; };
  const PvW<double> w = pv_weights_v<double>(d.pv_al81, PV, J, I);
  double x = dv[ix2(d, J, I)];
  x += ((-w.ga * uh[ix2(d, J, I + 1)] - w.de * uh[ix2(d, J, I)] - w.al * uh[ix2(d, J - 1, I)] -
         w.be * uh[ix2(d, J - 1, I + 1)] + w.e2 * vh[ix2(d, J - 1, I)] - w.e1 * vh[ix2(d, J + 1, I)]) *
            w.coef +
        kin[ix2(d, J - 1, I)] - kin[ix2(d, J, I)]) *
       d.n[2 * J];
  int kk = 2 * K + 1;
  double ds = d.sgth[kk - 2] - d.sgth[kk + 2];
  x -= .5 * (d.W3[ix3(d, K, J, I)] + d.W3[ix3(d, K, J - 1, I)]) *
       (d.V[ix3(d, K - 1, J, I)] - d.V[ix3(d, K + 1, J, I)]) / ds;
  dv[ix2(d, J, I)] = x;
}

// DUDT = (DUDT + PV*VH terms + dKIN/dx with filtered KIN) - vertical
// advection.  One block per (K,J) u-row: the KIN row is zonally filtered in
// LDS (rows of the filter band) and its pad refreshed, then one lane per I.
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(8))) k_dudt_core(const MemberDesc *D, int mb) {
  MEMBER;
  extern __shared__ double lds[];
  const RowMap m = row_map(d.ni);
  const int Jr = blockIdx.x * m.rpb + m.g, K = 1 + blockIdx.y, ni = d.ni;
  const bool live = Jr <= d.nj;
  const int J = live ? Jr : 0;
  const size_t o = (size_t)K * d.N2;
  const double *uh = d.sA + o, *vh = d.sB + o, *kin = d.sC + o, *pv = d.PV3 + o;
  double *kr = lds + (size_t)m.g * 2 * ni, *work = kr + ni;  // kr[I-1] = KIN(J,I)
  if (live)
    for (int I = 1 + m.lane; I <= ni; I += m.stride) kr[I - 1] = kin[ix2(d, J, I)];
  __syncthreads();
  filter_row_lds(d, kr, work, J, 0, live && d.rflag_h[J] != 0, m.lane);
  __syncthreads();
  if (!live) return;
  double *du = cur_tend(d, d.DUDT) + o;
  for (int I = 1 + m.lane; I <= ni; I += m.stride) {
    auto PV = [&](int j, int i) { return pv[ix2(d, j, i)]; };
    const PvW<double> w = pv_weights_u<double>(d.pv_al81, PV, J, I);
    const double kin_w = kr[I == 1 ? ni - 1 : I - 2], kin_c = kr[I - 1];
    double x = du[ix2(d, J, I)];
    x += ((w.al * vh[ix2(d, J + 1, I)] + w.be * vh[ix2(d, J + 1, I - 1)] + w.ga * vh[ix2(d, J, I - 1)] +
           w.de * vh[ix2(d, J, I)] + w.e2 * uh[ix2(d, J, I - 1)] - w.e1 * uh[ix2(d, J, I + 1)]) *
              w.coef +
          kin_w - kin_c) *
         d.m[2 * J + 1];
    int kk = 2 * K + 1;
    double ds = d.sgth[kk - 2] - d.sgth[kk + 2];
    x -= .5 * (d.W3[ix3(d, K, J, I)] + d.W3[ix3(d, K, J, I - 1)]) *
         (d.U[ix3(d, K - 1, J, I)] - d.U[ix3(d, K + 1, J, I)]) / ds;
    du[ix2(d, J, I)] = x;
  }
}

void L_uv_core(hipStream_t s, const Bucket &b) {
  if (b.mixed) { LM_uv_core(s, b); return; }
  const int nK = b.nk - 1;
  LAUNCH(k_uvh_kin, grid2(b.nj + 2, b.ni, nK, b.nm), dim3(TPB), 0, s, b.D, b.mb);
  LAUNCH(k_dvdt_core, grid2(b.nj, b.ni, nK, b.nm), dim3(TPB), 0, s, b.D, b.mb);
  const int rpb = rows_per_block(b.ni);
  LAUNCH(k_dudt_core, dim3((b.nj + 1 + rpb - 1) / rpb, nK, b.nm), dim3(256), sizeof(double) * rpb * 2 * b.ni, s,
         b.D, b.mb);
}

// ------------------------------------------------------------- uv_drag
// t_rad (src/shared/epic_sensible_heating.c:217-314): the Held-Suarez
// cooling time, or the tabulated one at the interface pressure; the hot
// Jupiter's relaxation time is its constant tau_rad
__device__ __forceinline__ double t_rad(const MemberDesc &d, int K, int J, int I) {
  if (d.is_hj) return d.tau_rad;
  if (d.is_tab) return tab_cool_time(d, -crm_log(d.P3[ix3(d, K, J, I)]));
  const double ka = 1. / (40. * 60. * 60. * 24.), ks = 1. / (4. * 60. * 60. * 24.);
  double p = d.P3[ix3(d, K, J, I)], kt = ka, pbot = d.P3[ix3(d, d.nk, J, I)];
  double amp = (p / pbot - 0.7) / (1. - 0.7);
  if (amp > 0.) kt += (ks - ka) * d.coslat4_h[J] * amp;  // host-libm table, as the oracle
  return 1. / kt;
}

// top sponge, K = 0..k_sponge (K = 0 runs even when k_sponge = 0)
__global__ void __launch_bounds__(256) k_sponge(const MemberDesc *D, int mb) {
  MEMBER;
  FLAT2(0, d.nj + 1, 1, d.ni);
  const int K = blockIdx.y;
  if (K > d.k_sponge) return;
  double *du = cur_tend(d, d.DUDT), *dv = cur_tend(d, d.DVDT);
  size_t q = ix3(d, K, J, I);
  double ts = d.t_sponge[K];
  const double x = du[q] - (ts) * (d.U[q] - d.USPIN[q]);
  du[q] = x;
  // zonal pads: nothing after uv_drag writes the interior of plane K = 0,
  // whose pads the reference fills in the bc_lateral that ends the 3-D
  // filter of the tendency (the filter itself starts at K = 1); the pads
  // of the other sponge planes are written again by the filter
  if (I == d.ni) du[ix3(d, K, J, 0)] = x;
  if (I == 1) du[ix3(d, K, J, d.ni + 1)] = x;
  if (J >= 1) {
    const double y = dv[q] - (ts) * (d.V[q] - 0.);
    dv[q] = y;
    if (I == d.ni) dv[ix3(d, K, J, 0)] = y;
    if (I == 1) dv[ix3(d, K, J, d.ni + 1)] = y;
  }
}

// Rayleigh drag (tau_drag, or prandtl / t_rad): one lane per
// (K,J) row, sequential zonal means as in the reference
__global__ void __launch_bounds__(256) k_rayleigh(const MemberDesc *D, int mb) {
  MEMBER;
  int t = blockIdx.x * blockDim.x + threadIdx.x;
  int nrow = (d.nk - 1 - d.k_sponge) * (d.nj + 1);
  if (t >= nrow || nrow <= 0) return;
  if (!((d.newt_cool_on && d.prandtl > 0.) || (d.tau_drag > 0. && d.tau_drag < 1.e+20))) return;
  const int K = d.k_sponge + 1 + t / (d.nj + 1), J = t % (d.nj + 1);
  double *du = cur_tend(d, d.DUDT), *dv = cur_tend(d, d.DVDT);
  const int ni = d.ni;
  const double za = (double)d.drag_zonal_avg, zb = (double)(1 - d.drag_zonal_avg);
  double uavg = 0.;
  for (int I = 1; I <= ni; I++) uavg += d.U[ix3(d, K, J, I)];
  uavg /= (ni);
  for (int I = 1; I <= ni; I++) {
    size_t q = ix3(d, K, J, I);
    double uray = za * uavg + zb * d.U[q];
    double nu0 = d.newt_cool_on ? d.prandtl / (.5 * (t_rad(d, K, J, I) + t_rad(d, K, J, I - 1))) : 1. / d.tau_drag;
    du[q] -= nu0 * (uray - d.USPIN[q]);
  }
  if (d.drag_v && J >= 1) {
    double vavg = 0.;
    for (int I = 1; I <= ni; I++) vavg += d.V[ix3(d, K, J, I)];
    vavg /= ni;
    for (int I = 1; I <= ni; I++) {
      size_t q = ix3(d, K, J, I);
      double vray = za * vavg + zb * d.V[q];
      double nu0 =
          d.newt_cool_on ? d.prandtl / (.5 * (t_rad(d, K, J, I) + t_rad(d, K, J - 1, I))) : 1. / d.tau_drag;
      dv[q] -= nu0 * vray;
    }
  }
}

// Held-Suarez boundary-layer drag, one lane per u- and v-column
__global__ void __launch_bounds__(256) k_hs_pbl(const MemberDesc *D, int mb) {
  MEMBER;
  FLAT2(0, d.nj + 1, 1, d.ni);
  if (!d.is_hs) return;
  const double nu00 = (1. / (24. * 60. * 60.));
  const int nk = d.nk;
  double *du = cur_tend(d, d.DUDT), *dv = cur_tend(d, d.DVDT);
  const double *P = d.P3;
  double pb = 0.5 * (P[ix3(d, nk, J, I)] + P[ix3(d, nk, J, I - 1)]);
  for (int K = nk - 1; K >= 1; K--) {
    double p3 = .5 * (P[ix3(d, K, J, I)] + P[ix3(d, K, J, I - 1)]);
    double amp = (p3 / pb - 0.7) / (1. - 0.7);
    if (amp < 0.) break;
    du[ix3(d, K, J, I)] -= nu00 * amp * d.U[ix3(d, K, J, I)];
  }
  if (J >= 1) {
    pb = 0.5 * (P[ix3(d, nk, J, I)] + P[ix3(d, nk, J - 1, I)]);
    for (int K = nk - 1; K >= 1; K--) {
      double p3 = .5 * (P[ix3(d, K, J, I)] + P[ix3(d, K, J - 1, I)]);
      double amp = (p3 / pb - 0.7) / (1. - 0.7);
      if (amp < 0.) break;
      dv[ix3(d, K, J, I)] -= nu00 * amp * d.V[ix3(d, K, J, I)];
    }
  }
}

void L_uv_drag(hipStream_t s, const Bucket &b) {
  if (b.mixed) { LM_uv_drag(s, b); return; }
  LAUNCH(k_sponge, grid2(b.nj + 1, b.ni, b.nk, b.nm), dim3(TPB), 0, s, b.D, b.mb);
  int nrow = b.nk * (b.nj + 1);
  if (b.any_ray) LAUNCH(k_rayleigh, dim3((nrow + 63) / 64, 1, b.nm), dim3(64), 0, s, b.D, b.mb);
  if (b.any_hs) LAUNCH(k_hs_pbl, grid2(b.nj + 1, b.ni, 1, b.nm), dim3(TPB), 0, s, b.D, b.mb);
}

// ------------------------------------------------------------- pgrad
// GZ3 rows -1 and nj+1 take the zonal mean of the adjacent row, K = 0..nk:
// one wavefront per (K, pole), the row staged in LDS and summed in order
__global__ void __launch_bounds__(256) k_gz_polefill(const MemberDesc *D, int mb) {
  MEMBER;
  if (d.pgrad_trad) return;  // the traditional form reads no pole row
  extern __shared__ double row[];
  const int K = blockIdx.x >> 1, north = blockIdx.x & 1, ni = d.ni;
  const int Js = north ? d.nj : 0, Jd = north ? d.nj + 1 : -1;
  for (int i = threadIdx.x; i < ni; i += blockDim.x) row[i] = d.GZ3[ix3(d, K, Js, i + 1)];
  __syncthreads();
  if (threadIdx.x == 0) {
    double tmp = seq_acc(row, ni);
    tmp /= ni;
    row[ni] = tmp;
  }
  __syncthreads();
  const double v = row[ni];
  for (int I = threadIdx.x; I <= ni + 1; I += blockDim.x) d.GZ3[ix3(d, K, Jd, I)] = v;
}

// finite-volume pressure-gradient force.  Longitudinal: the east face of
// cell I-1 is this cell's west face.  Latitudinal: the north face of row
// J-1 is this row's south face.
__device__ __forceinline__ double face_lon(const MemberDesc &d, int K, int J, int I, double &zbse, double &zbne,
                                           double &ztse, double &ztne, double &zbe, double &zte, double &pbe,
                                           double &pte) {
  const double *G = d.GZ3, *P2 = d.P2;
#define Z(k, j, i) G[ix3(d, k, j, i)]
  zbse = .25 * (Z(K, J, I) + Z(K, J - 1, I) + Z(K + 1, J, I) + Z(K + 1, J - 1, I));
  zbne = .25 * (Z(K, J, I) + Z(K, J + 1, I) + Z(K + 1, J, I) + Z(K + 1, J + 1, I));
  ztse = .25 * (Z(K, J, I) + Z(K, J - 1, I) + Z(K - 1, J, I) + Z(K - 1, J - 1, I));
  ztne = .25 * (Z(K, J, I) + Z(K, J + 1, I) + Z(K - 1, J, I) + Z(K - 1, J + 1, I));
  zbe = .5 * (Z(K, J, I) + Z(K + 1, J, I));
  zte = .5 * (Z(K, J, I) + Z(K - 1, J, I));
#undef Z
  pbe = P2[ix3(d, K + 1, J, I)];
  pte = P2[ix3(d, K, J, I)];
  double a1 = ztne + ztse, b1 = ztne * ztse, a2 = zbne + zbse, b2 = zbne * zbse;
  double t1 = a1 * a1 - b1 - (a2 * a2 - b2), t2 = 3. * (a1 - a2);
  return (-t1 * (pbe - pte) + t2 * (zte * pbe - zbe * pte)) / (zte - zbe);
}

__device__ __forceinline__ double face_lat(const MemberDesc &d, int K, int J, int I, double rlnn, double &zbnw,
                                           double &zbne, double &ztnw, double &ztne, double &zbn, double &ztn,
                                           double &pbn, double &ptn) {
  const double *G = d.GZ3, *P2 = d.P2;
#define Z(k, j, i) G[ix3(d, k, j, i)]
  zbnw = .25 * (Z(K, J, I) + Z(K, J, I - 1) + Z(K + 1, J, I) + Z(K + 1, J, I - 1));
  zbne = .25 * (Z(K, J, I) + Z(K, J, I + 1) + Z(K + 1, J, I) + Z(K + 1, J, I + 1));
  ztnw = .25 * (Z(K, J, I) + Z(K, J, I - 1) + Z(K - 1, J, I) + Z(K - 1, J, I - 1));
  ztne = .25 * (Z(K, J, I) + Z(K, J, I + 1) + Z(K - 1, J, I) + Z(K - 1, J, I + 1));
  zbn = .5 * (Z(K, J, I) + Z(K + 1, J, I));
  ztn = .5 * (Z(K, J, I) + Z(K - 1, J, I));
#undef Z
  pbn = P2[ix3(d, K + 1, J, I)];
  ptn = P2[ix3(d, K, J, I)];
  double a1 = ztne + ztnw, b1 = ztne * ztnw, a2 = zbne + zbnw, b2 = zbne * zbnw;
  double t1 = a1 * a1 - b1 - (a2 * a2 - b2), t2 = 3. * (a1 - a2);
  return rlnn * (-t1 * (pbn - ptn) + t2 * (ztn * pbn - zbn * ptn)) / (ztn - zbn);
}

// uv_pgrad_traditional (src/shared/epic_timestep.c:806-884, the reference's
// PGRAD_TRADITIONAL build): on K = 1..k_sigma-2 grad(MONT3) less the mean
// EXNER3 times grad(THETA); on K = k_sigma-1..nk-1 the h3-weighted GZ3
// difference less the p-weighted mean of the GZ3 differences of the
// interfaces above and below.  u at (K,J,I), v at J >= 1.
__device__ __forceinline__ void pgrad_traditional(const MemberDesc &d, int K, int J, int I, double *du, double *dv) {
  const double *MO = d.MONT3, *EX = d.EXNER3, *TH = d.THETA, *G = d.GZ3, *H = d.H3, *P = d.P3;
  const size_t q = ix3(d, K, J, I), w = ix3(d, K, J, I - 1), s = ix3(d, K, J - 1, I);
  if (K <= d.ks - 2) {
    du[q] -= d.m[2 * J + 1] * (MO[q] - MO[w] - .5 * (EX[q] + EX[w]) * (TH[q] - TH[w]));
    if (J >= 1) dv[q] -= d.n[2 * J] * (MO[q] - MO[s] - .5 * (EX[q] + EX[s]) * (TH[q] - TH[s]));
    return;
  }
  const size_t qu = q + d.N2, qd = q - d.N2;
  {
    const size_t wu = w + d.N2, wd = w - d.N2;
    const double hw = (G[q] * H[q] - G[w] * H[w]) * 2. / (H[q] + H[w]);
    const double up = (G[qu] + G[wu]) * (P[qu] - P[wu]) - (G[qd] + G[wd]) * (P[qd] - P[wd]);
    const double dn = P[qu] + P[wu] - P[qd] - P[wd];
    du[q] -= d.m[2 * J + 1] * (hw - up / dn);
  }
  if (J >= 1) {
    const size_t su = s + d.N2, sd = s - d.N2;
    const double hs = (G[q] * H[q] - G[s] * H[s]) * 2. / (H[q] + H[s]);
    const double up = (G[qu] + G[su]) * (P[qu] - P[su]) - (G[qd] + G[sd]) * (P[qd] - P[sd]);
    const double dn = P[qu] + P[su] - P[qd] - P[sd];
    dv[q] -= d.n[2 * J] * (hs - up / dn);
  }
}

__global__ void __launch_bounds__(256) k_pgrad(const MemberDesc *D, int mb) {
  MEMBER_X;
  FLAT2_X(0, d.nj + 1, 1, d.ni);
  const int K = 1 + bx_.y;
  const double TOL = 1.e-2, DG = M_PI / 180.;
  const double *G = d.GZ3, *P2 = d.P2;
  double *du = cur_tend(d, d.DUDT), *dv = cur_tend(d, d.DVDT);
  if (d.pgrad_trad) {
    pgrad_traditional(d, K, J, I, du, dv);
    return;
  }
#define Z(k, j, i) G[ix3(d, k, j, i)]
#define PZ(k, j, i) P2[ix3(d, k, j, i)]
  {
    // ---- longitudinal component at u-point (J,I)
    double rlns = d.rln[2 * J], rlnn = d.rln[2 * J + 2];
    double zbsw, zbnw, ztsw, ztnw, zbw, ztw, pbw, ptw;
    double west = face_lon(d, K, J, I - 1, zbsw, zbnw, ztsw, ztnw, zbw, ztw, pbw, ptw);
    double lnw = d.lon[2 * (I - 1) + 1] * DG;
    double zbse, zbne, ztse, ztne, zbe, zte, pbe, pte;
    double east = face_lon(d, K, J, I, zbse, zbne, ztse, ztne, zbe, zte, pbe, pte);
    double lne = d.lon[2 * I + 1] * DG;
    double rho = .5 * (d.RHO3[ix3(d, K, J, I)] + d.RHO3[ix3(d, K, J, I - 1)]);
    double top, bottom, a1, a2, b1, b2, t1, t2;
    a1 = ztne + ztse; a2 = ztnw + ztsw;
    t2 = 3. * (a1 - a2);
    if (fabs(2. * (zte - ztw) / (zte + ztw - zbe - zbw)) > TOL) {
      b1 = ztne * ztse; b2 = ztnw * ztsw;
      t1 = a1 * a1 - b1 - (a2 * a2 - b2);
      top = (-t1 * (ptw - pte) + t2 * (zte * ptw - ztw * pte)) / (zte - ztw);
    } else {
      top = .5 * (ptw + pte) * t2;
    }
    a1 = zbne + zbse; a2 = zbnw + zbsw;
    t2 = 3. * (a1 - a2);
    if (fabs(2. * (zbe - zbw) / (zte + ztw - zbe - zbw)) > TOL) {
      b1 = zbne * zbse; b2 = zbnw * zbsw;
      t1 = a1 * a1 - b1 - (a2 * a2 - b2);
      bottom = (-t1 * (pbw - pbe) + t2 * (zbe * pbw - zbw * pbe)) / (zbe - zbw);
    } else {
      bottom = .5 * (pbw + pbe) * t2;
    }
    double north = 0., south = 0.;
    double pgf = top - bottom + south - north + west - east;
    double dzn = ztne - zbne + ztnw - zbnw, dzs = ztse - zbse + ztsw - zbsw;
    pgf /= (lne - lnw) * rho * (rlns * (dzs + .5 * dzn) + rlnn * (dzn + .5 * dzs));
    du[ix3(d, K, J, I)] += pgf;
  }
  if (J >= 1) {
    // ---- latitudinal component at v-point (J,I)
    double rlns = d.rln[2 * J - 1], lts = d.lat[2 * J - 1] * DG;
    double zbsw, zbse, ztsw, ztse, zbs, zts, pbs, pts;
    double south = face_lat(d, K, J - 1, I, rlns, zbsw, zbse, ztsw, ztse, zbs, zts, pbs, pts);
    double rlnn = d.rln[2 * J + 1], ltn = d.lat[2 * J + 1] * DG;
    double zbnw, zbne, ztnw, ztne, zbn, ztn, pbn, ptn;
    double north = face_lat(d, K, J, I, rlnn, zbnw, zbne, ztnw, ztne, zbn, ztn, pbn, ptn);
    double rho = .5 * (d.RHO3[ix3(d, K, J, I)] + d.RHO3[ix3(d, K, J - 1, I)]);
    double rlt = d.rlt[2 * J];
    double a1, a2, b1, b2, b3, t1, t2;
    double zbe = .125 * (Z(K, J, I) + Z(K, J, I + 1) + Z(K, J - 1, I) + Z(K, J - 1, I + 1) + Z(K + 1, J, I) +
                         Z(K + 1, J, I + 1) + Z(K + 1, J - 1, I) + Z(K + 1, J - 1, I + 1));
    double zte = .125 * (Z(K - 1, J, I) + Z(K - 1, J, I + 1) + Z(K - 1, J - 1, I) + Z(K - 1, J - 1, I + 1) +
                         Z(K, J, I) + Z(K, J, I + 1) + Z(K, J - 1, I) + Z(K, J - 1, I + 1));
    double pbe = .25 * (PZ(K + 1, J, I) + PZ(K + 1, J, I + 1) + PZ(K + 1, J - 1, I) + PZ(K + 1, J - 1, I + 1));
    double pte = .25 * (PZ(K, J, I) + PZ(K, J, I + 1) + PZ(K, J - 1, I) + PZ(K, J - 1, I + 1));
    a1 = ztne + ztse; b1 = ztne * ztse; a2 = zbne + zbse; b2 = zbne * zbse;
    t1 = a1 * a1 - b1 - (a2 * a2 - b2);
    t2 = 3. * (a1 - a2);
    double east = .5 * (rlns - rlnn) * (-t1 * (pbe - pte) + t2 * (zte * pbe - zbe * pte)) / (zte - zbe);
    double zbw = .125 * (Z(K, J, I) + Z(K, J, I - 1) + Z(K, J - 1, I) + Z(K, J - 1, I - 1) + Z(K + 1, J, I) +
                         Z(K + 1, J, I - 1) + Z(K + 1, J - 1, I) + Z(K + 1, J - 1, I - 1));
    double ztw = .125 * (Z(K - 1, J, I) + Z(K - 1, J, I - 1) + Z(K - 1, J - 1, I) + Z(K - 1, J - 1, I - 1) +
                         Z(K, J, I) + Z(K, J, I - 1) + Z(K, J - 1, I) + Z(K, J - 1, I - 1));
    double pbw = .25 * (PZ(K + 1, J, I) + PZ(K + 1, J, I - 1) + PZ(K + 1, J - 1, I) + PZ(K + 1, J - 1, I - 1));
    double ptw = .25 * (PZ(K, J, I) + PZ(K, J, I - 1) + PZ(K, J - 1, I) + PZ(K, J - 1, I - 1));
    a1 = ztnw + ztsw; b1 = ztnw * ztsw; a2 = zbnw + zbsw; b2 = zbnw * zbsw;
    t1 = a1 * a1 - b1 - (a2 * a2 - b2);
    t2 = 3. * (a1 - a2);
    double west = .5 * (rlns - rlnn) * (-t1 * (pbw - ptw) + t2 * (ztw * pbw - zbw * ptw)) / (ztw - zbw);
    double top, bottom;
    a1 = ztne + ztnw; a2 = ztse + ztsw;
    t2 = 3. * (a1 - a2);
    if (fabs(2. * (ztn - zts) / (ztn + zts - zbn - zbs)) > TOL) {
      b1 = ztne * ztnw; b2 = ztse * ztsw; b3 = ztne * ztse + ztnw * ztsw;
      t1 = a1 * a1 - b1 - (a2 * a2 - b2);
      top = .5 * ((ptn - pts) * ((rlnn + rlns) * t1 + (rlnn - rlns) * (b1 + b2 - b3)) +
                  (rlnn + rlns) * t2 * (ztn * pts - zts * ptn)) / (ztn - zts);
    } else {
      top = .25 * (rlnn + rlns) * (pts + ptn) * t2;
    }
    a1 = zbne + zbnw; a2 = zbse + zbsw;
    t2 = 3. * (a1 - a2);
    if (fabs(2. * (zbn - zbs) / (ztn + zts - zbn - zbs)) > TOL) {
      b1 = zbne * zbnw; b2 = zbse * zbsw; b3 = zbne * zbse + zbnw * zbsw;
      t1 = a1 * a1 - b1 - (a2 * a2 - b2);
      bottom = .5 * ((pbn - pbs) * ((rlnn + rlns) * t1 + (rlnn - rlns) * (b1 + b2 - b3)) +
                     (rlnn + rlns) * t2 * (zbn * pbs - zbs * pbn)) / (zbn - zbs);
    } else {
      bottom = .25 * (rlnn + rlns) * (pbs + pbn) * t2;
    }
    double pgf = top - bottom + south - (north + west + east);
    double dzn = ztne - zbne + ztnw - zbnw, dzs = ztse - zbse + ztsw - zbsw;
    pgf /= (ltn - lts) * rho * rlt * (rlns * (dzs + .5 * dzn) + rlnn * (dzn + .5 * dzs));
    dv[ix3(d, K, J, I)] += pgf;
  }
#undef Z
#undef PZ
}

void L_pgrad(hipStream_t s, const Bucket &b) {
  LAUNCH(k_gz_polefill, dim3(2 * (b.nk + 1), 1, b.nm), dim3(64), sizeof(double) * (b.ni + 1), s, b.D, b.mb);
  // FP64 in both precision modes (k_mixed.hip says why)
  LAUNCH(k_pgrad, grid2(b.nj + 1, b.ni, b.nk - 1, b.nm), dim3(TPB), 0, s, b.D, b.mb);
}

// ------------------------------------------------------ divergence damping
__global__ void __launch_bounds__(256) k_divdamp(const MemberDesc *D, int mb) {
  MEMBER_X;
  FLAT2_X(0, d.nj + 1, 1, d.ni);
  const int K = 1 + bx_.y;
  if (d.nudiv == 0.) return;
  const double *div = d.sB + (size_t)K * d.N2;
  double *du = cur_tend(d, d.DUDT), *dv = cur_tend(d, d.DVDT);
  double coef = (d.lat[2 * J + 1] > 45.) ? d.nudiv * d.m0z : d.nudiv * d.m[2 * J + 1];
  du[ix3(d, K, J, I)] += coef * (div[ix2(d, J, I)] - div[ix2(d, J, I - 1)]);
  if (J >= 1) {
    double cv = d.nudiv * d.n[2 * J];
    dv[ix3(d, K, J, I)] += cv * (div[ix2(d, J, I)] - div[ix2(d, J - 1, I)]);
  }
}
void L_divdamp(hipStream_t s, const Bucket &b) {
  L_divergence(s, b, F_U, F_V, F_SB, 1, b.nk - 1);  // U filtered row by row inside
  if (b.mixed) { LM_divdamp_apply(s, b); return; }
  LAUNCH(k_divdamp, grid2(b.nj + 1, b.ni, b.nk - 1, b.nm), dim3(TPB), 0, s, b.D, b.mb);
}

// ------------------------------------------------------ Adams-Bashforth
// Coefficients from the FLAT_MAX start-up flags of the older planes,
// tested per wind at (KLO, jlo, ILO) as the reference does.
__global__ void __launch_bounds__(256) k_ab(const MemberDesc *D, int mb) {
  MEMBER;
  FLAT2(0, d.nj + 1, 1, d.ni);
  const int K = 1 + blockIdx.y;
  const MemberDyn dy = *d.dyn;
  const double FMAX = 1.7976931348623157e308;
  for (int w = 0; w < 2; w++) {
    const int jlo = w;  // u: JLO = 0, v: JFIRST = 1
    if (J < jlo) continue;
    double *val = w ? d.V : d.U;
    double *tb = w ? d.DVDT : d.DUDT;
    double ab0, ab1, ab2;
    if (tb[(size_t)dy.itm2 * d.N3 + ix3(d, 1, jlo, 1)] == FMAX) {
      if (tb[(size_t)dy.itm1 * d.N3 + ix3(d, 1, jlo, 1)] == FMAX) { ab0 = 1.; ab1 = 0.; ab2 = 0.; }
      else { ab0 = 3. / 2.; ab1 = -1. / 2.; ab2 = 0.; }
    } else {
      ab0 = 23. / 12.; ab1 = -16. / 12.; ab2 = 5. / 12.;
    }
    size_t q = ix3(d, K, J, I);
    const double y = val[q] + d.dt * (ab0 * tb[(size_t)dy.it0 * d.N3 + q] + ab1 * tb[(size_t)dy.itm1 * d.N3 + q] +
                                      ab2 * tb[(size_t)dy.itm2 * d.N3 + q]);
    val[q] = y;
    // zonal pads (bc_lateral after the update)
    if (I == d.ni) val[ix3(d, K, J, 0)] = y;
    if (I == 1) val[ix3(d, K, J, d.ni + 1)] = y;
  }
}
void L_ab(hipStream_t s, const Bucket &b) {
  if (b.mixed) { LM_ab(s, b); return; }
  LAUNCH(k_ab, grid2(b.nj + 1, b.ni, b.nk - 1, b.nm), dim3(TPB), 0, s, b.D, b.mb);
}

// monotone filtered geopotential, K = nk-1 down to 0 (src/shared/epic_timestep.c:176-196)
__global__ void __launch_bounds__(256) k_gz_mono(const MemberDesc *D, int mb) {
  MEMBER;
  FLAT2(0, d.nj + 1, 0, d.ni + 2);
  double *gz = d.gz;
  double below = gz[ix3(d, d.nk, J, I)];
  // levels go eight at a time: their loads are issued together ahead of the
  // chain (each level is read before it is written; the order of the chain
  // is unchanged)
  int K = d.nk - 1;
  for (; K >= 7; K -= 8) {
    double g[8];
#pragma unroll
    for (int u = 0; u < 8; u++) g[u] = gz[ix3(d, K - u, J, I)];
#pragma unroll
    for (int u = 0; u < 8; u++) {
      double v;
      if (fcmp_d(below, 0.) == 0) v = 1.e-3;
      else v = dmax(g[u], below + fabs(below) * 1.e-5);
      gz[ix3(d, K - u, J, I)] = v;
      below = v;
    }
  }
  for (; K >= 0; K--) {
    size_t q = ix3(d, K, J, I);
    double v;
    if (fcmp_d(below, 0.) == 0) v = 1.e-3;
    else v = dmax(gz[q], below + fabs(below) * 1.e-5);
    gz[q] = v;
    below = v;
  }
}
// gz = GZ3 with planes 0..nk-1 zonally filtered (rows of the h-grid band),
// one block per (K,J) row of the whole plane (J = -1..nj+1), pads included
__global__ void __launch_bounds__(256) k_gz_filtercopy(const MemberDesc *D, int mb) {
  MEMBER;
  extern __shared__ double lds[];
  const RowMap m = row_map(d.ni);
  const int Jr = (int)(blockIdx.x * m.rpb + m.g) - 1, K = blockIdx.y, ni = d.ni, W = ni + 2;
  const bool live = Jr <= d.nj + 1;
  const int J = live ? Jr : 0;
  double *x = lds + (size_t)m.g * (W + ni), *work = x + W;
  const double *src = d.GZ3 + (size_t)K * d.N2;
  double *dst = d.gz + (size_t)K * d.N2;
  if (live)
    for (int I = m.lane; I < W; I += m.stride) x[I] = src[ix2(d, J, I)];
  __syncthreads();
  const bool filt = live && K <= d.nk - 1 && J >= 0 && J <= d.nj && d.rflag_h[J] != 0;
  filter_row_lds(d, x + 1, work, J < 0 ? 0 : J, 0, filt, m.lane);
  __syncthreads();
  if (!live) return;
  for (int I = m.lane; I < W; I += m.stride) {
    double v = x[I];
    if (filt && I == 0) v = x[ni];
    if (filt && I == ni + 1) v = x[1];
    dst[ix2(d, J, I)] = v;
  }
}
void L_gz_filtercopy(hipStream_t s, const Bucket &b) {
  const int rpb = rows_per_block(b.ni);
  LAUNCH(k_gz_filtercopy, dim3((b.nj + 3 + rpb - 1) / rpb, b.nk + 2, b.nm), dim3(256),
         sizeof(double) * rpb * (2 * b.ni + 2), s, b.D, b.mb);
}

void L_gz_mono(hipStream_t s, const Bucket &b) {
  LAUNCH(k_gz_mono, grid2(b.nj + 1, b.ni + 2, 1, b.nm), dim3(TPB), 0, s, b.D, b.mb);
}

// ------------------------------------------------------ hyperviscosity
struct HvRange { int k0, k1, jbot, jh; };
__device__ __forceinline__ HvRange hv_range(const MemberDesc &d, int kind) {
  HvRange r;
  r.k0 = (kind == 3) ? d.ks - 1 : 1;
  r.k1 = (kind <= 1) ? d.nk - 1 : d.nk;
  r.jbot = (kind == 1) ? 1 : 0;
  r.jh = (kind == 1) ? 0 : 1;
  return r;
}

// field i of a set and its member
#define SET_MEMBER                                              \
  const int fi = (int)blockIdx.z / nm;                          \
  const MemberDesc &d = D[mb + (int)blockIdx.z - fi * nm];      \
  const int f = fs.f[fi], kind = fs.kind[fi]

// The two d2dy2 passes of del^4 (coef2, then coef4) in one launch, by
// column: a thread owns column I of a tile of TJ rows of one plane, loads
// the field on rows J0-2..J0+TJ+1 (clipped to 0..jmax) into registers at
// once (all the loads in flight together), and forms each first-pass value
// it needs from them (three per output row, rather than staging them: the
// same expressions, so the same values).  Rows 0, 1, jmax-1 and jmax, whose
// stencils are shifted, read the field again (from cache).
template <class T, int TJ>
__global__ void __launch_bounds__(256) k_d2dy2_cols(const MemberDesc *D, int mb, int nm, FieldSet fs) {
  SET_MEMBER;
  const int vg = (kind == 1);
  const int jmax = vg ? d.nj + 1 : d.nj;
  HvRange r = hv_range(d, kind);
  const int K = r.k0 + blockIdx.y;
  if (K > r.k1) return;  // uniform per block
  const int nchunk = d.ni / (int)blockDim.x, tile = blockIdx.x / nchunk, chunk = blockIdx.x - tile * nchunk;
  const int J0 = tile * TJ;
  if (J0 > jmax) return;  // uniform per block (the grid covers the v rows)
  const int I = 1 + chunk * (int)blockDim.x + (int)threadIdx.x;
  const double *hh = fld_ptr(d, f) + (size_t)K * d.N2;
  T *out = reinterpret_cast<T *>(d.sB) + (size_t)(fs.copy0 + fi) * d.N3 + (size_t)K * d.N2;
  const double *cx = vg ? d.vcx : d.hcx, *cy = vg ? d.vcy : d.hcy, *cz = vg ? d.vcz : d.hcz;
  const T c2 = T(d.coef2), c4 = T(d.coef4);
  T w[TJ + 4];  // rows J0-2 .. J0+TJ+1
#pragma unroll
  for (int t = 0; t < TJ + 4; t++) {
    const int J = J0 - 2 + t;
    w[t] = (J >= 0 && J <= jmax) ? T(hh[ix2(d, J, I)]) : T(0.);
  }
  // first pass at row j from rows (j0, j1, j2)
  auto p1 = [&](int j, T h0, T h1, T h2) {
    T a0 = c2 * T(cx[j]), a1 = c2 * T(cy[j]), a2 = c2 * T(cz[j]);
    return a0 * h0 + a1 * h1 + a2 * h2;
  };
  auto fld = [&](int j) { return T(hh[ix2(d, j, I)]); };
  auto p1_at = [&](int j) {  // any row, stencil shifted at the ends
    const int c = min(max(j, 1), jmax - 1);
    return p1(j, fld(c - 1), fld(c), fld(c + 1));
  };
#pragma unroll
  for (int t = 0; t < TJ; t++) {
    const int J = J0 + t;
    if (J > jmax) break;
    T q0, q1, q2;
    if (J >= 2 && J <= jmax - 2) {
      q0 = p1(J - 1, w[t], w[t + 1], w[t + 2]);
      q1 = p1(J, w[t + 1], w[t + 2], w[t + 3]);
      q2 = p1(J + 1, w[t + 2], w[t + 3], w[t + 4]);
    } else {
      const int c = min(max(J, 1), jmax - 1);
      q0 = p1_at(c - 1);
      q1 = p1_at(c);
      q2 = p1_at(c + 1);
    }
    T a0 = c4 * T(cx[J]), a1 = c4 * T(cy[J]), a2 = c4 * T(cz[J]);
    out[ix2(d, J, I)] = a0 * q0 + a1 * q1 + a2 * q2;
  }
}

// Rows of the order-2(pass+1) result D (copy i of sB, or of sC when src is
// 1; 1024/ni rows per block), for a member that applies that order
// (nu > 0): D is zonally filtered in LDS (rows of the filter band), then
// A += DT*sign*D on the kind's rows (sign -1 for del^4 and del^8, +1 for
// del^6) and the zonal pads of every row 0..nj+1 are written; when a higher
// order follows, the filtered D goes back to its copy, which the next
// order's meridional second difference reads (the reference filters it in
// place, src/shared/epic_subgrid.c:960-975).  Members that do not apply
// the order leave D and A as they are.
template <class T>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(8))) k_hv_apply(const MemberDesc *D, int mb, int nm, FieldSet fs, int pass, int src) {
  SET_MEMBER;
  extern __shared__ double lds_raw[];
  T *lds = reinterpret_cast<T *>(lds_raw);
  const RowMap m = row_map(d.ni);
  HvRange r = hv_range(d, kind);
  const int K = r.k0 + blockIdx.y, Jr = blockIdx.x * m.rpb + m.g;
  if (K > r.k1 || hv_nu(d, pass) == 0. || hv_passes(d) <= pass) return;  // uniform per block
  const bool live = Jr <= d.nj + 1;
  const int J = live ? Jr : 0, ni = d.ni, vg = (kind == 1);
  T *x = lds + (size_t)m.g * 2 * ni, *work = x + ni;
  double *a = fld_ptr(d, f) + (size_t)K * d.N2;
  T *dd = reinterpret_cast<T *>(src ? d.sC : d.sB) + (size_t)(fs.copy0 + fi) * d.N3 + (size_t)K * d.N2;
  const bool upd = live && J >= r.jbot && J <= d.nj;
  if (upd)
    for (int I = 1 + m.lane; I <= ni; I += m.stride) x[I - 1] = dd[ix2(d, J, I)];
  __syncthreads();
  filter_row_lds(d, x, work, J, vg, upd && (vg ? d.rflag_v[J] : d.rflag_h[J]) != 0, m.lane);
  __syncthreads();
  if (!live) return;
  const bool back = upd && hv_passes(d) > pass + 1;
  const double sign = (pass & 1) ? -1. : 1.;
  for (int I = 1 + m.lane; I <= ni; I += m.stride) {
    double v = a[ix2(d, J, I)];
    if (upd) {
      v += d.dt * sign * (double)x[I - 1];
      a[ix2(d, J, I)] = v;
    }
    if (back) dd[ix2(d, J, I)] = x[I - 1];
    if (I == d.ni) a[ix2(d, J, 0)] = v;
    if (I == 1) a[ix2(d, J, d.ni + 1)] = v;
  }
}

// One meridional second difference of the hyperviscosity chain past del^4:
// order 2(pass+1) = coef * d2dy2(order 2 pass) from copy i of sB (src 0) or
// sC (src 1) into the other, for members that run that pass
template <class T>
__global__ void __launch_bounds__(256) k_d2dy2_pass(const MemberDesc *D, int mb, int nm, FieldSet fs, int pass, int src) {
  SET_MEMBER;
  const int vg = (kind == 1);
  const int jmax = vg ? d.nj + 1 : d.nj;
  HvRange r = hv_range(d, kind);
  const int K = r.k0 + blockIdx.y;
  if (K > r.k1 || hv_passes(d) <= pass) return;  // uniform per block
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (jmax + 1) * d.ni) return;
  const int J = t / d.ni, I = 1 + t - J * d.ni;
  const size_t o = (size_t)(fs.copy0 + fi) * d.N3 + (size_t)K * d.N2;
  const T *hh = reinterpret_cast<const T *>(src ? d.sC : d.sB) + o;
  T *out = reinterpret_cast<T *>(src ? d.sB : d.sC) + o;
  const double *cx = vg ? d.vcx : d.hcx, *cy = vg ? d.vcy : d.hcy, *cz = vg ? d.vcz : d.hcz;
  int j0, j1, j2;
  if (J == 0) { j0 = 0; j1 = 1; j2 = 2; }
  else if (J == jmax) { j0 = J - 2; j1 = J - 1; j2 = J; }
  else { j0 = J - 1; j1 = J; j2 = J + 1; }
  const T c = T(hv_coef(d, pass));
  const T a0 = c * T(cx[J]), a1 = c * T(cy[J]), a2 = c * T(cz[J]);
  out[ix2(d, J, I)] = a0 * hh[ix2(d, j0, I)] + a1 * hh[ix2(d, j1, I)] + a2 * hh[ix2(d, j2, I)];
}

// the two d2dy2 passes of del^4 into copy i of sB, by column tiles of up to
// 256 columns and TJ rows (the rows_d2 knob; built in: 32 rows where that
// still gives 2048 blocks, else 16 or 8)
template <class T>
static void d2dy2_del4(hipStream_t s, const Bucket &b, const FieldSet &fs) {
  bool any_v = false;
  for (int i = 0; i < fs.n; i++) any_v |= (fs.kind[i] == 1);
  // rows 0..nj (+ nj+1 on the v grid); the kernel drops rows past its kind's
  const int jrows = any_v ? b.nj + 2 : b.nj + 1;
  const int CW = b.ni < 256 ? b.ni : 256;
  const int per = (b.ni / CW) * b.nk * fs.n * b.nm;  // blocks per row tile
  const int TJ = b.knob[KN_ROWS_D2]                           ? b.knob[KN_ROWS_D2]
                 : ((jrows + 31) / 32 * per >= 2048) ? 32
                 : ((jrows + 15) / 16 * per >= 2048) ? 16
                                                     : 8;
  dim3 g((jrows + TJ - 1) / TJ * (b.ni / CW), b.nk, fs.n * b.nm);
  if (TJ == 32) LAUNCH_AS("k_d2dy2_cols<T>", (k_d2dy2_cols<T, 32>), g, dim3(CW), 0, s, b.D, b.mb, b.nm, fs);
  else if (TJ == 16) LAUNCH_AS("k_d2dy2_cols<T>", (k_d2dy2_cols<T, 16>), g, dim3(CW), 0, s, b.D, b.mb, b.nm, fs);
  else LAUNCH_AS("k_d2dy2_cols<T>", (k_d2dy2_cols<T, 8>), g, dim3(CW), 0, s, b.D, b.mb, b.nm, fs);
}

// the mixed mode forms del^4 in FP32 (the field and its update stay FP64)
template <class T>
static void merid_hv_t(hipStream_t s, const Bucket &b, const FieldSet &fs) {
  bool any_v = false;
  for (int i = 0; i < fs.n; i++) any_v |= (fs.kind[i] == 1);
  const int jrows = any_v ? b.nj + 2 : b.nj + 1;
  d2dy2_del4<T>(s, b, fs);
  const int rpb = rows_per_block(b.ni);
  const dim3 ga((b.nj + 2 + rpb - 1) / rpb, b.nk, fs.n * b.nm);
  LAUNCH(k_hv_apply<T>, ga, dim3(256), sizeof(T) * rpb * 2 * b.ni, s, b.D, b.mb, b.nm, fs, 1, 0);
  // del^6 (pass 2, into sC) and del^8 (pass 3, back into sB)
  const dim3 gp((jrows * b.ni + TPB - 1) / TPB, b.nk, fs.n * b.nm);
  for (int pass = 2; pass < b.hv_passes; pass++) {
    const int src = (pass == 2) ? 0 : 1;
    LAUNCH(k_d2dy2_pass<T>, gp, dim3(TPB), 0, s, b.D, b.mb, b.nm, fs, pass, src);
    LAUNCH(k_hv_apply<T>, ga, dim3(256), sizeof(T) * rpb * 2 * b.ni, s, b.D, b.mb, b.nm, fs, pass, 1 - src);
  }
}
void L_merid_hv(hipStream_t s, const Bucket &b, const FieldSet &fs) {
  if (b.mixed) merid_hv_t<float>(s, b, fs);
  else merid_hv_t<double>(s, b, fs);
}

// zonal hyperviscosity: a block holds RB = 4*RW consecutive (K,J) rows in
// LDS, wavefront w owning rows w + 4u (u < RW), each lane the points
// I = 1 + lane + 64q (q < NQ) of them, whose second differences it keeps in
// registers; the row array is rewritten with them for the zonal means, which
// are sequential sums, as the reference sums them, lane r of the first
// wavefront summing row r (one chain of additions serves RB rows).  One row
// array per row (not two) lets a CU hold twice the rows, i.e. run twice the
// serial chains side by side.  Rows are strided by an odd number of doubles
// so the lanes of the serial phase fall in distinct LDS banks.
template <class T, int RW, int NQ>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(4))) k_zonal_hv(const MemberDesc *D, int mb, int nm, FieldSet fs) {
  SET_MEMBER;
  constexpr int RB = 4 * RW;
  extern __shared__ double lds_raw[];
  T *X = reinterpret_cast<T *>(lds_raw);
  const int ni = d.ni, W = ni + 2, WS = W | 1;
  const int wv = threadIdx.x >> 6, lane = threadIdx.x & 63;
  HvRange r = hv_range(d, kind);
  const int K = r.k0 + blockIdx.y;
  const int np = hv_passes(d);
  if (K > r.k1 || np == 0) return;  // uniform over the block
  T *meanv = X + (size_t)RB * WS;
  double *a = fld_ptr(d, f) + (size_t)K * d.N2;
  const int J0 = blockIdx.x * RB, jlast = d.nj + 1;
  for (int u = 0; u < RW; u++) {
    const int t = wv + 4 * u, J = J0 + t;
    T *x = X + t * WS;
    for (int I = lane; I < W; I += 64) x[I] = (J <= jlast) ? T(a[ix2(d, J, I)]) : T(0.);
  }
  // m (or m0 poleward of LAT0) of each row; pass p's coefficient of the
  // row is (coef[2(p+1)] mm) mm, as the reference forms it
  double mm[RW];
#pragma unroll
  for (int u = 0; u < RW; u++) {
    const int Jc = min(J0 + wv + 4 * u, d.nj);  // rows past nj are never applied
    const bool taper = fabs(d.lat[2 * Jc + r.jh]) > 45.;
    mm[u] = taper ? d.m0z : d.m[2 * Jc + r.jh];
  }
  __syncthreads();
  T y[RW][NQ];
  for (int pass = 0; pass < np; pass++) {
    const double cp = hv_coef(d, pass);
#pragma unroll
    for (int u = 0; u < RW; u++) {
      const T *x = X + (wv + 4 * u) * WS;
      const T dcu = T(cp * mm[u] * mm[u]);
#pragma unroll
      for (int q = 0; q < NQ; q++) {
        const int I = 1 + lane + 64 * q;
        if (I <= ni) y[u][q] = dcu * (x[I - 1] - T(2.) * x[I] + x[I + 1]);
      }
    }
    __syncthreads();
#pragma unroll
    for (int u = 0; u < RW; u++) {
      T *x = X + (wv + 4 * u) * WS;
#pragma unroll
      for (int q = 0; q < NQ; q++) {
        const int I = 1 + lane + 64 * q;
        if (I <= ni) x[I] = y[u][q];
      }
    }
    __syncthreads();
    if (wv == 0 && lane < RB) meanv[lane] = seq_acc(X + lane * WS + 1, ni) / T(ni);
    __syncthreads();
    // every point (and the zonal pads, copies of points ni and 1) less the
    // mean; the last pass's result stays in registers
#pragma unroll
    for (int u = 0; u < RW; u++) {
      const int t = wv + 4 * u;
      T *x = X + t * WS;
      const T fm = meanv[t];
#pragma unroll
      for (int q = 0; q < NQ; q++) {
        const int I = 1 + lane + 64 * q;
        if (I <= ni) {
          y[u][q] -= fm;
          if (pass < np - 1) {
            x[I] = y[u][q];
            if (I == ni) x[0] = y[u][q];
            if (I == 1) x[ni + 1] = y[u][q];
          }
        }
      }
    }
    __syncthreads();
    if (hv_nu(d, pass) == 0.) continue;
    // an order with nu > 0: A += DT*sign*D on the rows of the range (sign
    // -1 for del^4 and del^8, +1 for del^6), then pads
    const double sign = (pass & 1) ? -1. : 1.;
#pragma unroll
    for (int u = 0; u < RW; u++) {
      const int J = J0 + wv + 4 * u;
      if (J > jlast) break;
      const bool upd = J >= r.jbot && J <= d.nj;
#pragma unroll
      for (int q = 0; q < NQ; q++) {
        const int I = 1 + lane + 64 * q;
        if (I > ni) break;
        double v = a[ix2(d, J, I)];
        if (upd) v += d.dt * sign * (double)y[u][q];
        a[ix2(d, J, I)] = v;
        if (I == ni) a[ix2(d, J, 0)] = v;
        if (I == 1) a[ix2(d, J, ni + 1)] = v;
      }
    }
  }
}

// rows per block RB (4, 8 or 16) such that a lane's registers hold at most
// 16 points (RB/4 rows x ni/64 points)
static int zhv_rows(const Bucket &b) {
  const int nq = b.ni >= 64 ? b.ni / 64 : 1;
  const int rows = (b.nj + 2) * b.nk * b.nm;
  int RB = b.knob[KN_ROWS_ZHV] ? b.knob[KN_ROWS_ZHV] : (rows / 16 >= 2048) ? 16 : (rows / 8 >= 2048) ? 8 : 4;
  while (RB > 4 && RB / 4 * nq > 16) RB /= 2;
  return RB;
}
template <class T, int RW>
static void zonal_hv_rw(hipStream_t s, const Bucket &b, const FieldSet &fs, dim3 g, size_t shm) {
  switch (b.ni >= 64 ? b.ni / 64 : 1) {
    case 1: LAUNCH_AS("k_zonal_hv<T>", (k_zonal_hv<T, RW, 1>), g, dim3(256), shm, s, b.D, b.mb, b.nm, fs); break;
    case 2: LAUNCH_AS("k_zonal_hv<T>", (k_zonal_hv<T, RW, 2>), g, dim3(256), shm, s, b.D, b.mb, b.nm, fs); break;
    case 4: LAUNCH_AS("k_zonal_hv<T>", (k_zonal_hv<T, RW, 4>), g, dim3(256), shm, s, b.D, b.mb, b.nm, fs); break;
    case 8:
      if constexpr (RW <= 2) LAUNCH_AS("k_zonal_hv<T>", (k_zonal_hv<T, RW, 8>), g, dim3(256), shm, s, b.D, b.mb, b.nm, fs);
      break;
    default:
      if constexpr (RW == 1) LAUNCH_AS("k_zonal_hv<T>", (k_zonal_hv<T, 1, 16>), g, dim3(256), shm, s, b.D, b.mb, b.nm, fs);
      break;
  }
}
template <class T>
static void zonal_hv_t(hipStream_t s, const Bucket &b, const FieldSet &fs) {
  const int RB = zhv_rows(b), WS = (b.ni + 2) | 1;
  dim3 g((b.nj + 2 + RB - 1) / RB, b.nk, fs.n * b.nm);
  const size_t shm = sizeof(T) * ((size_t)RB * WS + RB);
  if (RB == 4) zonal_hv_rw<T, 1>(s, b, fs, g, shm);
  else if (RB == 8) zonal_hv_rw<T, 2>(s, b, fs, g, shm);
  else zonal_hv_rw<T, 4>(s, b, fs, g, shm);
}
void L_zonal_hv(hipStream_t s, const Bucket &b, const FieldSet &fs) {
  if (b.mixed) zonal_hv_t<float>(s, b, fs);
  else zonal_hv_t<double>(s, b, fs);
}

// Zonal mean of a row of 256 points or more in the mixed mode (FP32): the
// row's ni values are summed as 64 partial sums, partial l = y[l] + y[l+64]
// + y[l+128] + ... left to right, which then meet pairwise, l with l+32,
// then l+16, ..., l+1 (a wavefront butterfly; IEEE addition commutes, so
// every lane ends with the same sum), and the sum is divided by ni.  Called
// by every lane of a wavefront, lane l holding partial l.
__device__ __forceinline__ float mean_butterfly(float p, int ni) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) p += __shfl_xor(p, o);
  return p / float(ni);
}

// del^4 after its meridional second differences, one row at a time (rows
// of members whose highest order is del^4): the row of D (copy i of sB) is
// zonally filtered and applied, A += DT*(-1)*D, then the zonal del^4 of the
// updated row (two second differences, each less its zonal mean) is
// applied, A += DT*(-1)*D_zonal; the zonal pads of every row 0..nj+1 hold
// copies of points ni and 1.  The steps and expressions are those of
// k_hv_apply (pass 1) followed by k_zonal_hv, with the row read and written
// once instead of twice.  A row's ni/4 lanes (RowMap) each own the points
// I = 1 + lane + (ni/4) q, q < 4; three row buffers in LDS (ni + 2 values,
// pads included) carry the row, the first second difference and the input
// of the second.  The zonal means: BFLY, the mixed mode's order for rows of
// 256 points or more (mean_butterfly; a row's lanes are whole wavefronts);
// otherwise sums left to right, as the reference and k_zonal_hv form them,
// each lane of a row forming the same sum itself (which spares a barrier
// to hand it round).
template <class T, bool BFLY>
__global__ void __launch_bounds__(256) k_hv4_rows(const MemberDesc *D, int mb, int nm, FieldSet fs) {
  SET_MEMBER;
  extern __shared__ double lds_raw[];
  const RowMap m = row_map(d.ni);
  // a row's lanes exchange data only among themselves: within one
  // wavefront when ni <= 256, so a wavefront barrier suffices there
  auto row_bar = [&]() {
    if (m.stride <= 64) wave_sync();
    else __syncthreads();
  };
  const int ni = d.ni, WS = ni + 2;
  T *z0 = reinterpret_cast<T *>(lds_raw) + (size_t)m.g * 3 * WS, *z1 = z0 + WS, *z2 = z1 + WS;
  HvRange r = hv_range(d, kind);
  const int K = r.k0 + blockIdx.y, Jr = blockIdx.x * m.rpb + m.g;
  if (K > r.k1 || hv_passes(d) != 2) return;  // uniform per block
  const bool live = Jr <= d.nj + 1;
  const int J = live ? Jr : 0, vg = (kind == 1);
  double *a = fld_ptr(d, f) + (size_t)K * d.N2;
  const T *dd = reinterpret_cast<const T *>(d.sB) + (size_t)(fs.copy0 + fi) * d.N3 + (size_t)K * d.N2;
  const bool upd = live && J >= r.jbot && J <= d.nj;
  // the row of A and the row of D, both loads in flight together
  double v[4];
  T dq[4];
#pragma unroll
  for (int q = 0; q < 4; q++) {
    const int I = 1 + m.lane + m.stride * q;
    v[q] = live ? a[ix2(d, J, I)] : 0.;
    dq[q] = upd ? dd[ix2(d, J, I)] : T(0.);
  }
  // the meridional part: D filtered in place (x = z0[1..ni], work = z1) in
  // blocks that hold a row of the filter band (a block-uniform test)
  const int *rflag = vg ? d.rflag_v : d.rflag_h;
  bool anyf = false;
  for (int g = 0, J1 = blockIdx.x * m.rpb; g < m.rpb; g++, J1++)
    anyf |= (J1 >= r.jbot && J1 <= d.nj && rflag[J1] != 0);
  T *x = z0 + 1;
  if (anyf) {
#pragma unroll
    for (int q = 0; q < 4; q++) x[m.lane + m.stride * q] = dq[q];
    row_bar();
    filter_row_lds(d, x, z1, J, vg, upd && rflag[J] != 0, m.lane);
    row_bar();
#pragma unroll
    for (int q = 0; q < 4; q++) dq[q] = x[m.lane + m.stride * q];
  }
  const double sign = -1.;  // del^4
#pragma unroll
  for (int q = 0; q < 4; q++)
    if (upd) v[q] += d.dt * sign * (double)dq[q];
  // the updated row, pads included, as the zonal part reads it (each lane
  // rewrites only the points it read)
#pragma unroll
  for (int q = 0; q < 4; q++) {
    const int I = 1 + m.lane + m.stride * q;
    const T t = T(v[q]);
    z0[I] = t;
    if (I == ni) z0[0] = t;
    if (I == 1) z0[ni + 1] = t;
  }
  const int Jc = min(J, d.nj);
  const double mm = fabs(d.lat[2 * Jc + r.jh]) > 45. ? d.m0z : d.m[2 * Jc + r.jh];
  T y[4];
  row_bar();
  for (int pass = 0; pass < 2; pass++) {
    // pass 0 reads z0 and leaves its raw result in z1; pass 1 reads z2 (pass
    // 0 less its mean) and leaves its raw result in z0
    const T *xin = pass ? z2 : z0;
    T *raw = pass ? z0 : z1;
    const T dcu = T(hv_coef(d, pass) * mm * mm);
#pragma unroll
    for (int q = 0; q < 4; q++) {
      const int I = 1 + m.lane + m.stride * q;
      y[q] = dcu * (xin[I - 1] - T(2.) * xin[I] + xin[I + 1]);
      raw[I] = y[q];
    }
    row_bar();
    T fm;
    if constexpr (BFLY) {
      // each wavefront of the row's lanes forms the 64 partials itself
      const int l = m.lane & 63;
      T p = raw[1 + l];
      for (int k = l + 64; k < ni; k += 64) p += raw[1 + k];
      fm = mean_butterfly(p, ni);
    } else {
      fm = seq_acc(raw + 1, ni) / T(ni);
    }
#pragma unroll
    for (int q = 0; q < 4; q++) y[q] -= fm;
    if (pass == 0) {
#pragma unroll
      for (int q = 0; q < 4; q++) {
        const int I = 1 + m.lane + m.stride * q;
        z2[I] = y[q];
        if (I == ni) z2[0] = y[q];
        if (I == 1) z2[ni + 1] = y[q];
      }
      row_bar();
    }
  }
  if (!live) return;
#pragma unroll
  for (int q = 0; q < 4; q++) {
    const int I = 1 + m.lane + m.stride * q;
    if (upd) {
      v[q] += d.dt * sign * (double)y[q];
      a[ix2(d, J, I)] = v[q];
    }
    if (I == ni) a[ix2(d, J, 0)] = v[q];
    if (I == 1) a[ix2(d, J, ni + 1)] = v[q];
  }
}

// del^4 (and higher orders) of a field set: the meridional part, then the
// zonal part.  In the mixed mode with del^4 the highest order of the bucket
// (the engine keeps such members apart from those with higher orders), the
// filtered meridional update and the zonal part run as one row kernel
// (k_hv4_rows); the FP64 path keeps the separate kernels (faster there: its
// zonal means are the reference's serial sums, which k_zonal_hv spreads
// over more rows at once).
static void hv4_mixed(hipStream_t s, const Bucket &b, const FieldSet &fs) {
  using T = float;
  d2dy2_del4<T>(s, b, fs);
  const int rpb = rows_per_block(b.ni);
  const dim3 ga((b.nj + 2 + rpb - 1) / rpb, b.nk, fs.n * b.nm);
  const size_t shm = sizeof(T) * rpb * 3 * (b.ni + 2);
  if (b.ni >= 256) LAUNCH_AS("k_hv4_rows", (k_hv4_rows<T, true>), ga, dim3(256), shm, s, b.D, b.mb, b.nm, fs);
  else LAUNCH_AS("k_hv4_rows", (k_hv4_rows<T, false>), ga, dim3(256), shm, s, b.D, b.mb, b.nm, fs);
}
void L_hv(hipStream_t s, const Bucket &b, const FieldSet &fs) {
  if (b.mixed && b.hv_passes == 2 && b.ni >= 16 && b.ni <= 1024) {
    hv4_mixed(s, b, fs);
    return;
  }
  L_merid_hv(s, b, fs);
  L_zonal_hv(s, b, fs);
}

}  // namespace epic
