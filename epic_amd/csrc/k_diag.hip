// k_diag.hip -- diagnostic kernels of the EPIC step on gfx950 (FP64).
//
//   set_p2etc         src/shared/epic_funcs_diag.c:3352-3569  (column scan)
//   store_pgrad_vars  src/shared/epic_funcs_diag.c:3595-3950  (column scans)
//   gz_from_u         src/shared/epic_funcs_diag.c:4460-4610  (one block per member)
//   calc_heating      src/shared/epic_funcs_diag.c:4793-4869 +
//                     src/shared/epic_sensible_heating.c:51-463 (Held-Suarez)
//   calc_w            src/shared/epic_flux.c:123-335           (column scan)
//   vorticity         src/shared/epic_funcs_diag.c:4131-4445
//   divergence        src/shared/epic_funcs_diag.c:4064-4117
//
// Column kernels run one lane per (J,I) column with the K recurrence in the
// lane, so every load and store of a K-level is coalesced along I.  The
// arithmetic of each expression follows the reference's operation order.
#include "kernels.h"

namespace epic {


#define MEMBER const MemberDesc &d = D[mb + blockIdx.z]
// flat (J,I) over nJ rows starting at j0 and nI columns starting at i0
#define FLAT2(j0, nJ, i0, nI)                          \
  int t_ = blockIdx.x * TPB + threadIdx.x;             \
  if (t_ >= (nJ) * (nI)) return;                       \
  const int J = (j0) + t_ / (nI);                      \
  const int I = (i0) + t_ % (nI);

static inline dim3 grid2(int nJ, int nI, int nK, int nm) {
  return dim3((unsigned)((nJ * nI + TPB - 1) / TPB), (unsigned)nK, (unsigned)nm);
}

// ------------------------------------------------------------ utilities
__global__ void k_fill(const MemberDesc *D, int mb, int f, double v) {
  MEMBER;
  double *a = fld_ptr(d, f);
  size_t i = (size_t)blockIdx.x * TPB + threadIdx.x;
  size_t stride = (size_t)gridDim.x * TPB;
  for (; i < (size_t)d.N3; i += stride) a[i] = v;
}
void L_fill(hipStream_t s, const Bucket &b, int f, double v) {
  LAUNCH(k_fill, dim3(256, 1, b.nm), dim3(TPB), 0, s, b.D, b.mb, f, v);
}

// DZDT3 is zeroed whole while grid.itime <= 1 (src/shared/epic_funcs_diag.c:3674-3690)
__global__ void k_zero_dzdt(const MemberDesc *D, int mb) {
  MEMBER;
  if (d.dyn->itime > 1) return;
  size_t i = (size_t)blockIdx.x * TPB + threadIdx.x;
  size_t stride = (size_t)gridDim.x * TPB;
  for (; i < (size_t)d.N3; i += stride) d.DZDT3[i] = 0.;
}
void L_zero_dzdt(hipStream_t s, const Bucket &b) {
  LAUNCH(k_zero_dzdt, dim3(256, 1, b.nm), dim3(TPB), 0, s, b.D, b.mb);
}

// zonal periodicity of a whole 3-D array (bc_lateral THREEDIM,
// src/single/s_epic_functions.c:72-88): K = 0..nk+1, J = 0..nj+1
__global__ void k_bc3(const MemberDesc *D, int mb, int f) {
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
__global__ void k_copy_planes(const MemberDesc *D, int mb, int src, int dst, int k0) {
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

// ------------------------------------------------------------ set_p2etc
__device__ __forceinline__ double h_layer(const MemberDesc &d, const double *P3, int K, int J, int I) {
  return ((double)P3[ix3(d, K, J, I)] - (double)P3[ix3(d, K - 1, J, I)]) /
         ((double)d.g[2 * J + 1] * (double)d.dsgth[2 * K]);
}

__global__ void k_set_p2etc(const MemberDesc *D, int mb, int update_theta) {
  MEMBER;
  FLAT2(0, d.nj + 1, 0, d.ni + 2);
  const int nk = d.nk, ks = d.ks;
  double *P3 = d.P3, *H = d.HDRY;
  const double g = d.g[2 * J + 1];
  for (int K = 1; K <= nk; K++)
    P3[ix3(d, K, J, I)] = P3[ix3(d, K - 1, J, I)] + d.dsgth[2 * K] * g * H[ix3(d, K, J, I)];
  const double ptop = P3[ix3(d, 0, J, I)], pbot = P3[ix3(d, nk, J, I)];
  for (int K = ks - 1; K < nk; K++) {
    double sigma = (d.sgth[2 * K + 1] - d.zeta0) / (d.zeta1 - d.zeta0);
    P3[ix3(d, K, J, I)] = get_p_sigma(pbot, sigma, ptop);
  }
  for (int K = 1; K <= nk; K++) H[ix3(d, K, J, I)] = dmax(d.h_min[K], h_layer(d, P3, K, J, I));
  if (update_theta)
    for (int K = 0; K < ks - 1; K++) {
      double sigma = get_sigma(pbot, P3[ix3(d, K, J, I)], ptop);
      d.THETA[ix3(d, K, J, I)] = ((double)d.sgth[2 * K + 1] - f_sig(d, sigma)) / g_sig(d, sigma);
    }
  const double kp1 = d.kappa + 1.;
  for (int K = 1; K <= nk; K++) {
    double top = P3[ix3(d, K - 1, J, I)], bot = P3[ix3(d, K, J, I)], p2;
    if (fcmp_d(bot, top) == 0) p2 = .5 * (bot + top);
    else if (top < 0. || bot < 0.) { atomicOr(&d.dyn->err, 2); p2 = 0.; }
    else p2 = pow((pow(bot, kp1) - pow(top, kp1)) / (kp1 * (bot - top)), 1. / d.kappa);
    d.P2[ix3(d, K, J, I)] = p2;
  }
  d.P2[ix3(d, 0, J, I)] = P3[ix3(d, 0, J, I)] * P3[ix3(d, 0, J, I)] / d.P2[ix3(d, 1, J, I)];
  for (int K = 1; K <= nk; K++) {
    double h;
    if (K < nk) {
      int kk = 2 * K + 1;
      double topwt = d.dsgth[kk - 1], botwt = d.dsgth[kk + 1];
      h = (topwt * h_layer(d, P3, K, J, I) + botwt * h_layer(d, P3, K + 1, J, I)) / (topwt + botwt);
    } else {
      h = h_layer(d, P3, K, J, I);
    }
    d.H3[ix3(d, K, J, I)] = h;
  }
  d.H3[ix3(d, 0, J, I)] = d.H3[ix3(d, 1, J, I)];
  if (update_theta) {
    for (int K = 1; K < ks - 1; K++) {
      double sigma = get_sigma(pbot, d.P2[ix3(d, K, J, I)], ptop);
      d.THETA2[ix3(d, K, J, I)] = ((double)d.sgth[2 * K] - f_sig(d, sigma)) / g_sig(d, sigma);
    }
    d.THETA2[ix3(d, ks - 1, J, I)] = .5 * (d.THETA[ix3(d, ks - 1, J, I)] + d.THETA[ix3(d, ks - 2, J, I)]);
    // linear interpolation of theta against sigmatheta, table ordered from
    // the bottom interface up (src/shared/epic_funcs_diag.c:3522-3562)
    const int nkay = nk - (ks - 1) + 1;
    const double *tx = d.sgth + 2 * nk + 1;  // tx[kay*(-2)] = sgth[2*(nk-kay)+1]
    for (int K = ks; K <= nk; K++) {
      double sg = d.sgth[2 * K], h;
      int kay = table_place_s(nkay, tx, -2, sg, &h);
      double y0 = d.THETA[ix3(d, nk - kay, J, I)], y1 = d.THETA[ix3(d, nk - kay - 1, J, I)];
      d.THETA2[ix3(d, K, J, I)] = y0 + (y1 - y0) * (sg - tx[-2 * kay]) / h;
    }
  }
}
void L_set_p2etc(hipStream_t s, const Bucket &b, int update_theta) {
  LAUNCH(k_set_p2etc, grid2(b.nj + 1, b.ni + 2, 1, b.nm), dim3(TPB), 0, s, b.D, b.mb, update_theta);
}

// ------------------------------------------------- store_pgrad_vars parts
// T3, EXNER3, RHO3, T2, RHO2 and the top values; DZDT3 <- GZ3 when itime > 1
__global__ void k_thermo(const MemberDesc *D, int mb) {
  MEMBER;
  FLAT2(0, d.nj + 1, 0, d.ni + 2);
  const int nk = d.nk;
  if (d.dyn->itime > 1 && I >= 1 && I <= d.ni)
    for (int K = 1; K <= nk; K++) d.DZDT3[ix3(d, K, J, I)] = d.GZ3[ix3(d, K, J, I)];
  const double R_GAS = 8.314472e+3;
  for (int K = 0; K <= nk; K++) {
    size_t q = ix3(d, K, J, I);
    double th = d.THETA[q], p = d.P3[q];
    double t = th * pow(p / d.press0, d.kappa);
    d.T3[q] = t;
    d.EXNER3[q] = d.cp * t / th;
    d.RHO3[q] = p * d.mu3 / (R_GAS * t);
  }
  for (int K = 1; K <= nk; K++) {
    size_t q = ix3(d, K, J, I);
    double p = d.P2[q], th = d.THETA2[q];
    double t = th * pow(p / d.press0, d.kappa);
    d.T2[q] = t;
    d.RHO2[q] = p * d.mu2 / (R_GAS * t);
  }
  size_t q0 = ix3(d, 0, J, I);
  d.RHO2[q0] = d.RHO3[q0] * d.RHO3[q0] / d.RHO2[ix3(d, 1, J, I)];
  d.THETA2[q0] = d.sgth[0];
  d.T2[q0] = d.T3[q0];
}
void L_thermo(hipStream_t s, const Bucket &b) {
  LAUNCH(k_thermo, grid2(b.nj + 1, b.ni + 2, 1, b.nm), dim3(TPB), 0, s, b.D, b.mb);
}

// kinetic energy of a zonally uniform u row (get_kin with v = 0)
__device__ double kin_uniform(const MemberDesc &d, const double *u1, int J) {
  const double *mn = d.mn, *mni = d.mn_inv;
  int jj = 2 * J;
  double u2 = u1[J], u4 = u1[J], v2 = 0., v4 = 0.;
  double kc = (u2 * u2 + u4 * u4 + mn[jj + 1] * (v2 * v2 * mni[jj] + v4 * v4 * mni[jj + 2])) * .25;
  if (J == 0 || J == d.nj) return kc;
  u2 = .5 * (u1[J + 1] + u1[J - 1]);
  u4 = .5 * (u1[J + 1] + u1[J - 1]);
  v2 = .5 * (0. + 0.);
  v4 = .5 * (0. + 0.);
  double ks = (u2 * u2 + u4 * u4 + mn[jj + 1] * (v2 * v2 * mni[jj] + v4 * v4 * mni[jj + 2])) * .25;
  return (5. / 6.) * kc + (1. / 6.) * ks;
}

// Gas giants: GZ_SURFACE from gradient balance with the zonal-mean bottom
// wind.  Zonal means are sequential sums, as in the reference.
__global__ void k_gz_from_u(const MemberDesc *D, int mb) {
  MEMBER;
  if (!d.is_gas) return;
  extern __shared__ double sh[];
  const int nj = d.nj, ni = d.ni, nk = d.nk;
  double *u1 = sh, *p1 = sh + (nj + 2), *r1 = sh + 2 * (nj + 2), *pvr = sh + 3 * (nj + 2), *gz1 = sh + 4 * (nj + 3);
  for (int J = threadIdx.x; J <= nj; J += blockDim.x) {
    double su = 0., sp = 0., sr = 0.;
    for (int I = 1; I <= ni; I++) {
      su += d.U[ix3(d, nk, J, I)];
      sp += d.P3[ix3(d, nk, J, I)];
      sr += d.RHO3[ix3(d, nk, J, I)];
    }
    u1[J] = su / ni;
    p1[J] = sp / ni;
    r1[J] = sr / ni;
  }
  __syncthreads();
  // absolute vorticity of the uniform row field, stored in PV3(nk+1)
  double *pv = d.PV3 + (size_t)(nk + 1) * d.N2;
  for (int J = 1 + threadIdx.x; J <= nj; J += blockDim.x) {
    double f2j = d.f[2 * J], m2j = d.m[2 * J], n2j = d.n[2 * J];
    double mm1i = 1. / d.m[2 * J - 1], mp1i = 1. / d.m[2 * J + 1];
    double ze = m2j * ((0. - 0.) + n2j * (u1[J - 1] * mm1i - u1[J] * mp1i));
    pvr[J] = (ze + f2j);
  }
  if (threadIdx.x == 0) {
    double ze = 0.;
    for (int I = 1; I <= ni; I++) ze -= u1[0];
    ze *= d.mn[0] / ((d.m[1]) * (double)(ni));
    pvr[0] = ze + d.f[0];
    ze = 0.;
    for (int I = 1; I <= ni; I++) ze += u1[nj];
    ze *= d.mn[2 * (nj + 1)] / (d.m[2 * nj + 1] * (double)(ni));
    pvr[nj + 1] = ze + d.f[2 * (nj + 1)];
  }
  __syncthreads();
  for (int t = threadIdx.x; t < (nj + 2) * (ni + 2); t += blockDim.x) {
    int J = t / (ni + 2), I = t % (ni + 2);
    pv[ix2(d, J, I)] = pvr[J];
  }
  if (threadIdx.x == 0) {
    const double PVC = 1. / 12.;
    const int j0 = d.jtp;
    double pvh_prev = 0.;
    // pvhudy[J], then the meridional integral from j0 both ways
    double *pvh = gz1;  // reuse as pvhudy[0..nj+1]
    for (int J = 0; J <= nj + 1; J++) pvh[J] = 0.;
    for (int J = 1; J <= nj; J++) {
      double udyJ = u1[J] / d.n[2 * J + 1], udyM = u1[J - 1] / d.n[2 * J - 1];
      double ga = (pvr[J] + pvr[J] + pvr[J + 1]);
      double de = (pvr[J] + pvr[J + 1] + pvr[J]);
      double al = (pvr[J - 1] + pvr[J] + pvr[J]);
      double be = (pvr[J] + pvr[J] + pvr[J - 1]);
      pvh[J] = (ga * udyJ + de * udyJ + al * udyM + be * udyM) * PVC;
      pvh[J] += (p1[J] - p1[J - 1]) * 2. / (r1[J] + r1[J - 1]);
    }
    (void)pvh_prev;
    double *bern = sh + 5 * (nj + 3);
    double kin0 = 0.;
    if (j0 >= 0 && j0 <= nj) kin0 = kin_uniform(d, u1, j0);
    bern[j0] = 0. + kin0;
    for (int J = j0 + 1; J <= nj; J++) bern[J] = bern[J - 1] - pvh[J];
    for (int J = j0 - 1; J >= 0; J--) bern[J] = bern[J + 1] + pvh[J + 1];
    for (int J = 0; J <= nj; J++) bern[J] = bern[J] - kin_uniform(d, u1, J);
  }
  __syncthreads();
  double *gzr = sh + 5 * (nj + 3);
  for (int t = threadIdx.x; t < (nj + 1) * (ni + 2); t += blockDim.x) {
    int J = t / (ni + 2), I = t % (ni + 2);
    d.GZS[ix2(d, J, I)] = gzr[J];
  }
}
void L_gz_from_u(hipStream_t s, const Bucket &b) {
  size_t shm = sizeof(double) * 6 * (b.nj + 3);
  LAUNCH(k_gz_from_u, dim3(1, 1, b.nm), dim3(256), shm, s, b.D, b.mb);
}

// hydrostatic GZ3 and MONT3 (Leslie-Purser 4-point integral in ln p) and the
// monotonicity screen
__global__ void k_gz3(const MemberDesc *D, int mb) {
  MEMBER;
  FLAT2(0, d.nj + 1, 0, d.ni + 2);
  const int nk = d.nk;
  const double cpr = d.cpr, rgas = d.rgas;
  double *G = d.GZ3, *M = d.MONT3;
  const double *P = d.P3, *RH = d.RHO3, *T = d.T3;
#define PP(k) P[ix3(d, k, J, I)]
#define RR(k) RH[ix3(d, k, J, I)]
  G[ix3(d, nk, J, I)] = d.topo_scale * d.GZS[ix2(d, J, I)];
  M[ix3(d, nk, J, I)] = G[ix3(d, nk, J, I)] + cpr * T[ix3(d, nk, J, I)] * rgas;
  int K = nk;
  double gzk = G[ix3(d, K, J, I)];
  double gzm = gzk + ((PP(K) + PP(K - 1)) / (RR(K) + RR(K - 1))) * log(PP(K) / PP(K - 1));
  G[ix3(d, K - 1, J, I)] = gzm;
  M[ix3(d, K - 1, J, I)] = gzm + cpr * T[ix3(d, K - 1, J, I)] * rgas;
  int bad = 0;
  if (I >= 1 && I <= d.ni && gzk >= gzm) bad = 1;
  for (K = nk - 1; K >= 2; K--) {
    double dz10 = log(PP(K + 1) / PP(K)), dz21 = log(PP(K) / PP(K - 1)), dz32 = log(PP(K - 1) / PP(K - 2));
    double dz20 = dz21 + dz10, dz30 = dz32 + dz20, dz31 = dz32 + dz21;
    double df0 = PP(K + 1) / RR(K + 1), df1 = PP(K) / RR(K), df2 = PP(K - 1) / RR(K - 1), df3 = PP(K - 2) / RR(K - 2);
    double dgz = -(df0 * (dz21 * dz21 * dz21 * (dz31 + dz32)) / (dz10 * dz20 * dz30) +
                   df1 * (dz21 * (dz21 * (dz21 + 2. * (dz10 - dz31)) - 6. * dz10 * dz31)) / (dz10 * dz31) +
                   df2 * (dz21 * (dz21 * (dz21 + 2. * (dz32 - dz20)) - 6. * dz20 * dz32)) / (dz20 * dz32) +
                   df3 * (dz21 * dz21 * dz21 * (dz10 + dz20)) / (dz30 * dz31 * dz32)) / 12.;
    double gn = gzm + dgz;
    G[ix3(d, K - 1, J, I)] = gn;
    M[ix3(d, K - 1, J, I)] = gn + cpr * T[ix3(d, K - 1, J, I)] * rgas;
    if (I >= 1 && I <= d.ni && gzm >= gn) bad = 1;
    gzm = gn;
  }
  K = 1;
  double g0 = gzm + ((PP(K) + PP(K - 1)) / (RR(K) + RR(K - 1))) * log(PP(K) / PP(K - 1));
  G[ix3(d, 0, J, I)] = g0;
  M[ix3(d, 0, J, I)] = g0 + cpr * T[ix3(d, 0, J, I)] * rgas;
  if (I >= 1 && I <= d.ni && gzm >= g0) bad = 1;
  if (bad) atomicOr(&d.dyn->err, 1);
#undef PP
#undef RR
}
void L_gz3(hipStream_t s, const Bucket &b) {
  LAUNCH(k_gz3, grid2(b.nj + 1, b.ni + 2, 1, b.nm), dim3(TPB), 0, s, b.D, b.mb);
}

// finish DZDT3 (diagnostic vertical velocity)
__global__ void k_dzdt3(const MemberDesc *D, int mb) {
  MEMBER;
  FLAT2(0, d.nj + 1, 1, d.ni);
  const int K = 1 + blockIdx.y;
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
}
void L_dzdt3(hipStream_t s, const Bucket &b) {
  LAUNCH(k_dzdt3, grid2(b.nj + 1, b.ni, b.nk, b.nm), dim3(TPB), 0, s, b.D, b.mb);
}

// ------------------------------------------------------------ heating
__device__ double hs_newton(const MemberDesc &d, int K, int J, int I) {
  const double ka = 1. / (40. * 60. * 60. * 24.), ks = 1. / (4. * 60. * 60. * 24.);
  const double dt_y = 60., dth_z = 10., p0 = 1000. * 100.;
  double lat = d.lat[2 * J + 1];
  double p = d.P3[ix3(d, K, J, I)], t = d.T3[ix3(d, K, J, I)];
  double kt = ka, pbot = d.P3[ix3(d, d.nk, J, I)];
  double amp = (p / pbot - 0.7) / (1. - 0.7);
  if (amp > 0.) {
    double cl = cos(d.lat[2 * J + 1] * (M_PI / 180.));
    kt += (ks - ka) * pow(cl, 4.) * amp;
  }
  double t_cool = 1. / kt;
  double cpt = 1004. / t_cool;
  double c2 = cos(lat * (M_PI / 180.));
  c2 *= c2;
  double s2 = 1. - c2, pp0 = p / p0;
  double teq = (315. - dt_y * s2 - dth_z * log(pp0) * c2) * pow(pp0, d.kappa);
  teq = dmax(200., teq);
  return -cpt * (t - teq);
}

// HEAT3 = 0 everywhere, then the Held-Suarez relaxation on K = 1..nk with
// its zonal pads (the bc_lateral at the end of calc_heating)
__global__ void k_heating(const MemberDesc *D, int mb) {
  MEMBER;
  FLAT2(0, d.nj + 1, 1, d.ni);
  const int K = 1 + blockIdx.y;
  if (!(d.newt_cool_on && d.is_hs)) return;
  double h = 0. + hs_newton(d, K, J, I);
  d.HEAT3[ix3(d, K, J, I)] = h;
  if (I == d.ni) d.HEAT3[ix3(d, K, J, 0)] = h;
  if (I == 1) d.HEAT3[ix3(d, K, J, d.ni + 1)] = h;
}
// layer-mean removal (K <= k_sponge or newt_cool_adjust), sequential sums
__global__ void k_heating_adjust(const MemberDesc *D, int mb) {
  MEMBER;
  const int K = 1 + blockIdx.x * blockDim.x + threadIdx.x;
  if (K > d.nk) return;
  if (!(d.newt_cool_on && d.is_hs)) return;
  if (!(K <= d.k_sponge || d.newt_cool_adjust)) return;
  double area = 0., ha = 0.;
  for (int J = 0; J <= d.nj; J++) {
    double da = 1. / d.mn[2 * J + 1];
    for (int I = 1; I <= d.ni; I++) {
      area += da;
      ha += hs_newton(d, K, J, I) * da;
    }
  }
  double havg = ha / area;
  for (int J = 0; J <= d.nj; J++)
    for (int I = 0; I <= d.ni + 1; I++) {
      int Ii = (I == 0) ? d.ni : (I == d.ni + 1 ? 1 : I);
      d.HEAT3[ix3(d, K, J, I)] = 0. + (hs_newton(d, K, J, Ii) - havg);
    }
}
void L_heating(hipStream_t s, const Bucket &b, int adjust) {
  L_fill(s, b, F_HEAT3, 0.);
  LAUNCH(k_heating, grid2(b.nj + 1, b.ni, b.nk, b.nm), dim3(TPB), 0, s, b.D, b.mb);
  if (adjust) LAUNCH(k_heating_adjust, dim3(1, 1, b.nm), dim3(64), 0, s, b.D, b.mb);
}

// ------------------------------------------------------------ calc_w
// One lane per column.  SUM_DIV(1) and SUM_DIV(nk) are zero (the
// reference's top/bottom statements write stale-index pad cells), so the
// dF/dpbot term vanishes and is not formed.  With with_source the theta
// heating of source_sink (src/shared/epic_funcs_diag.c:4768-4778) is applied
// first.
__global__ void k_calc_w(const MemberDesc *D, int mb, int with_source) {
  MEMBER;
  FLAT2(0, d.nj + 1, 0, d.ni + 2);
  const int nk = d.nk, ks = d.ks;
  const double *P = d.P3;
  double *W = d.W3, *TH = d.THETA;
  if (with_source)
    for (int K = ks - 1; K <= nk; K++) {
      size_t q = ix3(d, K, J, I);
      TH[q] += d.dt * d.HEAT3[q] / d.EXNER3[q];
    }
  W[ix3(d, 0, J, I)] = 0.;
  W[ix3(d, nk, J, I)] = 0.;
  const double ptop = P[ix3(d, 0, J, I)], pbot = P[ix3(d, nk, J, I)], psig = P[ix3(d, ks - 1, J, I)];
  for (int K = 1; K < ks - 1; K++) {
    size_t q = ix3(d, K, J, I);
    double sigma = get_sigma(pbot, P[q], ptop);
    W[q] = g_sig(d, sigma) * d.HEAT3[q] / d.EXNER3[q];
  }
  for (int K = ks - 1; K < nk; K++) W[ix3(d, K, J, I)] = 0.;
  double S = 0.;
  int bad = 0;
  const double eps = d.eps13;
#define PP(k) P[ix3(d, k, J, I)]
#define DV(k) d.DIV[ix3(d, k, J, I)]
  for (int K = 1; K < nk; K++) {
    if (K >= 2) {
      double dz10 = log(PP(K - 1) / PP(K - 2)), dz21 = log(PP(K) / PP(K - 1)), dz32 = log(PP(K + 1) / PP(K));
      double dz20 = dz21 + dz10, dz30 = dz32 + dz20, dz31 = dz32 + dz21;
      double df0 = PP(K - 2) * DV(K - 2), df1 = PP(K - 1) * DV(K - 1), df2 = PP(K) * DV(K), df3 = PP(K + 1) * DV(K + 1);
      double ds = -(df0 * (dz21 * dz21 * dz21 * (dz31 + dz32)) / (dz10 * dz20 * dz30) +
                    df1 * (dz21 * (dz21 * (dz21 + 2. * (dz10 - dz31)) - 6. * dz10 * dz31)) / (dz10 * dz31) +
                    df2 * (dz21 * (dz21 * (dz21 + 2. * (dz32 - dz20)) - 6. * dz20 * dz32)) / (dz20 * dz32) +
                    df3 * (dz21 * dz21 * dz21 * (dz10 + dz20)) / (dz30 * dz31 * dz32)) / 12.;
      S = S + ds;
    }
    size_t q = ix3(d, K, J, I);
    double th = TH[q], p3 = P[q], pb3 = pbot;
    double dp = p3 * eps;
    double p2 = dmax(p3 - dp * .5, ptop), p4 = dmin(p3 + dp * .5, pbot);
    if (p2 < psig && p4 >= psig) p2 = p3;
    double dFdp = 0.;
    if (p2 < p4) dFdp = (sgth_of(d, th, p2, pb3, ptop, &bad) - sgth_of(d, th, p4, pb3, ptop, &bad)) / (p2 - p4);
    else bad = 1;
    W[q] -= S * dFdp;
  }
#undef PP
#undef DV
  if (bad) atomicOr(&d.dyn->err, 2);
}
void L_calc_w(hipStream_t s, const Bucket &b, int with_source) {
  LAUNCH(k_calc_w, grid2(b.nj + 1, b.ni + 2, 1, b.nm), dim3(TPB), 0, s, b.D, b.mb, with_source);
}

// ------------------------------------------------ vorticity / divergence
// PV3 planes k0..k1 from U (uu) and a filtered copy of V (vv), with H3
__global__ void k_vort_interior(const MemberDesc *D, int mb, int uf, int vf, int k0) {
  MEMBER;
  FLAT2(1, d.nj, 1, d.ni);
  const int K = k0 + blockIdx.y;
  const double *uu = fld_ptr(d, uf) + (size_t)K * d.N2, *vv = fld_ptr(d, vf) + (size_t)K * d.N2;
  const double *hh = d.H3 + (size_t)K * d.N2;
  double *pv = d.PV3 + (size_t)K * d.N2;
  double f2j = d.f[2 * J], m2j = d.m[2 * J], n2j = d.n[2 * J];
  double mm1i = 1. / d.m[2 * J - 1], mp1i = 1. / d.m[2 * J + 1];
  double mnm1i = 1. / d.mn[2 * J - 1], mnp1i = 1. / d.mn[2 * J + 1];
  double mn_pv = .5 / (mnm1i + mnp1i);
  double ze = m2j * ((vv[ix2(d, J, I)] - vv[ix2(d, J, I - 1)]) +
                     n2j * (uu[ix2(d, J - 1, I)] * mm1i - uu[ix2(d, J, I)] * mp1i));
  double v = (ze + f2j);
  double h_pv = ((hh[ix2(d, J, I)] + hh[ix2(d, J, I - 1)]) * mnp1i + (hh[ix2(d, J - 1, I)] + hh[ix2(d, J - 1, I - 1)]) * mnm1i) * mn_pv;
  v /= h_pv;
  pv[ix2(d, J, I)] = v;
  if (I == d.ni) pv[ix2(d, J, 0)] = v;
  if (I == 1) pv[ix2(d, J, d.ni + 1)] = v;
}
// pole PV from the circulation of the adjacent u-row, one lane per (pole, K)
__global__ void k_vort_poles(const MemberDesc *D, int mb, int uf, int k0, int nK) {
  MEMBER;
  int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= 2 * nK) return;
  const int K = k0 + t / 2, north = t & 1;
  const double *uu = fld_ptr(d, uf) + (size_t)K * d.N2;
  const double *hh = d.H3 + (size_t)K * d.N2;
  double *pv = d.PV3 + (size_t)K * d.N2;
  const int ni = d.ni, nj = d.nj;
  double ze = 0., h = 0., val;
  if (!north) {
    for (int I = 1; I <= ni; I++) ze -= uu[ix2(d, 0, I)];
    ze *= d.mn[0] / ((d.m[1]) * (double)(ni));
    for (int I = 1; I <= ni; I++) h += hh[ix2(d, 0, I)];
    h /= (double)(ni);
    val = (ze + d.f[0]) / h;
    for (int I = 0; I <= ni + 1; I++) pv[ix2(d, 0, I)] = val;
  } else {
    for (int I = 1; I <= ni; I++) ze += uu[ix2(d, nj, I)];
    ze *= d.mn[2 * (nj + 1)] / (d.m[2 * nj + 1] * (double)(ni));
    for (int I = 1; I <= ni; I++) h += hh[ix2(d, nj, I)];
    h /= (double)(ni);
    val = (ze + d.f[2 * (nj + 1)]) / h;
    for (int I = 0; I <= ni + 1; I++) pv[ix2(d, nj + 1, I)] = val;
  }
}
void L_vorticity(hipStream_t s, const Bucket &b, int uu, int vv, int k0, int k1) {
  int nK = k1 - k0 + 1;
  LAUNCH(k_vort_interior, grid2(b.nj, b.ni, nK, b.nm), dim3(TPB), 0, s, b.D, b.mb, uu, vv, k0);
  LAUNCH(k_vort_poles, dim3((2 * nK + 63) / 64, 1, b.nm), dim3(64), 0, s, b.D, b.mb, uu, k0, nK);
}

__global__ void k_divergence(const MemberDesc *D, int mb, int uf, int vf, int df, int k0) {
  MEMBER;
  FLAT2(0, d.nj + 1, 1, d.ni);
  const int K = k0 + blockIdx.y;
  const double *uu = fld_ptr(d, uf) + (size_t)K * d.N2, *vv = fld_ptr(d, vf) + (size_t)K * d.N2;
  double *di = fld_ptr(d, df) + (size_t)K * d.N2;
  int jj = 2 * J + 1;
  double n_inv = 1. / d.n[jj], mn = d.mn[jj];
  double m2j_inv = (J == 0) ? 0. : 1. / d.m[jj - 1];
  double m2jp2_inv = (J == d.nj) ? 0. : 1. / d.m[jj + 1];
  double v = mn * ((uu[ix2(d, J, I + 1)] * n_inv - uu[ix2(d, J, I)] * n_inv) +
                   (vv[ix2(d, J + 1, I)] * m2jp2_inv - vv[ix2(d, J, I)] * m2j_inv));
  di[ix2(d, J, I)] = v;
  if (I == d.ni) di[ix2(d, J, 0)] = v;
  if (I == 1) di[ix2(d, J, d.ni + 1)] = v;
}
void L_divergence(hipStream_t s, const Bucket &b, int uu, int vv, int dst, int k0, int k1) {
  LAUNCH(k_divergence, grid2(b.nj + 1, b.ni, k1 - k0 + 1, b.nm), dim3(TPB), 0, s, b.D, b.mb, uu, vv, dst,
                     k0);
}

}  // namespace epic
