// k_adv.hip -- flux-form predictor-corrector advection of Hsu & Arakawa as
// modified by Konor & Arakawa (hsu_advection, src/shared/epic_flux.c:652-1488)
// and its driver (advection, src/shared/epic_flux.c:346-637).
//
// Directions are split as in the reference: vertical column pass, then a
// meridional pass on every level, then a zonal pass on a zonally filtered
// copy.  The column pass runs one lane per (J,I) column.  The horizontal
// passes run one lane per cell and recompute the few neighbouring fluxes
// they need (stencil radius 2-3) instead of staging intermediate planes;
// every intermediate value is formed by the same expression as in the
// reference, so nothing changes but where it is computed.
//
// The reference keeps its column work arrays in statics across calls: the
// theta call starts at K = k_sigma-1 and reads wm[k_sigma-2] and
// w_a[k_sigma-3] as left by the last column of the preceding unity call.
// Those two values are handed over explicitly (wstale).
#include "kernels.h"

namespace epic {

#define MEMBER const MemberDesc &d = D[mb + blockIdx.z]

enum { AV_P3 = 0, AV_THETA = 1, AV_HDRY = 2 };

__device__ __forceinline__ double *adv_field(const MemberDesc &d, int which) {
  return which == AV_P3 ? d.unity : (which == AV_THETA ? d.THETA : d.HDRY);
}

// ------------------------------------------------------------ vertical
__global__ void k_hsu_vert(const MemberDesc *D, int mb, int which) {
  MEMBER;
  int t = blockIdx.x * TPB + threadIdx.x;
  const int nI = d.ni + 2;
  if (t >= (d.nj + 1) * nI) return;
  const int J = t / nI, I = t % nI;
  const int nk = d.nk;
  const int layer = (which == AV_HDRY);
  const int kf = (which == AV_THETA) ? d.ks - 1 : 1;
  const double ep = 1.e-10, DT = d.dt;
  const double *dsi = d.dsgth_inv;
  double *b = adv_field(d, which);
  double wa[KMAX + 2], wm[KMAX + 2], wp[KMAX + 2], pr[KMAX + 2], ff[KMAX + 2], d1[KMAX + 2], d2[KMAX + 2],
      aaa[KMAX + 2], amin[KMAX + 2];
  for (int K = 0; K <= nk + 1; K++) amin[K] = 0.;
  if (layer)
    for (int K = kf; K <= nk; K++) amin[K] = d.h_min[K];
  if (which == AV_THETA) {
    if (kf - 1 >= 1) wm[kf - 1] = d.wstale[0];
    if (kf - 2 >= 0) wa[kf - 2] = d.wstale[1];
  }
  for (int K = layer ? kf : kf - 1; K <= nk; K++) wa[K] = b[ix3(d, K, J, I)];
  if (layer) {
    for (int K = kf - 1; K <= nk; K++) {
      double w = d.W3[ix3(d, K, J, I)];
      wm[K] = dmin(0., w);
      wp[K] = dmax(0., w);
    }
  } else {
    for (int K = kf; K <= nk; K++) {
      double w = (K == 1 || K == nk) ? 0. : .5 * (d.W3[ix3(d, K, J, I)] + d.W3[ix3(d, K - 1, J, I)]);
      wm[K] = dmin(0., w);
      wp[K] = dmax(0., w);
    }
  }
  if (which == AV_P3 && J == d.nj && I == d.ni + 1) {
    // hand the statics the theta call will read to it
    const int kft = d.ks - 1;
    if (kft - 1 >= 1) d.wstale[0] = wm[kft - 1];
    if (kft - 2 >= 0) d.wstale[1] = wa[kft - 2];
  }
  if (layer) {
    ff[kf - 1] = 0.;
    ff[nk] = 0.;
    for (int K = kf; K < nk; K++) ff[K] = wp[K] * wa[K + 1] + wm[K] * wa[K];
    for (int K = kf; K <= nk; K++) pr[K] = wa[K] + (ff[K] - ff[K - 1]) * dsi[2 * K] * DT;
    for (int K = kf; K < nk; K++) ff[K] = .5 * (wp[K] * (pr[K] + wa[K + 1]) + wm[K] * (pr[K + 1] + wa[K]));
    for (int K = kf; K <= nk; K++) {
      double tt;
      if (K == 1) { d1[K] = pr[K] - wa[K + 1]; d2[K] = 0.; tt = fabs(wa[K + 1] - wa[K]) + ep; }
      else if (K == nk) { d1[K] = 0.; d2[K] = pr[K] - wa[K - 1]; tt = fabs(-wa[K] + wa[K - 1]) + ep; }
      else { d1[K] = pr[K] - wa[K + 1]; d2[K] = pr[K] - wa[K - 1]; tt = fabs(wa[K + 1] - 2. * wa[K] + wa[K - 1]) + ep; }
      aaa[K] = tt * tt;
    }
    for (int K = kf; K < nk; K++) {
      double whp = sqrt(wp[K] * wp[K + 1]), whm = -sqrt(wm[K] * wm[K - 1]);
      double mu = (wp[K] - wm[K]) * dsi[2 * K + 1] * DT, al = (1. + mu) / 6.;
      double gp = aaa[K + 1] / (aaa[K + 1] + (wa[K + 1] - amin[K + 1]) * (wa[K] - amin[K]));
      gp = gp * gp;
      double gm = aaa[K] / (aaa[K] + (wa[K + 1] - amin[K + 1]) * (wa[K] - amin[K]));
      gm = gm * gm;
      double bp = 1. + (1. / (2. * al) - 1.) * gp, bm = 1. + (1. / (2. * al) - 1.) * gm;
      double bhp = 1. - gp, bhm = 1. - gm;
      ff[K] += -al * (wp[K] * bp * d1[K] - whp * bhp * d1[K + 1] + wm[K] * bm * d2[K + 1] - whm * bhm * d2[K]);
    }
    for (int K = kf; K <= nk; K++) b[ix3(d, K, J, I)] += (ff[K] - ff[K - 1]) * dsi[2 * K] * DT;
  } else {
    for (int K = kf; K <= nk; K++) ff[K] = wp[K] * wa[K] + wm[K] * wa[K - 1];
    ff[nk + 1] = 0.;
    pr[kf - 1] = wa[kf - 1];
    for (int K = kf; K <= nk; K++) pr[K] = wa[K] + (ff[K + 1] - ff[K]) * dsi[2 * K + 1] * DT;
    for (int K = kf; K <= nk; K++) ff[K] = .5 * (wp[K] * (pr[K - 1] + wa[K]) + wm[K] * (pr[K] + wa[K - 1]));
    for (int K = kf - 1; K <= nk; K++) {
      double tt;
      if (K == 0) tt = fabs(wa[K + 1] - wa[K]) + ep;
      else if (K == nk) tt = fabs(-wa[K] + wa[K - 1]) + ep;
      else tt = fabs(wa[K + 1] - 2. * wa[K] + wa[K - 1]) + ep;
      aaa[K] = tt * tt;
    }
    d2[kf - 1] = 0.;
    for (int K = kf; K <= nk; K++) {
      d1[K] = pr[K - 1] - wa[K];
      d2[K] = pr[K] - wa[K - 1];
    }
    for (int K = kf; K <= nk; K++) {
      double whm = (K == 1) ? 0. : -sqrt(wm[K] * wm[K - 1]);
      double whp = (K == nk) ? 0. : sqrt(wp[K] * wp[K + 1]);
      double mu = (wp[K] - wm[K]) * dsi[2 * K] * DT, al = (1. + mu) / 6.;
      double gp = aaa[K] / (aaa[K] + (wa[K] - amin[K]) * (wa[K - 1] - amin[K - 1]));
      gp = gp * gp;
      double gm = aaa[K - 1] / (aaa[K - 1] + (wa[K] - amin[K]) * (wa[K - 1] - amin[K - 1]));
      gm = gm * gm;
      double bp = 1. + (1. / (2. * al) - 1.) * gp, bm = 1. + (1. / (2. * al) - 1.) * gm;
      double bhp = 1. - gp, bhm = 1. - gm, g;
      if (K < nk) g = -al * (wp[K] * bp * d1[K] - whp * bhp * d1[K + 1] + wm[K] * bm * d2[K] - whm * bhm * d2[K - 1]);
      else g = -al * (wp[K] * bp * d1[K] + wm[K] * bm * d2[K] - whm * bhm * d2[K - 1]);
      ff[K] += g;
    }
    ff[nk + 1] = 0.;
    for (int K = kf; K <= nk; K++) b[ix3(d, K, J, I)] += (ff[K + 1] - ff[K]) * dsi[2 * K + 1] * DT;
  }
}

// ---------------------------------------------------------- meridional
struct MerCol {
  const MemberDesc *d;
  const double *A;  // plane K of the field
  int K, I, which;
  double amin, DT;
  __device__ double a(int J) const { return A[ix2(*d, J, I)]; }
  __device__ double v2(int J) const {
    const MemberDesc &m = *d;
    if (which == AV_HDRY) {
      if (J < 1 || J > m.nj) return 0.;
      return .5 * (m.V[ix3(m, K, J, I)] + m.V[ix3(m, K - 1, J, I)]);
    }
    return m.V[ix3(m, K, J, I)];
  }
  __device__ double vm(int J) const {
    double mi = (fabs(d->lat[2 * J]) == 90.) ? 0. : 1. / d->m[2 * J];
    return dmin(0., v2(J) * mi);
  }
  __device__ double vp(int J) const {
    double mi = (fabs(d->lat[2 * J]) == 90.) ? 0. : 1. / d->m[2 * J];
    return dmax(0., v2(J) * mi);
  }
  __device__ double ffp(int J) const {
    if (J < 1 || J > d->nj) return 0.;
    return vp(J) * a(J - 1) + vm(J) * a(J);
  }
  __device__ double ap(int J) const { return a(J) + (ffp(J) - ffp(J + 1)) * d->mn[2 * J + 1] * DT; }
  __device__ double dif1(int J) const { return J == 0 ? 0. : ap(J) - a(J - 1); }
  __device__ double dif2(int J) const { return J == d->nj ? 0. : ap(J) - a(J + 1); }
  __device__ double aq(int J) const {
    const double ep = 1.e-10;
    double t;
    if (J == 0) t = fabs(-a(0) + a(1)) + ep;
    else if (J == d->nj) t = fabs(a(J - 1) - a(J)) + ep;
    else t = fabs(a(J - 1) - 2. * a(J) + a(J + 1)) + ep;
    return t * t;
  }
  __device__ double ff(int J) const {
    if (J < 1 || J > d->nj) return 0.;
    double VP = vp(J), VM = vm(J), APj = ap(J), APm = ap(J - 1), Am = a(J - 1), Aj = a(J);
    double fc = .5 * (VP * (APj + Am) + VM * (APm + Aj));
    double vhp = sqrt(VP * vp(J - 1)), vhm = -sqrt(VM * vm(J + 1));
    double mu = (VP - VM) * d->mn[2 * J] * DT, al = (1. + mu) / 6.;
    double q0 = aq(J - 1), q1 = aq(J);
    double gp = q0 / (q0 + (Am - amin) * (Aj - amin));
    gp = gp * gp;
    double gm = q1 / (q1 + (Am - amin) * (Aj - amin));
    gm = gm * gm;
    double bp = 1. + (1. / (2. * al) - 1.) * gp, bm = 1. + (1. / (2. * al) - 1.) * gm;
    double bhp = 1. - gp, bhm = 1. - gm;
    double d1J = (APj - Am), d1m = dif1(J - 1), d2m = dif2(J - 1), d2J = dif2(J);
    return fc + -al * (VP * bp * d1J - vhp * bhp * d1m + VM * bm * d2m - vhm * bhm * d2J);
  }
};

__global__ void k_hsu_merid(const MemberDesc *D, int mb, int which) {
  MEMBER;
  int t = blockIdx.x * TPB + threadIdx.x;
  if (t >= (d.nj + 1) * d.ni) return;
  const int J = t / d.ni, I = 1 + t % d.ni;
  const int K = 1 + blockIdx.y;
  if (which == AV_THETA && K <= d.ks - 2) return;
  MerCol c;
  c.d = &d;
  double *F = adv_field(d, which);
  c.A = F + (size_t)K * d.N2;
  c.K = K;
  c.I = I;
  c.which = which;
  c.amin = (which == AV_HDRY) ? d.h_min[K] : 0.;
  c.DT = d.dt;
  double v = c.a(J) + (c.ff(J) - c.ff(J + 1)) * d.mn[2 * J + 1] * c.DT;
  size_t q = ix3(d, K, J, I);
  d.sA[q] = v;
  d.sB[q] = v;
}

// --------------------------------------------------------------- zonal
struct ZonRow {
  const MemberDesc *d;
  const double *af;  // filtered copy, plane K, with pads
  int K, J, which, ni;
  double amin, DT, ninv, mn;
  __device__ int w(int I) const { return I < 1 ? I + ni : (I > ni ? I - ni : I); }
  __device__ double a(int I) const { return af[ix2(*d, J, I)]; }
  __device__ double u2(int I) const {
    const MemberDesc &m = *d;
    if (which == AV_HDRY) {
      int Ii = w(I);
      return .5 * (m.U[ix3(m, K, J, Ii)] + m.U[ix3(m, K - 1, J, Ii)]);
    }
    return m.U[ix3(m, K, J, I)];
  }
  __device__ double um(int I) const { return dmin(0., u2(I) * ninv); }
  __device__ double up(int I) const { return dmax(0., u2(I) * ninv); }
  // intermediates at c in 1..ni
  __device__ double ffp(int c) const { return up(c) * a(c - 1) + um(c) * a(c); }
  __device__ double ap(int c) const { return a(c) + (ffp(c) - ffp(w(c + 1))) * mn * DT; }
  __device__ double aq(int c) const {
    double t = fabs(a(c - 1) - 2. * a(c) + a(c + 1)) + 1.e-10;
    return t * t;
  }
  __device__ double ff(int c) const {
    double UP = up(c), UM = um(c), APc = ap(c), APm = ap(w(c - 1)), Am = a(c - 1), Ac = a(c);
    double fc = .5 * (UP * (APc + Am) + UM * (APm + Ac));
    double d1c = APc - Am, d2c = APc - a(c + 1);
    int cm = w(c - 1);
    double d1m = APm - a(cm - 1), d2m = APm - a(cm + 1);
    double uhp = sqrt(UP * up(c - 1)), uhm = -sqrt(UM * um(c + 1));
    double mu = (UP - UM) * mn * DT, al = (1. + mu) / 6.;
    double q0 = aq(cm), q1 = aq(c);
    double gp = q0 / (q0 + (Am - amin) * (Ac - amin));
    gp = gp * gp;
    double gm = q1 / (q1 + (Am - amin) * (Ac - amin));
    gm = gm * gm;
    double bp = 1. + (1. / (2. * al) - 1.) * gp, bm = 1. + (1. / (2. * al) - 1.) * gm;
    double bhp = 1. - gp, bhm = 1. - gm;
    return fc + -al * (UP * bp * d1c - uhp * bhp * d1m + UM * bm * d2m - uhm * bhm * d2c);
  }
};

__global__ void k_hsu_zonal(const MemberDesc *D, int mb, int which) {
  MEMBER;
  int t = blockIdx.x * TPB + threadIdx.x;
  if (t >= (d.nj + 1) * d.ni) return;
  const int J = t / d.ni, I = 1 + t % d.ni;
  const int K = 1 + blockIdx.y;
  if (which == AV_THETA && K <= d.ks - 2) return;
  ZonRow z;
  z.d = &d;
  z.af = d.sB + (size_t)K * d.N2;
  z.K = K;
  z.J = J;
  z.which = which;
  z.ni = d.ni;
  z.amin = (which == AV_HDRY) ? d.h_min[K] : 0.;
  z.DT = d.dt;
  z.ninv = 1. / d.n[2 * J + 1];
  z.mn = d.mn[2 * J + 1];
  double da = (z.ff(I) - z.ff(z.w(I + 1))) * z.mn * z.DT;
  size_t q = ix3(d, K, J, I);
  adv_field(d, which)[q] = d.sA[q] + da;
}

static void hsu(hipStream_t s, const Bucket &b, int which, int fld) {
  const int TP = TPB;
  int ncol = (b.nj + 1) * (b.ni + 2);
  LAUNCH(k_hsu_vert, dim3((ncol + TP - 1) / TP, 1, b.nm), dim3(TP), 0, s, b.D, b.mb, which);
  int ncell = (b.nj + 1) * b.ni;
  dim3 g((ncell + TP - 1) / TP, b.nk, b.nm);
  LAUNCH(k_hsu_merid, g, dim3(TP), 0, s, b.D, b.mb, which);
  L_filter(s, b, F_SB, 1, b.nk, 0, 0);
  LAUNCH(k_hsu_zonal, g, dim3(TP), 0, s, b.D, b.mb, which);
  L_bc3(s, b, fld);
}

// unity = 1 on K = 0..nk+1, J = 0..nj, I = 0..ni+1; OLD = THETA and the
// theta correction on K = k_sigma-1..nk+1 over the same rows
__global__ void k_adv_elem(const MemberDesc *D, int mb, int op) {
  MEMBER;
  int t = blockIdx.x * TPB + threadIdx.x;
  const int nI = d.ni + 2;
  if (t >= (d.nj + 1) * nI) return;
  const int J = t / nI, I = t % nI, K = blockIdx.y;
  if (K > d.nk + 1) return;
  size_t q = ix3(d, K, J, I);
  if (op == 0) {
    d.unity[q] = 1.;
  } else if (K >= d.ks - 1) {
    if (op == 1) d.old[q] = d.THETA[q];
    else d.THETA[q] -= d.old[q] * (d.unity[q] - 1.);
  }
}

void L_advection(hipStream_t s, const Bucket &b) {
  int ncol = (b.nj + 1) * (b.ni + 2);
  dim3 ge((ncol + TPB - 1) / TPB, b.nk + 2, b.nm);
  LAUNCH(k_adv_elem, ge, dim3(TPB), 0, s, b.D, b.mb, 0);
  hsu(s, b, AV_P3, F_UNITY);
  LAUNCH(k_adv_elem, ge, dim3(TPB), 0, s, b.D, b.mb, 1);
  hsu(s, b, AV_THETA, F_THETA);
  LAUNCH(k_adv_elem, ge, dim3(TPB), 0, s, b.D, b.mb, 2);
  hsu(s, b, AV_HDRY, F_HDRY);
  L_restore_mass(s, b);
  L_set_p2etc(s, b, 0);
}

}  // namespace epic
