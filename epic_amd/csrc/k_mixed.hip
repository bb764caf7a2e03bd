// k_mixed.hip -- the momentum kernels of the mixed-precision mode.
//
// The state (u, v, h, theta), the hydrostatic diagnostics (P3, P2, GZ3,
// MONT3, RHO3, ...) and the Adams-Bashforth accumulation stay FP64; the
// Adams-Bashforth history planes of DUDT, DVDT are FP32, and these tendency
// terms are formed in FP32: Coriolis and kinetic energy (uv_core), the
// sponge and drag (uv_drag) and divergence damping.  The finite-volume
// pressure gradient is the FP64 kernel of k_dyn.hip in both modes: its face
// integrals are products of geopotential-sized numbers whose difference is
// the small force that balances Coriolis, and formed in FP32 (even on
// geopotential shifted by the cell's own value) they left u of the hot
// Jupiter at 1024x512x64 8.4e-4 (relative L2) from the FP64 run after 100
// steps against 1.1e-5 in FP64, which costs no time (both kernels wait on
// the same loads).  The terms of the current step are summed into one FP64
// plane (DUDT, DVDT hold a single plane in this mode) and rounded to FP32
// once, when the step stores it in the history: in balanced flow the v
// tendency is the small residual of the Coriolis and pressure-gradient
// terms, which an FP32 running sum would lose.  Vertical differences of u
// and v are formed in FP64 before they are rounded.  This file is selected
// per bucket when the engine runs in the mixed mode.
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

// the current step's tendency plane (FP64) and the FP32 history ring
__device__ __forceinline__ double *curU(const MemberDesc &d) { return d.DUDT; }
__device__ __forceinline__ double *curV(const MemberDesc &d) { return d.DVDT; }
__device__ __forceinline__ float *sAf(const MemberDesc &d) { return reinterpret_cast<float *>(d.sA); }
__device__ __forceinline__ float *sBf(const MemberDesc &d) { return reinterpret_cast<float *>(d.sB); }
__device__ __forceinline__ float *sCf(const MemberDesc &d) { return reinterpret_cast<float *>(d.sC); }

// ------------------------------------------------------------ tendencies
// UH, VH, KIN (FP32 views of sA, sB, sC) on planes 1..nk-1 with zonal pads
__global__ void __launch_bounds__(256) km_uvh_kin(const MemberDesc *D, int mb) {
  MEMBER_X;
  FLAT2_X(0, d.nj + 2, 1, d.ni);
  const int K = 1 + bx_.y;
  const size_t o = (size_t)K * d.N2;
  const double *u = d.U + o, *v = d.V + o, *h3 = d.H3 + o;
  float *uh = sAf(d) + o, *vh = sBf(d) + o, *kin = sCf(d) + o;
  const int ni = d.ni, nj = d.nj;
  if (J <= nj) {
    float havg = .5f * ((float)h3[ix2(d, J, I)] + (float)h3[ix2(d, J, I - 1)]);
    float x = (float)u[ix2(d, J, I)] * havg * (float)d.ninv_h[J];
    uh[ix2(d, J, I)] = x;
    if (I == ni) uh[ix2(d, J, 0)] = x;
    if (I == 1) uh[ix2(d, J, ni + 1)] = x;
    const int jj = 2 * J;
    const float mn1 = (float)d.mn[jj + 1], mi0 = (float)d.mn_inv[jj], mi2 = (float)d.mn_inv[jj + 2];
    float u2 = (float)u[ix2(d, J, I)], u4 = (float)u[ix2(d, J, I + 1)];
    float v2 = (float)v[ix2(d, J, I)], v4 = (float)v[ix2(d, J + 1, I)];
    float kc = (u2 * u2 + u4 * u4 + mn1 * (v2 * v2 * mi0 + v4 * v4 * mi2)) * .25f, k;
    if (d.pv_al81 || J == 0 || J == nj) {
      k = kc;
    } else {
      u2 = .5f * ((float)u[ix2(d, J + 1, I)] + (float)u[ix2(d, J - 1, I)]);
      u4 = .5f * ((float)u[ix2(d, J + 1, I + 1)] + (float)u[ix2(d, J - 1, I + 1)]);
      v2 = .5f * ((float)v[ix2(d, J, I + 1)] + (float)v[ix2(d, J, I - 1)]);
      v4 = .5f * ((float)v[ix2(d, J + 1, I + 1)] + (float)v[ix2(d, J + 1, I - 1)]);
      float ks = (u2 * u2 + u4 * u4 + mn1 * (v2 * v2 * mi0 + v4 * v4 * mi2)) * .25f;
      k = (5.f / 6.f) * kc + (1.f / 6.f) * ks;
    }
    kin[ix2(d, J, I)] = k;
    if (I == ni) kin[ix2(d, J, 0)] = k;
    if (I == 1) kin[ix2(d, J, ni + 1)] = k;
  }
  float y = 0.f;
  if (J >= 1 && J <= nj) {
    float havg = .5f * ((float)h3[ix2(d, J, I)] + (float)h3[ix2(d, J - 1, I)]);
    y = (float)v[ix2(d, J, I)] * havg * (float)d.minv_v[J];
  }
  vh[ix2(d, J, I)] = y;
  if (I == ni) vh[ix2(d, J, 0)] = y;
  if (I == 1) vh[ix2(d, J, ni + 1)] = y;
}

__global__ void __launch_bounds__(256) km_dvdt_core(const MemberDesc *D, int mb) {
  MEMBER_X;
  FLAT2_X(1, d.nj, 1, d.ni);
  const int K = 1 + bx_.y;
  const size_t o = (size_t)K * d.N2;
  const float *uh = sAf(d) + o, *vh = sBf(d) + o, *kin = sCf(d) + o;
  const double *pv = d.PV3 + o;
  double *dv = curV(d) + o;
  auto PV = [&](int j, int i) { return (float)pv[ix2(d, j, i)]; };
  const PvW<float> w = pv_weights_v<float>(d.pv_al81, PV, J, I);
  float t = -w.ga * uh[ix2(d, J, I + 1)] - w.de * uh[ix2(d, J, I)] - w.al * uh[ix2(d, J - 1, I)] -
            w.be * uh[ix2(d, J - 1, I + 1)];
  if (d.pv_al81) t = t + w.e2 * vh[ix2(d, J - 1, I)] - w.e1 * vh[ix2(d, J + 1, I)];
  float x = (t * w.coef + (kin[ix2(d, J - 1, I)] - kin[ix2(d, J, I)])) * (float)d.n[2 * J];
  const int kk = 2 * K + 1;
  const float ds_inv = (float)(1. / (d.sgth[kk - 2] - d.sgth[kk + 2]));
  const float dvz = (float)(d.V[ix3(d, K - 1, J, I)] - d.V[ix3(d, K + 1, J, I)]);
  x -= .5f * (float)(d.W3[ix3(d, K, J, I)] + d.W3[ix3(d, K, J - 1, I)]) * dvz * ds_inv;
  dv[ix2(d, J, I)] += (double)x;
}

// One block per (K,J) u-row group: the row of KIN is zonally filtered in
// LDS (rows of the filter band, the FP32 filter of the 2-D zonal filter),
// then one lane per I forms the u tendency, as k_dudt_core does in FP64.
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(8))) km_dudt_core(const MemberDesc *D, int mb) {
  MEMBER;
  extern __shared__ double lds_raw[];
  float *lds = reinterpret_cast<float *>(lds_raw);
  const RowMap m = row_map(d.ni);
  const int Jr = blockIdx.x * m.rpb + m.g, K = 1 + blockIdx.y, ni = d.ni;
  const bool live = Jr <= d.nj;
  const int J = live ? Jr : 0;
  const size_t o = (size_t)K * d.N2;
  const float *uh = sAf(d) + o, *vh = sBf(d) + o, *kin = sCf(d) + o;
  float *kr = lds + (size_t)m.g * 2 * ni, *work = kr + ni;  // kr[I-1] = KIN(J,I)
  if (live)
    for (int I = 1 + m.lane; I <= ni; I += m.stride) kr[I - 1] = kin[ix2(d, J, I)];
  __syncthreads();
  filter_row_lds(d, kr, work, J, 0, live && d.rflag_h[J] != 0, m.lane);
  __syncthreads();
  if (!live) return;
  const double *pv = d.PV3 + o;
  double *du = curU(d) + o;
  const int kk = 2 * K + 1;
  const float ds_inv = (float)(1. / (d.sgth[kk - 2] - d.sgth[kk + 2])), mj = (float)d.m[2 * J + 1];
  for (int I = 1 + m.lane; I <= ni; I += m.stride) {
    auto PV = [&](int j, int i) { return (float)pv[ix2(d, j, i)]; };
    const PvW<float> w = pv_weights_u<float>(d.pv_al81, PV, J, I);
    float t = w.al * vh[ix2(d, J + 1, I)] + w.be * vh[ix2(d, J + 1, I - 1)] + w.ga * vh[ix2(d, J, I - 1)] +
              w.de * vh[ix2(d, J, I)];
    if (d.pv_al81) t = t + w.e2 * uh[ix2(d, J, I - 1)] - w.e1 * uh[ix2(d, J, I + 1)];
    const float kin_w = kr[I == 1 ? ni - 1 : I - 2], kin_c = kr[I - 1];
    float x = (t * w.coef + (kin_w - kin_c)) * mj;
    const float duz = (float)(d.U[ix3(d, K - 1, J, I)] - d.U[ix3(d, K + 1, J, I)]);
    x -= .5f * (float)(d.W3[ix3(d, K, J, I)] + d.W3[ix3(d, K, J, I - 1)]) * duz * ds_inv;
    du[ix2(d, J, I)] += (double)x;
  }
}

// ------------------------------------------------------------- uv_drag
__global__ void __launch_bounds__(256) km_sponge(const MemberDesc *D, int mb) {
  MEMBER;
  FLAT2(0, d.nj + 1, 1, d.ni);
  const int K = blockIdx.y;
  if (K > d.k_sponge) return;
  double *du = curU(d), *dv = curV(d);
  const size_t q = ix3(d, K, J, I);
  const float ts = (float)d.t_sponge[K];
  const double x = du[q] - (double)(ts * (float)(d.U[q] - d.USPIN[q]));
  du[q] = x;
  // zonal pads of plane K = 0, which no later kernel writes (the 3-D
  // filter's bc_lateral does in the reference)
  if (I == d.ni) du[ix3(d, K, J, 0)] = x;
  if (I == 1) du[ix3(d, K, J, d.ni + 1)] = x;
  if (J >= 1) {
    const double y = dv[q] - (double)(ts * (float)d.V[q]);
    dv[q] = y;
    if (I == d.ni) dv[ix3(d, K, J, 0)] = y;
    if (I == 1) dv[ix3(d, K, J, d.ni + 1)] = y;
  }
}

__device__ __forceinline__ float t_rad_f(const MemberDesc &d, int K, int J, int I) {
  if (d.is_hj) return (float)d.tau_rad;
  if (d.is_tab) return (float)tab_cool_time(d, -crm_log(d.P3[ix3(d, K, J, I)]));
  const float ka = (float)(1. / (40. * 60. * 60. * 24.)), ks = (float)(1. / (4. * 60. * 60. * 24.));
  float p = (float)d.P3[ix3(d, K, J, I)], kt = ka, pbot = (float)d.P3[ix3(d, d.nk, J, I)];
  float amp = (p / pbot - 0.7f) / (1.f - 0.7f);
  if (amp > 0.f) {
    float c = cosf((float)d.lat[2 * J + 1] * (float)(M_PI / 180.));
    kt += (ks - ka) * (c * c) * (c * c) * amp;
  }
  return 1.f / kt;
}

// Rayleigh drag: one wavefront per (K,J) row, the zonal mean by a wavefront
// reduction
__global__ void __launch_bounds__(256) km_rayleigh(const MemberDesc *D, int mb) {
  MEMBER;
  if (!((d.newt_cool_on && d.prandtl > 0.) || (d.tau_drag > 0. && d.tau_drag < 1.e+20))) return;
  const int K = d.k_sponge + 1 + blockIdx.y, J = blockIdx.x;
  if (K > d.nk - 1 || J > d.nj) return;
  double *du = curU(d), *dv = curV(d);
  const int ni = d.ni, lane = threadIdx.x;
  const float za = (float)d.drag_zonal_avg, zb = (float)(1 - d.drag_zonal_avg);
  for (int w = 0; w < 2; w++) {
    if (w == 1 && !(d.drag_v && J >= 1)) break;
    const double *X = w ? d.V : d.U;
    float s = 0.f;
    for (int I = 1 + lane; I <= ni; I += 64) s += (float)X[ix3(d, K, J, I)];
    for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off, 64);
    const float avg = s / (float)ni;
    for (int I = 1 + lane; I <= ni; I += 64) {
      const size_t q = ix3(d, K, J, I);
      float ray = za * avg + zb * (float)X[q];
      float nu0;
      if (d.newt_cool_on) {
        float t0 = t_rad_f(d, K, J, I), t1 = w ? t_rad_f(d, K, J - 1, I) : t_rad_f(d, K, J, I - 1);
        nu0 = (float)d.prandtl / (.5f * (t0 + t1));
      } else {
        nu0 = (float)(1. / d.tau_drag);
      }
      if (w == 0) du[q] -= (double)(nu0 * (ray - (float)d.USPIN[q]));
      else dv[q] -= (double)(nu0 * ray);
    }
  }
}

__global__ void __launch_bounds__(256) km_hs_pbl(const MemberDesc *D, int mb) {
  MEMBER;
  FLAT2(0, d.nj + 1, 1, d.ni);
  if (!d.is_hs) return;
  const float nu00 = (float)(1. / (24. * 60. * 60.));
  const int nk = d.nk;
  double *du = curU(d), *dv = curV(d);
  const double *P = d.P3;
  float pb = 0.5f * (float)(P[ix3(d, nk, J, I)] + P[ix3(d, nk, J, I - 1)]);
  for (int K = nk - 1; K >= 1; K--) {
    float p3 = .5f * (float)(P[ix3(d, K, J, I)] + P[ix3(d, K, J, I - 1)]);
    float amp = (p3 / pb - 0.7f) / (1.f - 0.7f);
    if (amp < 0.f) break;
    du[ix3(d, K, J, I)] -= (double)(nu00 * amp * (float)d.U[ix3(d, K, J, I)]);
  }
  if (J >= 1) {
    pb = 0.5f * (float)(P[ix3(d, nk, J, I)] + P[ix3(d, nk, J - 1, I)]);
    for (int K = nk - 1; K >= 1; K--) {
      float p3 = .5f * (float)(P[ix3(d, K, J, I)] + P[ix3(d, K, J - 1, I)]);
      float amp = (p3 / pb - 0.7f) / (1.f - 0.7f);
      if (amp < 0.f) break;
      dv[ix3(d, K, J, I)] -= (double)(nu00 * amp * (float)d.V[ix3(d, K, J, I)]);
    }
  }
}

// ------------------------------------------------------ divergence damping
__global__ void __launch_bounds__(256) km_divdamp(const MemberDesc *D, int mb) {
  MEMBER;
  FLAT2(0, d.nj + 1, 1, d.ni);
  const int K = 1 + blockIdx.y;
  if (d.nudiv == 0.) return;
  const double *div = d.sB + (size_t)K * d.N2;
  double *du = curU(d), *dv = curV(d);
  float coef = (float)((d.lat[2 * J + 1] > 45.) ? d.nudiv * d.m0z : d.nudiv * d.m[2 * J + 1]);
  du[ix3(d, K, J, I)] += (double)(coef * (float)(div[ix2(d, J, I)] - div[ix2(d, J, I - 1)]));
  if (J >= 1) {
    float cv = (float)(d.nudiv * d.n[2 * J]);
    dv[ix3(d, K, J, I)] += (double)(cv * (float)(div[ix2(d, J, I)] - div[ix2(d, J - 1, I)]));
  }
}

// ------------------------------------------------------- Adams-Bashforth
// FP64 accumulation: the current (FP64) tendency and the FP32 history
// planes IT_MINUS1, IT_MINUS2 (FLT_MAX flags start AB1/AB2); the current
// tendency is then stored, rounded to FP32, in the history slot IT_ZERO,
// which the rotation at the end of the step makes IT_MINUS1.  Lanes cover
// every cell of planes 1..nk-1 so the stored history is the whole plane.
__global__ void __launch_bounds__(256) km_ab(const MemberDesc *D, int mb) {
  MEMBER;
  FLAT2(0, d.nj + 1, 1, d.ni);
  const int K = 1 + blockIdx.y;
  const MemberDyn dy = *d.dyn;
  const float FMAX = 3.40282347e+38f;
  for (int w = 0; w < 2; w++) {
    const int jlo = w;
    double *val = w ? d.V : d.U;
    const double *cur = w ? d.DVDT : d.DUDT;
    float *hist = w ? d.DVDTf : d.DUDTf;
    const size_t q = ix3(d, K, J, I);
    const double t0 = cur[q];
    if (J >= jlo) {
      double ab0, ab1, ab2;
      if (hist[(size_t)dy.itm2 * d.N3 + ix3(d, 1, jlo, 1)] == FMAX) {
        if (hist[(size_t)dy.itm1 * d.N3 + ix3(d, 1, jlo, 1)] == FMAX) { ab0 = 1.; ab1 = 0.; ab2 = 0.; }
        else { ab0 = 3. / 2.; ab1 = -1. / 2.; ab2 = 0.; }
      } else {
        ab0 = 23. / 12.; ab1 = -16. / 12.; ab2 = 5. / 12.;
      }
      const double t1 = (ab1 != 0.) ? hist[(size_t)dy.itm1 * d.N3 + q] : 0.,
                   t2 = (ab2 != 0.) ? hist[(size_t)dy.itm2 * d.N3 + q] : 0.;
      const double y = val[q] + d.dt * (ab0 * t0 + ab1 * t1 + ab2 * t2);
      val[q] = y;
      if (I == d.ni) val[ix3(d, K, J, 0)] = y;
      if (I == 1) val[ix3(d, K, J, d.ni + 1)] = y;
    }
    hist[(size_t)dy.it0 * d.N3 + q] = (float)t0;
  }
}

// ------------------------------------------------------------- launchers
void LM_uv_core(hipStream_t s, const Bucket &b) {
  const int nK = b.nk - 1;
  LAUNCH(km_uvh_kin, grid2(b.nj + 2, b.ni, nK, b.nm), dim3(TPB), 0, s, b.D, b.mb);
  LAUNCH(km_dvdt_core, grid2(b.nj, b.ni, nK, b.nm), dim3(TPB), 0, s, b.D, b.mb);
  const int rpb = rows_per_block(b.ni);
  LAUNCH(km_dudt_core, dim3((b.nj + 1 + rpb - 1) / rpb, nK, b.nm), dim3(256), sizeof(float) * rpb * 2 * b.ni, s,
         b.D, b.mb);
}
void LM_uv_drag(hipStream_t s, const Bucket &b) {
  LAUNCH(km_sponge, grid2(b.nj + 1, b.ni, b.nk, b.nm), dim3(TPB), 0, s, b.D, b.mb);
  if (b.any_ray) LAUNCH(km_rayleigh, dim3(b.nj + 1, b.nk, b.nm), dim3(64), 0, s, b.D, b.mb);
  if (b.any_hs) LAUNCH(km_hs_pbl, grid2(b.nj + 1, b.ni, 1, b.nm), dim3(TPB), 0, s, b.D, b.mb);
}
void LM_divdamp_apply(hipStream_t s, const Bucket &b) {
  LAUNCH(km_divdamp, grid2(b.nj + 1, b.ni, b.nk - 1, b.nm), dim3(TPB), 0, s, b.D, b.mb);
}
void LM_ab(hipStream_t s, const Bucket &b) {
  LAUNCH(km_ab, grid2(b.nj + 1, b.ni, b.nk - 1, b.nm), dim3(TPB), 0, s, b.D, b.mb);
}

}  // namespace epic
