// k_adv.hip -- flux-form predictor-corrector advection of Hsu & Arakawa as
// modified by Konor & Arakawa (hsu_advection, src/shared/epic_flux.c:652-1488),
// the monotonized fourth-order centred scheme offered beside it (Akima,
// akima_advection, src/shared/epic_flux.c:1502-1924), and their driver
// (advection, src/shared/epic_flux.c:346-637).
//
// Directions are split as in the reference: vertical column pass, then a
// meridional pass on every level, then a zonal pass on a zonally filtered
// copy.  The column pass works on LDS column tiles, the meridional pass on
// LDS tiles of a few rows by up to 64 columns (each intermediate of the
// tile formed once), the zonal pass on LDS rows; every intermediate value
// is formed by the same expression as in the reference, so nothing changes
// but where it is computed.
//
// The three advected fields (unity on the pressure grid, theta, hdry) are
// independent of one another until the theta correction that follows them,
// so every launch carries all three: blockIdx.z = field * members + member,
// each field with its own scratch copies (sA..sD + field * N3).
//
// The reference keeps its column work arrays in statics across calls: the
// theta call starts at K = k_sigma-1 and reads wm[k_sigma-2] and
// w_a[k_sigma-3] as left by the last column (J = nj, I = ni+1) of the
// preceding unity call.  Both are functions of W3 and of unity before its
// advection (1 everywhere), so the theta blocks form them themselves.
#include "kernels.h"
#include "filter_row.h"

namespace epic {

#define MEMBER const MemberDesc &d = D[mb + blockIdx.z]
// field and member of a block; so: offset of the field's scratch copies
#define ADV_MEMBER                                          \
  const int which = (int)blockIdx.z / nm;                   \
  const MemberDesc &d = D[mb + (int)blockIdx.z - which * nm]; \
  const size_t so = (size_t)which * d.N3
constexpr int ADV_FIELDS = 3;

enum { AV_P3 = 0, AV_THETA = 1, AV_HDRY = 2 };

__device__ __forceinline__ double *adv_field(const MemberDesc &d, int which) {
  return which == AV_P3 ? d.unity : (which == AV_THETA ? d.THETA : d.HDRY);
}

// ------------------------------------------------------------ vertical
// Column tiles of C columns (J = 0..nj, I = 0..ni+1) with every level in LDS.
// The pass is a chain of K-stencils (first-guess flux, predictor, centred
// flux, limiter correction, update) with no recurrence, so each stage runs
// on all lanes over (column, level) pairs, separated by barriers.
//
// The advection passes are templated on the arithmetic type T of the fluxes:
// double reproduces the reference operation for operation; float is the
// mixed mode's (the fields stay FP64: each update adds the FP32 flux
// divergence to the FP64 value).
template <class T> __device__ __forceinline__ T tmin(T x, T y) { return x < y ? x : y; }
template <class T> __device__ __forceinline__ T tmax(T x, T y) { return x > y ? x : y; }

template <class T>
__device__ __forceinline__ T tt_layer(const T *wa, int K, int nk, T ep) {
  T tt;
  if (K == 1) tt = fabs(wa[K + 1] - wa[K]) + ep;
  else if (K == nk) tt = fabs(-wa[K] + wa[K - 1]) + ep;
  else tt = fabs(wa[K + 1] - T(2.) * wa[K] + wa[K - 1]) + ep;
  return tt * tt;
}
template <class T>
__device__ __forceinline__ T tt_iface(const T *wa, int K, int nk, T ep) {
  T tt;
  if (K == 0) tt = fabs(wa[K + 1] - wa[K]) + ep;
  else if (K == nk) tt = fabs(-wa[K] + wa[K - 1]) + ep;
  else tt = fabs(wa[K + 1] - T(2.) * wa[K] + wa[K - 1]) + ep;
  return tt * tt;
}

// The FP64 value a pass updates: the staged copy itself when the pass
// computes in FP64 (the same bits), else a re-read of the FP64 field
__device__ __forceinline__ double old_value(double staged, const double *) { return staged; }
__device__ __forceinline__ double old_value(float, const double *p) { return *p; }

template <class T>
__global__ void __launch_bounds__(1024) __attribute__((amdgpu_waves_per_eu(8))) k_hsu_vert(const MemberDesc *D, int mb, int nm, int C) {
  ADV_MEMBER;
  if (d.adv_akima[which]) return;  // the field takes the Akima scheme (uniform per block)
  extern __shared__ double lds_raw[];
  T *lds = reinterpret_cast<T *>(lds_raw);
  const int c = threadIdx.x % C, kk = threadIdx.x / C, KS = blockDim.x / C;
  const int nI = d.ni + 2, col = blockIdx.x * C + c;
  const bool live = col < (d.nj + 1) * nI;
  const int J = live ? col / nI : 0, I = live ? col % nI : 0;
  const int nk = d.nk, L = nk + 2;
  const int layer = (which == AV_HDRY);
  const int kf = (which == AV_THETA) ? d.ks - 1 : 1;
  const T ep = T(1.e-10), DT = T(d.dt);
  const double *dsi = d.dsgth_inv;
  double *b = adv_field(d, which);
  T *wa = lds + c * L, *wv = lds + (C + c) * L, *pr = lds + (2 * C + c) * L, *ff = lds + (3 * C + c) * L;
  // stage 1: field and vertical velocity (wv holds w on the points the
  // field lives on: W3 itself for layers, the interface average otherwise)
  if (live) {
    for (int K = kk; K <= nk + 1; K += KS) {
      double x = 0., w = 0.;
      if (layer) {
        if (K >= kf && K <= nk) x = b[ix3(d, K, J, I)];
        if (K >= kf - 1 && K <= nk) w = d.W3[ix3(d, K, J, I)];
      } else {
        if (K >= kf - 1 && K <= nk) x = b[ix3(d, K, J, I)];
        if (K >= kf && K <= nk) w = (K == 1 || K == nk) ? 0. : .5 * (d.W3[ix3(d, K, J, I)] + d.W3[ix3(d, K - 1, J, I)]);
        // statics left by the unity pass (wm[kf-1], wa[kf-2] of its last column)
        if (which == AV_THETA && K == kf - 1 && K >= 1) {
          const int Jl = d.nj, Il = d.ni + 1;
          w = dmin(0., (K == 1 || K == nk) ? 0. : .5 * (d.W3[ix3(d, K, Jl, Il)] + d.W3[ix3(d, K - 1, Jl, Il)]));
        }
        if (which == AV_THETA && K == kf - 2) x = 1.;
      }
      wa[K] = T(x);
      wv[K] = T(w);
    }
  }
  __syncthreads();
#define WM(k) tmin(T(0.), wv[k])
#define WP(k) tmax(T(0.), wv[k])
  if (layer) {
    // first-guess flux on interfaces kf..nk-1 (zero at kf-1 and nk)
    if (live)
      for (int K = kf - 1 + kk; K <= nk; K += KS)
        ff[K] = (K == kf - 1 || K == nk) ? T(0.) : WP(K) * wa[K + 1] + WM(K) * wa[K];
    __syncthreads();
    if (live)
      for (int K = kf + kk; K <= nk; K += KS) pr[K] = wa[K] + (ff[K] - ff[K - 1]) * T(dsi[2 * K]) * DT;
    __syncthreads();
    if (live)
      for (int K = kf + kk; K < nk; K += KS)
        ff[K] = T(.5) * (WP(K) * (pr[K] + wa[K + 1]) + WM(K) * (pr[K + 1] + wa[K]));
    __syncthreads();
    if (live)
      for (int K = kf + kk; K < nk; K += KS) {
        const T d1K = pr[K] - wa[K + 1];
        const T d1P = (K + 1 == nk) ? T(0.) : pr[K + 1] - wa[K + 2];
        const T d2K = (K == 1) ? T(0.) : pr[K] - wa[K - 1];
        const T d2P = pr[K + 1] - wa[K];
        const T aaK = tt_layer(wa, K, nk, ep), aaP = tt_layer(wa, K + 1, nk, ep);
        const T aminK = (K >= kf) ? T(d.h_min[K]) : T(0.), aminP = T(d.h_min[K + 1]);
        const T wpK = WP(K), wmK = WM(K);
        T whp = sqrt(wpK * WP(K + 1)), whm = -sqrt(wmK * WM(K - 1));
        T mu = (wpK - wmK) * T(dsi[2 * K + 1]) * DT, al = (T(1.) + mu) / T(6.);
        T gp = aaP / (aaP + (wa[K + 1] - aminP) * (wa[K] - aminK));
        gp = gp * gp;
        T gm = aaK / (aaK + (wa[K + 1] - aminP) * (wa[K] - aminK));
        gm = gm * gm;
        T bp = T(1.) + (T(1.) / (T(2.) * al) - T(1.)) * gp, bm = T(1.) + (T(1.) / (T(2.) * al) - T(1.)) * gm;
        T bhp = T(1.) - gp, bhm = T(1.) - gm;
        ff[K] += -al * (wpK * bp * d1K - whp * bhp * d1P + wmK * bm * d2P - whm * bhm * d2K);
      }
    __syncthreads();
    if (live)
      for (int K = kf + kk; K <= nk; K += KS)
        b[ix3(d, K, J, I)] = old_value(wa[K], b + ix3(d, K, J, I)) + (double)((ff[K] - ff[K - 1]) * T(dsi[2 * K]) * DT);
  } else {
    // first-guess flux on layers kf..nk (zero below the bottom)
    if (live)
      for (int K = kf + kk; K <= nk + 1; K += KS) ff[K] = (K == nk + 1) ? T(0.) : WP(K) * wa[K] + WM(K) * wa[K - 1];
    __syncthreads();
    if (live)
      for (int K = kf - 1 + kk; K <= nk; K += KS)
        pr[K] = (K == kf - 1) ? wa[K] : wa[K] + (ff[K + 1] - ff[K]) * T(dsi[2 * K + 1]) * DT;
    __syncthreads();
    if (live)
      for (int K = kf + kk; K <= nk; K += KS)
        ff[K] = T(.5) * (WP(K) * (pr[K - 1] + wa[K]) + WM(K) * (pr[K] + wa[K - 1]));
    __syncthreads();
    if (live)
      for (int K = kf + kk; K <= nk; K += KS) {
        const T d1K = pr[K - 1] - wa[K];
        const T d2K = pr[K] - wa[K - 1];
        const T d2M = (K == kf) ? T(0.) : pr[K - 1] - wa[K - 2];
        const T aaK = tt_iface(wa, K, nk, ep), aaM = tt_iface(wa, K - 1, nk, ep);
        const T wpK = WP(K), wmK = WM(K);
        T whm = (K == 1) ? T(0.) : -sqrt(wmK * WM(K - 1));
        T whp = (K == nk) ? T(0.) : sqrt(wpK * WP(K + 1));
        T mu = (wpK - wmK) * T(dsi[2 * K]) * DT, al = (T(1.) + mu) / T(6.);
        T gp = aaK / (aaK + (wa[K] - T(0.)) * (wa[K - 1] - T(0.)));
        gp = gp * gp;
        T gm = aaM / (aaM + (wa[K] - T(0.)) * (wa[K - 1] - T(0.)));
        gm = gm * gm;
        T bp = T(1.) + (T(1.) / (T(2.) * al) - T(1.)) * gp, bm = T(1.) + (T(1.) / (T(2.) * al) - T(1.)) * gm;
        T bhp = T(1.) - gp, bhm = T(1.) - gm, g;
        if (K < nk) {
          const T d1P = pr[K] - wa[K + 1];
          g = -al * (wpK * bp * d1K - whp * bhp * d1P + wmK * bm * d2K - whm * bhm * d2M);
        } else {
          g = -al * (wpK * bp * d1K + wmK * bm * d2K - whm * bhm * d2M);
        }
        ff[K] += g;
      }
    __syncthreads();
    if (live)
      for (int K = kf + kk; K <= nk; K += KS)
        b[ix3(d, K, J, I)] = old_value(wa[K], b + ix3(d, K, J, I)) + (double)((ff[K + 1] - ff[K]) * T(dsi[2 * K + 1]) * DT);
  }
#undef WM
#undef WP
}

// ---------------------------------------------------------- meridional
// One launch: a block holds a tile of TJ rows x CW columns of one plane of
// one field in LDS, with halos: the field on rows J0-2..J0+TJ+1, the
// meridional velocities vp, vm (v times 1/m, split by sign) on v-rows
// J0-1..J0+TJ+1, the smoothness terms aq and the predictor ap on rows
// J0-1..J0+TJ and the corrected flux on v-rows J0..J0+TJ, each formed once per
// cell by the reference's expression (the halo rows are formed again by the
// neighbouring tile: the same values).  The update goes to sA and to its copy
// sB (which the zonal pass filters), with sB's zonal pads.
template <class T>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(8))) k_hsu_mer(const MemberDesc *D, int mb, int nm, int CW, int MER_TJ) {
  ADV_MEMBER;
  const int K = 1 + blockIdx.y;
  if (d.adv_akima[which] || (which == AV_THETA && K <= d.ks - 2)) return;  // uniform per block
  extern __shared__ double sh_raw[];
  T *sh = reinterpret_cast<T *>(sh_raw);
  const int nj = d.nj, nchunk = d.ni / CW;
  const int tile = blockIdx.x / nchunk, chunk = blockIdx.x - tile * nchunk;
  const int J0 = tile * MER_TJ, c = threadIdx.x % CW, r0 = threadIdx.x / CW, RS = blockDim.x / CW;
  const int I = 1 + chunk * CW + c;
  const int RA = MER_TJ + 4, RV = MER_TJ + 3, RP = MER_TJ + 2;
  T *ta = sh, *tvp = ta + RA * CW, *tvm = tvp + RV * CW, *taq = tvm + RV * CW, *tap = taq + RP * CW;
  T *tff = tap + RP * CW;
  const size_t o = (size_t)K * d.N2;
  const double *A = adv_field(d, which) + o;
  const T amin = (which == AV_HDRY) ? T(d.h_min[K]) : T(0.), DT = T(d.dt), ep = T(1.e-10);
  // row index helpers: field rows from J0-2, v-rows and predictor rows from J0-1
  auto a = [&](int J) { return ta[(J - J0 + 2) * CW + c]; };
  auto vp = [&](int J) { return tvp[(J - J0 + 1) * CW + c]; };
  auto vm = [&](int J) { return tvm[(J - J0 + 1) * CW + c]; };
  auto aq = [&](int J) { return taq[(J - J0 + 1) * CW + c]; };
  auto ap = [&](int J) { return tap[(J - J0 + 1) * CW + c]; };
  // the field (rows outside -1..nj+1 are never read) and the velocities
  for (int r = r0; r < RA; r += RS) {
    const int J = J0 - 2 + r;
    ta[r * CW + c] = (J >= -1 && J <= nj + 1) ? T(A[ix2(d, J, I)]) : T(0.);
  }
  for (int r = r0; r < RV; r += RS) {
    const int J = J0 - 1 + r;
    if (J < 0 || J > nj + 1) continue;
    double v2;
    if (which == AV_HDRY) v2 = (J < 1 || J > nj) ? 0. : .5 * (d.V[ix3(d, K, J, I)] + d.V[ix3(d, K - 1, J, I)]);
    else v2 = d.V[ix3(d, K, J, I)];
    const double mi = d.minv_v[J];
    tvm[r * CW + c] = T(dmin(0., v2 * mi));
    tvp[r * CW + c] = T(dmax(0., v2 * mi));
  }
  __syncthreads();
  // smoothness terms and predictor on rows J0-1 .. J0+MER_TJ within 0..nj
  for (int r = r0; r < RP; r += RS) {
    const int J = J0 - 1 + r;
    if (J < 0 || J > nj) continue;
    T t;
    if (J == 0) t = fabs(-a(0) + a(1)) + ep;
    else if (J == nj) t = fabs(a(J - 1) - a(J)) + ep;
    else t = fabs(a(J - 1) - T(2.) * a(J) + a(J + 1)) + ep;
    taq[r * CW + c] = t * t;
    const T f0 = (J < 1) ? T(0.) : vp(J) * a(J - 1) + vm(J) * a(J);
    const T f1 = (J + 1 > nj) ? T(0.) : vp(J + 1) * a(J) + vm(J + 1) * a(J + 1);
    tap[r * CW + c] = a(J) + (f0 - f1) * T(d.mn[2 * J + 1]) * DT;
  }
  __syncthreads();
  // corrected flux on v-rows J0 .. J0+MER_TJ within 1..nj
  for (int r = r0; r <= MER_TJ; r += RS) {
    const int J = J0 + r;
    if (J < 1 || J > nj) continue;
    T VP = vp(J), VM = vm(J), APj = ap(J), APm = ap(J - 1), Am = a(J - 1), Aj = a(J);
    T fc = T(.5) * (VP * (APj + Am) + VM * (APm + Aj));
    T vhp = sqrt(VP * vp(J - 1)), vhm = -sqrt(VM * vm(J + 1));
    T mu = (VP - VM) * T(d.mn[2 * J]) * DT, al = (T(1.) + mu) / T(6.);
    T q0 = aq(J - 1), q1 = aq(J);
    T gp = q0 / (q0 + (Am - amin) * (Aj - amin));
    gp = gp * gp;
    T gm = q1 / (q1 + (Am - amin) * (Aj - amin));
    gm = gm * gm;
    T bp = T(1.) + (T(1.) / (T(2.) * al) - T(1.)) * gp, bm = T(1.) + (T(1.) / (T(2.) * al) - T(1.)) * gm;
    T bhp = T(1.) - gp, bhm = T(1.) - gm;
    T d1J = (APj - Am);
    T d1m = (J - 1 == 0) ? T(0.) : APm - a(J - 2);
    T d2m = APm - a(J);
    T d2J = (J == nj) ? T(0.) : APj - a(J + 1);
    tff[r * CW + c] = fc + -al * (VP * bp * d1J - vhp * bhp * d1m + VM * bm * d2m - vhm * bhm * d2J);
  }
  __syncthreads();
  // update on the tile's rows: the FP64 field plus the flux divergence; sB
  // (the copy the zonal pass filters) in the flux type
  double *sA = d.sA + so + o;
  T *sB = reinterpret_cast<T *>(d.sB) + so + o;
  for (int r = r0; r < MER_TJ; r += RS) {
    const int J = J0 + r;
    if (J > nj) break;
    const T f0 = (J >= 1) ? tff[r * CW + c] : T(0.), f1 = (J + 1 <= nj) ? tff[(r + 1) * CW + c] : T(0.);
    const double v = old_value(a(J), A + ix2(d, J, I)) + (double)((f0 - f1) * T(d.mn[2 * J + 1]) * DT);
    sA[ix2(d, J, I)] = v;
    sB[ix2(d, J, I)] = T(v);
    if (I == d.ni) sB[ix2(d, J, 0)] = T(v);
    if (I == 1) sB[ix2(d, J, d.ni + 1)] = T(v);
  }
}

// --------------------------------------------------------------- zonal
// Rows of the zonally filtered copy sB (1024/ni rows per block), staged in
// LDS with their pads: velocities, first-guess fluxes, predictor, corrected
// fluxes and the update sA + d(flux) each run on all lanes of the row,
// periodic in I.
template <class T>
__global__ void __launch_bounds__(256) k_hsu_zonal(const MemberDesc *D, int mb, int nm) {
  ADV_MEMBER;
  const RowMap m = row_map(d.ni);
  const int K = 1 + blockIdx.y, Jr = blockIdx.x * m.rpb + m.g;
  if (d.adv_akima[which] || (which == AV_THETA && K <= d.ks - 2)) return;  // uniform per block
  const bool live = Jr <= d.nj;
  const int J = live ? Jr : 0;
  extern __shared__ double sh_raw[];
  T *sh = reinterpret_cast<T *>(sh_raw);
  const int ni = d.ni, W = ni + 2;
  // three row arrays: the field, the velocity u/n (split by sign where it is
  // used; the corrected flux overwrites it once every lane has read it) and
  // the predictor; the first-guess flux is formed where the predictor needs
  // it (twice per cell, by the same expression)
  T *A = sh + (size_t)m.g * 3 * W;
  T *UV = A + W, *AP = A + 2 * W, *FF = UV;
  const T *af = reinterpret_cast<const T *>(d.sB) + so + (size_t)K * d.N2;
  const double ninv = d.ninv_h[J];
  const T mn = T(d.mn[2 * J + 1]), DT = T(d.dt);
  const T amin = (which == AV_HDRY) ? T(d.h_min[K]) : T(0.);
  auto wr = [ni](int I) { return I < 1 ? I + ni : (I > ni ? I - ni : I); };
  if (live)
    for (int I = m.lane; I < W; I += m.stride) {
      A[I] = af[ix2(d, J, I)];
      double u2;
      if (which == AV_HDRY) {
        int Ii = wr(I);
        u2 = .5 * (d.U[ix3(d, K, J, Ii)] + d.U[ix3(d, K - 1, J, Ii)]);
      } else {
        u2 = d.U[ix3(d, K, J, I)];
      }
      UV[I] = T(u2 * ninv);
    }
  __syncthreads();
  // T(min(0, x)) = min(0, T(x)): rounding is monotone and keeps zero
  auto UP = [UV](int I) { return tmax(T(0.), UV[I]); };
  auto UM = [UV](int I) { return tmin(T(0.), UV[I]); };
  auto FP = [&](int c) { return UP(c) * A[c - 1] + UM(c) * A[c]; };
  if (live)
    for (int c = 1 + m.lane; c <= ni; c += m.stride) AP[c] = A[c] + (FP(c) - FP(wr(c + 1))) * mn * DT;
  __syncthreads();
  // a lane's four corrected fluxes (c = 1 + lane + q*stride) in registers
  T ffr[4];
  if (live)
#pragma unroll
    for (int q = 0; q < 4; q++) {
      const int c = 1 + m.lane + q * m.stride;
      const int cm = wr(c - 1);
      T UPc = UP(c), UMc = UM(c), APc = AP[c], APm = AP[cm], Am = A[c - 1], Ac = A[c];
      T fc = T(.5) * (UPc * (APc + Am) + UMc * (APm + Ac));
      T d1c = APc - Am, d2c = APc - A[c + 1];
      T d1m = APm - A[cm - 1], d2m = APm - A[cm + 1];
      T uhp = sqrt(UPc * UP(c - 1)), uhm = -sqrt(UMc * UM(c + 1));
      T mu = (UPc - UMc) * mn * DT, al = (T(1.) + mu) / T(6.);
      T t0 = fabs(A[cm - 1] - T(2.) * A[cm] + A[cm + 1]) + T(1.e-10), q0 = t0 * t0;
      T t1 = fabs(A[c - 1] - T(2.) * A[c] + A[c + 1]) + T(1.e-10), q1 = t1 * t1;
      T gp = q0 / (q0 + (Am - amin) * (Ac - amin));
      gp = gp * gp;
      T gm = q1 / (q1 + (Am - amin) * (Ac - amin));
      gm = gm * gm;
      T bp = T(1.) + (T(1.) / (T(2.) * al) - T(1.)) * gp, bm = T(1.) + (T(1.) / (T(2.) * al) - T(1.)) * gm;
      T bhp = T(1.) - gp, bhm = T(1.) - gm;
      ffr[q] = fc + -al * (UPc * bp * d1c - uhp * bhp * d1m + UMc * bm * d2m - uhm * bhm * d2c);
    }
  __syncthreads();
  if (live)
#pragma unroll
    for (int q = 0; q < 4; q++) FF[1 + m.lane + q * m.stride] = ffr[q];
  __syncthreads();
  if (!live) return;
  double *out = adv_field(d, which);
  for (int I = 1 + m.lane; I <= ni; I += m.stride) {
    const size_t q = ix3(d, K, J, I);
    const double y = d.sA[so + q] + (double)((FF[I] - FF[wr(I + 1)]) * mn * DT);
    out[q] = y;
    if (I == ni) out[ix3(d, K, J, 0)] = y;
    if (I == 1) out[ix3(d, K, J, ni + 1)] = y;
  }
}

// ------------------------------------------------------------- Akima
// akima_advection (src/shared/epic_flux.c:1502-1924), the monotonized
// fourth-order centred scheme, in the same three passes and tiles as the
// predictor-corrector above: the difference dA of neighbouring values, the
// harmonic-mean slope dela = 2 dA- dA+ / (dA- + dA+) where both have one
// sign (0 otherwise), and the flux: the velocity times the edge value
// .5 (a- + a+) - (dela+ - dela-) / 6, each by the reference's expression.
// A block whose field takes the other scheme returns at once (uniform per
// block); the two schemes share the filtered copy sB and its filter launch.
// The reference also runs the column pass on the pad rows J = -1 and nj+1,
// where W3 is zero and the pass changes nothing.
template <class T>
__device__ __forceinline__ T ak_dela(T d0, T d1) {
  const T tmp = d0 * d1;
  return (tmp > T(0.)) ? T(2.) * tmp / (d0 + d1) : T(0.);
}

// column pass: columns J = 0..nj, I = 0..ni+1 in LDS tiles; the stages
// (differences, slopes, fluxes, update) run on all lanes over (column,
// level) pairs, separated by barriers (no recurrence in K)
template <class T>
__global__ void __launch_bounds__(1024) __attribute__((amdgpu_waves_per_eu(8))) k_ak_vert(const MemberDesc *D, int mb, int nm, int C) {
  ADV_MEMBER;
  if (!d.adv_akima[which]) return;  // uniform per block
  extern __shared__ double lds_raw[];
  T *lds = reinterpret_cast<T *>(lds_raw);
  const int c = threadIdx.x % C, kk = threadIdx.x / C, KS = blockDim.x / C;
  const int nI = d.ni + 2, col = blockIdx.x * C + c;
  const bool live = col < (d.nj + 1) * nI;
  const int J = live ? col / nI : 0, I = live ? col % nI : 0;
  const int nk = d.nk, L = nk + 2;
  const int layer = (which == AV_HDRY);
  const int kf = (which == AV_THETA) ? d.ks - 1 : 1;
  const T DT = T(d.dt);
  const double *dsi = d.dsgth_inv;
  double *b = adv_field(d, which);
  // the flux reuses the difference column once the slopes are formed
  T *wa = lds + c * L, *wv = lds + (C + c) * L, *da = lds + (2 * C + c) * L, *dl = lds + (3 * C + c) * L, *ff = da;
  if (live)
    for (int K = kk; K <= nk + 1; K += KS) {
      double x = 0., w = 0.;
      if (layer) {
        if (K >= kf && K <= nk) x = b[ix3(d, K, J, I)];
        if (K >= kf - 1 && K <= nk) w = d.W3[ix3(d, K, J, I)];
      } else {
        if (K >= kf - 1 && K <= nk) x = b[ix3(d, K, J, I)];
        if (K >= kf && K <= nk) w = (K == 1 || K == nk) ? 0. : .5 * (d.W3[ix3(d, K, J, I)] + d.W3[ix3(d, K - 1, J, I)]);
      }
      wa[K] = T(x);
      wv[K] = T(w);
    }
  __syncthreads();
  // differences: in a layer, da[K] = a[K] - a[K+1] on kf..nk-1 (0 at kf-1
  // and nk); on interfaces, da[K] = a[K-1] - a[K] on kf..nk
  if (live)
    for (int K = kf - 1 + kk; K <= nk; K += KS) {
      if (layer) da[K] = (K == kf - 1 || K == nk) ? T(0.) : wa[K] - wa[K + 1];
      else if (K >= kf) da[K] = wa[K - 1] - wa[K];
    }
  __syncthreads();
  if (live)
    for (int K = kf - 1 + kk; K <= nk; K += KS) {
      if (layer) {
        if (K >= kf) dl[K] = ak_dela(da[K], da[K - 1]);
      } else {
        dl[K] = (K == kf - 1 || K == nk) ? T(0.) : ak_dela(da[K], da[K + 1]);
      }
    }
  __syncthreads();
  // fluxes (mass_flux_top and mass_flux_bot are zero)
  if (live)
    for (int K = kf - 1 + kk; K <= nk + 1; K += KS) {
      if (layer) {
        if (K <= nk)
          ff[K] = (K == kf - 1 || K == nk) ? T(0.) : wv[K] * (T(.5) * (wa[K] + wa[K + 1]) - (dl[K] - dl[K + 1]) / T(6.));
      } else if (K >= kf) {
        ff[K] = (K == nk + 1) ? T(0.) : wv[K] * (T(.5) * (wa[K - 1] + wa[K]) - (dl[K - 1] - dl[K]) / T(6.));
      }
    }
  __syncthreads();
  if (live)
    for (int K = kf + kk; K <= nk; K += KS) {
      const size_t q = ix3(d, K, J, I);
      if (layer) b[q] = old_value(wa[K], b + q) + (double)((ff[K] - ff[K - 1]) * T(dsi[2 * K]) * DT);
      else b[q] = old_value(wa[K], b + q) + (double)((ff[K + 1] - ff[K]) * T(dsi[2 * K + 1]) * DT);
    }
}

// meridional pass: a tile of TJ rows x CW columns of one plane, with the
// field on rows J0-2..J0+TJ+1, the slopes on J0-1..J0+TJ, the velocity and
// the fluxes on v-rows J0..J0+TJ; the update goes to sA and its copy sB
// (with sB's zonal pads), as the predictor-corrector's does
template <class T>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(8))) k_ak_mer(const MemberDesc *D, int mb, int nm, int CW, int MER_TJ) {
  ADV_MEMBER;
  const int K = 1 + blockIdx.y;
  if (!d.adv_akima[which] || (which == AV_THETA && K <= d.ks - 2)) return;  // uniform per block
  extern __shared__ double sh_raw[];
  T *sh = reinterpret_cast<T *>(sh_raw);
  const int nj = d.nj, nchunk = d.ni / CW;
  const int tile = blockIdx.x / nchunk, chunk = blockIdx.x - tile * nchunk;
  const int J0 = tile * MER_TJ, c = threadIdx.x % CW, r0 = threadIdx.x / CW, RS = blockDim.x / CW;
  const int I = 1 + chunk * CW + c;
  const int RA = MER_TJ + 4, RV = MER_TJ + 1, RD = MER_TJ + 2;
  T *ta = sh, *tv = ta + RA * CW, *tdl = tv + RV * CW, *tff = tdl + RD * CW;
  const size_t o = (size_t)K * d.N2;
  const double *A = adv_field(d, which) + o;
  const T DT = T(d.dt);
  auto a = [&](int J) { return ta[(J - J0 + 2) * CW + c]; };
  auto dl = [&](int J) { return tdl[(J - J0 + 1) * CW + c]; };
  for (int r = r0; r < RA; r += RS) {
    const int J = J0 - 2 + r;
    ta[r * CW + c] = (J >= -1 && J <= nj + 1) ? T(A[ix2(d, J, I)]) : T(0.);
  }
  for (int r = r0; r < RV; r += RS) {
    const int J = J0 + r;
    double v2 = 0.;
    if (J >= 1 && J <= nj)
      v2 = (which == AV_HDRY) ? .5 * (d.V[ix3(d, K, J, I)] + d.V[ix3(d, K - 1, J, I)]) : d.V[ix3(d, K, J, I)];
    tv[r * CW + c] = T(v2);
  }
  __syncthreads();
  // slopes on rows 0..nj (zero elsewhere; the differences are zero outside
  // the rows 1..nj)
  auto dA = [&](int J) { return (J >= 1 && J <= nj) ? a(J) - a(J - 1) : T(0.); };
  for (int r = r0; r < RD; r += RS) {
    const int J = J0 - 1 + r;
    tdl[r * CW + c] = (J >= 0 && J <= nj) ? ak_dela(dA(J), dA(J + 1)) : T(0.);
  }
  __syncthreads();
  for (int r = r0; r < RV; r += RS) {
    const int J = J0 + r;
    T f = T(0.);
    if (J >= 1 && J <= nj) {
      const double mi = d.minv_v[J];
      f = tv[r * CW + c] * (T(.5) * (a(J) + a(J - 1)) - (dl(J) - dl(J - 1)) / T(6.)) * T(mi);
    }
    tff[r * CW + c] = f;
  }
  __syncthreads();
  double *sA = d.sA + so + o;
  T *sB = reinterpret_cast<T *>(d.sB) + so + o;
  for (int r = r0; r < MER_TJ; r += RS) {
    const int J = J0 + r;
    if (J > nj) break;
    const double v =
        old_value(a(J), A + ix2(d, J, I)) + (double)((tff[r * CW + c] - tff[(r + 1) * CW + c]) * T(d.mn[2 * J + 1]) * DT);
    sA[ix2(d, J, I)] = v;
    sB[ix2(d, J, I)] = T(v);
    if (I == d.ni) sB[ix2(d, J, 0)] = T(v);
    if (I == 1) sB[ix2(d, J, d.ni + 1)] = T(v);
  }
}

// zonal pass: rows of the filtered copy sB with its pads in LDS; slopes,
// fluxes and the update sA + d(flux) on all lanes of the row, periodic in I
template <class T>
__global__ void __launch_bounds__(256) k_ak_zonal(const MemberDesc *D, int mb, int nm) {
  ADV_MEMBER;
  const RowMap m = row_map(d.ni);
  const int K = 1 + blockIdx.y, Jr = blockIdx.x * m.rpb + m.g;
  if (!d.adv_akima[which] || (which == AV_THETA && K <= d.ks - 2)) return;  // uniform per block
  const bool live = Jr <= d.nj;
  const int J = live ? Jr : 0;
  extern __shared__ double sh_raw[];
  T *sh = reinterpret_cast<T *>(sh_raw);
  const int ni = d.ni, W = ni + 2;
  T *A = sh + (size_t)m.g * 3 * W, *DL = A + W, *FF = A + 2 * W;
  const T *af = reinterpret_cast<const T *>(d.sB) + so + (size_t)K * d.N2;
  const double ninv = d.ninv_h[J];
  const T mn = T(d.mn[2 * J + 1]), DT = T(d.dt);
  auto wr = [ni](int I) { return I < 1 ? I + ni : (I > ni ? I - ni : I); };
  if (live)
    for (int I = m.lane; I < W; I += m.stride) A[I] = af[ix2(d, J, I)];
  __syncthreads();
  // differences A(I) - A(I-1) on I = 1..ni, the pad of I = ni+1 that of I = 1
  auto dA = [&](int I) { return A[I] - A[I - 1]; };
  if (live)
    for (int I = 1 + m.lane; I <= ni; I += m.stride) DL[I] = ak_dela(dA(I), dA(wr(I + 1)));
  __syncthreads();
  if (live)
    for (int I = 1 + m.lane; I <= ni; I += m.stride) {
      const double u2 =
          (which == AV_HDRY) ? .5 * (d.U[ix3(d, K, J, I)] + d.U[ix3(d, K - 1, J, I)]) : d.U[ix3(d, K, J, I)];
      FF[I] = T(u2) * (T(.5) * (A[I] + A[I - 1]) - (DL[I] - DL[wr(I - 1)]) / T(6.)) * T(ninv);
    }
  __syncthreads();
  if (!live) return;
  double *out = adv_field(d, which);
  for (int I = 1 + m.lane; I <= ni; I += m.stride) {
    const size_t q = ix3(d, K, J, I);
    const double y = d.sA[so + q] + (double)((FF[I] - FF[wr(I + 1)]) * mn * DT);
    out[q] = y;
    if (I == ni) out[ix3(d, K, J, 0)] = y;
    if (I == 1) out[ix3(d, K, J, ni + 1)] = y;
  }
}

// one advection pass over all three fields of every member; the mixed mode
// forms the fluxes in FP32 (sB, the filtered copy, FP32 too).  Each scheme's
// launches run where some field of the bucket takes it.
template <class T>
static void hsu3_t(hipStream_t s, const Bucket &b) {
  const int ncol = (b.nj + 1) * (b.ni + 2), nz = ADV_FIELDS * b.nm;
  // columns per block: as many as keep the four LDS columns within 40 KB
  int C = 64;
  while (C > 16 && (size_t)4 * C * (b.nk + 2) * sizeof(double) > 40 * 1024) C /= 2;
  const int v = b.knob[KN_COLS_HSUV];
  if (v) C = knob_cols(v);
  const bool hsu = b.adv_schemes & 1, akima = b.adv_schemes & 2;
  const dim3 gv((ncol + C - 1) / C, 1, nz);
  if (hsu)
    LAUNCH(k_hsu_vert<T>, gv, dim3(knob_threads(v, b.nk + 2)), sizeof(T) * 4 * C * (b.nk + 2), s, b.D, b.mb, b.nm, C);
  if (akima) LAUNCH(k_ak_vert<T>, gv, dim3(knob_threads(v, b.nk + 2)), sizeof(T) * 4 * C * (b.nk + 2), s, b.D, b.mb, b.nm, C);
  const int TJ = b.knob[KN_ROWS_MER] ? b.knob[KN_ROWS_MER] : 4;
  const int CW = b.ni < 64 ? b.ni : 64, tiles = (b.nj + 1 + TJ - 1) / TJ;
  const dim3 gm(tiles * (b.ni / CW), b.nk, nz);
  if (hsu)
    LAUNCH(k_hsu_mer<T>, gm, dim3(256), sizeof(T) * CW * ((TJ + 4) + 2 * (TJ + 3) + 2 * (TJ + 2) + (TJ + 1)), s, b.D,
           b.mb, b.nm, CW, TJ);
  if (akima)
    LAUNCH(k_ak_mer<T>, gm, dim3(256), sizeof(T) * CW * ((TJ + 4) + 2 * (TJ + 1) + (TJ + 2)), s, b.D, b.mb, b.nm, CW,
           TJ);
  // the zonal pass reads a zonally filtered copy (the reference filters it
  // per level before the zonal sweep); a separate launch packs the rows of
  // the filter band, which is faster here than filtering inside the pass
  if (sizeof(T) == sizeof(double)) L_filter(s, b, F_SB, 1, b.nk, 0, 0, ADV_FIELDS);
  else L_filter_f32(s, b, F_SB, 1, b.nk, 0, 0, ADV_FIELDS);
  const int rpb = rows_per_block(b.ni);
  const dim3 gz((b.nj + 1 + rpb - 1) / rpb, b.nk, nz);
  if (hsu) LAUNCH(k_hsu_zonal<T>, gz, dim3(256), sizeof(T) * rpb * 3 * (b.ni + 2), s, b.D, b.mb, b.nm);
  if (akima) LAUNCH(k_ak_zonal<T>, gz, dim3(256), sizeof(T) * rpb * 3 * (b.ni + 2), s, b.D, b.mb, b.nm);
}
static void hsu3(hipStream_t s, const Bucket &b) {
  if (b.mixed) hsu3_t<float>(s, b);
  else hsu3_t<double>(s, b);
}

// op 0: unity = 1 on K = 0..nk+1, J = 0..nj, I = 0..ni+1, and OLD = THETA
// on K = k_sigma-1..nk+1 over the same rows; op 1: the theta correction
// there.  Both also apply restore_mass (HDRY = max(h_min, HDRY) on K = 1..nk
// over the same rows), which the reference runs right before advection
// (src/shared/epic_timestep.c) and right after it (src/shared/epic_flux.c)
// and which touches HDRY only.
__global__ void __launch_bounds__(256) k_adv_elem(const MemberDesc *D, int mb, int op) {
  MEMBER;
  int t = blockIdx.x * TPB + threadIdx.x;
  const int nI = d.ni + 2;
  if (t >= (d.nj + 1) * nI) return;
  const int J = t / nI, I = t % nI, K = blockIdx.y;
  if (K > d.nk + 1) return;
  size_t q = ix3(d, K, J, I);
  if (op == 0) {
    d.unity[q] = 1.;
    if (K >= d.ks - 1) d.old[q] = d.THETA[q];
  } else if (K >= d.ks - 1) {
    d.THETA[q] -= d.old[q] * (d.unity[q] - 1.);
  }
  if (K >= 1 && K <= d.nk) d.HDRY[q] = dmax(d.h_min[K], d.HDRY[q]);
}

// restore_mass, advection, restore_mass (the two restore_mass calls run
// inside the element-wise launches around the advection passes)
void L_advection(hipStream_t s, const Bucket &b) {
  int ncol = (b.nj + 1) * (b.ni + 2);
  dim3 ge((ncol + TPB - 1) / TPB, b.nk + 2, b.nm);
  // unity, theta and hdry advected together; the theta correction after
  // them (hdry's advection reads neither theta nor unity)
  LAUNCH(k_adv_elem, ge, dim3(TPB), 0, s, b.D, b.mb, 0);
  hsu3(s, b);
  LAUNCH(k_adv_elem, ge, dim3(TPB), 0, s, b.D, b.mb, 1);
  // set_p2etc(DONT_UPDATE_THETA) follows, with store_pgrad_vars
}

}  // namespace epic
