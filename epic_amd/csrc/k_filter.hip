// k_filter.hip -- the polar zonal filter (src/shared/epic_adjust.c:102-480).
//
// Each filtered latitude row (|lat| >= 45 deg) is transformed in LDS: the
// real FFT of length ni is the half-length complex radix-2 transform plus the
// split step, with the Numerical Recipes packing (data[1] = DC, data[2] =
// Nyquist; realft/four1, src/shared/epic_funcs_util.c:4023-4135).  The
// butterflies run in the same order and with the same twiddle values as the
// reference: the twiddles of the reference's trigonometric recurrences are
// data independent, so the host runs the recurrences once and uploads the
// values; each stage is then data parallel (one lane per butterfly) and the
// result is bit-identical to the sequential transform.  ni/4 lanes serve a
// row; a 256-lane block holds 1024/ni rows.
//
// The 3-D form removes a PCHIP profile of the variable against the filtered
// geopotential before filtering and adds it back afterwards, on every row
// (src/shared/epic_adjust.c:258-397, 463-474), fused with the transform.
#include "kernels.h"
#include "filter_row.h"

namespace epic {

#define MEMBER const MemberDesc &d = D[mb + blockIdx.z]

template <class T>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(8))) k_fft_filter(const MemberDesc *D, int mb, int f, int k0, int kn, int vgrid,
                                                    int pads) {
  MEMBER;
  extern __shared__ double lds_raw[];
  T *lds = reinterpret_cast<T *>(lds_raw);
  const int n = d.ni, nq = n >> 2, RB = TPB / nq;
  const int rs = threadIdx.x / nq, lane = threadIdx.x - rs * nq;
  const int nrows = vgrid ? d.nrow_v : d.nrow_h;
  const int r = blockIdx.x * RB + rs;
  const bool active = r < nrows;
  const int J = active ? (vgrid ? d.rows_v[r] : d.rows_h[r]) : 0;
  // planes k0..k0+kn-1 of each of gridDim.y/kn contiguous copies of the array
  const int copy = blockIdx.y / kn, K = k0 + blockIdx.y - copy * kn;
  T *a = fld_ptr_t<T>(d, f) + (size_t)copy * d.N3 + (size_t)K * d.N2;
  T *buf = lds + (size_t)rs * n;
  const int bits = d.nlogn - 1;
  const double *hl = (vgrid ? d.hl_pv : d.hl_h) + (size_t)J * (n / 2 + 2);
  LaneTw<T> tf, ti;
  load_tw(tf, tw_fwd<T>(d), tr_fwd<T>(d), n, lane);
  load_tw(ti, tw_inv<T>(d), tr_inv<T>(d), n, lane);
  T gain[4];
  if (active)
    for (int q = 0, e = lane; q < 4; q++, e += nq) {
      int gi = (e == 0) ? 1 : (e == 1 ? n / 2 + 1 : (e >> 1) + 1);
      gain[q] = T(hl[gi]);
      int k = e >> 1, c = e & 1;
      buf[(brev(k, bits) << 1) | c] = a[ix2(d, J, e + 1)];
    }
  row_sync(nq);
  four1_stages(buf, n, lane, tf, active);
  split_step(buf, n, lane, tf, T(-0.5), active);
  if (active && lane == 0) {
    T h1r;
    buf[0] = (h1r = buf[0]) + buf[1];
    buf[1] = h1r - buf[1];
  }
  row_sync(nq);
  if (active)
    for (int q = 0, e = lane; q < 4; q++, e += nq) buf[e] *= gain[q];
  row_sync(nq);
  split_step(buf, n, lane, ti, T(0.5), active);
  if (active && lane == 0) {
    T h1r;
    buf[0] = T(0.5) * ((h1r = buf[0]) + buf[1]);
    buf[1] = T(0.5) * (h1r - buf[1]);
  }
  row_sync(nq);
  T v[4];
  if (active)
    for (int q = 0, e = lane; q < 4; q++, e += nq) v[q] = buf[e];
  row_sync(nq);
  if (active)
    for (int q = 0, e = lane; q < 4; q++, e += nq) {
      int k = e >> 1, c = e & 1;
      buf[(brev(k, bits) << 1) | c] = v[q];
    }
  row_sync(nq);
  four1_stages(buf, n, lane, ti, active);
  if (active)
    for (int e = lane; e < n; e += nq) {
      const T y = buf[e] * T(2. / (n));
      a[ix2(d, J, e + 1)] = y;
      if (pads && e == n - 1) a[ix2(d, J, 0)] = y;
      if (pads && e == 0) a[ix2(d, J, n + 1)] = y;
    }
}

template <class T>
static void launch_fft(hipStream_t s, const Bucket &b, int f, int k0, int k1, int vgrid, int pads, int copies) {
  const int n = b.ni, nq = n / 4, RB = TPB / nq;
  const int rows = b.nj + 1;  // upper bound on the listed rows
  dim3 g((rows + RB - 1) / RB, (k1 - k0 + 1) * copies, b.nm);
  size_t shm = sizeof(T) * (size_t)RB * n;
  LAUNCH(k_fft_filter<T>, g, dim3(TPB), shm, s, b.D, b.mb, f, k0, k1 - k0 + 1, vgrid, pads);
}

// mode 0: also refresh the zonal pads of the filtered rows (the producers of
// every filtered copy write the pads of the rows they fill); mode 1: no pads
void L_filter(hipStream_t s, const Bucket &b, int f, int k0, int k1, int vgrid, int mode, int copies) {
  launch_fft<double>(s, b, f, k0, k1, vgrid, mode == 0, copies);
}
// the same on an FP32 array (mixed precision)
void L_filter_f32(hipStream_t s, const Bucket &b, int f, int k0, int k1, int vgrid, int mode, int copies) {
  launch_fft<float>(s, b, f, k0, k1, vgrid, mode == 0, copies);
}

// ------------------------------------------------------------- 3-D form
__device__ __forceinline__ double geez(const MemberDesc &d, int kind, int K, int J, int I) {
  const double *gz = d.gz;
  switch (kind) {
    case 0: return .5 * (gz[ix3(d, K, J, I)] + gz[ix3(d, K, J, I - 1)]);
    case 1: return .5 * (gz[ix3(d, K, J, I)] + gz[ix3(d, K, J - 1, I)]);
    case 2: return .5 * (gz[ix3(d, K, J, I)] + gz[ix3(d, K - 1, J, I)]);
    default: return gz[ix3(d, K, J, I)];
  }
}

__device__ __forceinline__ double sgn_prod(double a, double b) {
  if (a == 0.) return 0.;
  if (a < 0.) return b < 0. ? 1. : (b == 0. ? 0. : -1.);
  return b < 0. ? -1. : (b == 0. ? 0. : 1.);
}

struct F3 { int kstart, kend, jstart; };
__device__ __forceinline__ F3 f3_range(const MemberDesc &d, int kind) {
  F3 r;
  r.jstart = (kind == 1) ? 1 : 0;
  r.kstart = (kind == 3) ? d.ks - 1 : 1;
  r.kend = d.nk;
  return r;
}

// PCHIP derivatives with one lane per interior point: the serial sweep of
// the reference carries (h1, h2, del1, del2) from point to point, but each
// of them is a function of three neighbouring table entries, so every lane
// forms its own with the same expressions.  Lanes 0 and 1 form the two
// shape-preserving end derivatives.  Returns nonzero (in every lane) when
// x is not strictly increasing.
__device__ int pchip_wave(int n, const double *x, const double *y, double *dd, int lane) {
  int bad = 0;
  for (int i = 1 + lane; i < n; i += 64)
    if (x[i] <= x[i - 1]) bad = 1;
  if (__any(bad)) return -1;
  if (n == 2) {
    if (lane == 0) {
      double del1 = (y[1] - y[0]) / (x[1] - x[0]);
      dd[0] = dd[1] = del1;
    }
    return 0;
  }
  for (int i = 2 + lane; i <= n - 1; i += 64) {
    double h1 = x[i - 1] - x[i - 2], h2 = x[i] - x[i - 1], hsum = h1 + h2;
    double del1 = (y[i - 1] - y[i - 2]) / h1, del2 = (y[i] - y[i - 1]) / h2;
    double v = 0.0;
    if (sgn_prod(del1, del2) > 0.0) {
      double hsumt3 = 3.0 * hsum;
      double a1 = (hsum + h1) / hsumt3, a2 = (hsum + h2) / hsumt3;
      double dmx = dmax(fabs(del1), fabs(del2)), dmn = dmin(fabs(del1), fabs(del2));
      double r1 = del1 / dmx, r2 = del2 / dmx;
      v = dmn / (a1 * r1 + a2 * r2);
    }
    dd[i - 1] = v;
  }
  if (lane == 0) {
    double h1 = x[1] - x[0], del1 = (y[1] - y[0]) / h1;
    double h2 = x[2] - x[1], del2 = (y[2] - y[1]) / h2;
    double hsum = h1 + h2, w1 = (h1 + hsum) / hsum, w2 = -h1 / hsum;
    double v = w1 * del1 + w2 * del2;
    if (sgn_prod(v, del1) <= 0.0) v = 0.0;
    else if (sgn_prod(del1, del2) < 0.0) {
      double dm = 3.0 * del1;
      if (fabs(dm) < fabs(v)) v = dm;
    }
    dd[0] = v;
  } else if (lane == 1) {
    const int i = n - 1;
    double h1 = x[i - 1] - x[i - 2], h2 = x[i] - x[i - 1], hsum = h1 + h2;
    double del1 = (y[i - 1] - y[i - 2]) / h1, del2 = (y[i] - y[i - 1]) / h2;
    double w1 = -h2 / hsum, w2 = (h2 + hsum) / hsum;
    double v = w1 * del1 + w2 * del2;
    if (sgn_prod(v, del2) <= 0.0) v = 0.0;
    else if (sgn_prod(del1, del2) < 0.0) {
      double dm = 3.0 * del2;
      if (fabs(dm) < fabs(v)) v = dm;
    }
    dd[n - 1] = v;
  }
  return 0;
}

// one wavefront per row: the column of lowest bottom GZ (the first of
// equals, as the reference's strict-less scan finds it: each lane keeps the
// first minimum of its strided subset, the lanes reduce to the lowest value
// and, among equals, the lowest I), the profile table at that column and its
// PCHIP derivatives
// field i of a set and its member
#define SET_MEMBER                                              \
  const int fi = (int)blockIdx.z / nm;                          \
  const MemberDesc &d = D[mb + (int)blockIdx.z - fi * nm];      \
  const int f = fs.f[fi], kind = fs.kind[fi]

// F: the mixed mode's exp/log (m_exp)
template <class T, bool F>
__global__ void __launch_bounds__(256) k_spline_prep(const MemberDesc *D, int mb, int nm, FieldSet fs) {
  SET_MEMBER;
  F3 r = f3_range(d, kind);
  const int J = r.jstart + blockIdx.x;
  if (J > d.nj) return;  // uniform per block
  const int nk = d.nk, KT = nk + 2, n = r.kend - r.kstart + 1, lane = threadIdx.x;
  __shared__ double sx[KMAX + 2], sy[KMAX + 2], sz[KMAX + 2];
  double best = 1.7976931348623157e308;
  int bi = 0x7fffffff;
  for (int I = 1 + lane; I <= d.ni; I += 64) {
    double gv = d.gz[ix3(d, nk, J, I)];
    if (gv < best) { best = gv; bi = I; }
  }
  for (int off = 32; off > 0; off >>= 1) {
    double ov = __shfl_xor(best, off, 64);
    int oi = __shfl_xor(bi, off, 64);
    if (oi != 0x7fffffff && (bi == 0x7fffffff || ov < best || (ov == best && oi < bi))) {
      best = ov;
      bi = oi;
    }
  }
  const int Imin = (bi == 0x7fffffff) ? 1 : bi;
  const T *b = fld_ptr_t<T>(d, f);
  for (int e = lane; e < n; e += 64) {
    const int K = r.kend - e;
    sx[e] = geez(d, kind, K, J, Imin);
    double v = (double)b[ix3(d, K, J, Imin)];
    sy[e] = (kind == 2) ? m_log<F>(v) : v;
  }
  wave_sync();
  if (pchip_wave(n, sx, sy, sz, lane) && lane == 0) atomicOr(&d.dyn->err, 4);
  wave_sync();
  double *tx = d.spl + (size_t)(fi * (d.nj + 2) + J) * 3 * KT, *ty = tx + KT, *tz = ty + KT;
  for (int e = lane; e < n; e += 64) {
    tx[e] = sx[e];
    ty[e] = sy[e];
    tz[e] = sz[e];
  }
  // the table's last slot (n <= nk < KT - 1) says whether the filter may
  // place its points by walking from a guess
  if (lane == 0) tx[KT - 1] = table_well_separated(n, sx) ? 1. : 0.;
}

// The rest of the 3-D filter in one launch: on every row J = jstart..nj of
// planes kstart..kend the PCHIP profile v(K,J,I) is removed, the rows of the
// filter band are transformed, gain-multiplied and transformed back in LDS,
// the profile is added back and the zonal pads are written.  Rows outside the
// band come out as (a - v) + v, as in the reference, which subtracts and adds
// the profile on every row.  HDRY is stored with the restore_mass that
// follows the filters in the step.  RB rows per block, ni/4 lanes per row.
template <class T, bool F>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(6))) k_filter3(const MemberDesc *D, int mb, int nm, FieldSet fs) {
  SET_MEMBER;
  extern __shared__ double lds_raw[];
  double *tabs = lds_raw;
  F3 r = f3_range(d, kind);
  const int K = r.kstart + blockIdx.y;
  if (K > r.kend) return;  // uniform per block
  const int n = d.ni, nq = n >> 2, RB = blockDim.x / nq;
  const int rs = threadIdx.x / nq, lane = threadIdx.x - rs * nq;
  const int J = r.jstart + blockIdx.x * RB + rs;
  const bool live = J <= d.nj;
  const int vg = (kind == 1);
  const bool filt = live && (vg ? d.rflag_v[J] : d.rflag_h[J]);
  const int KT = d.nk + 2, nt = r.kend - r.kstart + 1;
  double *tab = tabs + (size_t)rs * 5 * KT;
  T *buf = reinterpret_cast<T *>(tabs + (size_t)RB * 5 * KT) + (size_t)rs * n;
  if (live)
    for (int e = lane; e < 3 * KT; e += nq) tab[e] = d.spl[(size_t)(fi * (d.nj + 2) + J) * 3 * KT + e];
  row_sync(nq);
  T *a = fld_ptr_t<T>(d, f) + (size_t)K * d.N2;
  const double *tx = tab, *ty = tab + KT, *tz = tab + 2 * KT;
  // the cubic of each interval k depends on the table alone (splint_pchip,
  // src/shared/epic_funcs_util.c:926-999): its coefficients are formed once
  // per row, by the same expressions, instead of once per cell
  double *tc2 = tab + 3 * KT, *tc3 = tab + 4 * KT;
  const bool near = live && tab[KT - 1] != 0.;
  if (live)
    for (int k = lane; k < nt - 1; k += nq) {
      const double h = tx[k + 1] - tx[k];
      const double delta = (ty[k + 1] - ty[k]) / h;
      const double e1 = (tz[k] - delta) / h, e2 = (tz[k + 1] - delta) / h;
      tc3[k] = (e1 + e2) / h;
      tc2[k] = -(e1 + e1 + e2);
    }
  row_sync(nq);
  double v[4];
  T x[4];
  if (live)
    for (int q = 0; q < 4; q++) {
      const int I = lane + q * nq + 1;
      double xx = geez(d, kind, K, J, I), h;
      // entry e of the table is level kend - e of the column at Imin
      int k = near ? table_place_near(nt, tx, xx, r.kend - K, &h) : table_place(nt, tx, xx, &h);
      double xr = xx - tx[k];
      double val = ty[k] + xr * (tz[k] + xr * (tc2[k] + xr * tc3[k]));
      if (kind == 2) val = m_exp<F>(val);
      v[q] = val;
      x[q] = T((double)a[ix2(d, J, I)] - val);
    }
  LaneTw<T> tf, ti;
  load_tw(tf, tw_fwd<T>(d), tr_fwd<T>(d), n, lane);
  load_tw(ti, tw_inv<T>(d), tr_inv<T>(d), n, lane);
  const int bits = d.nlogn - 1;
  const double *hl = (vg ? d.hl_pv : d.hl_h) + (size_t)(live ? J : 0) * (n / 2 + 2);
  T gain[4];
  if (filt)
    for (int q = 0, e = lane; q < 4; q++, e += nq) {
      int gi = (e == 0) ? 1 : (e == 1 ? n / 2 + 1 : (e >> 1) + 1);
      gain[q] = T(hl[gi]);
      int kq = e >> 1, c = e & 1;
      buf[(brev(kq, bits) << 1) | c] = x[q];
    }
  row_sync(nq);
  four1_stages(buf, n, lane, tf, filt);
  split_step(buf, n, lane, tf, T(-0.5), filt);
  if (filt && lane == 0) {
    T h1r;
    buf[0] = (h1r = buf[0]) + buf[1];
    buf[1] = h1r - buf[1];
  }
  row_sync(nq);
  if (filt)
    for (int q = 0, e = lane; q < 4; q++, e += nq) buf[e] *= gain[q];
  row_sync(nq);
  split_step(buf, n, lane, ti, T(0.5), filt);
  if (filt && lane == 0) {
    T h1r;
    buf[0] = T(0.5) * ((h1r = buf[0]) + buf[1]);
    buf[1] = T(0.5) * (h1r - buf[1]);
  }
  row_sync(nq);
  T w[4];
  if (filt)
    for (int q = 0, e = lane; q < 4; q++, e += nq) w[q] = buf[e];
  row_sync(nq);
  if (filt)
    for (int q = 0, e = lane; q < 4; q++, e += nq) {
      int kq = e >> 1, c = e & 1;
      buf[(brev(kq, bits) << 1) | c] = w[q];
    }
  row_sync(nq);
  four1_stages(buf, n, lane, ti, filt);
  if (filt)
    for (int q = 0, e = lane; q < 4; q++, e += nq) x[q] = buf[e] * T(2. / (n));
  // hdry leaves with restore_mass applied (max(h_min, h), every cell of
  // K = 1..nk, J = 0..nj, I = 0..ni+1 is written here)
  const double hmin = (kind == 2) ? d.h_min[K] : 0.;
  if (live)
    for (int q = 0, e = lane; q < 4; q++, e += nq) {
      T out = T((double)x[q] + v[q]);
      if (kind == 2) out = T(dmax(hmin, (double)out));
      a[ix2(d, J, e + 1)] = out;
      if (e == n - 1) a[ix2(d, J, 0)] = out;
      if (e == 0) a[ix2(d, J, n + 1)] = out;
    }
}

// the rows of a kind: v rows 1..nj for V, 0..nj otherwise (the launches
// cover 0..nj and the kernels drop the rows outside the kind's range)
void L_filter3(hipStream_t s, const Bucket &b, const FieldSet &fs) {
  const int nrow = b.nj + 1, nz = fs.n * b.nm;
  if (b.mixed) LAUNCH_AS("k_spline_prep<double>", (k_spline_prep<double, true>), dim3(nrow, 1, nz), dim3(64), 0, s, b.D, b.mb, b.nm, fs);
  else LAUNCH_AS("k_spline_prep<double>", (k_spline_prep<double, false>), dim3(nrow, 1, nz), dim3(64), 0, s, b.D, b.mb, b.nm, fs);
  const int nq = b.ni / 4, RB = (256 / nq) < 16 ? (256 / nq) : 16;
  dim3 g((nrow + RB - 1) / RB, b.nk, nz);
  size_t shm = sizeof(double) * (size_t)RB * 5 * (b.nk + 2) + sizeof(double) * (size_t)RB * b.ni;
  if (b.mixed) LAUNCH_AS("k_filter3<double>", (k_filter3<double, true>), g, dim3(RB * nq), shm, s, b.D, b.mb, b.nm, fs);
  else LAUNCH_AS("k_filter3<double>", (k_filter3<double, false>), g, dim3(RB * nq), shm, s, b.D, b.mb, b.nm, fs);
}

}  // namespace epic
