// filter_row.h -- the polar-filter transform shared by the kernels that
// filter rows: the NR real FFT of src/shared/epic_funcs_util.c:4023-4135
// (four1 stages and the realft split step, with the twiddles of the
// reference's recurrences tabulated on the host), one butterfly per lane,
// and filter_row_lds(), which filters one latitude row held in LDS inside
// a consumer kernel (src/shared/epic_adjust.c:102-242, TWODIM).
#pragma once
#include "dev_common.h"

namespace epic {

__device__ __forceinline__ int brev(int k, int bits) { return bits ? (int)(__brev((unsigned)k) >> (32 - bits)) : 0; }

// Barrier between the transform stages of one row.  A row's ni/4 lanes lie
// in one wavefront when ni <= 256, and rows never exchange data, so a
// wavefront-level barrier suffices there; larger rows span waves.
__device__ __forceinline__ void row_sync(int nq) {
  if (nq <= 64) wave_sync();
  else __syncthreads();
}

// The stage twiddles stay in their table (global, cache-resident: the
// stage-s twiddle m is complex entry 2^s - 1 + m); the lane's split-step
// twiddle is held in registers.
constexpr int MAX_STAGES = 9;  // n <= 1024: stages with 2 << s < n

template <class T>
struct LaneTw {
  const T *tw;  // stage twiddles
  T sr, si;     // split-step twiddle of this lane
};

template <class T>
__device__ __forceinline__ void load_tw(LaneTw<T> &t, const T *tw, const T *tr, int n, int lane) {
  t.tw = tw;
  const bool on = lane < n / 4 - 1;
  t.sr = on ? tr[2 * lane] : T(0);
  t.si = on ? tr[2 * lane + 1] : T(0);
}

// one radix-2 butterfly of the reference's four1 on (xr, xi) and (yr, yi)
// with twiddle (wr, wi), in its operation order
template <class T>
__device__ __forceinline__ void bfly(T &xr, T &xi, T &yr, T &yi, T wr, T wi) {
  const T tr = wr * yr - wi * yi;
  const T ti = wr * yi + wi * yr;
  yr = xr - tr;
  yi = xi - ti;
  xr += tr;
  xi += ti;
}

// twiddle tables and field views of a storage type
template <class T> __device__ __forceinline__ const T *tw_fwd(const MemberDesc &d);
template <class T> __device__ __forceinline__ const T *tw_inv(const MemberDesc &d);
template <class T> __device__ __forceinline__ const T *tr_fwd(const MemberDesc &d);
template <class T> __device__ __forceinline__ const T *tr_inv(const MemberDesc &d);
template <> __device__ __forceinline__ const double *tw_fwd<double>(const MemberDesc &d) { return d.tw_f; }
template <> __device__ __forceinline__ const double *tw_inv<double>(const MemberDesc &d) { return d.tw_i; }
template <> __device__ __forceinline__ const double *tr_fwd<double>(const MemberDesc &d) { return d.tr_f; }
template <> __device__ __forceinline__ const double *tr_inv<double>(const MemberDesc &d) { return d.tr_i; }
template <> __device__ __forceinline__ const float *tw_fwd<float>(const MemberDesc &d) { return d.tw_f32; }
template <> __device__ __forceinline__ const float *tw_inv<float>(const MemberDesc &d) { return d.tw_i32; }
template <> __device__ __forceinline__ const float *tr_fwd<float>(const MemberDesc &d) { return d.tr_f32; }
template <> __device__ __forceinline__ const float *tr_inv<float>(const MemberDesc &d) { return d.tr_i32; }

// complex radix-2 stages on buf[0..n) (n doubles = n/2 complex), in the
// reference's stage order.  Stages go two at a time: a lane loads the four
// points p, p+h, p+2h, p+3h (h = 2^s) into registers, applies stage s to
// (p, p+h) and (p+2h, p+3h) and stage s+1 to (p, p+2h) and (p+h, p+3h), and
// stores them: every point sees the same butterflies in the same order as
// with one stage per pass, with half the LDS round trips and barriers.  n/8
// lanes take part in a pair of stages; an odd last stage runs one butterfly
// per lane on n/4 lanes.
template <class T>
__device__ __forceinline__ void four1_stages(T *buf, int n, int lane, const LaneTw<T> &t, bool active) {
  const T *tw = t.tw;
  int S = 0;
  while ((2 << S) < n) S++;  // stages s with 2 << s < n
  int s = 0;
  for (; s + 1 < S; s += 2) {
    const int half = 1 << s;
    if (active && lane < (n >> 3)) {
      const int mm = lane & (half - 1), g = lane >> s;
      const int i0 = 2 * (mm + g * 4 * half), i1 = i0 + 2 * half, i2 = i1 + 2 * half, i3 = i2 + 2 * half;
      T x0r = buf[i0], x0i = buf[i0 + 1], x1r = buf[i1], x1i = buf[i1 + 1];
      T x2r = buf[i2], x2i = buf[i2 + 1], x3r = buf[i3], x3i = buf[i3 + 1];
      const int ka = half - 1 + mm, kb = 2 * half - 1 + mm;
      const T ar = tw[2 * ka], ai = tw[2 * ka + 1];
      const T br = tw[2 * kb], bi = tw[2 * kb + 1], cr = tw[2 * (kb + half)], ci = tw[2 * (kb + half) + 1];
      bfly(x0r, x0i, x1r, x1i, ar, ai);
      bfly(x2r, x2i, x3r, x3i, ar, ai);
      bfly(x0r, x0i, x2r, x2i, br, bi);
      bfly(x1r, x1i, x3r, x3i, cr, ci);
      buf[i0] = x0r;
      buf[i0 + 1] = x0i;
      buf[i1] = x1r;
      buf[i1 + 1] = x1i;
      buf[i2] = x2r;
      buf[i2 + 1] = x2i;
      buf[i3] = x3r;
      buf[i3 + 1] = x3i;
    }
    row_sync(n >> 2);
  }
  if (s < S) {
    const int mmax = 2 << s, half = 1 << s;
    if (active) {
      const int g = lane >> s, mm = lane & (half - 1);
      const int i0 = 2 * mm + g * 2 * mmax, j0 = i0 + mmax, k = half - 1 + mm;
      T xr = buf[i0], xi = buf[i0 + 1], yr = buf[j0], yi = buf[j0 + 1];
      bfly(xr, xi, yr, yi, tw[2 * k], tw[2 * k + 1]);
      buf[i0] = xr;
      buf[i0 + 1] = xi;
      buf[j0] = yr;
      buf[j0 + 1] = yi;
    }
    row_sync(n >> 2);
  }
}

// split step of the real transform; c2 = -0.5 forward, +0.5 inverse
template <class T>
__device__ __forceinline__ void split_step(T *buf, int n, int lane, const LaneTw<T> &t, T c2, bool active) {
  const T c1 = T(0.5);
  if (active && lane < n / 4 - 1) {
    int i = lane + 2;
    int i1 = i + i - 1, i2 = 1 + i1, i3 = n + 3 - i2, i4 = 1 + i3;
    i1--; i2--; i3--; i4--;
    T wr = t.sr, wi = t.si;
    T h1r = c1 * (buf[i1] + buf[i3]);
    T h1i = c1 * (buf[i2] - buf[i4]);
    T h2r = -c2 * (buf[i2] + buf[i4]);
    T h2i = c2 * (buf[i1] - buf[i3]);
    buf[i1] = h1r + wr * h2r - wi * h2i;
    buf[i2] = h1i + wr * h2i + wi * h2r;
    buf[i3] = h1r - wr * h2r + wi * h2i;
    buf[i4] = -h1i + wr * h2i + wi * h2r;
  }
  row_sync(n >> 2);
}

// Rows per block of the row kernels: a row takes ni/4 lanes (the transform
// team: one butterfly per lane per stage), each lane holding every
// (ni/4)-th point of the row, and 1024/ni rows share a 256-lane block
// (one row when ni = 1024).
struct RowMap {
  int rpb;     // rows per block
  int g;       // this lane's row within the block
  int lane;    // this lane's index within its row
  int stride;  // lanes per row
};
__device__ __forceinline__ RowMap row_map(int ni) {
  RowMap m;
  m.stride = ni >> 2;
  m.rpb = 256 / m.stride;
  m.g = threadIdx.x / m.stride;
  m.lane = threadIdx.x - m.g * m.stride;
  return m;
}
static inline int rows_per_block(int ni) { return 1024 / ni; }

// Filter one latitude row held in LDS, in place: x[0..ni) holds I = 1..ni
// in natural order; work is ni values of scratch LDS; lane is the caller's
// index within its row (RowMap::lane), the row's ni/4 lanes being the
// transform team.  When ni <= 256 the team lies in one wavefront and the
// stages are separated by wavefront barriers; when ni > 256 by block
// barriers, so every lane of the block must call.  The caller places a
// block barrier after the call before other lanes read x.  filt = whether row J is in
// the filter band; the row is untouched otherwise.  The arithmetic is that
// of k_fft_filter, operation for operation.
template <class T>
__device__ __forceinline__ void filter_row_lds(const MemberDesc &d, T *x, T *work, int J, int vgrid, bool filt,
                                               int lane) {
  const int n = d.ni, nq = n >> 2;
  if (nq <= 64 && lane >= nq) return;
  const bool active = filt && lane < nq;
  const int bits = d.nlogn - 1;
  LaneTw<T> tf, ti;
  load_tw(tf, tw_fwd<T>(d), tr_fwd<T>(d), n, lane);
  load_tw(ti, tw_inv<T>(d), tr_inv<T>(d), n, lane);
  const double *hl = (vgrid ? d.hl_pv : d.hl_h) + (size_t)J * (n / 2 + 2);
  T gain[4];
  if (active)
    for (int q = 0, e = lane; q < 4; q++, e += nq) {
      int gi = (e == 0) ? 1 : (e == 1 ? n / 2 + 1 : (e >> 1) + 1);
      gain[q] = T(hl[gi]);
      int k = e >> 1, c = e & 1;
      work[(brev(k, bits) << 1) | c] = x[e];
    }
  row_sync(nq);
  four1_stages(work, n, lane, tf, active);
  split_step(work, n, lane, tf, T(-0.5), active);
  if (active && lane == 0) {
    T h1r;
    work[0] = (h1r = work[0]) + work[1];
    work[1] = h1r - work[1];
  }
  row_sync(nq);
  if (active)
    for (int q = 0, e = lane; q < 4; q++, e += nq) work[e] *= gain[q];
  row_sync(nq);
  split_step(work, n, lane, ti, T(0.5), active);
  if (active && lane == 0) {
    T h1r;
    work[0] = T(0.5) * ((h1r = work[0]) + work[1]);
    work[1] = T(0.5) * (h1r - work[1]);
  }
  row_sync(nq);
  T v[4];
  if (active)
    for (int q = 0, e = lane; q < 4; q++, e += nq) v[q] = work[e];
  row_sync(nq);
  if (active)
    for (int q = 0, e = lane; q < 4; q++, e += nq) {
      int k = e >> 1, c = e & 1;
      work[(brev(k, bits) << 1) | c] = v[q];
    }
  row_sync(nq);
  four1_stages(work, n, lane, ti, active);
  if (active)
    for (int e = lane; e < n; e += nq) x[e] = work[e] * T(2. / (n));
  row_sync(nq);
}

}  // namespace epic
