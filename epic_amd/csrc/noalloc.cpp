// noalloc.cpp -- the engine's FP64 step of one member for callers that
// cannot allocate device memory (code under HIP graph capture, serving
// loops, harnesses that forbid hipMalloc).  The caller holds every field of
// the member (the epic_field arrays) in its own device memory; the library's
// working memory (descriptor, time planes, tables, u_spinup, the filtered
// geopotential copy, advection and hyperviscosity scratch, spline tables)
// lives in the fixed static device array `scratch` below, of
// EPIC_NOALLOC_SCRATCH_BYTES bytes per GPU, which each GPU the library runs
// on holds until the process ends.  No call makes a device allocation.
//
// The same kernels as epic_engine_add / epic_engine_step run, in the same
// order, on the same kind of descriptor, so the results equal the engine's
// bit for bit.  Every call clears the working memory first (with a fill
// kernel) and then writes its tables: a call keeps nothing from the one
// before.  Calls that overlap on one GPU share its scratch array and race.
#include <stdio.h>
#include <string.h>

#include <mutex>
#include <vector>

#include "engine_internal.h"

using namespace epic;

namespace {
constexpr int kMaxDevices = 64;
}

static __device__ __align__(256) unsigned char scratch[EPIC_NOALLOC_SCRATCH_BYTES];

namespace {

// the device address of `scratch` on each GPU, looked up once per device
// with that device current
std::mutex g_mu;
unsigned char *g_scratch[kMaxDevices] = {nullptr};

int scratch_of_current_device(unsigned char **out, char *err, size_t cap) {
  int dev = -1;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kMaxDevices) {
    snprintf(err, cap, "no current HIP device");
    return -5;
  }
  std::lock_guard<std::mutex> lk(g_mu);
  if (!g_scratch[dev]) {
    void *p = nullptr;
    hipError_t rc = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (rc != hipSuccess || !p) {
      snprintf(err, cap, "hipGetSymbolAddress(scratch) on device %d: %s", dev, hipGetErrorString(rc));
      return -5;
    }
    g_scratch[dev] = (unsigned char *)p;
  }
  *out = g_scratch[dev];
  return 0;
}

// zero n bytes (a multiple of 16) from p; a kernel rather than
// hipMemsetAsync, so a caller may capture the call in a graph
__global__ void __launch_bounds__(256) k_clear(uint4 *p, size_t n16) {
  const uint4 z = make_uint4(0u, 0u, 0u, 0u);
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n16; i += (size_t)gridDim.x * blockDim.x)
    p[i] = z;
}

size_t align256(size_t x) { return (x + 255) & ~(size_t)255; }

}  // namespace

#define NACHK(x)                                                              \
  do {                                                                        \
    hipError_t e_ = (x);                                                      \
    if (e_ != hipSuccess) {                                                   \
      snprintf(err, cap, "%s: %s", #x, hipGetErrorString(e_));                \
      return -5;                                                              \
    }                                                                         \
  } while (0)

extern "C" int epic_step_noalloc(const epic_case *c, double *const *fields, int *it4, int action, int nsteps,
                                 void *stream, char *errbuf, size_t errcap) {
  char local[256];
  char *err = (errbuf && errcap) ? errbuf : local;
  const size_t cap = (errbuf && errcap) ? errcap : sizeof(local);
  err[0] = 0;
  if (!c || !fields || !it4) {
    snprintf(err, cap, "null case, field table or time planes");
    return -1;
  }
  if (action != EPIC_STEP_PROGS_AND_SYNC_DIAGS && action != EPIC_SYNC_DIAGS_ONLY) {
    snprintf(err, cap, "unrecognized action=%d", action);
    return -1;
  }
  if (action == EPIC_STEP_PROGS_AND_SYNC_DIAGS && nsteps < 1) {
    snprintf(err, cap, "nsteps=%d: a stepping call needs at least one step", nsteps);
    return -1;
  }
  if (member_check(c, EPIC_FP64, err, cap)) return -1;
  const size_t N3 = (size_t)c->Nelem3d;
  for (int f = 0; f < EPIC_F_COUNT; f++)
    if (!fields[f]) {
      snprintf(err, cap, "field %d has no device array", f);
      return -1;
    }
  // the kernels address the time planes as plane * Nelem3d from plane 0
  for (int p = 1; p < 3; p++)
    if (fields[EPIC_F_DUDT0 + p] != fields[EPIC_F_DUDT0] + p * N3 ||
        fields[EPIC_F_DVDT0 + p] != fields[EPIC_F_DVDT0] + p * N3) {
      snprintf(err, cap, "the three u (and v) tendency planes must be consecutive Nelem3d-element planes");
      return -1;
    }
  unsigned char *base = nullptr;
  int rc = scratch_of_current_device(&base, err, cap);
  if (rc) return rc;
  hipStream_t s = (hipStream_t)stream;

  epic_tables T;
  epic_tables_build(c, &T);
  struct TablesGuard {
    epic_tables *t;
    ~TablesGuard() { epic_tables_free(t); }
  } tg{&T};
  MemberDesc d;
  std::vector<MemberPart> parts;
  MemberHold hold;
  member_plan(c, false, T, d, parts, hold);
  // working memory: the descriptor first, then every part that is not one of
  // the caller's fields
  size_t off = align256(sizeof(MemberDesc));
  for (auto &p : parts) {
    if (p.field >= 0) {
      *p.dst = fields[p.field];
    } else {
      *p.dst = base + off;
      off += align256(p.bytes);
    }
  }
  if (off > (size_t)EPIC_NOALLOC_SCRATCH_BYTES) {
    snprintf(err, cap, "a %dx%dx%d member needs %zu bytes of working memory; the static buffer holds %zu",
             c->ni, c->nj, c->nk, off, (size_t)EPIC_NOALLOC_SCRATCH_BYTES);
    return -2;
  }
  // clear, then tables and u_spinup, the time planes and the descriptor
  hipLaunchKernelGGL(k_clear, dim3(1024), dim3(256), 0, s, (uint4 *)base, off / 16);
  NACHK(hipGetLastError());
  for (auto &p : parts)
    if (p.field < 0 && p.src) NACHK(hipMemcpyAsync(*p.dst, p.src, p.bytes, hipMemcpyHostToDevice, s));
  MemberDyn dy;
  memset(&dy, 0, sizeof(dy));
  dy.it0 = it4[0];
  dy.itm1 = it4[1];
  dy.itm2 = it4[2];
  dy.itime = it4[3];
  NACHK(hipMemcpyAsync(d.dyn, &dy, sizeof(dy), hipMemcpyHostToDevice, s));
  NACHK(hipMemcpyAsync(base, &d, sizeof(d), hipMemcpyHostToDevice, s));

  Bucket b;
  memset(&b, 0, sizeof(b));
  b.D = (const MemberDesc *)base;
  b.mb = 0;
  b.nm = 1;
  b.ni = c->ni;
  b.nj = c->nj;
  b.nk = c->nk;
  b.mixed = 0;
  b.any_gas = d.is_gas;
  b.any_hs = d.is_hs;
  b.any_ray = (c->newt_cool_on && c->prandtl > 0.) || (c->tau_drag > 0. && c->tau_drag < 1.e+20);
  b.hv_passes = hv_passes(d);
  b.adv_schemes = adv_schemes(d);
  const bool hv = b.hv_passes > 0, divd = T.nudiv != 0.;
  const bool adjust = (c->newt_cool_adjust || c->k_sponge >= 1) && c->newt_cool_on;
  if (action == EPIC_SYNC_DIAGS_ONLY) {
    member_record(s, b, true, hv, divd, adjust);
  } else {
    for (int n = 0; n < nsteps; n++) member_record(s, b, false, hv, divd, adjust);
  }
  // RI2 of the final state (the engine forms it when it is read)
  L_richardson(s, b);
  NACHK(hipGetLastError());
  // the host tables are read by the copies above: wait before they go
  NACHK(hipStreamSynchronize(s));
  NACHK(hipMemcpy(&dy, d.dyn, sizeof(dy), hipMemcpyDeviceToHost));
  it4[0] = dy.it0;
  it4[1] = dy.itm1;
  it4[2] = dy.itm2;
  it4[3] = dy.itime;
  if (dy.err) {
    snprintf(err, cap, "fatal condition %d (non-monotone GZ3=1, p range=2, spline=4)", dy.err);
    return -4;
  }
  return 0;
}
