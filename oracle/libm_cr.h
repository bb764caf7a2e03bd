/* libm_cr.h -- TEST INFRASTRUCTURE.  Included by cpuref.c only when it is
 * compiled with -DCPUREF_CRMATH (oracle/libcpuref_cr.so).
 *
 * Routes the oracle's exp(), log() and pow() through the correctly rounded
 * functions of epic_amd/csrc/crmath.h, the same source the GPU's FP64
 * kernels evaluate them with (calls written as (exp)(...) etc. keep the host
 * libm: those values the GPU side also forms on the host).  Every other
 * operation of the oracle is unchanged, so this copy of the oracle
 * attributes GPU-vs-oracle differences: the GPU must equal it bit for bit,
 * and its own distance from the plain oracle is the whole effect of the
 * host libm's last-bit rounding.  crmath.h itself is pinned against
 * mpmath's correctly rounded values in tests/test_crmath.py.  Each call
 * also counts whether glibc's result differs from the correctly rounded
 * one (cpuref_libm_stats). */
#include <math.h>

#include "../epic_amd/csrc/crmath.h"

static long cpuref_cr_calls[3], cpuref_cr_diffs[3];
/* which of exp (bit 0), log (bit 1), pow (bit 2) return crmath.h's value
 * (the others glibc's): attributes the step's sensitivity to each function */
static int cpuref_cr_mask = 7;

static inline double cpuref_cr_exp(double x) {
  double a = crm_exp(x), b = (exp)(x);
  cpuref_cr_calls[0]++;
  cpuref_cr_diffs[0] += (a != b);
  return (cpuref_cr_mask & 1) ? a : b;
}
static inline double cpuref_cr_log(double x) {
  double a = crm_log(x), b = (log)(x);
  cpuref_cr_calls[1]++;
  cpuref_cr_diffs[1] += (a != b);
  return (cpuref_cr_mask & 2) ? a : b;
}
static inline double cpuref_cr_pow(double x, double y) {
  double a = crm_pow(x, y), b = (pow)(x, y);
  cpuref_cr_calls[2]++;
  cpuref_cr_diffs[2] += (a != b);
  return (cpuref_cr_mask & 4) ? a : b;
}
#define exp(x) cpuref_cr_exp(x)
#define log(x) cpuref_cr_log(x)
#define pow(x, y) cpuref_cr_pow(x, y)
