/* libm_nudge.h -- TEST INFRASTRUCTURE.  Included by cpuref.c only when it is
 * compiled with -DCPUREF_LIBM_NUDGE (oracle/libcpuref_nudge.so).
 *
 * Moves a deterministic pseudo-random 10% of the results of exp(), log()
 * and pow() by one ulp.  The GPU evaluates these three functions with the
 * device math library, whose results may differ from the host libm's in the
 * last bit; every other operation of the step follows the reference's order
 * exactly.  Running the oracle against this perturbed copy of itself
 * measures how far the reference algorithm carries last-bit differences of
 * its transcendentals (the finite-volume pressure gradient amplifies them
 * by 1e4-1e5 in near-balanced states): the sensitivity floor below which a
 * GPU-vs-oracle difference carries no information. */
#include <math.h>
#include <stdint.h>
#include <string.h>

static inline double cpuref_nudge(double y, double key) {
  uint64_t b;
  memcpy(&b, &key, 8);
  b ^= b >> 29;
  b *= 0xbf58476d1ce4e5b9ULL;
  b ^= b >> 32;
  if ((double)(b >> 11) * 0x1.0p-53 < 0.1) return nextafter(y, (b & 1) ? INFINITY : -INFINITY);
  return y;
}
static inline double cpuref_exp(double x) { return cpuref_nudge(exp(x), x); }
static inline double cpuref_log(double x) { return cpuref_nudge(log(x), x); }
static inline double cpuref_pow(double x, double y) { return cpuref_nudge(pow(x, y), x * 1.3 + y); }
#define exp(x) cpuref_exp(x)
#define log(x) cpuref_log(x)
#define pow(x, y) cpuref_pow(x, y)
