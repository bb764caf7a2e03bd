/*
 * cpuref.h -- CPU FP64 restatement of EPIC's dynamical-core timestep.
 * TEST INFRASTRUCTURE ONLY: the oracle of the parity tests.  Only tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it, and
 * only as the checker or as the timed CPU baseline; the product never links
 * or calls it.
 */
#ifndef EPIC_CPUREF_H
#define EPIC_CPUREF_H
#include "../include/epic_amd.h"

typedef struct cpuref cpuref;

/* Build the case from the spec (epic_case_build) and take ownership. */
cpuref *cpuref_create(const epic_case_spec *spec);
/* the state of an epic.nc file (epic_nc_read), Adams-Bashforth history included */
cpuref *cpuref_create_nc(const char *path);
/* the traditional pressure gradient alone into a cleared plane IT_ZERO */
void cpuref_pgrad_traditional(cpuref *R);
void    cpuref_destroy(cpuref *r);
/* timestep(SYNC_DIAGS_ONLY) and timestep(STEP_PROGS_AND_SYNC_DIAGS) */
int     cpuref_sync(cpuref *r);
int     cpuref_step(cpuref *r);
/* Field copy-out in epic_field order; tendencies by time plane. */
int     cpuref_read(cpuref *r, int field, double *out);
/* Overwrite a 3-D field (not GZ_SURFACE, not a tendency plane) and form
 * RI2 again from the current U, V, P2, T2, RHO2, P3, RHO3: pins a
 * diagnostic of the GPU to the oracle's function of the same inputs. */
int     cpuref_write(cpuref *r, int field, const double *in);
void    cpuref_richardson(cpuref *r);
/* Current IT_ZERO, IT_MINUS1, IT_MINUS2 and the number of steps done. */
void    cpuref_time_planes(cpuref *r, int *it3, long *steps);
/* Non-zero once a fatal condition of the reference was met
 * (non-monotone GZ3, p out of range, non-monotone spline table). */
int     cpuref_error(cpuref *r);
/* Extents: ni, nj, nk, Nelem3d, Nelem2d. */
void    cpuref_dims(cpuref *r, int *dims5);

/* Stand-alone kernels exposed for unit tests of the restatement. */
/* NR-packed real FFT of length n (power of two), in place, 1-based data. */
void    cpuref_realft(double *data_1based, unsigned long n, int isign);
/* PCHIP derivative setup on (x, y) -> d, n points; returns 0 or -1. */
int     cpuref_pchip(int n, const double *x, const double *y, double *d);
double  cpuref_pchip_eval(int n, const double *x, const double *y, const double *d, double xx);
int     cpuref_fcmp(double a, double b);
/* exp (fn 0), log (1), pow (2) as this build evaluates them (the host libm,
 * or the correctly rounded crmath.h in libcpuref_cr.so), and that build's
 * count of calls and of glibc results that differ from crmath.h's. */
double  cpuref_math(int fn, double x, double y);
void    cpuref_libm_stats(long *calls3, long *diffs3, int reset);
/* libcpuref_cr.so: which of exp (bit 0), log (1), pow (2) take crmath.h's
 * result (default 7, all three); the others take the host libm's. */
void    cpuref_cr_select(int mask);
#endif
