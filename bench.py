#!/usr/bin/env python3
"""Throughput of the EPIC dynamical-core step on one MI355X.

Workload (default): the Jupiter-like globe ni=256 nj=128 nk=20, dt=60 s,
ideal gas, eight seeded zonal jets plus 0.5 m/s noise, del^4
hyperviscosity, polar filter on.  A "step" is one
timestep(STEP_PROGS_AND_SYNC_DIAGS) of every member, with the state resident
in HBM.  The unit is grid-point-steps/s with N_pts = ni*(nj+1)*nk per member.

Prints ONE JSON line.  `roofline` is for the kernel that takes the most
device time, from HIP events around every launch of a profiled pass:
achieved = its algorithmic bytes per launch / its mean launch duration.
`cpu_baseline` is the FP64 CPU restatement of the step (oracle/) timed on
this host, one thread, on a bounded number of steps of the same workload.
"""
import argparse
import json
import math
import os
import sys
import time

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)

PEAK_HBM = 8.0e12  # MI355X HBM3E, B/s

WORKLOADS = {
    # name: (planet, ni, nj, nk, dt, members)
    "jupiter_256x128x20": (1, 256, 128, 20, 60, 1),
    "held_suarez_64x32x10": (0, 64, 32, 10, 600, 1),
    "ensemble_64x128x64x16": (0, 128, 64, 16, 600, 64),
    "hot_jupiter_1024x512x64": (1, 1024, 512, 64, 30, 1),
}

# Algorithmic traffic of each kernel, in full passes over one member's
# interior cells (8 bytes each): arrays read plus arrays written.
KERNEL_PASSES = {
    "k_hsu_vert": 3, "k_hsu_merid": 4, "k_hsu_zonal": 4, "k_pgrad": 7, "k_set_p2etc": 7,
    "k_thermo": 10, "k_gz3": 5, "k_calc_w": 7, "k_zonal_hv": 2, "k_d2dy2": 2, "k_hv_apply": 3,
    "k_uvh_kin": 6, "k_dvdt_core": 8, "k_dudt_core": 8, "k_spline_sub": 4, "k_spline_add": 3,
    "k_vort_interior": 4, "k_divergence": 3, "k_copy_planes": 2, "k_ab": 10, "k_dzdt3": 8,
    "k_restore_mass": 2, "k_gz_mono": 2, "k_zero_tend": 2, "k_fill": 1, "k_adv_elem": 2,
    "k_heating": 3, "k_sponge": 3, "k_hs_pbl": 4,
}


def filtered_fraction(nj):
    """Fraction of h-rows with |lat| >= 45 deg on EPIC's globe grid."""
    dlt = 180.0 / ((nj + 1) + 2 * math.sqrt(0.5))
    lat0 = -90.0 + dlt * math.sqrt(0.5)
    rows = 0
    for J in range(nj + 1):
        lat = lat0 + (2 * J + 1) * dlt * 0.5
        rows += abs(lat) >= 45.0
    return rows / (nj + 1)


def kernel_bytes(name, ni, nj, nk, members):
    cells = ni * (nj + 1) * nk * members
    if name == "k_fft_filter":
        return 2 * 8 * cells * filtered_fraction(nj)
    return KERNEL_PASSES.get(name, 0) * 8 * cells


def cpu_baseline(planet, ni, nj, nk, dt, budget_s):
    from oracle.cpuref import CpuRef
    import epic_amd as ea
    sp = ea.spec(planet, ni, nj, nk, dt, seed=1, nu4=0.5,
                 noise_theta=0.1 if planet in (0, 3) else 0.0, noise_uv=0.0 if planet == 0 else 0.5)
    r = CpuRef(sp)
    r.sync()
    r.step(1)
    n, t0 = 0, time.perf_counter()
    while True:
        r.step(1)
        n += 1
        el = time.perf_counter() - t0
        if el > budget_s or n >= 200:
            break
    r.close()
    return ni * (nj + 1) * nk * n / el, n, el


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--workload", default="jupiter_256x128x20", choices=sorted(WORKLOADS))
    ap.add_argument("--precision", default="fp64", choices=["fp64", "mixed"])
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--no-cpu", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one MI355X; --gpus {args.gpus} is not supported here "
              "(members shard across GPUs with no exchange; see DESIGN.md)", file=sys.stderr)
        sys.exit(2)

    import epic_amd as ea
    planet, ni, nj, nk, dt, members = WORKLOADS[args.workload]
    prec = ea.FP64 if args.precision == "fp64" else ea.MIXED
    eng = ea.Engine(prec, members)
    for m in range(members):
        eng.add(ea.spec(planet, ni, nj, nk, dt, seed=1 + m, nu4=0.5,
                        noise_theta=0.1 if planet in (0, 3) else 0.0,
                        noise_uv=0.0 if planet == 0 else 0.5))
    eng.step(args.warmup)
    eng.sync()
    t0 = time.perf_counter()
    eng.step(args.steps)
    eng.sync()
    el = time.perf_counter() - t0
    npts = ni * (nj + 1) * nk * members
    value = npts * args.steps / el

    prof = eng.profile(3)
    tot = sum(ms for _, ms, _ in prof)
    name, ms, cnt = max(prof, key=lambda x: x[1])
    avg_s = ms / cnt / 1e3
    nbytes = kernel_bytes(name, ni, nj, nk, members)
    achieved = nbytes / avg_s if avg_s > 0 else 0.0
    step_bytes = 112 if prec == ea.FP64 else 88

    out = {
        "metric": "grid-pt-steps/s",
        "value": value,
        "unit": "grid-pt-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": el / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if prec == ea.FP64 else "f32+f64",
        "data": "synthetic (seeded initial state of the workload's planet)",
        "config": {"workload": args.workload, "ni": ni, "nj": nj, "nk": nk, "members": members, "dt": dt,
                   "precision": args.precision, "n_pts": npts},
        "step_hbm_fraction": value * step_bytes / PEAK_HBM,
        "roofline": {"bound": "hbm", "kernel": name, "achieved": achieved / 1e9, "peak": PEAK_HBM / 1e9,
                     "unit": "GB/s", "frac": achieved / PEAK_HBM, "traffic": None,
                     "launch_us": avg_s * 1e6, "bytes_per_launch": nbytes,
                     "share_of_step": ms / tot if tot else None},
    }
    eng.close()
    if not args.no_cpu:
        v, n, cel = cpu_baseline(planet, ni, nj, nk, dt, args.cpu_seconds)
        out["cpu_baseline"] = {"value": v, "unit": "grid-pt-steps/s", "cores": 1, "kind": "port",
                               "sample": f"{n} FP64 steps of one {ni}x{nj}x{nk} member in {cel:.1f} s"}
    print(json.dumps(out))


if __name__ == "__main__":
    main()
