#!/usr/bin/env python3
"""Throughput of the EPIC dynamical-core step on one MI355X.

Workload (default): the Jupiter-like globe ni=256 nj=128 nk=20, dt=60 s,
ideal gas, eight seeded zonal jets plus 0.5 m/s u,v noise, del^4
hyperviscosity, polar filter on.  A "step" is one
timestep(STEP_PROGS_AND_SYNC_DIAGS) of every member, with the state resident
in HBM.  The unit is grid-point-steps/s with N_pts = ni*(nj+1)*nk per member.

Prints ONE JSON line.  `roofline` is for the kernel that takes the most
device time in a profiled pass (HIP events around every launch, on the
engine's stream): achieved = its algorithmic bytes per launch / its mean
launch duration, with the per-cell byte counts of KERNEL_BYTES.
`step_roofline` prices the whole step: N_pts*steps/s * 112 B (FP64; 88 B in
the mixed mode) against 8 TB/s.  `cpu_baseline` is the reference's own
timestep() built for one CPU (oracle/_ref/epic_ref, where it was built) or
else the oracle's FP64 CPU restatement, timed on this host, one thread, on a
bounded number of steps of the same workload.
"""
import argparse
import json
import math
import os
import subprocess
import sys
import time

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)

PEAK_HBM = 8.0e12  # MI355X HBM3E, B/s

WORKLOADS = {
    # name: (planet, ni, nj, nk, dt, members)
    "jupiter_256x128x20": (1, 256, 128, 20, 60, 1),
    "held_suarez_64x32x10": (0, 64, 32, 10, 600, 1),
    "hs_ensemble_64x128x64x16": (0, 128, 64, 16, 300, 64),
    "hot_jupiter_1024x512x64": (4, 1024, 512, 64, 30, 1),
}

# Algorithmic bytes per cell of each kernel: the distinct arrays it must
# read and write once (8 B each in FP64), over the cells of its launch
# (interior h-cells ni*(nj+1) per plane, times the planes it covers), and
# the number of fields one launch carries (the advection passes carry
# unity, theta and hdry; the 3-D filter U, V, hdry and theta; the
# hyperviscosity kernels two fields).
KERNEL_BYTES = {
    "k_set_p2etc": (56, 1), "k_thermo": (88, 1), "k_gz3": (40, 1), "k_calc_w": (56, 1), "k_dzdt3": (72, 1),
    "k_hsu_vert": (24, 3), "k_hsu_mer": (40, 3),
    "k_hsu_zonal": (32, 3), "k_fft_filter": (16, 3),
    "k_filter3": (24, 4), "k_zonal_hv": (16, 2), "k_d2dy2x2": (16, 2), "k_hv_apply": (24, 2),
    "k_pgrad": (48, 1), "k_uvh_kin": (48, 1), "k_dvdt_core": (40, 1), "k_dudt_core": (40, 1), "k_ab": (80, 1),
    "k_divdamp": (40, 1), "km_uvh_kin": (36, 1), "km_dvdt_core": (32, 1), "km_dudt_core": (32, 1), "km_ab": (64, 1),
    "km_pgrad": (48, 1), "km_divdamp": (40, 1), "k_gz_mono": (16, 1),
    "k_zero_tend": (16, 1), "k_adv_elem": (16, 1), "k_heating": (24, 1), "k_sponge": (24, 1), "k_hs_pbl": (32, 1),
    "k_divergence": (24, 1), "k_vort_interior": (32, 1),
}


def measured_traffic(workload, precision, name):
    """HBM bytes per launch of `name` from the committed counter summary of
    this workload (profiles/traffic_<workload>_<precision>.json: FETCH_SIZE
    and WRITE_SIZE per dispatch, each scaled by what the counter reads for a
    plain copy of known size), or (None, None) if none is committed."""
    fn = os.path.join(HERE, "profiles", f"traffic_{workload}_{precision}.json")
    if not os.path.exists(fn):
        return None, None
    k = json.load(open(fn))["kernels"]
    for cand in (name, name.replace("<T>", "<double>")):
        if cand in k:
            return k[cand]["traffic_bytes"], os.path.relpath(fn, HERE)
    # a kernel with more template arguments than its launch name shows
    base = name.split("<")[0]
    same = [c for c in k if c.split("<")[0] == base]
    if len(same) == 1:
        return k[same[0]]["traffic_bytes"], os.path.relpath(fn, HERE)
    return None, None


def filtered_rows(nj):
    """h-rows with |lat| >= 45 deg on EPIC's globe grid."""
    dlt = 180.0 / ((nj + 1) + 2 * math.sqrt(0.5))
    lat0 = -90.0 + dlt * math.sqrt(0.5)
    return sum(abs(lat0 + (2 * J + 1) * dlt * 0.5) >= 45.0 for J in range(nj + 1))


def kernel_bytes(name, ni, nj, nk, members):
    """Algorithmic bytes of one launch of `name` (None if not tabulated)."""
    base = name.split("<")[0]  # template arguments do not change the bytes
    if base not in KERNEL_BYTES:
        return None
    per_cell, fields = KERNEL_BYTES[base]
    if base == "k_fft_filter":
        cells = ni * filtered_rows(nj) * nk
    else:
        cells = ni * (nj + 1) * nk
    return per_cell * cells * fields * members


REF_EXE = os.path.join(HERE, "oracle", "_ref", "epic_ref")


def cpu_baseline_reference(planet, ni, nj, nk, dt, budget_s):
    """The reference's own timestep() on one core: its C sources built
    single-processor in FP64 (EPIC_PRECISION=8, gcc -O2) by oracle/Makefile
    and driven by oracle/ref_harness.c from the same seeded case, which
    times its step loop itself.  A 2-step run sizes the sample to about
    budget_s of stepping."""
    import re
    import tempfile
    import epic_amd as ea
    sp = spec_of(ea, planet, ni, nj, nk, dt, 1)

    def run(n):
        with tempfile.TemporaryDirectory() as td:
            args = [REF_EXE] + [str(x) for x in (sp.planet, sp.ni, sp.nj, sp.nk, sp.dt, sp.seed, sp.nu4_nondim,
                                                   sp.nudiv_nondim, sp.noise_theta, sp.noise_uv, n)]
            r = subprocess.run(args + [os.path.join(td, "out")], cwd=td, capture_output=True, text=True,
                               timeout=4 * budget_s + 300)
        m = re.search(r"timed (\d+) steps in ([0-9.eE+-]+) s", r.stdout)
        if r.returncode != 0 or not m:
            raise RuntimeError(f"{REF_EXE} failed ({r.returncode}): {r.stderr[-400:]}")
        return int(m.group(1)), float(m.group(2))

    n, el = run(2)
    n, el = run(max(2, min(200, int(budget_s / (el / n)))))
    return ni * (nj + 1) * nk * n / el, n, el


def cpu_baseline(planet, ni, nj, nk, dt, budget_s):
    from oracle.cpuref import CpuRef
    import epic_amd as ea
    sp = spec_of(ea, planet, ni, nj, nk, dt, 1)
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


def copy_probe(nbytes=256 << 20, reps=20):
    """Bandwidth of a plain device-to-device streaming copy (k_copy_probe):
    (read + write bytes) / s."""
    import epic_amd as ea
    s = ea.copy_probe(nbytes, reps)
    return {"bytes_per_copy": 2 * nbytes, "GBps": 2 * nbytes / s / 1e9, "copy_us": s * 1e6}


def spec_of(ea, planet, ni, nj, nk, dt, seed):
    """The seeded case of a workload: jets plus 0.5 m/s u, v noise on the
    Jupiter/Saturn-like planets, 1 K theta noise on the hot Jupiter (at rest,
    day-night forcing), 0.1 K on Held-Suarez."""
    if planet == ea.HOT_JUPITER:
        return ea.spec(planet, ni, nj, nk, dt, seed=seed, nu4=0.5, noise_theta=1.0, noise_uv=0.0)
    gas = planet in (ea.JUPITER_LIKE, ea.SATURN_LIKE)
    return ea.spec(planet, ni, nj, nk, dt, seed=seed, nu4=0.5,
                   noise_theta=0.0 if gas else 0.1, noise_uv=0.5 if gas else 0.0)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--workload", default="jupiter_256x128x20", choices=sorted(WORKLOADS))
    ap.add_argument("--precision", default="fp64", choices=["fp64", "mixed"])
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--serve", type=int, default=16,
                    help="requests of the serving mix to run after the step benchmark (0: skip)")
    ap.add_argument("--no-autotune", action="store_true",
                    help="keep the built-in launch shapes instead of timing the candidates first")
    ap.add_argument("--copy-probe", action="store_true",
                    help="also time a plain 256 MiB device copy (the measured stream-copy peak)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one MI355X; --gpus {args.gpus} is not supported "
              "(this build runs on one GPU; members would shard across GPUs with no exchange)",
              file=sys.stderr)
        sys.exit(2)

    import epic_amd as ea
    planet, ni, nj, nk, dt, members = WORKLOADS[args.workload]
    prec = ea.FP64 if args.precision == "fp64" else ea.MIXED
    eng = ea.Engine(prec, members)
    for m in range(members):
        eng.add(spec_of(ea, planet, ni, nj, nk, dt, 1 + m))
    eng.step(args.warmup)
    eng.sync()
    tuned = None if args.no_autotune else eng.autotune(2).strip().splitlines()
    eng.step(1)  # captures the step's graph with the chosen shapes
    eng.sync()
    t0 = time.perf_counter()
    eng.step(args.steps)
    eng.sync()
    el = time.perf_counter() - t0
    npts = ni * (nj + 1) * nk * members
    value = npts * args.steps / el

    prof = eng.profile(3)
    tot = sum(ms for _, ms, _ in prof)
    ranked = sorted(prof, key=lambda x: -x[1])
    name, ms, cnt = next(x for x in ranked if kernel_bytes(x[0], ni, nj, nk, members) is not None)
    avg_s = ms / cnt / 1e3
    nbytes = kernel_bytes(name, ni, nj, nk, members)
    achieved = nbytes / avg_s
    step_bytes = 112 if prec == ea.FP64 else 88
    traffic, traffic_src = measured_traffic(args.workload, args.precision, name)

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
                   "precision": args.precision, "n_pts": npts, "launch_shapes": tuned},
        "roofline": {"bound": "hbm", "kernel": name, "achieved": achieved / 1e9, "peak": PEAK_HBM / 1e9,
                     "unit": "GB/s", "frac": achieved / PEAK_HBM, "traffic": traffic, "traffic_source": traffic_src,
                     "launch_us": avg_s * 1e6, "bytes_per_launch": nbytes,
                     "share_of_step": ms / tot if tot else None,
                     "top_kernels": [[n, round(m / 3 * 1e3, 1), c // 3] for n, m, c in ranked[:8]]},
        "step_roofline": {"bytes_per_pt": step_bytes, "achieved": value * step_bytes / 1e9,
                          "frac": value * step_bytes / PEAK_HBM, "unit": "GB/s"},
    }
    eng.close()
    if args.copy_probe:
        out["copy_probe"] = copy_probe()
    if args.serve > 0:
        from epic_amd import serving as sv
        reqs = sv.poisson_mix(args.serve, rate=4.0, seed=2024)
        prec_s = ea.FP64 if args.precision == "fp64" else ea.MIXED
        res, summ = sv.serve(reqs, precision=prec_s, hbm_budget=128 << 30, max_members=32, chunk=50)
        summ["mix"] = "Poisson arrivals at 4/s of {64x32x10,128x64x16,256x128x20,512x256x32} x " \
                      "{held_suarez, jupiter, saturn, venus} x U(200,2000) steps, seed 2024"
        out["serving"] = summ
    if not args.no_cpu:
        ref_ok = os.path.exists(REF_EXE)
        if ref_ok:
            try:
                v, n, cel = cpu_baseline_reference(planet, ni, nj, nk, dt, args.cpu_seconds)
                kind, what = "reference", "the reference's own timestep() (single-processor FP64 build)"
            except (RuntimeError, OSError, subprocess.TimeoutExpired) as exc:
                print(f"reference CPU leg failed, timing the oracle instead: {exc}", file=sys.stderr)
                ref_ok = False
        if not ref_ok:
            v, n, cel = cpu_baseline(planet, ni, nj, nk, dt, args.cpu_seconds)
            kind, what = "port", "the oracle's CPU restatement"
        out["cpu_baseline"] = {"value": v, "unit": "grid-pt-steps/s", "cores": 1, "kind": kind,
                               "sample": f"{n} FP64 steps of one {ni}x{nj}x{nk} member by {what} in {cel:.1f} s"}
    print(json.dumps(out))


if __name__ == "__main__":
    main()
