#!/usr/bin/env python3
"""Throughput of the EPIC dynamical-core step on one MI355X, mixed
FP32/FP64 beside FP64, at a stated field error.

Workload (default): the Jupiter-like globe ni=256 nj=128 nk=20, dt=60 s,
ideal gas, eight seeded zonal jets plus 0.5 m/s u,v noise, del^4
hyperviscosity, polar filter on, the top sponge as the reference's
`initial` program sets it.  A "step" is one
timestep(STEP_PROGS_AND_SYNC_DIAGS) of every member, with the state resident
in HBM.  The unit is grid-point-steps/s with N_pts = ni*(nj+1)*nk per member.

Prints ONE JSON line.  `value` is the mixed-precision step's rate (FP32
tendency arithmetic and history; FP64 state, hydrostatics and
Adams-Bashforth accumulation); `fp64` holds the FP64 step's rate, and
`mixed_over_fp64` their ratio.  `field_error` is the mixed run against the
FP64 run after 1000 steps of member 0 (relative L2 and relative max per
field, global mass and total-energy drift of each).  `roofline` is for the
kernel that takes the most device time in a profiled pass of the mixed step
(HIP events around every launch, on the engine's stream): achieved = its
algorithmic bytes per launch / its mean launch duration, with the per-cell
byte counts of KERNEL_BYTES; `traffic` is that kernel's HBM bytes per launch
from the committed counter summary named in `traffic_source` (the counters
cannot be read inside this run).  `step_roofline` prices the whole step:
N_pts*steps/s * 112 B (FP64) or 88 B (mixed) against 8 TB/s.
`cpu_baseline` is the reference's own timestep() built for one CPU
(oracle/_ref/epic_ref, where it was built; else the oracle's FP64 CPU
restatement) on a bounded number of steps of the same workload: one
process on each of the host's cores at once (`value`, `cores`), and one
process alone (`single_core`).
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
    # the ensemble: 64 members of the Jupiter-like planet, each with its own
    # seed (noise and jet coefficients)
    "jupiter_ensemble_64x128x64x16": (1, 128, 64, 16, 60, 64),
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
    # set_p2etc<1> reads h, theta and writes p3, h, p2, h3, theta2, theta;
    # <0> also forms T3 and RHO3: reads h, theta, writes p3, h, t3, rho3;
    # k_thermo reads theta, p3, p2, theta2 and writes t3, rho3, exner3, t2,
    # rho2, heat3, dzdt3
    "k_set_p2etc<1>": (56, 1), "k_set_p2etc<0>": (48, 1), "k_thermo": (88, 1),
    "k_gz3": (40, 1), "k_calc_w": (56, 1), "k_dzdt3": (72, 1),
    "k_hsu_vert": (24, 3), "k_hsu_mer": (40, 3),
    "k_hsu_zonal": (32, 3), "k_fft_filter": (16, 3),
    "k_filter3": (24, 4), "k_zonal_hv": (16, 2), "k_d2dy2_cols": (16, 2), "k_hv_apply": (24, 2),
    # the mixed mode's fused del^4 rows: D (FP32) read, the field read and written
    "k_hv4_rows": (20, 2),
    "k_pgrad": (48, 1), "k_uvh_kin": (48, 1), "k_dvdt_core": (40, 1), "k_dudt_core": (40, 1), "k_ab": (80, 1),
    "k_divdamp": (40, 1), "km_uvh_kin": (36, 1), "km_dvdt_core": (32, 1), "km_dudt_core": (32, 1), "km_ab": (64, 1),
    "km_divdamp": (40, 1), "k_gz_mono": (16, 1),
    "k_zero_tend": (16, 1), "k_adv_elem": (16, 1), "k_heating": (24, 1), "k_sponge": (24, 1), "k_hs_pbl": (32, 1),
    "k_divergence": (24, 1), "k_store_diag": (40, 1),
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
    if name.endswith(">"):
        more = [c for c in k if c.startswith(name[:-1] + ",")]
        if len(more) == 1:
            return k[more[0]]["traffic_bytes"], os.path.relpath(fn, HERE)
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
    if name in KERNEL_BYTES:  # ... except where the table names them
        per_cell, fields = KERNEL_BYTES[name]
    elif base in KERNEL_BYTES:
        per_cell, fields = KERNEL_BYTES[base]
    else:
        return None
    if base == "k_fft_filter":
        cells = ni * filtered_rows(nj) * nk
    else:
        cells = ni * (nj + 1) * nk
    return per_cell * cells * fields * members


REF_EXE = os.path.join(HERE, "oracle", "_ref", "epic_ref")


def _ref_args(sp, n, out):
    return [REF_EXE] + [f"--{k}={v}" for k, v in sp.options().items()] + \
        [str(x) for x in sp.args()] + [str(n), out]


def _ref_timed(stdout):
    import re
    m = re.search(r"timed (\d+) steps in ([0-9.eE+-]+) s", stdout)
    return (int(m.group(1)), float(m.group(2))) if m else None


def cpu_baseline_reference(planet, ni, nj, nk, dt, budget_s):
    """The reference's own timestep() on one core: its C sources built
    single-processor in FP64 (EPIC_PRECISION=8, gcc -O2) by oracle/Makefile
    and driven by oracle/ref_harness.c from the same seeded case, which
    times its step loop itself.  A 2-step run sizes the sample to about
    budget_s of stepping."""
    import tempfile
    import epic_amd as ea
    sp = spec_of(ea, planet, ni, nj, nk, dt, 1)

    def run(n):
        with tempfile.TemporaryDirectory() as td:
            r = subprocess.run(_ref_args(sp, n, os.path.join(td, "out")), cwd=td, capture_output=True, text=True,
                               timeout=4 * budget_s + 300)
        t = _ref_timed(r.stdout)
        if r.returncode != 0 or not t:
            raise RuntimeError(f"{REF_EXE} failed ({r.returncode}): {r.stderr[-400:]}")
        return t

    n, el = run(2)
    n, el = run(max(2, min(200, int(budget_s / (el / n)))))
    return ni * (nj + 1) * nk * n / el, n, el


def host_cores():
    """Cores this process may use (its affinity set), at most 16: the CPU
    share of one GPU on the boxes this runs on."""
    try:
        n = len(os.sched_getaffinity(0))
    except AttributeError:
        n = os.cpu_count() or 1
    return max(1, min(16, n))


def cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline_reference_all_cores(planet, ni, nj, nk, dt, nsteps, cores):
    """The same build, one process per core at once, each stepping its own
    member (seed 1..cores) nsteps times: the aggregate rate of the host
    (the reference's MPI build needs an MPI library this image lacks;
    independent members are what the ensemble and serving runs are)."""
    import tempfile
    import epic_amd as ea
    with tempfile.TemporaryDirectory() as td:
        procs = []
        for c in range(cores):
            sp = spec_of(ea, planet, ni, nj, nk, dt, 1 + c)
            d = os.path.join(td, str(c))
            os.makedirs(d)
            procs.append(subprocess.Popen(_ref_args(sp, nsteps, os.path.join(d, "out")), cwd=d,
                                          stdout=subprocess.PIPE, stderr=subprocess.PIPE, text=True))
        times = []
        for p in procs:
            out, err = p.communicate(timeout=3600)
            t = _ref_timed(out)
            if p.returncode != 0 or not t:
                raise RuntimeError(f"{REF_EXE} failed ({p.returncode}): {err[-400:]}")
            times.append(t[1])
    return cores * ni * (nj + 1) * nk * nsteps / max(times), max(times)


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
    day-night forcing), 0.1 K on Held-Suarez.  The top sponge is the
    reference `initial` program's choice, except on the hot Jupiter, whose
    stated forcing is the day-night relaxation alone (a sponge would also
    remove the relaxation's layer mean in its layers)."""
    if planet == ea.HOT_JUPITER:
        return ea.spec(planet, ni, nj, nk, dt, seed=seed, nu4=0.5, noise_theta=1.0, noise_uv=0.0, k_sponge=0)
    gas = planet in (ea.JUPITER_LIKE, ea.SATURN_LIKE)
    return ea.spec(planet, ni, nj, nk, dt, seed=seed, nu4=0.5,
                   noise_theta=0.0 if gas else 0.1, noise_uv=0.5 if gas else 0.0)


def run_engine(ea, prec, workload, steps, warmup, autotune):
    """Warm up, tune, then time `steps` steps of a workload; returns the
    engine (still holding its members) and the figures."""
    planet, ni, nj, nk, dt, members = WORKLOADS[workload]
    eng = ea.Engine(prec, members)
    for m in range(members):
        eng.add(spec_of(ea, planet, ni, nj, nk, dt, 1 + m))
    eng.step(warmup)
    eng.sync()
    tuned = eng.autotune(2).strip().splitlines() if autotune else None
    eng.step(1)  # captures the step's graph with the chosen shapes
    eng.sync()
    t0 = time.perf_counter()
    eng.step(steps)
    eng.sync()
    el = time.perf_counter() - t0
    npts = ni * (nj + 1) * nk * members
    step_bytes = 112 if prec == ea.FP64 else 88
    value = npts * steps / el
    fig = {"value": value, "ms_per_step": el / steps * 1e3, "launch_shapes": tuned,
           "step_roofline": {"bytes_per_pt": step_bytes, "achieved": value * step_bytes / 1e9,
                             "frac": value * step_bytes / PEAK_HBM, "unit": "GB/s"}}
    return eng, fig


def field_error(ea, workload, steps=1000):
    """Member 0 of the workload in the mixed mode against the FP64 mode
    after `steps` steps: relative L2 and relative max of each field's
    interior, and each run's global mass (sum of h) and total-energy
    (sum of h (u^2/2 + T)) drift."""
    import numpy as np
    planet, ni, nj, nk, dt, _ = WORKLOADS[workload]
    sp = spec_of(ea, planet, ni, nj, nk, dt, 1)
    fields = ("U", "V", "HDRY", "THETA", "P3", "GZ3", "PV3")

    def inner(a):
        return a[1:nk + 1, 1:nj + 2, 1:ni + 1]

    def mass_energy(e, s):
        h, u, t = (inner(e.read(s, f)) for f in ("HDRY", "U", "T3"))
        return float(h.sum()), float((h * (0.5 * u * u + t)).sum())

    out = {}
    for prec in (ea.FP64, ea.MIXED):
        e = ea.Engine(prec, 1)
        s = e.add(sp)
        m0 = mass_energy(e, s)
        e.step(steps)
        e.sync()
        m1 = mass_energy(e, s)
        out[prec] = ({f: inner(e.read(s, f)) for f in fields}, (m1[0] - m0[0]) / m0[0], (m1[1] - m0[1]) / m0[1])
        e.close()
    ref, mix = out[ea.FP64], out[ea.MIXED]
    err = {}
    for f in fields:
        a, b = mix[0][f], ref[0][f]
        err[f] = [float(np.linalg.norm(a - b) / np.linalg.norm(b)), float(np.abs(a - b).max() / np.abs(b).max())]
    return {"steps": steps, "member": 0, "rel_l2_and_rel_max": err,
            "mass_drift": {"fp64": ref[1], "mixed": mix[1]},
            "energy_drift": {"fp64": ref[2], "mixed": mix[2]}}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--workload", default="jupiter_256x128x20", choices=sorted(WORKLOADS))
    ap.add_argument("--precision", default="both", choices=["both", "fp64", "mixed"],
                    help="both: time the mixed step (the value) and the FP64 step beside it")
    ap.add_argument("--error-steps", type=int, default=1000,
                    help="steps of the mixed-vs-FP64 field comparison (0: skip)")
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--serve", type=int, default=16,
                    help="requests of the serving mix to run after the step benchmark (0: skip)")
    ap.add_argument("--serve-policy", choices=["shortest", "lockstep"], default="shortest",
                    help="serving ticks step the shape ending soonest, or every shape together")
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
    npts = ni * (nj + 1) * nk * members
    figs = {}
    if args.precision in ("both", "fp64"):
        eng, figs["fp64"] = run_engine(ea, ea.FP64, args.workload, args.steps, args.warmup, not args.no_autotune)
        if args.precision == "fp64":
            head, prec, pname = eng, ea.FP64, "fp64"
        else:
            eng.close()
    if args.precision in ("both", "mixed"):
        head, figs["mixed"] = run_engine(ea, ea.MIXED, args.workload, args.steps, args.warmup, not args.no_autotune)
        prec, pname = ea.MIXED, "mixed"
    main_fig = figs[pname]

    prof = head.profile(3)
    tot = sum(ms for _, ms, _ in prof)
    ranked = sorted(prof, key=lambda x: -x[1])
    name, ms, cnt = next(x for x in ranked if kernel_bytes(x[0], ni, nj, nk, members) is not None)
    avg_s = ms / cnt / 1e3
    nbytes = kernel_bytes(name, ni, nj, nk, members)
    achieved = nbytes / avg_s
    traffic, traffic_src = measured_traffic(args.workload, pname, name)
    head.close()

    out = {
        "metric": "grid-pt-steps/s mixed vs FP64 at stated field error",
        "value": main_fig["value"],
        "unit": "grid-pt-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": main_fig["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if prec == ea.FP64 else "f32+f64",
        "data": "synthetic (seeded initial state of the workload's planet)",
        "config": {"workload": args.workload, "ni": ni, "nj": nj, "nk": nk, "members": members, "dt": dt,
                   "precision": pname, "n_pts": npts, "launch_shapes": main_fig["launch_shapes"]},
        "roofline": {"bound": "hbm", "kernel": name, "achieved": achieved / 1e9, "peak": PEAK_HBM / 1e9,
                     "unit": "GB/s", "frac": achieved / PEAK_HBM, "traffic": traffic,
                     "traffic_source": (f"committed counter summary {traffic_src} (FETCH_SIZE, WRITE_SIZE of a "
                                        "rocprofv3 run of this workload)") if traffic_src else None,
                     "launch_us": avg_s * 1e6, "bytes_per_launch": nbytes,
                     "share_of_step": ms / tot if tot else None,
                     "top_kernels": [[n, round(m / 3 * 1e3, 1), c // 3] for n, m, c in ranked[:8]]},
        "step_roofline": main_fig["step_roofline"],
    }
    if "fp64" in figs and "mixed" in figs:
        out["fp64"] = figs["fp64"]
        out["mixed_over_fp64"] = figs["mixed"]["value"] / figs["fp64"]["value"]
    if args.error_steps > 0 and args.precision == "both":
        out["field_error"] = field_error(ea, args.workload, args.error_steps)
    if args.copy_probe:
        out["copy_probe"] = copy_probe()
    if args.serve > 0:
        from epic_amd import serving as sv
        reqs = sv.poisson_mix(args.serve, rate=4.0, seed=2024)
        res, summ = sv.serve(reqs, precision=prec, hbm_budget=128 << 30, max_members=32, chunk=50,
                             policy=args.serve_policy)
        summ["mix"] = "Poisson arrivals at 4/s of {64x32x10,128x64x16,256x128x20,512x256x32} x " \
                      "{held_suarez, jupiter, saturn, venus} x U(200,2000) steps, seed 2024"
        summ["precision"] = pname
        out["serving"] = summ
    if not args.no_cpu:
        wl = f"{ni}x{nj}x{nk}"
        ref_ok = os.path.exists(REF_EXE)
        if ref_ok:
            try:
                v, n, cel = cpu_baseline_reference(planet, ni, nj, nk, dt, args.cpu_seconds)
                kind, what = "reference", "the reference's own timestep() (single-processor FP64 build)"
                cores = host_cores()
                va, wall = cpu_baseline_reference_all_cores(planet, ni, nj, nk, dt, n, cores)
            except (RuntimeError, OSError, subprocess.TimeoutExpired) as exc:
                print(f"reference CPU leg failed, timing the oracle instead: {exc}", file=sys.stderr)
                ref_ok = False
        if not ref_ok:
            v, n, cel = cpu_baseline(planet, ni, nj, nk, dt, args.cpu_seconds)
            kind, what = "port", "the oracle's CPU restatement"
            cores, va, wall = 1, v, cel
        single = {"value": v, "unit": "grid-pt-steps/s", "cores": 1,
                  "sample": f"{n} FP64 steps of one {wl} member by {what} in {cel:.1f} s"}
        out["cpu_baseline"] = {"value": va, "unit": "grid-pt-steps/s", "cores": cores, "kind": kind,
                               "cpu": cpu_model(),
                               "sample": f"{cores} processes at once, each {n} FP64 steps of its own {wl} member "
                                         f"by {what}, {wall:.1f} s" if cores > 1 else single["sample"],
                               "single_core": single}
    print(json.dumps(out))


if __name__ == "__main__":
    main()
