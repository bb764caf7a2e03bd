"""The drop-in timestep() on the GPU, called the way an EPIC executable
calls it: from a host program that owns the globals grid, var and IT_*.

tests/native/dropin_host.c is a minimal such program; oracle/_ref/
epic_ref_dropin is the reference's own program (its globals, metric set-up,
set_sponge and solar_longitude) with the reference's timestep() renamed away
so that the call binds to libepic_amd.so (present where the reference tree
was available at build time)."""
import os
import subprocess

import numpy as np
import pytest

import epic_amd as ea
from oracle.cpuref import CpuRef
from tests.golden.make_golden import read_dump
from tests.gpu_util import DIAG, DIAG_TOL, FLOOR_FACTOR, PROG, PROG_TOL, bitwise_diff, rel_l2, say

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HOST = os.path.join(ROOT, "tests", "native", "_bin", "dropin_host")
REF = os.path.join(ROOT, "oracle", "_ref", "epic_ref")
REF_DROPIN = os.path.join(ROOT, "oracle", "_ref", "epic_ref_dropin")

# (positional spec values, reference switches)
HS = ((0, 32, 16, 8, 600, 21, 0.5, 0.1, 0.3, 0.3), {})  # Held-Suarez with divergence damping
# a gas giant: gz_from_u rewrites gz_surface every call; the top sponge as
# `initial` sets it and Rayleigh drag about the zonal mean
JUP = ((1, 32, 16, 10, 60, 23, 0.5, 0.0, 0.0, 1.0),
       {"k_sponge": -1, "tau_drag": 2.0e4, "drag_v": 1, "drag_zonal_avg": 1})
# Newtonian cooling toward the t_vs_p sounding with t_cool_table's times,
# and the prandtl drag that reads those times
SAT_NEWT = ((2, 32, 12, 8, 60, 28, 0.4, 0.0, 0.0, 0.5), {"k_sponge": 2, "newt_cool": 1, "prandtl": 0.5, "drag_v": 1})
# the Akima advection scheme for hdry and theta (var.*.advection_scheme)
AKIMA = ((0, 32, 20, 8, 900, 31, 0.5, 0.1, 0.3, 0.4), {"adv_mass": 1, "adv_nonmass": 1})
# a caller built with the Arakawa-Lamb PV averaging (epic_amd_set_pv_scheme)
PV_AL81 = ((1, 32, 16, 10, 60, 35, 0.5, 0.0, 0.0, 1.0), {"k_sponge": -1, "pv_scheme": 1})
NSTEPS = 12


def opts_args(opts):
    return [f"--{k}={v}" for k, v in opts.items()]


def spec_of(case):
    args, opts = case
    sp = ea.spec(*args, k_sponge=opts.get("k_sponge", 0))
    for k, v in opts.items():
        if k != "k_sponge":
            setattr(sp, k, v)
    return sp


def run(exe, case, extra, tmp_path, name):
    args, opts = case
    fn = str(tmp_path / name)
    cmd = [exe] + opts_args(opts) + [str(a) for a in args] + [str(NSTEPS), fn] + [str(x) for x in extra]
    say("run", os.path.basename(exe), *cmd[1:])
    r = subprocess.run(cmd, cwd=str(tmp_path), capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    (hdr, f), = [d for d in read_dump(fn) if int(d[0][0]) == NSTEPS]
    return hdr, f


def cr_equal(case, f, tag):
    """Every array the drop-in wrote back equals the oracle built with the
    GPU's correctly rounded exp/log/pow, bit for bit."""
    _, cr = oracle_state(case, nudged="cr")
    bad = bitwise_diff(lambda q: f[q], lambda q: cr[q])
    say(tag, "arrays differing from the correctly rounded oracle:", bad)
    assert not bad, (tag, bad)


def oracle_state(case, nudged=False):
    r = CpuRef(spec_of(case), nudged=nudged)
    r.sync()
    r.step(NSTEPS)
    out = {f: r.read(f) for f in ea.FIELDS}
    tp = r.time_planes()[:3]
    r.close()
    return tp, out


def compare(case, f, ref, tag):
    """Tolerances of tests/gpu_util.py: the base bound or ten times the
    libm-nudged oracle's distance from the oracle, whichever is larger."""
    _, exact = oracle_state(case)
    _, nud = oracle_state(case, nudged=True)
    for q in PROG + DIAG + ("GZ_SURFACE",):
        e, fl = rel_l2(f[q], ref[q]), rel_l2(nud[q], exact[q])
        tol = max(PROG_TOL if q in PROG else DIAG_TOL, FLOOR_FACTOR * fl)
        say(tag, q, f"{e:.1e}/{fl:.1e}")
        assert e <= tol, (tag, q, e, fl)


def ri2_of(case, f):
    """The oracle's store_diag RI2 of the U, V, P2, T2, RHO2, P3, RHO3 in f."""
    fn = CpuRef(spec_of(case))
    try:
        return fn.richardson_of(f)
    finally:
        fn.close()


@pytest.mark.skipif(not os.path.exists(HOST), reason="tests/native/_bin/dropin_host not built")
@pytest.mark.parametrize("case", [HS, JUP, SAT_NEWT, AKIMA, PV_AL81],
                         ids=["held_suarez", "jupiter_sponge_drag", "saturn_sounding", "held_suarez_akima",
                              "jupiter_pv_al81"])
def test_dropin_eager_and_lazy_match_oracle(case, tmp_path):
    h_e, f_e = run(HOST, case, [0], tmp_path, "eager.bin")
    h_l, f_l = run(HOST, case, [1], tmp_path, "lazy.bin")
    for q in ea.FIELDS:
        assert np.array_equal(f_e[q], f_l[q]), q
    tp, ref = oracle_state(case)
    assert [int(x) for x in h_e[1:4]] == tp
    cr_equal(case, f_e, "dropin_host")
    compare(case, f_e, ref, "dropin_host")
    # var.ri2 comes back formed from the written-back state
    assert f_e["RI2"].tobytes() == ri2_of(case, f_e).tobytes()


@pytest.mark.skipif(not (os.path.exists(REF_DROPIN) and os.path.exists(REF)),
                    reason="reference programs (oracle/_ref) not built")
@pytest.mark.parametrize("case", [HS, JUP, SAT_NEWT, AKIMA],
                         ids=["held_suarez", "jupiter_sponge_drag", "saturn_sounding", "held_suarez_akima"])
def test_reference_program_with_dropin_matches_reference(case, tmp_path):
    h_g, f_g = run(REF_DROPIN, case, [NSTEPS], tmp_path, "gpu.bin")
    h_c, f_c = run(REF, case, [NSTEPS], tmp_path, "cpu.bin")
    assert [int(x) for x in h_g[1:4]] == [int(x) for x in h_c[1:4]]
    cr_equal(case, f_g, "epic_ref_dropin")
    compare(case, f_g, f_c, "epic_ref_dropin")
    # a gas giant's gz_surface moves with the zonal-mean bottom wind; the
    # caller's array must hold the step's value, not the initial one
    if case is JUP:
        assert not np.array_equal(f_c["GZ_SURFACE"], np.zeros_like(f_c["GZ_SURFACE"]))
    # RI2: the drop-in's from its own state, the reference's from its own,
    # each equal to the oracle's formula of that state
    assert f_g["RI2"].tobytes() == ri2_of(case, f_g).tobytes()
    assert f_c["RI2"].tobytes() == ri2_of(case, f_c).tobytes()
