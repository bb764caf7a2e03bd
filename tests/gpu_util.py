"""Helpers shared by the GPU parity tests: run a case on the HIP engine and
on the oracle side by side, and measure their difference.

Tolerance.  The GPU evaluates exp, log and pow with the device math library,
whose results may differ from the host libm's in the last bit; every other
operation of the step is performed in the reference's order with no
contraction, so those last-bit differences are the only source of
disagreement, and the reference algorithm carries them forward (its
finite-volume pressure gradient amplifies them by 1e4-1e5 in near-balanced
states).  The oracle is therefore run twice: as is, and with 10% of its own
exp/log/pow results moved by one ulp (oracle/libm_nudge.h).  The second
run's distance from the first is the sensitivity floor of the case.  A
field passes when its GPU-vs-oracle relative L2 error is at most
    max(base, FLOOR_FACTOR * floor)
with base = 1e-12 for u, v, h, theta, p and 1e-10 for the diagnostics and
tendencies.  Any index, sign or missing-term mistake gives errors many
orders of magnitude above this bound."""
import sys

import numpy as np

import epic_amd as ea
from oracle.cpuref import CpuRef

PROG = ("U", "V", "HDRY", "THETA", "P3")
DIAG = ("P2", "H3", "THETA2", "T2", "T3", "RHO2", "RHO3", "EXNER3", "GZ3", "MONT3", "PV3",
        "DIV_UV3", "HEAT3")
PROG_TOL = 1e-12
DIAG_TOL = 1e-10
FLOOR_FACTOR = 10.0


def rel_l2(a, b):
    nb = float(np.sqrt((b * b).sum()))
    d = float(np.sqrt(((a - b) ** 2).sum()))
    if nb == 0.0:
        return d
    return d / nb


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


class Pair:
    """One case advanced on the GPU engine, on the oracle and on the
    libm-nudged oracle."""

    def __init__(self, sp, precision=ea.FP64):
        self.sp = sp
        self.eng = ea.Engine(precision, 1)
        self.slot = self.eng.add(sp)
        self.ref = CpuRef(sp)
        self.ref.sync()
        self.nud = CpuRef(sp, nudged=True)
        self.nud.sync()
        self.fn = None  # oracle instance used as a function of the GPU's fields
        self.steps = 0

    def step(self, n):
        self.eng.step(n)
        self.ref.step(n)
        self.nud.step(n)
        self.steps += n
        self.eng.sync()

    def field(self, f):
        return self.eng.read(self.slot, f), self.ref.read(f), self.nud.read(f)

    def tendency_field(self, wind):
        it = self.ref.time_planes()
        return self.field(f"D{wind}DT{it[1]}")

    def report(self, tag):
        """{name: (gpu error, floor, tolerance)} for every checked field."""
        out = {}
        items = [(f, self.field(f), PROG_TOL if f in PROG else DIAG_TOL) for f in PROG + DIAG]
        if self.steps:
            items += [(f"D{w}DT", self.tendency_field(w), DIAG_TOL) for w in ("U", "V")]
        for name, (g, r, n), base in items:
            e, fl = rel_l2(g, r), rel_l2(n, r)
            out[name] = (e, fl, max(base, FLOOR_FACTOR * fl))
        say(tag, " ".join(f"{k}={e:.1e}/{fl:.1e}" for k, (e, fl, _) in out.items()))
        return out

    def check(self, tag):
        rep = self.report(tag)
        bad = {k: v for k, v in rep.items() if v[0] > v[2]}
        assert not bad, (tag, bad)
        self.check_ri2(tag)
        return rep

    def check_ri2(self, tag):
        """RI2 (a ratio of a near-cancelling buoyancy term to the squared
        shear, with no libm call) is pinned bit for bit as a function: the
        GPU's RI2 must equal the oracle's store_diag formula applied to the
        GPU's own U, V, P2, T2, RHO2, P3, RHO3.  Its distance from the
        oracle's RI2 is reported (one-ulp differences in P2, T2, RHO2 move it
        by up to about 1e-9 in relative L2 through the cancellation)."""
        if self.fn is None:
            self.fn = CpuRef(self.sp)
        g = {f: self.eng.read(self.slot, f) for f in ("U", "V", "P2", "T2", "RHO2", "P3", "RHO3", "RI2")}
        want = self.fn.richardson_of(g)
        say(tag, f"RI2 vs oracle {rel_l2(g['RI2'], self.ref.read('RI2')):.1e}")
        assert g["RI2"].tobytes() == want.tobytes(), (tag, "RI2", int((g["RI2"] != want).sum()))

    def close(self):
        self.eng.close()
        self.ref.close()
        self.nud.close()
        if self.fn is not None:
            self.fn.close()
