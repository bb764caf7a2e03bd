"""Helpers shared by the GPU parity tests: run a case on the HIP engine and
on the oracle side by side, and measure their difference.

Two oracles.  The GPU's FP64 kernels evaluate exp, log and pow with the
correctly rounded functions of epic_amd/csrc/crmath.h and perform every
other operation in the reference's order with no contraction.  The oracle
built with the same crmath.h (oracle/libcpuref_cr.so) must therefore equal
the GPU bit for bit in every field, and that is the parity check proper:
any index, sign, order or missing-term mistake breaks it.

Against the plain oracle (host libm, bit-exact with the reference's own
build) the GPU differs only through glibc's own last-bit roundings of exp,
log and pow (it rounds about 5e-4 of them differently from the correctly
rounded value), which the dynamics carry forward: one ulp of theta at a
single cell grows to ~3e-12 (u) and ~5e-12 (v) relative after 100 steps of
Held-Suarez 64x32x10 (tests/test_crmath.py).  That distance is reported
beside the distance of a second copy of the oracle with 10% of its
exp/log/pow results moved by one ulp (the floor); h, theta and p are held
to 1e-12 (diagnostics 1e-10) or ten times the floor, u and v to ten times
the floor."""
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


def bitwise_diff(get_a, get_b, fields=None):
    """[(field, cells differing)] of two sources of fields."""
    out = []
    for f in (fields or ea.FIELDS):
        a, b = get_a(f), get_b(f)
        if a.tobytes() != b.tobytes():
            out.append((f, int((a != b).sum())))
    return out


class Pair:
    """One case advanced on the GPU engine, on the oracle with the GPU's
    correctly rounded exp/log/pow, on the plain oracle and on the
    libm-nudged oracle."""

    def __init__(self, sp, precision=ea.FP64):
        self.sp = sp
        self.eng = ea.Engine(precision, 1)
        self.slot = self.eng.add(sp)
        self.cr = CpuRef(sp, nudged="cr")
        self.cr.sync()
        self.ref = CpuRef(sp)
        self.ref.sync()
        self.nud = CpuRef(sp, nudged=True)
        self.nud.sync()
        self.steps = 0

    def step(self, n):
        self.eng.step(n)
        self.cr.step(n)
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
        """{name: (gpu error vs the plain oracle, floor, tolerance)}."""
        out = {}
        items = [(f, self.field(f), PROG_TOL if f in ("HDRY", "THETA", "P3") else (0.0 if f in PROG else DIAG_TOL))
                 for f in PROG + DIAG]
        if self.steps:
            items += [(f"D{w}DT", self.tendency_field(w), DIAG_TOL) for w in ("U", "V")]
        for name, (g, r, n), base in items:
            e, fl = rel_l2(g, r), rel_l2(n, r)
            out[name] = (e, fl, max(base, FLOOR_FACTOR * fl))
        say(tag, " ".join(f"{k}={e:.1e}/{fl:.1e}" for k, (e, fl, _) in out.items()))
        return out

    def check(self, tag):
        """Bit for bit against the oracle with the GPU's exp/log/pow (every
        field, the tendency planes, GZ_SURFACE and RI2 included, and the
        time planes), then within tolerance of the plain oracle."""
        bad = bitwise_diff(lambda f: self.eng.read(self.slot, f), self.cr.read)
        say(tag, "fields differing from the correctly rounded oracle:", bad)
        assert not bad, (tag, bad)
        assert self.eng.time_planes(self.slot)[:3] == self.cr.time_planes()[:3], tag
        rep = self.report(tag)
        bad = {k: v for k, v in rep.items() if v[0] > v[2]}
        assert not bad, (tag, bad)
        return rep

    def close(self):
        self.eng.close()
        self.cr.close()
        self.ref.close()
        self.nud.close()
