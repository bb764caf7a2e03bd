"""Helpers shared by the GPU parity tests: run a case on the HIP engine and
on the oracle side by side, and measure their difference."""
import sys

import numpy as np

import epic_amd as ea
from oracle.cpuref import CpuRef

PROG = ("U", "V", "HDRY", "THETA", "P3")
DIAG = ("P2", "H3", "THETA2", "T2", "T3", "RHO2", "RHO3", "EXNER3", "GZ3", "MONT3", "PV3",
        "DIV_UV3", "HEAT3")


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
    """One case advanced on the GPU engine and on the CPU oracle."""

    def __init__(self, sp, precision=ea.FP64):
        self.sp = sp
        self.eng = ea.Engine(precision, 1)
        self.slot = self.eng.add(sp)
        self.ref = CpuRef(sp)
        self.ref.sync()
        self.steps = 0

    def step(self, n):
        self.eng.step(n)
        self.ref.step(n)
        self.steps += n
        self.eng.sync()

    def gpu(self, f):
        return self.eng.read(self.slot, f)

    def cpu(self, f):
        return self.ref.read(f)

    def err(self, f):
        return rel_l2(self.gpu(f), self.cpu(f))

    def tendency_err(self, wind):
        """Newest tendency plane (IT_MINUS1 after a step) of U or V."""
        it = self.ref.time_planes()
        f = f"D{wind}DT{it[1]}"
        return rel_l2(self.gpu(f), self.cpu(f))

    def close(self):
        self.eng.close()
        self.ref.close()
