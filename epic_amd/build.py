"""Builds libepic_amd.so in-tree for gfx950 (hipcc for the HIP and C++
sources, gcc for the C sources that must round like the host libm path)."""
import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OUT = os.path.join(HERE, "libepic_amd.so")
OBJ = os.path.join(HERE, "_obj")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("EPIC_AMD_ARCH", "gfx950")

HIP_FLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-fPIC", "-std=c++17", "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-command-line-argument", "-I", CSRC,
]
C_FLAGS = ["-O2", "-fPIC", "-std=gnu11", "-ffp-contract=off"]

HIP_SOURCES = ["k_diag.hip", "k_dyn.hip", "k_filter.hip", "k_adv.hip", "k_mixed.hip", "engine.cpp", "dropin.cpp",
               "noalloc.cpp"]
C_SOURCES = ["epic_case.c", "epic_tables.c", "ncio.c", "epic_nc.c"]


def _run(cmd):
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if r.returncode != 0:
        raise RuntimeError("build step failed: " + " ".join(cmd) + "\n" + r.stdout)
    return r.stdout


def build(verbose=False):
    os.makedirs(OBJ, exist_ok=True)
    jobs = []
    for src in HIP_SOURCES:
        path = os.path.join(CSRC, src)
        if not os.path.exists(path):
            continue
        obj = os.path.join(OBJ, src + ".o")
        lang = [] if src.endswith(".hip") else ["-x", "hip"]
        jobs.append([HIPCC] + HIP_FLAGS + lang + ["-fgpu-rdc"] * 0 + ["-c", path, "-o", obj])
    for src in C_SOURCES:
        obj = os.path.join(OBJ, src + ".o")
        jobs.append(["gcc"] + C_FLAGS + ["-c", os.path.join(CSRC, src), "-o", obj])
    with ThreadPoolExecutor(max_workers=min(8, len(jobs))) as ex:
        outs = list(ex.map(_run, jobs))
    if verbose:
        for o in outs:
            if o.strip():
                print(o)
    objs = [j[-1] for j in jobs]
    tmp = OUT + ".tmp"
    _run([HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp] + objs + ["-lm"])
    os.replace(tmp, OUT)
    return OUT


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
