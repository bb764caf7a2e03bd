#!/usr/bin/env python3
"""Writes the golden fixtures of the parity tests from the reference's own
FP64 build (oracle/_ref/epic_ref, built by `make -C oracle ref` from the
reference sources).  Runs only where the reference tree is present.

Fixtures (inputs are the seeded case specs recorded in golden.json):
  <name>.npz        padded u, v, hdry, theta, p3, gz3, pv3, w3 of the 16x8x6
                    grids after the SYNC call and after 1 and 10 steps
  golden.json       the case specs, and sha256 of every field of every case
                    and dump, including
                    the 64x32x10 Held-Suarez case at steps 1, 10 and 100,
                    plus per-level L2 norms of u, v, hdry, theta, p3
"""
import hashlib
import json
import os
import subprocess
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
EXE = os.path.join(ROOT, "oracle", "_ref", "epic_ref")
sys.path.insert(0, ROOT)
from epic_amd import FIELDS  # noqa: E402

CASES = {
    # name: (planet, ni, nj, nk, dt, seed, nu4, nudiv, noise_theta, noise_uv, dumps[, options])
    # options: the reference's switches as --name=value of the harness
    # (k_sponge -1 is the reference `initial` program's choice)
    "hs_tiny": (0, 16, 8, 6, 600, 7, 0.3, 0.2, 0.5, 0.5, [0, 1, 10], {"k_sponge": 0}),
    "jup_tiny": (1, 16, 8, 6, 60, 5, 0.5, 0.0, 0.0, 0.5, [0, 1, 10], {"k_sponge": 0}),
    "venus_tiny": (3, 16, 8, 6, 600, 9, 0.5, 0.1, 0.1, 0.5, [0, 1, 10], {"k_sponge": 0}),
    "hs_c1": (0, 64, 32, 10, 600, 1, 0.5, 0.0, 0.1, 0.0, [1, 10, 100], {"k_sponge": 0}),
    # the top sponge as `initial` sets it, Rayleigh drag on u and v about
    # the zonal mean
    "jup_drag_tiny": (1, 16, 8, 6, 60, 6, 0.5, 0.0, 0.0, 0.5, [0, 1, 10],
                      {"k_sponge": -1, "tau_drag": 2.0e4, "drag_v": 1, "drag_zonal_avg": 1}),
    # Held-Suarez with the layer-mean heating removal and prandtl/t_rad drag
    "hs_adjust_tiny": (0, 16, 8, 6, 600, 8, 0.3, 0.2, 0.5, 0.5, [0, 1, 10],
                       {"newt_cool_adjust": 1, "prandtl": 0.3, "drag_v": 1}),
    # the reference `initial` program's hyperviscosity default: del^8 alone
    "jup_del8_tiny": (1, 16, 8, 6, 60, 9, 0.0, 0.0, 0.0, 0.5, [0, 1, 10],
                      {"k_sponge": -1, "nu8_nondim": 0.5}),
    # Newtonian cooling toward the sounding with the tabulated cooling time,
    # its layer-mean removal and the prandtl drag that reads that time
    "jup_newt_tiny": (1, 16, 8, 6, 60, 10, 0.5, 0.0, 0.0, 0.5, [0, 1, 10],
                      {"k_sponge": -1, "newt_cool": 1, "newt_cool_adjust": 1, "prandtl": 0.3, "drag_v": 1}),
}


def spec_of(entry):
    """The case spec of a golden.json case entry: its positional values and
    the reference switches it set (k_sponge 0 where none is given)."""
    from epic_amd import spec
    opts = dict(entry.get("options", {}))
    sp = spec(*entry["spec"], k_sponge=opts.pop("k_sponge", 0))
    for k, v in opts.items():
        setattr(sp, k, v)
    return sp


def read_dump(fn):
    b = open(fn, "rb").read()
    magic, ni, nj, nk, nd, nf, n3, n2, mism = np.frombuffer(b[:36], dtype=np.int32)
    assert magic == 0x45504943 and mism == 0
    off, out = 36, []
    for _ in range(nd):
        hdr = np.frombuffer(b[off:off + 16], dtype=np.int32).copy()
        off += 16
        f = {}
        for q in range(26):
            f[FIELDS[q]] = np.frombuffer(b[off:off + 8 * n3], dtype=np.float64).reshape(nk + 2, nj + 3, ni + 2).copy()
            off += 8 * n3
        f["GZ_SURFACE"] = np.frombuffer(b[off:off + 8 * n2], dtype=np.float64).reshape(nj + 3, ni + 2).copy()
        off += 8 * n2
        f["RI2"] = np.frombuffer(b[off:off + 8 * n3], dtype=np.float64).reshape(nk + 2, nj + 3, ni + 2).copy()
        off += 8 * n3
        out.append((hdr, f))
    return out


def main():
    if not os.path.exists(EXE):
        sys.exit("reference build missing: make -C oracle ref")
    hashes, meta = {}, {}
    for name, c in CASES.items():
        args, dumps, opts = list(c[:10]), c[10], (c[11] if len(c) > 11 else {})
        fn = f"/tmp/golden_{name}.bin"
        cmd = ([EXE] + [f"--{k}={v}" for k, v in opts.items()] + [str(x) for x in args] +
               [str(max(dumps)), fn] + [str(d) for d in dumps])
        subprocess.run(cmd, check=True, cwd="/tmp")
        meta[name] = {"spec": args, "dumps": dumps, "options": opts}
        arrays = {}
        for hdr, f in read_dump(fn):
            step = int(hdr[0])
            key = f"{name}/step{step}"
            hashes[key] = {"time_planes": [int(x) for x in hdr[1:4]]}
            for nm, a in f.items():
                hashes[key][nm] = hashlib.sha256(a.tobytes()).hexdigest()
            for nm in ("U", "V", "HDRY", "THETA", "P3"):
                hashes[key]["l2_" + nm] = [float(np.sqrt((f[nm][k] ** 2).sum())) for k in range(f[nm].shape[0])]
            if name.endswith("_tiny"):
                for nm in ("U", "V", "HDRY", "THETA", "P3", "GZ3", "PV3", "W3"):
                    arrays[f"s{step}_{nm}"] = f[nm]
        if arrays:
            np.savez_compressed(os.path.join(HERE, f"{name}.npz"), **arrays)
    json.dump({"cases": meta, "hashes": hashes}, open(os.path.join(HERE, "golden.json"), "w"), indent=1)


if __name__ == "__main__":
    main()
