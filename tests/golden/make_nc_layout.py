#!/usr/bin/env python3
"""Writes tests/golden/epic_nc_layout.json: the global attributes the
reference's var_read reads from an epic.nc file for ALL_DATA (name, the
READ macro -- C text, I int, F floating, D double, TIME int --, the value
count as written there, and the switch it depends on), and the variables it
reads, listed from src/shared/epic_funcs_io.c (var_read, :730-1368).  Runs
only where the reference tree is present; the JSON is the fixture."""
import json
import os
import re
import sys

REF = "/root/reference/src/shared/epic_funcs_io.c"
HERE = os.path.dirname(os.path.abspath(__file__))


def main():
    if not os.path.exists(REF):
        sys.exit("reference tree missing")
    lines = open(REF).read().splitlines()
    start = next(i for i, l in enumerate(lines) if l.startswith("void var_read("))
    end = next(i for i in range(start + 1, len(lines)) if lines[i].startswith("void var_write("))
    atts, cond = [], None
    for l in lines[start:end]:
        s = l.strip()
        if "nmt_physics_on == 1" in s:
            cond = "nmt_physics_on"
        elif "var.ntp > 0" in s:
            cond = "ntp"
        elif "var.n_t_cool > 0" in s:
            cond = "n_t_cool"
        m = re.match(r"READ(C|I|F|D|TIME)\(\s*&?([^,]+),\s*([A-Za-z0-9_]+)\s*(?:,\s*(.+?))?\);", s)
        if m:
            kind, _, name, count = m.groups()
            atts.append({"name": name, "kind": kind, "count": count or "1", "if": cond})
        if s == "}" and cond and l.startswith("    }"):
            cond = None
    variables = ["u", "v", "p3", "theta", "u_spinup", "gz_surface", "du_IT_MINUS1dt", "du_IT_MINUS2dt",
                 "dv_IT_MINUS1dt", "dv_IT_MINUS2dt"]
    json.dump({"source": "src/shared/epic_funcs_io.c var_read", "attributes": atts, "variables": variables},
              open(os.path.join(HERE, "epic_nc_layout.json"), "w"), indent=1)
    print(len(atts), "attributes")


if __name__ == "__main__":
    main()
