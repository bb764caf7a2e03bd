"""CPU checks of the C ABI: the library loads, exports every symbol that
include/epic_amd.h declares, and the restated reference structs have the
reference's sizes and offsets (compile-time assertions in abi_check.c)."""
import ctypes
import os
import re
import subprocess

import pytest

import epic_amd as ea

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HDR = os.path.join(ROOT, "include", "epic_amd.h")


def declared_functions():
    src = open(HDR).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    names = re.findall(r"^\s*[A-Za-z_][\w\s\*]*?\b(\w+)\s*\([^;{]*\)\s*;", src, flags=re.M)
    return sorted(set(n for n in names if n not in ("if", "return")))


def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(ea.LIB_PATH)
    names = declared_functions()
    assert "timestep" in names and "epic_engine_step" in names
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def test_struct_layout_assertions_compile():
    r = subprocess.run(["gcc", "-fsyntax-only", "-std=gnu11", os.path.join(ROOT, "epic_amd", "csrc", "abi_check.c")],
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stderr


def test_case_builder_and_footprint():
    sp = ea.spec(ea.HELD_SUAREZ, 64, 32, 10)
    assert ea.footprint(64, 32, 10, ea.FP64) > 30 * 8 * 66 * 35 * 12
    assert ea.footprint(64, 32, 10, ea.MIXED) < ea.footprint(64, 32, 10, ea.FP64)


def _needed(exe):
    r = subprocess.run(["readelf", "-d", "--dyn-syms", "-W", exe], capture_output=True, text=True)
    return r.stdout


@pytest.mark.parametrize("exe", ["tests/native/_bin/dropin_host", "oracle/_ref/epic_ref_dropin"])
def test_host_programs_bind_timestep_to_the_library(exe):
    """An EPIC-shaped executable leaves timestep() undefined (it binds to
    libepic_amd.so) and exports the globals the drop-in reads."""
    path = os.path.join(ROOT, exe)
    if not os.path.exists(path):
        pytest.skip(f"{exe} not built")
    out = _needed(path)
    assert "libepic_amd.so" in out
    syms = {}
    for line in out.splitlines():
        p = line.split()
        if len(p) >= 8 and p[0].rstrip(":").isdigit():
            syms[p[7]] = p[6]
    assert syms.get("timestep") == "UND"
    for g in ("grid", "var", "IT_ZERO", "IT_MINUS1", "IT_MINUS2"):
        assert g in syms and syms[g] != "UND", g


def test_library_leaves_the_callers_globals_weak():
    out = _needed(ea.LIB_PATH)
    for line in out.splitlines():
        p = line.split()
        if len(p) >= 8 and p[7] in ("grid", "var", "IT_ZERO", "IT_MINUS1", "IT_MINUS2", "solar_longitude",
                                    "epic_error"):
            assert p[4] == "WEAK" and p[6] == "UND", line


@pytest.mark.parametrize("name,opts,msg", [
    ("jupiter", [], "needs the t_vs_p sounding"),
    ("venus", [], "needs the t_vs_p sounding"),
    ("hot_jupiter", [], "needs the t_vs_p sounding"),
    # Venus (LLR05) takes its temp_eq in the reference, not the sounding
    ("venus_llr05", ["--newt_cool=1"], "venus_llr05 and titan"),
])
def test_dropin_refuses_newtonian_cooling_it_does_not_have(tmp_path, name, opts, msg):
    """With grid.newt_cool_on, a planet other than held_suarez relaxes toward
    its t_vs_p sounding with t_cool_table's times in the reference (a planet
    named hot_jupiter too: the engine's day-night relaxation is its own case,
    not the reference's), so a caller without the sounding is refused, and
    so are the two planets whose relaxation is their temp_eq; timestep()
    stops with an EPIC-style error before any GPU work instead of stepping
    with another forcing."""
    exe = os.path.join(ROOT, "tests", "native", "_bin", "dropin_host")
    if not os.path.exists(exe):
        pytest.skip("dropin_host not built")
    # the Held-Suarez case (newt_cool_on = 1, no sounding) under another
    # planet's name; with --newt_cool=1 a Jupiter case with its sounding
    planet = "1" if opts else "0"
    args = ([exe] + opts + [planet, "32", "16", "8", "600", "1", "0.5", "0", "0.1", "0", "1",
                            str(tmp_path / "out"), "0", name])
    r = subprocess.run(args, cwd=str(tmp_path), capture_output=True, text=True, timeout=60)
    assert r.returncode == 1, r.stdout + r.stderr
    assert msg in r.stderr
    assert not os.path.exists(tmp_path / "out")


@pytest.mark.parametrize("cname,pyname", [("epic_case_spec", "CaseSpec"), ("epic_case", "Case")])
def test_python_binding_layouts_match_the_header(cname, pyname, tmp_path):
    """The ctypes restatements of epic_case_spec and epic_case have the C
    header's size and member offsets (gcc, this machine's ABI)."""
    st = getattr(ea, pyname)
    src = tmp_path / "lay.c"
    lines = ['#include <stdio.h>', '#include <stddef.h>', f'#include "{HDR}"', "int main(void) {",
             f'  printf("size %zu\\n", sizeof({cname}));']
    for name, _ in st._fields_:
        lines.append(f'  printf("{name} %zu\\n", offsetof({cname}, {name}));')
    lines.append("  return 0;\n}")
    src.write_text("\n".join(lines))
    exe = tmp_path / "lay"
    r = subprocess.run(["gcc", "-std=gnu11", "-o", str(exe), str(src)], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    out = dict(line.split() for line in subprocess.run([str(exe)], capture_output=True, text=True).stdout.splitlines())
    assert int(out["size"]) == ctypes.sizeof(st)
    for name, _ in st._fields_:
        assert int(out[name]) == getattr(st, name).offset, name


def test_case_builder_through_the_binding():
    """epic_case_build through ctypes: the padded extents and the sponge of
    the reference `initial` program's choice (2 layers at nk = 10)."""
    c = ea.build_case(ea.spec(ea.JUPITER_LIKE, 32, 16, 10, 60, 3))
    try:
        assert (c.ni, c.nj, c.nk, c.Iadim, c.Jadim, c.Kadim) == (32, 16, 10, 34, 19, 12)
        assert c.Nelem3d == 34 * 19 * 12 and c.k_sponge == 2
        ts = c.array("t_sponge_inv", 3)
        assert ts[0] == 1.0 / (30.0 * 60) and ts[2] == 0.0 and 0 < ts[1] < ts[0]
    finally:
        ea.free_case(c)
