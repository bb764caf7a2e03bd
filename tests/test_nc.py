"""epic.nc: the reference's netCDF-3 model-state file (epic_amd/csrc/ncio.c,
epic_nc.c).

The netCDF classic codec is pinned against netCDF's own test files
(tests/golden/ref_nctest_classic.nc and ref_nctest_64bit_offset.nc, from the
reference's vendored netCDF 3.6.3 nctest suite) and against scipy's
independent netCDF-3 reader and writer; the EPIC layout against the list of
attributes the reference's var_read reads (tests/golden/epic_nc_layout.json,
written by make_nc_layout.py from src/shared/epic_funcs_io.c).  The
reference's own netCDF library needs generated build files and is not built
here, so no file written by the reference's var_write is available."""
import ctypes
import json
import os

import numpy as np
import pytest
from scipy.io import netcdf_file

import epic_amd as ea
from oracle.cpuref import CpuRef

HERE = os.path.dirname(os.path.abspath(__file__))
GOLD = os.path.join(HERE, "golden")


# ---------------------------------------------------------------- the codec
class _Att(ctypes.Structure):
    _fields_ = [("name", ctypes.c_char_p), ("type", ctypes.c_int), ("n", ctypes.c_size_t), ("val", ctypes.c_void_p)]


class _Dim(ctypes.Structure):
    _fields_ = [("name", ctypes.c_char_p), ("len", ctypes.c_size_t)]


class _Var(ctypes.Structure):
    _fields_ = [("name", ctypes.c_char_p), ("ndims", ctypes.c_int), ("dimid", ctypes.c_int * 8),
                ("natts", ctypes.c_int), ("atts", ctypes.POINTER(_Att)), ("type", ctypes.c_int),
                ("data", ctypes.c_void_p), ("nval", ctypes.c_size_t), ("vsize", ctypes.c_size_t),
                ("begin", ctypes.c_size_t)]


class _File(ctypes.Structure):
    _fields_ = [("version", ctypes.c_int), ("numrecs", ctypes.c_size_t), ("ndims", ctypes.c_int),
                ("ngatts", ctypes.c_int), ("nvars", ctypes.c_int), ("dims", ctypes.POINTER(_Dim)),
                ("gatts", ctypes.POINTER(_Att)), ("vars", ctypes.POINTER(_Var))]


_NP = {1: np.int8, 2: np.uint8, 3: np.int16, 4: np.int32, 5: np.float32, 6: np.float64}


def _vals(ptr, typ, n):
    if n == 0:
        return np.zeros(0, dtype=_NP[typ])
    buf = (ctypes.c_char * (n * np.dtype(_NP[typ]).itemsize)).from_address(ptr)
    return np.frombuffer(bytes(buf), dtype=_NP[typ])


def ncio_parse(path):
    """The file as ncio_read sees it: version, numrecs, dims, global
    attributes and variables (dims, attributes, values)."""
    L = ea.lib()
    L.ncio_read.argtypes = [ctypes.POINTER(_File), ctypes.c_char_p]
    L.ncio_free.argtypes = [ctypes.POINTER(_File)]
    f = _File()
    rc = L.ncio_read(ctypes.byref(f), str(path).encode())
    assert rc == 0, rc
    try:
        dims = [(f.dims[i].name.decode(), f.dims[i].len) for i in range(f.ndims)]
        gatts = {f.gatts[i].name.decode(): _vals(f.gatts[i].val, f.gatts[i].type, f.gatts[i].n)
                 for i in range(f.ngatts)}
        vs = {}
        for i in range(f.nvars):
            v = f.vars[i]
            vs[v.name.decode()] = {
                "dims": [dims[v.dimid[d]][0] for d in range(v.ndims)],
                "atts": {v.atts[a].name.decode(): _vals(v.atts[a].val, v.atts[a].type, v.atts[a].n)
                         for a in range(v.natts)},
                "values": _vals(v.data, v.type, v.nval)}
        return f.version, f.numrecs, dims, gatts, vs
    finally:
        L.ncio_free(ctypes.byref(f))


def _same_att(mine, theirs):
    theirs = np.atleast_1d(theirs)
    if theirs.dtype.kind in "SU":
        return bytes(mine).rstrip(b"\0") == (theirs.tobytes() if theirs.dtype.kind == "S" else
                                             "".join(theirs).encode()).rstrip(b"\0")
    return np.array_equal(mine.astype(np.float64), theirs.astype(np.float64))


@pytest.mark.parametrize("name", ["ref_nctest_classic.nc", "ref_nctest_64bit_offset.nc"])
def test_ncio_reads_netcdfs_own_test_files(name):
    """netCDF's reference files (every type, several record variables, a
    scalar, CDF-1 and CDF-2) read as scipy reads them."""
    path = os.path.join(GOLD, name)
    version, numrecs, dims, gatts, vs = ncio_parse(path)
    f = netcdf_file(path, "r", mmap=False)
    assert version == f.version_byte
    assert [(n, l or None) for n, l in dims] == list(f.dimensions.items())
    assert numrecs == f._recs
    assert set(gatts) == set(f._attributes)
    for k, v in f._attributes.items():
        assert _same_att(gatts[k], v), k
    assert list(vs) == list(f.variables)
    for k, v in f.variables.items():
        assert vs[k]["dims"] == list(v.dimensions), k
        want = np.asarray(v.data)
        got = vs[k]["values"]
        if want.dtype.kind == "S":
            want = np.frombuffer(want.tobytes(), dtype=np.uint8)
        assert got.size == want.size, k
        assert np.array_equal(got.reshape(-1).view(got.dtype), want.reshape(-1).astype(got.dtype)), k
        for a, av in v._attributes.items():
            assert _same_att(vs[k]["atts"][a], av), (k, a)
    f.close()


# ------------------------------------------------------------- the layout
def _case_fields(c):
    n3, n2 = c.Nelem3d, c.Nelem2d
    F = {k: c.array(k.lower(), n3) for k in ("U", "V", "P3", "THETA", "U_SPINUP")}
    F["GZ_SURFACE"] = c.array("gz_surface", n2)
    return F


def _write_case(path, sp, **kw):
    c = ea.build_case(sp)
    try:
        F = _case_fields(c)
        rng = np.random.default_rng(5)
        # distinct Adams-Bashforth planes, to see each land in its place
        for k in ("DUDT_M1", "DUDT_M2", "DVDT_M1", "DVDT_M2"):
            F[k] = rng.standard_normal(c.Nelem3d) * 1e-4
        ea.nc_write(path, c, F, **kw)
        return {k: v.copy() for k, v in F.items()}
    finally:
        ea.free_case(c)


def _interior(a, sp):
    """K = 0..nk, J = 0..nj, I = 1..ni of a padded array."""
    return a.reshape(sp.nk + 2, sp.nj + 3, sp.ni + 2)[:sp.nk + 1, 1:sp.nj + 2, 1:sp.ni + 1]


JUP_NEWT = ea.spec(ea.JUPITER_LIKE, 32, 16, 10, 60, 3, 0.5, 0.0, 0.0, 0.5, newt_cool=1, tau_drag=2e4, drag_v=1)
HS = ea.spec(ea.HELD_SUAREZ, 32, 16, 8, 600, 4, 0.5, 0.1, 0.3, 0.3)


@pytest.mark.parametrize("sp", [JUP_NEWT, HS], ids=["jupiter_sounding", "held_suarez"])
def test_nc_file_holds_what_var_read_reads(sp, tmp_path):
    """Every global attribute the reference's var_read reads for ALL_DATA is
    present with the type its READ macro demands (a double-precision
    build: floating values as NC_DOUBLE) and at least the count it reads;
    every variable it reads is there with its dimensions, read back by
    scipy's independent reader."""
    path = tmp_path / "epic.nc"
    F = _write_case(path, sp, model_time=7200.)
    lay = json.load(open(os.path.join(GOLD, "epic_nc_layout.json")))
    f = netcdf_file(path, "r", mmap=False)
    c = ea.build_case(sp)
    consts = {"GEOM_STR": 16, "N_STR": 256, "TOPDIM": 3, "LAST_INDEX-FIRST_INDEX+1": 54,
              "MAX_NU_ORDER+1": 9, "2*(nk+1)+1": 2 * (sp.nk + 1) + 1, "nk+1": sp.nk + 1,
              "var.ntp": c.ntp, "var.n_t_cool": c.n_t_cool}
    kind_dtype = {"C": "S", "I": "i", "TIME": "i", "F": "f", "D": "f"}
    try:
        for a in lay["attributes"]:
            if a["if"] == "nmt_physics_on" or (a["if"] == "ntp" and c.ntp == 0) or \
                    (a["if"] == "n_t_cool" and c.n_t_cool == 0):
                assert a["name"] not in f._attributes or a["if"] != "nmt_physics_on"
                continue
            assert a["name"] in f._attributes, a["name"]
            v = np.atleast_1d(f._attributes[a["name"]])
            if a["kind"] == "C":
                assert isinstance(f._attributes[a["name"]], bytes), a["name"]
                continue
            assert v.dtype.kind == kind_dtype[a["kind"]], a
            if a["kind"] in ("F", "D"):
                assert v.dtype.itemsize == 8, a
            n = int(a["count"]) if a["count"].isdigit() else consts[a["count"]]
            assert v.size >= n, a
        g = f._attributes
        assert (g["grid_ni"], g["grid_nj"], g["grid_nk"], g["var_model_time"]) == (sp.ni, sp.nj, sp.nk, 7200)
        assert g["planet_name"].rstrip(b"\0") == c.planet_name
        assert np.array_equal(g["grid_sigmatheta"], c.array("sigmatheta", 2 * (sp.nk + 1) + 1))
        for name in lay["variables"]:
            if name == "gz_surface" and c.planet_type != b"terrestrial":
                assert name not in f.variables
                continue
            assert name in f.variables, name
        grids = {"u": "u", "v": "v", "p3": "p3", "theta": "p3", "u_spinup": "u"}
        for name, g_ in grids.items():
            var = f.variables[name]
            assert var.dimensions == ("time", f"sigmatheta_{g_}", f"lat_{g_}", f"lon_{g_}")
            key = {"u_spinup": "U_SPINUP"}.get(name, name.upper())
            assert np.array_equal(var.data[0], _interior(F[key], sp)), name
        for w, g_ in (("u", "u"), ("v", "v")):
            for p in (1, 2):
                var = f.variables[f"d{w}_IT_MINUS{p}dt"]
                assert var.dimensions == (f"sigmatheta_{g_}", f"lat_{g_}", f"lon_{g_}")
                assert np.array_equal(var.data, _interior(F[f"D{w.upper()}DT_M{p}"], sp))
        # the staggered coordinates
        lon, lat = c.array("lon", 2 * (sp.ni + 1) + 1), c.array("lat", 2 * (sp.nj + 1) + 1)
        assert np.array_equal(f.variables["lon_u"].data, lon[2:2 * sp.ni + 1:2])
        assert np.array_equal(f.variables["lon_p3"].data, lon[3:2 * sp.ni + 2:2])
        assert np.array_equal(f.variables["lat_v"].data, lat[0:2 * sp.nj + 1:2])
        assert np.array_equal(f.variables["lat_p3"].data, lat[1:2 * sp.nj + 2:2])
        assert f.variables["time"].data[0] == 7200. / 86400.
    finally:
        f.close()
        ea.free_case(c)


def _case_arrays(c, sp):
    nJ, nK = 2 * (sp.nj + 1) + 1, 2 * (sp.nk + 1) + 1
    out = {k: c.array(k, nJ) for k in ("lat", "f", "m", "n", "mn", "g", "rln", "rlt")}
    out.update({k: c.array(k, nK) for k in ("sigmatheta", "dsgth", "dsgth_inv", "p_ref", "theta_ref")})
    out["lon"] = c.array("lon", 2 * (sp.ni + 1) + 1)
    out["h_min"] = c.array("h_min", sp.nk + 1)
    out["t_sponge_inv"] = c.array("t_sponge_inv", c.k_sponge + 1)
    return out


@pytest.mark.parametrize("sp", [JUP_NEWT, HS], ids=["jupiter_sounding", "held_suarez"])
def test_nc_read_rebuilds_the_case(sp, tmp_path):
    """Reading a written file gives the case back: every scalar the step
    uses, every metric array (formed again from the scalars) and the state
    bit for bit, the Adams-Bashforth planes in dudt_hist/dvdt_hist, and the
    sounding with the reference's hPa round trip (p/100*100, within an ulp)."""
    path = tmp_path / "epic.nc"
    F = _write_case(path, sp, model_time=3600.)
    c, r = ea.build_case(sp), ea.nc_read(path)
    try:
        for k in ("ni", "nj", "nk", "dt", "jtp", "k_sigma", "k_sponge", "newt_cool_on", "newt_cool_adjust",
                  "drag_v", "drag_zonal_avg", "planet_index", "ntp", "n_t_cool"):
            assert getattr(r, k) == getattr(c, k), k
        for k in ("re", "rp", "omega", "cp", "rgas", "cpr", "kappa", "GM", "J2", "kinvisc", "dln", "dlt",
                  "press0", "pbot", "zeta0", "zeta1", "hybrid_alpha", "sigma_sigma", "topo_scale", "prandtl",
                  "tau_drag", "nudiv_nondim", "dy0"):
            assert getattr(r, k) == getattr(c, k), k
        assert list(r.nu) == list(c.nu) and list(r.nu_nondim) == list(c.nu_nondim)
        assert r.planet_name == c.planet_name and r.planet_type == c.planet_type
        a, b = _case_arrays(c, sp), _case_arrays(r, sp)
        for k in a:
            assert a[k].tobytes() == b[k].tobytes(), k
        n3 = c.Nelem3d
        for k in ("u", "v", "p3", "theta"):
            assert r.array(k, n3).tobytes() == c.array(k, n3).tobytes(), k
        # u_spinup above the top from the layer below it, as var_read sets it
        us = r.array("u_spinup", n3).reshape(sp.nk + 2, sp.nj + 3, sp.ni + 2)
        assert np.array_equal(us[0], us[1])
        hist = {"DUDT": r.array("dudt_hist", 2 * n3), "DVDT": r.array("dvdt_hist", 2 * n3)}
        for w in ("DUDT", "DVDT"):
            for p in (1, 2):
                got = hist[w][(p - 1) * n3:p * n3]
                assert np.array_equal(_interior(got, sp), _interior(F[f"{w}_M{p}"], sp)), (w, p)
                g3 = got.reshape(sp.nk + 2, sp.nj + 3, sp.ni + 2)
                assert np.array_equal(g3[:, :, 0], g3[:, :, sp.ni]) and np.array_equal(g3[:, :, sp.ni + 1], g3[:, :, 1])
        assert r.model_time == 3600.
        if c.ntp:
            pd, pr = c.array("pdat", c.ntp), r.array("pdat", c.ntp)
            assert np.all(np.abs(pd - pr) <= np.spacing(pd))
            assert np.array_equal(pd / 100. * 100., pr)
            assert r.array("tdat", c.ntp).tobytes() == c.array("tdat", c.ntp).tobytes()
            tx = c.array("t_cool_x", c.n_t_cool)
            assert np.array_equal(r.array("t_cool_x", c.n_t_cool), -np.log(np.exp(-tx) / 100. * 100.))
            assert r.array("t_cool_y", c.n_t_cool).tobytes() == c.array("t_cool_y", c.n_t_cool).tobytes()
        if c.planet_type == b"terrestrial":
            assert r.array("gz_surface", c.Nelem2d).tobytes() == c.array("gz_surface", c.Nelem2d).tobytes()
    finally:
        ea.free_case(c)
        ea.free_case(r)


def _rewrite_with_scipy(src, dst, edit=None):
    """The file as scipy's writer lays it out, optionally with a global
    attribute changed."""
    f = netcdf_file(src, "r", mmap=False)
    g = netcdf_file(dst, "w", version=f.version_byte)
    for k, v in f._attributes.items():
        setattr(g, k, v)
    if edit:
        for k, v in edit.items():
            setattr(g, k, v)
    for k, n in f.dimensions.items():
        g.createDimension(k, n)
    for k, v in f.variables.items():
        w = g.createVariable(k, v.typecode(), v.dimensions)
        for a, av in v._attributes.items():
            setattr(w, a, av)
        if v.dimensions and v.dimensions[0] == "time":
            w[:] = v.data
        else:
            w[...] = v.data
    g.close()
    f.close()


def test_nc_reads_a_file_of_another_writer(tmp_path):
    """A file written by scipy's independent netCDF-3 writer (its own
    header and data placement) reads to the same case."""
    src, dst = tmp_path / "a.nc", tmp_path / "b.nc"
    _write_case(src, JUP_NEWT, model_time=60.)
    _rewrite_with_scipy(src, dst)
    a, b = ea.nc_read(src), ea.nc_read(dst)
    try:
        n3 = a.Nelem3d
        sizes = {"sigmatheta": 2 * (a.nk + 1) + 1, "dsgth": 2 * (a.nk + 1) + 1, "f": 2 * (a.nj + 1) + 1,
                 "g": 2 * (a.nj + 1) + 1}
        for k in ("u", "v", "p3", "theta", "u_spinup", "sigmatheta", "dsgth", "f", "g"):
            n = sizes.get(k, n3)
            assert a.array(k, n).tobytes() == b.array(k, n).tobytes(), k
        assert a.array("dudt_hist", 2 * n3).tobytes() == b.array("dudt_hist", 2 * n3).tobytes()
    finally:
        ea.free_case(a)
        ea.free_case(b)


@pytest.mark.parametrize("edit,msg", [
    ({"grid_turbulence_scheme": b"Spalart-Allmaras DES"}, "grid_turbulence_scheme"),
    ({"grid_uv_timestep_scheme": b"Leapfrog (Asselin filtered)"}, "grid_uv_timestep_scheme"),
    ({"grid_microphysics_on": np.int32(1)}, "microphysics"),
    ({"var_theta_advection_scheme": b"Upwind"}, "var_theta_advection_scheme"),
])
def test_nc_read_refuses_what_the_step_does_not_implement(edit, msg, tmp_path):
    src, dst = tmp_path / "a.nc", tmp_path / "b.nc"
    _write_case(src, JUP_NEWT)
    _rewrite_with_scipy(src, dst, edit)
    with pytest.raises(ea.EpicError, match=msg):
        ea.nc_read(dst)


@pytest.mark.parametrize("mass,nonmass", [(1, 0), (0, 1), (1, 1)])
def test_nc_carries_the_advection_schemes(mass, nonmass, tmp_path):
    """The Akima scheme of hdry and of theta goes into var_hdry_ and
    var_theta_advection_scheme by the reference's name and comes back."""
    src, dst = tmp_path / "a.nc", tmp_path / "b.nc"
    sp = ea.spec(ea.HELD_SUAREZ, 16, 8, 5, 600, 3, adv_mass=mass, adv_nonmass=nonmass)
    _write_case(src, sp)
    akima, hsu = b"Monotonized, 4th-order centered (Akima)", b"Predictor-corrector (Hsu, Konor, and Arakawa)"
    f = netcdf_file(src, "r", mmap=False)
    try:
        assert f._attributes["var_hdry_advection_scheme"].rstrip(b"\0") == (akima if mass else hsu)
        assert f._attributes["var_theta_advection_scheme"].rstrip(b"\0") == (akima if nonmass else hsu)
    finally:
        f.close()
    _rewrite_with_scipy(src, dst)
    r = ea.nc_read(dst)
    try:
        assert (r.adv_mass, r.adv_nonmass) == (mass, nonmass)
    finally:
        ea.free_case(r)


def test_nc_read_names_a_missing_attribute(tmp_path):
    src, dst = tmp_path / "a.nc", tmp_path / "b.nc"
    _write_case(src, HS)
    f = netcdf_file(src, "r", mmap=False)
    atts = dict(f._attributes)
    f.close()
    del atts["grid_k_sigma"]
    _rewrite_with_scipy(src, dst)
    g = netcdf_file(dst, "a", mmap=False)
    del g._attributes["grid_k_sigma"]
    g.close()
    with pytest.raises(ea.EpicError, match="grid_k_sigma"):
        ea.nc_read(dst)


def test_oracle_restart_from_nc_continues_the_run(tmp_path):
    """The oracle's state after 4 steps, written to epic.nc and taken up as
    the reference's epic_main takes it up (HDRY formed again from P3),
    continues to within 1e-12 of the uninterrupted run after 4 more steps
    (the rounding of HDRY -> P3 -> HDRY is the only difference), and a file
    written at start-up (FLOAT_MAX tendency planes) restarts bit for bit."""
    sp = JUP_NEWT
    a = CpuRef(sp)
    a.sync()
    a.step(4)
    tp = a.time_planes()[:3]
    c = ea.build_case(sp)
    try:
        F = {k: a.read(k).reshape(-1) for k in ("U", "V", "P3", "THETA")}
        F["U_SPINUP"] = c.array("u_spinup", c.Nelem3d)
        F["GZ_SURFACE"] = a.read("GZ_SURFACE").reshape(-1)
        for w in ("U", "V"):
            F[f"D{w}DT_M1"] = a.read(f"D{w}DT{tp[1]}").reshape(-1)
            F[f"D{w}DT_M2"] = a.read(f"D{w}DT{tp[2]}").reshape(-1)
        ea.nc_write(tmp_path / "s4.nc", c, F, model_time=4 * sp.dt)
        ea.nc_write(tmp_path / "s0.nc", c, _case_fields(c))
    finally:
        ea.free_case(c)
    b = CpuRef(sp, nc_path=tmp_path / "s4.nc")
    b.sync()
    a.step(4)
    b.step(4)
    for f in ("U", "V", "HDRY", "THETA", "P3"):
        x, y = a.read(f), b.read(f)
        e = float(np.linalg.norm(x - y) / np.linalg.norm(x))
        assert e <= 1e-12, (f, e)
    s = CpuRef(sp)
    s.sync()
    s.step(3)
    z = CpuRef(sp, nc_path=tmp_path / "s0.nc")
    z.sync()
    z.step(3)
    for f in ea.FIELDS:
        assert s.read(f).tobytes() == z.read(f).tobytes(), f
    for r in (a, b, s, z):
        r.close()
