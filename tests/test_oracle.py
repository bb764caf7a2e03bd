"""CPU tests that pin the oracle (oracle/cpuref.c) before it is trusted:
bit-for-bit against the reference's own FP64 build (oracle/_ref, where the
reference tree is present) and against the fixtures written from that build
(tests/golden), plus checks of its numerical kernels against what the
mathematics fixes (numpy's real FFT, scipy's PCHIP)."""
import ctypes
import hashlib
import json
import os
import subprocess

import numpy as np
import pytest

import epic_amd as ea
from oracle import cpuref as cr

HERE = os.path.dirname(os.path.abspath(__file__))
GOLD = os.path.join(HERE, "golden")
REF_EXE = os.path.join(os.path.dirname(HERE), "oracle", "_ref", "epic_ref")
# the reference's step built with its Arakawa-Lamb PV averaging
REF_AL81_EXE = os.path.join(os.path.dirname(HERE), "oracle", "_ref", "epic_ref_al81")
GJ = json.load(open(os.path.join(GOLD, "golden.json")))


def spec_of(name):
    """The case spec a golden fixture was written from (its positional
    values and the reference switches it set)."""
    from tests.golden.make_golden import spec_of as golden_spec
    return golden_spec(GJ["cases"][name])


def same_bits(a, b):
    return a.tobytes() == b.tobytes()


@pytest.mark.parametrize("name", ["hs_tiny", "jup_tiny", "venus_tiny", "jup_drag_tiny", "hs_adjust_tiny", "jup_del8_tiny"])
def test_oracle_matches_golden_fields(name):
    z = np.load(os.path.join(GOLD, f"{name}.npz"))
    r = cr.CpuRef(spec_of(name))
    r.sync()
    done = 0
    for step in GJ["cases"][name]["dumps"]:
        r.step(step - done)
        done = step
        for nm in ("U", "V", "HDRY", "THETA", "P3", "GZ3", "PV3", "W3"):
            assert same_bits(r.read(nm), z[f"s{step}_{nm}"]), (name, step, nm)
        h = GJ["hashes"][f"{name}/step{step}"]
        for nm in ea.FIELDS:
            assert hashlib.sha256(r.read(nm).tobytes()).hexdigest() == h[nm], (name, step, nm)
        assert r.time_planes()[:3] == h["time_planes"]
    assert r.error() == 0


def test_oracle_matches_golden_c1_100_steps():
    """Held-Suarez 64x32x10, dt = 600 s: every field bit for bit at steps 1, 10, 100."""
    r = cr.CpuRef(spec_of("hs_c1"))
    r.sync()
    done = 0
    for step in GJ["cases"]["hs_c1"]["dumps"]:
        r.step(step - done)
        done = step
        h = GJ["hashes"][f"hs_c1/step{step}"]
        for nm in ea.FIELDS:
            assert hashlib.sha256(r.read(nm).tobytes()).hexdigest() == h[nm], (step, nm)


def _ref_run(args, dumps, tmp_path, opts=None):
    fn = str(tmp_path / "ref.bin")
    exe = REF_AL81_EXE if (opts or {}).get("pv_scheme") else REF_EXE
    cmd = ([exe] + [f"--{k}={v}" for k, v in (opts or {}).items()] + [str(a) for a in args] +
           [str(max(dumps)), fn] + [str(d) for d in dumps])
    rc = subprocess.run(cmd, cwd=str(tmp_path)).returncode
    assert rc == 0, "reference run failed or its metric arrays (t_sponge_inv included) differ from the case builder's"
    from tests.golden.make_golden import read_dump
    return read_dump(fn)


# (positional spec values, reference switches).  k_sponge -1 is the
# reference `initial` program's choice (src/shared/epic_initial.c:1249-1266):
# the reference's own set_sponge() then fills t_sponge_inv in the harness.
REF_CASES = [
    ((0, 32, 20, 8, 900, 11, 0.5, 0.15, 0.3, 0.4), {}),           # HS with divergence damping
    ((1, 32, 16, 10, 60, 12, 0.5, 0.0, 0.0, 1.0), {"k_sponge": 0}),  # gas giant: gz_from_u every call
    ((2, 64, 24, 12, 60, 13, 0.4, 0.1, 0.0, 0.5), {"k_sponge": 0}),  # Saturn-like
    ((3, 16, 12, 5, 600, 14, 0.5, 0.0, 0.2, 0.2), {"k_sponge": 0}),  # slow rotator, nk = 5
    # the top sponge as `initial` sets it (2 layers at nk = 10)
    ((1, 32, 16, 10, 60, 15, 0.5, 0.0, 0.0, 1.0), {"k_sponge": -1}),
    ((2, 32, 12, 20, 60, 16, 0.4, 0.0, 0.0, 0.5), {"k_sponge": -1}),   # 4 layers
    # Rayleigh drag: u only, pointwise; u and v about the zonal mean
    ((1, 32, 16, 10, 60, 17, 0.5, 0.0, 0.0, 1.0), {"k_sponge": 1, "tau_drag": 1.0e5}),
    ((3, 16, 12, 6, 600, 18, 0.5, 0.0, 0.2, 0.5),
     {"k_sponge": -1, "tau_drag": 3.0e4, "drag_v": 1, "drag_zonal_avg": 1}),
    ((2, 32, 12, 8, 60, 19, 0.4, 0.0, 0.0, 0.5), {"k_sponge": 2, "tau_drag": 5.0e4, "drag_v": 1}),
    # Held-Suarez: layer-mean removal of the heating; prandtl/t_rad drag
    ((0, 32, 16, 8, 600, 20, 0.5, 0.0, 0.3, 0.4), {"newt_cool_adjust": 1}),
    ((0, 32, 16, 8, 600, 21, 0.5, 0.1, 0.3, 0.4), {"prandtl": 0.5, "drag_v": 1}),
    ((0, 32, 16, 8, 600, 22, 0.5, 0.0, 0.3, 0.4), {"prandtl": 0.2, "drag_zonal_avg": 1, "newt_cool_adjust": 1}),
    # higher-order hyperviscosity: del^8 alone (the reference `initial`
    # program's default, nu_nondim[8] = 0.5), del^6 alone, and del^4, del^6
    # and del^8 together (each order filtered and applied in turn)
    ((1, 32, 16, 10, 60, 23, 0.0, 0.0, 0.0, 1.0), {"k_sponge": -1, "nu8_nondim": 0.5}),
    ((3, 16, 12, 6, 600, 24, 0.0, 0.1, 0.2, 0.5), {"k_sponge": -1, "nu6_nondim": 0.4}),
    ((0, 32, 16, 8, 600, 25, 0.3, 0.0, 0.3, 0.4), {"nu6_nondim": 0.2, "nu8_nondim": 0.5}),
    # Newtonian cooling toward a sounding with a tabulated cooling time
    # (get_sounding, t_rad's table branch), alone, with the layer-mean removal
    # and with the prandtl drag that reads the tabulated time
    ((1, 32, 16, 10, 60, 26, 0.5, 0.0, 0.0, 1.0), {"k_sponge": -1, "newt_cool": 1}),
    ((3, 16, 12, 6, 600, 27, 0.5, 0.0, 0.2, 0.5), {"newt_cool": 1, "newt_cool_adjust": 1}),
    ((2, 32, 12, 8, 60, 28, 0.4, 0.0, 0.0, 0.5),
     {"k_sponge": 2, "newt_cool": 1, "prandtl": 0.5, "drag_v": 1}),
    # the Akima advection scheme: for hdry and theta; for theta (and the
    # unity field of its correction) beside the predictor-corrector's hdry;
    # for hdry beside the predictor-corrector's theta (64-point rows)
    ((0, 32, 20, 8, 900, 31, 0.5, 0.1, 0.3, 0.4), {"adv_mass": 1, "adv_nonmass": 1}),
    ((1, 32, 16, 10, 60, 32, 0.5, 0.0, 0.0, 1.0), {"k_sponge": -1, "adv_nonmass": 1}),
    ((2, 64, 24, 12, 60, 33, 0.4, 0.0, 0.0, 0.5), {"k_sponge": 0, "adv_mass": 1}),
    # the Arakawa-Lamb PV averaging (the reference built with that
    # PV_SCHEME): Held-Suarez; a gas giant, where gz_from_u takes the
    # weights and get_kin too
    ((0, 32, 20, 8, 900, 34, 0.5, 0.1, 0.3, 0.4), {"pv_scheme": 1}),
    ((1, 32, 16, 10, 60, 35, 0.5, 0.0, 0.0, 1.0), {"k_sponge": -1, "pv_scheme": 1}),
    ((2, 64, 24, 12, 60, 36, 0.4, 0.0, 0.0, 0.5), {"k_sponge": 0, "pv_scheme": 1, "adv_nonmass": 1}),
]


@pytest.mark.skipif(not (os.path.exists(REF_EXE) and os.path.exists(REF_AL81_EXE)),
                    reason="reference build (oracle/_ref) not present")
def test_reference_pv_schemes_differ(tmp_path):
    """The two reference builds average PV differently (so the Arakawa-Lamb
    cases above pin that choice), and each refuses the other's scheme."""
    args = (1, 32, 16, 10, 60, 35, 0.5, 0.0, 0.0, 1.0)
    a = _ref_run(args, [3], tmp_path, {"k_sponge": -1})
    b = _ref_run(args, [3], tmp_path, {"k_sponge": -1, "pv_scheme": 1})
    for nm in ("U", "V", "GZ_SURFACE"):
        x, y = a[0][1][nm], b[0][1][nm]
        assert np.abs(x - y).max() > 1e-9 * np.abs(x).max(), nm
    for exe, pv in ((REF_EXE, 1), (REF_AL81_EXE, 0)):
        cmd = [exe, f"--pv_scheme={pv}"] + [str(v) for v in args] + ["1", str(tmp_path / "x.bin"), "1"]
        assert subprocess.run(cmd, cwd=str(tmp_path), capture_output=True).returncode == 2


def case_spec(args, opts):
    sp = ea.spec(*args, k_sponge=opts.get("k_sponge", 0))
    for k, v in opts.items():
        if k != "k_sponge":
            setattr(sp, k, v)
    return sp


@pytest.mark.skipif(not (os.path.exists(REF_EXE) and os.path.exists(REF_AL81_EXE)),
                    reason="reference build (oracle/_ref) not present")
@pytest.mark.parametrize("args,opts", REF_CASES, ids=[f"{a[:4]}{o}" for a, o in REF_CASES])
def test_oracle_bitwise_vs_reference_build(args, opts, tmp_path):
    dumps = [0, 1, 2, 3, 7]
    ref = _ref_run(args, dumps, tmp_path, opts)
    r = cr.CpuRef(case_spec(args, opts))
    r.sync()
    done = 0
    for hdr, f in ref:
        step = int(hdr[0])
        r.step(step - done)
        done = step
        for nm in ea.FIELDS:
            assert same_bits(r.read(nm), f[nm]), (args, step, nm)
        assert r.time_planes()[:3] == [int(x) for x in hdr[1:4]]


@pytest.mark.skipif(not os.path.exists(REF_EXE), reason="reference build (oracle/_ref) not present")
@pytest.mark.parametrize("args,opts", [
    ((0, 32, 16, 8, 600, 31, 0.5, 0.1, 0.3, 0.4), {}),
    ((1, 32, 16, 10, 60, 32, 0.5, 0.0, 0.0, 1.0), {"k_sponge": -1}),
    ((2, 32, 12, 12, 60, 33, 0.4, 0.0, 0.0, 0.5), {"k_sponge": 2}),
], ids=["held_suarez", "jupiter", "saturn"])
def test_traditional_pressure_gradient_is_the_reference_function(args, opts, tmp_path):
    """The reference's build compiles uv_pgrad_traditional
    (src/shared/epic_timestep.c:806-884) beside the finite-volume pressure
    gradient it calls (its PGRAD_TYPE is fixed in the source), so the
    traditional form is checked as a function: after 5 steps of the build,
    the reference's uv_pgrad_traditional alone on the final state equals the
    oracle's on the same state, bit for bit, in both tendency planes."""
    ref = _ref_run(args, [5], tmp_path, dict(opts, pgrad_traditional=1))
    (h5, f5), (hp, fp) = ref
    assert int(h5[0]) == 5 and int(hp[0]) == -1
    r = cr.CpuRef(case_spec(args, opts))
    r.sync()
    r.step(5)
    it0 = r.time_planes()[0]
    assert it0 == int(hp[1])
    for nm in ("U", "V", "GZ3", "MONT3"):
        assert same_bits(r.read(nm), f5[nm]), nm
    r.pgrad_traditional()
    for w in ("DUDT", "DVDT"):
        got, want = r.read(f"{w}{it0}"), fp[f"{w}{it0}"]
        assert same_bits(got, want), w
        assert np.count_nonzero(want) > 0
    r.close()


@pytest.mark.skipif(not os.path.exists(REF_EXE), reason="reference build (oracle/_ref) not present")
@pytest.mark.parametrize("args", [(1, 32, 16, 10, 60, 5, 0.5, 0.0, 0.0, 1.0), (0, 32, 16, 8, 600, 5, 0.5, 0.0, 0.3, 0.3)],
                         ids=["jupiter", "held_suarez"])
def test_reference_leapfrog_does_not_survive_its_start_up_flags(args, tmp_path):
    """Why the GPU step refuses "Leapfrog (Asselin filtered)": the reference's
    own build, started as its `initial` program starts that scheme (U and V
    of IT_MINUS1 flagged FLOAT_MAX on K = 1..nk), leaves flags in the
    current plane after its first (forward) step -- the bottom interface
    K = nk, which no step updates, and the v pole row; they spread through
    the second step, and the third stops with a NaN pressure (calc_w's
    check)."""
    cmd = [REF_EXE, "--leapfrog=1"] + [str(a) for a in args] + ["4", str(tmp_path / "lf.bin")]
    r = subprocess.run(cmd, cwd=str(tmp_path), capture_output=True, text=True, timeout=120)
    out = r.stdout + r.stderr
    bad = [int(out.split(f"leapfrog step {s}: ")[1].split()[0]) for s in (1, 2)]
    assert 0 < bad[0] < bad[1], bad
    assert r.returncode != 0 and "nan" in out.lower() and "leapfrog step 4:" not in out


def test_richardson_of_given_fields_reproduces_store_diag():
    """The function form the GPU tests use for RI2 (oracle fields written in,
    RI2 formed again) gives store_diag's RI2 bit for bit, and the golden
    hashes above pin store_diag's RI2 to the reference build's."""
    sp = spec_of("jup_tiny")
    r = cr.CpuRef(sp)
    r.sync()
    r.step(3)
    f = {k: r.read(k) for k in ("U", "V", "P2", "T2", "RHO2", "P3", "RHO3", "RI2")}
    fn = cr.CpuRef(sp)
    assert same_bits(fn.richardson_of(f), f["RI2"])
    # a changed shear changes RI2 where it enters
    f["U"][2, 4, 5] += 1.0
    assert not same_bits(fn.richardson_of(f), f["RI2"])


def test_hot_jupiter_relaxation_is_the_stated_closed_form():
    """The day-night relaxation has no reference counterpart, so it is pinned
    to its definition, evaluated here independently (lat/lon of the h-points
    from the grid spacing, the reference temperature through exp/log):
    HEAT3 = cp (T_eq - T)/tau, T_eq = T_ref(p) + dT(p)(max(0, cos lat cos lon)
    - 1/4), dT = 1000 K for p <= 1 bar, linear in ln p to 0 at 10 bar.
    With a top sponge of k layers the layers K <= k lose their area-weighted
    mean (the cell area 1/mn), as the reference's newtonian_cooling does
    (src/shared/epic_sensible_heating.c:139-167)."""
    ni, nj, nk = 32, 16, 8
    r = cr.CpuRef(ea.spec(ea.HOT_JUPITER, ni, nj, nk, 30, 3, 0.5, 0.0, 1.0, 0.0, k_sponge=0))
    r.sync()
    heat, t3, p3 = r.read("HEAT3"), r.read("T3"), r.read("P3")
    rs = cr.CpuRef(ea.spec(ea.HOT_JUPITER, ni, nj, nk, 30, 3, 0.5, 0.0, 1.0, 0.0, k_sponge=3))
    rs.sync()
    heat_s = rs.read("HEAT3")
    cp, rgas, g, n2, thb, p0, pb = 14300.0, 3523.0, 9.42, 1.5e-6, 1500.0, 1.0e5, 1.0e6
    kap = rgas / cp
    aok = n2 * rgas / g ** 2 / kap
    dlt = 180.0 / ((nj + 1) + 2 * np.sqrt(0.5))
    lat = np.radians(-90.0 + dlt * np.sqrt(0.5) + (np.arange(nj + 1) + 0.5) * dlt)
    lon = np.radians(-180.0 + (np.arange(1, ni + 1) - 0.5) * 360.0 / ni)
    p, t = p3[1:nk + 1, 1:nj + 2, 1:ni + 1], t3[1:nk + 1, 1:nj + 2, 1:ni + 1]
    lnx = kap * np.log(p / p0)
    tref = np.exp(lnx) / (1.0 / thb + aok * (np.exp(lnx) - np.exp(kap * np.log(pb / p0))))
    dT = 1000.0 * np.clip(np.log(pb / p) / np.log(pb / 1.0e5), 0.0, 1.0)
    day = np.maximum(0.0, np.cos(lat)[None, :, None] * np.cos(lon)[None, None, :])
    want = cp * (tref + dT * (day - 0.25) - t) / 1.0e5
    got = heat[1:nk + 1, 1:nj + 2, 1:ni + 1]
    assert np.abs(got - want).max() <= 1e-9 * np.abs(want).max()
    # the sponge layers 1..3 lose their area-weighted mean, the rest do not.
    # Cell area 1/mn (set_fmn on a sphere): re^2 cos(lat) dlon dlat on the
    # interior rows; on the two pole rows n is divided by 1 + sqrt(1/2) and
    # mn = n^2 / (dlon/2), i.e. area = re^2 dlon dlat^2 (1 + sqrt(1/2))^2 / 2
    dlnr, dltr = np.radians(360.0 / ni), np.radians(dlt)
    w = np.cos(lat) * dlnr * dltr
    w[0] = w[-1] = 0.5 * dlnr * dltr ** 2 * (1.0 + np.sqrt(0.5)) ** 2
    diff = heat[1:nk + 1, 1:nj + 2, 1:ni + 1] - heat_s[1:nk + 1, 1:nj + 2, 1:ni + 1]
    for k in range(nk):
        if k + 1 <= 3:
            havg = (want[k] * w[:, None]).sum() / (w.sum() * ni)
            assert np.abs(diff[k] - havg).max() <= 1e-9 * np.abs(want[k]).max(), k
            assert abs(havg) > 1e-6 * np.abs(want[k]).max()
        else:
            assert not diff[k].any()
    # the pads are the zonal periodic copies, the top interface is not heated
    assert np.array_equal(heat[:, :, 0], heat[:, :, ni]) and np.array_equal(heat[:, :, ni + 1], heat[:, :, 1])
    assert not heat[0].any()


# ---------------------------------------------------- numerical pins
def test_realft_is_the_real_dft_in_nr_packing():
    L = cr.lib()
    rng = np.random.default_rng(3)
    for n in (8, 64, 256, 1024):
        x = rng.standard_normal(n)
        d = np.concatenate([[0.], x])
        L.cpuref_realft(d.ctypes.data, n, 1)
        X = np.fft.rfft(x)
        # NR forward uses exp(+i...): its bins are the conjugates of numpy's
        got_re = np.concatenate([[d[1]], d[3::2], [d[2]]])
        got_im = np.concatenate([[0.], d[4::2], [0.]])
        assert np.allclose(got_re, X.real, atol=1e-12 * n)
        assert np.allclose(got_im, -X.imag, atol=1e-12 * n)
        L.cpuref_realft(d.ctypes.data, n, -1)
        assert np.allclose(d[1:] * 2. / n, x, atol=1e-13)


def test_pchip_matches_scipy_and_is_monotone():
    from scipy.interpolate import PchipInterpolator
    L = cr.lib()
    rng = np.random.default_rng(5)
    for n in (3, 5, 12, 40):
        x = np.cumsum(rng.uniform(0.1, 2.0, n))
        y = np.cumsum(rng.uniform(-1, 2, n))
        dd = np.zeros(n)
        assert L.cpuref_pchip(n, x.ctypes.data, y.ctypes.data, dd.ctypes.data) == 0
        ref = PchipInterpolator(x, y)
        for xx in np.linspace(x[0], x[-1], 57):
            v = L.cpuref_pchip_eval(n, x.ctypes.data, y.ctypes.data, dd.ctypes.data, xx)
            assert abs(v - ref(xx)) <= 1e-12 * max(1., abs(ref(xx)))
    # monotone data give a monotone interpolant
    x = np.arange(10.); y = np.array([0, 0, 1, 1, 1, 5, 5, 6, 9, 9.])
    dd = np.zeros(10)
    L.cpuref_pchip(10, x.ctypes.data, y.ctypes.data, dd.ctypes.data)
    v = [L.cpuref_pchip_eval(10, x.ctypes.data, y.ctypes.data, dd.ctypes.data, t) for t in np.linspace(0, 9, 400)]
    assert np.all(np.diff(v) >= -1e-15)


def test_fcmp_is_a_one_ulp_binade_comparison():
    L = cr.lib()
    assert L.cpuref_fcmp(1.0, 1.0 + 2.2e-16) == 0
    assert L.cpuref_fcmp(1.0, 1.0 + 1e-15) == -1
    assert L.cpuref_fcmp(0.0, 0.0) == 0
    assert L.cpuref_fcmp(1e-300, 0.0) == 1
    assert L.cpuref_fcmp(-3.0, 2.0) == -1


def test_zonal_filter_gain_keeps_the_zonal_mean():
    """The polar filter multiplies bin 0 by 1: a filtered row keeps its mean."""
    L = cr.lib()
    n = 64
    x = np.random.default_rng(0).standard_normal(n)
    d = np.concatenate([[0.], x])
    L.cpuref_realft(d.ctypes.data, n, 1)
    d[3:] *= 0.3  # damp every non-DC bin
    d[2] *= 0.3
    L.cpuref_realft(d.ctypes.data, n, -1)
    y = d[1:] * 2. / n
    assert abs(y.mean() - x.mean()) < 1e-14


def test_libm_sensitivity_floor_of_the_step():
    """The nudged oracle (10% of exp/log/pow results moved by one ulp) stays
    within a few ulps on h, theta, p but carries the perturbation into u, v
    at 1e-12..1e-10 relative over 100 steps of Held-Suarez 64x32x10: the
    reference algorithm's own floor for any FP64 implementation whose
    transcendentals are not bit-identical to the host libm's."""
    sp = spec_of("hs_c1")
    a, b = cr.CpuRef(sp), cr.CpuRef(sp, nudged=True)
    a.sync(); b.sync()
    a.step(100); b.step(100)
    e = {f: float(np.linalg.norm(a.read(f) - b.read(f)) / np.linalg.norm(a.read(f)))
         for f in ("U", "V", "HDRY", "THETA", "P3")}
    print(e)
    assert 0 < e["U"] < 1e-10 and 0 < e["V"] < 1e-10
    assert e["HDRY"] < 1e-12 and e["THETA"] < 1e-12 and e["P3"] < 1e-12
