"""CPU tests of the correctly rounded FP64 exp, log and pow of
epic_amd/csrc/crmath.h (the GPU's FP64 kernels use them; the oracle's
libcpuref_cr.so evaluates them from the same source with gcc), pinned
against mpmath's correctly rounded values, and of what the step's
sensitivity to last-bit differences means for parity with the host libm."""
import numpy as np
import pytest

import epic_amd as ea
from oracle import cpuref as cr

mpmath = pytest.importorskip("mpmath")
mpmath.mp.prec = 200


def cr_math(fn, x, y=0.0):
    return cr.lib("cr").cpuref_math(fn, x, y)


def glibc_math(fn, x, y=0.0):
    return cr.lib().cpuref_math(fn, x, y)


def exact(fn, x, y=0.0):
    """mpmath at 200 bits, rounded to the nearest double."""
    X = mpmath.mpf(x)
    if fn == 0:
        v = mpmath.exp(X)
    elif fn == 1:
        v = mpmath.log(X)
    else:
        v = mpmath.power(X, mpmath.mpf(y))
    return float(v)


def arguments(fn, n, rng):
    if fn == 0:
        x = np.concatenate([rng.uniform(-700, 700, n // 2), rng.uniform(-1e-3, 1e-3, n // 4),
                            rng.uniform(-60, 0, n - n // 2 - n // 4)])
        return [(float(v), 0.0) for v in x]
    if fn == 1:
        x = np.concatenate([10.0 ** rng.uniform(-300, 300, n // 2), 1.0 + rng.uniform(-2e-3, 2e-3, n // 4),
                            10.0 ** rng.uniform(1, 7, n - n // 2 - n // 4)])
        return [(float(v), 0.0) for v in x]
    # pow: pressure ratios to the powers the step takes (kappa, kappa + 1,
    # 1/kappa for a range of planets), and general exponents
    kap = rng.uniform(0.2, 0.3, n)
    y = np.where(rng.uniform(size=n) < 0.75, np.choose(rng.integers(0, 3, n), [kap, kap + 1.0, 1.0 / kap]),
                 rng.uniform(-4, 4, n))
    x = 10.0 ** rng.uniform(-4, 2, n)
    return [(float(a), float(b)) for a, b in zip(x, y)]


@pytest.mark.parametrize("fn,name", [(0, "exp"), (1, "log"), (2, "pow")])
def test_crmath_is_correctly_rounded(fn, name):
    rng = np.random.default_rng(100 + fn)
    args = arguments(fn, 6000, rng)
    wrong = [(x, y) for x, y in args if cr_math(fn, x, y) != exact(fn, x, y)]
    glibc = sum(glibc_math(fn, x, y) != exact(fn, x, y) for x, y in args)
    print(f"{name}: {len(wrong)} of {len(args)} not correctly rounded; glibc: {glibc}")
    assert not wrong, wrong[:5]


def test_crmath_special_values():
    assert cr_math(2, 1.0, 0.37) == 1.0 and cr_math(2, 3.1, 0.0) == 1.0
    assert cr_math(2, 4.0, 0.5) == 2.0 and cr_math(2, 2.0, 10.0) == 1024.0
    assert cr_math(1, 1.0) == 0.0 and cr_math(0, 0.0) == 1.0
    assert cr_math(1, 2.0 ** -1070) == exact(1, 2.0 ** -1070)  # subnormal argument
    assert np.isnan(cr_math(1, -1.0)) and cr_math(0, 800.0) == float("inf")


def test_one_ulp_grows_to_the_u_v_floor_of_held_suarez():
    """Why u and v cannot meet 1e-12 against the host-libm oracle after 100
    steps of Held-Suarez 64x32x10: one ulp of theta at one cell of the
    prognostic (isentropic) layers, and nothing else, moves u by more than
    1e-12 and v by more than 2e-12 (relative L2) after 100 steps; the
    correctly rounded exp, log and pow (each alone, or all three) move them
    by the same amount, since glibc rounds about 5e-4 of the step's calls
    differently from the correct rounding."""
    sp = ea.spec(ea.HELD_SUAREZ, 64, 32, 10, 600, 1, 0.5, 0.0, 0.1, 0.0, k_sponge=0)
    a = cr.CpuRef(sp)
    a.sync()
    a.step(100)
    want = {f: a.read(f) for f in ("U", "V", "HDRY", "THETA", "P3")}

    def dist(r):
        return {f: float(np.linalg.norm(r.read(f) - want[f]) / np.linalg.norm(want[f])) for f in want}

    b = cr.CpuRef(sp)
    b.sync()
    th = b.read("THETA")
    th[9, 3, 40] = np.nextafter(th[9, 3, 40], np.inf)
    b.write("THETA", th)
    b.step(100)
    e_ulp = dist(b)
    L = cr.lib("cr")
    cr.libm_stats("cr")
    e_fn = {}
    try:
        for mask, name in ((1, "exp"), (2, "log"), (4, "pow"), (7, "all")):
            L.cpuref_cr_select(mask)
            c = cr.CpuRef(sp, nudged="cr")
            c.sync()
            c.step(100)
            e_fn[name] = dist(c)
            c.close()
    finally:
        L.cpuref_cr_select(7)
    calls, diffs = cr.libm_stats("cr")
    print("one ulp:", e_ulp)
    print("correctly rounded:", e_fn)
    print("calls", calls, "glibc differs", diffs)
    assert e_ulp["U"] > 1e-12 and e_ulp["V"] > 2e-12
    for name, e in e_fn.items():
        assert 0.3 < e["U"] / e_ulp["U"] < 3 and 0.3 < e["V"] / e_ulp["V"] < 3, name
        assert e["HDRY"] < 1e-12 and e["THETA"] < 1e-12 and e["P3"] < 1e-12, name
    for q in range(3):
        assert 1e-4 < diffs[q] / calls[q] < 3e-3
