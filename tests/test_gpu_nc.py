"""Restart through epic.nc on the GPU: a member's state written by
epic_engine_write_nc and taken up again (epic_nc_read + epic_engine_add, as
the reference's epic_main takes up var_read's state: HDRY formed again from
P3) continues bit for bit as the oracle built with the GPU's correctly
rounded exp/log/pow continues from the same file."""
import numpy as np
import pytest

import epic_amd as ea
from oracle.cpuref import CpuRef
from tests.gpu_util import bitwise_diff, say

pytestmark = pytest.mark.gpu

CASES = {
    # a gas giant with the sounding relaxation, the default sponge and drag
    "jupiter_sounding_32x16x10": ea.spec(ea.JUPITER_LIKE, 32, 16, 10, 60, 3, 0.5, 0.0, 0.0, 0.5, newt_cool=1,
                                         tau_drag=2e4, drag_v=1),
    # a terrestrial planet: gz_surface travels in the file
    "held_suarez_32x16x8": ea.spec(ea.HELD_SUAREZ, 32, 16, 8, 600, 4, 0.5, 0.1, 0.3, 0.3),
}


@pytest.mark.parametrize("name", sorted(CASES))
def test_restart_from_epic_nc_matches_the_oracle(name, tmp_path):
    sp = CASES[name]
    say("case", name)
    e = ea.Engine(ea.FP64, 2)
    s = e.add(sp)
    e.step(5)
    path = tmp_path / "epic.nc"
    e.write_nc(s, path)
    c = ea.nc_read(path)
    try:
        assert c.model_time == 5 * sp.dt
        s2 = e.add_case(c)
        # the winds and the history arrive unchanged (p3 and the sigma
        # region's theta are formed again by the sync call from HDRY, as in
        # the reference, and are compared with the oracle below)
        for f in ("U", "V"):
            assert e.read(s2, f).tobytes() == e.read(s, f).tobytes(), f
        it = e.time_planes(s)
        inner = (slice(0, sp.nk + 1), slice(1, sp.nj + 2), slice(1, sp.ni + 1))  # what the file holds
        for w in ("U", "V"):
            for p in (1, 2):
                a, b = e.read(s2, f"D{w}DT{p}"), e.read(s, f"D{w}DT{it[p]}")
                assert a[inner].tobytes() == b[inner].tobytes(), (w, p)
        ref = CpuRef(sp, nudged="cr", nc_path=path)
        ref.sync()
        bad = bitwise_diff(lambda f: e.read(s2, f), ref.read)
        assert not bad, ("after the sync call", bad)
        e.step_members([s2], 4)
        ref.step(4)
        bad = bitwise_diff(lambda f: e.read(s2, f), ref.read)
        say(name, "restart + 4 steps, fields differing from the oracle:", bad)
        assert not bad
        assert e.time_planes(s2)[:3] == ref.time_planes()[:3]
        ref.close()
    finally:
        ea.free_case(c)
        e.close()


def test_mixed_member_writes_the_start_up_flags_as_float_max(tmp_path):
    """A mixed-precision member holds its Adams-Bashforth history in FP32,
    the start-up flag as FP32's largest value; in the file the flag is the
    reference's FLOAT_MAX, and a case read from it starts up as a new one."""
    sp = CASES["held_suarez_32x16x8"]
    e = ea.Engine(ea.MIXED, 2)
    s = e.add(sp)
    path = tmp_path / "m.nc"
    e.write_nc(s, path)
    c = ea.nc_read(path)
    try:
        n3 = c.Nelem3d
        h = c.array("dudt_hist", 2 * n3).reshape(2, sp.nk + 2, sp.nj + 3, sp.ni + 2)
        assert np.all(h[:, 1:sp.nk + 1, 1:sp.nj + 2, 1:sp.ni + 1] == np.finfo(np.float64).max)
        s2 = e.add_case(c)
        e.step(3)
        for f in ("U", "V", "THETA", "P3"):
            assert e.read(s2, f).tobytes() == e.read(s, f).tobytes(), f
    finally:
        ea.free_case(c)
        e.close()
