"""Engine behaviour on the GPU: a member's result does not depend on what
it is batched with, members come and go, the HBM budget is enforced."""
import numpy as np
import pytest

import epic_amd as ea
from tests.gpu_util import PROG, say

pytestmark = pytest.mark.gpu

A = ea.spec(ea.HELD_SUAREZ, 64, 32, 10, 600, 1, 0.5, 0.1, 0.1, 0.3)
B = ea.spec(ea.JUPITER_LIKE, 32, 16, 8, 60, 2, 0.5, 0.0, 0.0, 0.5)
C = ea.spec(ea.HELD_SUAREZ, 64, 32, 10, 600, 3, 0.3, 0.0, 0.2, 0.1)


def run_alone(sp, n):
    e = ea.Engine(ea.FP64, 1)
    s = e.add(sp)
    e.step(n)
    out = {f: e.read(s, f) for f in PROG + ("PV3", "GZ3")}
    e.close()
    return out


def test_member_alone_equals_member_in_mixed_batch():
    say("alone vs batch")
    n = 7
    ref = {k: run_alone(sp, n) for k, sp in (("A", A), ("B", B), ("C", C))}
    e = ea.Engine(ea.FP64, 4)
    slots = {"A": e.add(A), "B": e.add(B), "C": e.add(C)}
    e.step(n)
    for k, s in slots.items():
        for f, a in ref[k].items():
            assert np.array_equal(e.read(s, f), a), (k, f)
    e.close()


def test_join_leave_mid_run():
    """A member that joins late evolves exactly as if it had run alone."""
    e = ea.Engine(ea.FP64, 3)
    sa = e.add(A)
    e.step(3)
    sc = e.add(C)
    e.step(4)
    e.remove(sa)
    e.step(2)
    assert e.steps_done(sc) == 6
    ref = run_alone(C, 6)
    for f, a in ref.items():
        assert np.array_equal(e.read(sc, f), a), f
    e.close()


def test_step_members_subset():
    e = ea.Engine(ea.FP64, 2)
    sa, sc = e.add(A), e.add(C)
    e.step_members([sc], 5)
    assert e.steps_done(sa) == 0 and e.steps_done(sc) == 5
    ref = run_alone(C, 5)
    assert np.array_equal(e.read(sc, "THETA"), ref["THETA"])
    e.close()


@pytest.mark.parametrize("precision", [ea.FP64, ea.MIXED])
def test_step_shape_steps_one_shape_with_the_same_bits(precision):
    """Stepping one grid shape at a time (its own graph, captured on first
    use and again after a member leaves) gives each member the bits it
    gets from the whole engine's steps."""
    say("step_shape", precision)
    e = ea.Engine(precision, 4)
    sa, sb, sc = e.add(A), e.add(B), e.add(C)
    e.step_shape(64, 32, 10, 5)      # A and C
    e.step_shape(32, 16, 8, 3)       # B
    e.step_shape(64, 32, 10, 2)
    assert [e.steps_done(s) for s in (sa, sb, sc)] == [7, 3, 7]
    with pytest.raises(RuntimeError):
        e.step_shape(16, 8, 4, 1)
    r = ea.Engine(precision, 4)
    ra, rb, rc = r.add(A), r.add(B), r.add(C)
    r.step(3)
    r.step_members([ra, rc], 4)
    for k, (s, t) in {"A": (sa, ra), "B": (sb, rb), "C": (sc, rc)}.items():
        for f in PROG:
            assert e.read(s, f).tobytes() == r.read(t, f).tobytes(), (k, f)
    # after a member leaves, the shape's graph is captured again
    e.remove(sa)
    r.remove(ra)
    e.step_shape(64, 32, 10, 3)
    r.step_members([rc], 3)
    for f in PROG:
        assert e.read(sc, f).tobytes() == r.read(rc, f).tobytes(), f
    e.close()
    r.close()


def test_hbm_budget_is_enforced():
    fp = ea.footprint(64, 32, 10, ea.FP64)
    e = ea.Engine(ea.FP64, 4, hbm_budget=int(fp * 1.5))
    e.add(A)
    with pytest.raises(ea.EpicError):
        e.add(C)
    assert e.bytes_in_use() <= fp * 1.5
    e.close()


def test_write_then_read_round_trip():
    e = ea.Engine(ea.FP64, 1)
    s = e.add(A)
    th = e.read(s, "THETA")
    th2 = th.copy()
    th2[5, 10, 10] += 0.25
    e.write_field(s, "THETA", th2)
    assert np.array_equal(e.read(s, "THETA"), th2)
    e.close()


def test_rejects_unsupported_shapes():
    e = ea.Engine(ea.FP64, 1)
    with pytest.raises(ea.EpicError):
        e.add(ea.spec(ea.HELD_SUAREZ, 2048, 32, 10))
    with pytest.raises(ea.EpicError):
        e.add(ea.spec(ea.HELD_SUAREZ, 64, 32, 80))
    e.close()


def test_autotune_changes_no_state_and_no_result():
    """Tuning the launch shapes leaves every member's state and step count
    as it was, and the tuned step gives the same bits as the built-in shapes
    (a 40-level member starts from the narrow column tiles)."""
    say("autotune")
    D = ea.spec(ea.JUPITER_LIKE, 64, 32, 40, 60, 4, 0.5, 0.0, 0.0, 0.5)
    ref = {k: run_alone(sp, 6) for k, sp in (("A", A), ("B", B), ("D", D))}
    e = ea.Engine(ea.FP64, 4)
    slots = {"A": e.add(A), "B": e.add(B), "D": e.add(D)}
    e.step(2)
    before = {k: {f: e.read(s, f) for f in PROG} for k, s in slots.items()}
    rep = e.autotune(1)
    say(rep)
    assert rep.count("\n") == 3 * 7  # three buckets, seven shapes each
    for k, s in slots.items():
        assert e.steps_done(s) == 2
        for f in PROG:
            assert e.read(s, f).tobytes() == before[k][f].tobytes(), (k, f)
    e.step(4)
    for k, s in slots.items():
        for f, a in ref[k].items():
            assert np.isfinite(a).all(), (k, f)
            assert e.read(s, f).tobytes() == a.tobytes(), (k, f)
    e.close()


@pytest.mark.parametrize("precision", [ea.FP64, ea.MIXED])
def test_every_column_tile_shape_gives_the_same_bits(precision):
    """Candidate shapes of the column-tile kernels (16-64 columns per block
    of 256, 512 or 1024 threads, or 4-12 columns with one thread per level),
    set by hand on one knob at a time, and 8-32 rows per column tile of the
    del^4 meridional second differences, step a Held-Suarez and a
    Jupiter-like member to the same bits as the built-in shapes."""
    members = (A, ea.spec(ea.JUPITER_LIKE, 64, 32, 20, 60, 5, 0.5, 0.0, 0.0, 0.5))

    def run(knob=None, v=0):
        e = ea.Engine(precision, 2)
        sl = [e.add(sp) for sp in members]
        if knob is not None:
            e.set_shape(knob, v)
        e.step(3)
        out = [{f: e.read(s, f) for f in PROG + ("GZ3", "W3")} for s in sl]
        e.close()
        return out

    ref = run()
    bad = []
    for knob in ("cols_p2etc", "cols_gz3", "cols_calcw", "cols_hsuv"):
        for v in (16, 64, 2032, 2064, 4064, 10004, 10012):
            say(precision, knob, v)
            got = run(knob, v)
            for m, (g, r) in enumerate(zip(got, ref)):
                bad += [(knob, v, m, f) for f in r if g[f].tobytes() != r[f].tobytes()]
    # rows per column tile of the del^4 meridional second differences
    for v in (8, 16, 32):
        say(precision, "rows_d2", v)
        got = run("rows_d2", v)
        for m, (g, r) in enumerate(zip(got, ref)):
            bad += [("rows_d2", v, m, f) for f in r if g[f].tobytes() != r[f].tobytes()]
    assert not bad


def test_set_shape_refuses_what_is_not_a_candidate():
    e = ea.Engine(ea.FP64, 1)
    e.add(A)
    # not candidates; a knob that does not exist; 32 lanes per 10-level
    # column (more than twice its 12 points)
    for knob, v in (("cols_p2etc", 48), ("cols_gz3", 3016), ("rows_zhv", 32), (9, 8), ("cols_p2etc", 8)):
        with pytest.raises(ea.EpicError):
            e.set_shape(knob, v)
    e.set_shape("cols_gz3", 4064)
    e.set_shape("cols_gz3", 0)
    e.step(1)
    e.close()


@pytest.mark.parametrize("precision", [ea.FP64, ea.MIXED])
def test_checkpoint_restart_is_exact(tmp_path, precision):
    """A member restored from a checkpoint (state, Adams-Bashforth history,
    time planes) continues bit for bit as the original, also when it lands
    in a bucket beside another member."""
    say("checkpoint", precision)
    e = ea.Engine(precision, 2)
    s = e.add(B)
    e.step(9)
    fn = tmp_path / "member.ckpt"
    e.checkpoint(s, fn)
    e.step(6)
    want = {f: e.read(s, f) for f in PROG + ("PV3", "GZ3", "DUDT0", "DUDT1", "DUDT2")}
    tp = e.time_planes(s)
    e.close()
    r = ea.Engine(precision, 3)
    r.add(C)
    t = r.restore(B, fn)
    assert r.steps_done(t) == 9
    r.step(6)
    assert r.time_planes(t) == tp
    for f, a in want.items():
        assert r.read(t, f).tobytes() == a.tobytes(), f
    with pytest.raises(ea.EpicError):
        r.restore(A, fn)  # another grid
    r.close()
