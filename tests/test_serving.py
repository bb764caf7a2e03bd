"""CPU tests of the request loop's scheduling: admission by HBM footprint
and free slots, continuous batching (work joins while other work is on
the device), retirement, latency bookkeeping.  The engine is a host-side
stand-in with the same methods; the GPU test runs the real one."""
import time

import numpy as np

import epic_amd as ea
from epic_amd import serving as sv


class FakeEngine:
    """Advances step counters only; the clock charges each step by the
    resident grid points."""

    def __init__(self, clock):
        self.clock = clock
        self.members = {}
        self.next = 0
        self.max_resident = 0
        self.joined_while_busy = 0
        self.shape_ticks = 0

    def add(self, sp):
        if self.members:
            self.joined_while_busy += 1
        s = self.next
        self.next += 1
        self.members[s] = [sp, 0]
        self.max_resident = max(self.max_resident, len(self.members))
        return s

    def step(self, n):
        pts = sum(sp.ni * (sp.nj + 1) * sp.nk for sp, _ in self.members.values())
        for m in self.members.values():
            m[1] += n
        self.clock.t += n * pts * 1e-10

    def step_shape(self, ni, nj, nk, n):
        hit = [m for m in self.members.values() if (m[0].ni, m[0].nj, m[0].nk) == (ni, nj, nk)]
        assert hit, "no member of that shape"
        for m in hit:
            m[1] += n
        self.clock.t += n * sum(sp.ni * (sp.nj + 1) * sp.nk for sp, _ in hit) * 1e-10
        self.shape_ticks += 1

    def wait(self):
        pass

    def member_error(self, slot):
        return 0

    def read(self, slot, f):
        sp, n = self.members[slot]
        return np.full((sp.nk + 2, sp.nj + 3, sp.ni + 2), float(n))

    def remove(self, slot):
        del self.members[slot]


class Clock:
    def __init__(self):
        self.t = 0.0

    def __call__(self):
        return self.t

    def sleep(self, dt):
        self.t += dt


def footprint(r):
    return (r.ni + 2) * (r.nj + 3) * (r.nk + 2) * 8 * 36


def test_admission_respects_budget_and_everything_completes():
    reqs = sv.poisson_mix(40, rate=20.0, seed=3, steps=(20, 200))
    clk = Clock()
    eng = FakeEngine(clk)
    budget = 3 * footprint(sv.Request(0, 0, 512, 256, 32, 1, 0, 0.))
    srv = sv.Server(eng, budget, max_members=8, chunk=25, footprint=footprint, clock=clk, sleep=clk.sleep)
    peak = [0]
    orig_add = eng.add

    def add(sp):
        slot = orig_add(sp)
        peak[0] = max(peak[0], srv.in_use + footprint(sv.Request(0, 0, sp.ni, sp.nj, sp.nk, 1, 0, 0.)))
        return slot

    eng.add = add
    res = srv.run(reqs)
    assert len(res) == 40 and all(r.status == "ok" for r in res)
    assert peak[0] <= budget
    assert eng.max_resident <= 8
    assert eng.joined_while_busy > 0  # continuous batching, not a queue
    for r in res:
        req = reqs[r.rid]
        assert r.steps == req.steps and r.checksum == req.steps * req.ni * (req.nj + 1) * req.nk
        assert r.admitted >= r.arrival and r.finished >= r.admitted
    summ = sv.summarize(res, srv.elapsed, srv.point_steps)
    assert summ["p99_latency_s"] >= summ["p50_latency_s"] > 0
    assert srv.point_steps == sum(r.steps * r.points for r in reqs)


def test_oversized_request_is_rejected_not_waited_on():
    clk = Clock()
    srv = sv.Server(FakeEngine(clk), 10 * 2 ** 20, 4, footprint=footprint, clock=clk, sleep=clk.sleep)
    reqs = [sv.Request(0, 0, 512, 256, 32, 10, 1, 0.0), sv.Request(1, 0, 64, 32, 10, 10, 2, 0.1)]
    res = srv.run(reqs)
    st = {r.rid: r.status for r in res}
    assert st[0].startswith("rejected") and st[1] == "ok"


def test_poisson_mix_is_seeded():
    a = sv.poisson_mix(10, 5.0, 7)
    b = sv.poisson_mix(10, 5.0, 7)
    assert a == b and all(x.arrival < y.arrival for x, y in zip(a, a[1:]))


class FakeCaseEngine(FakeEngine):
    """Takes cases built ahead (on the server's worker thread) as the real
    engine's add_case does."""

    def __init__(self, clock):
        super().__init__(clock)
        self.from_cases = 0

    def step(self, n):
        super().step(n)
        time.sleep(n * 1e-3)  # device time passes while the worker builds

    def step_shape(self, ni, nj, nk, n):
        super().step_shape(ni, nj, nk, n)
        time.sleep(n * 1e-3)

    def add_case(self, c, spec=None):
        assert (c.ni, c.nj, c.nk, c.dt) == (spec.ni, spec.nj, spec.nk, spec.dt)
        self.from_cases += 1
        return self.add(spec)


def test_cases_built_ahead_join_while_others_step():
    reqs = sv.poisson_mix(12, rate=5000.0, seed=5, grids=[(64, 32, 10), (128, 64, 16)], steps=(20, 200))
    clk = Clock()
    eng = FakeCaseEngine(clk)
    srv = sv.Server(eng, 1 << 40, max_members=4, chunk=25, footprint=footprint, clock=clk, sleep=clk.sleep)
    res = srv.run(reqs)
    assert len(res) == 12 and all(r.status == "ok" for r in res)
    assert eng.from_cases == 12 and not srv.built and not srv.pending
    assert eng.max_resident <= 4 and eng.joined_while_busy > 0
    assert srv.point_steps == sum(r.steps * r.points for r in reqs)
    for r in res:
        req = reqs[r.rid]
        assert r.steps == req.steps and r.checksum == req.steps * req.ni * (req.nj + 1) * req.nk


def test_arrival_joins_within_a_step_not_a_chunk():
    clk = Clock()
    eng = FakeEngine(clk)
    first = sv.Request(0, ea.HELD_SUAREZ, 64, 32, 10, 100000, 1, 0.0)
    late = sv.Request(1, ea.HELD_SUAREZ, 64, 32, 10, 10, 2, 0.01)
    srv = sv.Server(eng, 1 << 40, max_members=4, chunk=500, footprint=footprint, clock=clk, sleep=clk.sleep)
    res = {r.rid: r for r in srv.run([first, late])}
    step = first.points * 1e-10  # the fake engine's time per step with one member
    assert 0 <= res[1].admitted - late.arrival <= 1.01 * step
    assert res[0].steps == first.steps and res[1].steps == late.steps


def _run(policy, reqs):
    clk = Clock()
    eng = FakeEngine(clk)
    srv = sv.Server(eng, 1 << 40, max_members=16, chunk=50, footprint=footprint, clock=clk, sleep=clk.sleep,
                    policy=policy)
    res = srv.run(reqs)
    return eng, srv, {r.rid: r for r in res}


def test_shortest_first_steps_one_shape_and_cuts_mean_latency():
    """Stepping the shape whose first member ends soonest finishes the same
    work (every request, all its steps, the same point-steps) with a lower
    mean latency than stepping every member together; the last one ends at
    about the same time (the device does the same work)."""
    reqs = sv.poisson_mix(24, rate=400.0, seed=11, steps=(50, 500))
    eng_s, srv_s, res_s = _run("shortest", reqs)
    eng_l, srv_l, res_l = _run("lockstep", reqs)
    assert eng_s.shape_ticks == srv_s.ticks > 0 and eng_l.shape_ticks == 0
    for res in (res_s, res_l):
        assert len(res) == 24 and all(r.status == "ok" for r in res.values())
        for r in res.values():
            req = reqs[r.rid]
            assert r.steps == req.steps and r.checksum == req.steps * req.ni * (req.nj + 1) * req.nk
    assert srv_s.point_steps == srv_l.point_steps == sum(r.steps * r.points for r in reqs)
    mean = lambda res: np.mean([r.latency for r in res.values()])
    assert mean(res_s) < 0.8 * mean(res_l)
    last = lambda res: max(r.finished for r in res.values())
    assert abs(last(res_s) - last(res_l)) <= 0.05 * last(res_l)


def test_short_request_of_a_small_shape_passes_a_long_large_one():
    big = sv.Request(0, ea.JUPITER_LIKE, 512, 256, 32, 2000, 1, 0.0)
    small = sv.Request(1, ea.HELD_SUAREZ, 64, 32, 10, 200, 2, 0.001)
    eng, srv, res = _run("shortest", [big, small])
    step_small = small.points * 1e-10
    # the small one waits at most about one step of the big one to join,
    # then runs alone: its latency is its own work, not the big one's
    assert res[1].latency <= 200 * step_small + 2 * big.points * 1e-10
    assert res[1].finished < res[0].finished
    assert res[0].steps == 2000 and res[1].steps == 200


def test_unknown_policy_is_refused():
    try:
        sv.Server(FakeEngine(Clock()), 1 << 30, 4, policy="fifo")
    except ValueError as e:
        assert "fifo" in str(e)
    else:
        raise AssertionError("accepted an unknown policy")
