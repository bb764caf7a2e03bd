"""The ensemble and serving configurations at their stated sizes.

Ensemble: 64 members of the Jupiter-like planet at 128x64x16 (dt = 60 s),
each with its own seed (noise and jet coefficients), advanced 500 steps in
one engine in the mixed precision mode: members 0, 31 and 63 equal the same
member run alone bit for bit, and every member's h, theta and p stay within
1e-5 relative L2 of the FP64 run of the same ensemble.

Serving: a mix that includes 256x128x20 and 512x256x32 integrations of 200
steps and more, arriving while others run; each result equals the same
request served alone bit for bit."""
import numpy as np
import pytest

import epic_amd as ea
from epic_amd import serving as sv
from tests.gpu_util import say

pytestmark = pytest.mark.gpu

NM, NI, NJ, NK, DT, STEPS = 64, 128, 64, 16, 60, 500
FIELDS = ("U", "V", "HDRY", "THETA", "P3")


def member_spec(m):
    return ea.spec(ea.JUPITER_LIKE, NI, NJ, NK, DT, 1 + m, 0.5, 0.0, 0.0, 0.5)


def inner(a):
    return a[1:NK + 1, 1:NJ + 2, 1:NI + 1]


def rel_l2(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def test_ensemble_64_members_mixed_500_steps():
    say(f"ensemble {NM} x {NI}x{NJ}x{NK}, {STEPS} steps, mixed and fp64")
    mix = ea.Engine(ea.MIXED, NM)
    ms = [mix.add(member_spec(m)) for m in range(NM)]
    mix.step(STEPS)
    mix.sync()
    f64 = ea.Engine(ea.FP64, NM)
    fs = [f64.add(member_spec(m)) for m in range(NM)]
    f64.step(STEPS)
    f64.sync()
    worst = {f: 0.0 for f in FIELDS}
    kept = {}
    for m in range(NM):
        for f in FIELDS:
            a, b = mix.read(ms[m], f), f64.read(fs[m], f)
            assert np.isfinite(a).all(), (m, f)
            worst[f] = max(worst[f], rel_l2(inner(a), inner(b)))
            if m in (0, 31, 63):
                kept[(m, f)] = a
    say("worst member, mixed vs fp64 after", STEPS, "steps:", {f: f"{e:.2e}" for f, e in worst.items()})
    mix.close()
    f64.close()
    for f in ("HDRY", "THETA", "P3"):
        assert worst[f] <= 1e-5, (f, worst[f])
    for m in (0, 31, 63):
        e = ea.Engine(ea.MIXED, 1)
        s = e.add(member_spec(m))
        e.step(STEPS)
        for f in FIELDS:
            assert np.array_equal(e.read(s, f), kept[(m, f)]), ("alone vs ensemble", m, f)
        e.close()


def test_serving_mix_with_the_large_grids():
    grids = [(256, 128, 20), (512, 256, 32), (64, 32, 10), (128, 64, 16)]
    reqs = sv.poisson_mix(6, rate=20.0, seed=31, grids=grids, steps=(200, 260))
    # at least one of each large grid
    reqs[0] = sv.Request(0, ea.JUPITER_LIKE, 512, 256, 32, 220, 7, reqs[0].arrival)
    reqs[1] = sv.Request(1, ea.SATURN_LIKE, 256, 128, 20, 240, 8, reqs[1].arrival)
    say("requests", [(r.rid, r.planet, r.ni, r.steps) for r in reqs])
    mix, summ = sv.serve(reqs, precision=ea.MIXED, hbm_budget=64 << 30, max_members=6, chunk=40,
                         keep_fields=True)
    say("mix", summ)
    assert all(r.status == "ok" for r in mix), [r.status for r in mix]
    by_id = {r.rid: r for r in mix}
    for q in reqs:
        if q.ni < 256:
            continue
        alone, _ = sv.serve([sv.Request(q.rid, q.planet, q.ni, q.nj, q.nk, q.steps, q.seed, 0.0)],
                            precision=ea.MIXED, keep_fields=True)
        for f, a in alone[0].fields.items():
            assert np.array_equal(by_id[q.rid].fields[f], a), (q, f)
