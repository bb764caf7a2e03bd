"""The request loop on the GPU: every integration's result is the same
served alone and served in a mix with others arriving and leaving
(bit for bit, FP64 and mixed precision, stepping the shortest shape
first or every shape together)."""
import numpy as np
import pytest

import epic_amd as ea
from epic_amd import serving as sv
from tests.gpu_util import say

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("policy", ["shortest", "lockstep"])
@pytest.mark.parametrize("precision", [ea.FP64, ea.MIXED])
def test_result_alone_equals_result_in_mix(precision, policy):
    reqs = sv.poisson_mix(8, rate=50.0, seed=11, grids=[(64, 32, 10), (128, 64, 16)], steps=(20, 80))
    mix, summ = sv.serve(reqs, precision=precision, hbm_budget=8 << 30, max_members=4, chunk=15,
                         keep_fields=True, policy=policy)
    say("mix", summ)
    assert all(r.status == "ok" for r in mix)
    by_id = {r.rid: r for r in mix}
    for q in reqs[:4]:
        alone, _ = sv.serve([sv.Request(q.rid, q.planet, q.ni, q.nj, q.nk, q.steps, q.seed, 0.0)],
                            precision=precision, keep_fields=True)
        for f, a in alone[0].fields.items():
            assert np.array_equal(by_id[q.rid].fields[f], a), (q, f)
