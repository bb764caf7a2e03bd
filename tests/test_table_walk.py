"""The 3-D filter's walk search (table_place_near in
epic_amd/csrc/dev_common.h) returns the bisection's interval (table_place,
the reference's find_place_in_table with its ulp comparison fcmp,
src/shared/epic_funcs_util.c) on every table that table_well_separated
accepts: a property check of that argument on random ascending and
descending tables, with arguments on, next to (one and two ulps) and
between the entries and outside the range.  Host restatements of the three
device functions; the GPU path itself is checked bit for bit by the parity
tests."""
import math
import random

import numpy as np

EPS = 2.220446049250313e-16


def fcmp(x1, x2):
    e = math.frexp(x1 if abs(x1) > abs(x2) else x2)[1]
    delta = math.ldexp(EPS, e)
    d = x1 - x2
    return 1 if d > delta else (-1 if d < -delta else 0)


def bisect(x, xx):
    n = len(x)
    ascend = x[n - 1] >= x[0]
    lo, hi = fcmp(xx, x[0]), fcmp(xx, x[n - 1])
    if ascend:
        if lo <= 0:
            return 0
        if hi >= 0:
            return n - 2
    else:
        if lo >= 0:
            return 0
        if hi <= 0:
            return n - 2
    il, iu = -1, n
    while iu - il > 1:
        im = (iu + il) >> 1
        if (fcmp(xx, x[im]) >= 0) == ascend:
            il = im
        else:
            iu = im
    return il


def walk(x, xx, g):
    n = len(x)
    ascend = x[n - 1] >= x[0]
    lo, hi = fcmp(xx, x[0]), fcmp(xx, x[n - 1])
    if ascend:
        if lo <= 0:
            return 0
        if hi >= 0:
            return n - 2
    else:
        if lo >= 0:
            return 0
        if hi <= 0:
            return n - 2
    g = min(max(g, 0), n - 2)
    p = lambda i: (fcmp(xx, x[i]) >= 0) == ascend  # noqa: E731
    if p(g):
        while g + 1 <= n - 2 and p(g + 1):
            g += 1
    else:
        while g > 0 and not p(g):
            g -= 1
    return g


def well_separated(x):
    n = len(x)
    mx = max(abs(v) for v in x)
    gap = 2.0 * math.ldexp(EPS, math.frexp(mx)[1] + 1)
    ascend = x[n - 1] >= x[0]
    for e in range(n - 1):
        dd = x[e + 1] - x[e] if ascend else x[e] - x[e + 1]
        if not dd > gap:
            return False
    return n >= 2


def test_walk_equals_bisection_on_separated_tables():
    rng = random.Random(5)
    checked = 0
    for trial in range(150):
        n = rng.randint(2, 40)
        scale = 10.0 ** rng.uniform(-3, 7)
        steps = [rng.uniform(0.0, 1.0) ** 3 * scale for _ in range(n)]
        if trial % 7 == 0:  # some nearly-touching entries
            steps[rng.randrange(n)] = math.ldexp(EPS, math.frexp(scale)[1]) * rng.uniform(0.0, 6.0)
        x = list(np.cumsum(steps) + rng.uniform(-2, 2) * scale)
        if trial % 2:
            x = x[::-1]
        if not well_separated(x):
            continue
        probes = []
        for v in x:
            probes += [v, np.nextafter(v, np.inf), np.nextafter(v, -np.inf),
                       np.nextafter(np.nextafter(v, np.inf), np.inf)]
        probes += [rng.uniform(min(x), max(x)) for _ in range(20)]
        probes += [min(x) - scale, max(x) + scale]
        for xx in probes:
            want = bisect(x, float(xx))
            for g in (-3, 0, n // 2, n - 2, n + 3, rng.randrange(n)):
                assert walk(x, float(xx), g) == want, (x, xx, g)
                checked += 1
    assert checked > 3000
