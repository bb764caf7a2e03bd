"""The allocation-free step (epic_step_noalloc) through the C ABI, against
the engine (epic_engine_add / epic_engine_step) on the same case: every
field bit for bit, and the time planes, over calls in a row on the same
GPU -- the priming call and a stepping call on one grid, then a stepping
call on a smaller grid (the static working memory keeps the last call's
contents and must be cleared by each call) -- and the refusal of a member
too large for the static working memory.

Device memory for the fields is the caller's: torch allocates it here; the
library allocates none."""
import ctypes

import numpy as np
import pytest

import epic_amd as ea
from tests.gpu_util import say

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

NF = len(ea.FIELDS)
FLT_MAX64 = np.finfo(np.float64).max


def _noalloc():
    L = ea.lib()
    f = L.epic_step_noalloc
    f.argtypes = [ctypes.POINTER(ea.Case), ctypes.POINTER(ctypes.c_void_p), ctypes.POINTER(ctypes.c_int),
                  ctypes.c_int, ctypes.c_int, ctypes.c_void_p, ctypes.c_char_p, ctypes.c_size_t]
    f.restype = ctypes.c_int
    return f


class DevFields:
    """The caller's device arrays of one member: one per epic_field, the
    three u and the three v tendency planes consecutive."""

    def __init__(self, sp):
        self.shape3 = (sp.nk + 2, sp.nj + 3, sp.ni + 2)
        self.shape2 = self.shape3[1:]
        n3 = int(np.prod(self.shape3))
        self.t = {}
        for f in ea.FIELDS:
            if f in ("DUDT1", "DUDT2", "DVDT1", "DVDT2"):
                continue
            if f in ("DUDT0", "DVDT0"):
                self.t[f] = torch.zeros(3 * n3, dtype=torch.float64, device="cuda")
            elif f == "GZ_SURFACE":
                self.t[f] = torch.zeros(int(np.prod(self.shape2)), dtype=torch.float64, device="cuda")
            else:
                self.t[f] = torch.zeros(n3, dtype=torch.float64, device="cuda")
        self.n3 = n3

    def ptrs(self):
        arr = (ctypes.c_void_p * NF)()
        for q, f in enumerate(ea.FIELDS):
            if f[:4] in ("DUDT", "DVDT"):
                base, p = self.t[f[:4] + "0"], int(f[4])
                arr[q] = base.data_ptr() + 8 * p * self.n3
            else:
                arr[q] = self.t[f].data_ptr()
        return arr

    def _view(self, f):
        if f[:4] in ("DUDT", "DVDT"):
            p = int(f[4])
            return self.t[f[:4] + "0"][p * self.n3:(p + 1) * self.n3]
        return self.t[f]

    def put(self, f, a):
        self._view(f).copy_(torch.from_numpy(np.ascontiguousarray(a, dtype=np.float64).ravel()))

    def get(self, f):
        shape = self.shape2 if f == "GZ_SURFACE" else self.shape3
        return self._view(f).cpu().numpy().reshape(shape)


def call(case, dev, it4, action, nsteps):
    it = (ctypes.c_int * 4)(*it4)
    err = ctypes.create_string_buffer(512)
    torch.cuda.synchronize()
    rc = _noalloc()(ctypes.byref(case), dev.ptrs(), it, action, nsteps, None, err, len(err))
    return rc, list(it), err.value.decode()


def initial_state(sp, case, dev):
    """The state epic_engine_add starts from: the case's u, v, theta, p3,
    gz_surface, h from the layer pressure thickness (as epic_main forms it),
    and the FLOAT_MAX Adams-Bashforth start-up flags in planes 1 and 2."""
    Ka, Ja, Ia = dev.shape3
    n3, n2 = dev.n3, Ja * Ia
    get3 = lambda name: case.array(name, n3).reshape(Ka, Ja, Ia)
    for f, name in (("U", "u"), ("V", "v"), ("THETA", "theta"), ("P3", "p3")):
        dev.put(f, get3(name))
    dev.put("GZ_SURFACE", case.array("gz_surface", n2))
    p3 = get3("p3")
    g = case.array("g", 2 * (sp.nj + 1) + 1)
    dsg = case.array("dsgth", 2 * (sp.nk + 1) + 1)
    h = np.zeros((Ka, Ja, Ia))
    ni, nj, nk = sp.ni, sp.nj, sp.nk
    for K in range(1, nk + 1):
        for J in range(0, nj + 1):
            h[K, J + 1, 1:ni + 1] = (p3[K, J + 1, 1:ni + 1] - p3[K - 1, J + 1, 1:ni + 1]) / (g[2 * J + 1] * dsg[2 * K])
    h[:, 1:nj + 3, 0] = h[:, 1:nj + 3, ni]
    h[:, 1:nj + 3, ni + 1] = h[:, 1:nj + 3, 1]
    dev.put("HDRY", h)
    flags = np.zeros((Ka, Ja, Ia))
    flags[1:nk + 1, 1:nj + 2, 1:ni + 1] = FLT_MAX64
    for w in ("DUDT", "DVDT"):
        for p in (1, 2):
            dev.put(f"{w}{p}", flags)


def engine_fields(eng, slot):
    return {f: eng.read(slot, f) for f in ea.FIELDS}


def same(dev, want, tag):
    bad = [f for f in ea.FIELDS if dev.get(f).tobytes() != want[f].tobytes()]
    say(tag, "fields differing from the engine:", bad)
    assert not bad, (tag, bad)


def test_noalloc_equals_engine_on_calls_in_a_row():
    a = ea.spec(ea.HELD_SUAREZ, 64, 32, 10, 600, 1, 0.5, 0.1, 0.1, 0.2)
    b = ea.spec(ea.JUPITER_LIKE, 32, 16, 8, 60, 9, 0.5, 0.0, 0.0, 0.5, tau_drag=2.0e4, drag_v=1)
    # the engine's results
    eng_a = ea.Engine(ea.FP64, 1)
    sa = eng_a.add(a)
    primed_a, tp_a0 = engine_fields(eng_a, sa), eng_a.time_planes(sa)
    eng_a.step(4)
    stepped_a, tp_a1 = engine_fields(eng_a, sa), eng_a.time_planes(sa)
    eng_a.close()
    eng_b = ea.Engine(ea.FP64, 1)
    sb = eng_b.add(b)
    primed_b, tp_b0 = engine_fields(eng_b, sb), eng_b.time_planes(sb)
    eng_b.step(3)
    stepped_b, tp_b1 = engine_fields(eng_b, sb), eng_b.time_planes(sb)
    eng_b.close()

    ca, cb = ea.build_case(a), ea.build_case(b)
    try:
        # call 1: prime grid A from its initial state
        da = DevFields(a)
        initial_state(a, ca, da)
        say("noalloc call 1: prime held_suarez 64x32x10")
        rc, it, err = call(ca, da, [0, 1, 2, 0], ea.SYNC_DIAGS_ONLY, 0)
        assert rc == 0, err
        assert it == tp_a0
        same(da, primed_a, "prime A")
        # call 2: four steps of grid A from the primed state
        say("noalloc call 2: 4 steps of held_suarez 64x32x10")
        rc, it, err = call(ca, da, tp_a0, ea.STEP_PROGS_AND_SYNC_DIAGS, 4)
        assert rc == 0, err
        assert it == tp_a1
        same(da, stepped_a, "step A")
        # call 3: a smaller grid, from the engine's primed state of B
        db = DevFields(b)
        for f in ea.FIELDS:
            db.put(f, primed_b[f])
        say("noalloc call 3: 3 steps of jupiter 32x16x8 (sponge, Rayleigh drag)")
        rc, it, err = call(cb, db, tp_b0, ea.STEP_PROGS_AND_SYNC_DIAGS, 3)
        assert rc == 0, err
        assert it == tp_b1
        same(db, stepped_b, "step B")
    finally:
        ea.free_case(ca)
        ea.free_case(cb)


def test_noalloc_refuses_a_member_beyond_its_working_memory():
    big = ea.spec(ea.JUPITER_LIKE, 512, 256, 32, 60, 1)
    c = ea.build_case(big)
    try:
        dev = DevFields(ea.spec(ea.JUPITER_LIKE, 16, 8, 6, 60, 1))  # never touched: refused first
        it = (ctypes.c_int * 4)(0, 1, 2, 0)
        err = ctypes.create_string_buffer(512)
        ptrs = (ctypes.c_void_p * NF)(*([dev.t["U"].data_ptr()] * NF))
        n3 = c.Nelem3d
        base_u, base_v = dev.t["DUDT0"].data_ptr(), dev.t["DVDT0"].data_ptr()
        for p in range(3):
            ptrs[ea.FIELD[f"DUDT{p}"]] = base_u + 8 * p * n3
            ptrs[ea.FIELD[f"DVDT{p}"]] = base_v + 8 * p * n3
        rc = _noalloc()(ctypes.byref(c), ptrs, it, ea.STEP_PROGS_AND_SYNC_DIAGS, 1, None, err, len(err))
        say("refusal:", rc, err.value.decode())
        assert rc == -2 and b"working memory" in err.value
    finally:
        ea.free_case(c)
