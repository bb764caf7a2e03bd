"""ctypes wrapper of oracle/libcpuref.so, the CPU FP64 restatement of the
EPIC timestep.  TEST INFRASTRUCTURE ONLY: imported by tests/,
__graft_entry__.smoke() and bench.py's cpu_baseline leg, as the checker or
the timed CPU baseline, never by the product."""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "libcpuref.so")
LIB_NUDGE = os.path.join(HERE, "libcpuref_nudge.so")
LIB_CR = os.path.join(HERE, "libcpuref_cr.so")
_L = {}


def lib(nudged=False):
    """The oracle library.  nudged=True (or "nudge") gives the copy whose
    exp/log/pow results are moved by one ulp 10% of the time
    (libm_nudge.h); "cr" the copy that evaluates exp/log/pow with the
    correctly rounded crmath.h of the GPU's FP64 kernels (libm_cr.h)."""
    key = "nudge" if nudged is True else (nudged or "plain")
    if key not in _L:
        path = {"plain": LIB, "nudge": LIB_NUDGE, "cr": LIB_CR}[key]
        if not os.path.exists(path):
            raise RuntimeError(f"{path} not built (make -C oracle)")
        L = ctypes.CDLL(path)
        vp = ctypes.c_void_p
        L.cpuref_create.restype = vp
        L.cpuref_create.argtypes = [vp]
        L.cpuref_create_nc.restype = vp
        L.cpuref_create_nc.argtypes = [ctypes.c_char_p]
        L.cpuref_pgrad_traditional.argtypes = [vp]
        for f in ("cpuref_sync", "cpuref_step", "cpuref_error"):
            getattr(L, f).argtypes = [vp]
        L.cpuref_destroy.argtypes = [vp]
        L.cpuref_read.argtypes = [vp, ctypes.c_int, vp]
        L.cpuref_write.argtypes = [vp, ctypes.c_int, vp]
        L.cpuref_richardson.argtypes = [vp]
        L.cpuref_time_planes.argtypes = [vp, ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_long)]
        L.cpuref_realft.argtypes = [vp, ctypes.c_ulong, ctypes.c_int]
        L.cpuref_pchip.argtypes = [ctypes.c_int, vp, vp, vp]
        L.cpuref_pchip_eval.argtypes = [ctypes.c_int, vp, vp, vp, ctypes.c_double]
        L.cpuref_pchip_eval.restype = ctypes.c_double
        L.cpuref_fcmp.argtypes = [ctypes.c_double, ctypes.c_double]
        L.cpuref_math.argtypes = [ctypes.c_int, ctypes.c_double, ctypes.c_double]
        L.cpuref_math.restype = ctypes.c_double
        L.cpuref_libm_stats.argtypes = [ctypes.POINTER(ctypes.c_long), ctypes.POINTER(ctypes.c_long), ctypes.c_int]
        L.cpuref_cr_select.argtypes = [ctypes.c_int]
        _L[key] = L
    return _L[key]


def libm_stats(variant="cr", reset=True):
    """(calls, differing) of exp, log, pow in a build: how many calls were
    made and how many of them glibc rounded differently from crmath.h."""
    c, d = (ctypes.c_long * 3)(), (ctypes.c_long * 3)()
    lib(variant).cpuref_libm_stats(c, d, 1 if reset else 0)
    return list(c), list(d)


class CpuRef:
    """One integration on the CPU oracle, same field order as epic_amd.FIELDS."""

    def __init__(self, sp, nudged=False, nc_path=None):
        """sp: a case spec; or, with nc_path, the state of that epic.nc file
        (sp then only needs ni, nj, nk)."""
        self.sp = sp
        self._L = lib(nudged)
        if nc_path is not None:
            self._r = self._L.cpuref_create_nc(str(nc_path).encode())
        else:
            self._r = self._L.cpuref_create(ctypes.addressof(sp))
        if not self._r:
            raise RuntimeError("cpuref_create failed")
        self.shape3 = (sp.nk + 2, sp.nj + 3, sp.ni + 2)

    def sync(self):
        return self._L.cpuref_sync(self._r)

    def pgrad_traditional(self):
        """The traditional pressure gradient alone into a cleared plane
        IT_ZERO of the tendencies."""
        self._L.cpuref_pgrad_traditional(self._r)

    def step(self, n=1):
        rc = 0
        for _ in range(n):
            rc = self._L.cpuref_step(self._r)
        return rc

    def read(self, f):
        from epic_amd import FIELD
        q = FIELD[f] if isinstance(f, str) else f
        shape = self.shape3[1:] if q == FIELD["GZ_SURFACE"] else self.shape3
        out = np.empty(shape, dtype=np.float64)
        self._L.cpuref_read(self._r, q, out.ctypes.data)
        return out

    def write(self, f, a):
        from epic_amd import FIELD
        a = np.ascontiguousarray(a, dtype=np.float64)
        assert a.shape == self.shape3
        if self._L.cpuref_write(self._r, FIELD[f], a.ctypes.data) != 0:
            raise ValueError(f"cpuref_write({f}) refused")

    def richardson_of(self, fields):
        """RI2 as store_diag forms it from the given U, V, P2, T2, RHO2, P3,
        RHO3 (written into this oracle instance, which is left holding them)."""
        for f in ("U", "V", "P2", "T2", "RHO2", "P3", "RHO3"):
            self.write(f, fields[f])
        self._L.cpuref_richardson(self._r)
        return self.read("RI2")

    def time_planes(self):
        it = (ctypes.c_int * 3)()
        st = ctypes.c_long()
        self._L.cpuref_time_planes(self._r, it, ctypes.byref(st))
        return list(it) + [st.value]

    def error(self):
        return self._L.cpuref_error(self._r)

    def close(self):
        if self._r:
            self._L.cpuref_destroy(self._r)
            self._r = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
