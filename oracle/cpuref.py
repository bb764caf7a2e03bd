"""ctypes wrapper of oracle/libcpuref.so, the CPU FP64 restatement of the
EPIC timestep.  TEST INFRASTRUCTURE ONLY: imported by tests/,
__graft_entry__.smoke() and bench.py's cpu_baseline leg, as the checker or
the timed CPU baseline, never by the product."""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "libcpuref.so")
_L = None


def lib():
    global _L
    if _L is None:
        if not os.path.exists(LIB):
            raise RuntimeError(f"{LIB} not built (make -C oracle)")
        L = ctypes.CDLL(LIB)
        vp = ctypes.c_void_p
        L.cpuref_create.restype = vp
        L.cpuref_create.argtypes = [vp]
        for f in ("cpuref_sync", "cpuref_step", "cpuref_error"):
            getattr(L, f).argtypes = [vp]
        L.cpuref_destroy.argtypes = [vp]
        L.cpuref_read.argtypes = [vp, ctypes.c_int, vp]
        L.cpuref_time_planes.argtypes = [vp, ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_long)]
        L.cpuref_realft.argtypes = [vp, ctypes.c_ulong, ctypes.c_int]
        L.cpuref_pchip.argtypes = [ctypes.c_int, vp, vp, vp]
        L.cpuref_pchip_eval.argtypes = [ctypes.c_int, vp, vp, vp, ctypes.c_double]
        L.cpuref_pchip_eval.restype = ctypes.c_double
        L.cpuref_fcmp.argtypes = [ctypes.c_double, ctypes.c_double]
        _L = L
    return _L


class CpuRef:
    """One integration on the CPU oracle, same field order as epic_amd.FIELDS."""

    def __init__(self, sp):
        self.sp = sp
        self._r = lib().cpuref_create(ctypes.addressof(sp))
        if not self._r:
            raise RuntimeError("cpuref_create failed")
        self.shape3 = (sp.nk + 2, sp.nj + 3, sp.ni + 2)

    def sync(self):
        return lib().cpuref_sync(self._r)

    def step(self, n=1):
        rc = 0
        for _ in range(n):
            rc = lib().cpuref_step(self._r)
        return rc

    def read(self, f):
        from epic_amd import FIELD
        q = FIELD[f] if isinstance(f, str) else f
        shape = self.shape3[1:] if q == FIELD["GZ_SURFACE"] else self.shape3
        out = np.empty(shape, dtype=np.float64)
        lib().cpuref_read(self._r, q, out.ctypes.data)
        return out

    def time_planes(self):
        it = (ctypes.c_int * 3)()
        st = ctypes.c_long()
        lib().cpuref_time_planes(self._r, it, ctypes.byref(st))
        return list(it) + [st.value]

    def error(self):
        return lib().cpuref_error(self._r)

    def close(self):
        if self._r:
            lib().cpuref_destroy(self._r)
            self._r = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
