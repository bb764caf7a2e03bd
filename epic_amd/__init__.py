"""epic_amd -- MI355X (gfx950) implementation of EPIC's dynamical-core
timestep, driven through the C ABI of libepic_amd.so (include/epic_amd.h).

The Python layer only marshals arguments: every computation of the step runs
in the HIP kernels of the shared library.  If the library is missing the
import fails loudly; there is no fallback path.
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libepic_amd.so")

# planets (enum epic_planet_kind)
HELD_SUAREZ, JUPITER_LIKE, SATURN_LIKE, VENUS_LIKE, HOT_JUPITER = 0, 1, 2, 3, 4
FP64, MIXED = 0, 1
# timestep() actions (include/epic.h:294-295)
STEP_PROGS_AND_SYNC_DIAGS, SYNC_DIAGS_ONLY = 0, 1

FIELDS = ["U", "V", "HDRY", "THETA", "P3", "P2", "H3", "THETA2", "T2", "T3", "RHO2", "RHO3",
          "EXNER3", "GZ3", "MONT3", "PV3", "DIV_UV3", "W3", "HEAT3", "DZDT3", "DUDT0", "DUDT1",
          "DUDT2", "DVDT0", "DVDT1", "DVDT2", "GZ_SURFACE", "RI2"]
FIELD = {n: i for i, n in enumerate(FIELDS)}
# launch-shape knobs of the autotuner (epic_engine_set_shape's numbering)
SHAPE_KNOBS = ["cols_p2etc", "cols_gz3", "cols_calcw", "cols_hsuv", "rows_zhv", "rows_mer", "rows_d2"]


K_SPONGE_DEFAULT = -1  # the reference `initial` program's choice of k_sponge


class CaseSpec(ctypes.Structure):
    """epic_case_spec of include/epic_amd.h."""
    _fields_ = [("planet", ctypes.c_int), ("ni", ctypes.c_int), ("nj", ctypes.c_int),
                ("nk", ctypes.c_int), ("dt", ctypes.c_int), ("seed", ctypes.c_uint64),
                ("nu4_nondim", ctypes.c_double), ("nudiv_nondim", ctypes.c_double),
                ("noise_theta", ctypes.c_double), ("noise_uv", ctypes.c_double),
                ("k_sponge", ctypes.c_int), ("newt_cool", ctypes.c_int),
                ("newt_cool_adjust", ctypes.c_int), ("drag_v", ctypes.c_int),
                ("drag_zonal_avg", ctypes.c_int), ("tau_drag", ctypes.c_double),
                ("prandtl", ctypes.c_double), ("nu6_nondim", ctypes.c_double),
                ("nu8_nondim", ctypes.c_double), ("pgrad_traditional", ctypes.c_int),
                ("adv_mass", ctypes.c_int), ("adv_nonmass", ctypes.c_int), ("pv_scheme", ctypes.c_int)]

    # the reference's switches as keyword arguments of spec(), in the order
    # of the command-line options of oracle/ref_harness.c
    OPTIONS = ("k_sponge", "newt_cool", "newt_cool_adjust", "drag_v", "drag_zonal_avg", "tau_drag",
               "prandtl", "nu6_nondim", "nu8_nondim", "pgrad_traditional", "adv_mass", "adv_nonmass",
               "pv_scheme")

    def options(self):
        """{option: value} of the switches that differ from a zeroed spec."""
        return {k: getattr(self, k) for k in self.OPTIONS if getattr(self, k)}

    def args(self):
        """The ten positional values (planet .. noise_uv)."""
        return (self.planet, self.ni, self.nj, self.nk, self.dt, self.seed, self.nu4_nondim,
                self.nudiv_nondim, self.noise_theta, self.noise_uv)

    def __repr__(self):
        extra = "".join(f", {k}={v}" for k, v in self.options().items())
        return f"spec{self.args()}{extra}"


def spec(planet=HELD_SUAREZ, ni=64, nj=32, nk=10, dt=600, seed=1, nu4=0.5, nudiv=0.0,
         noise_theta=0.1, noise_uv=0.0, k_sponge=K_SPONGE_DEFAULT, newt_cool=0, newt_cool_adjust=0,
         drag_v=0, drag_zonal_avg=0, tau_drag=0.0, prandtl=0.0, nu6=0.0, nu8=0.0, pgrad_traditional=0,
         adv_mass=0, adv_nonmass=0, pv_scheme=0):
    """A case: planet, grid, dt and seed, the del^4 and divergence-damping
    fractions, the initial noise, and the reference's sponge, drag and
    Newtonian-cooling switches (k_sponge defaults to the reference `initial`
    program's choice: none for Held-Suarez, 1..4 layers otherwise)."""
    return CaseSpec(planet, ni, nj, nk, dt, seed, nu4, nudiv, noise_theta, noise_uv, k_sponge, newt_cool,
                    newt_cool_adjust, drag_v, drag_zonal_avg, tau_drag, prandtl, nu6, nu8, pgrad_traditional,
                    adv_mass, adv_nonmass, pv_scheme)


_D, _I, _P = ctypes.c_double, ctypes.c_int, ctypes.POINTER(ctypes.c_double)


class Case(ctypes.Structure):
    """epic_case of include/epic_amd.h (built by epic_case_build)."""
    _fields_ = ([("planet_name", ctypes.c_char * 32), ("planet_type", ctypes.c_char * 16), ("planet_index", _I)] +
                [(n, _D) for n in ("re", "rp", "omega", "cp", "rgas", "cpr", "kappa", "GM", "J2", "x_h2",
                                   "kinvisc")] +
                [(n, _I) for n in ("ni", "nj", "nk", "dt", "jtp", "k_sigma", "k_sponge", "newt_cool_on",
                                   "newt_cool_adjust", "drag_v", "drag_zonal_avg")] +
                [(n, _D) for n in ("dln", "dlt", "dy0", "globe_lonbot", "globe_lontop", "globe_latbot",
                                   "globe_lattop", "press0", "pbot", "ptop", "zeta0", "zeta1", "hybrid_alpha",
                                   "sigma_sigma", "topo_scale", "prandtl", "tau_drag", "nudiv_nondim")] +
                [("nu", _D * 9), ("nu_nondim", _D * 9)] +
                [(n, _D) for n in ("teq_contrast", "teq_p_contrast", "tau_rad", "ref_theta_b", "ref_pb",
                                   "ref_a_over_kappa")] +
                [(n, _I) for n in ("Iadim", "Jadim", "Kadim", "Nelem2d", "Nelem3d")] +
                [(n, _P) for n in ("lon", "lat", "rln", "rlt", "f", "m", "n", "mn", "g", "sigmatheta", "dsgth",
                                   "dsgth_inv", "p_ref", "theta_ref", "h_min", "t_sponge_inv", "u", "v", "hdry",
                                   "theta", "p3", "u_spinup", "gz_surface")] +
                [("ntp", _I), ("n_t_cool", _I)] +
                [(n, _P) for n in ("pdat", "tdat", "t_cool_x", "t_cool_y")] +
                [("dudt_hist", _P), ("dvdt_hist", _P), ("model_time", _D), ("start_time", _D),
                 ("nc_header", ctypes.c_void_p), ("pgrad_traditional", _I), ("adv_mass", _I),
                 ("adv_nonmass", _I), ("pv_scheme", _I)])

    def array(self, name, n):
        """A numpy copy of the first n values of one of the case's arrays."""
        return np.ctypeslib.as_array(getattr(self, name), shape=(n,)).copy()


def build_case(sp):
    """epic_case_build of a spec; free it with free_case."""
    c = Case()
    rc = lib().epic_case_build(ctypes.byref(sp), ctypes.byref(c))
    if rc:
        lib().epic_case_free(ctypes.byref(c))
        raise EpicError(f"epic_case_build failed ({rc})")
    return c


def free_case(c):
    lib().epic_case_free(ctypes.byref(c))


class NcState(ctypes.Structure):
    """epic_nc_state of include/epic_amd.h."""
    _fields_ = ([(n, ctypes.c_void_p) for n in ("u", "v", "p3", "theta", "u_spinup", "gz_surface",
                                                "dudt_m1", "dudt_m2", "dvdt_m1", "dvdt_m2")] +
                [("model_time", _D), ("l_s_deg", _D)])


def nc_read(path):
    """A case from an epic.nc file (epic_nc_read); free it with free_case."""
    c = Case()
    rc = lib().epic_nc_read(str(path).encode(), ctypes.byref(c))
    if rc:
        raise EpicError(f"epic_nc_read failed ({rc}): {lib().epic_nc_error().decode()}")
    return c


def nc_write(path, case, fields, model_time=None, l_s_deg=0.0):
    """Write an epic.nc file from padded arrays: fields maps U, V, P3,
    THETA, U_SPINUP, GZ_SURFACE and optionally DUDT_M1, DUDT_M2, DVDT_M1,
    DVDT_M2 to numpy arrays (epic_nc_write)."""
    keep = {k: np.ascontiguousarray(v, dtype=np.float64) for k, v in fields.items()}
    st = NcState()
    for k, nm in (("U", "u"), ("V", "v"), ("P3", "p3"), ("THETA", "theta"), ("U_SPINUP", "u_spinup"),
                  ("GZ_SURFACE", "gz_surface"), ("DUDT_M1", "dudt_m1"), ("DUDT_M2", "dudt_m2"),
                  ("DVDT_M1", "dvdt_m1"), ("DVDT_M2", "dvdt_m2")):
        if k in keep:
            setattr(st, nm, keep[k].ctypes.data)
    st.model_time = case.model_time if model_time is None else model_time
    st.l_s_deg = l_s_deg
    rc = lib().epic_nc_write(str(path).encode(), ctypes.byref(case), ctypes.byref(st))
    if rc:
        raise EpicError(f"epic_nc_write failed ({rc}): {lib().epic_nc_error().decode()}")


def dims(sp):
    """(Iadim, Jadim, Kadim) of a spec: the reference's padded extents."""
    return sp.ni + 2, sp.nj + 3, sp.nk + 2


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is not built; run __graft_entry__.build()")
        L = ctypes.CDLL(LIB_PATH)
        vp, ci, cd = ctypes.c_void_p, ctypes.c_int, ctypes.c_double
        L.epic_engine_create.restype = vp
        L.epic_engine_create.argtypes = [ci, ci, ctypes.c_size_t]
        L.epic_engine_destroy.argtypes = [vp]
        L.epic_engine_add_spec.argtypes = [vp, ctypes.POINTER(CaseSpec)]
        L.epic_engine_add_spec.restype = ci
        L.epic_engine_checkpoint.argtypes = [vp, ci, ctypes.c_char_p]
        L.epic_engine_restore_spec.argtypes = [vp, ctypes.POINTER(CaseSpec), ctypes.c_char_p]
        L.epic_engine_restore_spec.restype = ci
        L.epic_engine_remove.argtypes = [vp, ci]
        L.epic_engine_step.argtypes = [vp, ci]
        L.epic_engine_step_members.argtypes = [vp, ctypes.POINTER(ci), ci, ci]
        L.epic_engine_step_shape.argtypes = [vp, ci, ci, ci, ci]
        L.epic_engine_read.argtypes = [vp, ci, ci, vp]
        L.epic_engine_write.argtypes = [vp, ci, ci, vp]
        L.epic_engine_sync.argtypes = [vp]
        L.epic_engine_wait.argtypes = [vp]
        L.epic_engine_member_error.argtypes = [vp, ci]
        L.epic_engine_steps_done.argtypes = [vp, ci]
        L.epic_engine_steps_done.restype = ctypes.c_long
        L.epic_engine_bytes_in_use.argtypes = [vp]
        L.epic_engine_bytes_in_use.restype = ctypes.c_size_t
        L.epic_engine_last_error.argtypes = [vp]
        L.epic_engine_last_error.restype = ctypes.c_char_p
        L.epic_engine_time_planes.argtypes = [vp, ci, ctypes.POINTER(ci)]
        L.epic_engine_profile.argtypes = [vp, ci, vp, ctypes.POINTER(cd), ctypes.POINTER(ci), ci]
        L.epic_engine_autotune.argtypes = [vp, ci, vp, ci]
        L.epic_engine_set_shape.argtypes = [vp, ci, ci]
        L.epic_copy_probe.argtypes = [ctypes.c_size_t, ci, ctypes.POINTER(cd)]
        L.epic_member_footprint.argtypes = [ci, ci, ci, ci]
        L.epic_member_footprint.restype = ctypes.c_size_t
        L.epic_case_build.argtypes = [ctypes.POINTER(CaseSpec), ctypes.POINTER(Case)]
        L.epic_case_build.restype = ci
        L.epic_case_free.argtypes = [ctypes.POINTER(Case)]
        L.epic_engine_add.argtypes = [vp, ctypes.POINTER(Case)]
        L.epic_engine_add.restype = ci
        L.epic_nc_read.argtypes = [ctypes.c_char_p, ctypes.POINTER(Case)]
        L.epic_nc_read.restype = ci
        L.epic_nc_write.argtypes = [ctypes.c_char_p, ctypes.POINTER(Case), ctypes.POINTER(NcState)]
        L.epic_nc_write.restype = ci
        L.epic_nc_error.restype = ctypes.c_char_p
        L.epic_engine_write_nc.argtypes = [vp, ci, ctypes.POINTER(Case), ctypes.c_char_p]
        L.epic_engine_write_nc.restype = ci
        _lib = L
    return _lib


class EpicError(RuntimeError):
    pass


class Engine:
    """Device-resident members advanced together (epic_engine_* of the C ABI)."""

    def __init__(self, precision=FP64, max_members=1, hbm_budget=0):
        L = lib()
        self._e = L.epic_engine_create(precision, max_members, hbm_budget)
        if not self._e:
            raise EpicError("epic_engine_create failed")
        self.specs = {}

    def _check(self, rc, what):
        if rc < 0:
            msg = lib().epic_engine_last_error(self._e).decode()
            raise EpicError(f"{what} failed ({rc}): {msg}")
        return rc

    def add(self, sp):
        slot = self._check(lib().epic_engine_add_spec(self._e, ctypes.byref(sp)), "add")
        self.specs[slot] = sp
        return slot

    def add_case(self, c, spec=None):
        """A member from a built or read case (epic_engine_add); the engine
        copies what it needs, the caller keeps and frees the case.  With
        `spec` (the spec the case was built from) the engine keeps that
        instead of the case, which the caller may then free at once."""
        slot = self._check(lib().epic_engine_add(self._e, ctypes.byref(c)), "add")
        self.specs[slot] = c if spec is None else spec
        return slot

    def write_nc(self, slot, path, case=None):
        """The member's state as an epic.nc file; the header from `case`
        (default: the case the member was added from, built again from its
        spec)."""
        c, built = case, False
        if c is None:
            c = self.specs[slot]
            if isinstance(c, CaseSpec):
                c, built = build_case(c), True
        try:
            self._check(lib().epic_engine_write_nc(self._e, slot, ctypes.byref(c), str(path).encode()),
                        "write_nc")
        finally:
            if built:
                free_case(c)

    def remove(self, slot):
        self._check(lib().epic_engine_remove(self._e, slot), "remove")
        self.specs.pop(slot, None)

    def step(self, n=1):
        self._check(lib().epic_engine_step(self._e, n), "step")

    def step_members(self, slots, n=1):
        arr = (ctypes.c_int * len(slots))(*slots)
        self._check(lib().epic_engine_step_members(self._e, arr, len(slots), n), "step_members")

    def step_shape(self, ni, nj, nk, n=1):
        """Advance every member of grid shape ni x nj x nk by n steps (the
        other members wait): the same bits as stepping the whole engine."""
        self._check(lib().epic_engine_step_shape(self._e, ni, nj, nk, n), "step_shape")

    def sync(self):
        self._check(lib().epic_engine_sync(self._e), "sync")

    def wait(self):
        """Wait for the queued work without checking the members' error bits."""
        self._check(lib().epic_engine_wait(self._e), "wait")

    def member_error(self, slot):
        """Sticky fatal-condition bits of one member (0: healthy)."""
        return self._check(lib().epic_engine_member_error(self._e, slot), "member_error")

    def read(self, slot, field):
        sp = self.specs[slot]
        Ia, Ja, Ka = dims(sp)
        f = FIELD[field] if isinstance(field, str) else field
        shape = (Ja, Ia) if f == FIELD["GZ_SURFACE"] else (Ka, Ja, Ia)
        out = np.empty(shape, dtype=np.float64)
        self._check(lib().epic_engine_read(self._e, slot, f, out.ctypes.data), "read")
        return out

    def write_field(self, slot, field, arr):
        """Overwrite a prognostic field (U, V, HDRY, THETA) of a member."""
        sp = self.specs[slot]
        a = np.ascontiguousarray(arr, dtype=np.float64)
        if a.shape != tuple(reversed(dims(sp))):
            raise ValueError(f"expected shape {tuple(reversed(dims(sp)))}, got {a.shape}")
        f = FIELD[field] if isinstance(field, str) else field
        self._check(lib().epic_engine_write(self._e, slot, f, a.ctypes.data), "write")

    def time_planes(self, slot):
        arr = (ctypes.c_int * 4)()
        self._check(lib().epic_engine_time_planes(self._e, slot, arr), "time_planes")
        return list(arr)

    def tendency(self, slot, wind, back=0):
        """Tendency plane of U or V: back=0 is the newest (IT_MINUS1 after a step)."""
        it0, itm1, itm2, _ = self.time_planes(slot)
        plane = [itm1, itm2, it0][back]
        return self.read(slot, f"D{wind}DT{plane}")

    def steps_done(self, slot):
        return lib().epic_engine_steps_done(self._e, slot)

    def bytes_in_use(self):
        return lib().epic_engine_bytes_in_use(self._e)

    def profile(self, nsteps, cap=128):
        names = ctypes.create_string_buffer(48 * cap)
        ms = (ctypes.c_double * cap)()
        cnt = (ctypes.c_int * cap)()
        n = self._check(lib().epic_engine_profile(self._e, nsteps, names, ms, cnt, cap), "profile")
        raw = names.raw
        out = []
        for i in range(n):
            nm = raw[48 * i:48 * (i + 1)].split(b"\0", 1)[0].decode()
            out.append((nm, ms[i], cnt[i]))
        return out

    def checkpoint(self, slot, path):
        """Write the member's state for an exact restart to `path`."""
        self._check(lib().epic_engine_checkpoint(self._e, slot, str(path).encode()), "checkpoint")

    def restore(self, sp, path):
        """A new member from its case spec and a checkpoint of it; returns
        the slot.  It continues bit for bit as the checkpointed member."""
        slot = self._check(lib().epic_engine_restore_spec(self._e, ctypes.byref(sp), str(path).encode()),
                           "restore")
        self.specs[slot] = sp
        return slot

    def autotune(self, steps=2):
        """Pick the launch shapes of the tunable kernels for every bucket
        (state and results unchanged); returns the choices as text."""
        buf = ctypes.create_string_buffer(4096)
        self._check(lib().epic_engine_autotune(self._e, steps, buf, len(buf)), "autotune")
        return buf.value.decode()

    def set_shape(self, knob, value):
        """One launch shape for every bucket by hand (epic_engine_set_shape);
        knob by name or number."""
        k = SHAPE_KNOBS.index(knob) if isinstance(knob, str) else knob
        self._check(lib().epic_engine_set_shape(self._e, k, value), "set_shape")

    def close(self):
        if self._e:
            lib().epic_engine_destroy(self._e)
            self._e = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def copy_probe(nbytes=256 << 20, reps=20):
    """Seconds per plain device copy of nbytes (nbytes read + nbytes written)."""
    sec = ctypes.c_double()
    rc = lib().epic_copy_probe(nbytes, reps, ctypes.byref(sec))
    if rc:
        raise EpicError(f"copy probe failed ({rc})")
    return sec.value


def footprint(ni, nj, nk, precision=FP64):
    return lib().epic_member_footprint(ni, nj, nk, precision)
