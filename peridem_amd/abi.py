"""ctypes view of the C-ABI in include/peridem_gpu.h.

`System` (peridem_amd.gen) describes a run; `Engine` drives one context of
a library exporting that ABI.  The HIP library (`libperidem_gpu.so`) exports
pd_*; the CPU oracle used by the tests exports the same entry points under
the or_ prefix, so both are driven by the same class.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
# PERIDEM_GPU_LIB points the binding at another build of the same library
# (the sanitizer builds of peridem_amd/sanitize.py)
GPU_LIB = os.environ.get("PERIDEM_GPU_LIB") or os.path.join(HERE, "libperidem_gpu.so")

PD_OK = 0
ERRORS = {-1: "PD_E_ARG", -2: "PD_E_STATE", -3: "PD_E_HIP", -4: "PD_E_UNSUPPORTED",
          -5: "PD_E_OVERFLOW", -6: "PD_E_NOMEM"}

PD_FLAG_ENABLE_FRACTURE = 1
PD_FLAG_SELF_CONTACT = 2
PD_FLAG_ALLOW_DAMPING = 3
PD_FLAG_ALLOW_FRICTION = 4
PD_FLAG_ENABLE_TORQUE = 5
PD_FLAG_COMPUTE_WALL_REACTION = 6
PD_FLAG_VELOCITY_CONSTRAINT = 7
PD_PROFILE_PHASES = 1   # a mark around every phase of a profiled step
PD_PROFILE_BOND = 2     # marks around the bond kernel only
PD_PFLAG_MOVABLE = 1
PD_PFLAG_STOPPABLE = 2
PD_PFLAG_BREAKABLE = 3

_u32, _u64, _i32, _dbl = C.c_uint32, C.c_uint64, C.c_int32, C.c_double
_pd = C.POINTER(C.c_double)
_pu = C.POINTER(C.c_uint32)


class RunDesc(C.Structure):
    _fields_ = [("dt", _dbl), ("enable_fracture", _u32), ("self_contact", _u32),
                ("allow_damping", _u32), ("allow_friction", _u32), ("nl_bdry_only", _u32),
                ("compute_wall_reaction", _u32), ("gradient_extforce", _u32),
                ("extforce_maxstep", _u32), ("enable_torque", _u32), ("allow_wall", _u32),
                ("contact_rad", _dbl), ("normal_stiffness", _dbl),
                ("friction_coefficient", _dbl), ("damping_ratio", _dbl),
                ("wall_lrtb", _dbl * 6), ("wall_speed", _dbl * 6),
                ("first_step", _u32), ("reserved", _u32)]


class ParticleDesc(C.Structure):
    _fields_ = [("nnodes", _u32), ("instance", _u32), ("pos", _pd), ("vol", _pd),
                ("disp", _pd), ("vel", _pd), ("acc", _pd), ("extforce", _pd),
                ("nconn", _u64), ("conn", _pu), ("nclamped", _u32), ("clamped", _pu),
                ("nbdry", _u32), ("bdry", _pu), ("delta", _dbl), ("rho", _dbl),
                ("cnot", _dbl), ("snot", _dbl), ("movable", _i32), ("breakable", _i32),
                ("stoppable", _i32), ("torque_axis", _u32), ("torque_val", _dbl)]


class PoolDesc(C.Structure):
    _fields_ = [("max_nodes", _u64), ("max_bond_entries", _u64), ("max_particles", _u64),
                ("max_jobs", _u64)]


class Counters(C.Structure):
    _fields_ = [(n, _u64) for n in ("steps", "bond_updates", "pair_computations",
                                    "alive_bonds", "total_bonds", "nodes", "particles",
                                    "cell_rebuilds", "list_capacity", "list_regrows")]

    def as_dict(self):
        return {n: int(getattr(self, n)) for n, _ in self._fields_}


SYMBOLS = ["create", "add_particle", "commit", "step", "set_flag", "set_wall_speed",
           "set_wall", "set_particle_flag", "set_snot", "get_state", "set_state",
           "get_connectivity", "get_wall", "get_counters", "last_error", "destroy"]


class Profile(C.Structure):
    _fields_ = [("ms", _dbl * 5), ("launches", _u64 * 5)]
    PHASES = ("position", "contact_list", "bond", "contact", "update")

    def as_dict(self):
        return {n: {"ms": self.ms[i], "launches": int(self.launches[i])}
                for i, n in enumerate(self.PHASES)}


class ServeEvent(C.Structure):
    _fields_ = [("request", _u64), ("job", _u32), ("status", C.c_int32), ("t_submit", C.c_double),
                ("t_admit", C.c_double), ("t_done", C.c_double)]


class PDError(RuntimeError):
    """A failed library call; `code` is the PD_E* status (None when not from a call)."""

    def __init__(self, msg, code=None):
        super().__init__(msg)
        self.code = code


PD_E_OVERFLOW = -5
PD_ARITH_BITWISE = 0
PD_ARITH_TOLERANCE = 1


def _bind(lib, prefix):
    P = C.c_void_p
    sig = {
        "create": (C.c_int, [C.POINTER(P), C.POINTER(RunDesc)]),
        "add_particle": (C.c_int, [P, C.POINTER(ParticleDesc), _pu]),
        "commit": (C.c_int, [P]),
        "step": (C.c_int, [P, _u32]),
        "set_flag": (C.c_int, [P, C.c_int, C.c_int]),
        "set_wall_speed": (C.c_int, [P, _pd]),
        "set_wall": (C.c_int, [P, _pd]),
        "set_particle_flag": (C.c_int, [P, _u32, C.c_int, C.c_int]),
        "set_snot": (C.c_int, [P, _u32, _dbl]),
        "reset_partzero_y": (C.c_int, [P, _dbl]),
        "get_state": (C.c_int, [P, _u32, _pd, _pd, _pd, _pd, _pd]),
        "set_state": (C.c_int, [P, _u32, _pd, _pd, _pd]),
        "get_connectivity": (C.c_int, [P, _u32, _pu, _u64, C.POINTER(_u64)]),
        "get_wall": (C.c_int, [P, _u32, _pd, _pd]),
        "get_counters": (C.c_int, [P, C.POINTER(Counters)]),
        "set_velocity_constraint": (C.c_int, [P, _u32, C.c_int, _dbl, C.c_int, _dbl]),
        "get_applied_force": (C.c_int, [P, _u32, _pd]),
        "last_error": (C.c_char_p, [P]),
        "destroy": (None, [P]),
    }
    fns = {}
    for name, (res, args) in sig.items():
        f = getattr(lib, prefix + name)
        f.restype = res
        f.argtypes = args
        fns[name] = f
    return fns


def _read_connectivity(call, initial_pairs, check):
    """Intact bonds of a particle as (i, q) pairs, i < q, in row order: one call
    with the particle's initial bond count as capacity (bonds only break), the
    count-then-fill protocol only if a caller's system holds fewer pairs."""
    m = _u64()
    out = np.empty((max(initial_pairs, 1), 2), dtype=np.uint32)
    if call(_ptr(out, C.c_uint32), initial_pairs, C.byref(m)) == 0:
        return out[:m.value]
    check(call(None, 0, C.byref(m)), "get_connectivity")
    out = np.empty((m.value, 2), dtype=np.uint32)
    check(call(_ptr(out, C.c_uint32), m.value, C.byref(m)), "get_connectivity")
    return out


_LIBS = {}


def load_gpu_lib(path=GPU_LIB):
    """Load the HIP library; raises if it is not built (no CPU fallback)."""
    if path not in _LIBS:
        if not os.path.exists(path):
            raise PDError(f"{path} is not built: run __graft_entry__.build()")
        lib = C.CDLL(path)
        lib.pd_version.restype = C.c_char_p
        lib.pd_last_step_ms.restype = C.c_int
        lib.pd_last_step_ms.argtypes = [C.c_void_p, C.POINTER(C.c_double)]
        lib.pd_set_profiling.argtypes = [C.c_void_p, C.c_int]
        lib.pd_set_profiling.restype = C.c_int
        lib.pd_get_profile.argtypes = [C.c_void_p, C.POINTER(Profile)]
        lib.pd_get_profile.restype = C.c_int
        lib.pd_math_mode.argtypes = [C.c_void_p, C.POINTER(C.c_int)]
        lib.pd_math_mode.restype = C.c_int
        lib.pd_selftest_math.argtypes = [_u64, _u64, C.POINTER(_u64), C.POINTER(_u64)]
        lib.pd_selftest_math.restype = C.c_int
        lib.pd_selftest_math_noalloc.argtypes = [_u64, _u64, C.POINTER(_u64), C.POINTER(_u64),
                                                 C.POINTER(_u64)]
        lib.pd_selftest_math_noalloc.restype = C.c_int
        P = C.c_void_p
        for name, res, args in (
                ("pd_reserve", C.c_int, [P, C.POINTER(PoolDesc)]),
                ("pd_job_footprint", C.c_int, [C.POINTER(ParticleDesc), _u32, C.POINTER(_u64),
                                               C.POINTER(_u64), C.POINTER(_u64)]),
                ("pd_add_job", C.c_int, [P, C.POINTER(RunDesc), C.POINTER(ParticleDesc), _u32, _u64,
                                         _pu]),
                ("pd_remove_job", C.c_int, [P, _u32]),
                ("pd_job_remaining", C.c_int, [P, _u32, C.POINTER(_u64)]),
                ("pd_job_get_state", C.c_int, [P, _u32, _u32, _pd, _pd, _pd, _pd, _pd]),
                ("pd_job_get_connectivity", C.c_int, [P, _u32, _u32, _pu, _u64, C.POINTER(_u64)]),
                ("pd_job_set_flag", C.c_int, [P, _u32, C.c_int, C.c_int]),
                ("pd_set_graphs", C.c_int, [P, C.c_int]),
                ("pd_set_arithmetic", C.c_int, [P, C.c_int]),
                ("pd_get_arithmetic", C.c_int, [P, C.POINTER(C.c_int)]),
                ("pd_bond_shape_count", C.c_int, []),
                ("pd_bond_shape_info", C.c_int, [C.c_int, C.POINTER(C.c_int), C.POINTER(C.c_int),
                                                 C.POINTER(C.c_int)]),
                ("pd_set_bond_shape", C.c_int, [P, C.c_int]),
                ("pd_get_bond_shape", C.c_int, [P, C.POINTER(C.c_int)]),
                ("pd_autotune", C.c_int, [P, _u32, C.POINTER(C.c_int), _pd]),
                ("pd_job_prepare", C.c_int, [C.POINTER(RunDesc), C.POINTER(ParticleDesc), _u32,
                                             C.POINTER(P)]),
                ("pd_job_blob_footprint", C.c_int, [P, C.POINTER(_u64), C.POINTER(_u64),
                                                    C.POINTER(_u64)]),
                ("pd_add_prepared", C.c_int, [P, P, _u64, _pu]),
                ("pd_job_blob_free", None, [P]),
                ("pd_serve_create", C.c_int, [P, _u32, _u32, C.POINTER(P)]),
                ("pd_serve_submit", C.c_int, [P, C.POINTER(RunDesc), C.POINTER(ParticleDesc), _u32, _u64,
                                              C.POINTER(_u64)]),
                ("pd_serve_poll", C.c_int, [P, C.POINTER(ServeEvent), _u32, C.POINTER(_u32)]),
                ("pd_serve_pending", C.c_int, [P, C.POINTER(_u64), C.POINTER(_u64), C.POINTER(_u64)]),
                ("pd_serve_release", C.c_int, [P, _u32]),
                ("pd_serve_destroy", None, [P])):
            f = getattr(lib, name)
            f.restype, f.argtypes = res, args
        _LIBS[path] = (lib, _bind(lib, "pd_"))
    return _LIBS[path]


def _ptr(a, t=C.c_double):
    return None if a is None else a.ctypes.data_as(C.POINTER(t))


class Engine:
    """One context of a library exporting the pd_ ABI (or the oracle's or_)."""

    def __init__(self, lib, fns, system, prefix="pd_"):
        self.lib, self.f, self.sys, self.prefix = lib, fns, system, prefix
        self.ctx = C.c_void_p()
        self._keep = []
        rd = system.run_desc()
        self._check(self.f["create"](C.byref(self.ctx), C.byref(rd)), "create")
        for p in system.particles:
            d = p.desc(self._keep)
            self._check(self.f["add_particle"](self.ctx, C.byref(d), None), "add_particle")
        self._check(self.f["commit"](self.ctx), "commit")
        self._keep = []

    def _check(self, rc, what):
        if rc != PD_OK:
            msg = self.f["last_error"](self.ctx) if self.ctx else b""
            raise PDError(f"{self.prefix}{what}: {ERRORS.get(rc, rc)}: "
                          f"{(msg or b'').decode(errors='replace')}", rc)

    def step(self, n=1):
        self._check(self.f["step"](self.ctx, n), "step")

    def set_flag(self, flag, value):
        self._check(self.f["set_flag"](self.ctx, flag, int(value)), "set_flag")

    def set_wall_speed(self, speed):
        a = np.ascontiguousarray(speed, dtype=np.float64)
        self._check(self.f["set_wall_speed"](self.ctx, _ptr(a)), "set_wall_speed")

    def set_wall(self, lrtb):
        a = np.ascontiguousarray(lrtb, dtype=np.float64)
        self._check(self.f["set_wall"](self.ctx, _ptr(a)), "set_wall")

    def set_particle_flag(self, p, which, value):
        self._check(self.f["set_particle_flag"](self.ctx, p, which, int(value)), "set_particle_flag")

    def set_snot(self, p, snot):
        self._check(self.f["set_snot"](self.ctx, p, float(snot)), "set_snot")

    def reset_partzero_y(self, wheel_rad):
        """Wheel reset of particle 0 before the next step (timeloop.h:496-535)."""
        self._check(self.f["reset_partzero_y"](self.ctx, float(wheel_rad)), "reset_partzero_y")

    def set_velocity_constraint(self, p, fix_mean_x=0, mean_x=0.0, fix_angular=0, angular_vel=0.0):
        self._check(self.f["set_velocity_constraint"](self.ctx, p, int(fix_mean_x), float(mean_x),
                                                      int(fix_angular), float(angular_vel)),
                    "set_velocity_constraint")

    def applied_force(self, p):
        out = np.empty((self.sys.particles[p].n, 3))
        self._check(self.f["get_applied_force"](self.ctx, p, _ptr(out)), "get_applied_force")
        return out

    def state(self, p):
        n = self.sys.particles[p].n
        out = {k: np.empty((n, 3)) for k in ("disp", "currpos", "vel", "acc", "force")}
        self._check(self.f["get_state"](self.ctx, p, *[_ptr(out[k]) for k in
                                                      ("disp", "currpos", "vel", "acc", "force")]),
                    "get_state")
        return out

    def set_state(self, p, disp=None, vel=None, acc=None):
        arrs = [None if a is None else np.ascontiguousarray(a, dtype=np.float64)
                for a in (disp, vel, acc)]
        self._check(self.f["set_state"](self.ctx, p, *[_ptr(a) for a in arrs]), "set_state")

    def connectivity(self, p):
        return _read_connectivity(lambda buf, cap, m: self.f["get_connectivity"](self.ctx, p, buf, cap, m),
                                  len(self.sys.particles[p].conn), self._check)

    def wall(self, instance=0):
        lrtb, reac = np.empty(6), np.empty(18)
        self._check(self.f["get_wall"](self.ctx, instance, _ptr(lrtb), _ptr(reac)), "get_wall")
        return lrtb, reac.reshape(6, 3)

    def counters(self):
        c = Counters()
        self._check(self.f["get_counters"](self.ctx, C.byref(c)), "get_counters")
        return c.as_dict()

    def last_step_ms(self):
        ms = C.c_double()
        self._check(self.lib.pd_last_step_ms(self.ctx, C.byref(ms)), "last_step_ms")
        return ms.value

    def set_profiling(self, on=True):
        self._check(self.lib.pd_set_profiling(self.ctx, int(on)), "set_profiling")

    def profile(self):
        pr = Profile()
        self._check(self.lib.pd_get_profile(self.ctx, C.byref(pr)), "get_profile")
        return pr.as_dict()

    def set_graphs(self, on=True):
        self._check(self.lib.pd_set_graphs(self.ctx, int(on)), "set_graphs")

    def set_arithmetic(self, mode):
        self._check(self.lib.pd_set_arithmetic(self.ctx, mode), "set_arithmetic")

    def set_bond_shape(self, shape):
        """Force a bond-kernel shape (see bond_shapes()); -1 = autotune."""
        self._check(self.lib.pd_set_bond_shape(self.ctx, int(shape)), "set_bond_shape")

    def bond_shape(self):
        v = C.c_int()
        self._check(self.lib.pd_get_bond_shape(self.ctx, C.byref(v)), "get_bond_shape")
        return v.value

    def autotune(self, reps=3):
        """Time every bond-kernel shape on the resident state; returns (best, ms per shape)."""
        n = self.lib.pd_bond_shape_count()
        ms = np.zeros(n)
        best = C.c_int()
        self._check(self.lib.pd_autotune(self.ctx, reps, C.byref(best), _ptr(ms)), "autotune")
        return best.value, ms

    def math_mode_safe(self):
        v = C.c_int()
        self._check(self.lib.pd_math_mode(self.ctx, C.byref(v)), "math_mode")
        return bool(v.value)

    def close(self):
        if self.ctx:
            self.f["destroy"](self.ctx)
            self.ctx = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


UNLIMITED = (1 << 64) - 1


def _descs(system, keep):
    arr = (ParticleDesc * len(system.particles))()
    for i, p in enumerate(system.particles):
        arr[i] = p.desc(keep)
    return arr


def footprint(system):
    """(node slots, bond entries, device bytes) a system takes in a job pool."""
    lib, _ = load_gpu_lib()
    keep = []
    arr = _descs(system, keep)
    ns, ne, nb = _u64(), _u64(), _u64()
    rc = lib.pd_job_footprint(arr, len(system.particles), C.byref(ns), C.byref(ne), C.byref(nb))
    if rc != PD_OK:
        raise PDError(f"pd_job_footprint: {ERRORS.get(rc, rc)}")
    return ns.value, ne.value, nb.value


class JobBlob:
    """A job laid out on the host by pd_job_prepare, ready to be placed."""

    def __init__(self, lib, ptr):
        self.lib, self.ptr = lib, ptr

    def footprint(self):
        ns, ne, nb = _u64(), _u64(), _u64()
        self.lib.pd_job_blob_footprint(self.ptr, C.byref(ns), C.byref(ne), C.byref(nb))
        return ns.value, ne.value, nb.value

    def free(self):
        if self.ptr:
            self.lib.pd_job_blob_free(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def prepare_job(system):
    """Lay a system out for a pool without any context (thread-safe; the C
    call runs without the GIL, so it overlaps a pool stepping on another thread)."""
    lib, _ = load_gpu_lib()
    keep = []
    arr = _descs(system, keep)
    rd = system.run_desc()
    ptr = C.c_void_p()
    rc = lib.pd_job_prepare(C.byref(rd), arr, len(system.particles), C.byref(ptr))
    if rc != PD_OK:
        msg = lib.pd_last_error(None) or b""
        raise PDError(f"pd_job_prepare: {ERRORS.get(rc, rc)}: {msg.decode(errors='replace')}")
    return JobBlob(lib, ptr)


class Pool:
    """A job pool of the HIP library: independent systems advanced together."""

    def __init__(self, max_nodes, max_bond_entries, max_particles, max_jobs):
        self.lib, self.f = load_gpu_lib()
        self.ctx = C.c_void_p()
        self.systems = {}
        rd = RunDesc()
        rd.dt = 1.0
        self._check(self.f["create"](C.byref(self.ctx), C.byref(rd)), "create")
        pool = PoolDesc(max_nodes, max_bond_entries, max_particles, max_jobs)
        self._check(self.lib.pd_reserve(self.ctx, C.byref(pool)), "reserve")

    def _check(self, rc, what):
        if rc != PD_OK:
            msg = self.f["last_error"](self.ctx) if self.ctx else b""
            raise PDError(f"pd_{what}: {ERRORS.get(rc, rc)}: "
                          f"{(msg or b'').decode(errors='replace')}", rc)

    def add_job(self, system, nsteps=UNLIMITED):
        """Admit a system with a step budget; returns its job slot."""
        keep = []
        arr = _descs(system, keep)
        rd = system.run_desc()
        job = _u32()
        self._check(self.lib.pd_add_job(self.ctx, C.byref(rd), arr, len(system.particles),
                                        nsteps, C.byref(job)), "add_job")
        self.systems[job.value] = system
        return job.value

    def try_add_job(self, system, nsteps=UNLIMITED):
        """add_job, or None when the pool has no room (PD_E_NOMEM)."""
        keep = []
        arr = _descs(system, keep)
        rd = system.run_desc()
        job = _u32()
        rc = self.lib.pd_add_job(self.ctx, C.byref(rd), arr, len(system.particles), nsteps,
                                 C.byref(job))
        if rc == -6:
            return None
        self._check(rc, "add_job")
        self.systems[job.value] = system
        return job.value

    def add_prepared(self, blob, system, nsteps=UNLIMITED):
        """Place a prepared job (see prepare_job); None when it does not fit."""
        job = _u32()
        rc = self.lib.pd_add_prepared(self.ctx, blob.ptr, nsteps, C.byref(job))
        if rc == -6:
            return None
        self._check(rc, "add_prepared")
        self.systems[job.value] = system
        return job.value

    def remove_job(self, job):
        self._check(self.lib.pd_remove_job(self.ctx, job), "remove_job")
        self.systems.pop(job, None)

    def remaining(self, job):
        r = _u64()
        self._check(self.lib.pd_job_remaining(self.ctx, job, C.byref(r)), "job_remaining")
        return r.value

    def status(self, job):
        """(steps remaining, None) or, for a job the library stopped, (0, why)."""
        r = _u64()
        rc = self.lib.pd_job_remaining(self.ctx, job, C.byref(r))
        if rc == PD_E_OVERFLOW:
            return 0, (self.f["last_error"](self.ctx) or b"").decode(errors="replace")
        self._check(rc, "job_remaining")
        return r.value, None

    def step(self, n=1):
        self._check(self.f["step"](self.ctx, n), "step")

    def state(self, job, p):
        n = self.systems[job].particles[p].n
        out = {k: np.empty((n, 3)) for k in ("disp", "currpos", "vel", "acc", "force")}
        self._check(self.lib.pd_job_get_state(self.ctx, job, p, *[
            _ptr(out[k]) for k in ("disp", "currpos", "vel", "acc", "force")]), "job_get_state")
        return out

    def connectivity(self, job, p):
        return _read_connectivity(
            lambda buf, cap, m: self.lib.pd_job_get_connectivity(self.ctx, job, p, buf, cap, m),
            len(self.systems[job].particles[p].conn), self._check)

    def wall(self, job):
        lrtb, reac = np.empty(6), np.empty(18)
        self._check(self.f["get_wall"](self.ctx, job, _ptr(lrtb), _ptr(reac)), "get_wall")
        return lrtb, reac.reshape(6, 3)

    def set_flag(self, job, flag, value):
        self._check(self.lib.pd_job_set_flag(self.ctx, job, flag, int(value)), "job_set_flag")

    def counters(self):
        c = Counters()
        self._check(self.f["get_counters"](self.ctx, C.byref(c)), "get_counters")
        return c.as_dict()

    def set_graphs(self, on=True):
        self._check(self.lib.pd_set_graphs(self.ctx, int(on)), "set_graphs")

    def set_bond_shape(self, shape):
        """Force a bond-kernel shape (see bond_shapes()); -1 = autotune."""
        self._check(self.lib.pd_set_bond_shape(self.ctx, int(shape)), "set_bond_shape")

    def bond_shape(self):
        v = C.c_int()
        self._check(self.lib.pd_get_bond_shape(self.ctx, C.byref(v)), "get_bond_shape")
        return v.value

    def autotune(self, reps=3):
        """Time every bond-kernel shape on the resident state; returns (best, ms per shape)."""
        n = self.lib.pd_bond_shape_count()
        ms = np.zeros(n)
        best = C.c_int()
        self._check(self.lib.pd_autotune(self.ctx, reps, C.byref(best), _ptr(ms)), "autotune")
        return best.value, ms

    def counters(self):
        c = Counters()
        self._check(self.f["get_counters"](self.ctx, C.byref(c)), "get_counters")
        return c.as_dict()

    def last_step_ms(self):
        ms = C.c_double()
        self._check(self.lib.pd_last_step_ms(self.ctx, C.byref(ms)), "last_step_ms")
        return ms.value

    def set_profiling(self, on=True):
        self._check(self.lib.pd_set_profiling(self.ctx, int(on)), "set_profiling")

    def profile(self):
        pr = Profile()
        self._check(self.lib.pd_get_profile(self.ctx, C.byref(pr)), "get_profile")
        return pr.as_dict()

    def close(self):
        if self.ctx:
            self.f["destroy"](self.ctx)
            self.ctx = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def bond_shapes():
    """[(unroll, pipeline, staged)] of the bond-kernel shapes the library was
    built with (pipeline: 0 none, 1 software-pipelined, 2 pipelined and unrolled
    twice, 3 pipelined with four lanes per node; staged: block node tables and
    static records in LDS)."""
    lib, _ = load_gpu_lib()
    return [bond_shape_info(k) for k in range(lib.pd_bond_shape_count())]


def bond_shape_info(k):
    """(unroll, pipeline, staged) of bond shape k."""
    lib, _ = load_gpu_lib()
    u, p, st = C.c_int(), C.c_int(), C.c_int()
    if lib.pd_bond_shape_info(k, C.byref(u), C.byref(p), C.byref(st)) != PD_OK:
        raise ValueError(f"no bond shape {k}")
    return (u.value, p.value, bool(st.value))


def wide_bond_shape():
    """Shape index of the unstaged bond kernel with 64-bit record offsets
    (what pools of 2^27 node slots or more use); never tuned, but it can be
    forced with set_bond_shape on any pool."""
    lib, _ = load_gpu_lib()
    return lib.pd_bond_shape_count()


def selftest_math(n, seed=1):
    """(division mismatches, sqrt mismatches) of the fast-path arithmetic."""
    lib, _ = load_gpu_lib()
    md, ms = _u64(), _u64()
    rc = lib.pd_selftest_math(n, seed, C.byref(md), C.byref(ms))
    if rc != PD_OK:
        raise PDError(f"pd_selftest_math: {ERRORS.get(rc, rc)}")
    return md.value, ms.value


def selftest_math_noalloc(n, seed=1):
    """(division mismatches, sqrt mismatches, samples evaluated) from the
    allocation-free entry point (counters in the library's static scratch)."""
    lib, _ = load_gpu_lib()
    md, ms, ev = _u64(), _u64(), _u64()
    rc = lib.pd_selftest_math_noalloc(n, seed, C.byref(md), C.byref(ms), C.byref(ev))
    if rc != PD_OK:
        lib.pd_last_error.restype = C.c_char_p
        raise PDError(f"pd_selftest_math_noalloc: {ERRORS.get(rc, rc)}")
    return md.value, ms.value, ev.value


def gpu_engine(system):
    lib, fns = load_gpu_lib()
    return Engine(lib, fns, system, "pd_")


class CServer:
    """The library's own request loop (pd_serve_*) over a Pool: submit
    requests, poll until they finish, read a finished one's results from its
    pool slot (Pool.state / connectivity / wall), release it."""

    def __init__(self, pool, chunk=256, prepare_threads=2):
        self.pool = pool
        self.lib = pool.lib
        self.ptr = C.c_void_p()
        pool._check(self.lib.pd_serve_create(pool.ctx, chunk, prepare_threads, C.byref(self.ptr)), "serve_create")
        self.systems = {}

    def submit(self, system, nsteps):
        keep = []
        arr = _descs(system, keep)
        rd = system.run_desc()
        rid = _u64()
        self.pool._check(self.lib.pd_serve_submit(self.ptr, C.byref(rd), arr, len(system.particles), nsteps,
                                                  C.byref(rid)), "serve_submit")
        self.systems[rid.value] = system
        return rid.value

    def poll(self, cap=64):
        ev = (ServeEvent * cap)()
        n = _u32()
        self.pool._check(self.lib.pd_serve_poll(self.ptr, ev, cap, C.byref(n)), "serve_poll")
        out = []
        for k in range(n.value):
            e = ev[k]
            d = {f: getattr(e, f) for f, _ in ServeEvent._fields_}
            if e.job != 0xFFFFFFFF:
                self.pool.systems[e.job] = self.systems[e.request]
            out.append(d)
        return out

    def pending(self):
        q, r, u = _u64(), _u64(), _u64()
        self.pool._check(self.lib.pd_serve_pending(self.ptr, C.byref(q), C.byref(r), C.byref(u)), "serve_pending")
        return q.value, r.value, u.value

    def release(self, job):
        self.pool._check(self.lib.pd_serve_release(self.ptr, job), "serve_release")
        self.pool.systems.pop(job, None)

    def close(self):
        if self.ptr:
            self.lib.pd_serve_destroy(self.ptr)
            self.ptr = C.c_void_p()
