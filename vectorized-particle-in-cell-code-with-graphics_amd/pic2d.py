"""pic2d -- Python binding of the MI355X-native 2-D electrostatic PIC timestep.

A thin ctypes layer over ``libpic2d.so`` (the C ABI declared in
``include/pic2d.h``).  The classes mirror the reference's C++ objects so that
driving code reads like the reference's ``main.cpp`` / ``pic.cpp``:

* :class:`Params`  -- ``parameters.h`` (parameters.h:41-366) as a runtime struct;
* :class:`Pic`     -- class ``pic`` (pic.h:89-214): ``pic_run``, ``dump_data``,
  ``setRuntime`` and the per-phase methods of ``particle``/``fields``;
* :func:`mglin`    -- ``Mglin(u, ncycle, length)`` (mglin.cpp:21-55).

Every call raises :class:`PicError` (carrying the library's status code and
message) on failure; there is no CPU fallback.
"""
from __future__ import annotations

import ctypes as C
import os
import re

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libpic2d.so")
HEADER_PATH = os.path.join(os.path.dirname(_HERE), "include", "pic2d.h")

PIC_OK = 0
ERRORS = {
    -1: "PIC_ERR_ARG",
    -2: "PIC_ERR_GRID",
    -3: "PIC_ERR_FILE",
    -4: "PIC_ERR_FORMAT",
    -5: "PIC_ERR_DEVICE",
    -6: "PIC_ERR_CAPACITY",
    -7: "PIC_ERR_UNSUPPORTED",
    -8: "PIC_ERR_ALLOC",
    -9: "PIC_ERR_NUMERIC",
    -10: "PIC_ERR_TIMEOUT",
}
MODE_REF, MODE_PHYSICAL = 0, 1
MG_FIXED, MG_TOL = 0, 1
LOAD_RANDOM, LOAD_LATTICE = 0, 1
WALL_ABSORB, WALL_STEADY = 0, 1
GAS_ARGON, GAS_HYDROGEN = 0, 1
SIGMA_CONSTANT, SIGMA_ENERGY = 0, 1
ELECTRONS, IONS = 0, 1


class PicError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"{ERRORS.get(code, code)}: {msg}")
        self.code = code


class Params(C.Structure):
    """Mirror of ``pic_params`` (include/pic2d.h)."""

    _fields_ = [
        ("nodes", C.c_int32),
        ("init_particle_no", C.c_int32),
        ("mode", C.c_int32),
        ("loader", C.c_int32),
        ("wall", C.c_int32),
        ("have_collisions", C.c_int32),
        ("gas", C.c_int32),
        ("sigma_cx", C.c_int32),
        ("sigma_elastic", C.c_int32),
        ("sigma_ionization", C.c_int32),
        ("mg_mode", C.c_int32),
        ("mg_ncycle", C.c_int32),
        ("mg_max_iter", C.c_int32),
        ("dt_dti", C.c_int32),
        ("run_time", C.c_int32),
        ("number_to_inject", C.c_int32),
        ("no_ion_species", C.c_int32),
        ("corner8", C.c_int32),
        ("seed", C.c_uint64),
        ("length", C.c_double),
        ("dt", C.c_double),
        ("dti", C.c_double),
        ("electron_mass", C.c_double),
        ("ion_mass", C.c_double),
        ("electron_charge", C.c_double),
        ("ion_charge", C.c_double),
        ("qce", C.c_double),
        ("qci", C.c_double),
        ("vte", C.c_double),
        ("vti", C.c_double),
        ("vtn", C.c_double),
        ("sigma_load_e", C.c_double),
        ("sigma_load_i", C.c_double),
        ("epsilon", C.c_double),
        ("neutral_density", C.c_double),
        ("threshold", C.c_double),
        ("sigma_value", C.c_double),
        ("elastic_sigma_value", C.c_double),
        ("ionization_sigma_value", C.c_double),
        ("magnetic_field", C.c_double),
        ("angle", C.c_double),
        ("phi_left", C.c_double),
        ("phi_right", C.c_double),
        ("phi_bottom", C.c_double),
        ("phi_top", C.c_double),
        ("phi_period", C.c_double),
        ("mg_tol", C.c_double),
        ("sort_interval", C.c_int32),
        ("graph_start", C.c_int32),
        ("graph_steps", C.c_int32),
        ("time_data_freq", C.c_int32),
        ("no_bins", C.c_int32),
        ("divide_bins_by", C.c_int32),
        ("spread", C.c_int32),
        ("dist_loc", C.c_int32 * 4),
        ("probe", C.c_int32 * 6),
        ("write_graphs", C.c_int32),
        ("reduce", C.c_int32),
        ("no_groups", C.c_int32),
        ("size_to_check", C.c_int32 * 2),
        ("group_limit", (C.c_int32 * 8) * 2),
        ("new_group_size", (C.c_int32 * 8) * 2),
        ("v_limit", (C.c_double * 8) * 2),
        ("e_impact_see", C.c_int32),
        ("ion_impact_see", C.c_int32),
        ("reflection", C.c_int32),
        ("backscatter", C.c_int32),
        ("see_max_yield", C.c_double),
        ("see_E0", C.c_double),
        ("see_Emax0", C.c_double),
        ("work_function", C.c_double),
        ("sort_key", C.c_int32),
        ("mg_sweeps", C.c_int32),
        ("null_bound", C.c_int32),
        ("graphs", C.c_int32),
    ]

    @property
    def dx(self):
        return self.length / (self.nodes - 1)

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class ServeRequest(C.Structure):
    """Mirror of ``pic_serve_request``."""
    _fields_ = [("params", Params), ("steps", C.c_int32), ("arrival_s", C.c_double)]


class ServeResult(C.Structure):
    """Mirror of ``pic_serve_result``."""
    _fields_ = [("admit_s", C.c_double), ("finish_s", C.c_double),
                ("particle_steps", C.c_int64), ("step_p50_ms", C.c_double),
                ("step_p99_ms", C.c_double), ("status", C.c_int32), ("phi_error", C.c_double),
                ("e_error", C.c_double)]


class ServeStats(C.Structure):
    """Mirror of ``pic_serve_stats``."""
    _fields_ = [("wall_s", C.c_double), ("particle_steps", C.c_int64),
                ("throughput", C.c_double), ("step_p50_ms", C.c_double),
                ("step_p99_ms", C.c_double), ("peak_bytes", C.c_int64),
                ("max_concurrent", C.c_int32)]


PARTICLE_DTYPE = np.dtype(
    [("x", "f8"), ("y", "f8"), ("vx", "f8"), ("vy", "f8"), ("vz", "f8"), ("weight", "f8")]
)

_lib = None


class ServeStatus(C.Structure):
    _fields_ = [("state", C.c_int32), ("error", C.c_int32), ("steps_done", C.c_int32),
                ("steps", C.c_int32), ("submit_s", C.c_double), ("admit_s", C.c_double),
                ("finish_s", C.c_double), ("step_p50_ms", C.c_double),
                ("step_p99_ms", C.c_double), ("particle_steps", C.c_int64),
                ("phi_error", C.c_double), ("e_error", C.c_double), ("fingerprint", C.c_uint64),
                ("device_bytes", C.c_int64)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


REQ_QUEUED, REQ_RUNNING, REQ_DONE, REQ_CANCELLED, REQ_FAILED = 0, 1, 2, 3, 4
PIC_ERR_ARG, PIC_ERR_FILE, PIC_ERR_FORMAT = -1, -3, -4
PIC_ERR_CAPACITY, PIC_ERR_NUMERIC, PIC_ERR_TIMEOUT = -6, -9, -10


def declared_functions(header=HEADER_PATH):
    """Names of every function declared in include/pic2d.h."""
    text = open(header).read()
    text = re.sub(r"/\*.*?\*/", " ", text, flags=re.S)
    return sorted(set(re.findall(r"^\s*(?:const\s+)?\w+\s*\*?\s*(pic_\w+)\s*\(", text, flags=re.M)))


def load_library(path=LIB_PATH):
    """Load libpic2d.so; raises if it is missing (no fallback exists)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise PicError(-5, f"{path} is not built; run __graft_entry__.build()")
    lib = C.CDLL(path)
    P = C.POINTER
    vp = C.c_void_p
    i32, i64, dbl = C.c_int32, C.c_int64, C.c_double
    sig = {
        "pic_params_default": [P(Params)],
        "pic_params_load": [C.c_char_p, P(Params)],
        "pic_create": [P(Params), C.c_int, C.c_char_p, P(vp)],
        "pic_destroy": [vp],
        "pic_set_runtime": [vp, i32],
        "pic_run": [vp],
        "pic_dump_data": [vp, C.c_char_p, i32],
        "pic_checkpoint_write": [vp, C.c_char_p],
        "pic_checkpoint_params": [C.c_char_p, vp],
        "pic_checkpoint_read": [vp, C.c_char_p],
        "pic_set_tuning": [vp, i32, i32, i32],
        "pic_get_tuning": [vp, vp],
        "pic_species_gather": [vp, i32],
        "pic_species_move": [vp, i32],
        "pic_species_collect_charge": [vp, i32],
        "pic_fields_grid_rho": [vp],
        "pic_fields_grid_potential": [vp, dbl, dbl, dbl, dbl, dbl, dbl],
        "pic_fields_grid_electric_field": [vp],
        "pic_e_collisions": [vp],
        "pic_ion_collisions": [vp],
        "pic_boundary": [vp, i32],
        "pic_mglin": [vp, i32, i32, dbl],
        "pic_particle_size": [vp, i32, P(i64)],
        "pic_get_particles": [vp, i32, vp, i64],
        "pic_get_phi": [vp, vp],
        "pic_get_rho": [vp, vp],
        "pic_get_E": [vp, vp, vp],
        "pic_get_q": [vp, i32, vp],
        "pic_get_step": [vp, P(i64)],
        "pic_get_loss_counts": [vp, i32, vp],
        "pic_get_nodes": [vp, P(i32)],
        "pic_get_state": [vp, i32, vp, vp, vp, vp, vp, vp, vp, i64],
        "pic_set_state": [vp, i32, vp, vp, vp, vp, vp, vp, vp, i64],
        "pic_get_qfixed": [vp, i32, vp],
        "pic_get_E32": [vp, vp],
        "pic_set_E32": [vp, vp],
        "pic_profile_enable": [vp, i32],
        "pic_profile_read": [vp, vp, C.c_int32],
        "pic_profile_steps": [vp, vp, i32, P(i32)],
        "pic_advance": [vp, i32],
        "pic_poisson_tol": [vp, vp, i32, dbl, dbl, i32, P(i32), P(dbl)],
        "pic_set_step": [vp, i64],
        "pic_get_collision_counts": [vp, i32, vp],
        "pic_get_overflow": [vp, P(i64)],
        "pic_get_mg_stats": [vp, P(i32), P(i64), P(dbl)],
        "pic_sync": [vp],
        "pic_set_output_dir": [vp, C.c_char_p],
        "pic_get_field_averages": [vp, vp, vp, vp, vp],
        "pic_get_density_average": [vp, i32, vp],
        "pic_get_velocity_slices": [vp, i32, vp, vp],
        "pic_get_distributions": [vp, i32, vp],
        "pic_get_time_data": [vp, vp, i32, P(i32)],
        "pic_write_graphs": [vp, C.c_char_p],
        "pic_write_time_data": [vp, C.c_char_p],
        "pic_get_graph_max": [vp, vp],
        "pic_get_particle_subset": [vp, i32, i32, vp, i64, P(i64)],
        "pic_write_snapshot": [vp, C.c_char_p, i32],
        "pic_get_weights": [vp, i32, vp, i64],
        "pic_set_weights": [vp, i32, vp, i64],
        "pic_inject_particles": [vp, i32],
        "pic_contain_particle_growth": [vp, i32],
        "pic_get_population_stats": [vp, i32, vp],
        "pic_get_wall_counts": [vp, i32, vp],
        "pic_particle_count": [vp, i32, P(i64)],
        "pic_get_device_bytes": [vp, P(i64)],
        "pic_estimate_bytes": [P(Params), P(i64)],
        "pic_batch_advance": [P(vp), i32, i32],
        "pic_serve": [P(ServeRequest), i32, i64, i32, P(ServeResult), P(ServeStats)],
        "pic_cross_section": [i32, i32, i32, dbl, P(dbl)],
        "pic_field_error": [vp, dbl, vp],
        "pic_mglin_noalloc": [vp, i32, i32, dbl, vp],
        "pic_field_finite": [vp, P(i32)],
        "pic_server_create": [i64, i32, P(vp)],
        "pic_server_destroy": [vp],
        "pic_serve_submit": [vp, P(Params), i32, P(i64)],
        "pic_serve_poll": [vp, i64, P(ServeStatus)],
        "pic_serve_cancel": [vp, i64],
        "pic_serve_wait": [vp, i64, dbl, P(ServeStatus)],
        "pic_set_accuracy_report": [vp, i32, dbl],
        "pic_get_accuracy_report": [vp, vp, i32, P(i32)],
    }
    for name, args in sig.items():
        f = getattr(lib, name)
        f.argtypes = args
        f.restype = C.c_int
    lib.pic_last_error.argtypes = [vp]
    lib.pic_last_error.restype = C.c_char_p
    _lib = lib
    return lib


def _ptr(a):
    return a.ctypes.data_as(C.c_void_p) if a is not None else None


def _check(rc, h=None):
    if rc != PIC_OK:
        msg = load_library().pic_last_error(h)
        raise PicError(rc, msg.decode() if msg else "")


def default_params() -> Params:
    p = Params()
    _check(load_library().pic_params_default(C.byref(p)))
    return p


def load_params(path) -> Params:
    """Evaluate a parameters.h-style file (parameters.h:41-366)."""
    p = Params()
    _check(load_library().pic_params_load(str(path).encode(), C.byref(p)))
    return p


def mglin(u, ncycle=2, length=1.0):
    """GPU FP64 full multigrid (mglin.cpp:21-55); returns the solution."""
    a = np.ascontiguousarray(u, dtype=np.float64).copy()
    n = a.shape[0]
    _check(load_library().pic_mglin(_ptr(a), n, ncycle, length))
    return a


XS_ELASTIC, XS_EXCITATION, XS_IONIZATION, XS_CX, XS_ION_ISO = 0, 1, 2, 3, 4


def checkpoint_params(path) -> Params:
    """the parameters a checkpoint was written with"""
    p = Params()
    _check(load_library().pic_checkpoint_params(str(path).encode(), C.byref(p)))
    return p


def cross_section(gas, channel, energy, mode=MODE_PHYSICAL):
    """Cross section [1e-20 m^2] the collision kernels use at `energy` eV
    (e_collisions::sigma_Ar_elastic and its siblings, e_collisions.h:192-195);
    host-side FP64, no GPU needed.  Accepts a scalar or an array."""
    lib = load_library()
    out = C.c_double()
    vals = []
    for e in np.atleast_1d(np.asarray(energy, dtype=np.float64)):
        _check(lib.pic_cross_section(gas, channel, mode, float(e), C.byref(out)))
        vals.append(out.value)
    return vals[0] if np.ndim(energy) == 0 else np.array(vals)


def poisson_tol(f, u0=None, length=1.0, tol=1e-10, max_iter=60):
    """GPU mixed-precision multigrid to a relative residual; returns (u, cycles, relres)."""
    f = np.ascontiguousarray(f, dtype=np.float64)
    u = np.zeros_like(f) if u0 is None else np.ascontiguousarray(u0, dtype=np.float64).copy()
    cyc = C.c_int32()
    rr = C.c_double()
    _check(load_library().pic_poisson_tol(_ptr(f), _ptr(u), f.shape[0], length, tol, max_iter,
                                          C.byref(cyc), C.byref(rr)))
    return u, cyc.value, rr.value


def estimate_bytes(params: Params) -> int:
    """upper estimate of the device memory a run of these parameters takes"""
    b = C.c_int64()
    _check(load_library().pic_estimate_bytes(C.byref(params), C.byref(b)))
    return b.value


def batch_advance(sims, steps):
    """advance every run by `steps` steps at once (one host thread per run)"""
    arr = (C.c_void_p * len(sims))(*[s.h.value for s in sims])
    _check(load_library().pic_batch_advance(arr, len(sims), int(steps)))


def serve(requests, hbm_budget, max_active=64):
    """continuous-batching serving loop; requests = [(Params, steps, arrival_s)].
    Returns (list of result dicts, stats dict)."""
    n = len(requests)
    reqs = (ServeRequest * n)()
    for k, (p, steps, t) in enumerate(requests):
        reqs[k].params = p
        reqs[k].steps = int(steps)
        reqs[k].arrival_s = float(t)
    res = (ServeResult * n)()
    st = ServeStats()
    rc = load_library().pic_serve(reqs, n, int(hbm_budget), int(max_active), res, C.byref(st))
    out = [{k: getattr(r, k) for k, _ in ServeResult._fields_} for r in res]
    stats = {k: getattr(st, k) for k, _ in ServeStats._fields_}
    _check(rc)
    return out, stats


class Server:
    """the serving loop (pic_server_create ...): submit runs at any time, poll,
    cancel, wait; continuous batching under an HBM budget"""

    def __init__(self, hbm_budget, max_active=64):
        self.lib = load_library()
        self.s = C.c_void_p()
        _check(self.lib.pic_server_create(int(hbm_budget), int(max_active), C.byref(self.s)))

    def submit(self, params: Params, steps) -> int:
        rid = C.c_int64()
        _check(self.lib.pic_serve_submit(self.s, C.byref(params), int(steps), C.byref(rid)))
        return rid.value

    def poll(self, rid) -> dict:
        st = ServeStatus()
        _check(self.lib.pic_serve_poll(self.s, int(rid), C.byref(st)))
        return st.as_dict()

    def cancel(self, rid):
        _check(self.lib.pic_serve_cancel(self.s, int(rid)))

    def wait(self, rid, timeout=None) -> dict:
        """the request's status once DONE, CANCELLED or FAILED; raises
        PicError(PIC_ERR_TIMEOUT) if timeout (s) passes first"""
        st = ServeStatus()
        _check(self.lib.pic_serve_wait(self.s, int(rid), -1.0 if timeout is None else float(timeout),
                                       C.byref(st)))
        return st.as_dict()

    def close(self):
        if self.s:
            self.lib.pic_server_destroy(self.s)
            self.s = C.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Pic:
    """One simulation (the reference's ``pic`` object, pic.cpp:55-221)."""

    def __init__(self, params: Params, restart=False, dump_dir="."):
        self.lib = load_library()
        self.params = params
        self.h = C.c_void_p()
        _check(self.lib.pic_create(C.byref(params), int(restart), str(dump_dir).encode(),
                                   C.byref(self.h)))
        n = C.c_int32()
        _check(self.lib.pic_get_nodes(self.h, C.byref(n)), self.h)
        self.N = n.value

    # -- lifecycle
    def close(self):
        if self.h:
            self.lib.pic_destroy(self.h)
            self.h = C.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _c(self, rc):
        _check(rc, self.h)

    # -- pic.h
    def setRuntime(self, run_time):
        self._c(self.lib.pic_set_runtime(self.h, int(run_time)))

    def pic_run(self):
        self._c(self.lib.pic_run(self.h))

    def run(self, steps):
        self.setRuntime(steps)
        self.pic_run()

    def advance(self, steps):
        """the step loop without pic_run's pre-loop wall pass"""
        self._c(self.lib.pic_advance(self.h, int(steps)))

    def dump_data(self, directory=".", precision=17):
        self._c(self.lib.pic_dump_data(self.h, str(directory).encode(), int(precision)))

    def checkpoint(self, path):
        """exact-resume file: Pic.resume(path) continues bit for bit"""
        self._c(self.lib.pic_checkpoint_write(self.h, str(path).encode()))

    def set_tuning(self, sort_interval_e=0, sort_interval_i=0, mg_sweeps=0):
        """passes per counting sort of each species, tolerance-solve sweeps
        per leg (0: leave as is); results do not depend on them"""
        self._c(self.lib.pic_set_tuning(self.h, int(sort_interval_e), int(sort_interval_i),
                                        int(mg_sweeps)))

    def tuning(self):
        """{sort_interval_e, sort_interval_i, sort_key, mg_sweeps} in effect"""
        a = (C.c_int32 * 4)()
        self._c(self.lib.pic_get_tuning(self.h, a))
        return dict(zip(("sort_interval_e", "sort_interval_i", "sort_key", "mg_sweeps"), a))

    @classmethod
    def resume(cls, path, params=None):
        """a run created from a checkpoint's parameters (or `params`, which
        may differ in run_time, write_graphs and graphs) and filled from it"""
        p = checkpoint_params(path) if params is None else params
        sim = cls(p)
        try:
            sim._c(sim.lib.pic_checkpoint_read(sim.h, str(path).encode()))
        except Exception:
            sim.close()
            raise
        return sim

    def sync(self):
        self._c(self.lib.pic_sync(self.h))

    # -- per-phase methods (particle.h / fields.h / collisions / boundary)
    def get_particle_electric_field(self, s):
        self._c(self.lib.pic_species_gather(self.h, s))

    def move_particles(self, s):
        self._c(self.lib.pic_species_move(self.h, s))

    def collect_charge(self, s):
        self._c(self.lib.pic_species_collect_charge(self.h, s))

    def grid_rho(self):
        self._c(self.lib.pic_fields_grid_rho(self.h))

    def grid_potential(self, phi0X, phiLX, phi0Y, phiLY, timestep, phi_period):
        self._c(self.lib.pic_fields_grid_potential(self.h, phi0X, phiLX, phi0Y, phiLY,
                                                   float(timestep), phi_period))

    def grid_electric_field(self):
        self._c(self.lib.pic_fields_grid_electric_field(self.h))

    def e_collisions(self):
        self._c(self.lib.pic_e_collisions(self.h))

    def ion_collisions(self):
        self._c(self.lib.pic_ion_collisions(self.h))

    def boundary(self, which):
        """which = 0: ion wall (implement_i_boundary), 1: electron wall."""
        self._c(self.lib.pic_boundary(self.h, which))

    # -- getters
    def particle_size(self, s):
        n = C.c_int64()
        self._c(self.lib.pic_particle_size(self.h, s, C.byref(n)))
        return n.value

    def get_particles(self, s):
        n = self.particle_size(s)
        out = np.zeros(n, dtype=PARTICLE_DTYPE)
        self._c(self.lib.pic_get_particles(self.h, s, _ptr(out), n))
        return out

    def get_state(self, s):
        n = self.particle_size(s)
        d = {k: np.zeros(n, dtype=np.uint32 if k in ("cell", "id") else np.float32)
             for k in ("cell", "ox", "oy", "vx", "vy", "vz", "id")}
        self._c(self.lib.pic_get_state(self.h, s, *[_ptr(d[k]) for k in
                                                    ("cell", "ox", "oy", "vx", "vy", "vz", "id")], n))
        return d

    def set_state(self, s, d):
        arrs = [np.ascontiguousarray(d[k], dtype=np.uint32 if k in ("cell", "id") else np.float32)
                for k in ("cell", "ox", "oy", "vx", "vy", "vz", "id")]
        self._c(self.lib.pic_set_state(self.h, s, *[_ptr(a) for a in arrs], len(arrs[0])))

    def get_phi(self):
        a = np.zeros((self.N, self.N))
        self._c(self.lib.pic_get_phi(self.h, _ptr(a)))
        return a

    def get_rho(self):
        a = np.zeros((self.N, self.N))
        self._c(self.lib.pic_get_rho(self.h, _ptr(a)))
        return a

    def get_E(self):
        ex = np.zeros((self.N + 1) * (self.N + 1))
        ey = np.zeros_like(ex)
        self._c(self.lib.pic_get_E(self.h, _ptr(ex), _ptr(ey)))
        return ex, ey

    def get_E32(self):
        a = np.zeros(2 * ((self.N + 1) * self.N + 1), dtype=np.float32)
        self._c(self.lib.pic_get_E32(self.h, _ptr(a)))
        return a

    def set_E32(self, a):
        a = np.ascontiguousarray(a, dtype=np.float32)
        self._c(self.lib.pic_set_E32(self.h, _ptr(a)))

    def get_q(self, s):
        a = np.zeros((self.N, self.N))
        self._c(self.lib.pic_get_q(self.h, s, _ptr(a)))
        return a

    def get_qfixed(self, s):
        a = np.zeros((self.N, self.N), dtype=np.int64)
        self._c(self.lib.pic_get_qfixed(self.h, s, _ptr(a)))
        return a

    def get_step(self):
        v = C.c_int64()
        self._c(self.lib.pic_get_step(self.h, C.byref(v)))
        return v.value

    def mg_stats(self):
        """(V-cycles of the last solve, cumulative V-cycles, last relative residual)"""
        a, b, c = C.c_int32(), C.c_int64(), C.c_double()
        self._c(self.lib.pic_get_mg_stats(self.h, C.byref(a), C.byref(b), C.byref(c)))
        return a.value, b.value, c.value

    def get_weights(self, s):
        n = self.particle_size(s)
        w = np.zeros(n, dtype=np.float32)
        self._c(self.lib.pic_get_weights(self.h, s, _ptr(w), n))
        return w

    def set_weights(self, s, w):
        w = np.ascontiguousarray(w, dtype=np.float32)
        self._c(self.lib.pic_set_weights(self.h, s, _ptr(w), len(w)))

    def inject_particles(self, s):
        """particle::inject_more_particles"""
        self._c(self.lib.pic_inject_particles(self.h, s))

    def contain_particle_growth(self, s):
        """particle::contain_particle_growth"""
        self._c(self.lib.pic_contain_particle_growth(self.h, s))

    def population_stats(self, s):
        """[reductions performed, particles removed by them, particles injected]"""
        a = np.zeros(3, dtype=np.int64)
        self._c(self.lib.pic_get_population_stats(self.h, s, _ptr(a)))
        return a

    def particle_count(self, s):
        """live particles without compacting the store"""
        n = C.c_int64()
        self._c(self.lib.pic_particle_count(self.h, s, C.byref(n)))
        return n.value

    def device_bytes(self):
        n = C.c_int64()
        self._c(self.lib.pic_get_device_bytes(self.h, C.byref(n)))
        return n.value

    def wall_counts(self, s):
        """[reached a wall, absorbed, reflected, backscattered, emitting,
        secondary electrons created]"""
        a = np.zeros(6, dtype=np.int64)
        self._c(self.lib.pic_get_wall_counts(self.h, s, _ptr(a)))
        return a

    def loss_counts(self, s):
        """[left, right, bottom, top, removed by collision]"""
        a = np.zeros(5, dtype=np.int64)
        self._c(self.lib.pic_get_loss_counts(self.h, s, _ptr(a)))
        return a

    def collision_counts(self, s):
        """events by kind: [-, elastic, excitation, ionisation, charge exchange,
        isotropic ion elastic, -, -]"""
        a = np.zeros(8, dtype=np.int64)
        self._c(self.lib.pic_get_collision_counts(self.h, s, _ptr(a)))
        return a

    def field_error(self, tol=1e-12):
        """relative L2 errors of phi, Ex, Ey against an FP64 multigrid solve of
        the last right-hand side converged to tol; returns a dict"""
        a = np.zeros(5)
        self._c(self.lib.pic_field_error(self.h, float(tol), _ptr(a)))
        return dict(phi=a[0], ex=a[1], ey=a[2], ref_relres=a[3], ref_cycles=int(a[4]))

    def set_accuracy_report(self, every, tol=1e-12):
        self._c(self.lib.pic_set_accuracy_report(self.h, int(every), float(tol)))

    def accuracy_report(self, cap=1 << 14):
        """rows (step, solver relres, err phi, err Ex, err Ey, reference relres)"""
        out = np.zeros((cap, 6))
        n = C.c_int32()
        self._c(self.lib.pic_get_accuracy_report(self.h, _ptr(out), cap, C.byref(n)))
        return out[: min(n.value, cap)]

    def field_finite(self):
        f = C.c_int32()
        self._c(self.lib.pic_field_finite(self.h, C.byref(f)))
        return bool(f.value)

    def overflow(self):
        v = C.c_int64()
        self._c(self.lib.pic_get_overflow(self.h, C.byref(v)))
        return v.value

    def set_step(self, step):
        self._c(self.lib.pic_set_step(self.h, int(step)))

    # -- diagnostics (graphs.cpp / timeData)
    def set_output_dir(self, directory):
        self._c(self.lib.pic_set_output_dir(self.h, None if directory is None
                                            else str(directory).encode()))

    def field_averages(self):
        n = self.N
        a = [np.zeros((n, n)) for _ in range(4)]
        self._c(self.lib.pic_get_field_averages(self.h, *[_ptr(x) for x in a]))
        return dict(rho=a[0], phi=a[1], ex=a[2], ey=a[3])

    def density_average(self, s):
        a = np.zeros((self.N, self.N))
        self._c(self.lib.pic_get_density_average(self.h, s, _ptr(a)))
        return a

    def velocity_slices(self, s):
        vx, vy = np.zeros(self.N), np.zeros(self.N)
        self._c(self.lib.pic_get_velocity_slices(self.h, s, _ptr(vx), _ptr(vy)))
        return vx, vy

    def distributions(self, s, no_bins):
        a = np.zeros(4 * no_bins)
        self._c(self.lib.pic_get_distributions(self.h, s, _ptr(a)))
        return a.reshape(4, no_bins)

    def time_data(self, cap=1 << 16):
        rows = C.c_int32()
        self._c(self.lib.pic_get_time_data(self.h, None, 0, C.byref(rows)))
        a = np.zeros((max(rows.value, 1), 8))
        self._c(self.lib.pic_get_time_data(self.h, _ptr(a), rows.value, C.byref(rows)))
        return a[: rows.value]

    def write_graphs(self, directory="."):
        self._c(self.lib.pic_write_graphs(self.h, str(directory).encode()))

    def write_time_data(self, directory="."):
        self._c(self.lib.pic_write_time_data(self.h, str(directory).encode()))

    def graph_max(self):
        """[phi max, Ex max, Ey max, electron density max, ion density max]"""
        a = np.zeros(5)
        self._c(self.lib.pic_get_graph_max(self.h, _ptr(a)))
        return a

    # -- snapshot export
    def particle_subset(self, s, stride):
        cap = self.particle_size(s) // stride + 64
        a = np.zeros(cap, dtype=PARTICLE_DTYPE)
        n = C.c_int64()
        self._c(self.lib.pic_get_particle_subset(self.h, s, stride, _ptr(a), cap, C.byref(n)))
        return a[: n.value]

    def write_snapshot(self, directory=".", stride=1):
        self._c(self.lib.pic_write_snapshot(self.h, str(directory).encode(), int(stride)))

    # -- profiling
    def profile(self, on=True):
        self._c(self.lib.pic_profile_enable(self.h, int(on)))

    def profile_read(self):
        """[step ms, push-phase ms, scatter-kernel ms, field ms, count-kernel ms,
        scan ms, in-place-kernel ms, scatter launches, in-place launches, 0, 0, 0]
        summed over the profiled steps"""
        a = np.zeros(12)
        self._c(self.lib.pic_profile_read(self.h, _ptr(a), len(a)))
        return a

    def profile_steps(self, cap=1 << 16):
        a = np.zeros(cap)
        n = C.c_int32()
        self._c(self.lib.pic_profile_steps(self.h, _ptr(a), cap, C.byref(n)))
        return a[: n.value]


def cell_unpack(cell):
    """packed int16 cell -> (jx, jy) int arrays"""
    c = np.asarray(cell, dtype=np.uint32)
    jx = (c & 0xFFFF).astype(np.uint16).view(np.int16).astype(np.int32)
    jy = (c >> 16).astype(np.uint16).view(np.int16).astype(np.int32)
    return jx, jy


def cell_pack(jx, jy):
    jx = np.asarray(jx).astype(np.int16).view(np.uint16).astype(np.uint32)
    jy = np.asarray(jy).astype(np.int16).view(np.uint16).astype(np.uint32)
    return jx | (jy << 16)
