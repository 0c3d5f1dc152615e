"""ctypes binding of the CPU oracle (oracle/liboracle.so) -- test infrastructure.

The oracle is the checker: tests compare the HIP library against it.  Nothing
in the product imports this module.
"""
import ctypes as C
import os

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_SO = os.path.join(ROOT, "oracle", "liboracle.so")
REF_MGLIN_SO = os.path.join(ROOT, "oracle", "_ref", "libref_mglin.so")
# fixed-point charge: a unit-weight particle deposits 2^FX_TOTAL_BITS in all
FX_TOTAL_BITS = 48

vp = C.c_void_p
MODE_REF, MODE_PHYS, MODE_REF_STEADY = 0, 1, 2
ST_KEPT, ST_LEFT, ST_RIGHT, ST_BOTTOM, ST_TOP = range(5)

PARTICLE_DTYPE = np.dtype(
    [("x", "f8"), ("y", "f8"), ("vx", "f8"), ("vy", "f8"), ("vz", "f8"), ("weight", "f8")]
)


class PushParams(C.Structure):
    _fields_ = [("N", C.c_int), ("mode", C.c_int), ("cE", C.c_float), ("dtdx", C.c_float),
                ("dtdy", C.c_float), ("bfield", C.c_int), ("cos_a", C.c_float),
                ("sin_a", C.c_float), ("cos_w", C.c_float), ("sin_w", C.c_float)]


class Cfg32(C.Structure):
    _fields_ = [("N", C.c_int), ("mode", C.c_int), ("corner8", C.c_int), ("mg_mode", C.c_int),
                ("mg_ncycle", C.c_int), ("mg_max_cycles", C.c_int), ("dt_dti", C.c_int),
                ("preloop", C.c_int), ("L", C.c_double), ("dx", C.c_double),
                ("eps", C.c_double), ("mg_tol", C.c_double), ("phi_left", C.c_double),
                ("phi_right", C.c_double), ("phi_bottom", C.c_double), ("phi_top", C.c_double),
                ("phi_period", C.c_double), ("scale_e", C.c_double), ("scale_i", C.c_double),
                ("start_step", C.c_long), ("pe", PushParams), ("pi", PushParams),
                ("key0", C.c_uint32), ("key1", C.c_uint32), ("steady_se", C.c_float),
                ("steady_si", C.c_float)]


class CfgRef(C.Structure):
    _fields_ = [("N", C.c_int), ("ncycle", C.c_int), ("dt_dti", C.c_int), ("pad", C.c_int)] + [
        (k, C.c_double) for k in ("L", "dx", "dy", "eps", "coef_e", "coef_i", "dt_e", "dt_i",
                                  "qc_dxdy_e", "qc_dxdy_i", "cos_a", "sin_a", "cos_we", "sin_we",
                                  "cos_wi", "sin_wi", "bfield", "phi_left", "phi_right",
                                  "phi_bottom", "phi_top", "phi_period")
    ] + [("start_step", C.c_long), ("mg_tol", C.c_double)]


class MccParams(C.Structure):
    _fields_ = [("gas", C.c_int), ("electron", C.c_int)] + [
        (k, C.c_double) for k in ("ng", "dt", "mass", "M", "q", "vtn", "e_exc", "e_iz", "opal_B",
                                  "kev_amu_div", "nu_max", "p_null")
    ] + [("key0", C.c_uint32), ("key1", C.c_uint32), ("binned", C.c_int), ("pad_", C.c_int),
         ("bin_p", C.c_double * 64), ("bin_nu", C.c_double * 64)]


class RefCxParams(C.Structure):
    _fields_ = [("gas", C.c_int), ("sig_energy", C.c_int), ("sig_const", C.c_float),
                ("q", C.c_float), ("mass", C.c_float), ("dt", C.c_float), ("ng", C.c_float),
                ("key0", C.c_uint32), ("key1", C.c_uint32)]


class ReduceParams(C.Structure):
    _fields_ = [("ng", C.c_int), ("size_check", C.c_uint32), ("species", C.c_uint32),
                ("key0", C.c_uint32), ("key1", C.c_uint32), ("vlim", C.c_double * 8),
                ("glimit", C.c_uint32 * 8), ("gnew", C.c_uint32 * 8)]


class WallParams(C.Structure):
    _fields_ = [("mode", C.c_int), ("N", C.c_int)] + [
        (k, C.c_double) for k in ("dx", "Lw", "dt", "me", "mi", "q", "max_yield", "E0", "Emax0",
                                  "thr", "wf")] + [
        (k, C.c_int) for k in ("see_e", "refl", "back", "see_i")] + [
        ("key0", C.c_uint32), ("key1", C.c_uint32)]


class WallOut(C.Structure):
    _fields_ = [("cell", vp)] + [(k, vp) for k in ("ox", "oy", "vx", "vy", "vz", "w")] + [
        ("key", vp)]


_lib = None


def lib():
    global _lib
    if _lib is None:
        _lib = C.CDLL(ORACLE_SO)
        _lib.orc_u01.restype = C.c_float
        _lib.orc_u01.argtypes = [C.c_uint32]
        _lib.orc_mglin.argtypes = [vp, C.c_int, C.c_int, C.c_double]
        _lib.orc_mg_tol.argtypes = [vp, C.c_int, C.c_double, C.c_double, C.c_int, vp]
        _lib.orc_mg_relres.argtypes = [vp, vp, C.c_int, C.c_double]
        _lib.orc_mg_relres.restype = C.c_double
        _lib.orc_ref_gather.argtypes = [vp, C.c_long, vp, vp, C.c_double, C.c_double, C.c_int,
                                        C.c_int, vp, vp]
        _lib.orc_ref_push.argtypes = [vp, C.c_long, vp, vp, C.c_double, C.c_double, C.c_int,
                                      C.c_int] + [C.c_double] * 7
        _lib.orc_ref_deposit.argtypes = [vp, C.c_long, C.c_double, C.c_double, C.c_double,
                                         C.c_int, C.c_int, vp]
        _lib.orc_ref_charge_postscale.argtypes = [vp, C.c_int, C.c_int]
        _lib.orc_ref_grid_rho.argtypes = [vp, vp, C.c_int, vp]
        _lib.orc_ref_grid_potential.argtypes = [vp, C.c_int] + [C.c_double] * 9 + [C.c_int,
                                                                                  C.c_double, vp]
        _lib.orc_ref_grid_efield.argtypes = [vp, C.c_int, C.c_double, vp, vp]
        _lib.orc_ref_absorb.argtypes = [vp, vp, C.c_long, C.c_double, C.c_double, C.c_int,
                                        C.c_int]
        _lib.orc_ref_absorb.restype = C.c_long
        _lib.orc32_push.argtypes = [C.c_long] + [vp] * 7 + [C.POINTER(PushParams), vp]
        _lib.orc32_deposit.argtypes = [C.c_long, vp, vp, vp, C.c_int, C.c_int, vp]
        _lib.orc_fixed_to_rhs.argtypes = [vp, vp, C.c_int, C.c_double, C.c_double, C.c_int] + \
            [C.c_double] * 6 + [vp, vp]
        _lib.orc_efield_pack.argtypes = [vp, C.c_int, C.c_double, vp, vp, vp]
        _lib.orc_ref_run.argtypes = [C.POINTER(CfgRef), C.c_int] + [vp] * 12
        _lib.orc32_run.argtypes = [C.POINTER(Cfg32), C.c_int] + [vp] * 23
        _lib.orc_set_threads.argtypes = [C.c_int]
        _lib.orc_sigma.argtypes = [C.c_int, C.c_double, C.c_double]
        _lib.orc_sigma.restype = C.c_double
        _lib.orc_mcc_setup.argtypes = [C.POINTER(MccParams)]
        _lib.orc_energy_bin.argtypes = [C.POINTER(MccParams), C.c_double]
        _lib.orc_energy_bin.restype = C.c_int
        _lib.orc_mcc_apply.argtypes = [C.POINTER(MccParams), C.c_long, vp, vp, vp, vp, C.c_uint32,
                                       vp, vp]
        _lib.orc_ref_run_mcc.argtypes = [C.POINTER(CfgRef), C.POINTER(MccParams),
                                         C.POINTER(MccParams), C.c_int, C.c_double, C.c_double,
                                         C.c_int, vp, C.c_long, vp, vp, vp, C.c_long, vp, vp, vp,
                                         vp, vp, vp, vp, vp, vp, vp]
        _lib.orc32_steady.argtypes = [C.c_long] + [vp] * 7 + [C.c_long] + [vp] * 7 + \
            [C.c_int, C.c_uint32, C.c_uint32, C.c_uint32, C.c_float, C.c_float]
        _lib.orc32_steady.restype = C.c_long
        _lib.orc32_ref_cx_prob.argtypes = [C.POINTER(RefCxParams), C.c_float, C.c_float,
                                           C.c_float]
        _lib.orc32_ref_cx_prob.restype = C.c_float
        _lib.orc32_ref_cx.argtypes = [C.c_long] + [vp] * 7 + [C.c_int, C.POINTER(RefCxParams), vp,
                                                              vp, C.c_int, C.c_uint32]
        _lib.orc32_ref_cx.restype = C.c_long
        _lib.orc_neutral_table.argtypes = [C.c_uint64, C.c_double, vp]
        _lib.orc32_diag_particles.argtypes = [C.c_long] + [vp] * 5 + [C.c_int, C.c_double, C.c_int,
                                                                       C.c_int, C.c_int] + [vp] * 6
        _lib.orc_diag_vfinal.argtypes = [C.c_int] + [vp] * 6
        _lib.orc_dump_write.argtypes = [C.c_char_p, vp, C.c_long, C.c_int]
        _lib.orc_dump_read.argtypes = [C.c_char_p, vp, C.c_long]
        _lib.orc_dump_read.restype = C.c_long
        _lib.orc_inject.argtypes = [C.c_long, C.c_uint32, C.c_uint32, C.c_uint32, C.c_uint32,
                                    C.c_double, C.c_double, C.c_double] + [vp] * 6
        _lib.orc_reduce.argtypes = [C.POINTER(ReduceParams), C.c_uint32, C.c_long] + [vp] * 7
        _lib.orc_reduce.restype = C.c_long
        _lib.orc32_deposit_w.argtypes = [C.c_long, vp, vp, vp, vp, C.c_int, C.c_int, vp]
        _lib.orc_vaughan.argtypes = [C.c_double] * 5
        _lib.orc_vaughan.restype = C.c_double
        _lib.orc_ar_ion_yield.argtypes = [C.c_double]
        _lib.orc_ar_ion_yield.restype = C.c_double
        _lib.orc_wall.argtypes = [C.POINTER(WallParams), C.c_int, C.c_uint32, C.c_long] + \
            [vp] * 8 + [C.POINTER(WallOut), C.c_long, vp]
        _lib.orc_wall.restype = C.c_long
    return _lib


def P(a):
    return a.ctypes.data_as(vp) if a is not None else None


def philox2(ctr, key):
    """Philox2x32-10 of the oracle"""
    out = (C.c_uint32 * 2)()
    lib().orc_philox2x32_10((C.c_uint32 * 2)(*ctr), C.c_uint32(key), out)
    return tuple(out)


def philox(ctr, key):
    out = (C.c_uint32 * 4)()
    lib().orc_philox4x32_10((C.c_uint32 * 4)(*ctr), (C.c_uint32 * 2)(*key), out)
    return list(out)


def mglin(u, ncycle=2, length=1.0):
    a = np.ascontiguousarray(u, dtype=np.float64).copy()
    rc = lib().orc_mglin(P(a), a.shape[0], ncycle, length)
    if rc != 0:
        raise ValueError("n-1 must be a power of 2")
    return a


def mg_tol(f, length, tol, max_cycles=200):
    a = np.ascontiguousarray(f, dtype=np.float64).copy()
    rr = C.c_double()
    c = lib().orc_mg_tol(P(a), a.shape[0], length, tol, max_cycles, C.byref(rr))
    return a, c, rr.value


def relres(u, f, length):
    u = np.ascontiguousarray(u, dtype=np.float64)
    f = np.ascontiguousarray(f, dtype=np.float64)
    return lib().orc_mg_relres(P(u), P(f), u.shape[0], length)


def ref_mglin_available():
    return os.path.exists(REF_MGLIN_SO)


def ref_mglin(u, ncycle=2, length=1.0):
    """The reference's own Mglin (mglin.cpp), built into oracle/_ref."""
    r = C.CDLL(REF_MGLIN_SO)
    r.ref_mglin.argtypes = [vp, C.c_int, C.c_int, C.c_double]
    a = np.ascontiguousarray(u, dtype=np.float64).copy()
    if r.ref_mglin(P(a), a.shape[0], ncycle, length) != 0:
        raise ValueError("reference Mglin rejected n")
    return a


def cell_pack(jx, jy):
    jx = np.asarray(jx).astype(np.int16).view(np.uint16).astype(np.uint32)
    jy = np.asarray(jy).astype(np.int16).view(np.uint16).astype(np.uint32)
    return jx | (jy << 16)


def cell_unpack(cell):
    c = np.asarray(cell, dtype=np.uint32)
    return ((c & 0xFFFF).astype(np.uint16).view(np.int16).astype(np.int32),
            (c >> 16).astype(np.uint16).view(np.int16).astype(np.int32))


def push32(state, E32, pp):
    """orc32_push on copies of the SoA arrays; returns (new_state, status)."""
    s = {k: np.array(v, copy=True) for k, v in state.items()}
    n = len(s["cell"])
    st = np.zeros(n, dtype=np.uint8)
    lib().orc32_push(n, P(s["cell"]), P(s["ox"]), P(s["oy"]), P(s["vx"]), P(s["vy"]), P(s["vz"]),
                     P(np.ascontiguousarray(E32, dtype=np.float32)), C.byref(pp), P(st))
    return s, st


def deposit32(state, N, mode):
    q = np.zeros(N * N, dtype=np.int64)
    n = len(state["cell"])
    lib().orc32_deposit(n, P(state["cell"]), P(state["ox"]), P(state["oy"]), N, mode, P(q))
    return q.reshape(N, N)


def efield_pack(phi, dx):
    N = phi.shape[0]
    ex = np.zeros(N * N)
    ey = np.zeros(N * N)
    e32 = np.zeros(2 * ((N + 1) * N + 1), dtype=np.float32)
    lib().orc_efield_pack(P(np.ascontiguousarray(phi)), N, dx, P(ex), P(ey), P(e32))
    return ex.reshape(N, N), ey.reshape(N, N), e32


def ref_efield(phi, dx, ex=None, ey=None):
    N = phi.shape[0]
    ex = np.zeros((N + 1) * (N + 1)) if ex is None else ex
    ey = np.zeros((N + 1) * (N + 1)) if ey is None else ey
    lib().orc_ref_grid_efield(P(np.ascontiguousarray(phi)), N, dx, P(ex), P(ey))
    return ex, ey


def run32(cfg, nsteps, e, i, E32, mg_cycles=False):
    """orc32_run on copies; returns dict with the new states and fields."""
    N = cfg.N
    e = {k: np.array(v, copy=True) for k, v in e.items()}
    i = {k: np.array(v, copy=True) for k, v in i.items()}
    ne = C.c_long(len(e["cell"]))
    ni = C.c_long(len(i["cell"]))
    E = np.array(E32, dtype=np.float32, copy=True)
    qe = np.zeros(N * N, dtype=np.int64)
    qi = np.zeros(N * N, dtype=np.int64)
    rho = np.zeros(N * N)
    phi = np.zeros(N * N)
    lost = np.zeros(8, dtype=np.int64)
    cyc = np.zeros(max(nsteps, 1), dtype=np.int32)
    keys = ("cell", "ox", "oy", "vx", "vy", "vz", "id")
    rc = lib().orc32_run(C.byref(cfg), nsteps, C.byref(ne), *[P(e[k]) for k in keys], C.byref(ni),
                         *[P(i[k]) for k in keys], P(E), P(qe), P(qi), P(rho), P(phi), P(lost),
                         P(cyc))
    if rc != 0:
        raise RuntimeError("orc32_run failed")
    e = {k: v[: ne.value] for k, v in e.items()}
    i = {k: v[: ni.value] for k, v in i.items()}
    return dict(e=e, i=i, E32=E, qe=qe.reshape(N, N), qi=qi.reshape(N, N), rho=rho.reshape(N, N),
                phi=phi.reshape(N, N), lost=lost, mg_cycles=cyc)


def run_ref(cfg, nsteps, e, i, eid, iid, Ex, Ey):
    """orc_ref_run (FP64 AoS reference semantics) on copies."""
    N = cfg.N
    e = np.array(e, dtype=PARTICLE_DTYPE, copy=True)
    i = np.array(i, dtype=PARTICLE_DTYPE, copy=True)
    eid = np.array(eid, dtype=np.uint32, copy=True)
    iid = np.array(iid, dtype=np.uint32, copy=True)
    ne = C.c_long(len(e))
    ni = C.c_long(len(i))
    Ex = np.array(Ex, dtype=np.float64, copy=True)
    Ey = np.array(Ey, dtype=np.float64, copy=True)
    qe = np.zeros(N * N)
    qi = np.zeros(N * N)
    rho = np.zeros(N * N)
    phi = np.zeros(N * N)
    rc = lib().orc_ref_run(C.byref(cfg), nsteps, C.byref(ne), P(e), P(eid), C.byref(ni), P(i),
                           P(iid), P(Ex), P(Ey), P(qe), P(qi), P(rho), P(phi))
    if rc != 0:
        raise RuntimeError("orc_ref_run failed")
    return dict(e=e[: ne.value], i=i[: ni.value], eid=eid[: ne.value], iid=iid[: ni.value], Ex=Ex,
                Ey=Ey, qe=qe.reshape(N, N), qi=qi.reshape(N, N), rho=rho.reshape(N, N),
                phi=phi.reshape(N, N))


def configs_from_params(p, mode=None):
    """(Cfg32, CfgRef) matching a pic2d.Params the way the library derives them."""
    import math
    N = p.nodes
    dx = p.length / (N - 1)
    if mode is None:
        mode = p.mode
        if p.mode == MODE_REF and getattr(p, "wall", 0) == 1:   # reference steady-state walls
            mode = MODE_REF_STEADY

    def pp(q_over_m, dt):
        coef = 0.5 * dt * q_over_m
        w = dt * q_over_m * p.magnetic_field
        return PushParams(N=N, mode=mode, cE=coef, dtdx=dt / dx, dtdy=dt / dx,
                          bfield=int(p.magnetic_field > 0), cos_a=math.cos(p.angle),
                          sin_a=math.sin(p.angle), cos_w=math.cos(w), sin_w=math.sin(w))

    qme = p.electron_charge / p.electron_mass
    qmi = p.ion_charge / p.ion_mass
    c32 = Cfg32(N=N, mode=mode, corner8=p.corner8, mg_mode=0, mg_ncycle=p.mg_ncycle,
                mg_max_cycles=p.mg_max_iter, dt_dti=p.dt_dti, preloop=0, L=p.length, dx=dx,
                eps=p.epsilon, mg_tol=p.mg_tol, phi_left=p.phi_left, phi_right=p.phi_right,
                phi_bottom=p.phi_bottom, phi_top=p.phi_top, phi_period=p.phi_period,
                scale_e=math.ldexp(p.qce / (dx * dx) * 1.0, -FX_TOTAL_BITS),
                scale_i=math.ldexp(p.qci / (dx * dx) * 1.0, -FX_TOTAL_BITS), start_step=0,
                pe=pp(qme, p.dt), pi=pp(qmi, p.dti), key0=getattr(p, "seed", 0) & 0xFFFFFFFF,
                key1=(getattr(p, "seed", 0) >> 32) & 0xFFFFFFFF, steady_se=getattr(p, "vte", 0.0),
                steady_si=getattr(p, "vti", 0.0))
    we = p.dt * qme * p.magnetic_field
    wi = p.dti * qmi * p.magnetic_field
    cref = CfgRef(N=N, ncycle=p.mg_ncycle, dt_dti=p.dt_dti, L=p.length, dx=dx, dy=dx,
                  eps=p.epsilon, coef_e=0.5 * p.dt * qme, coef_i=0.5 * p.dti * qmi, dt_e=p.dt,
                  dt_i=p.dti, qc_dxdy_e=p.qce / (dx * dx), qc_dxdy_i=p.qci / (dx * dx),
                  cos_a=math.cos(p.angle), sin_a=math.sin(p.angle), cos_we=math.cos(we),
                  sin_we=math.sin(we), cos_wi=math.cos(wi), sin_wi=math.sin(wi),
                  bfield=p.magnetic_field, phi_left=p.phi_left, phi_right=p.phi_right,
                  phi_bottom=p.phi_bottom, phi_top=p.phi_top, phi_period=p.phi_period,
                  start_step=0)
    return c32, cref


def state_to_aos(state, dx, weight=1.0):
    jx, jy = cell_unpack(state["cell"])
    a = np.zeros(len(jx), dtype=PARTICLE_DTYPE)
    a["x"] = (jx + state["ox"].astype(np.float64)) * dx
    a["y"] = (jy + state["oy"].astype(np.float64)) * dx
    a["vx"] = state["vx"]
    a["vy"] = state["vy"]
    a["vz"] = state["vz"]
    a["weight"] = weight
    return a


def random_state(n, N, sigma, seed, rng=None):
    """uniform positions in the domain and a Maxwellian, fp32 SoA (for CPU tests)"""
    rng = np.random.default_rng(seed) if rng is None else rng
    t = rng.random((2, n)) * (N - 1)
    j = np.floor(t)
    o = (t - j).astype(np.float32)
    o[o >= 1.0] = 0.0
    v = (rng.standard_normal((3, n)) * sigma).astype(np.float32)
    return dict(cell=cell_pack(j[0].astype(np.int32), j[1].astype(np.int32)), ox=o[0].copy(),
                oy=o[1].copy(), vx=v[0].copy(), vy=v[1].copy(), vz=v[2].copy(),
                id=np.arange(n, dtype=np.uint32))


SIG = dict(AR_EL_REF=0, AR_EL_PHYS=1, AR_IZ_REF=2, AR_IZ_PHYS=3, AR_EXC=4, H_EL=5, H_IZ=6,
           AR_CX=7, AR_ISO=8, H_CX=9)


def sigma(kind, E, thr=0.0):
    """cross section [1e-20 m^2] of the oracle (restated reference fits)"""
    f = lib().orc_sigma
    return np.array([f(SIG[kind], float(e), thr) for e in np.atleast_1d(E)])


def e_nu_parts(E, g, ng, gas="AR", e_exc=11.55, e_iz=15.76):
    """electron collision frequencies (elastic, excitation, ionisation) [1/s]"""
    f = ng * 1e-20 * g
    if gas == "AR":
        return (f * sigma("AR_EL_PHYS", E), f * sigma("AR_EXC", E, e_exc),
                f * sigma("AR_IZ_PHYS", E, e_iz))
    return f * np.maximum(sigma("H_EL", E), 0), 0 * E, f * 1000 * sigma("H_IZ", E)


def i_nu_parts(E, g, ng):
    f = ng * 1e-20 * g
    return f * sigma("AR_ISO", E), f * sigma("AR_CX", E)


def nu_max(mass, ng, electron, q=1.602177e-19):
    """null-collision bound; the energy scan ends at 3e4 eV (electrons) / 1e3 eV (ions)"""
    E = 10.0 ** np.linspace(-6, 4.5 if electron else 3.0, 20001)
    g = np.sqrt(2 * q * E / mass)
    parts = e_nu_parts(E, g, ng) if electron else i_nu_parts(E, g, ng)
    return float(np.max(sum(parts))) * 1.02


Q_E = 1.602177e-19
AMU = 1.66054e-27


def mcc_params(p, electron):
    """oracle MCC parameters for one species of a pic2d.Params (physical mode)"""
    argon = p.gas == 0
    c = MccParams(gas=0 if argon else 1, electron=int(electron), ng=p.neutral_density,
                  dt=p.dt if electron else p.dti,
                  mass=p.electron_mass if electron else p.ion_mass, M=p.ion_mass, q=Q_E,
                  vtn=p.vtn, e_exc=11.55, e_iz=15.76 if argon else 13.6,
                  opal_B=10.0 if argon else 8.3, kev_amu_div=p.ion_mass / AMU,
                  key0=p.seed & 0xFFFFFFFF, key1=(p.seed >> 32) & 0xFFFFFFFF)
    lib().orc_mcc_setup(C.byref(c))
    if getattr(p, "null_bound", 0) == 1:  # PIC_NULL_BOUND_GLOBAL
        c.binned = 0
    return c


def e_bound(c, E):
    """(p, nu) of the null-collision bound an electron of energy E [eV] gets:
    its energy bin's (binned parameters) or the global one"""
    if not c.binned:
        return c.p_null, c.nu_max
    b = lib().orc_energy_bin(C.byref(c), 2.0 * Q_E * E / c.mass)
    return c.bin_p[b], c.bin_nu[b]


def mcc_apply(c, vx, vy, vz, ids, step):
    """FP64 collisions of arrays of particles; returns (vx, vy, vz, event, child6)"""
    vx = np.array(vx, dtype=np.float64, copy=True)
    vy = np.array(vy, dtype=np.float64, copy=True)
    vz = np.array(vz, dtype=np.float64, copy=True)
    ids = np.ascontiguousarray(ids, dtype=np.uint32)
    n = len(vx)
    ev = np.zeros(n, dtype=np.int32)
    ch = np.zeros((n, 6))
    lib().orc_mcc_apply(C.byref(c), n, P(vx), P(vy), P(vz), P(ids), step, P(ev), P(ch))
    return vx, vy, vz, ev, ch


def run_ref_mcc(cfg, me, mi, plop, vte, vti, nsteps, e, i, eid, iid, next_id, Ex, Ey,
                grow=2.0):
    """orc_ref_run_mcc on copies; capacity grow x the initial counts"""
    N = cfg.N
    cap_e = int(len(e) * grow) + 1024
    cap_i = int(len(i) * grow) + 1024
    ea = np.zeros(cap_e, dtype=PARTICLE_DTYPE)
    ea[: len(e)] = e
    ia = np.zeros(cap_i, dtype=PARTICLE_DTYPE)
    ia[: len(i)] = i
    ea_id = np.zeros(cap_e, dtype=np.uint32)
    ea_id[: len(e)] = eid
    ia_id = np.zeros(cap_i, dtype=np.uint32)
    ia_id[: len(i)] = iid
    ne = C.c_long(len(e))
    ni = C.c_long(len(i))
    nid = np.array(next_id, dtype=np.uint32)
    Ex = np.array(Ex, dtype=np.float64, copy=True)
    Ey = np.array(Ey, dtype=np.float64, copy=True)
    qe = np.zeros(N * N)
    qi = np.zeros(N * N)
    rho = np.zeros(N * N)
    phi = np.zeros(N * N)
    ev = np.zeros(8, dtype=np.int64)
    rc = lib().orc_ref_run_mcc(C.byref(cfg), C.byref(me), C.byref(mi), int(plop), vte, vti,
                               nsteps, C.byref(ne), cap_e, P(ea), P(ea_id), C.byref(ni), cap_i,
                               P(ia), P(ia_id), P(nid), P(Ex), P(Ey), P(qe), P(qi), P(rho),
                               P(phi), P(ev))
    if rc != 0:
        raise RuntimeError("orc_ref_run_mcc failed")
    return dict(e=ea[: ne.value], i=ia[: ni.value], eid=ea_id[: ne.value], iid=ia_id[: ni.value],
                next_id=nid, Ex=Ex, Ey=Ey, qe=qe.reshape(N, N), qi=qi.reshape(N, N),
                rho=rho.reshape(N, N), phi=phi.reshape(N, N), events=ev)


def refcx_params(p):
    """oracle parameters of the reference ion charge exchange for a pic2d.Params"""
    return RefCxParams(gas=int(p.gas), sig_energy=int(p.sigma_cx == 1), sig_const=p.sigma_value,
                       q=Q_E, mass=p.ion_mass, dt=p.dti, ng=p.neutral_density,
                       key0=p.seed & 0xFFFFFFFF, key1=(p.seed >> 32) & 0xFFFFFFFF)


def neutral_table(seed, vtn):
    t = np.zeros(30000, dtype=np.float32)
    lib().orc_neutral_table(seed, vtn, P(t))
    return t


def ref_cx(state, N, c, pcol, table, step):
    """orc32_ref_cx on a copy of a species state; returns (state, events)"""
    st = {k: np.array(v, copy=True) for k, v in state.items()}
    ev = lib().orc32_ref_cx(len(st["id"]), P(st["cell"]), P(st["ox"]), P(st["oy"]), P(st["vx"]),
                            P(st["vy"]), P(st["vz"]), P(st["id"]), N, C.byref(c), P(pcol),
                            P(table), 10000, step)
    return st, ev


class Diag:
    """oracle accumulators of the batch diagnostics of one species"""

    def __init__(self, N, no_bins):
        self.N, self.nb = N, no_bins
        self.vsx = np.zeros(N, np.int64)
        self.vsy = np.zeros(N, np.int64)
        self.cx = np.zeros(N, np.uint32)
        self.cy = np.zeros(N, np.uint32)
        self.vpx = np.zeros(N)
        self.vpy = np.zeros(N)
        self.dist = np.zeros(4 * no_bins, np.uint64)

    def step(self, st, vt, divide_bins_by, spread, loc):
        loc = np.asarray(loc, dtype=np.int32)
        lib().orc32_diag_particles(len(st["cell"]), P(st["cell"]), P(st["ox"]), P(st["oy"]),
                                   P(st["vx"]), P(st["vy"]), self.N, vt, self.nb, divide_bins_by,
                                   spread, P(loc), P(self.vsx), P(self.vsy), P(self.cx),
                                   P(self.cy), P(self.dist))
        lib().orc_diag_vfinal(self.N, P(self.vsx), P(self.vsy), P(self.cx), P(self.cy),
                              P(self.vpx), P(self.vpy))


# ---------------------------------------------------------- particle-count control
def inject(n, species, step, seed, dx, L, vt):
    """orc_inject: the source's particles (unmoved), fp32 SoA"""
    d = {k: np.zeros(n, dtype=np.uint32 if k == "cell" else np.float32)
         for k in ("cell", "ox", "oy", "vx", "vy", "vz")}
    lib().orc_inject(n, species, step, seed & 0xFFFFFFFF, (seed >> 32) & 0xFFFFFFFF, dx, L, vt,
                     *[P(d[k]) for k in ("cell", "ox", "oy", "vx", "vy", "vz")])
    return d


def reduce_params(p, species, seed=None):
    """ReduceParams of species 0 (electrons) / 1 (ions) from a pic2d.Params"""
    seed = p.seed if seed is None else seed
    c = ReduceParams(ng=p.no_groups, size_check=p.size_to_check[species], species=species,
                     key0=seed & 0xFFFFFFFF, key1=(seed >> 32) & 0xFFFFFFFF)
    for g in range(8):
        c.vlim[g] = p.v_limit[species][g]
        c.glimit[g] = p.group_limit[species][g]
        c.gnew[g] = p.new_group_size[species][g]
    return c


def reduce(c, step, vx, vy, ids, w):
    """orc_reduce on copies: (keep mask, new weights, Ngroup[8], Nremove[8], removed)"""
    n = len(ids)
    w = np.array(w, dtype=np.float32, copy=True)
    keep = np.zeros(n, dtype=np.uint8)
    ng = np.zeros(8, dtype=np.uint32)
    nr = np.zeros(8, dtype=np.uint32)
    vx = np.ascontiguousarray(vx, dtype=np.float32)
    vy = np.ascontiguousarray(vy, dtype=np.float32)
    ids = np.ascontiguousarray(ids, dtype=np.uint32)
    r = lib().orc_reduce(C.byref(c), step, n, P(vx), P(vy), P(ids), P(w), P(keep), P(ng), P(nr))
    return keep.astype(bool), w, ng, nr, r


def deposit32_w(state, w, N, mode):
    q = np.zeros(N * N, dtype=np.int64)
    n = len(state["cell"])
    w = np.ascontiguousarray(w, dtype=np.float32)
    lib().orc32_deposit_w(n, P(state["cell"]), P(state["ox"]), P(state["oy"]), P(w), N, mode,
                          P(q))
    return q.reshape(N, N)


# ---------------------------------------------------------- surface physics
CHARGE = 1.602177e-19


def wall_params(p, seed=None):
    """WallParams as the library derives them from a pic2d.Params"""
    seed = p.seed if seed is None else seed
    dx = p.length / (p.nodes - 1)
    phys = p.mode == MODE_PHYS
    return WallParams(mode=MODE_PHYS if phys else MODE_REF, N=p.nodes, dx=dx,
                      Lw=p.length if phys else dx * p.nodes, dt=p.dt, me=p.electron_mass,
                      mi=p.ion_mass, q=CHARGE, max_yield=p.see_max_yield, E0=p.see_E0,
                      Emax0=p.see_Emax0, thr=p.threshold, wf=p.work_function,
                      see_e=int(bool(p.e_impact_see)), refl=int(bool(p.reflection)),
                      back=int(bool(p.backscatter)),
                      see_i=int(bool(p.ion_impact_see) and p.gas == 0),
                      key0=seed & 0xFFFFFFFF, key1=(seed >> 32) & 0xFFFFFFFF)


def wall(c, species, step, state, w=None, cap=None):
    """orc_wall on the particles of `state` (all assumed to have reached a
    wall); returns (outputs dict with 'key', counts[6])"""
    n = len(state["cell"])
    cap = 8 * n + 16 if cap is None else cap
    out = {k: np.zeros(cap, dtype=np.uint32 if k == "cell" else np.float32)
           for k in ("cell", "ox", "oy", "vx", "vy", "vz", "w")}
    out["key"] = np.zeros(cap, dtype=np.uint64)
    wo = WallOut(**{k: out[k].ctypes.data for k in out})
    counts = np.zeros(6, dtype=np.int64)
    w = np.ones(n, dtype=np.float32) if w is None else np.ascontiguousarray(w, dtype=np.float32)
    arr = [np.ascontiguousarray(state[k], dtype=np.uint32 if k in ("cell", "id") else np.float32)
           for k in ("cell", "ox", "oy", "vx", "vy", "vz")]
    ids = np.ascontiguousarray(state["id"], dtype=np.uint32)
    m = lib().orc_wall(C.byref(c), species, step, n, *[P(a) for a in arr], P(w), P(ids),
                       C.byref(wo), cap, P(counts))
    assert m <= cap
    return {k: v[:m] for k, v in out.items()}, counts
