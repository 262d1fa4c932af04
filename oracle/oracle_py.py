"""ctypes binding of the CPU oracle (oracle/liboracle.so).

TEST INFRASTRUCTURE ONLY: used by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py as the checker. Arrays use the reference layout
(nx, ny, nz) C order.
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")
FIELDS = ("u", "v", "w", "rho", "p", "theta", "qv", "qc", "qr")
_dp = ctypes.POINTER(ctypes.c_double)
_fp = ctypes.POINTER(ctypes.c_float)


class OcGrid(ctypes.Structure):
    _fields_ = [("nx", ctypes.c_int), ("ny", ctypes.c_int), ("nz", ctypes.c_int),
                ("dx", ctypes.c_double), ("dy", ctypes.c_double), ("dz", ctypes.c_double)]


class OcConfig(ctypes.Structure):
    _fields_ = [("dt", ctypes.c_double), ("substeps", ctypes.c_int), ("acoustic_mode", ctypes.c_int),
                ("hevi_beta", ctypes.c_double), ("coriolis_f", ctypes.c_double),
                ("bc_west", ctypes.c_int), ("bc_east", ctypes.c_int), ("bc_south", ctypes.c_int),
                ("bc_north", ctypes.c_int), ("bc_bottom", ctypes.c_int), ("bc_top", ctypes.c_int),
                ("qc0", ctypes.c_double), ("c_auto", ctypes.c_double), ("c_accr", ctypes.c_double),
                ("evap_eff", ctypes.c_double), ("vt_max", ctypes.c_double),
                ("sponge", ctypes.c_int), ("sponge_z_d", ctypes.c_double), ("sponge_alpha", ctypes.c_double),
                ("sponge_w_cells", ctypes.c_int), ("sponge_w_alpha", ctypes.c_double),
                ("sponge_lateral_cells", ctypes.c_int), ("sponge_lateral_alpha", ctypes.c_double),
                ("sgs", ctypes.c_int), ("sgs_mode", ctypes.c_int), ("sgs_filter", ctypes.c_int),
                ("sgs_cs", ctypes.c_double), ("sgs_pr_t", ctypes.c_double), ("sgs_sc_t", ctypes.c_double),
                ("sgs_nu_max", ctypes.c_double), ("sgs_k_max", ctypes.c_double), ("sgs_dt", ctypes.c_double),
                ("diffusion", ctypes.c_int), ("diffusion_kh", ctypes.c_double), ("diffusion_kv", ctypes.c_double)]


class OcBase(ctypes.Structure):
    _fields_ = [(n, _dp) for n in ("rho0", "p0", "theta0", "qv0", "u0", "v0")]


class OcState(ctypes.Structure):
    _fields_ = [(n, ctypes.c_void_p) for n in FIELDS] + [("rain_mm", _dp)]


_LIB = None


def lib():
    global _LIB
    if _LIB is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} missing: run `make -C oracle`")
        L = ctypes.CDLL(LIB_PATH)
        P = ctypes.POINTER
        for name in ("oracle_large_steps_f32", "oracle_large_steps_f64"):
            getattr(L, name).argtypes = [P(OcGrid), P(OcBase), P(OcConfig), P(OcState), ctypes.c_int]
        L.oracle_resolve_substeps.argtypes = [P(OcGrid), P(OcBase), P(OcConfig)]
        L.oracle_apply_sponges_f64.argtypes = [P(OcGrid), P(OcBase), P(OcConfig), P(OcState), ctypes.c_double]
        for name in ("oracle_sgs_terms_f64", "oracle_sgs_terms_f32"):
            getattr(L, name).argtypes = [P(OcGrid), P(OcConfig), P(OcState)] + [_dp] * 8
        for name in ("oracle_apply_sgs_f64", "oracle_apply_sgs_f32"):
            getattr(L, name).argtypes = [P(OcGrid), P(OcConfig), P(OcState)]
        for name in ("oracle_acoustic_substep_f32", "oracle_acoustic_substep_f64"):
            getattr(L, name).argtypes = [P(OcGrid), P(OcBase), P(OcState), ctypes.c_double]
        for name in ("oracle_hevi_substep_f64", "oracle_hevi_substep_f32"):
            getattr(L, name).argtypes = [P(OcGrid), P(OcBase), P(OcState), ctypes.c_double, ctypes.c_double, _dp]
        L.oracle_acoustic_substeps_batched_f32.argtypes = [P(OcGrid), P(OcBase), P(OcConfig), P(OcState),
                                                           ctypes.c_int, ctypes.c_double, ctypes.c_double,
                                                           _fp, _fp, _fp]
        L.oracle_acoustic_substeps_batched_f64.argtypes = [P(OcGrid), P(OcBase), P(OcConfig), P(OcState),
                                                           ctypes.c_int, ctypes.c_double, ctypes.c_double,
                                                           _dp, _dp, _dp]
        L.oracle_apply_boundaries_f32.argtypes = [P(OcGrid), P(OcConfig), P(OcState), ctypes.c_double]
        L.oracle_slow_tendencies_f32.argtypes = [P(OcGrid), P(OcBase), P(OcConfig), P(OcState), _fp, _fp, _fp]
        L.oracle_slow_tendencies_f64.argtypes = [P(OcGrid), P(OcBase), P(OcConfig), P(OcState), _dp, _dp, _dp]
        L.oracle_apply_boundaries_f64.argtypes = [P(OcGrid), P(OcConfig), P(OcState), ctypes.c_double]
        L.oracle_slow_rk2_f32.argtypes = [P(OcGrid), P(OcBase), P(OcConfig), P(OcState)]
        L.oracle_slow_rk2_f64.argtypes = [P(OcGrid), P(OcBase), P(OcConfig), P(OcState)]
        L.oracle_transport_f32.argtypes = [P(OcGrid), P(OcConfig), _fp, _fp, _fp, _fp, _fp, _fp, _fp,
                                           ctypes.c_double, ctypes.c_int]
        L.oracle_transport_f64.argtypes = [P(OcGrid), P(OcConfig), _dp, _dp, _dp, _dp, _dp, _dp, _dp,
                                           ctypes.c_double, ctypes.c_int]
        L.oracle_kessler_tendencies_f32.argtypes = [P(OcGrid), P(OcConfig)] + [_fp] * 6 + [ctypes.c_double] + [_fp] * 4
        L.oracle_microphysics_step_f32.argtypes = [P(OcGrid), P(OcConfig), P(OcState), ctypes.c_double, _dp,
                                                   ctypes.c_int]
        L.oracle_microphysics_step_f64.argtypes = [P(OcGrid), P(OcConfig), P(OcState), ctypes.c_double, _dp,
                                                   ctypes.c_int]
        for name, n in (("oracle_weno5_face", 5), ("oracle_mc_slope", 3), ("oracle_upwind_fd5", 9),
                        ("oracle_upwind_fd3", 7), ("oracle_qvs_water", 2), ("oracle_theta_to_temperature", 2),
                        ("oracle_eos_pressure", 2), ("oracle_kessler_rain_speed", 2)):
            fn = getattr(L, name)
            fn.argtypes = [ctypes.c_double] * n
            fn.restype = ctypes.c_double
        L.oracle_upwind_point.argtypes = [_fp, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                          ctypes.c_int, ctypes.c_int, ctypes.c_double, ctypes.c_double]
        L.oracle_upwind_point.restype = ctypes.c_double
        L.oracle_saturation_adjust.argtypes = [ctypes.c_double, ctypes.c_double, _dp, _dp]
        L.oracle_saturation_adjust.restype = ctypes.c_double
        L.oracle_thomas.argtypes = [ctypes.c_int, _dp, _dp, _dp, _dp, _dp]
        L.oracle_hevi_column_system.argtypes = [ctypes.c_int, ctypes.c_double, ctypes.c_double, ctypes.c_double,
                                                _dp, _dp, _dp, _dp, _dp, _dp, _dp, _dp, _dp]
        _i = ctypes.c_int
        _d = ctypes.c_double
        L.oracle_diffusion_tendency_f32.argtypes = [_i, _i, _i, _d, _d, _dp, _fp, _fp, _fp, _d, _d, _i, _i, _i, _i,
                                                    _fp]
        L.oracle_diffusion_tendency_f64.argtypes = [_i, _i, _i, _d, _d, _dp, _dp, _dp, _dp, _d, _d, _i, _i, _i, _i,
                                                    _dp]
        L.oracle_apply_diffusion_f64.argtypes = [P(OcGrid), P(OcConfig), P(OcState)]
        L.oracle_set_threads.argtypes = [ctypes.c_int]
        L.oracle_max_threads.restype = ctypes.c_int
        _LIB = L
    return _LIB


def set_threads(n):
    """OpenMP threads of the oracle's loops from now on; returns the previous count."""
    prev = lib().oracle_max_threads()
    lib().oracle_set_threads(int(n))
    return prev


def grid(nx, ny, nz, dx, dy, dz):
    return OcGrid(int(nx), int(ny), int(nz), float(dx), float(dy), float(dz))


def config(dt, substeps, acoustic_mode=0, hevi_beta=0.1, coriolis_f=0.0, lateral=1, bottom=0, top=0,
           qc0=1.0e-3, c_auto=1.0e-3, c_accr=2.2, evap_eff=1.0, vt_max=20.0, sponge=None, sgs=None,
           diffusion=None):
    lat = lateral if isinstance(lateral, (tuple, list)) else (lateral,) * 4
    sp = sponge or {}
    sg = sgs or {}
    return OcConfig(float(dt), int(substeps), int(acoustic_mode), float(hevi_beta), float(coriolis_f),
                    lat[0], lat[1], lat[2], lat[3], int(bottom), int(top), qc0, c_auto, c_accr, evap_eff, vt_max,
                    int(bool(sponge)), float(sp.get("z_d", 12000.0)), float(sp.get("alpha", 0.05)),
                    int(sp.get("w_cells", 8)), float(sp.get("w_alpha", 0.2)), int(sp.get("lateral_cells", 0)),
                    float(sp.get("lateral_alpha", 0.05)),
                    int(bool(sgs)), int(sg.get("mode", 0)), int(sg.get("filter", 0)), float(sg.get("cs", 0.18)),
                    float(sg.get("pr_t", 1.0 / 3.0)), float(sg.get("sc_t", 1.0 / 3.0)), float(sg.get("nu_max", 1000.0)),
                    float(sg.get("k_max", 1000.0)), float(sg.get("dt", 1.0)),
                    *diffusion_fields(diffusion))


def diffusion_fields(diffusion):
    """(apply code, K_h, K_v) of an explicit-diffusion setting {"apply": "scalars" | "momentum" | "all",
    "kh": .., "kv": ..}, or (0, 0, 0) for None."""
    if not diffusion:
        return 0, 0.0, 0.0
    code = {"scalars": 1, "momentum": 2, "all": 3}[diffusion.get("apply", "all")]
    return code, float(diffusion.get("kh", 0.0)), float(diffusion.get("kv", 0.0))


def from_step_config(c):
    """Oracle config with the same values as a nimbus_amd.lib.StepConfig."""
    return OcConfig(*[getattr(c, n) for n, _ in OcConfig._fields_])


def base_struct(b):
    return OcBase(*[np.ascontiguousarray(b[n], dtype=np.float64).ctypes.data_as(_dp)
                    for n in ("rho0", "p0", "theta0", "qv0", "u0", "v0")])


def _state_struct(st):
    return OcState(*[st[n].ctypes.data_as(ctypes.c_void_p) for n in FIELDS],
                   st["rain_mm"].ctypes.data_as(_dp))


def copy_state(st, dtype):
    out = {n: np.ascontiguousarray(st[n], dtype=dtype).copy() for n in FIELDS}
    rain = st.get("rain_mm")
    out["rain_mm"] = (np.zeros(st["u"].shape[:2]) if rain is None else np.array(rain, dtype=np.float64))
    return out


def large_steps(g, base, cfg, st, n_steps, dtype=np.float64):
    """Runs n_steps large steps on a copy of st; returns (state, substeps)."""
    s = copy_state(st, dtype)
    b = {k: np.ascontiguousarray(v, dtype=np.float64) for k, v in base.items()}
    fn = lib().oracle_large_steps_f64 if np.dtype(dtype) == np.float64 else lib().oracle_large_steps_f32
    n = fn(ctypes.byref(g), ctypes.byref(base_struct(b)), ctypes.byref(cfg), ctypes.byref(_state_struct(s)),
           int(n_steps))
    return s, n


def acoustic_substep(g, base, st, dts, dtype=np.float32):
    s = copy_state(st, dtype)
    b = {k: np.ascontiguousarray(v, dtype=np.float64) for k, v in base.items()}
    fn = lib().oracle_acoustic_substep_f32 if np.dtype(dtype) == np.float32 else lib().oracle_acoustic_substep_f64
    fn(ctypes.byref(g), ctypes.byref(base_struct(b)), ctypes.byref(_state_struct(s)), float(dts))
    return s


def hevi_substep(g, base, st, dts, beta, dtype=np.float64):
    s = copy_state(st, dtype)
    b = {k: np.ascontiguousarray(v, dtype=np.float64) for k, v in base.items()}
    wz = np.zeros(st["u"].shape, dtype=np.float64)
    fn = lib().oracle_hevi_substep_f64 if np.dtype(dtype) == np.float64 else lib().oracle_hevi_substep_f32
    fn(ctypes.byref(g), ctypes.byref(base_struct(b)), ctypes.byref(_state_struct(s)), float(dts), float(beta),
       wz.ctypes.data_as(_dp))
    return s, wz


def acoustic_substeps_batched(g, base, cfg, st, n, dts, weight, mf=None, dtype=np.float32):
    """Reference batched substep contract: n substeps with per-substep boundaries."""
    s = copy_state(st, dtype)
    b = {k: np.ascontiguousarray(v, dtype=np.float64) for k, v in base.items()}
    shape = st["u"].shape
    mf = [np.zeros(shape, dtype) for _ in range(3)] if mf is None else [np.array(x, dtype) for x in mf]
    if np.dtype(dtype) == np.float32:
        fn, ptr = lib().oracle_acoustic_substeps_batched_f32, _fp
    else:
        fn, ptr = lib().oracle_acoustic_substeps_batched_f64, _dp
    fn(ctypes.byref(g), ctypes.byref(base_struct(b)), ctypes.byref(cfg), ctypes.byref(_state_struct(s)), int(n),
       float(dts), float(weight), *[x.ctypes.data_as(ptr) for x in mf])
    return s, mf


def apply_boundaries(g, cfg, st, dt_rad, dtype=np.float32):
    """One boundary closure on a copy of st, the radiative stash seeded from
    the input faces; dt_rad > 0 radiates the KW78 faces over dt_rad."""
    s = copy_state(st, dtype)
    fn = lib().oracle_apply_boundaries_f32 if np.dtype(dtype) == np.float32 else lib().oracle_apply_boundaries_f64
    fn(ctypes.byref(g), ctypes.byref(cfg), ctypes.byref(_state_struct(s)), float(dt_rad))
    return s


def slow_tendencies(g, base, cfg, st, dtype=np.float32):
    s = copy_state(st, dtype)
    b = {k: np.ascontiguousarray(v, dtype=np.float64) for k, v in base.items()}
    du, dv, dw = (np.zeros(st["u"].shape, dtype=dtype) for _ in range(3))
    if np.dtype(dtype) == np.float32:
        fn, ptr = lib().oracle_slow_tendencies_f32, _fp
    else:
        fn, ptr = lib().oracle_slow_tendencies_f64, _dp
    fn(ctypes.byref(g), ctypes.byref(base_struct(b)), ctypes.byref(cfg), ctypes.byref(_state_struct(s)),
       du.ctypes.data_as(ptr), dv.ctypes.data_as(ptr), dw.ctypes.data_as(ptr))
    return du, dv, dw


def slow_rk2(g, base, cfg, st, dtype=np.float32):
    s = copy_state(st, dtype)
    b = {k: np.ascontiguousarray(v, dtype=np.float64) for k, v in base.items()}
    fn = lib().oracle_slow_rk2_f32 if np.dtype(dtype) == np.float32 else lib().oracle_slow_rk2_f64
    fn(ctypes.byref(g), ctypes.byref(base_struct(b)), ctypes.byref(cfg), ctypes.byref(_state_struct(s)))
    return s


def transport(g, cfg, q, mfx, mfy, mfz, rho_begin, rho_end, dt, positive_definite, dtype=np.float32):
    a = [np.ascontiguousarray(x, dtype=dtype) for x in (q, mfx, mfy, mfz, rho_begin, rho_end)]
    out = np.zeros_like(a[0])
    if np.dtype(dtype) == np.float32:
        p = [x.ctypes.data_as(_fp) for x in a]
        lib().oracle_transport_f32(ctypes.byref(g), ctypes.byref(cfg), p[0], out.ctypes.data_as(_fp), *p[1:],
                                   float(dt), int(bool(positive_definite)))
    else:
        p = [x.ctypes.data_as(_dp) for x in a]
        lib().oracle_transport_f64(ctypes.byref(g), ctypes.byref(cfg), p[0], out.ctypes.data_as(_dp), *p[1:],
                                   float(dt), int(bool(positive_definite)))
    return out


def kessler_tendencies(g, cfg, p, theta, rho, qv, qc, qr, dt):
    a = [np.ascontiguousarray(x, dtype=np.float32) for x in (p, theta, rho, qv, qc, qr)]
    outs = [np.zeros_like(a[0]) for _ in range(4)]
    lib().oracle_kessler_tendencies_f32(ctypes.byref(g), ctypes.byref(cfg), *[x.ctypes.data_as(_fp) for x in a],
                                        float(dt), *[x.ctypes.data_as(_fp) for x in outs])
    return outs


def microphysics_step(g, cfg, st, dt, dtype=np.float32, lengths=False):
    """The microphysics interval dt on a copy of st; with lengths=True also
    the lengths of the steps the state-dependent subcycling took."""
    s = copy_state(st, dtype)
    fn = lib().oracle_microphysics_step_f32 if np.dtype(dtype) == np.float32 else lib().oracle_microphysics_step_f64
    cap = 4096
    h = np.zeros(cap)
    n = fn(ctypes.byref(g), ctypes.byref(cfg), ctypes.byref(_state_struct(s)), float(dt), h.ctypes.data_as(_dp), cap)
    return (s, h[:min(n, cap)].copy()) if lengths else s


def apply_sponges(g, base, cfg, st, dt):
    """The Rayleigh sponge pass alone (FP64) on a copy of st."""
    s = copy_state(st, np.float64)
    b = {k: np.ascontiguousarray(v, dtype=np.float64) for k, v in base.items()}
    lib().oracle_apply_sponges_f64(ctypes.byref(g), ctypes.byref(base_struct(b)), ctypes.byref(cfg),
                                   ctypes.byref(_state_struct(s)), float(dt))
    return s


def sgs_terms(g, cfg, st, dtype=np.float64):
    """Smagorinsky closure and tendencies of a state (no commit): dict nu,
    k_theta, k_q, du, dv, dw, dtheta, dqv. dtype float32 keeps the state,
    coefficients and tendencies in FP32 storage as the reference's fields."""
    s = copy_state(st, dtype)
    names = ("nu", "k_theta", "k_q", "du", "dv", "dw", "dtheta", "dqv")
    out = {n: np.zeros(st["u"].shape) for n in names}
    fn = lib().oracle_sgs_terms_f64 if np.dtype(dtype) == np.float64 else lib().oracle_sgs_terms_f32
    fn(ctypes.byref(g), ctypes.byref(cfg), ctypes.byref(_state_struct(s)),
       *[out[n].ctypes.data_as(_dp) for n in names])
    return out


def apply_sgs(g, cfg, st, dtype=np.float64):
    """The split SGS integration of one large step on a copy of st."""
    s = copy_state(st, dtype)
    fn = lib().oracle_apply_sgs_f64 if np.dtype(dtype) == np.float64 else lib().oracle_apply_sgs_f32
    fn(ctypes.byref(g), ctypes.byref(cfg), ctypes.byref(_state_struct(s)))
    return s


def weno5_face(a, b, c, d, e):
    return lib().oracle_weno5_face(a, b, c, d, e)


def saturation_adjust(T, p, qv, qc):
    qv_ = ctypes.c_double(qv)
    qc_ = ctypes.c_double(qc)
    Ta = lib().oracle_saturation_adjust(T, p, ctypes.byref(qv_), ctypes.byref(qc_))
    return Ta, qv_.value, qc_.value


def diffusion_tendency(field, rho, dz, dx, dy, K_h, K_v, location, c_grid=True, periodic_x=False, periodic_y=False,
                       kv=None, dtype=np.float32):
    """Explicit diffusion tendency d(field)/dt of one [nx][ny][nz] field
    (location 0 scalar, 1 u face, 2 v face, 3 w face); kv: optional per-cell
    vertical diffusivity. Arithmetic FP64; storage `dtype`."""
    nx, ny, nz = field.shape
    f = np.ascontiguousarray(field, dtype)
    r = np.ascontiguousarray(rho, dtype)
    k = None if kv is None else np.ascontiguousarray(kv, dtype)
    dzv = np.ascontiguousarray(dz, np.float64)
    out = np.empty_like(f)
    name = "oracle_diffusion_tendency_f32" if dtype == np.float32 else "oracle_diffusion_tendency_f64"
    ptr = _fp if dtype == np.float32 else _dp
    getattr(lib(), name)(nx, ny, nz, float(dx), float(dy), dzv.ctypes.data_as(_dp), f.ctypes.data_as(ptr),
                         r.ctypes.data_as(ptr), None if k is None else k.ctypes.data_as(ptr), float(K_h), float(K_v),
                         int(location), int(bool(c_grid)), int(bool(periodic_x)), int(bool(periodic_y)),
                         out.ctypes.data_as(ptr))
    return out


def apply_diffusion(g, cfg, st):
    """The large step's explicit diffusion pass alone on a copy of st (FP64)."""
    s = copy_state(st, np.float64)
    lib().oracle_apply_diffusion_f64(ctypes.byref(g), ctypes.byref(cfg), ctypes.byref(_state_struct(s)))
    return s
