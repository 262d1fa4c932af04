"""ctypes binding of libnimbus_hip.so (the C ABI in include/nimbus_hip.h).

Host arrays use the reference layout: numpy arrays of shape (nx, ny, nz) in C
order, i.e. index (i*ny + j)*nz + k with k contiguous
(include/core/field/field3d.hpp:543-574 of the Nimbus reference).

The library is loaded from the package directory only; a missing or
unloadable library raises, there is no fallback path.
"""
import ctypes
import os
import struct

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libnimbus_hip.so")

PRECISION_FP64 = 0
PRECISION_MIXED = 1
ACOUSTIC_EXPLICIT = 0
ACOUSTIC_HEVI = 1
BC_PERIODIC, BC_RADIATIVE, BC_ZERO_GRADIENT, BC_RIGID, BC_RIGID_SLIP = 0, 1, 2, 3, 4
COLUMN_SOLVE_AUTO, COLUMN_SOLVE_SEQUENTIAL, COLUMN_SOLVE_PARTITIONED = 0, 1, 2
BOTTOM_FREE_SLIP, BOTTOM_NO_SLIP = 0, 1
TOP_RIGID, TOP_OPEN = 0, 1
SGS_3D, SGS_HORIZONTAL, SGS_VERTICAL = 0, 1, 2
SGS_FILTER_CUBIC_ROOT, SGS_FILTER_DX = 0, 1

FIELDS = ("u", "v", "w", "rho", "p", "theta", "qv", "qc", "qr")
_c_double_p = ctypes.POINTER(ctypes.c_double)
_c_float_p = ctypes.POINTER(ctypes.c_float)


class NimbusError(RuntimeError):
    pass


class Grid(ctypes.Structure):
    _fields_ = [("nx", ctypes.c_int), ("ny", ctypes.c_int), ("nz", ctypes.c_int),
                ("dx", ctypes.c_double), ("dy", ctypes.c_double), ("dz", ctypes.c_double)]


class StepConfig(ctypes.Structure):
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


DIFFUSE_SCALARS, DIFFUSE_MOMENTUM = 1, 2


class BaseState(ctypes.Structure):
    _fields_ = [(n, _c_double_p) for n in ("rho0", "p0", "theta0", "qv0", "u0", "v0")]


class HostState(ctypes.Structure):
    _fields_ = [(n, ctypes.c_void_p) for n in FIELDS] + [("rain_mm", _c_double_p)]


class Sounding(ctypes.Structure):
    _fields_ = [(n, ctypes.c_double) for n in (
        "theta_sfc", "theta_tropo", "z_tropo", "T_tropo", "rh_sfc", "rh_top", "qv_cap", "p_sfc",
        "shear", "shear_depth", "curve_depth", "curvature")]


class Bubble(ctypes.Structure):
    _fields_ = [(n, ctypes.c_double) for n in ("xc", "yc", "zc", "rh", "rz", "dtheta")]


class ICParams(ctypes.Structure):
    _fields_ = [("sounding", Sounding), ("n_bubbles", ctypes.c_int), ("bubbles", Bubble * 4),
                ("noise_amp", ctypes.c_double), ("seed", ctypes.c_uint64), ("member", ctypes.c_uint32)]


class Cm1Sounding(ctypes.Structure):
    _fields_ = ([("n", ctypes.c_int), ("capacity", ctypes.c_int)] +
                [(n, ctypes.c_double) for n in ("p_sfc", "theta_sfc", "qv_sfc")] +
                [(n, _c_double_p) for n in ("z", "theta", "qv", "u", "v", "p", "T")])


class Interp(ctypes.Structure):
    _fields_ = [("method", ctypes.c_int), ("extrapolate_below", ctypes.c_int), ("extrapolate_above", ctypes.c_int)]


class MemberDesc(ctypes.Structure):
    _fields_ = [("grid", Grid), ("config", StepConfig), ("precision", ctypes.c_int), ("base", BaseState)]


class Diag(ctypes.Structure):
    _fields_ = [(n, ctypes.c_double) for n in (
        "max_w", "min_w", "rms_theta_p", "rain_total_mm", "max_rain_mm", "mean_qc", "mean_qr")] + [
        ("steps", ctypes.c_int64)]


class AcousticFusedDispatch(ctypes.Structure):
    _fields_ = ([(n, _c_float_p) for n in ("u", "v", "w", "rho", "p")] +
                [("theta_data", _c_float_p), ("rho0_base_host", _c_double_p), ("p0_base_host", _c_double_p),
                 ("base_len", ctypes.c_int), ("nx", ctypes.c_int), ("ny", ctypes.c_int), ("nz", ctypes.c_int)] +
                [(n, ctypes.c_double) for n in ("dx_host", "dy_host", "dz_host", "dt_small_host", "gamma_host")] +
                [(n, ctypes.c_float) for n in ("rho_min", "rho_max", "p_min", "p_max", "p_ref_fallback",
                                               "wind_clamp_h", "wind_clamp_v")])


class AcousticBatchedDispatch(ctypes.Structure):
    _fields_ = ([(n, _c_float_p) for n in ("u", "v", "w", "rho", "p")] +
                [("theta_data", _c_float_p), ("rho0_base_host", _c_double_p), ("p0_base_host", _c_double_p),
                 ("nx", ctypes.c_int), ("ny", ctypes.c_int), ("nz", ctypes.c_int)] +
                [(n, ctypes.c_double) for n in ("dx", "dy", "dz", "dt_small")] +
                [("n_substeps", ctypes.c_int), ("acoustic_mode", ctypes.c_int), ("hevi_beta", ctypes.c_double)] +
                [(n, ctypes.c_int) for n in ("bc_west", "bc_east", "bc_south", "bc_north", "bc_bottom", "bc_top")] +
                [(n, _c_float_p) for n in ("mass_flux_x", "mass_flux_y", "mass_flux_z")] +
                [("accumulate_mass_flux", ctypes.c_int), ("mass_flux_weight", ctypes.c_double)])


class SlowTendenciesDispatch(ctypes.Structure):
    _fields_ = ([(n, _c_float_p) for n in ("u_data", "v_data", "w_data", "rho_data", "qv_data", "qc_data",
                                          "qr_data")] +
                [(n, _c_double_p) for n in ("rho0_base", "qv0_base", "u0_base", "v0_base")] +
                [(n, _c_float_p) for n in ("du_dt_data", "dv_dt_data", "dw_dt_data")] +
                [("nr", ctypes.c_int), ("nth", ctypes.c_int), ("nz", ctypes.c_int)] +
                [(n, ctypes.c_double) for n in ("dx", "dy", "dz", "coriolis_f")])


class SlowRk2Dispatch(ctypes.Structure):
    _fields_ = ([(n, _c_float_p) for n in ("u", "v", "w", "rho", "qv", "qc", "qr")] +
                [(n, _c_double_p) for n in ("rho0_base", "qv0_base", "u0_base", "v0_base")] +
                [("nx", ctypes.c_int), ("ny", ctypes.c_int), ("nz", ctypes.c_int)] +
                [(n, ctypes.c_double) for n in ("dx", "dy", "dz", "dt_large", "coriolis_f")])


class AdvectScalarsDispatch(ctypes.Structure):
    _fields_ = [("n_scalars", ctypes.c_int), ("q", _c_float_p * 4), ("positive_definite", ctypes.c_int * 4),
                ("mass_flux_x", _c_float_p), ("mass_flux_y", _c_float_p), ("mass_flux_z", _c_float_p),
                ("rho_begin", _c_float_p), ("rho_end", _c_float_p),
                ("nx", ctypes.c_int), ("ny", ctypes.c_int), ("nz", ctypes.c_int),
                ("dx", ctypes.c_double), ("dy", ctypes.c_double), ("dz", ctypes.c_double), ("dt", ctypes.c_double),
                ("bc_top", ctypes.c_int)]


class KesslerDispatch(ctypes.Structure):
    _fields_ = ([(n, _c_float_p) for n in ("p", "theta", "rho", "qv", "qc", "qr")] +
                [(n, _c_float_p) for n in ("dtheta_dt", "dqv_dt", "dqc_dt", "dqr_dt")] +
                [("nx", ctypes.c_int), ("ny", ctypes.c_int), ("nz", ctypes.c_int)] +
                [(n, ctypes.c_double) for n in ("dz", "dt", "qc0", "c_auto", "c_accr", "evap_eff", "vt_max")])


# ---- reference dispatch contracts as they are (field lists of
# include/compute/dispatch/{dynamics,microphysics,advection}.hpp) ----
_f = ctypes.c_float
_i = ctypes.c_int
FourFieldObserver = ctypes.CFUNCTYPE(None, ctypes.c_void_p, _c_float_p, _c_float_p, _c_float_p, _c_float_p)


class CgridSlowTendenciesDispatch(ctypes.Structure):
    _fields_ = ([(n, _c_float_p) for n in ("u_data", "v_data", "w_data", "rho_data", "p_data", "rho0_base_data")] +
                [("rho0_base_len", _i), ("qv0_base_data", _c_float_p), ("qv0_base_len", _i),
                 ("u0_base_data", _c_float_p), ("v0_base_data", _c_float_p), ("wind0_base_len", _i),
                 ("qv_data", _c_float_p), ("loading_data", _c_float_p)] +
                [(n, _c_float_p) for n in ("du_dt_data", "dv_dt_data", "dw_dt_data", "drho_dt_data", "dp_dt_data")] +
                [("nr", _i), ("nth", _i), ("nz", _i)] +
                [(n, _f) for n in ("dx", "dy", "dz", "g", "coriolis_f_val")] +
                [(n, _i) for n in ("has_moisture", "has_loading", "slow_rk2_invariant_input_residency")] +
                [(n, _c_float_p) for n in ("du_advection_dt_data", "dv_advection_dt_data", "dw_advection_dt_data")] +
                [(n, _f) for n in ("frame_velocity_u_ms", "frame_velocity_v_ms", "coriolis_beta")])


class SlowApplyDispatch(ctypes.Structure):
    _fields_ = ([(n, _c_float_p) for n in ("u", "v", "w", "p_data", "du_dt", "dv_dt", "dw_dt", "dp_dt",
                                          "du_dt_pbl", "dv_dt_pbl")] +
                [("nx", _i), ("ny", _i), ("nz", _i), ("dt_large", _f), ("completed_observer", FourFieldObserver),
                 ("completed_observer_context", ctypes.c_void_p), ("sanitize", _i)])


class StateRelaxationDispatch(ctypes.Structure):
    _fields_ = ([(n, _c_float_p) for n in ("u", "v", "w", "theta_data", "profiles")] +
                [(n, _i) for n in ("profiles_len", "nx", "ny", "nz", "lateral_width", "do_upper", "do_w_only",
                                   "do_lateral")] +
                [("dt_step", _f), ("completed_observer", FourFieldObserver),
                 ("completed_observer_context", ctypes.c_void_p)])


MICRO_SPECIES = ("theta_data", "qv_data", "qc_data", "qr_data", "qi_data", "qs_data", "qg_data", "qh_data")
MICRO_TENDENCIES = ("dtheta", "dqv", "dqc", "dqr", "dqi", "dqs", "dqg", "dqh")


class MicrophysicsApplyDispatch(ctypes.Structure):
    _fields_ = ([(n, _c_float_p) for n in MICRO_SPECIES + MICRO_TENDENCIES + ("dtheta_rad", "dtheta_pbl", "dqv_pbl")] +
                [(n, _i) for n in ("nx", "ny", "nz", "i_begin", "i_end", "j_begin", "j_end")] +
                [(n, _f) for n in ("dt_step", "theta_fallback", "theta_min", "theta_max", "max_theta_step",
                                   "qv_max_val", "hydro_max_val")] +
                [("stats_out", ctypes.POINTER(ctypes.c_uint)), ("stats_len", _i)])


class KesslerPointwiseDispatch(ctypes.Structure):
    _fields_ = ([(n, _c_float_p) for n in ("temperature_data", "p_data", "qv_data", "qc_data", "qr_data",
                                          "rho_d_data", "qh_data", "dtheta_dt_data", "dqv_dt_data", "dqc_dt_data",
                                          "dqr_dt_data", "dqg_dt_data", "dqh_dt_data")] +
                [(n, _i) for n in ("nr", "nth", "nz", "i_begin", "i_end", "j_begin", "j_end", "k_begin", "k_end")] +
                [(n, _f) for n in ("qc0", "c_auto", "c_accr", "evaporation_efficiency", "Lv_cp", "T0", "dt_step")])


class KesslerSedimentationDispatch(ctypes.Structure):
    _fields_ = ([(n, _c_float_p) for n in ("qr_data", "qg_data", "qh_data", "rho_d_data", "dqr_dt_data",
                                          "dqg_dt_data", "dqh_dt_data")] +
                [(n, _i) for n in ("nr", "nth", "nz", "i_begin", "i_end", "j_begin", "j_end", "k_begin", "k_end")] +
                [(n, _f) for n in ("dz_val", "Vt_max_rain", "vt_cap", "dt_step")] +
                [("layer_thickness_m", _c_float_p)])


class ConservativeTransportDispatch(ctypes.Structure):
    _fields_ = ([(n, _c_float_p) for n in ("u", "v", "w", "rho_data")] +
                [("q_in", ctypes.POINTER(_c_float_p)), ("q_out", ctypes.POINTER(_c_float_p)),
                 ("dt_kappas", _c_float_p), ("positive_definite", ctypes.POINTER(_i)), ("n_scalars", _i),
                 ("nx", _i), ("ny", _i), ("nz", _i)] +
                [(n, _f) for n in ("inv_dx", "inv_dy", "inv_dz", "inv_dx2", "inv_dy2", "inv_dz2", "dt_step",
                                   "rho_floor")] +
                [(n, _c_float_p) for n in ("density_begin", "mass_flux_x", "mass_flux_y", "mass_flux_z")] +
                [("use_supplied_mass_flux", _i), ("frame_velocity_u_ms", _f), ("frame_velocity_v_ms", _f),
                 ("weno", _i), ("preserve_positive_support", _i)])


class DiffusionBatchDispatch(ctypes.Structure):
    _fields_ = [("fields", ctypes.POINTER(_c_float_p)), ("tendencies", ctypes.POINTER(_c_float_p)),
                ("locations", ctypes.POINTER(_i)), ("count", _i), ("rho", _c_float_p), ("dz", _c_float_p),
                ("nx", _i), ("ny", _i), ("nz", _i), ("dx", _f), ("dy", _f), ("K_h", _f), ("K_v", _f),
                ("c_grid", _i), ("periodic_x", _i), ("periodic_y", _i),
                ("vertical_K", ctypes.POINTER(_c_float_p))]


DIFFUSION_SCALAR, DIFFUSION_U_FACE, DIFFUSION_V_FACE, DIFFUSION_W_FACE = 0, 1, 2, 3

REFERENCE_DISPATCH = {
    "nimbus_hip_dispatch_cartesian_cgrid_slow_tendencies": CgridSlowTendenciesDispatch,
    "nimbus_hip_dispatch_slow_apply": SlowApplyDispatch,
    "nimbus_hip_dispatch_state_relaxation_cartesian": StateRelaxationDispatch,
    "nimbus_hip_dispatch_microphysics_apply": MicrophysicsApplyDispatch,
    "nimbus_hip_dispatch_kessler_pointwise": KesslerPointwiseDispatch,
    "nimbus_hip_dispatch_kessler_sedimentation": KesslerSedimentationDispatch,
    "nimbus_hip_dispatch_advect_scalars_conservative_cartesian_cgrid": ConservativeTransportDispatch,
    "nimbus_hip_dispatch_diffusion_batch": DiffusionBatchDispatch,
}

_LIB = None


def load_library(path=LIB_PATH):
    """Loads libnimbus_hip.so and declares argument types. Raises if absent."""
    global _LIB
    if _LIB is not None:
        return _LIB
    if not os.path.exists(path):
        raise NimbusError(f"{path} is missing: build it with __graft_entry__.build() "
                          f"(python -m nimbus_amd._build)")
    lib = ctypes.CDLL(path)
    P = ctypes.POINTER
    vp = ctypes.c_void_p
    decl = {
        "nimbus_hip_last_error": (ctypes.c_char_p, []),
        "nimbus_hip_create": (ctypes.c_int, [ctypes.c_int, P(vp)]),
        "nimbus_hip_destroy": (ctypes.c_int, [vp]),
        "nimbus_hip_mem_info": (ctypes.c_int, [vp, P(ctypes.c_uint64), P(ctypes.c_uint64)]),
        "nimbus_hip_dispatch_count": (ctypes.c_uint64, [vp]),
        "nimbus_default_sounding": (None, [P(Sounding)]),
        "nimbus_ic_base_state": (ctypes.c_int, [P(Grid), P(Sounding)] + [_c_double_p] * 6),
        "nimbus_ic_fill_f32": (ctypes.c_int, [P(Grid), P(ICParams), P(HostState)]),
        "nimbus_ic_fill_f64": (ctypes.c_int, [P(Grid), P(ICParams), P(HostState)]),
        "nimbus_ic_fill_base_f32": (ctypes.c_int, [P(Grid), P(BaseState), P(ICParams), P(HostState)]),
        "nimbus_ic_fill_base_f64": (ctypes.c_int, [P(Grid), P(BaseState), P(ICParams), P(HostState)]),
        "nimbus_ic_sounding_levels": (ctypes.c_int, [P(Sounding), ctypes.c_int, _c_double_p, P(Cm1Sounding)]),
        "nimbus_npy_write": (ctypes.c_int, [ctypes.c_char_p, vp, ctypes.c_int, ctypes.c_int, P(ctypes.c_int64)]),
        "nimbus_hip_member_export_npy": (ctypes.c_int, [vp, ctypes.c_uint32, ctypes.c_char_p]),
        "nimbus_hip_member_snapshot_size": (ctypes.c_int, [vp, ctypes.c_uint32, P(ctypes.c_uint64)]),
        "nimbus_hip_member_snapshot": (ctypes.c_int, [vp, ctypes.c_uint32, vp, ctypes.c_uint64]),
        "nimbus_hip_member_restore": (ctypes.c_int, [vp, vp, ctypes.c_uint64, P(ctypes.c_uint32)]),
        "nimbus_hip_member_save": (ctypes.c_int, [vp, ctypes.c_uint32, ctypes.c_char_p]),
        "nimbus_hip_member_load": (ctypes.c_int, [vp, ctypes.c_char_p, P(ctypes.c_uint32)]),
        "nimbus_cm1_parse": (ctypes.c_int, [ctypes.c_char_p, ctypes.c_size_t, P(Cm1Sounding)]),
        "nimbus_cm1_read": (ctypes.c_int, [ctypes.c_char_p, P(Cm1Sounding)]),
        "nimbus_cm1_format": (ctypes.c_int, [P(Cm1Sounding), ctypes.c_char_p, ctypes.c_size_t,
                                             P(ctypes.c_size_t)]),
        "nimbus_cm1_column": (ctypes.c_int, [P(Cm1Sounding), P(Interp), ctypes.c_int, _c_double_p, ctypes.c_double]
                              + [_c_double_p] * 7),
        "nimbus_cm1_base_state": (ctypes.c_int, [P(Grid), P(Cm1Sounding), P(Interp)] + [_c_double_p] * 6),
        "nimbus_ic_uniform": (ctypes.c_double, [ctypes.c_uint64, ctypes.c_uint32, ctypes.c_int, ctypes.c_int,
                                                ctypes.c_int]),
        "nimbus_resolve_substeps": (ctypes.c_int, [P(Grid), P(BaseState), P(StepConfig)]),
        "nimbus_hip_member_create": (ctypes.c_int, [vp, P(MemberDesc), P(ctypes.c_uint32)]),
        "nimbus_hip_member_destroy": (ctypes.c_int, [vp, ctypes.c_uint32]),
        "nimbus_hip_member_bytes": (ctypes.c_uint64, [P(MemberDesc)]),
        "nimbus_hip_member_upload": (ctypes.c_int, [vp, ctypes.c_uint32, P(HostState)]),
        "nimbus_hip_member_download": (ctypes.c_int, [vp, ctypes.c_uint32, P(HostState)]),
        "nimbus_hip_member_diag": (ctypes.c_int, [vp, ctypes.c_uint32, P(Diag)]),
        "nimbus_hip_step": (ctypes.c_int, [vp, P(ctypes.c_uint32), ctypes.c_int, ctypes.c_int]),
        "nimbus_hip_step_async": (ctypes.c_int, [vp, P(ctypes.c_uint32), ctypes.c_int, ctypes.c_int]),
        "nimbus_hip_synchronize": (ctypes.c_int, [vp]),
        "nimbus_hip_set_graphs": (ctypes.c_int, [vp, ctypes.c_int]),
        "nimbus_hip_member_set_tiling": (ctypes.c_int, [vp, ctypes.c_uint32, ctypes.c_int, ctypes.c_int,
                                                        ctypes.c_int]),
        "nimbus_hip_member_set_column_solve": (ctypes.c_int, [vp, ctypes.c_uint32, ctypes.c_int]),
        "nimbus_hip_member_column_solve": (ctypes.c_int, [vp, ctypes.c_uint32, P(ctypes.c_int)]),
        "nimbus_hip_member_autotune": (ctypes.c_int, [vp, ctypes.c_uint32, ctypes.c_int, P(ctypes.c_int),
                                                      P(ctypes.c_int), P(ctypes.c_int), P(ctypes.c_double)]),
        "nimbus_hip_set_kernel_timing": (ctypes.c_int, [vp, ctypes.c_int]),
        "nimbus_hip_kernel_timing": (ctypes.c_int, [vp, ctypes.c_int, vp, P(ctypes.c_uint64),
                                                    P(ctypes.c_double)]),
        "nimbus_hip_reset_kernel_timing": (ctypes.c_int, [vp]),
        "nimbus_hip_last_step_ms": (ctypes.c_double, [vp]),
        "nimbus_hip_dispatch_acoustic_substep_fused": (ctypes.c_int, [vp, P(AcousticFusedDispatch)]),
        "nimbus_hip_dispatch_acoustic_substeps_batched": (ctypes.c_int, [vp, P(AcousticBatchedDispatch)]),
        "nimbus_hip_dispatch_slow_tendencies": (ctypes.c_int, [vp, P(SlowTendenciesDispatch)]),
        "nimbus_hip_dispatch_slow_rk2": (ctypes.c_int, [vp, P(SlowRk2Dispatch)]),
        "nimbus_hip_dispatch_advect_scalars": (ctypes.c_int, [vp, P(AdvectScalarsDispatch)]),
        "nimbus_hip_dispatch_kessler": (ctypes.c_int, [vp, P(KesslerDispatch)]),
    }
    decl["nimbus_hip_successful_dispatch_count"] = (ctypes.c_uint64, [vp, ctypes.c_char_p])
    decl["nimbus_hip_device_arch"] = (ctypes.c_int, [vp, ctypes.c_char_p, ctypes.c_size_t])
    decl["nimbus_hip_batch_stats"] = (ctypes.c_int, [vp, P(ctypes.c_uint64), P(ctypes.c_double), P(ctypes.c_uint64),
                                                     P(ctypes.c_double)])
    for name, st in REFERENCE_DISPATCH.items():
        decl[name] = (ctypes.c_int, [vp, P(st)])
    decl["nimbus_hip_advect_scalars_conservative_noalloc"] = (ctypes.c_int, [vp, P(ConservativeTransportDispatch)])
    for name, (res, args) in decl.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _LIB = lib
    return lib


def check(rc, what=""):
    if rc != 0:
        msg = load_library().nimbus_hip_last_error().decode()
        raise NimbusError(f"{what} failed ({rc}): {msg}")


def fptr(a):
    return a.ctypes.data_as(_c_float_p)


def dptr(a):
    return a.ctypes.data_as(_c_double_p)


# ---------------------------------------------------------------- host-side helpers

def default_sounding(**overrides):
    s = Sounding()
    load_library().nimbus_default_sounding(ctypes.byref(s))
    for k, v in overrides.items():
        setattr(s, k, v)
    return s


def make_grid(nx, ny, nz, dx, dy, dz):
    return Grid(int(nx), int(ny), int(nz), float(dx), float(dy), float(dz))


def step_config(dt, substeps, acoustic_mode=ACOUSTIC_EXPLICIT, hevi_beta=0.1, coriolis_f=0.0,
                lateral=BC_RADIATIVE, bottom=BOTTOM_FREE_SLIP, top=TOP_RIGID,
                qc0=1.0e-3, c_auto=1.0e-3, c_accr=2.2, evap_eff=1.0, vt_max=20.0, sponge=None, sgs=None,
                diffusion=None):
    """Large-step configuration. `sponge` is None (off) or a dict of the
    Rayleigh sponge settings (see sponge_preset); `sgs` is None (off) or a
    dict of Smagorinsky settings (see sgs_preset); `diffusion` is None (off)
    or {"apply": "scalars" | "momentum" | "all", "kh": K_h, "kv": K_v} for the
    explicit Laplacian diffusion (m^2/s)."""
    dif = diffusion or {}
    dcode = {"scalars": DIFFUSE_SCALARS, "momentum": DIFFUSE_MOMENTUM,
             "all": DIFFUSE_SCALARS | DIFFUSE_MOMENTUM}[dif.get("apply", "all")] if diffusion else 0
    lat = lateral if isinstance(lateral, (tuple, list)) else (lateral,) * 4
    sp = sponge or {}
    sg = sgs or {}
    return StepConfig(float(dt), int(substeps), int(acoustic_mode), float(hevi_beta), float(coriolis_f),
                      lat[0], lat[1], lat[2], lat[3], int(bottom), int(top),
                      qc0, c_auto, c_accr, evap_eff, vt_max,
                      int(bool(sponge)), float(sp.get("z_d", 12000.0)), float(sp.get("alpha", 0.05)),
                      int(sp.get("w_cells", 8)), float(sp.get("w_alpha", 0.2)),
                      int(sp.get("lateral_cells", 0)), float(sp.get("lateral_alpha", 0.05)),
                      int(bool(sgs)), int(sg.get("mode", SGS_3D)), int(sg.get("filter", SGS_FILTER_CUBIC_ROOT)),
                      float(sg.get("cs", 0.18)), float(sg.get("pr_t", 1.0 / 3.0)), float(sg.get("sc_t", 1.0 / 3.0)),
                      float(sg.get("nu_max", 1000.0)), float(sg.get("k_max", 1000.0)), float(sg.get("dt", 1.0)),
                      dcode, float(dif.get("kh", 0.0)), float(dif.get("kv", 0.0)))


def sponge_preset(dx):
    """The reference Kessler presets' sponges (configs/*meterResolution/kessler/baseline.yaml:
    damping): upper sponge from 12 km at 0.05 1/s, 8-cell w-only absorber at 0.2 1/s under
    the lid, and a 10 km lateral band at 0.05 1/s."""
    return {"z_d": 12000.0, "alpha": 0.05, "w_cells": 8, "w_alpha": 0.2,
            "lateral_cells": int(round(10000.0 / dx)), "lateral_alpha": 0.05}


def sgs_preset():
    """The reference Kessler presets' turbulence block (configs/*meterResolution/kessler/
    baseline.yaml: turbulence): Smagorinsky, mode 2d (horizontal only), Cs 0.18,
    Pr_t = Sc_t = 1/3, cube-root filter width, caps 1000 m^2/s, 1 s SGS substeps."""
    return {"mode": SGS_HORIZONTAL, "filter": SGS_FILTER_CUBIC_ROOT, "cs": 0.18, "pr_t": 1.0 / 3.0,
            "sc_t": 1.0 / 3.0, "nu_max": 1000.0, "k_max": 1000.0, "dt": 1.0}


def base_state(grid, sounding=None):
    """FP64 base-state profiles (dict of numpy arrays, each nz long)."""
    lib = load_library()
    s = sounding if sounding is not None else default_sounding()
    out = {n: np.zeros(grid.nz) for n in ("rho0", "p0", "theta0", "qv0", "u0", "v0")}
    check(lib.nimbus_ic_base_state(ctypes.byref(grid), ctypes.byref(s),
                                   *[dptr(out[n]) for n in ("rho0", "p0", "theta0", "qv0", "u0", "v0")]),
          "nimbus_ic_base_state")
    return out


def base_struct(base):
    return BaseState(*[dptr(base[n]) for n in ("rho0", "p0", "theta0", "qv0", "u0", "v0")])


def ic_params(grid, sounding=None, bubbles=None, noise_amp=0.1, seed=0, member=0):
    """Initial-condition parameters; the default is one +2 K bubble of 10 km x 1.4 km
    radius, centred 1.4 km above the domain centre."""
    p = ICParams()
    p.sounding = sounding if sounding is not None else default_sounding()
    if bubbles is None:
        xc = (grid.nx - 2) * grid.dx / 2.0
        yc = (grid.ny - 2) * grid.dy / 2.0
        bubbles = [(xc, yc, 1400.0, 10000.0, 1400.0, 2.0)]
    p.n_bubbles = len(bubbles)
    for q, b in enumerate(bubbles):
        p.bubbles[q] = Bubble(*b)
    p.noise_amp = noise_amp
    p.seed = seed
    p.member = member
    return p


def new_state(grid, dtype):
    st = {n: np.zeros((grid.nx, grid.ny, grid.nz), dtype=dtype) for n in FIELDS}
    st["rain_mm"] = np.zeros((grid.nx, grid.ny), dtype=np.float64)
    return st


def host_state_struct(st):
    return HostState(*[st[n].ctypes.data_as(ctypes.c_void_p) if st.get(n) is not None else None for n in FIELDS],
                     dptr(st["rain_mm"]) if st.get("rain_mm") is not None else None)


def initial_state(grid, params, dtype=np.float32):
    """Seeded initial state in the reference host layout."""
    lib = load_library()
    st = new_state(grid, dtype)
    hs = host_state_struct(st)
    fill = lib.nimbus_ic_fill_f32 if np.dtype(dtype) == np.float32 else lib.nimbus_ic_fill_f64
    check(fill(ctypes.byref(grid), ctypes.byref(params), ctypes.byref(hs)), "nimbus_ic_fill")
    return st


def initial_state_from_base(grid, base, params, dtype=np.float32):
    """Seeded initial state (bubbles, noise) over given base profiles, e.g. a
    CM1 sounding's cm1_base_state."""
    lib = load_library()
    st = new_state(grid, dtype)
    hs = host_state_struct(st)
    b = {k: np.ascontiguousarray(v, dtype=np.float64) for k, v in base.items()}
    fill = lib.nimbus_ic_fill_base_f32 if np.dtype(dtype) == np.float32 else lib.nimbus_ic_fill_base_f64
    check(fill(ctypes.byref(grid), ctypes.byref(base_struct(b)), ctypes.byref(params), ctypes.byref(hs)),
          "nimbus_ic_fill_base")
    return st


# ---------------------------------------------------------------- state I/O

NPY_F4, NPY_F8 = 0, 1
# member snapshot header: magic, version, endian mark, grid and config sizes,
# precision, tile_w, k_chunks, xcd_remap, column-solve form, reserved, steps,
# field bytes, rain bytes
SNAP_HEAD = "<8s4I6iq2Q"


def npy_write(path, array):
    """Write a float32/float64 C-order array (1-3 dims) with the library's NPY writer."""
    a = np.ascontiguousarray(array)
    if a.dtype not in (np.float32, np.float64) or not 1 <= a.ndim <= 3:
        raise NimbusError("npy_write: float32/float64 arrays of 1 to 3 dimensions")
    shape = (ctypes.c_int64 * a.ndim)(*a.shape)
    check(load_library().nimbus_npy_write(os.fsencode(path), a.ctypes.data_as(ctypes.c_void_p),
                                          NPY_F4 if a.dtype == np.float32 else NPY_F8, a.ndim, shape),
          "nimbus_npy_write")


# ---------------------------------------------------------------- CM1 input_sounding

CM1_COLUMNS = ("z", "theta", "qv", "u", "v", "p", "T")
INTERP_LINEAR, INTERP_SPLINE, INTERP_LOG_PRESSURE = 0, 1, 2


def _cm1_struct(col):
    """Cm1Sounding over the arrays of a column dict (keeps no copies: the
    caller keeps `col` alive while the struct is in use)."""
    n = len(col["z"])
    s = Cm1Sounding()
    s.n = n
    s.capacity = n
    s.p_sfc, s.theta_sfc, s.qv_sfc = (float(col[k]) for k in ("p_sfc", "theta_sfc", "qv_sfc"))
    for k in CM1_COLUMNS:
        if col.get(k) is not None:
            setattr(s, k, dptr(col[k]))
    return s


def _cm1_empty(capacity):
    col = {k: np.zeros(capacity) for k in CM1_COLUMNS}
    col.update(p_sfc=0.0, theta_sfc=0.0, qv_sfc=0.0)
    s = _cm1_struct(col)
    s.n = 0
    return col, s


def _cm1_result(col, s):
    out = {k: col[k][:s.n].copy() for k in CM1_COLUMNS}
    out.update(p_sfc=s.p_sfc, theta_sfc=s.theta_sfc, qv_sfc=s.qv_sfc)
    return out


def cm1_parse(text):
    """CM1 input_sounding text -> column dict (SI units; p and T reconstructed)."""
    lib = load_library()
    data = text.encode() if isinstance(text, str) else bytes(text)
    col, s = _cm1_empty(max(8, data.count(b"\n") + 2))
    check(lib.nimbus_cm1_parse(data, len(data), ctypes.byref(s)), "nimbus_cm1_parse")
    return _cm1_result(col, s)


def cm1_read(path):
    """CM1 input_sounding file -> column dict."""
    lib = load_library()
    size = os.path.getsize(path) if os.path.isfile(path) else 0
    col, s = _cm1_empty(max(8, size // 8 + 2))
    check(lib.nimbus_cm1_read(os.fsencode(path), ctypes.byref(s)), "nimbus_cm1_read")
    return _cm1_result(col, s)


def cm1_format(col):
    """Column dict -> CM1 input_sounding text."""
    lib = load_library()
    c = {k: (np.ascontiguousarray(col[k], dtype=np.float64) if k in CM1_COLUMNS and col.get(k) is not None
             else col.get(k)) for k in list(CM1_COLUMNS) + ["p_sfc", "theta_sfc", "qv_sfc"]}
    s = _cm1_struct(c)
    need = ctypes.c_size_t(0)
    lib.nimbus_cm1_format(ctypes.byref(s), None, 0, ctypes.byref(need))
    buf = ctypes.create_string_buffer(need.value + 1)
    check(lib.nimbus_cm1_format(ctypes.byref(s), buf, len(buf), ctypes.byref(need)), "nimbus_cm1_format")
    return buf.value.decode()


def _interp(method, extrapolate_below, extrapolate_above):
    return Interp(int(method), int(bool(extrapolate_below)), int(bool(extrapolate_above)))


def cm1_column(col, z, dz, method=INTERP_LINEAR, extrapolate_below=False, extrapolate_above=False):
    """Model column at heights z from a CM1 column dict: rho0, p0, theta0,
    qv0, u0, v0 and T0."""
    lib = load_library()
    c = {k: (np.ascontiguousarray(col[k], dtype=np.float64) if k in CM1_COLUMNS else col[k])
         for k in list(CM1_COLUMNS) + ["p_sfc", "theta_sfc", "qv_sfc"]}
    s = _cm1_struct(c)
    z = np.ascontiguousarray(z, dtype=np.float64)
    names = ("rho0", "p0", "theta0", "qv0", "u0", "v0", "T0")
    out = {n: np.zeros(len(z)) for n in names}
    how = _interp(method, extrapolate_below, extrapolate_above)
    check(lib.nimbus_cm1_column(ctypes.byref(s), ctypes.byref(how), len(z), dptr(z), float(dz),
                                *[dptr(out[n]) for n in names]), "nimbus_cm1_column")
    return out


def cm1_base_state(grid, col, method=INTERP_LINEAR, extrapolate_below=False, extrapolate_above=False):
    """Base-state profiles at the cell centres of `grid` from a CM1 column dict."""
    lib = load_library()
    c = {k: (np.ascontiguousarray(col[k], dtype=np.float64) if k in CM1_COLUMNS else col[k])
         for k in list(CM1_COLUMNS) + ["p_sfc", "theta_sfc", "qv_sfc"]}
    s = _cm1_struct(c)
    names = ("rho0", "p0", "theta0", "qv0", "u0", "v0")
    out = {n: np.zeros(grid.nz) for n in names}
    how = _interp(method, extrapolate_below, extrapolate_above)
    check(lib.nimbus_cm1_base_state(ctypes.byref(grid), ctypes.byref(s), ctypes.byref(how),
                                    *[dptr(out[n]) for n in names]), "nimbus_cm1_base_state")
    return out


def sounding_levels(z, sounding=None):
    """The analytic sounding sampled at heights z (z[0] = 0) as a CM1 column dict."""
    lib = load_library()
    z = np.ascontiguousarray(z, dtype=np.float64)
    snd = sounding if sounding is not None else default_sounding()
    col, s = _cm1_empty(len(z))
    check(lib.nimbus_ic_sounding_levels(ctypes.byref(snd), len(z), dptr(z), ctypes.byref(s)),
          "nimbus_ic_sounding_levels")
    return _cm1_result(col, s)


def member_bytes(grid, config, base, precision=PRECISION_MIXED):
    """Device bytes a member of this shape holds (admission planning; no device needed)."""
    b = {k: np.ascontiguousarray(v, dtype=np.float64) for k, v in base.items()}
    desc = MemberDesc(grid, config, precision, base_struct(b))
    return int(load_library().nimbus_hip_member_bytes(ctypes.byref(desc)))


def resolve_substeps(grid, base, config):
    return load_library().nimbus_resolve_substeps(ctypes.byref(grid), ctypes.byref(base_struct(base)),
                                                  ctypes.byref(config))


# ---------------------------------------------------------------- device context

class Context:
    """One HIP device context (one per host thread)."""

    def __init__(self, device=0):
        self.lib = load_library()
        self._h = ctypes.c_void_p()
        check(self.lib.nimbus_hip_create(int(device), ctypes.byref(self._h)), "nimbus_hip_create")

    @property
    def handle(self):
        return self._h

    def close(self):
        if self._h:
            self.lib.nimbus_hip_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def mem_info(self):
        f, t = ctypes.c_uint64(), ctypes.c_uint64()
        check(self.lib.nimbus_hip_mem_info(self._h, ctypes.byref(f), ctypes.byref(t)), "mem_info")
        return f.value, t.value

    def member(self, grid, config, base, precision=PRECISION_MIXED):
        return Member(self, grid, config, base, precision)

    def restore(self, blob):
        """A new member from a snapshot (Member.snapshot or the file Member.save wrote)."""
        if isinstance(blob, (str, os.PathLike)):
            with open(blob, "rb") as f:
                blob = f.read()
        if not isinstance(blob, (bytes, bytearray)):
            blob = bytes(blob)
        head = struct.unpack_from(SNAP_HEAD, blob, 0)
        off = struct.calcsize(SNAP_HEAD)
        grid = Grid.from_buffer_copy(blob, off)
        config = StepConfig.from_buffer_copy(blob, off + ctypes.sizeof(Grid))
        off += ctypes.sizeof(Grid) + ctypes.sizeof(StepConfig)
        prof = np.frombuffer(blob, dtype="<f8", count=6 * grid.nz, offset=off).reshape(6, grid.nz)
        base = {n: prof[q].copy() for q, n in enumerate(("rho0", "p0", "theta0", "qv0", "u0", "v0"))}
        # the library reads the snapshot in place: no host copy of the fields
        if isinstance(blob, bytes):
            ptr = ctypes.cast(ctypes.c_char_p(blob), ctypes.c_void_p)
        else:
            ptr = ctypes.cast((ctypes.c_char * len(blob)).from_buffer(blob), ctypes.c_void_p)
        mid = ctypes.c_uint32()
        check(self.lib.nimbus_hip_member_restore(self._h, ptr, len(blob), ctypes.byref(mid)), "member_restore")
        return Member(self, grid, config, base, head[5], existing_id=mid.value)

    def step(self, members, n_steps=1, wait=True):
        ids = (ctypes.c_uint32 * len(members))(*[m.id for m in members])
        fn = self.lib.nimbus_hip_step if wait else self.lib.nimbus_hip_step_async
        check(fn(self._h, ids, len(members), int(n_steps)), "nimbus_hip_step")

    def synchronize(self):
        check(self.lib.nimbus_hip_synchronize(self._h), "synchronize")

    def set_graphs(self, enable):
        check(self.lib.nimbus_hip_set_graphs(self._h, int(bool(enable))), "set_graphs")

    def set_kernel_timing(self, enable):
        check(self.lib.nimbus_hip_set_kernel_timing(self._h, int(bool(enable))), "set_kernel_timing")

    def reset_kernel_timing(self):
        check(self.lib.nimbus_hip_reset_kernel_timing(self._h), "reset_kernel_timing")

    def kernel_timing(self, max_rows=64):
        names = (ctypes.c_char * 48 * max_rows)()
        launches = (ctypes.c_uint64 * max_rows)()
        ms = (ctypes.c_double * max_rows)()
        n = self.lib.nimbus_hip_kernel_timing(self._h, max_rows, ctypes.cast(names, ctypes.c_void_p), launches, ms)
        check(0 if n >= 0 else n, "kernel_timing")
        return {bytes(names[r]).split(b"\0")[0].decode(): (launches[r], ms[r]) for r in range(n)}

    def last_step_ms(self):
        return self.lib.nimbus_hip_last_step_ms(self._h)

    def device_arch(self):
        """Architecture name of the context's device, without feature suffixes (e.g. 'gfx950')."""
        buf = ctypes.create_string_buffer(128)
        check(self.lib.nimbus_hip_device_arch(self._h, buf, 128), "device_arch")
        return buf.value.decode().split(":")[0]

    def batch_stats(self):
        """Membership-change costs so far: batch builds and graph captures (counts, host ms)."""
        b, c = ctypes.c_uint64(), ctypes.c_uint64()
        bm, cm = ctypes.c_double(), ctypes.c_double()
        check(self.lib.nimbus_hip_batch_stats(self._h, ctypes.byref(b), ctypes.byref(bm), ctypes.byref(c),
                                              ctypes.byref(cm)), "batch_stats")
        return {"batch_builds": b.value, "batch_build_ms": bm.value, "graph_captures": c.value,
                "graph_capture_ms": cm.value}

    def dispatch_count(self):
        return self.lib.nimbus_hip_dispatch_count(self._h)


class Member:
    """A device-resident ensemble member."""

    def __init__(self, ctx, grid, config, base, precision, existing_id=None):
        self.ctx = ctx
        self.grid = grid
        self.config = config
        self.precision = precision
        self.dtype = np.float32 if precision == PRECISION_MIXED else np.float64
        self._base = {k: np.ascontiguousarray(v, dtype=np.float64) for k, v in base.items()}
        desc = MemberDesc(grid, config, precision, base_struct(self._base))
        self.bytes = ctx.lib.nimbus_hip_member_bytes(ctypes.byref(desc))
        if existing_id is None:
            mid = ctypes.c_uint32()
            check(ctx.lib.nimbus_hip_member_create(ctx.handle, ctypes.byref(desc), ctypes.byref(mid)),
                  "member_create")
            existing_id = mid.value
        self.id = existing_id
        self.substeps = resolve_substeps(grid, self._base, config)

    def upload(self, st):
        st = {k: (np.ascontiguousarray(v, dtype=self.dtype) if k in FIELDS else
                  (np.ascontiguousarray(v, dtype=np.float64) if v is not None else None)) for k, v in st.items()}
        hs = host_state_struct(st)
        check(self.ctx.lib.nimbus_hip_member_upload(self.ctx.handle, self.id, ctypes.byref(hs)), "upload")

    def download(self):
        st = new_state(self.grid, self.dtype)
        hs = host_state_struct(st)
        check(self.ctx.lib.nimbus_hip_member_download(self.ctx.handle, self.id, ctypes.byref(hs)), "download")
        return st

    def diag(self):
        d = Diag()
        check(self.ctx.lib.nimbus_hip_member_diag(self.ctx.handle, self.id, ctypes.byref(d)), "diag")
        return {n: getattr(d, n) for n, _ in Diag._fields_}

    def export_npy(self, directory):
        """u, v, w, theta, p, rho, qv, qc, qr and accumulated_rainfall as .npy files in `directory`."""
        check(self.ctx.lib.nimbus_hip_member_export_npy(self.ctx.handle, self.id, os.fsencode(directory)),
              "export_npy")

    def snapshot(self):
        """The member's snapshot as a bytearray, written in place by the library
        (restore with Context.restore)."""
        n = ctypes.c_uint64()
        check(self.ctx.lib.nimbus_hip_member_snapshot_size(self.ctx.handle, self.id, ctypes.byref(n)),
              "snapshot_size")
        blob = bytearray(n.value)
        ptr = ctypes.cast((ctypes.c_char * n.value).from_buffer(blob), ctypes.c_void_p)
        check(self.ctx.lib.nimbus_hip_member_snapshot(self.ctx.handle, self.id, ptr, n.value), "snapshot")
        return blob

    def save(self, path):
        check(self.ctx.lib.nimbus_hip_member_save(self.ctx.handle, self.id, os.fsencode(path)), "member_save")

    def set_tiling(self, tile_w, k_chunks, xcd_remap):
        """Launch geometry: tile_w in {32, 64} (column tiles), or 0 (per-cell work in runs of 256
        consecutive cells of a level, per-column work in runs of 256 columns); k_chunks in 1..nz;
        XCD-contiguous order on/off."""
        check(self.ctx.lib.nimbus_hip_member_set_tiling(self.ctx.handle, self.id, int(tile_w), int(k_chunks),
                                                        int(xcd_remap)), "set_tiling")

    def set_column_solve(self, form):
        """HE-VI column-solve form (COLUMN_SOLVE_*), fixed for the member until changed."""
        check(self.ctx.lib.nimbus_hip_member_set_column_solve(self.ctx.handle, self.id, int(form)),
              "set_column_solve")

    def column_solve(self):
        f = ctypes.c_int()
        check(self.ctx.lib.nimbus_hip_member_column_solve(self.ctx.handle, self.id, ctypes.byref(f)),
              "column_solve")
        return f.value

    def autotune(self, trials=2):
        """Times each launch-geometry candidate on a saved copy of the state; returns (tile_w, k_chunks, remap, ms)."""
        w, c, r, ms = ctypes.c_int(), ctypes.c_int(), ctypes.c_int(), ctypes.c_double()
        check(self.ctx.lib.nimbus_hip_member_autotune(self.ctx.handle, self.id, int(trials), ctypes.byref(w),
                                                      ctypes.byref(c), ctypes.byref(r), ctypes.byref(ms)),
              "autotune")
        return w.value, c.value, r.value, ms.value

    def destroy(self):
        if self.id:
            check(self.ctx.lib.nimbus_hip_member_destroy(self.ctx.handle, self.id), "member_destroy")
            self.id = 0
