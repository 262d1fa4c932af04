"""Explicit Laplacian diffusion on the device.

The batch contract (the reference's DiffusionBatchDispatch as it is) through
the C ABI: every
tendency of the six golden cases written by the reference's own
ExplicitDiffusionScheme (tests/golden/diffusion.npz: C- and A-grid, closed
and periodic sides, stretched layers, per-cell vertical diffusivity,
horizontal-only and vertical-only) bitwise, all five fields of a case in one
batch; a batch of one field against the oracle on a larger grid; and the
refusals leave the caller's tendencies untouched.

In the resident large step (after the SGS, before the EOS projection and the
sponges): FP64 members against the FP64 oracle with momentum and scalar
diffusion, periodic and closed sides, with and without SGS and sponges; a
mixed member against the FP64 oracle at its FP32 rounding bound; zero
coefficients leave the step bitwise as without diffusion."""
import ctypes
import os

import numpy as np
import pytest

from nimbus_amd import lib as L
from oracle import oracle_py as O
from tests.helpers import max_abs, oracle_grid, small_case
from tests.test_gpu_parity import SMALL_SPONGE, _run_member, check_fp64, check_mixed

pytestmark = pytest.mark.gpu

F32 = np.float32
HERE = os.path.dirname(os.path.abspath(__file__))
GOLDEN = np.load(os.path.join(HERE, "golden", "diffusion.npz"))
LOCS = (("u", 1, "du"), ("v", 2, "dv"), ("w", 3, "dw"), ("theta", 0, "dtheta"), ("qv", 0, "dqv"))
FP = ctypes.POINTER(ctypes.c_float)


def fp(a):
    return a.ctypes.data_as(FP) if a is not None else None


def dispatch(ctx, fields, locs, rho, dz, dx, dy, kh, kv, c_grid, px, py, kvs=None, tend=None):
    nx, ny, nz = rho.shape
    fields = [np.ascontiguousarray(f, F32) for f in fields]
    tend = tend if tend is not None else [np.full((nx, ny, nz), np.nan, F32) for _ in fields]
    fa = (FP * len(fields))(*[fp(f) for f in fields])
    ta = (FP * len(fields))(*[fp(t) for t in tend])
    la = (ctypes.c_int * len(fields))(*locs)
    ka = None if kvs is None else (FP * len(fields))(*[fp(k) for k in kvs])
    dzf = np.ascontiguousarray(dz, F32)
    rhof = np.ascontiguousarray(rho, F32)
    d = L.DiffusionBatchDispatch(fa, ta, la, len(fields), fp(rhof), fp(dzf), nx, ny, nz, dx, dy, kh, kv,
                                 int(c_grid), int(px), int(py), ka)
    rc = ctx.lib.nimbus_hip_dispatch_diffusion_batch(ctx.handle, ctypes.byref(d))
    return rc, tend


def _cases():
    return sorted({k.split("/")[0] for k in GOLDEN.files if "/" in k})


@pytest.mark.parametrize("name", _cases())
def test_golden_cases_bitwise(hip_ctx, name):
    NX, NY, NZ, DX, DY = GOLDEN["grid"]
    kh, kv, cg, px, py, var, apply, _ = GOLDEN[f"{name}/settings"]
    print(f"case: diffusion golden {name}", flush=True)
    rho = GOLDEN[f"{name}/in/rho"]
    dz = GOLDEN[f"{name}/in/dz"]
    sel = [(f, loc, o) for f, loc, o in LOCS if apply == 2 or (apply == 1 and loc > 0) or (apply == 0 and loc == 0)]
    kvf = GOLDEN[f"{name}/in/kv"] if var else None
    n0 = hip_ctx.lib.nimbus_hip_successful_dispatch_count(hip_ctx.handle, b"diffusion_batch")
    rc, tend = dispatch(hip_ctx, [GOLDEN[f"{name}/in/{f}"] for f, _, _ in sel], [loc for _, loc, _ in sel], rho, dz,
                        float(DX), float(DY), kh, kv, cg, px, py, kvs=None if kvf is None else [kvf] * len(sel))
    assert rc == 0, hip_ctx.lib.nimbus_hip_last_error().decode()
    assert hip_ctx.lib.nimbus_hip_successful_dispatch_count(hip_ctx.handle, b"diffusion_batch") == n0 + 1
    for (f, _, o), t in zip(sel, tend):
        ref = GOLDEN[f"{name}/out/{o}"]
        print(f"  {f}: max diff {float(np.max(np.abs(t.astype(np.float64) - ref))):.3e}", flush=True)
        assert np.array_equal(t, ref), (name, f)


@pytest.mark.parametrize("loc", [0, 1, 2, 3])
def test_single_field_matches_oracle_on_a_larger_grid(hip_ctx, loc):
    nx, ny, nz = 40, 36, 30
    rng = np.random.default_rng(loc)
    dz = (150.0 * 1.05 ** np.arange(nz)).astype(F32)
    rho = (1.1 * np.exp(-np.cumsum(dz)[None, None, :] / 9000.0) * (1.0 + 1e-3 * rng.standard_normal((nx, ny, nz))))
    rho = rho.astype(F32)
    f = (5.0 * rng.standard_normal((nx, ny, nz))).astype(F32)
    ref = O.diffusion_tendency(f, rho, dz.astype(np.float64), 600.0, 550.0, 75.0, 12.0, loc, True, True, False)
    rc, (t,) = dispatch(hip_ctx, [f], [loc], rho, dz, 600.0, 550.0, 75.0, 12.0, True, True, False)
    assert rc == 0, hip_ctx.lib.nimbus_hip_last_error().decode()
    assert np.array_equal(t, ref)


def test_refusals_leave_tendencies_untouched(hip_ctx):
    nx, ny, nz = 10, 9, 8
    rng = np.random.default_rng(3)
    rho = np.ones((nx, ny, nz), F32)
    dz = np.full(nz, 200.0, F32)
    f = rng.standard_normal((nx, ny, nz)).astype(F32)
    bad_rho = rho.copy()
    bad_rho[4, 4, 3] = 0.0
    bad_f = f.copy()
    bad_f[5, 5, 2] = np.nan
    bad_dz = dz.copy()
    bad_dz[3] = -1.0
    for label, args in (("density", (f, 0, bad_rho, dz, 10.0, 1.0)), ("field", (bad_f, 0, rho, dz, 10.0, 1.0)),
                        ("layer", (f, 0, rho, bad_dz, 10.0, 1.0)), ("location", (f, 4, rho, dz, 10.0, 1.0)),
                        ("diffusivity", (f, 0, rho, dz, -1.0, 1.0))):
        ff, loc, r, z, kh, kv = args
        tend = [np.full((nx, ny, nz), 7.0, F32)]
        rc, tend = dispatch(hip_ctx, [ff], [loc], r, z, 500.0, 500.0, kh, kv, True, False, False, tend=tend)
        assert rc == -1, label  # NIMBUS_EINVAL
        assert np.all(tend[0] == 7.0), label


@pytest.mark.parametrize("apply,lateral,sgs,sponge", [
    ("all", L.BC_RIGID_SLIP, False, False),
    ("momentum", L.BC_PERIODIC, True, True),
    ("scalars", L.BC_RADIATIVE, False, True),
    ("all", (L.BC_PERIODIC, L.BC_PERIODIC, L.BC_RIGID_SLIP, L.BC_RIGID_SLIP), True, False),
])
@pytest.mark.parametrize("mode", [L.ACOUSTIC_EXPLICIT, L.ACOUSTIC_HEVI])
def test_large_step_with_diffusion_fp64_matches_fp64_oracle(hip_ctx, mode, apply, lateral, sgs, sponge):
    g, base, st = small_case(nx=20, ny=18, nz=14, cloudy=True, wind_noise=2.0, dtype=np.float64)
    dif = {"apply": apply, "kh": 2000.0, "kv": 50.0}
    kw = dict(acoustic_mode=mode, lateral=lateral, sponge=SMALL_SPONGE if sponge else None,
              sgs=L.sgs_preset() if sgs else None)
    cfg = L.step_config(3.0, 6, diffusion=dif, **kw)
    n_steps = 3
    print(f"case: diffusion step {apply} lateral {lateral} sgs {sgs} sponge {sponge} mode {mode}", flush=True)
    ref, _ = O.large_steps(oracle_grid(g), base, O.from_step_config(cfg), st, n_steps, np.float64)
    out, _ = _run_member(hip_ctx, g, base, cfg, st, L.PRECISION_FP64, n_steps)
    check_fp64(out, ref, f"diffusion {apply}", lateral == L.BC_RADIATIVE, diffusive=True)
    plain, _ = O.large_steps(oracle_grid(g), base, O.from_step_config(L.step_config(3.0, 6, **kw)), st, n_steps,
                             np.float64)
    moved = "u" if apply != "scalars" else "theta"
    assert max_abs(plain[moved], ref[moved]) > 1e-4  # the diffusion did work in this case


def test_large_step_with_diffusion_mixed_matches_fp64_oracle(hip_ctx):
    g, base, st64 = small_case(nx=20, ny=18, nz=14, cloudy=True, wind_noise=2.0, dtype=np.float64)
    st = {k: (v.astype(np.float32) if k in L.FIELDS else v) for k, v in st64.items()}
    cfg = L.step_config(3.0, 6, acoustic_mode=L.ACOUSTIC_HEVI, lateral=L.BC_RIGID_SLIP, sponge=SMALL_SPONGE,
                        sgs=L.sgs_preset(), diffusion={"apply": "all", "kh": 2000.0, "kv": 50.0})
    ref, _ = O.large_steps(oracle_grid(g), base, O.from_step_config(cfg), st64, 3, np.float64)
    out, _ = _run_member(hip_ctx, g, base, cfg, st, L.PRECISION_MIXED, 3)
    check_mixed(out, ref, "diffusion mixed")


@pytest.mark.parametrize("precision", [L.PRECISION_MIXED, L.PRECISION_FP64])
def test_zero_coefficients_leave_the_step_as_without_diffusion(hip_ctx, precision):
    dtype = np.float32 if precision == L.PRECISION_MIXED else np.float64
    g, base, st = small_case(nx=20, ny=18, nz=14, cloudy=True, wind_noise=1.0, dtype=dtype)
    kw = dict(acoustic_mode=L.ACOUSTIC_HEVI, lateral=L.BC_RIGID_SLIP, sponge=SMALL_SPONGE, sgs=L.sgs_preset())
    a, _ = _run_member(hip_ctx, g, base, L.step_config(3.0, 6, **kw), st, precision, 2)
    b, _ = _run_member(hip_ctx, g, base, L.step_config(3.0, 6, diffusion={"apply": "all", "kh": 0.0, "kv": 0.0},
                                                       **kw), st, precision, 2)
    for f in list(L.FIELDS) + ["rain_mm"]:
        assert np.array_equal(a[f], b[f]), f
