"""The explicit diffusion oracle (oracle/nimbus_oracle.hpp diffusion_tendency,
restating src/numerics/diffusion/schemes/explicit/explicit.cpp) against the
reference's own scheme and against what the mathematics fixes.

- Golden vectors: tests/golden/diffusion.npz, written by the reference's
  ExplicitDiffusionScheme (make -C oracle ref; tests/golden/
  make_diffusion_golden.py) for six cases (C- and A-grid, closed and periodic
  sides, stretched layers, per-cell vertical diffusivity, horizontal-only and
  vertical-only): every tendency bitwise.
- Closed forms: a constant field has no tendency; with constant density and
  uniform spacing the discrete Laplacian of x^2, y^2 and z^2 is exactly 2, so
  interior tendencies are 2 K; a periodic sine is an eigenvector with
  eigenvalue -4 K sin^2(k d / 2) / d^2; the flux form conserves the control
  mass times the field on a closed domain; variable diffusivity equal to K_v
  everywhere gives the constant-K result.
"""
import os

import numpy as np
import pytest

from oracle import oracle_py as O

HERE = os.path.dirname(os.path.abspath(__file__))
GOLDEN = np.load(os.path.join(HERE, "golden", "diffusion.npz"))
LOCS = (("u", 1, "du"), ("v", 2, "dv"), ("w", 3, "dw"), ("theta", 0, "dtheta"), ("qv", 0, "dqv"))


def _cases():
    return sorted({k.split("/")[0] for k in GOLDEN.files if "/" in k})


@pytest.mark.parametrize("name", _cases())
def test_golden_vectors_bitwise(name):
    NX, NY, NZ, DX, DY = GOLDEN["grid"]
    kh, kv, cg, px, py, var, apply, _ = GOLDEN[f"{name}/settings"]
    rho = GOLDEN[f"{name}/in/rho"]
    dz = GOLDEN[f"{name}/in/dz"]
    kvf = GOLDEN[f"{name}/in/kv"] if var else None
    checked = 0
    for f, loc, o in LOCS:
        ref = GOLDEN[f"{name}/out/{o}"]
        if not (apply == 2 or (apply == 1 and loc > 0) or (apply == 0 and loc == 0)):
            assert not np.any(ref)
            continue
        got = O.diffusion_tendency(GOLDEN[f"{name}/in/{f}"], rho, dz, DX, DY, kh, kv, loc, bool(cg), bool(px),
                                   bool(py), kv=kvf)
        assert np.array_equal(got, ref), (name, f)
        assert np.any(ref != 0.0)
        checked += 1
    assert checked >= 2


def _grid(nx=12, ny=11, nz=10):
    return nx, ny, nz, 500.0, 400.0, np.full(nz, 250.0)


@pytest.mark.parametrize("loc", [0, 1, 2, 3])
@pytest.mark.parametrize("c_grid", [True, False])
def test_constant_field_has_no_tendency(loc, c_grid):
    nx, ny, nz, dx, dy, dz = _grid()
    rng = np.random.default_rng(1)
    rho = (1.0 + 0.1 * rng.random((nx, ny, nz))).astype(np.float64)
    f = np.full((nx, ny, nz), 3.5)
    t = O.diffusion_tendency(f, rho, dz, dx, dy, 70.0, 12.0, loc, c_grid, dtype=np.float64)
    assert np.all(t == 0.0)


@pytest.mark.parametrize("axis", [0, 1, 2])
def test_quadratic_profiles_give_twice_the_diffusivity(axis):
    nx, ny, nz, dx, dy, dz = _grid()
    d = (dx, dy, dz[0])[axis]
    idx = np.indices((nx, ny, nz))[axis].astype(np.float64) * d
    f = 0.003 * idx ** 2
    rho = np.ones((nx, ny, nz))
    kh, kv = (40.0, 0.0) if axis < 2 else (0.0, 25.0)
    t = O.diffusion_tendency(f, rho, dz, dx, dy, kh, kv, 0, True, dtype=np.float64)
    # interior cells: two cells away from the control-volume edges on every axis
    inner = t[3:nx - 3, 3:ny - 3, 2:nz - 2]
    assert np.allclose(inner, 2.0 * 0.003 * (kh if axis < 2 else kv), rtol=1e-9)


@pytest.mark.parametrize("loc,px,py", [(0, True, False), (1, True, False), (2, False, True), (0, False, True)])
def test_periodic_sine_is_an_eigenvector(loc, px, py):
    nx, ny, nz, dx, dy, dz = _grid(nx=18, ny=18, nz=6)
    axis = 0 if px else 1
    d = dx if axis == 0 else dy
    n_per = (nx - 2) if axis == 0 else (ny - 2)  # the periodic ring: columns / rows 1 .. n-2
    pos = np.indices((nx, ny, nz))[axis]
    kw = 2.0 * np.pi * 2.0 / n_per
    f = np.sin(kw * (pos - 1))
    rho = np.ones((nx, ny, nz))
    kh = 60.0
    t = O.diffusion_tendency(f, rho, dz, dx, dy, kh, 0.0, loc, True, periodic_x=px, periodic_y=py,
                             dtype=np.float64)
    lam = -4.0 * kh * np.sin(kw / 2.0) ** 2 / d ** 2
    # cells away from the closed direction's edges
    sl = (slice(1, nx - 1), slice(2, ny - 3), slice(0, nz - 1)) if px else (slice(2, nx - 3), slice(1, ny - 1),
                                                                           slice(0, nz - 1))
    assert np.allclose(t[sl], lam * f[sl], rtol=1e-9, atol=1e-12)


def test_flux_form_conserves_mass_weighted_field_on_a_closed_domain():
    nx, ny, nz, dx, dy, _ = _grid()
    dz = 200.0 * 1.2 ** np.arange(nz)
    rng = np.random.default_rng(5)
    rho = 1.0 + 0.2 * rng.random((nx, ny, nz))
    f = rng.standard_normal((nx, ny, nz))
    # vertical and horizontal parts separately: each is a sum of flux differences
    tv = O.diffusion_tendency(f, rho, dz, dx, dy, 0.0, 30.0, 0, True, dtype=np.float64)
    th = O.diffusion_tendency(f, rho, dz, dx, dy, 90.0, 0.0, 0, True, dtype=np.float64)
    mass = rho * dz[None, None, :]
    scale = np.sum(np.abs(mass * tv))
    assert abs(np.sum(mass * tv)) <= 1e-12 * scale
    assert abs(np.sum(rho * th)) <= 1e-12 * np.sum(np.abs(rho * th))


@pytest.mark.parametrize("loc", [0, 1, 2, 3])
def test_variable_diffusivity_equal_to_k_v_matches_constant_k(loc):
    nx, ny, nz, dx, dy, _ = _grid()
    dz = 200.0 * 1.1 ** np.arange(nz)
    rng = np.random.default_rng(6)
    rho = 1.0 + 0.2 * rng.random((nx, ny, nz))
    f = rng.standard_normal((nx, ny, nz))
    a = O.diffusion_tendency(f, rho, dz, dx, dy, 0.0, 17.0, loc, True, dtype=np.float64)
    b = O.diffusion_tendency(f, rho, dz, dx, dy, 0.0, 0.0, loc, True, kv=np.full(f.shape, 17.0), dtype=np.float64)
    assert np.allclose(a, b, rtol=1e-13, atol=1e-16)


def _step_case():
    from nimbus_amd import lib as L
    from tests.helpers import oracle_grid, small_case
    g, base, st = small_case(nx=18, ny=16, nz=10, cloudy=True, wind_noise=2.0, dtype=np.float64)
    return L, oracle_grid(g), g, st


@pytest.mark.parametrize("apply", ["scalars", "momentum", "all"])
def test_step_pass_commits_dt_times_the_pinned_tendencies(apply):
    """The large step's diffusion pass: every field it owns becomes field + dt
    tendency (the tendencies of the state before the pass, sides periodic
    only where both faces of an axis are), every other field is untouched."""
    L, og, g, st = _step_case()
    dt, kh, kv = 3.0, 2000.0, 50.0
    lat = (L.BC_PERIODIC, L.BC_PERIODIC, L.BC_RIGID_SLIP, L.BC_RIGID_SLIP)
    cfg = O.from_step_config(L.step_config(dt, 6, lateral=lat, diffusion={"apply": apply, "kh": kh, "kv": kv}))
    out = O.apply_diffusion(og, cfg, st)
    dz = np.full(g.nz, g.dz)
    owned = {"u": 1, "v": 2, "w": 3} if apply != "scalars" else {}
    if apply != "momentum":
        owned.update({"theta": 0, "qv": 0})
    for f in L.FIELDS:
        if f not in owned:
            assert np.array_equal(out[f], st[f]), f
            continue
        t = O.diffusion_tendency(st[f], st["rho"], dz, g.dx, g.dy, kh, kv, owned[f], True, True, False,
                                 dtype=np.float64)
        assert np.any(t != 0.0), f
        assert np.allclose(out[f], st[f] + dt * t, rtol=0, atol=1e-12 * max(1.0, float(np.max(np.abs(st[f]))))), f


def test_step_pass_is_off_without_positive_coefficients():
    L, og, g, st = _step_case()
    for dif in (None, {"apply": "all", "kh": 0.0, "kv": 0.0}):
        out = O.apply_diffusion(og, O.from_step_config(L.step_config(3.0, 6, diffusion=dif)), st)
        for f in L.FIELDS:
            assert np.array_equal(out[f], st[f]), f
