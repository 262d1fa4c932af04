"""Pins of the oracle's composed operators: the slow tendencies, the Heun RK2,
the boundary closure with KW78 radiation, the transport composition, the
microphysics subcycling and the large step, each against an answer that does
not come from the oracle itself:

- the reference's own known-answer tests, restated on the oracle's entry
  points (cited per test);
- closed forms of the mathematics (polynomial exactness of the fifth-order
  upwind derivative, the Heun update of a linear oscillator, the KW78 face
  update, a 1-D flux-form transport step built from the independently pinned
  WENO5 face function);
- invariants of the composition (mass and equilibrium over many steps).

Each test names the mistakes it catches; every one was checked on the CPU to
fail when that mistake is planted in the oracle.
"""
import numpy as np
import pytest

from nimbus_amd import lib as L
from oracle import oracle_py as O

G = 9.81


def grid(nx, ny, nz, dx, dy, dz):
    return O.grid(nx, ny, nz, dx, dy, dz)


def uniform_base(nz, rho0=1.0, p0=1.0e5, theta0=300.0, qv0=0.0, u0=0.0, v0=0.0):
    return {"rho0": np.full(nz, rho0), "p0": np.full(nz, p0), "theta0": np.full(nz, theta0),
            "qv0": np.full(nz, qv0), "u0": np.full(nz, u0), "v0": np.full(nz, v0)}


def state_from_base(shape, base, dtype=np.float64):
    nz = shape[2]
    st = {n: np.zeros(shape, dtype) for n in O.FIELDS}
    st["rho"][:] = np.broadcast_to(base["rho0"], shape)
    st["p"][:] = np.broadcast_to(base["p0"], shape)
    st["theta"][:] = np.broadcast_to(base["theta0"], shape)
    st["qv"][:] = np.broadcast_to(base["qv0"], shape)
    st["u"][:] = np.broadcast_to(base["u0"][:nz], shape)
    st["v"][:] = np.broadcast_to(base["v0"][:nz], shape)
    st["rain_mm"] = np.zeros(shape[:2])
    return st


# ------------------------------------------------------------------ slow tendencies

def test_slow_tendencies_hydrostatic_state_is_at_rest():
    """A resting state on its own hydrostatic base has no momentum tendency
    (tests/dynamics/test_cartesian_cgrid_slow_tendencies_gpu_parity.cpp:732-787,
    'Hydrostatic: zero tendencies'): du = dv = 0 exactly, dw only the FP32
    rounding of the stored density (g 2^-24). Catches a buoyancy that does not
    subtract the base state or uses full instead of face densities."""
    g = L.make_grid(14, 13, 12, 1000.0, 1000.0, 500.0)
    base = L.base_state(g)
    base = dict(base, u0=np.zeros(g.nz), v0=np.zeros(g.nz))
    base["qv0"] = base["qv0"].astype(np.float32).astype(np.float64)
    st = state_from_base((g.nx, g.ny, g.nz), base, np.float32)
    du, dv, dw = O.slow_tendencies(grid(g.nx, g.ny, g.nz, g.dx, g.dy, g.dz), base, O.config(6.0, 6), st)
    assert np.all(du == 0.0) and np.all(dv == 0.0)
    assert float(np.max(np.abs(dw))) <= G * 2.0 ** -23


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_slow_tendencies_coriolis_acts_on_the_perturbation_wind(dtype):
    """tests/dynamics/test_cartesian_cgrid_slow_tendencies_gpu_parity.cpp:946-1001:
    on a calm uniform base (rho0 = 1, p0 = 1e5) with a uniform u = 10 m/s,
    dv/dt = -f u' at an interior cell (rel 1e-5), du/dt = 0, and nothing when f
    = 0 or when u0 = u (the environmental wind feels no Coriolis force).
    Catches Coriolis on the full wind, a sign error or a swapped component."""
    nx = ny = nz = 12
    og = grid(nx, ny, nz, 500.0, 500.0, 250.0)
    base = uniform_base(nz)
    st = state_from_base((nx, ny, nz), base, dtype)
    st["u"][:] = 10.0
    f = 8.365e-5
    _, dv, _ = O.slow_tendencies(og, base, O.config(0.1, 6, coriolis_f=f), st, dtype)
    du, _, _ = O.slow_tendencies(og, base, O.config(0.1, 6, coriolis_f=f), st, dtype)
    assert float(dv[5, 5, 5]) == pytest.approx(-f * 10.0, rel=1e-5)
    assert abs(float(du[5, 5, 5])) <= 1e-12
    _, dv0, _ = O.slow_tendencies(og, base, O.config(0.1, 6, coriolis_f=0.0), st, dtype)
    assert abs(float(dv0[5, 5, 5])) <= 1e-12
    env = dict(base, u0=np.full(nz, 10.0))
    _, dve, _ = O.slow_tendencies(og, env, O.config(0.1, 6, coriolis_f=f), st, dtype)
    assert abs(float(dve[5, 5, 5])) <= 1e-12


# Polynomial winds of total degree <= 5 (x, y, z in units of 5 km): along any
# grid line each is a polynomial of degree <= 5, where the fifth-order upwind
# derivative (-2 m3 + 15 m2 - 60 m1 + 20 c + 30 p1 - 3 p2) / 60 h and its
# mirror are exact, whichever side the advecting velocity selects.
def _U(x, y, z):
    return 4.0 + 3.0 * x - 2.0 * x ** 2 + 0.5 * x ** 3 + 1.5 * y - 0.8 * y ** 2 + 2.0 * z - 1.2 * z ** 2 + 0.7 * x * y \
        + 0.3 * x ** 2 * z + 0.05 * x ** 5


def _V(x, y, z):
    return -3.0 + 1.0 * x + 0.6 * x ** 2 - 2.5 * y + 0.9 * y ** 3 - 1.0 * z + 0.4 * y * z + 0.04 * y ** 5


def _W(x, y, z):
    return 0.5 - 0.3 * x + 0.2 * y + 0.6 * z - 0.5 * z ** 2 + 0.1 * x * z - 0.02 * z ** 4


def _grad(fn, x, y, z, h=1e-4):
    return [(fn(x + h, y, z) - fn(x - h, y, z)) / (2 * h), (fn(x, y + h, z) - fn(x, y - h, z)) / (2 * h),
            (fn(x, y, z + h) - fn(x, y, z - h)) / (2 * h)]


def test_slow_tendencies_polynomial_winds_are_exact():
    """Winds sampled at their C-grid positions (u at x_{i+1/2}, v at y_{j+1/2},
    w at z_{k+1/2}; cartesian_cgrid_3d.cpp:1064-1356): at interior points the
    momentum tendencies equal -(a . grad) of the analytic field, the advecting
    velocities being the 4-point averages of the analytic neighbours at the
    staggered positions, plus Coriolis f (v_avg - v0) / -f (u_avg - u0) and, for
    w, the buoyancy of a 1 % density deficit, 1 g/kg of vapour and 0.7 g/kg of
    condensate. Catches a wrong averaging neighbour, a derivative along the
    wrong axis or with the wrong spacing, a Coriolis sign, a missing term."""
    nx, ny, nz = 18, 17, 16
    dx, dy, dz = 1000.0, 800.0, 400.0
    og = grid(nx, ny, nz, dx, dy, dz)
    S = 5000.0
    I, J, K = np.meshgrid(np.arange(nx), np.arange(ny), np.arange(nz), indexing="ij")
    xu, yu, zu = (I + 0.5) * dx / S, J * dy / S, (K + 0.5) * dz / S
    xv, yv, zv = I * dx / S, (J + 0.5) * dy / S, (K + 0.5) * dz / S
    xw, yw, zw = I * dx / S, J * dy / S, (K + 1.0) * dz / S
    rho0 = 1.1 * np.exp(-(np.arange(nz) + 0.5) * dz / 8000.0)
    qv0 = np.float32(0.01) * np.exp(-(np.arange(nz) + 0.5) * dz / 3000.0).astype(np.float32).astype(np.float64)
    base = uniform_base(nz)
    base.update(rho0=rho0, qv0=qv0, u0=np.linspace(1.0, 3.0, nz), v0=np.linspace(-2.0, 0.5, nz))
    st = state_from_base((nx, ny, nz), base)
    st["u"] = _U(xu, yu, zu)
    st["v"] = _V(xv, yv, zv)
    st["w"] = _W(xw, yw, zw)
    st["rho"] = np.broadcast_to(0.99 * rho0, (nx, ny, nz)).copy()
    st["qv"] = np.broadcast_to(qv0 + 1e-3, (nx, ny, nz)).copy()
    st["qc"][:] = 5e-4
    st["qr"][:] = 2e-4
    f = 1.0e-4
    du, dv, dw = O.slow_tendencies(og, base, O.config(6.0, 6, coriolis_f=f), st, np.float64)
    checked = 0
    for i in range(3, nx - 4):
        for j in range(3, ny - 4):
            for k in range(3, nz - 4):
                # u point (i+1/2, j, k)
                X, Y, Z = (i + 0.5) * dx / S, j * dy / S, (k + 0.5) * dz / S
                va = 0.25 * (_V(i * dx / S, (j - 0.5) * dy / S, Z) + _V(i * dx / S, (j + 0.5) * dy / S, Z)
                             + _V((i + 1) * dx / S, (j - 0.5) * dy / S, Z) + _V((i + 1) * dx / S, (j + 0.5) * dy / S, Z))
                wa = 0.25 * (_W(i * dx / S, Y, k * dz / S) + _W(i * dx / S, Y, (k + 1) * dz / S)
                             + _W((i + 1) * dx / S, Y, k * dz / S) + _W((i + 1) * dx / S, Y, (k + 1) * dz / S))
                gx, gy, gz = _grad(_U, X, Y, Z)
                exp_u = -(_U(X, Y, Z) * gx + va * gy + wa * gz) / S + f * (va - base["v0"][k])
                # v point (i, j+1/2, k)
                X, Y = i * dx / S, (j + 0.5) * dy / S
                ua = 0.25 * (_U((i - 0.5) * dx / S, j * dy / S, Z) + _U((i + 0.5) * dx / S, j * dy / S, Z)
                             + _U((i - 0.5) * dx / S, (j + 1) * dy / S, Z) + _U((i + 0.5) * dx / S, (j + 1) * dy / S, Z))
                wa = 0.25 * (_W(X, j * dy / S, k * dz / S) + _W(X, j * dy / S, (k + 1) * dz / S)
                             + _W(X, (j + 1) * dy / S, k * dz / S) + _W(X, (j + 1) * dy / S, (k + 1) * dz / S))
                gx, gy, gz = _grad(_V, X, Y, Z)
                exp_v = -(ua * gx + _V(X, Y, Z) * gy + wa * gz) / S - f * (ua - base["u0"][k])
                # w point (i, j, k+1/2)
                X, Y, Z = i * dx / S, j * dy / S, (k + 1.0) * dz / S
                ua = 0.25 * (_U((i - 0.5) * dx / S, Y, (k + 0.5) * dz / S) + _U((i + 0.5) * dx / S, Y, (k + 0.5) * dz / S)
                             + _U((i - 0.5) * dx / S, Y, (k + 1.5) * dz / S)
                             + _U((i + 0.5) * dx / S, Y, (k + 1.5) * dz / S))
                va = 0.25 * (_V(X, (j - 0.5) * dy / S, (k + 0.5) * dz / S) + _V(X, (j + 0.5) * dy / S, (k + 0.5) * dz / S)
                             + _V(X, (j - 0.5) * dy / S, (k + 1.5) * dz / S)
                             + _V(X, (j + 0.5) * dy / S, (k + 1.5) * dz / S))
                gx, gy, gz = _grad(_W, X, Y, Z)
                buoy = G * 0.01 / 0.99 + 0.608 * G * 1e-3 - G * 7e-4
                exp_w = -(ua * gx + va * gy + _W(X, Y, Z) * gz) / S + buoy
                for got, exp in ((du[i, j, k], exp_u), (dv[i, j, k], exp_v), (dw[i, j, k], exp_w)):
                    assert got == pytest.approx(exp, rel=1e-7, abs=1e-9)
                checked += 1
    assert checked > 500


def test_slow_rk2_is_heun_for_a_linear_oscillation():
    """Uniform perturbation winds (U', V') on a calm base with an f-plane: the
    slow tendencies are exactly (f V', -f U'), a linear oscillator, for which
    Heun's method (dynamics.cpp:5123-5292: k1 at u_n, k2 at the predictor
    u_n + dt k1, state restored, u_n + dt (k1 + k2)/2) gives the closed form
    U'_1 = U' + dt f V' - (dt f)^2 U'/2 and V'_1 = V' - dt f U' - (dt f)^2 V'/2.
    Catches Euler instead of Heun, stage 2 at u_n, a state left at the
    predictor, or a factor error in the average or the apply."""
    nx, ny, nz = 16, 15, 12
    og = grid(nx, ny, nz, 1000.0, 1000.0, 500.0)
    base = uniform_base(nz, u0=5.0, v0=2.0)
    st = state_from_base((nx, ny, nz), base)
    Up, Vp = 7.0, -3.0
    st["u"][:] = 5.0 + Up
    st["v"][:] = 2.0 + Vp
    f, dt = 2.0e-3, 6.0
    out = O.slow_rk2(og, base, O.config(dt, 6, coriolis_f=f), st, np.float64)
    a = dt * f
    exp_u = 5.0 + Up + a * Vp - 0.5 * a * a * Up
    exp_v = 2.0 + Vp - a * Up - 0.5 * a * a * Vp
    inner = (slice(4, nx - 4), slice(4, ny - 4), slice(0, nz - 1))
    assert np.allclose(out["u"][inner], exp_u, rtol=0, atol=1e-12)
    assert np.allclose(out["v"][inner], exp_v, rtol=0, atol=1e-12)
    assert np.all(out["w"][inner] == 0.0)
    # Euler would differ by (dt f)^2 / 2 * U' = 5e-4 m/s: well above the tolerance
    assert abs(exp_u - (5.0 + Up + a * Vp)) > 1e-4


# ------------------------------------------------------------------ boundary closure

C_STAR = 30.0


def _bc_state(rng, nx=10, ny=9, nz=6):
    base = uniform_base(nz)
    st = state_from_base((nx, ny, nz), base)
    for f in ("u", "v", "w"):
        st[f] = 5.0 * rng.standard_normal((nx, ny, nz))
    st["rho"] = 1.0 + 0.01 * rng.standard_normal((nx, ny, nz))
    st["p"] = 1e5 + 10.0 * rng.standard_normal((nx, ny, nz))
    return st


def test_kw78_radiative_faces_follow_the_radiation_equation(rng):
    """boundary_conditions_cartesian_cgrid.cpp:26-40 (and
    include/dynamics/radiative_face.hpp): the boundary-normal face velocity
    obeys du_b/dt = -(u_b +- c*) du/dn with c* = 30 m/s, one forward step of
    dt per substep from the stashed face value, held on inflow. East face
    u[NX-2]: u - dt (u + c*)(u - u[NX-3]) / dx when u + c* > 0; west face u[0]:
    u - dt (u - c*)(u[1] - u) / dx when u - c* < 0; north/south alike for v.
    Ghosts copy the closed face (east) or the interior (tangential, rho, p).
    Catches a wrong phase sign, a one-sided difference on the wrong side, the
    wrong spacing, or radiation on inflow."""
    nx, ny, nz = 10, 9, 6
    dx, dy, dt = 1000.0, 700.0, 0.75
    og = grid(nx, ny, nz, dx, dy, 500.0)
    st = _bc_state(rng, nx, ny, nz)
    # a few inflow faces: west u - c* >= 0, east u + c* <= 0
    st["u"][0, 2, :] = 35.0
    st["u"][nx - 2, 3, :] = -34.0
    st["v"][4, 0, :] = 31.0
    st["v"][5, ny - 2, :] = -32.0
    out = O.apply_boundaries(og, O.config(dt, 6, lateral=L.BC_RADIATIVE), st, dt, np.float64)
    u, v = st["u"], st["v"]
    for j in range(1, ny - 1):  # rows 0 and ny-1 are the south/north ghosts, closed after
        for k in range(nz - 1):
            ub = u[nx - 2, j, k]
            exp = ub - dt * (ub + C_STAR) * (ub - u[nx - 3, j, k]) / dx if ub + C_STAR > 0 else ub
            assert out["u"][nx - 2, j, k] == pytest.approx(exp, rel=2e-7, abs=1e-6)
            assert out["u"][nx - 1, j, k] == out["u"][nx - 2, j, k]
            ub = u[0, j, k]
            exp = ub - dt * (ub - C_STAR) * (u[1, j, k] - ub) / dx if ub - C_STAR < 0 else ub
            assert out["u"][0, j, k] == pytest.approx(exp, rel=2e-7, abs=1e-6)
    for i in range(1, nx - 1):
        for k in range(nz - 1):
            vb = v[i, ny - 2, k]
            exp = vb - dt * (vb + C_STAR) * (vb - v[i, ny - 3, k]) / dy if vb + C_STAR > 0 else vb
            assert out["v"][i, ny - 2, k] == pytest.approx(exp, rel=2e-7, abs=1e-6)
            vb = v[i, 0, k]
            exp = vb - dt * (vb - C_STAR) * (v[i, 1, k] - vb) / dy if vb - C_STAR < 0 else vb
            assert out["v"][i, 0, k] == pytest.approx(exp, rel=2e-7, abs=1e-6)
    # inflow faces held
    assert np.all(out["u"][0, 2, :-1] == 35.0) and np.all(out["u"][nx - 2, 3, :-1] == -34.0)
    # zero-gradient tangential wind and rho/p in the ghosts
    assert np.array_equal(out["rho"][0, 1:-1], st["rho"][1, 1:-1])
    assert np.array_equal(out["p"][nx - 1, 1:-1], st["p"][nx - 2, 1:-1])
    # rigid lid: w = 0 at the top face, u and v zero-gradient there
    assert np.all(out["w"][1:-1, 1:-1, nz - 1] == 0.0)
    # without time (the zero-length closure) no face radiates
    same = O.apply_boundaries(og, O.config(dt, 6, lateral=L.BC_RADIATIVE), st, 0.0, np.float64)
    assert np.array_equal(same["u"][nx - 2, 1:-1, :-1], st["u"][nx - 2, 1:-1, :-1])


@pytest.mark.parametrize("bc", [L.BC_RIGID, L.BC_RIGID_SLIP, L.BC_ZERO_GRADIENT, L.BC_PERIODIC])
def test_lateral_closures_follow_their_definitions(rng, bc):
    """boundary_conditions_cartesian_cgrid.cpp:26-44: rigid (no-slip) walls
    zero the normal face and the tangential ghost winds; rigid-slip zeros the
    normal face and copies the tangential winds; zero-gradient copies the
    adjacent interior values; periodic wraps the ghost from the opposite
    interior (ghost 0 <- NX-2, ghost NX-1 <- 1). rho and p copy or wrap."""
    nx, ny, nz = 10, 9, 6
    og = grid(nx, ny, nz, 1000.0, 1000.0, 500.0)
    st = _bc_state(rng, nx, ny, nz)
    out = O.apply_boundaries(og, O.config(1.0, 6, lateral=bc), st, 0.0, np.float64)
    j, k = 4, 2
    if bc == L.BC_PERIODIC:
        for f in ("u", "v", "w", "rho", "p"):
            assert out[f][0, j, k] == st[f][nx - 2, j, k]
            assert out[f][nx - 1, j, k] == st[f][1, j, k]
    else:
        if bc in (L.BC_RIGID, L.BC_RIGID_SLIP):
            assert out["u"][0, j, k] == 0.0 and out["u"][nx - 2, j, k] == 0.0
        else:
            assert out["u"][0, j, k] == st["u"][1, j, k]
            assert out["u"][nx - 2, j, k] == st["u"][nx - 3, j, k]
        if bc == L.BC_RIGID:
            assert out["v"][0, j, k] == 0.0 and out["w"][0, j, k] == 0.0
            assert out["v"][nx - 1, j, k] == 0.0 and out["w"][nx - 1, j, k] == 0.0
        else:
            assert out["v"][0, j, k] == st["v"][1, j, k] and out["w"][0, j, k] == st["w"][1, j, k]
        for f in ("rho", "p"):
            assert out[f][0, j, k] == st[f][1, j, k] and out[f][nx - 1, j, k] == st[f][nx - 2, j, k]


# ------------------------------------------------------------------ transport composition

def _weno_or_mc_face(q, i, n, mf):
    """Face value at face i (between cells i and i+1) of a 1-D line, as the
    reference selects it (advection_cartesian.cpp:1966-2009): the upwind WENO5
    reconstruction where its five cells exist, else the MC-limited slope of the
    upwind cell (zero slope at the line ends)."""
    def mc(a, b, c):
        r, l = c - b, b - a
        cands = (0.5 * (r + l), 2.0 * r, 2.0 * l)
        if all(x > 0 for x in cands):
            return min(cands)
        if all(x < 0 for x in cands):
            return max(cands)
        return 0.0
    if mf >= 0.0:
        if i >= 2 and i + 2 <= n - 1:
            return O.weno5_face(q[i - 2], q[i - 1], q[i], q[i + 1], q[i + 2])
        return q[i] + 0.5 * (mc(q[i - 1], q[i], q[i + 1]) if i >= 1 else 0.0)
    if i - 1 >= 0 and i + 3 <= n - 1:
        return O.weno5_face(q[i + 3], q[i + 2], q[i + 1], q[i], q[i - 1])
    return q[i + 1] - 0.5 * (mc(q[i], q[i + 1], q[i + 2]) if i + 1 <= n - 2 else 0.0)


@pytest.mark.parametrize("axis", [0, 1, 2])
@pytest.mark.parametrize("sign", [1.0, -1.0])
def test_transport_one_dimensional_step_is_the_flux_difference(rng, axis, sign):
    """With a mass flux along one axis only, the conservative update
    (advection_cartesian.cpp:2465-2650) of an interior cell is
    (rho_b q - dt (F_{+} - F_{-}) / h) / rho_e, F = mf * face value, the face
    value from the independently pinned WENO5 function (bit-exact with the
    reference) or the MC fallback near the line ends; faces include the wall
    faces; z has no flux through the ground or the rigid lid. Catches a face
    index shift, the wrong upwind side, a wrong fallback near the walls, a
    wrong spacing, or a density in the wrong place."""
    n_axis = 13
    shape = [7, 6, 5]
    shape[axis] = n_axis
    nx, ny, nz = shape
    h = (1000.0, 800.0, 300.0)
    og = grid(nx, ny, nz, *h)
    q = 300.0 + np.cumsum(rng.standard_normal(shape), axis=axis) + 0.5 * rng.standard_normal(shape)
    mf = [np.zeros(shape) for _ in range(3)]
    mf[axis][:] = sign * (0.8 + 0.4 * rng.random(shape))
    if axis == 2:
        mf[2][:, :, nz - 1] = 0.0  # the lid
    rho_b = 1.0 + 0.05 * rng.random(shape)
    rho_e = rho_b * (1.0 + 0.02 * rng.standard_normal(shape))
    dt = 2.0
    out = O.transport(og, O.config(dt, 6), q, mf[0], mf[1], mf[2], rho_b, rho_e, dt, False, np.float64)
    for i in range(1, nx - 1):
        for j in range(1, ny - 1):
            for k in range(nz):
                idx = [i, j, k]
                pos = idx[axis]
                line = [slice(None)] * 3
                for a in range(3):
                    if a != axis:
                        line[a] = idx[a]
                ql = q[tuple(line)]
                ml = mf[axis][tuple(line)]
                n = shape[axis]
                up = 0.0
                if pos <= n - 2 and not (axis == 2 and pos == n - 1):
                    up = ml[pos] * _weno_or_mc_face(ql, pos, n, ml[pos]) if (axis != 2 or pos < n - 1) else 0.0
                lo = ml[pos - 1] * _weno_or_mc_face(ql, pos - 1, n, ml[pos - 1]) if pos >= 1 else 0.0
                c = (i, j, k)
                exp = (rho_b[c] * q[c] - dt * (up - lo) / h[axis]) / rho_e[c]
                assert out[c] == pytest.approx(exp, rel=1e-13), (c, out[c], exp)


def test_transport_keeps_a_constant_under_divergent_flow():
    """tests/numerics/test_advection_cartesian.cpp:683-745 (Gate 4): a constant
    300 K field under a strongly divergent, sheared flow (u = 10 sin(2 pi i/nr),
    v = 8 cos(2 pi j/nth), w = 6 k/nz, rho = 1.1 exp(-z/8 km), 24 x 24 x 12,
    dx = 400 m, dz = 250 m, dt = 1.5 s) stays within 0.01 K over 60 steps
    when the end density carries the same mass fluxes. Catches an advective
    instead of a flux form, or an end density that does not match the fluxes."""
    nr, nth, nz = 24, 24, 12
    dx, dz, dt = 400.0, 250.0, 1.5
    og = grid(nr, nth, nz, dx, dx, dz)
    I, J, K = np.meshgrid(np.arange(nr), np.arange(nth), np.arange(nz), indexing="ij")
    u = 10.0 * np.sin(2 * np.pi * I / nr)
    v = 8.0 * np.cos(2 * np.pi * J / nth)
    w = 6.0 * K / nz
    rho = (1.1 * np.exp(-K * dz / 8000.0)).astype(np.float32).astype(np.float64)
    mfx = 0.5 * (rho + np.roll(rho, -1, 0)) * u
    mfy = 0.5 * (rho + np.roll(rho, -1, 1)) * v
    mfz = 0.5 * (rho + np.roll(rho, -1, 2)) * w
    mfx[-1] = 0.0
    mfy[:, -1] = 0.0
    mfz[:, :, -1] = 0.0
    div = np.zeros_like(rho)
    div[1:] += (mfx[1:] - mfx[:-1]) / dx
    div[0] += mfx[0] / dx
    div[:, 1:] += (mfy[:, 1:] - mfy[:, :-1]) / dx
    div[:, 0] += mfy[:, 0] / dx
    div[:, :, 1:] += (mfz[:, :, 1:] - mfz[:, :, :-1]) / dz
    div[:, :, 0] += mfz[:, :, 0] / dz
    rho_e = rho - dt * div
    theta = np.full((nr, nth, nz), 300.0, np.float32)
    for _ in range(60):
        theta = O.transport(og, O.config(dt, 6), theta, mfx, mfy, mfz, rho, rho_e, dt, False, np.float32)
    assert float(np.max(np.abs(theta.astype(np.float64) - 300.0))) < 0.01


def test_transport_translates_a_bump_at_the_flow_speed():
    """tests/numerics/test_advection_cartesian.cpp:749-790 (Gate 5): uniform
    rho = 1 and u = 30 m/s on 200 x 60 x 8 cells of 100 m, dt = 1.5 s (CFL
    0.45): a Gaussian bump (sigma 800 m) at x0 = 5 km moves its centroid to
    within 2 % of x0 + 30 m/s * 100.5 s after 67 steps, by more than 2.5 km,
    and keeps its mass to 1e-3. Catches a scheme that does not move the
    scalar, moves it at the wrong speed or does not conserve it."""
    nr, nth, nz = 200, 60, 8
    dx, dt = 100.0, 1.5
    og = grid(nr, nth, nz, dx, dx, dx)
    x = (np.arange(nr) - 0.5) * dx
    y = (np.arange(nth) - 0.5) * dx
    X, Y = np.meshgrid(x, y, indexing="ij")
    bump2 = np.exp(-((X - 5000.0) ** 2 + (Y - 3000.0) ** 2) / (2 * 800.0 ** 2))
    q = np.repeat(bump2[:, :, None], nz, axis=2).astype(np.float32)
    rho = np.ones((nr, nth, nz))
    mfx = np.full((nr, nth, nz), 30.0)
    mfx[-1] = 0.0
    mfx[0] = 0.0
    mfx[-2] = 0.0
    zero = np.zeros_like(rho)

    def mass(a):
        return float(np.sum(a[1:-1, 1:-1].astype(np.float64)))

    def cx(a):
        w = np.sum(a[1:-1, 1:-1].astype(np.float64), axis=(1, 2))
        return float(np.sum(w * x[1:-1]) / np.sum(w))

    m0, c0 = mass(q), cx(q)
    assert c0 == pytest.approx(5000.0, abs=2.0)
    for _ in range(67):
        q = O.transport(og, O.config(dt, 6), q, mfx, zero, zero, rho, rho, dt, True, np.float32)
    m1, c1 = mass(q), cx(q)
    expected = 5000.0 + 30.0 * 67 * dt
    assert c1 - c0 > 2500.0
    assert abs(c1 - expected) / expected < 0.02
    assert abs(m1 - m0) / m0 < 1e-3


def test_pd_limiter_scales_the_outflow_of_a_draining_cell():
    """The positive-definite donor limiter (advection_cartesian.cpp:2174-2240):
    a cell whose outgoing flux would remove more than rho_b q / dt has all its
    outgoing fluxes scaled by r = rho_b q / (dt * outgoing) (stored just below,
    advection_cartesian.hpp:101-108), so it ends at (almost exactly) zero and
    its neighbour receives r times the unlimited flux. Catches a limiter on
    the wrong cell, incoming fluxes scaled, or no limiter."""
    nx, ny, nz = 12, 5, 4
    dx, dt = 100.0, 10.0
    og = grid(nx, ny, nz, dx, dx, dx)
    q = np.zeros((nx, ny, nz))
    q[6] = 1e-3  # one wet plane, dry neighbours
    rho = np.ones((nx, ny, nz))
    mfx = np.zeros((nx, ny, nz))
    mfx[:nx - 1] = 100.0  # 100 kg m^-2 s^-1 eastward: far more than the wet cell holds
    zero = np.zeros_like(rho)
    out = O.transport(og, O.config(dt, 6), q, mfx, zero, zero, rho, rho, dt, True, np.float64)
    assert np.all(out >= 0.0)
    # the wet plane drains to (almost exactly) zero: its outflow limited to what it holds
    assert np.all(out[6, 1:-1] <= 1e-3 * 1e-5)
    # the dry plane upstream has nothing to give: its (WENO-positive) outflow is limited to zero
    assert np.all(out[5, 1:-1] == 0.0)
    # the next plane receives what the wet plane held, and cannot pass it on in the same step
    assert np.allclose(out[7, 1:-1], 1e-3, rtol=1e-5)
    assert np.all(out[8, 1:-1] == 0.0)
    assert float(np.sum(out[1:-1, 1:-1])) == pytest.approx(float(np.sum(q[1:-1, 1:-1])), rel=1e-12)


# ------------------------------------------------------------------ microphysics subcycling

def _column_case(nz=20, dz=100.0, qr=0.0, qc=0.0, rh=0.5):
    g = L.make_grid(6, 5, nz, 1000.0, 1000.0, dz)
    base = L.base_state(g)
    shape = (g.nx, g.ny, g.nz)
    st = state_from_base(shape, base)
    st["qv"][:] = np.broadcast_to(rh * base["qv0"], shape)
    st["qc"][:] = qc
    st["qr"][:] = qr
    return g, base, st


def _reference_limit(g, st, vt_max=20.0):
    """sedimentation_step_limit (microphysics_step.cpp:3501-3572) from the
    pinned fall-speed function: min over physical cells of 0.8 dz / speed,
    speed = min(Vt_max, kessler_rain(qv + qc + qr, rho) sqrt(1.2 / rho))."""
    best = np.inf
    for i in range(1, g.nx - 1):
        for j in range(1, g.ny - 1):
            for k in range(g.nz):
                r = float(st["rho"][i, j, k])
                qt = float(st["qv"][i, j, k]) + float(st["qc"][i, j, k]) + float(st["qr"][i, j, k])
                sp = min(vt_max, O.lib().oracle_kessler_rain_speed(qt, r) * np.sqrt(1.2 / r))
                if sp > 0:
                    best = min(best, 0.8 * g.dz / sp)
    return best


def test_microphysics_interval_under_the_limit_is_one_step():
    g, base, st = _column_case(dz=500.0)
    og = grid(g.nx, g.ny, g.nz, g.dx, g.dy, g.dz)
    lim = _reference_limit(g, st)
    assert lim > 6.0
    _, h = O.microphysics_step(og, O.config(6.0, 6), st, 6.0, np.float64, lengths=True)
    assert list(h) == [6.0]


def test_microphysics_subcycles_by_the_state_limit_with_a_remainder():
    """step_microphysics (microphysics_step.cpp:3619,3700-3750): an interval
    longer than the state's limit L runs steps of min(remaining, L), L
    recomputed after each step. A subsaturated state without condensate has
    no conversions, so L stays the same and the steps are L, L, ..., and the
    remainder. Catches equal-length subcycles, a missing remainder step, or a
    limit from Vt_max instead of the state."""
    g, base, st = _column_case(dz=100.0, rh=0.5)
    og = grid(g.nx, g.ny, g.nz, g.dx, g.dy, g.dz)
    lim = _reference_limit(g, st)
    assert 0.8 * g.dz / 20.0 < lim < 30.0  # the state sets it, not Vt_max
    dt = 3.5 * lim
    out, h = O.microphysics_step(og, O.config(dt, 6), st, dt, np.float64, lengths=True)
    assert len(h) == 4
    assert h[0] == lim and h[1] == lim and h[2] == lim
    assert h[3] == pytest.approx(dt - 3 * lim, rel=1e-12)
    assert float(np.sum(h)) == pytest.approx(dt, rel=1e-14)
    for f in ("qv", "theta"):
        assert np.array_equal(out[f], st[f])


def test_microphysics_subcycling_of_a_raining_column_closes_the_water_budget():
    """With rain the limit changes between steps (the rain falls out): the
    first step is the initial state's limit, the steps add up to the
    interval, and total water plus the surface rain is conserved between
    closed walls (precipitation flux sink * rho * dz per step,
    microphysics_step.cpp:1917-2024)."""
    g, base, st = _column_case(dz=100.0, rh=0.9, qr=3e-3, qc=1e-3)
    og = grid(g.nx, g.ny, g.nz, g.dx, g.dy, g.dz)
    lim = _reference_limit(g, st)
    dt = 2.6 * lim
    out, h = O.microphysics_step(og, O.config(dt, 6), st, dt, np.float64, lengths=True)
    assert h[0] == lim and len(h) >= 3
    assert float(np.sum(h)) == pytest.approx(dt, rel=1e-14)
    inner = (slice(1, -1), slice(1, -1))

    def water(s):
        rho = s["rho"][inner]
        q = s["qv"][inner] + s["qc"][inner] + s["qr"][inner]
        return float(np.sum(rho * q) * g.dz) + float(np.sum(s["rain_mm"][inner]))

    assert water(out) == pytest.approx(water(st), rel=1e-12)
    assert float(np.sum(out["rain_mm"])) > 0.0


# ------------------------------------------------------------------ large step

def _rest_case(nx=14, ny=13, nz=16, dz=500.0, wind=(0.0, 0.0)):
    g = L.make_grid(nx, ny, nz, 1000.0, 1000.0, dz)
    base = L.base_state(g)
    base = dict(base, u0=np.full(nz, wind[0]), v0=np.full(nz, wind[1]), qv0=np.zeros(nz))
    shape = (g.nx, g.ny, g.nz)
    st = state_from_base(shape, base)
    # p from the EOS of the stored rho and theta (the acoustic substep re-diagnoses it anyway)
    return g, base, st


@pytest.mark.parametrize("mode", [L.ACOUSTIC_EXPLICIT, L.ACOUSTIC_HEVI])
def test_large_steps_keep_hydrostatic_mass_over_60_seconds(mode):
    """tests/dynamics/test_cartesian_boundary_conditions.cpp:337-374 (Gate 2):
    a resting hydrostatic atmosphere with open (radiative) lateral boundaries
    keeps its total mass to 1e-4 over 60 simulated seconds. Catches a
    boundary closure or an operator that creates or drains mass at rest."""
    g, base, st = _rest_case()
    og = grid(g.nx, g.ny, g.nz, g.dx, g.dy, g.dz)
    cfg = O.config(6.0, 6, acoustic_mode=mode, lateral=L.BC_RADIATIVE)
    m0 = float(np.sum(st["rho"][1:-1, 1:-1]))
    out, _ = O.large_steps(og, base, cfg, st, 10, np.float64)
    m1 = float(np.sum(out["rho"][1:-1, 1:-1]))
    assert abs(m1 - m0) / m0 <= 1e-4
    assert float(np.max(np.abs(out["w"]))) < 1e-6


@pytest.mark.parametrize("mode", [L.ACOUSTIC_EXPLICIT, L.ACOUSTIC_HEVI])
def test_large_steps_with_a_uniform_wind_keep_equilibrium(mode):
    """tests/dynamics/test_cartesian_boundary_conditions.cpp:376-408 (Gate 3):
    the hydrostatic atmosphere under a uniform wind (10, 5) m/s through open
    boundaries keeps |w| <= 1e-2 m/s over 10 steps. Catches a boundary or
    advection operator that forces the balanced state."""
    g, base, st = _rest_case(wind=(10.0, 5.0))
    og = grid(g.nx, g.ny, g.nz, g.dx, g.dy, g.dz)
    cfg = O.config(3.0, 6, acoustic_mode=mode, lateral=L.BC_RADIATIVE)
    out, _ = O.large_steps(og, base, cfg, st, 10, np.float64)
    assert float(np.max(np.abs(out["w"]))) <= 1e-2
    assert np.allclose(out["u"][1:-1, 1:-1, :-1], 10.0, atol=1e-6)
