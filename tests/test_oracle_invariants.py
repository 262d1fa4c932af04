"""Closed forms and invariants pinning the oracle where no reference vector exists.

- WENO5 linear exactness / free-stream / ENO behaviour at a jump and UP5 polynomial
  exactness (tests/numerics/test_weno5_reconstruction.cpp and
  test_staggered_derivatives.cpp of the reference test the same properties).
- Saturation adjustment equilibrium and latent-energy invariants
  (tests/physics/test_microphysics.cpp:479-567 of the reference).
- The HE-VI vertically implicit acoustic step (no reference counterpart): Thomas
  against a dense solve, the rest state, and the amplification factor of the
  off-centred trapezoidal scheme for a vertical acoustic mode (closed form).
- Transport free-stream preservation and positivity; whole-step mass and water
  conservation between rigid lateral walls.
"""
import ctypes

import numpy as np
import pytest

from nimbus_amd import lib as L
from nimbus_amd import workloads as W
from oracle import oracle_py as O

DP = ctypes.POINTER(ctypes.c_double)


def dp(a):
    return a.ctypes.data_as(DP)


# ------------------------------------------------------------------ reconstruction / derivatives

@pytest.mark.parametrize("c,k", [(0.0, 1.0), (300.0, 0.25), (1e-3, -2e-4), (-5.0, 3.0)])
def test_weno5_linear_data_is_exact(c, k):
    vals = [c + k * o for o in (-2, -1, 0, 1, 2)]
    assert O.weno5_face(*vals) == pytest.approx(c + 0.5 * k, rel=1e-14, abs=1e-15)


def test_weno5_free_stream_and_eno():
    assert O.weno5_face(*([301.5] * 5)) == 301.5
    # across a jump between c and d the face value must stay in [c, d]
    for jump in (1.0, 1e-3, 50.0):
        f = O.weno5_face(0.0, 0.0, 0.0, jump, jump)
        assert 0.0 <= f <= jump
        f = O.weno5_face(0.0, 0.0, jump, jump, jump)
        assert abs(f - jump) < 1e-6 * jump


def test_upwind5_exact_for_quintic():
    L = O.lib()
    h = 0.37
    for deg in range(0, 6):
        for a in (1.0, -1.0):
            xs = np.array([-3, -2, -1, 0, 1, 2, 3], dtype=np.float64) * h
            fv = xs ** deg + 2.0 * xs
            exact = (deg * 0.0 ** (deg - 1) if deg >= 1 else 0.0) + 2.0
            got = L.oracle_upwind_fd5(*fv, a, 1.0 / h)
            assert got == pytest.approx(exact, abs=1e-9), (deg, a)


# ------------------------------------------------------------------ thermodynamics

@pytest.mark.parametrize("T,p,qv,qc", [(290.0, 90000.0, 0.02, 0.0), (280.0, 80000.0, 0.004, 0.002),
                                       (300.0, 100000.0, 0.03, 0.001), (250.0, 50000.0, 0.0005, 0.0)])
def test_saturation_adjustment_invariants(T, p, qv, qc):
    Ta, qva, qca = O.saturation_adjust(T, p, qv, qc)
    Lcp = 2.5e6 / 1004.0
    assert qva + qca == pytest.approx(qv + qc, rel=1e-14)
    assert Ta - T == pytest.approx(Lcp * (qca - qc), abs=2e-12 * T)
    qs = O.lib().oracle_qvs_water(Ta, p)
    if qca > 0.0:
        assert qva == pytest.approx(qs, rel=1e-12)      # saturated equilibrium
    else:
        assert qva <= qs * (1 + 1e-12)                 # subsaturated, all cloud gone


# ------------------------------------------------------------------ HE-VI

def test_thomas_matches_dense_solve(rng):
    L = O.lib()
    for n in (2, 5, 17, 60):
        a = -rng.uniform(0.0, 1.0, n)
        c = -rng.uniform(0.0, 1.0, n)
        a[0] = 0.0
        c[-1] = 0.0
        d = 1.0 + np.abs(a) + np.abs(c) + rng.uniform(0.0, 1.0, n)
        r = rng.standard_normal(n)
        x = np.zeros(n)
        assert L.oracle_thomas(n, dp(a), dp(d), dp(c), dp(r), dp(x)) == 0
        M = np.diag(d) + np.diag(a[1:], -1) + np.diag(c[:-1], 1)
        assert np.allclose(x, np.linalg.solve(M, r), rtol=1e-12, atol=1e-14)


def test_hevi_column_system_is_the_implicit_scheme(rng):
    """Coefficients of the w system equal the elimination of rho and p from the
    off-centred implicit equations, checked by assembling the 2n-1 coupled
    equations densely and solving them directly."""
    nz, dt, dz, alpha = 9, 0.7, 150.0, 0.55
    rb = rng.uniform(0.8, 1.2, nz)
    rb[-1] = 0.0
    C = rng.uniform(9e4, 1.2e5, nz + 1)
    G = 1e-3 * rng.standard_normal(nz)
    wo = rng.standard_normal(nz)
    e = np.zeros(nz + 1)
    for k in range(nz):
        ft = rb[k] * wo[k] if k < nz - 1 else 0.0
        fb = rb[k - 1] * wo[k - 1] if k > 0 else 0.0
        e[k] = -(1 - alpha) * dt / dz * (ft - fb)
    A, D, U, R = (np.zeros(nz) for _ in range(4))
    O.lib().oracle_hevi_column_system(nz, dt, dz, alpha, dp(rb), dp(C), dp(G), dp(e), dp(wo), dp(A), dp(D), dp(U),
                                      dp(R))
    n = nz - 1
    x = np.zeros(n)
    assert O.lib().oracle_thomas(n, dp(A), dp(D), dp(U), dp(R), dp(x)) == 0
    # dense: unknowns w_0..w_{n-1}, delta_0..delta_{nz-1}
    m = n + nz
    M = np.zeros((m, m))
    rhs = np.zeros(m)
    for k in range(n):  # momentum at face k: w + dt/(rb dz) alpha (C_{k+1} d_{k+1} - C_k d_k) = wo - dt/rb G
        M[k, k] = 1.0
        M[k, n + k + 1] += dt / (rb[k] * dz) * alpha * C[k + 1]
        M[k, n + k] -= dt / (rb[k] * dz) * alpha * C[k]
        rhs[k] = wo[k] - dt / rb[k] * G[k]
    for k in range(nz):  # continuity: d_k + alpha dt/dz (rb_k w_k - rb_{k-1} w_{k-1}) = e_k
        row = n + k
        M[row, n + k] = 1.0
        if k < n:
            M[row, k] += alpha * dt / dz * rb[k]
        if k > 0:
            M[row, k - 1] -= alpha * dt / dz * rb[k - 1]
        rhs[row] = e[k]
    sol = np.linalg.solve(M, rhs)
    assert np.allclose(x, sol[:n], rtol=1e-10, atol=1e-12)


def _uniform_column(nz, dz, rho0=1.0, theta0=300.0, n=3):
    g = O.grid(n, n, nz, 1000.0, 1000.0, dz)
    p0 = O.lib().oracle_eos_pressure(rho0, theta0)
    base = {"rho0": np.full(nz, rho0), "p0": np.full(nz, p0), "theta0": np.full(nz, theta0),
            "qv0": np.zeros(nz), "u0": np.zeros(nz), "v0": np.zeros(nz)}
    shape = (n, n, nz)
    st = {f: np.zeros(shape) for f in ("u", "v", "w", "qv", "qc", "qr")}
    st["rho"] = np.full(shape, rho0)
    st["theta"] = np.full(shape, theta0)
    st["p"] = np.full(shape, p0)
    return g, base, st, p0


def test_hevi_rest_state_is_exact():
    g, base, st, _ = _uniform_column(12, 200.0)
    out, _ = O.hevi_substep(g, base, st, 2.0, 0.1)
    assert np.max(np.abs(out["w"])) == 0.0
    assert np.max(np.abs(out["rho"] - st["rho"])) == 0.0


@pytest.mark.parametrize("beta,courant", [(0.0, 0.4), (0.1, 3.0), (0.2, 8.0)])
def test_hevi_vertical_mode_amplification_matches_theta_method(beta, courant):
    """A standing vertical acoustic mode evolves as Re(g^n) with the off-centred
    trapezoidal amplification g = (1 + i(1-a)W) / (1 - i a W), a = (1+beta)/2,
    W = omega dt, omega = (2c/dz) sin(m pi / (2 nz)); stable for any Courant number."""
    nz, dz, m, eps = 16, 200.0, 3, 1e-7
    g, base, st, p0 = _uniform_column(nz, dz, n=5)
    c = np.sqrt(1004.0 / 717.0 * p0 / 1.0)
    dt = courant * dz / c
    k = np.arange(nz)
    st["rho"][:, :, :] = 1.0 + eps * np.cos(m * np.pi * (k + 0.5) / nz)
    # the substep-start pressure is the EOS of the state, as the large step leaves it
    eos = np.vectorize(O.lib().oracle_eos_pressure)
    st["p"] = eos(st["rho"], st["theta"])
    n_steps = 4
    # horizontally uniform columns with zero-gradient lateral closure between
    # substeps: only the vertical acoustic mode is active
    cfg = O.config(n_steps * dt, n_steps, acoustic_mode=1, hevi_beta=beta, lateral=2)
    s, _ = O.acoustic_substeps_batched(g, base, cfg, st, n_steps, dt, 0.0, dtype=np.float64)
    alpha = 0.5 * (1.0 + beta)
    W = 2.0 * c / dz * np.sin(m * np.pi / (2 * nz)) * dt
    gfac = (1 + 1j * (1 - alpha) * W) / (1 - 1j * alpha * W)
    expect = eps * np.real(gfac ** n_steps) * np.cos(m * np.pi * (k + 0.5) / nz)
    got = s["rho"][2, 2, :] - 1.0
    assert np.max(np.abs(s["u"][1:-1, 1:-1])) == 0.0
    assert np.max(np.abs(got - expect)) < 1e-3 * eps, (got[:4], expect[:4])
    if beta > 0:
        assert abs(gfac) < 1.0



def _acoustic_pulse(nx=40, ny=5, nz=40, dx=1000.0, dz=500.0, amp=1e-3):
    g = L.make_grid(nx, ny, nz, dx, dx, dz)
    base = L.base_state(g)
    st = L.initial_state(g, L.ic_params(g, bubbles=[], noise_amp=0.0), np.float64)
    x = (np.arange(nx) - nx / 2)[:, None, None]
    z = (np.arange(nz) - nz / 3)[None, None, :]
    st["rho"] = st["rho"] * (1.0 + amp * np.exp(-(x ** 2 + z ** 2) / 8.0))
    return O.grid(nx, ny, nz, dx, dx, dz), base, st


def test_hevi_two_dimensional_acoustics_stay_bounded():
    """An x-z density pulse under HE-VI substeps alone (horizontal Courant 0.35,
    vertical 0.7, beta = 0.1) for 600 substeps: the off-centring acts on the
    substep-start pressure, so the coupled horizontal-vertical acoustic modes
    do not grow (off-centring against the horizontally updated p* would grow
    them by up to 24 % per substep)."""
    g, base, st = _acoustic_pulse()
    dts, n = 1.0, 100
    cfg = O.config(6 * n * dts, 6 * n, acoustic_mode=1, hevi_beta=0.1, lateral=2)
    peaks = []
    for _ in range(6):
        st, _ = O.acoustic_substeps_batched(g, base, cfg, st, n, dts, 0.0, dtype=np.float64)
        peaks.append(float(np.max(np.abs(st["w"]))))
    assert np.all(np.isfinite(peaks))
    assert max(peaks[1:]) < 1.5 * peaks[0], peaks


def test_hevi_large_steps_track_the_explicit_scheme():
    """Method change only: on a coarse warm-bubble case the HE-VI and explicit
    large steps give the same updraft to within 10 % after 20 steps (6 vs 10
    acoustic substeps per 6 s step)."""
    wl = W.Workload("t", 24, 24, 40, 1000.0, 500.0, 6.0, 6, 20)
    g = wl.grid
    og = O.grid(g.nx, g.ny, g.nz, g.dx, g.dy, g.dz)
    peaks = {}
    for mode in (0, 1):
        cfg = O.from_step_config(wl.config(acoustic_mode=mode))
        st, _ = O.large_steps(og, wl.base(), cfg, wl.initial_state(np.float64), 20, np.float64)
        peaks[mode] = float(np.max(st["w"]))
    assert 0.5 < peaks[0] < 5.0
    assert abs(peaks[1] - peaks[0]) < 0.1 * peaks[0], peaks


# ------------------------------------------------------------------ transport

def _flux_case(rng, nx=12, ny=10, nz=9, dt=5.0, dx=1000.0, dz=400.0):
    g = O.grid(nx, ny, nz, dx, dx, dz)
    shape = (nx, ny, nz)
    rb = 1.0 + 0.1 * rng.random(shape)
    u = 10.0 * rng.standard_normal(shape)
    v = 10.0 * rng.standard_normal(shape)
    w = 3.0 * rng.standard_normal(shape)
    mfx = 0.5 * (rb + np.roll(rb, -1, 0)) * u
    mfy = 0.5 * (rb + np.roll(rb, -1, 1)) * v
    mfz = 0.5 * (rb + np.roll(rb, -1, 2)) * w
    mfx[-1] = 0
    mfy[:, -1] = 0
    mfz[:, :, -1] = 0
    div = np.zeros(shape)
    div[1:] += (mfx[1:] - mfx[:-1]) / dx
    div[0] += mfx[0] / dx
    div[:, 1:] += (mfy[:, 1:] - mfy[:, :-1]) / dx
    div[:, 0] += mfy[:, 0] / dx
    div[:, :, 1:] += (mfz[:, :, 1:] - mfz[:, :, :-1]) / dz
    div[:, :, 0] += mfz[:, :, 0] / dz
    re = rb - dt * div
    return g, O.config(dt, 6), mfx, mfy, mfz, rb, re, dt


def test_transport_free_stream_preserved(rng):
    g, cfg, mfx, mfy, mfz, rb, re, dt = _flux_case(rng)
    q = np.full(rb.shape, 301.25)
    out = O.transport(g, cfg, q, mfx, mfy, mfz, rb, re, dt, False, np.float64)
    assert np.max(np.abs(out - 301.25)) < 1e-10


def test_transport_positive_definite_stays_nonnegative(rng):
    g, cfg, mfx, mfy, mfz, rb, re, dt = _flux_case(rng, dt=20.0)
    q = np.where(rng.random(rb.shape) < 0.3, 1e-3 * rng.random(rb.shape), 0.0)
    out = O.transport(g, cfg, q, mfx, mfy, mfz, rb, re, dt, True, np.float64)
    assert out.min() >= -1e-18
    noPD = O.transport(g, cfg, q, mfx, mfy, mfz, rb, re, dt, False, np.float64)
    assert noPD.min() < 0.0  # the limiter is doing work in this case


# ------------------------------------------------------------------ whole step conservation

def _periodic_case(dtype):
    from nimbus_amd import lib as L
    g = L.make_grid(16, 14, 12, 1000.0, 1000.0, 500.0)
    base = L.base_state(g)
    prm = L.ic_params(g, bubbles=[(7000.0, 6000.0, 1400.0, 3000.0, 1400.0, 2.0)], noise_amp=0.1, seed=2)
    st = L.initial_state(g, prm, np.float64)
    st["qc"][:, :, 3:6] = 1e-3
    st["qr"][:, :, 2:6] = 1e-3
    # rigid free-slip walls: no mass or water crosses the lateral boundary
    cfg = L.step_config(3.0, 6, lateral=L.BC_RIGID_SLIP)
    return g, base, st, cfg


def test_large_step_conserves_mass_and_water_closed_walls():
    g, base, st, cfg = _periodic_case(np.float64)
    og = O.grid(g.nx, g.ny, g.nz, g.dx, g.dy, g.dz)
    inner = (slice(1, -1), slice(1, -1), slice(None))
    m0 = st["rho"][inner].sum()
    w0 = (st["rho"] * (st["qv"] + st["qc"] + st["qr"]))[inner].sum() * g.dz
    out, _ = O.large_steps(og, base, O.from_step_config(cfg), st, 3, np.float64)
    m1 = out["rho"][inner].sum()
    w1 = (out["rho"] * (out["qv"] + out["qc"] + out["qr"]))[inner].sum() * g.dz + out["rain_mm"][1:-1, 1:-1].sum()
    assert abs(m1 - m0) < 1e-12 * m0
    assert abs(w1 - w0) < 1e-9 * w0


def test_fp32_oracle_tracks_fp64_oracle():
    g, base, st, cfg = _periodic_case(np.float64)
    og = O.grid(g.nx, g.ny, g.nz, g.dx, g.dy, g.dz)
    a, _ = O.large_steps(og, base, O.from_step_config(cfg), st, 3, np.float64)
    b, _ = O.large_steps(og, base, O.from_step_config(cfg), st, 3, np.float32)
    assert np.max(np.abs(a["w"] - b["w"])) < 1e-3
    assert np.max(np.abs(a["theta"] - b["theta"])) < 1e-3


def test_substep_promotion_follows_reference_rule():
    from nimbus_amd import lib as L
    from nimbus_amd import workloads as W
    for wl, mode, expect in ((W.C2, L.ACOUSTIC_EXPLICIT, 8), (W.C0, L.ACOUSTIC_EXPLICIT, 10),
                             (W.C0, L.ACOUSTIC_HEVI, 6), (W.C1, L.ACOUSTIC_EXPLICIT, 10), (W.C1, L.ACOUSTIC_HEVI, 6)):
        g = wl.grid
        base = wl.base()
        cfg = wl.config(acoustic_mode=mode)
        n_oracle = O.lib().oracle_resolve_substeps(ctypes.byref(O.grid(g.nx, g.ny, g.nz, g.dx, g.dy, g.dz)),
                                                   ctypes.byref(O.base_struct(base)),
                                                   ctypes.byref(O.from_step_config(cfg)))
        assert n_oracle == expect == L.resolve_substeps(g, base, cfg)


def test_oracle_result_does_not_depend_on_thread_count():
    # every OpenMP loop of the oracle writes disjoint cells from inputs no
    # iteration writes, so one thread and many give the same bits
    g, base, st, cfg = _periodic_case(np.float32)
    og = O.grid(g.nx, g.ny, g.nz, g.dx, g.dy, g.dz)
    c = O.from_step_config(cfg)
    many, _ = O.large_steps(og, base, c, st, 2, np.float32)
    prev = O.set_threads(1)
    try:
        one, _ = O.large_steps(og, base, c, st, 2, np.float32)
    finally:
        O.set_threads(prev)
    for k in many:
        assert np.array_equal(many[k], one[k]), k
