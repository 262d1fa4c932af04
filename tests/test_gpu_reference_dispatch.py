"""The reference-contract entry points (the reference's dispatch structs as
they are) through the C ABI on the device, against the oracle where it
computes the same operator, else against the contract's closed form:

- cartesian C-grid slow tendencies: du, dv, dw against the oracle's slow
  tendencies (reference Stencil tier), dp/dt = -u.grad p centred, drho = 0;
- Kessler pointwise + sedimentation composed as the reference's caller does
  (host saturation adjustment booked as rates, the temperature tendency
  converted to theta, kessler.cpp:150-339) against the oracle's Kessler
  tendencies (Microphysics tiers);
- conservative transport of four scalars with supplied mass fluxes, and in
  the standalone mode, against the oracle's transport;
- slow apply, state relaxation and microphysics apply against their
  definitions, including the non-finite, limiter and clamp paths.
"""
import ctypes

import numpy as np
import pytest

from nimbus_amd import lib as L
from oracle import oracle_py as O
from tests.helpers import oracle_grid, small_case

pytestmark = pytest.mark.gpu

F32 = np.float32


def fp(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_float)) if a is not None else None


def within(got, ref, rel, ab):
    got = np.asarray(got, np.float64)
    ref = np.asarray(ref, np.float64)
    bad = np.abs(got - ref) > ab + rel * np.abs(ref)
    return int(bad.sum()), float(np.max(np.abs(got - ref)))


def call(ctx, name, d):
    L.check(getattr(ctx.lib, name)(ctx.handle, ctypes.byref(d)), name)


def test_cgrid_slow_tendencies_match_oracle(hip_ctx):
    g, base, st = small_case(nx=16, ny=14, nz=12, cloudy=True, wind_noise=2.0)
    base = dict(base, u0=np.linspace(2.0, 8.0, g.nz), v0=np.linspace(-1.0, 3.0, g.nz))
    f = 1.0e-4
    ref_du, ref_dv, ref_dw = O.slow_tendencies(oracle_grid(g), base, O.config(3.0, 6, coriolis_f=f), st)
    a = {k: np.ascontiguousarray(st[k], F32) for k in ("u", "v", "w", "rho", "p", "qv", "qc", "qr")}
    load = (a["qc"].astype(np.float64) + a["qr"]).astype(F32)
    prof = {k: np.ascontiguousarray(base[k], F32) for k in ("rho0", "qv0", "u0", "v0")}
    out = {k: np.full(a["u"].shape, np.nan, F32) for k in ("du", "dv", "dw", "drho", "dp", "adu", "adv", "adw")}
    d = L.CgridSlowTendenciesDispatch(
        fp(a["u"]), fp(a["v"]), fp(a["w"]), fp(a["rho"]), fp(a["p"]), fp(prof["rho0"]), g.nz, fp(prof["qv0"]), g.nz,
        fp(prof["u0"]), fp(prof["v0"]), g.nz, fp(a["qv"]), fp(load),
        fp(out["du"]), fp(out["dv"]), fp(out["dw"]), fp(out["drho"]), fp(out["dp"]),
        g.nx, g.ny, g.nz, g.dx, g.dy, g.dz, 9.81, f, 1, 1, 0, fp(out["adu"]), fp(out["adv"]), fp(out["adw"]),
        0.0, 0.0, 0.0)
    n0 = hip_ctx.lib.nimbus_hip_successful_dispatch_count(hip_ctx.handle, b"cartesian_cgrid_slow_tendencies")
    call(hip_ctx, "nimbus_hip_dispatch_cartesian_cgrid_slow_tendencies", d)
    assert hip_ctx.lib.nimbus_hip_successful_dispatch_count(hip_ctx.handle,
                                                            b"cartesian_cgrid_slow_tendencies") == n0 + 1
    for name, got, ref in (("du", out["du"], ref_du), ("dv", out["dv"], ref_dv), ("dw", out["dw"], ref_dw)):
        nbad, err = within(got, ref, 1e-4, 5e-5)  # Stencil tier
        print(f"cgrid slow {name}: max diff {err:.3e}")
        assert nbad == 0, name
    assert np.all(out["drho"] == 0.0)
    # dp/dt = -u.grad p at cell centres, centred differences (cartesian_cgrid_3d.cpp:1358-1399)
    u, v, w, p = (a[k].astype(np.float64) for k in ("u", "v", "w", "p"))
    exp = np.zeros_like(p)
    sl = (slice(1, -1), slice(1, -1), slice(1, -1))
    uc = 0.5 * (u[1:-1, 1:-1, 1:-1] + u[:-2, 1:-1, 1:-1])
    vc = 0.5 * (v[1:-1, 1:-1, 1:-1] + v[1:-1, :-2, 1:-1])
    wc = 0.5 * (w[1:-1, 1:-1, 1:-1] + w[1:-1, 1:-1, :-2])
    exp[sl] = -(uc * (p[2:, 1:-1, 1:-1] - p[:-2, 1:-1, 1:-1]) / (2 * g.dx)
                + vc * (p[1:-1, 2:, 1:-1] - p[1:-1, :-2, 1:-1]) / (2 * g.dy)
                + wc * (p[1:-1, 1:-1, 2:] - p[1:-1, 1:-1, :-2]) / (2 * g.dz))
    nbad, err = within(out["dp"], exp, 1e-6, 1e-6)
    assert nbad == 0, err
    # the advection components are the tendencies without Coriolis and buoyancy
    v4 = 0.25 * (v[1:-1, :-2] + v[1:-1, 1:-1] + v[2:, :-2] + v[2:, 1:-1])
    cor = f * (v4 - base["v0"][None, None, :])
    diff = out["du"][1:-1, 1:-1, :-1].astype(np.float64) - out["adu"][1:-1, 1:-1, :-1] - cor[:, :, :-1]
    assert float(np.max(np.abs(diff))) < 1e-5


def test_slow_apply_follows_its_definition(hip_ctx, rng):
    shape = (9, 8, 7)
    u, v, w, p = (rng.standard_normal(shape).astype(F32) * s for s in (10, 10, 3, 100))
    du, dv, dw, dp, dup, dvp = (rng.standard_normal(shape).astype(F32) for _ in range(6))
    du[1, 2, 3] = np.nan
    dv[4, 4, 4] = np.inf
    dt = 6.0
    seen = {}

    @L.FourFieldObserver
    def observer(_ctx, f0, f1, f2, f3):
        seen["u"] = np.ctypeslib.as_array(f0, shape=(u.size,)).copy().reshape(shape)
        seen["old_u"] = u_dev.copy()

    u_dev, v_dev, w_dev, p_dev = (x.copy() for x in (u, v, w, p))
    d = L.SlowApplyDispatch(fp(u_dev), fp(v_dev), fp(w_dev), fp(p_dev), fp(du), fp(dv), fp(dw), fp(dp), fp(dup),
                            fp(dvp), *shape, dt, observer, None, 1)
    call(hip_ctx, "nimbus_hip_dispatch_slow_apply", d)

    def expect(x, t):
        t = np.where(np.isfinite(t), t, 0.0)
        return (x.astype(np.float64) + t * dt).astype(F32)

    assert np.array_equal(u_dev, expect(u, du.astype(np.float64) + dup))
    assert np.array_equal(v_dev, expect(v, dv.astype(np.float64) + dvp))
    assert np.array_equal(w_dev, expect(w, dw.astype(np.float64)))
    assert np.array_equal(p_dev, expect(p, dp.astype(np.float64)))
    assert u_dev[1, 2, 3] == u[1, 2, 3]  # the non-finite tendency integrated as zero
    assert np.array_equal(seen["u"], u_dev) and np.array_equal(seen["old_u"], u)


def test_state_relaxation_follows_its_definition(hip_ctx, rng):
    nx, ny, nz = 12, 10, 9
    shape = (nx, ny, nz)
    u, v, w, th = (rng.standard_normal(shape).astype(F32) * s + o for s, o in ((5, 3), (5, -1), (2, 0), (3, 300)))
    prof = np.zeros(7 * nz, F32)
    prof[0:nz] = np.linspace(0.0, 0.05, nz)           # upper alpha
    prof[nz:2 * nz] = np.r_[np.zeros(nz - 3), [0.05, 0.1, 0.2]]  # w-only alpha
    prof[2 * nz:3 * nz] = 4.0                          # u_ref
    prof[3 * nz:4 * nz] = -2.0                         # v_ref
    prof[4 * nz:5 * nz] = np.linspace(300.0, 330.0, nz)  # theta0
    prof[5 * nz:5 * nz + 3] = [0.05, 0.03, 0.01]       # lateral alpha by distance
    prof[6 * nz:7 * nz] = np.linspace(0.0, 0.06, nz)   # upper alpha at w faces
    dt = 3.0
    a = [x.copy() for x in (u, v, w, th)]
    d = L.StateRelaxationDispatch(fp(a[0]), fp(a[1]), fp(a[2]), fp(a[3]), fp(prof), 7 * nz, nx, ny, nz, 3, 1, 1, 1,
                                  dt, L.FourFieldObserver(), None)
    call(hip_ctx, "nimbus_hip_dispatch_state_relaxation_cartesian", d)
    P = prof.astype(np.float64).reshape(7, nz)
    uu, vv, ww, tt = (x.astype(np.float64) for x in (u, v, w, th))
    fa = -np.expm1(-P[0] * dt)
    fw = -np.expm1(-P[6] * dt)
    uu = uu + fa * (P[2] - uu)
    vv = vv + fa * (P[3] - vv)
    ww = ww + fw * (0.0 - ww)
    tt = tt + fa * (P[4] - tt)
    ww = ww + (-np.expm1(-P[1] * dt)) * (0.0 - ww)
    I, J = np.meshgrid(np.arange(nx), np.arange(ny), indexing="ij")
    dist = np.minimum(np.minimum(I, nx - 1 - I), np.minimum(J, ny - 1 - J))
    fl = np.where(dist < 3, -np.expm1(-P[5][np.minimum(dist, nz - 1)] * dt), 0.0)[:, :, None]
    uu = uu + fl * (P[2] - uu)
    vv = vv + fl * (P[3] - vv)
    ww = ww + fl * (0.0 - ww)
    tt = tt + fl * (P[4] - tt)
    for got, exp in zip(a, (uu, vv, ww, tt)):
        assert within(got, exp, 2e-7, 1e-6)[0] == 0


def test_microphysics_apply_follows_its_definition(hip_ctx, rng):
    nx, ny, nz = 10, 9, 8
    shape = (nx, ny, nz)
    th = (300.0 + rng.standard_normal(shape)).astype(F32)
    qs = [np.abs(rng.standard_normal(shape)).astype(F32) * 1e-3 for _ in range(7)]
    dth = rng.standard_normal(shape).astype(F32)
    dth[3, 3, 3] = 100.0     # a limited step
    dth[4, 4, 4] = np.nan    # a non-finite tendency
    th[5, 5, 5] = 699.9
    dth[5, 5, 5] = 1.0       # clamped at theta_max
    dq = [rng.standard_normal(shape).astype(F32) * 1e-4 for _ in range(7)]
    dq[0][2, 2, 2] = -1.0    # drives qv negative: clamped to 0
    stats = np.zeros(10 + nz + nx, np.uint32)
    fields = [th.copy()] + [q.copy() for q in qs]
    zero = np.zeros(shape, F32)
    dt = 2.0
    args = [fp(x) for x in fields] + [fp(dth)] + [fp(x) for x in dq] + [fp(zero), fp(zero), fp(zero)]
    d = L.MicrophysicsApplyDispatch(*args, nx, ny, nz, 1, nx - 1, 1, ny - 1, dt, 300.0, 200.0, 700.0, 50.0, 0.05,
                                    0.05, stats.ctypes.data_as(ctypes.POINTER(ctypes.c_uint)), stats.size)
    call(hip_ctx, "nimbus_hip_dispatch_microphysics_apply", d)
    inner = (slice(1, nx - 1), slice(1, ny - 1))
    t = np.where(np.isfinite(dth), dth, 0.0).astype(np.float64) * dt
    lim = np.clip(t, -50.0, 50.0)
    exp_th = np.clip(th.astype(np.float64) + lim, 200.0, 700.0)
    assert within(fields[0][inner], exp_th[inner], 1e-7, 1e-5)[0] == 0
    for q in range(7):
        top = 0.05
        exp = np.clip(qs[q].astype(np.float64) + dq[q].astype(np.float64) * dt, 0.0, top)
        assert within(fields[q + 1][inner], exp[inner], 1e-6, 1e-9)[0] == 0
        assert np.array_equal(fields[q + 1][0], qs[q][0])  # outside the column range untouched
    assert fields[1][2, 2, 2] == 0.0
    assert stats[0] == 1 and stats[2] >= 1 and stats[3] >= 1 and stats[4] == 1
    assert stats[10 + 5] >= 1 and stats[10 + nz + 5] >= 1


def _theta_tendency(dT, p):
    pp = np.clip(p.astype(np.float64), 100.0, 200000.0)
    return dT * (1.0e5 / pp) ** (287.0 / 1004.0)


def test_kessler_pointwise_and_sedimentation_compose_to_the_oracle(hip_ctx):
    g, base, st = small_case(cloudy=True)
    dt = 3.0
    a = {k: np.ascontiguousarray(st[k], F32) for k in ("p", "theta", "rho", "qv", "qc", "qr")}
    shape = a["p"].shape
    # the reference caller's host part (kessler.cpp:150-210): theta -> T, saturation adjustment
    T = np.zeros(shape, F32)
    qv_adj = np.zeros(shape, F32)
    qc_adj = np.zeros(shape, F32)
    lib = O.lib()
    for idx in np.ndindex(*shape):
        T0 = np.float32(lib.oracle_theta_to_temperature(float(a["theta"][idx]), float(a["p"][idx])))
        Ta, qv1, qc1 = O.saturation_adjust(float(T0), float(a["p"][idx]), float(a["qv"][idx]), float(a["qc"][idx]))
        T[idx], qv_adj[idx], qc_adj[idx] = Ta, qv1, qc1
    outs = {k: np.full(shape, np.nan, F32) for k in ("dth", "dqv", "dqc", "dqr", "dqg", "dqh")}
    lo = (1, g.nx - 1, 1, g.ny - 1, 0, g.nz)
    d = L.KesslerPointwiseDispatch(fp(T), fp(a["p"]), fp(qv_adj), fp(qc_adj), fp(a["qr"]), fp(a["rho"]), None,
                                   *[fp(outs[k]) for k in ("dth", "dqv", "dqc", "dqr", "dqg", "dqh")],
                                   g.nx, g.ny, g.nz, *lo, 1e-3, 1e-3, 2.2, 1.0, 2.5e6 / 1004.0, 273.15, dt)
    call(hip_ctx, "nimbus_hip_dispatch_kessler_pointwise", d)
    assert np.all(outs["dqg"] == 0.0) and np.all(outs["dqh"] == 0.0)
    assert np.all(outs["dqr"][0] == 0.0)  # outside the i range
    # the adjustment booked as rates on the host (kessler.cpp:232-255)
    rate = (qc_adj.astype(np.float64) - a["qc"]) / dt
    phys = np.zeros(shape, bool)
    phys[1:-1, 1:-1, :] = True
    for k, sgn in (("dqc", 1.0), ("dqv", -1.0), ("dth", 2.5e6 / 1004.0)):
        outs[k] = np.where(phys & (rate != 0.0), (outs[k] + (sgn * rate).astype(F32)).astype(F32), outs[k])
    s = L.KesslerSedimentationDispatch(fp(a["qr"]), None, None, fp(a["rho"]), fp(outs["dqr"]), None, None,
                                       g.nx, g.ny, g.nz, *lo, g.dz, 20.0, np.finfo(F32).max, dt, None)
    call(hip_ctx, "nimbus_hip_dispatch_kessler_sedimentation", s)
    dth = _theta_tendency(outs["dth"].astype(np.float64), a["p"]).astype(F32)
    ref = O.kessler_tendencies(oracle_grid(g), O.config(dt, 6), a["p"], a["theta"], a["rho"], a["qv"], a["qc"],
                               a["qr"], dt)
    tiers = {"dth": (1e-4, 6e-4), "dqv": (1e-4, 1e-7), "dqc": (1e-4, 1e-7), "dqr": (1e-4, 1e-7)}
    for q, name in enumerate(("dth", "dqv", "dqc", "dqr")):
        got = dth if name == "dth" else outs[name]
        nbad, err = within(got[1:-1, 1:-1], ref[q][1:-1, 1:-1], *tiers[name])
        print(f"kessler contract {name}: max diff {err:.3e}, outside {nbad}")
        assert nbad == 0, name


def _scalars(st):
    return [np.ascontiguousarray(st[k], F32) for k in ("theta", "qv", "qc", "qr")]


@pytest.mark.parametrize("supplied", [True, False])
def test_conservative_transport_matches_oracle(hip_ctx, rng, supplied):
    from tests.test_gpu_parity import _transport_inputs
    g, base, st = small_case(nx=16, ny=14, nz=12, cloudy=True)
    dt = 4.0
    (mfx, mfy, mfz), rho_b, rho_e = _transport_inputs(g, base, st, dt, rng)
    shape = rho_b.shape
    if supplied:
        u = v = w = None
    else:
        # standalone: fluxes from the winds and the live density; the end density carries them
        u = (st["u"] + 5.0 * rng.standard_normal(shape)).astype(F32)
        v = (st["v"] + 5.0 * rng.standard_normal(shape)).astype(F32)
        w = (2.0 * rng.standard_normal(shape)).astype(F32)
        w[:, :, -1] = 0.0
        rb = rho_b.astype(np.float64)
        mfx = (0.5 * (rb + np.roll(rb, -1, 0)) * u).astype(F32)
        mfy = (0.5 * (rb + np.roll(rb, -1, 1)) * v).astype(F32)
        mfz = (0.5 * (rb + np.roll(rb, -1, 2)) * w).astype(F32)
        mfx[-1] = 0.0
        mfy[:, -1] = 0.0
        mfz[:, :, -1] = 0.0
        div = np.zeros(shape)
        m64 = [x.astype(np.float64) for x in (mfx, mfy, mfz)]
        div += (m64[0] - np.pad(m64[0], ((1, 0), (0, 0), (0, 0)))[:-1]) / g.dx
        div += (m64[1] - np.pad(m64[1], ((0, 0), (1, 0), (0, 0)))[:, :-1]) / g.dy
        div += (m64[2] - np.pad(m64[2], ((0, 0), (0, 0), (1, 0)))[:, :, :-1]) / g.dz
        rho_e = (rb - dt * div).astype(F32)
    qs = _scalars(st)
    outs = [np.zeros_like(q) for q in qs]
    pd = (ctypes.c_int * 4)(0, 1, 1, 1)
    qin = (ctypes.POINTER(ctypes.c_float) * 4)(*[fp(q) for q in qs])
    qout = (ctypes.POINTER(ctypes.c_float) * 4)(*[fp(o) for o in outs])
    kap = np.zeros(4, F32)
    d = L.ConservativeTransportDispatch(fp(u), fp(v), fp(w), fp(rho_e if supplied else rho_b), qin, qout, fp(kap), pd,
                                        4, g.nx, g.ny, g.nz, 1.0 / g.dx, 1.0 / g.dy, 1.0 / g.dz, 0.0, 0.0, 0.0, dt,
                                        0.05, fp(rho_b) if supplied else None, fp(mfx) if supplied else None,
                                        fp(mfy) if supplied else None, fp(mfz) if supplied else None, int(supplied),
                                        0.0, 0.0, 1, 0)
    call(hip_ctx, "nimbus_hip_dispatch_advect_scalars_conservative_cartesian_cgrid", d)
    cfg = O.config(dt, 6)
    for q, (qq, name) in enumerate(zip(qs, ("theta", "qv", "qc", "qr"))):
        ref = O.transport(oracle_grid(g), cfg, qq, mfx, mfy, mfz, rho_b, rho_e, dt, q > 0)
        nbad, err = within(outs[q], ref, 1e-4, 5e-5 if q == 0 else 1e-9)
        print(f"transport contract {name} supplied={supplied}: max diff {err:.3e}")
        assert nbad == 0, name


def test_reference_contract_refusals_leave_host_fields_untouched(hip_ctx):
    shape = (8, 7, 6)
    q = np.ones(shape, F32)
    out = np.full(shape, 7.0, F32)
    kap = np.array([0.5], F32)  # chained point diffusion: refused, as the reference's device backends do
    qin = (ctypes.POINTER(ctypes.c_float) * 1)(fp(q))
    qout = (ctypes.POINTER(ctypes.c_float) * 1)(fp(out))
    pd = (ctypes.c_int * 1)(0)
    d = L.ConservativeTransportDispatch(None, None, None, fp(q), qin, qout, fp(kap), pd, 1, *shape, 1e-3, 1e-3, 1e-3,
                                        0.0, 0.0, 0.0, 1.0, 0.05, fp(q), fp(q), fp(q), fp(q), 1, 0.0, 0.0, 1, 0)
    rc = hip_ctx.lib.nimbus_hip_dispatch_advect_scalars_conservative_cartesian_cgrid(hip_ctx.handle, ctypes.byref(d))
    assert rc == -1 and np.all(out == 7.0)
