"""C-ABI library and initial-condition generator (CPU only: no device call).

- libnimbus_hip.so loads and exports every entry point declared in include/nimbus_hip.h.
- Argument validation refuses bad inputs without touching the device.
- The analytic sounding / base state / bubble / noise generator: hydrostatic
  balance, EOS consistency, humidity law and cap, bubble amplitude and
  position, noise bounds and determinism.
"""
import ctypes
import os
import re

import numpy as np
import pytest

from nimbus_amd import lib as L
from nimbus_amd import workloads as W

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    text = open(os.path.join(ROOT, "include", "nimbus_hip.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(nimbus_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_entry_point():
    lib = L.load_library()
    names = declared_functions()
    assert len(names) >= 30
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def test_member_bytes_and_validation_without_device():
    lib = L.load_library()
    g = L.make_grid(400, 400, 100, 250.0, 250.0, 200.0)
    base = L.base_state(g)
    desc = L.MemberDesc(g, L.step_config(1.5, 8), L.PRECISION_MIXED, L.base_struct(base))
    mixed = lib.nimbus_hip_member_bytes(ctypes.byref(desc))
    desc.precision = L.PRECISION_FP64
    fp64 = lib.nimbus_hip_member_bytes(ctypes.byref(desc))
    cells = 400 * 400 * 100
    assert 30 * 4 * cells < mixed < 45 * 4 * cells
    assert fp64 > 1.9 * mixed
    desc.grid = L.make_grid(3, 400, 100, 250.0, 250.0, 200.0)
    assert lib.nimbus_hip_member_bytes(ctypes.byref(desc)) == 0
    # a null context is refused before any device work
    assert lib.nimbus_hip_step(None, None, 0, 1) == -1


def test_base_state_hydrostatic_and_eos_consistent():
    g = L.make_grid(10, 10, 100, 250.0, 250.0, 200.0)
    b = L.base_state(g)
    from oracle import oracle_py as O
    z = (np.arange(g.nz) + 0.5) * g.dz
    # EOS consistency p0 = p00 (rho0 R theta0 / p00)^gamma
    pe = np.array([O.lib().oracle_eos_pressure(r, t) for r, t in zip(b["rho0"], b["theta0"])])
    assert np.allclose(pe, b["p0"], rtol=1e-12)
    # hydrostatic: dp/dz = -g rho_v with rho_v = p / (Rd T_v)
    kappa = 287.0 / 1004.0
    T = b["theta0"] * (b["p0"] / 1e5) ** kappa
    Tv = T * (1.0 + 0.608 * b["qv0"])
    rhov = b["p0"] / (287.0 * Tv)
    dpdz = np.diff(b["p0"]) / g.dz
    assert np.allclose(dpdz, -9.81 * 0.5 * (rhov[1:] + rhov[:-1]), rtol=2e-3)
    # sounding shape: theta 300 K at the ground rising to 343 K at 12 km, isothermal above
    assert abs(b["theta0"][0] - 300.0) < 0.5
    k12 = np.argmin(np.abs(z - 12000.0))
    assert abs(b["theta0"][k12] - 343.0) < 0.5
    assert np.all(np.diff(b["theta0"]) > 0)
    assert np.max(b["qv0"]) <= 0.014 + 1e-15
    # quarter-circle hodograph: 30 m/s of shear over 0-6 km, constant above
    du = np.hypot(b["u0"] - b["u0"][0], b["v0"] - b["v0"][0])
    k6 = np.argmin(np.abs(z - 6000.0))
    assert abs(du[k6] - 30.0) < 1.0
    assert np.allclose(b["u0"][z > 6100.0], b["u0"][-1])


def test_initial_state_bubble_noise_and_balance():
    wl = W.C0
    g = wl.grid
    b = wl.base()
    st = wl.initial_state(np.float64)
    th_p = st["theta"] - b["theta0"][None, None, :]
    # one +2 K bubble centred at 1.4 km above the domain centre plus +-0.1 K noise
    assert 1.9 < th_p.max() < 2.11
    i, j, k = np.unravel_index(np.argmax(th_p), th_p.shape)
    assert abs((i - 0.5) * g.dx - (g.nx - 2) * g.dx / 2) <= 1.5 * g.dx
    assert abs((k + 0.5) * g.dz - 1400.0) <= g.dz
    far = th_p[1:-1, 1:-1, -5:]
    assert far.min() >= -0.1 and far.max() < 0.1 and abs(far.mean()) < 0.01
    # pressure balanced: p = p0 and rho from the EOS of the perturbed theta
    assert np.allclose(st["p"], b["p0"][None, None, :])
    from oracle import oracle_py as O
    c = (i, j, k)
    assert O.lib().oracle_eos_pressure(st["rho"][c], st["theta"][c]) == pytest.approx(st["p"][c], rel=1e-12)
    # relative-humidity preserving moistening inside the bubble, capped
    assert st["qv"][c] >= b["qv0"][k] - 1e-15 and st["qv"].max() <= 0.014 + 1e-12
    assert np.all(st["qc"] == 0) and np.all(st["qr"] == 0) and np.all(st["w"] == 0)


def test_noise_hash_is_deterministic_and_uniform():
    lib = L.load_library()
    vals = np.array([lib.nimbus_ic_uniform(0, 0, i, j, k) for i in range(20) for j in range(10) for k in range(10)])
    again = np.array([lib.nimbus_ic_uniform(0, 0, i, j, k) for i in range(20) for j in range(10) for k in range(10)])
    assert np.array_equal(vals, again)
    assert vals.min() >= 0.0 and vals.max() < 1.0
    assert abs(vals.mean() - 0.5) < 0.03
    other = np.array([lib.nimbus_ic_uniform(0, 1, i, j, k) for i in range(20) for j in range(10) for k in range(10)])
    assert not np.array_equal(vals, other)


def test_ensemble_draws_cover_the_specified_ranges():
    members = W.ensemble_c3()
    assert len(members) == 32
    assert {m.nx for m in members} <= {48, 64, 96, 128, 160}
    assert {m.nz for m in members} <= {40, 60, 80}
    assert {m.dx for m in members} <= {500.0, 1000.0}
    for m in members:
        s = m.sounding
        assert 298.0 <= s["theta_sfc"] <= 302.0 and 0.011 <= s["qv_cap"] <= 0.016
        assert 10.0 <= s["shear"] <= 40.0 and 0.0 <= s["curvature"] <= 1.0
        assert m.steps * m.dt == pytest.approx(1800.0, abs=m.dt)
    assert [m.nx for m in W.ensemble_c3()] == [m.nx for m in members]  # seeded


def test_hevi_coefficient_scratch_only_for_the_sequential_sweep():
    # FP64 members of at most 129 levels solve their columns in the partitioned
    # form and carry no FP64 Thomas coefficient arrays; FP32-storage members
    # and deeper FP64 columns keep two FP64 arrays over the padded cells (rows
    # padded to 64 B of the storage type)
    def pair(nz, precision):
        g = L.make_grid(24, 20, nz, 1000.0, 1000.0, 20000.0 / nz)
        base = L.base_state(g)
        b = {m: L.member_bytes(g, L.step_config(3.0, 6, acoustic_mode=m), base, precision)
             for m in (L.ACOUSTIC_EXPLICIT, L.ACOUSTIC_HEVI)}
        return b[L.ACOUSTIC_HEVI] - b[L.ACOUSTIC_EXPLICIT]

    def scratch(nz, elem):
        per_row = 64 // elem
        cells = (24 + per_row - 1) // per_row * per_row * 20 * nz
        return 2 * ((cells * 8 + 255) // 256 * 256)

    assert pair(100, L.PRECISION_FP64) == 0
    assert pair(129, L.PRECISION_FP64) == 0
    assert pair(130, L.PRECISION_FP64) == scratch(130, 8)
    assert pair(100, L.PRECISION_MIXED) == scratch(100, 4)
