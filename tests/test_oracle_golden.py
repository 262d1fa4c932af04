"""The CPU oracle against golden vectors produced by the reference's own code.

The fixtures under tests/golden/ come from oracle/_ref/ref_harness, which
compiles the reference's translation units and headers where they lie
(see tests/golden/make_golden.py). FP64 functions that restate the same
operation sequence must agree bit for bit; FP32-storage operators agree bit
for bit unless a libm call differs in the last place, which the tolerances
below (a few FP32 ulps) allow for.
"""
import os

import numpy as np
import pytest

from oracle import oracle_py as O

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def load(name):
    return np.load(os.path.join(GOLDEN, name), allow_pickle=False)


def test_weno5_face_bit_exact():
    d = load("weno5.npz")
    got = np.array([O.weno5_face(*row) for row in d["in5"]])
    assert np.array_equal(got, d["weno"]), np.max(np.abs(got - d["weno"]))


def test_saturation_adjustment_and_thermo_match_reference():
    d = load("thermo.npz")
    L = O.lib()
    for (T, p, qv, qc), ref, qvs, tt, vt in zip(d["tpqq"], d["satadj"], d["qvs"], d["t_of_theta"], d["vt"]):
        assert L.oracle_qvs_water(T, p) == qvs
        assert L.oracle_theta_to_temperature(T + 20.0, p) == tt
        assert L.oracle_kessler_rain_speed(qv, 0.5 + qc * 100.0) == vt
        Ta, qva, qca = O.saturation_adjust(T, p, qv, qc)
        assert (Ta, qva, qca) == tuple(ref)


def test_upwind_derivatives_bit_exact():
    d = load("upwind.npz")
    f = np.ascontiguousarray(d["f"])
    nx, ny, nz, dx, dy, dz = d["params"]
    inv = (1.0 / dx, 1.0 / dy, 1.0 / dz)
    fp = f.ctypes.data_as(O._fp)
    got = np.array([O.lib().oracle_upwind_point(fp, int(nx), int(ny), int(nz), int(i), int(j), int(k), int(ax),
                                                float(a), inv[ax]) for (i, j, k, ax), a in zip(d["pts"], d["a"])])
    assert np.array_equal(got, d["upwind"])


def _kessler_case(name):
    d = load(name)
    nx, ny, nz = (int(x) for x in d["shape"])
    g = O.grid(nx, ny, nz, 1000.0, 1000.0, float(d["dz"]))
    cfg = O.config(float(d["dt"]), 6)
    got = O.kessler_tendencies(g, cfg, d["p"], d["theta"], d["rho"], d["qv"], d["qc"], d["qr"], float(d["dt"]))
    return d, got


@pytest.mark.parametrize("fixture", ["kessler.npz", "kessler_column.npz"])
def test_kessler_tendencies_match_reference(fixture):
    d, got = _kessler_case(fixture)
    for name, g in zip(("dth", "dqv", "dqc", "dqr"), got):
        ref = d[name]
        scale = max(float(np.max(np.abs(ref))), 1e-30)
        err = float(np.max(np.abs(g - ref)))
        assert err <= 4 * np.finfo(np.float32).eps * scale, (name, err, scale)


def test_kessler_column_water_budget_closes_with_surface_sink():
    """Column total-water tendency equals minus the surface sink (conservative receiver weighting)."""
    d, (dth, dqv, dqc, dqr) = _kessler_case("kessler_column.npz")
    rho = d["rho"].astype(np.float64)
    dz = float(d["dz"])
    terms = rho * (dqv.astype(np.float64) + dqc + dqr)
    col = terms.sum(axis=2)[1, 1] * dz
    sink = float(d["sink"][1, 1]) * rho[1, 1, 0] * dz
    # tendencies are stored in FP32: bound the residual by FP32 rounding of the gross terms
    gross = (rho * (np.abs(dqv) + np.abs(dqc) + np.abs(dqr))).sum(axis=2)[1, 1] * dz
    assert abs(col + sink) <= 8 * np.finfo(np.float32).eps * gross, (col, sink, gross)
    assert sink > 0.0


def test_acoustic_substep_matches_reference():
    d = load("acoustic.npz")
    nx, ny, nz = (int(x) for x in d["shape"])
    dx, dy, dz, dts = (float(x) for x in d["spacing"])
    g = O.grid(nx, ny, nz, dx, dy, dz)
    st = {n: d[n] for n in ("u", "v", "w", "rho", "p", "theta")}
    for n in ("qv", "qc", "qr"):
        st[n] = np.zeros_like(d["u"])
    base = {"rho0": d["rho0"], "p0": d["p0"], "theta0": np.zeros(nz), "qv0": np.zeros(nz), "u0": np.zeros(nz),
            "v0": np.zeros(nz)}
    out = O.acoustic_substep(g, base, st, dts, np.float32)
    for n in ("u", "v", "w", "rho", "p"):
        ref = d[n + "_out"]
        assert np.array_equal(out[n], ref), (n, float(np.max(np.abs(out[n] - ref))))
