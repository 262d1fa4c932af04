"""The backend's CM1 input_sounding entry points (host code, no device)
against the reference's own reader and column builder (golden vectors from
tests/golden/make_cm1_golden.py), plus the writer round trip and the
analytic-sounding export."""
import ctypes
import json
import os

import numpy as np
import pytest

from nimbus_amd import lib as L

HERE = os.path.dirname(os.path.abspath(__file__))
GOLD_DIR = os.path.join(HERE, "golden")
GOLD = json.load(open(os.path.join(GOLD_DIR, "cm1_golden.json")))
NAMES = sorted(GOLD["files"])


def path_of(name):
    return os.path.join(GOLD_DIR, "cm1", name)


def reader_rows(col):
    return np.stack([col["z"], col["p"] / 100.0, col["T"], col["theta"], col["qv"], col["u"], col["v"]], 1)


# Which check refused, by the phrase the reference's message uses -> the tag
# this library's message carries for the same check.
REFUSAL_TAGS = (
    ("failed quality control", "[quality]"),
    ("silently shortened profile", "[qc-shortened]"),
    ("columns; expected", "[columns]"),
    ("is not numeric", "[not-numeric]"),
    ("first profile level must be at z=0", "[first-z0]"),
    ("heights must increase strictly", "[z-order]"),
    ("at least five vertical levels", "[too-few-levels]"),
    ("surface header theta does not match", "[header-theta]"),
    ("wind component magnitude exceeds", "[wind-limit]"),
    ("model top interface", "[top-coverage]"),
)


def same_refusal(ref_msg, ours):
    """Our refusal is the same check the reference's message names."""
    tags = [tag for phrase, tag in REFUSAL_TAGS if phrase in ref_msg]
    assert len(tags) == 1, f"no single check for the reference message {ref_msg!r}"
    return tags[0] in ours


def test_every_reference_refusal_maps_to_one_check():
    msgs = [c["error"] for f in GOLD["files"].values() for c in f["columns"] if c["error"]]
    msgs += [e["read"]["error"] for e in GOLD["synthetic"].values() if e["read"]["error"]]
    assert msgs
    for m in msgs:
        assert len([t for p, t in REFUSAL_TAGS if p in m]) == 1, m


@pytest.mark.parametrize("name", NAMES)
def test_reader_bit_exact(name):
    col = L.cm1_read(path_of(name))
    assert np.array_equal(reader_rows(col), np.array(GOLD["files"][name]["read"]["rows"]))
    text = open(path_of(name)).read()
    assert np.array_equal(reader_rows(L.cm1_parse(text)), reader_rows(col))


@pytest.mark.parametrize("name", NAMES)
def test_columns_bit_exact(name):
    col = L.cm1_read(path_of(name))
    for case in GOLD["files"][name]["columns"]:
        ex = bool(case["extrapolate"])
        if case["error"]:
            with pytest.raises(L.NimbusError) as e:
                L.cm1_column(col, case["z"], case["dz"], case["method"], ex, ex)
            assert same_refusal(case["error"], str(e.value))
            continue
        out = L.cm1_column(col, case["z"], case["dz"], case["method"], ex, ex)
        rows = np.stack([np.array(case["z"])] + [out[k] for k in ("theta0", "T0", "p0", "rho0", "qv0", "u0", "v0")], 1)
        assert np.array_equal(rows, np.array(case["rows"])), (case["z"][:3], case["method"], ex)


@pytest.mark.parametrize("name", sorted(GOLD["synthetic"]))
def test_synthetic_texts(name):
    e = GOLD["synthetic"][name]
    if e["read"]["error"]:
        with pytest.raises(L.NimbusError) as err:
            L.cm1_parse(e["text"])
        assert same_refusal(e["read"]["error"], str(err.value))
        return
    col = L.cm1_parse(e["text"])
    assert np.array_equal(reader_rows(col), np.array(e["read"]["rows"]))
    z = [250.0, 750.0, 1250.0, 1750.0]
    out = L.cm1_column(col, z, 250.0)
    rows = np.stack([np.array(z)] + [out[k] for k in ("theta0", "T0", "p0", "rho0", "qv0", "u0", "v0")], 1)
    assert np.array_equal(rows, np.array(e["column"]["rows"]))


def test_base_state_is_the_column_at_cell_centres():
    col = L.cm1_read(path_of("MEDCAPENEW_SEMICIRCLE_06"))
    g = L.make_grid(16, 16, 38, 500.0, 500.0, 500.0)
    b = L.cm1_base_state(g, col, L.INTERP_SPLINE)
    c = L.cm1_column(col, (np.arange(38) + 0.5) * 500.0, 500.0, L.INTERP_SPLINE)
    for k in ("rho0", "p0", "theta0", "qv0", "u0", "v0"):
        assert np.array_equal(b[k], c[k]), k
    with pytest.raises(L.NimbusError, match=r"\[top-coverage\]"):
        L.cm1_base_state(L.make_grid(16, 16, 40, 500.0, 500.0, 500.0), col)


@pytest.mark.parametrize("name", NAMES)
def test_writer_round_trip(name):
    col = L.cm1_read(path_of(name))
    back = L.cm1_parse(L.cm1_format(col))
    for k in ("z", "theta", "u", "v"):
        assert np.array_equal(back[k], col[k]), k
    # qv is written in g/kg: x/1000*1000 may move the last bit
    assert np.allclose(back["qv"], col["qv"], rtol=4e-16, atol=0)
    assert np.allclose(back["p"], col["p"], rtol=1e-14, atol=0)
    assert np.allclose(back["T"], col["T"], rtol=1e-14, atol=0)
    assert abs(back["p_sfc"] - col["p_sfc"]) <= 1e-15 * col["p_sfc"]


def test_capacity_and_argument_errors():
    lib = L.load_library()
    text = open(path_of("HIGHCAPENEW_QUARTERCIRCLE")).read().encode()
    col, s = L._cm1_empty(10)
    assert lib.nimbus_cm1_parse(text, len(text), ctypes.byref(s)) == -1
    assert s.n == 96 and "capacity" in lib.nimbus_hip_last_error().decode()
    assert lib.nimbus_cm1_parse(None, 0, ctypes.byref(s)) == -1
    assert lib.nimbus_cm1_read(b"/nonexistent/sounding", ctypes.byref(s)) == -1
    assert "not a regular file" in lib.nimbus_hip_last_error().decode()
    need = ctypes.c_size_t(0)
    full = L.cm1_read(path_of("HIGHCAPENEW_QUARTERCIRCLE"))
    c = {k: (np.ascontiguousarray(full[k]) if k in L.CM1_COLUMNS else full[k]) for k in full}
    st = L._cm1_struct(c)
    buf = ctypes.create_string_buffer(16)
    assert lib.nimbus_cm1_format(ctypes.byref(st), buf, 16, ctypes.byref(need)) == -1 and need.value > 16
    with pytest.raises(L.NimbusError, match="increase strictly"):
        L.cm1_column(full, [100.0, 50.0], 50.0)
    with pytest.raises(L.NimbusError, match="vertical spacing"):
        L.cm1_column(full, [100.0], 0.0)


def test_oversized_file_refused(tmp_path):
    p = tmp_path / "big"
    with open(p, "wb") as f:
        f.truncate(9 * 1024 * 1024)
    with pytest.raises(L.NimbusError, match="8 MiB"):
        L.cm1_read(str(p))


def test_analytic_sounding_export_round_trip():
    """The seeded workloads' analytic sounding written as CM1 text and read
    back: the column at the sample heights keeps theta, qv and the winds
    exactly; the pressure differs from the analytic base state only by the
    reference path's dry re-integration (no 0.608 qv term) and quadrature."""
    g = L.make_grid(16, 16, 36, 500.0, 500.0, 500.0)
    zc = (np.arange(g.nz) + 0.5) * g.dz
    z = np.concatenate([[0.0], zc, [g.nz * g.dz]])
    col = L.sounding_levels(z)
    base = L.base_state(g)
    assert np.array_equal(col["theta"][1:-1], base["theta0"])
    assert np.array_equal(col["u"][1:-1], base["u0"]) and np.array_equal(col["v"][1:-1], base["v0"])
    assert np.allclose(col["p"][1:-1], base["p0"], rtol=1e-12)
    back = L.cm1_parse(L.cm1_format(col))
    b2 = L.cm1_base_state(g, back)
    assert np.allclose(b2["theta0"], base["theta0"], rtol=1e-15, atol=0)
    assert np.allclose(b2["qv0"], base["qv0"], rtol=1e-15, atol=0)
    assert np.array_equal(b2["u0"], base["u0"]) and np.array_equal(b2["v0"], base["v0"])
    # dry theta instead of theta_v: the Exner function falls faster aloft, so
    # the re-integrated pressure is lower by up to ~0.6 % at 18 km
    assert abs(b2["p0"][0] / base["p0"][0] - 1.0) < 1e-5
    assert np.all(b2["p0"] <= base["p0"] * (1.0 + 1e-5)) and np.allclose(b2["p0"], base["p0"], rtol=1e-2)
    assert np.allclose(b2["rho0"], base["rho0"], rtol=1e-2)
    # and equals the dry trapezoid integral of the analytic theta from its own first pressure
    kappa = 287.0 / 1004.0
    pi = (b2["p0"][0] / 1e5) ** kappa - np.concatenate(
        [[0.0], np.cumsum(9.81 / 1004.0 * g.dz * 0.5 * (1.0 / base["theta0"][:-1] + 1.0 / base["theta0"][1:]))])
    assert np.allclose(b2["p0"], 1e5 * pi ** (1.0 / kappa), rtol=1e-12)
    assert np.all(np.diff(b2["p0"]) < 0)


def test_fill_over_base_equals_analytic_fill():
    g = L.make_grid(12, 11, 20, 1000.0, 1000.0, 500.0)
    prm = L.ic_params(g, noise_amp=0.2, seed=3)
    for dt in (np.float32, np.float64):
        a = L.initial_state(g, prm, dt)
        b = L.initial_state_from_base(g, L.base_state(g), prm, dt)
        for f in L.FIELDS:
            assert np.array_equal(a[f], b[f]), f
