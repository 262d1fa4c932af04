"""The CM1 sounding oracle against the reference's own reader and column
builder (golden vectors from tests/golden/make_cm1_golden.py), the reference
tests' known answers and closed forms."""
import json
import math
import os

import pytest

from oracle import sounding_oracle as SO

HERE = os.path.dirname(os.path.abspath(__file__))
GOLD_DIR = os.path.join(HERE, "golden")
GOLD = json.load(open(os.path.join(GOLD_DIR, "cm1_golden.json")))
NAMES = sorted(GOLD["files"])


def fixture_text(name):
    with open(os.path.join(GOLD_DIR, "cm1", name)) as f:
        return f.read()


def reader_rows(col):
    return [[col[k][i] for k in ("z", "p_hpa", "T", "theta", "qv", "u", "v")] for i in range(len(col["z"]))]


@pytest.mark.parametrize("name", NAMES)
def test_reader_bit_exact_against_reference(name):
    col = SO.read_cm1(fixture_text(name))
    assert reader_rows(col) == GOLD["files"][name]["read"]["rows"]


@pytest.mark.parametrize("name", NAMES)
def test_columns_bit_exact_against_reference(name):
    col = SO.read_cm1(fixture_text(name))
    for case in GOLD["files"][name]["columns"]:
        ex = bool(case["extrapolate"])
        if case["error"]:
            with pytest.raises(SO.SoundingError):
                SO.column(col, case["z"], case["dz"], case["method"], ex, ex)
            continue
        out = SO.column(col, case["z"], case["dz"], case["method"], ex, ex)
        rows = [[case["z"][i]] + [out[k][i] for k in ("theta", "T", "p", "rho", "qv", "u", "v")]
                for i in range(len(case["z"]))]
        assert rows == case["rows"], (case["z"][:3], case["method"], ex)


@pytest.mark.parametrize("name", sorted(GOLD["synthetic"]))
def test_synthetic_texts_against_reference(name):
    e = GOLD["synthetic"][name]
    if e["read"]["error"]:
        with pytest.raises(SO.SoundingError):
            SO.read_cm1(e["text"])
        return
    col = SO.read_cm1(e["text"])
    assert reader_rows(col) == e["read"]["rows"]
    out = SO.column(col, [250.0, 750.0, 1250.0, 1750.0], 250.0)
    rows = [[z] + [out[k][i] for k in ("theta", "T", "p", "rho", "qv", "u", "v")]
            for i, z in enumerate([250.0, 750.0, 1250.0, 1750.0])]
    assert rows == e["column"]["rows"]


def test_reference_known_answers():
    # tests/data/test_soundings.cpp:708-776: every fixture parses to >= 90 levels
    # from z = 0 to >= 19 km, p strictly decreasing, theta = T (p0/p)^kappa to 1e-12
    for name in NAMES:
        col = SO.read_cm1(fixture_text(name))
        n = len(col["z"])
        assert n >= 90 and col["z"][0] == 0.0 and col["z"][-1] >= 19000.0
        for k in range(n):
            assert col["T"][k] > 0.0 and col["qv"][k] >= 0.0
            if k:
                assert col["z"][k] > col["z"][k - 1] and col["p_hpa"][k] < col["p_hpa"][k - 1]
            th = col["T"][k] * (SO.P00 / (col["p_hpa"][k] * 100.0)) ** SO.KAPPA
            assert math.isclose(th, col["theta"][k], rel_tol=1e-12)
    col = SO.read_cm1(fixture_text("HIGHCAPENEW_QUARTERCIRCLE"))
    assert math.isclose(col["qv"][0], 0.0147394) and col["u"][0] == -15.798 and col["v"][0] == 0.0
    # tests/data/test_soundings.cpp:778-799 and test_file_sounding.cpp:401-421: at 25 m
    # (halfway between the 0 and 50 m rows) every primary field is the midpoint
    r = SO.interpolate(col, [25.0])
    assert abs(r["theta"][0] - 0.5 * (306.465 + 306.494)) <= 1e-12
    assert abs(r["qv"][0] - 0.5 * (14.7394 + 14.710) / 1000.0) <= 1e-12
    assert abs(r["u"][0] - 0.5 * (-15.798 - 15.388)) <= 1e-12
    assert abs(r["v"][0] - 0.5 * (0.0 + 0.848)) <= 1e-12
    c = SO.column(col, [25.0, 75.0, 125.0, 175.0, 225.0], 50.0)
    assert abs(c["theta"][0] - 0.5 * (306.465 + 306.494)) <= 1e-12
    # test_file_sounding.cpp:556-593: 32 levels from 250 m to 15.75 km, finite, p decreasing
    z = [250.0 + 500.0 * k for k in range(32)]
    for name in NAMES:
        c = SO.column(SO.read_cm1(fixture_text(name)), z, 500.0)
        assert all(math.isfinite(x) for k in c for x in c[k])
        assert all(c["p"][k] < c["p"][k - 1] for k in range(1, 32))
    # test_file_sounding.cpp:642-661: top-interface coverage
    for zz in ([19500.0], [18800.0]):
        with pytest.raises(SO.SoundingError, match="upper half-cell"):
            SO.column(col, zz, 500.0)


def test_dry_isentropic_exner_is_linear():
    # constant theta, no vapour: d(pi)/dz = -g/(cp theta) exactly, and the
    # trapezoid rule on a constant integrand is exact
    text = "1000 300 0\n" + "".join(f"{z} 300 0 5 -3\n" for z in range(0, 6001, 1000))
    col = SO.read_cm1(text)
    for z, p_hpa, T in zip(col["z"], col["p_hpa"], col["T"]):
        pi = 1.0 - SO.G * z / (SO.CP * 300.0)
        assert math.isclose(T, 300.0 * pi, rel_tol=1e-13)
        assert math.isclose(p_hpa * 100.0, SO.P00 * pi ** (1.0 / SO.KAPPA), rel_tol=1e-12)
    c = SO.column(col, [100.0 + 200.0 * k for k in range(25)], 200.0)
    pi0 = (c["p"][0] / SO.P00) ** SO.KAPPA
    for k, z in enumerate(100.0 + 200.0 * k for k in range(25)):
        pi = pi0 - SO.G * (z - 100.0) / (SO.CP * 300.0)
        assert math.isclose(c["T"][k], 300.0 * pi, rel_tol=1e-12)
        assert math.isclose(c["rho"][k], c["p"][k] / (SO.RD * c["T"][k]), rel_tol=1e-15)
        assert c["u"][k] == 5.0 and c["v"][k] == -3.0 and c["qv"][k] == 0.0


def test_interpolants_reproduce_linear_and_exponential_data():
    zs = [0.0, 300.0, 700.0, 1500.0, 2200.0, 4000.0]
    lin = [2.5 - 0.003 * z for z in zs]
    targets = [0.0, 150.0, 300.0, 512.5, 1999.0, 3999.0, 4000.0]
    assert all(math.isclose(SO.linear(zs, lin, t), 2.5 - 0.003 * t, rel_tol=1e-13, abs_tol=1e-13) for t in targets)
    assert all(math.isclose(v, 2.5 - 0.003 * t, rel_tol=1e-12, abs_tol=1e-12)
               for v, t in zip(SO.spline(zs, lin, targets), targets))
    # extrapolation continues the end segments; without it the end values are held
    assert math.isclose(SO.linear(zs, lin, 5000.0, above=True), 2.5 - 15.0, rel_tol=1e-13)
    assert SO.linear(zs, lin, 5000.0) == lin[-1]
    # log-linear pressure is exact on p = p0 exp(-z / H)
    p = [1000.0 * math.exp(-z / 8000.0) for z in zs]
    col = {"z": zs, "p_hpa": p, "T": lin, "theta": lin, "qv": lin, "u": lin, "v": lin}
    r = SO.interpolate(col, targets, SO.LOG_PRESSURE)
    assert all(math.isclose(v, 1000.0 * math.exp(-t / 8000.0), rel_tol=1e-13) for v, t in zip(r["p_hpa"], targets))


def test_pchip_is_monotone_and_flat_at_extrema():
    zs = [0.0, 1.0, 2.0, 3.0, 4.0, 5.0, 6.0]
    ys = [0.0, 0.1, 0.15, 3.0, 3.05, 3.1, 10.0]          # monotone with a sharp step
    dense = [0.05 * k for k in range(121)]
    v = SO.spline(zs, ys, dense)
    assert all(b >= a - 1e-15 for a, b in zip(v, v[1:]))
    assert min(v) >= 0.0 and max(v) <= 10.0
    m = SO.pchip_slopes([0.0, 1.0, 2.0, 3.0], [0.0, 1.0, 0.5, 2.0])
    assert m[1] == 0.0 and m[2] == 0.0                   # local extrema get zero slope


@pytest.mark.parametrize("bad,match", [
    ("", "no numeric rows"),
    ("1000 300 10\n0 300 10 0 0\n", "at least five"),
    ("1000 300 10 1\n" + "".join(f"{z} 300 10 0 0\n" for z in range(0, 2501, 500)), "columns"),
    ("1000 300 10\n" + "".join(f"{z} 300 10 0 0\n" for z in range(0, 2501, 500)).replace("1000 300", "1000 9e999"),
     "finite"),
])
def test_reader_refusals(bad, match):
    with pytest.raises(SO.SoundingError, match=match):
        SO.read_cm1(bad)
