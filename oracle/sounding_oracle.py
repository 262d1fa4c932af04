"""TEST INFRASTRUCTURE ONLY: the CM1 input_sounding reader and the model
column built from a sounding, restated in plain Python (IEEE double
arithmetic in the reference's operation order) as the checker for the
backend's nimbus_cm1_* entry points. Only tests/ import it.

Reference algorithm, function by function:
  read_cm1        src/soundings/schemes/cm1/cm1_sounding.cpp:31-134 (comments,
                  numeric rows), :159-281 (checks, hydrostatic reconstruction);
                  quality control src/soundings/base/sounding_algorithms.cpp:153-353
  interpolate     linear sounding_algorithms.cpp:427-641 (linear_interpolate :142-145);
                  monotone spline src/soundings/schemes/sharpy/sharpy_interpolation.cpp:114-318;
                  log-linear pressure sharpy_interpolation.cpp:27-104
  column          src/init/sounding/sources/file_sounding.cpp:344-438 (target checks,
                  top-interface coverage), :63-327 (theta-primary re-integration);
                  self-consistency src/init/sounding/sounding_source.cpp:34-70

Pins (tests/test_oracle_sounding.py): bit-exact agreement with the
reference's own reader and column builder on the four CM1 fixtures the
reference tests hold (tests/golden/cm1/, vectors from
tests/golden/make_cm1_golden.py); the reference tests' known answers
(tests/data/test_soundings.cpp:708-805, tests/init/test_file_sounding.cpp:
401-421,556-593,647-664); closed forms (a dry constant-theta column has a
linear Exner function; linear and PCHIP interpolation reproduce linear data;
log-linear pressure reproduces an exponential profile; PCHIP stays monotone).
"""
import math

G, RD, CP, P00 = 9.81, 287.0, 1004.0, 100000.0
KAPPA = RD / CP
MAX_LEVELS = 100000
QC = {"p_min_hpa": 1.0, "p_max_hpa": 1100.0, "t_min": 200.0, "t_max": 350.0}
LINEAR, SPLINE, LOG_PRESSURE = 0, 1, 2


class SoundingError(ValueError):
    pass


def _fail(msg):
    raise SoundingError("CM1 input_sounding: " + msg)


def _rows(text):
    """(line number, stripped text) of every line with content after '!'/'#' comments."""
    out = []
    for number, line in enumerate(text.split("\n"), start=1):
        cut = min([p for p in (line.find("!"), line.find("#")) if p >= 0], default=len(line))
        body = line[:cut].strip(" \t\r\n")
        if body:
            out.append((number, body))
            if len(out) > MAX_LEVELS + 1:
                _fail("exceeds the 100,000-level safety limit")
    return out


def _numbers(number, body, want):
    tok = body.split()
    if len(tok) != want:
        _fail(f"line {number} has {len(tok)} columns; expected {want}")
    vals = []
    for c, t in enumerate(tok, start=1):
        try:
            x = float(t.replace("d", "e").replace("D", "E"))
        except ValueError:
            _fail(f"line {number}, column {c} is not numeric: '{t}'")
        if not math.isfinite(x):
            _fail(f"line {number}, column {c} is not a finite number: '{t}'")
        vals.append(x)
    return vals


def read_cm1(text):
    """Parsed column: dict of lists z, p_hpa, T, theta, qv (kg/kg), u, v and
    the header (p_sfc_hpa, theta_sfc, qv_sfc_gkg)."""
    rows = _rows(text)
    if not rows:
        _fail("contains no numeric rows")
    if len(rows) < 6:
        _fail("requires at least five vertical levels after the surface header")
    ps, ths, qs = _numbers(*rows[0], 3)
    if not 100.0 <= ps <= 1100.0:
        _fail("surface pressure must be in [100, 1100] hPa")
    if not 150.0 <= ths <= 700.0:
        _fail("surface potential temperature must be in [150, 700] K")
    if not 0.0 <= qs <= 100.0:
        _fail("surface water-vapor mixing ratio must be in [0, 100] g/kg")
    col = {k: [] for k in ("z", "p_hpa", "T", "theta", "qv", "u", "v")}
    exner = (ps * 100.0 / P00) ** KAPPA
    z_prev = 0.0
    thv_prev = ths * (1.0 + 0.608 * qs / 1000.0)
    for r, (number, body) in enumerate(rows[1:], start=1):
        z, th, q, u, v = _numbers(number, body, 5)
        if z < 0.0:
            _fail(f"height cannot be negative at line {number}")
        if r == 1 and abs(z) > 1.0e-9:
            _fail("first profile level must be at z=0 m")
        if r > 1 and not z > col["z"][-1]:
            _fail(f"heights must increase strictly at line {number}")
        if not 150.0 <= th <= 700.0:
            _fail(f"potential temperature must be in [150, 700] K at line {number}")
        if not 0.0 <= q <= 100.0:
            _fail(f"water-vapor mixing ratio must be in [0, 100] g/kg at line {number}")
        if abs(u) > 300.0 or abs(v) > 300.0:
            _fail(f"wind component magnitude exceeds 300 m/s at line {number}")
        if r == 1 and (abs(th - ths) > 1.0e-3 or abs(q - qs) > 1.0e-3):
            _fail("surface header theta/qv does not match the z=0 row")
        qk = q / 1000.0
        thv = th * (1.0 + 0.608 * qk)
        exner -= (G / CP) * (z - z_prev) * 0.5 * ((1.0 / thv_prev) + (1.0 / thv))
        if not (math.isfinite(exner) and exner > 0.0):
            _fail(f"hydrostatic Exner pressure became invalid at line {number}")
        p = P00 * exner ** (1.0 / KAPPA)
        T = th * exner
        for k, x in (("z", z), ("p_hpa", p / 100.0), ("T", T), ("theta", th), ("qv", qk), ("u", u), ("v", v)):
            col[k].append(x)
        z_prev, thv_prev = z, thv
    good = 0
    for k in range(len(col["z"])):
        ok = (QC["p_min_hpa"] <= col["p_hpa"][k] <= QC["p_max_hpa"] and QC["t_min"] <= col["T"][k] <= QC["t_max"]
              and col["qv"][k] >= 0.0 and col["theta"][k] > 0.0)
        if k > 0:
            ok = ok and col["z"][k] > col["z"][k - 1] and col["p_hpa"][k] < col["p_hpa"][k - 1]
        good += ok
    if good < 5:
        _fail("failed quality control")
    if good != len(col["z"]):
        _fail("quality control rejected one or more levels; refusing a silently shortened profile")
    col.update(p_sfc_hpa=ps, theta_sfc=ths, qv_sfc_gkg=qs)
    return col


def _lerp(x1, y1, x2, y2, x):
    if x1 == x2:
        return y1
    return y1 + (y2 - y1) * (x - x1) / (x2 - x1)


def _log_lerp(x1, y1, x2, y2, x):
    if x1 == x2:
        return y1
    if not (math.isfinite(y1) and math.isfinite(y2)) or y1 <= 0.0 or y2 <= 0.0:
        return _lerp(x1, y1, x2, y2, x)
    return math.exp(_lerp(x1, math.log(y1), x2, math.log(y2), x))


def _lower_bound(zs, t):
    lo, hi = 0, len(zs)
    while lo < hi:
        mid = (lo + hi) // 2
        if zs[mid] < t:
            lo = mid + 1
        else:
            hi = mid
    return lo


def linear(zs, ys, t, below=False, above=False, interp=_lerp):
    n = len(zs)
    i2 = _lower_bound(zs, t)
    if i2 == 0:
        return interp(zs[0], ys[0], zs[1], ys[1], t) if below and n > 1 else ys[0]
    if i2 == n:
        return interp(zs[n - 2], ys[n - 2], zs[n - 1], ys[n - 1], t) if above and n > 1 else ys[n - 1]
    return interp(zs[i2 - 1], ys[i2 - 1], zs[i2], ys[i2], t)


def _pchip_end(h0, h1, d0, d1):
    den = h0 + h1
    if not math.isfinite(den) or den == 0.0:
        return d0
    m = ((2.0 * h0 + h1) * d0 - h0 * d1) / den
    if m * d0 <= 0.0:
        return 0.0
    if abs(m) > 3.0 * abs(d0):
        return 3.0 * d0
    return m


def pchip_slopes(x, y):
    n = len(x)
    if n < 2:
        return [0.0] * n
    h = [x[i + 1] - x[i] for i in range(n - 1)]
    d = [(y[i + 1] - y[i]) / h[i] for i in range(n - 1)]
    if n == 2:
        return [d[0], d[0]]
    m = [0.0] * n
    m[0] = _pchip_end(h[0], h[1], d[0], d[1])
    m[n - 1] = _pchip_end(h[n - 2], h[n - 3], d[n - 2], d[n - 3])
    for i in range(1, n - 1):
        if d[i - 1] == 0.0 or d[i] == 0.0 or d[i - 1] * d[i] < 0.0:
            continue
        w1, w2 = 2.0 * h[i] + h[i - 1], h[i] + 2.0 * h[i - 1]
        m[i] = (w1 + w2) / ((w1 / d[i - 1]) + (w2 / d[i]))
    return m


def _hermite(x0, x1, y0, y1, m0, m1, x):
    h = x1 - x0
    if h == 0.0:
        return y0
    t = (x - x0) / h
    t2 = t * t
    t3 = t2 * t
    return (((2.0 * t3) - (3.0 * t2) + 1.0) * y0 + (t3 - (2.0 * t2) + t) * h * m0
            + ((-2.0 * t3) + (3.0 * t2)) * y1 + (t3 - t2) * h * m1)


def spline(zs, ys, targets, below=False, above=False):
    n = len(zs)
    ok = n >= 2 and all(math.isfinite(a) and math.isfinite(b) for a, b in zip(zs, ys)) and \
        all(zs[i] > zs[i - 1] for i in range(1, n))
    m = pchip_slopes(zs, ys) if ok else None
    out = []
    for t in targets:
        i2 = _lower_bound(zs, t)
        if i2 == 0:
            out.append(_lerp(zs[0], ys[0], zs[1], ys[1], t) if below else ys[0])
        elif i2 == n:
            out.append(_lerp(zs[n - 2], ys[n - 2], zs[n - 1], ys[n - 1], t) if above else ys[n - 1])
        elif zs[i2] == t:
            out.append(ys[i2])
        elif ok:
            out.append(_hermite(zs[i2 - 1], zs[i2], ys[i2 - 1], ys[i2], m[i2 - 1], m[i2], t))
        else:
            out.append(_lerp(zs[i2 - 1], ys[i2 - 1], zs[i2], ys[i2], t))
    return out


def interpolate(col, targets, method=LINEAR, below=False, above=False):
    keys = ("p_hpa", "T", "theta", "qv", "u", "v")
    if method == SPLINE:
        r = {k: spline(col["z"], col[k], targets, below, above) for k in keys}
    else:
        r = {k: [linear(col["z"], col[k], t, below, above) for t in targets] for k in keys}
    if method == LOG_PRESSURE:
        r["p_hpa"] = [linear(col["z"], col["p_hpa"], t, below, above, _log_lerp) for t in targets]
    return r


def column(col, z, dz, method=LINEAR, below=False, above=False):
    """Model column at heights z: dict of lists theta, T, p, rho, qv, u, v."""
    nz = len(z)
    if nz == 0:
        _fail("target model-height column is empty")
    if not (math.isfinite(dz) and dz > 0.0):
        _fail("target vertical spacing must be positive and finite")
    if not math.isfinite(z[0]) or z[0] < 0.0:
        _fail("target model heights must be finite and non-negative")
    for k in range(1, nz):
        if not math.isfinite(z[k]) or z[k] <= z[k - 1]:
            _fail("target model heights must increase strictly")
    half = 0.5 * (z[-1] - z[-2]) if nz > 1 else 0.5 * dz
    top = z[-1] + half
    tol = 1.0e-10 * max(1.0, abs(top), abs(col["z"][-1]))
    if top > col["z"][-1] + tol and not above:
        _fail("model top interface exceeds sounding top; refusing to extrapolate an uncovered upper half-cell")
    r = interpolate(col, list(z), method, below, above)
    exner = (r["p_hpa"][0] * 100.0 / P00) ** KAPPA
    out = {k: [] for k in ("theta", "T", "p", "rho", "qv", "u", "v")}
    for k in range(nz):
        if k > 0:
            exner -= (G / CP) * (z[k] - z[k - 1]) * 0.5 * ((1.0 / r["theta"][k - 1]) + (1.0 / r["theta"][k]))
            if not (math.isfinite(exner) and exner > 0.0):
                _fail("hydrostatic Exner pressure became non-positive")
        th = r["theta"][k]
        T = th * exner
        p = P00 * exner ** (1.0 / KAPPA)
        for key, x in (("theta", th), ("T", T), ("p", p), ("rho", p / (RD * T)), ("qv", r["qv"][k]),
                       ("u", r["u"][k]), ("v", r["v"][k])):
            out[key].append(x)
    return out
