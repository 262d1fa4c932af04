// CM1 input_sounding text I/O and the model column built from a sounding
// (host code, no device).
//
// Reading follows the reference's native CM1 reader
// (src/soundings/schemes/cm1/cm1_sounding.cpp:22-281): comment stripping,
// strict column counts, range checks, the hydrostatic reconstruction of p and
// T from the header pressure, then the shared quality control
// (src/soundings/base/sounding_algorithms.cpp:153-353), with a shortened
// profile refused. The model column follows FileSoundingSource::build and
// sounding_from_data (src/init/sounding/sources/file_sounding.cpp:63-327,
// 344-438): interpolation to the model heights (linear, monotone PCHIP or
// log-linear pressure: sounding_algorithms.cpp:427-641,
// src/soundings/schemes/sharpy/sharpy_interpolation.cpp:13-407), the
// coverage check of the model top interface, and a dry Exner re-integration
// that keeps theta exact. The arithmetic of the two hydrostatic recurrences
// and of the interpolants is written in the reference's operation order, so
// the columns agree bit for bit.
#include "nimbus_hip.h"
#include "nimbus_internal.hpp"

#include <algorithm>
#include <cerrno>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <iterator>
#include <limits>
#include <stdexcept>
#include <string>
#include <sys/stat.h>
#include <vector>

namespace nimbus
{
namespace
{

constexpr double kVirtualCoef = 0.608;  // theta_v = theta (1 + 0.608 qv)
// Quality-control bounds of the file-sounding path (include/init/sounding/file_sounding.hpp:26-29).
constexpr double kQcPMinHpa = 1.0, kQcPMaxHpa = 1100.0;
constexpr double kQcTMin = 200.0, kQcTMax = 350.0;

struct Bad : std::runtime_error
{
    using std::runtime_error::runtime_error;
};

[[noreturn]] void bad(const std::string& what) { throw Bad("CM1 input_sounding: " + what); }

std::string at_line(std::size_t line) { return " (line " + std::to_string(line) + ")"; }

// One substantive (non-blank after comment removal) line of the file.
struct Line
{
    std::size_t number;
    std::string text;
};

std::vector<Line> substantive(const char* text, std::size_t len)
{
    std::vector<Line> out;
    std::size_t pos = 0, number = 0;
    while (pos < len)
    {
        std::size_t end = pos;
        while (end < len && text[end] != '\n') ++end;
        ++number;
        std::string line(text + pos, end - pos);
        pos = end + 1;
        const std::size_t cut = line.find_first_of("!#");
        if (cut != std::string::npos) line.resize(cut);
        const std::size_t a = line.find_first_not_of(" \t\r\n");
        if (a == std::string::npos) continue;
        const std::size_t b = line.find_last_not_of(" \t\r\n");
        out.push_back({number, line.substr(a, b - a + 1)});
        if (out.size() > std::size_t(NIMBUS_CM1_MAX_LEVELS) + 1) bad("[level-limit] more than 100000 levels");
    }
    return out;
}

// A whitespace-separated numeric row with exactly `want` entries; Fortran
// 'd'/'D' exponents are read as 'e'/'E'.
std::vector<double> numbers(const Line& ln, std::size_t want)
{
    std::vector<std::string> tok;
    const char* ws = " \t\n\v\f\r";
    std::size_t p = ln.text.find_first_not_of(ws);
    while (p != std::string::npos)
    {
        const std::size_t q = ln.text.find_first_of(ws, p);
        tok.push_back(ln.text.substr(p, q == std::string::npos ? std::string::npos : q - p));
        p = q == std::string::npos ? q : ln.text.find_first_not_of(ws, q);
    }
    if (tok.size() != want)
        bad("[columns] line " + std::to_string(ln.number) + ": " + std::to_string(tok.size()) + " values where " +
            std::to_string(want) + " belong");
    std::vector<double> v;
    for (std::size_t c = 0; c < tok.size(); ++c)
    {
        std::string t = tok[c];
        for (char& ch : t)
        {
            if (ch == 'd') ch = 'e';
            if (ch == 'D') ch = 'E';
        }
        errno = 0;
        char* end = nullptr;
        const double x = std::strtod(t.c_str(), &end);
        const std::string where = " line " + std::to_string(ln.number) + ", value " + std::to_string(c + 1);
        if (end == t.c_str() || errno == ERANGE) bad("[not-numeric]" + where + ": cannot read '" + tok[c] + "'");
        if (std::size_t(end - t.c_str()) != t.size() || !std::isfinite(x))
            bad("[not-finite]" + where + ": '" + tok[c] + "' is not a finite value");
        v.push_back(x);
    }
    return v;
}

void need(bool ok, const std::string& what)
{
    if (!ok) bad(what);
}

// A column in the units of the reference's SoundingData (pressure in hPa);
// p_pa keeps the reader's pressure before the division, so that the Pa value
// handed out divides back to exactly the reader's hPa value.
struct Column
{
    std::vector<double> z, p_hpa, T, theta, qv, u, v;
    std::vector<double> p_pa;
};

// Parse, validate and reconstruct (cm1_sounding.cpp:159-281).
Column parse(const char* text, std::size_t len, double hdr[3])
{
    const std::vector<Line> lines = substantive(text, len);
    need(!lines.empty(), "[empty] no data rows");
    need(lines.size() >= 6, "[too-few-levels] fewer than 5 levels follow the header");
    const std::vector<double> h = numbers(lines[0], 3);
    const double ps_hpa = h[0], ths = h[1], qvs_gkg = h[2];
    need(ps_hpa >= 100.0 && ps_hpa <= 1100.0, "[header-p] header pressure outside 100..1100 hPa");
    need(ths >= 150.0 && ths <= 700.0, "[header-theta-range] header theta outside 150..700 K");
    need(qvs_gkg >= 0.0 && qvs_gkg <= 100.0, "[header-qv-range] header qv outside 0..100 g/kg");
    hdr[0] = ps_hpa * 100.0;
    hdr[1] = ths;
    hdr[2] = qvs_gkg / 1000.0;

    Column c;
    const double kappa = kRd / kCp;
    // Exner integration of d(pi)/dz = -g / (cp theta_v) by the trapezoid
    // rule on 1/theta_v, starting from the header values at z = 0.
    double exner = std::pow(ps_hpa * 100.0 / kP00, kappa);
    double z_prev = 0.0;
    double thv_prev = ths * (1.0 + kVirtualCoef * qvs_gkg / 1000.0);
    for (std::size_t r = 1; r < lines.size(); ++r)
    {
        const std::vector<double> row = numbers(lines[r], 5);
        const std::size_t ln = lines[r].number;
        const double z = row[0], th = row[1], q_gkg = row[2], u = row[3], v = row[4];
        need(z >= 0.0, "[z-negative] negative height" + at_line(ln));
        if (r == 1)
            need(std::abs(z) <= 1.0e-9, "[first-z0] the first level is not at z = 0");
        else
            need(z > c.z.back(), "[z-order] height does not increase" + at_line(ln));
        need(th >= 150.0 && th <= 700.0, "[theta-range] theta outside 150..700 K" + at_line(ln));
        need(q_gkg >= 0.0 && q_gkg <= 100.0, "[qv-range] qv outside 0..100 g/kg" + at_line(ln));
        need(std::abs(u) <= 300.0 && std::abs(v) <= 300.0, "[wind-limit] a wind component beyond 300 m/s" + at_line(ln));
        if (r == 1)
        {
            // header values may be printed with fewer digits than the z = 0 row
            need(std::abs(th - ths) <= 1.0e-3, "[header-theta] header theta differs from the z = 0 level");
            need(std::abs(q_gkg - qvs_gkg) <= 1.0e-3, "[header-qv] header qv differs from the z = 0 level");
        }
        const double q = q_gkg / 1000.0;
        const double thv = th * (1.0 + kVirtualCoef * q);
        const double dz = z - z_prev;
        exner -= (kG / kCp) * dz * 0.5 * ((1.0 / thv_prev) + (1.0 / thv));
        need(std::isfinite(exner) && exner > 0.0, "[exner] Exner integration left the valid range" + at_line(ln));
        const double p = kP00 * std::pow(exner, 1.0 / kappa);
        const double T = th * exner;
        need(std::isfinite(p) && p > 0.0 && std::isfinite(T) && T > 0.0,
             "[thermo] pressure and temperature reconstruction failed" + at_line(ln));
        c.z.push_back(z);
        c.p_pa.push_back(p);
        c.p_hpa.push_back(p / 100.0);
        c.T.push_back(T);
        c.theta.push_back(th);
        c.qv.push_back(q);
        c.u.push_back(u);
        c.v.push_back(v);
        z_prev = z;
        thv_prev = thv;
    }
    // Quality control (sounding_algorithms.cpp:233-310): each level against
    // the bounds and its predecessor; a reader that would drop any level
    // refuses the file instead.
    std::size_t good = 0;
    for (std::size_t k = 0; k < c.z.size(); ++k)
    {
        bool ok = std::isfinite(c.z[k]) && std::isfinite(c.p_hpa[k]) && std::isfinite(c.T[k]);
        ok = ok && c.p_hpa[k] >= kQcPMinHpa && c.p_hpa[k] <= kQcPMaxHpa;
        ok = ok && c.T[k] >= kQcTMin && c.T[k] <= kQcTMax;
        ok = ok && std::isfinite(c.qv[k]) && c.qv[k] >= 0.0;
        ok = ok && std::isfinite(c.theta[k]) && c.theta[k] > 0.0;
        ok = ok && std::isfinite(c.u[k]) && std::isfinite(c.v[k]);
        if (k > 0) ok = ok && c.z[k] > c.z[k - 1] && c.p_hpa[k] < c.p_hpa[k - 1];
        good += ok;
    }
    if (good < 5) bad("[quality] only " + std::to_string(good) + " levels pass quality control");
    if (good != c.z.size())
        bad("[qc-shortened] quality control would drop " + std::to_string(c.z.size() - good) +
            " levels (theta outside 200-350 K, pressure outside 1-1100 hPa or not monotone); the profile is not "
            "shortened silently");
    return c;
}

void store(const Column& c, const double hdr[3], nimbus_cm1_sounding& out)
{
    const int n = int(c.z.size());
    out.n = n;
    if (out.capacity < n) bad("column capacity " + std::to_string(out.capacity) + " < " + std::to_string(n) + " levels");
    if (!out.z || !out.theta || !out.qv || !out.u || !out.v || !out.p || !out.T) bad("null output array");
    out.p_sfc = hdr[0];
    out.theta_sfc = hdr[1];
    out.qv_sfc = hdr[2];
    for (int k = 0; k < n; ++k)
    {
        out.z[k] = c.z[k];
        out.p[k] = c.p_pa[k];
        out.T[k] = c.T[k];
        out.theta[k] = c.theta[k];
        out.qv[k] = c.qv[k];
        out.u[k] = c.u[k];
        out.v[k] = c.v[k];
    }
}

// The caller's column back in SoundingData units (p / 100 is the reader's hPa
// value bit for bit when p came from the reader).
Column load(const nimbus_cm1_sounding& s)
{
    if (s.n < 2 || s.n > NIMBUS_CM1_MAX_LEVELS) bad("a column needs 2 to 100,000 levels");
    if (!s.z || !s.theta || !s.qv || !s.u || !s.v || !s.p || !s.T) bad("null column array");
    Column c;
    for (int k = 0; k < s.n; ++k)
    {
        if (k > 0 && !(s.z[k] > s.z[k - 1])) bad("heights must increase strictly");
        c.z.push_back(s.z[k]);
        c.p_hpa.push_back(s.p[k] / 100.0);
        c.T.push_back(s.T[k]);
        c.theta.push_back(s.theta[k]);
        c.qv.push_back(s.qv[k]);
        c.u.push_back(s.u[k]);
        c.v.push_back(s.v[k]);
    }
    return c;
}

double lerp(double x1, double y1, double x2, double y2, double x)
{
    if (x1 == x2) return y1;
    return y1 + (y2 - y1) * (x - x1) / (x2 - x1);
}

// Log-linear in y for positive y, else linear (sharpy_interpolation.cpp:27-42).
double log_lerp(double x1, double y1, double x2, double y2, double x)
{
    if (x1 == x2) return y1;
    if (!std::isfinite(y1) || !std::isfinite(y2) || y1 <= 0.0 || y2 <= 0.0) return lerp(x1, y1, x2, y2, x);
    return std::exp(lerp(x1, std::log(y1), x2, std::log(y2), x));
}

// Where a target height falls in the source column: below the first level
// (or on it), above the last, or inside (i2 - 1, i2] with i2 the first source
// height >= target.
struct Bracket
{
    enum Kind
    {
        Below,
        Above,
        Inside
    } kind;
    std::size_t i2;
};

Bracket bracket(const std::vector<double>& zs, double t)
{
    const std::size_t i2 = std::size_t(std::lower_bound(zs.begin(), zs.end(), t) - zs.begin());
    if (i2 == 0) return {Bracket::Below, 0};
    if (i2 == zs.size()) return {Bracket::Above, i2};
    return {Bracket::Inside, i2};
}

// One series, piecewise linear, with the endpoint rules of
// interpolate_sounding_linear (sounding_algorithms.cpp:538-636).
double linear_at(const std::vector<double>& zs, const std::vector<double>& ys, double t, const nimbus_interp& how,
                 bool log_space = false)
{
    const std::size_t n = zs.size();
    auto f = [&](std::size_t a, std::size_t b) {
        return log_space ? log_lerp(zs[a], ys[a], zs[b], ys[b], t) : lerp(zs[a], ys[a], zs[b], ys[b], t);
    };
    const Bracket br = bracket(zs, t);
    if (br.kind == Bracket::Below) return how.extrapolate_below && n > 1 ? f(0, 1) : ys[0];
    if (br.kind == Bracket::Above) return how.extrapolate_above && n > 1 ? f(n - 2, n - 1) : ys[n - 1];
    return f(br.i2 - 1, br.i2);
}

// Monotone PCHIP (Fritsch-Carlson weighted harmonic mean of the adjacent
// secants, zero at extrema; one-sided three-point end slopes limited to the
// secant's sign and three times its size), evaluated as a cubic Hermite
// segment (sharpy_interpolation.cpp:108-318).
double pchip_end_slope(double h0, double h1, double d0, double d1)
{
    const double den = h0 + h1;
    if (!std::isfinite(den) || den == 0.0) return d0;
    double m = ((2.0 * h0 + h1) * d0 - h0 * d1) / den;
    if (m * d0 <= 0.0)
        m = 0.0;
    else if (std::abs(m) > 3.0 * std::abs(d0))
        m = 3.0 * d0;
    return m;
}

std::vector<double> pchip_slopes(const std::vector<double>& x, const std::vector<double>& y)
{
    const std::size_t n = x.size();
    std::vector<double> m(n, 0.0);
    if (n < 2) return m;
    std::vector<double> h(n - 1), d(n - 1);
    for (std::size_t i = 0; i + 1 < n; ++i)
    {
        h[i] = x[i + 1] - x[i];
        d[i] = (y[i + 1] - y[i]) / h[i];
    }
    if (n == 2)
    {
        m[0] = m[1] = d[0];
        return m;
    }
    m[0] = pchip_end_slope(h[0], h[1], d[0], d[1]);
    m[n - 1] = pchip_end_slope(h[n - 2], h[n - 3], d[n - 2], d[n - 3]);
    for (std::size_t i = 1; i + 1 < n; ++i)
    {
        if (d[i - 1] == 0.0 || d[i] == 0.0 || d[i - 1] * d[i] < 0.0) continue;
        const double w1 = 2.0 * h[i] + h[i - 1], w2 = h[i] + 2.0 * h[i - 1];
        m[i] = (w1 + w2) / ((w1 / d[i - 1]) + (w2 / d[i]));
    }
    return m;
}

double hermite(double x0, double x1, double y0, double y1, double m0, double m1, double x)
{
    const double h = x1 - x0;
    if (h == 0.0) return y0;
    const double t = (x - x0) / h, t2 = t * t, t3 = t2 * t;
    const double h00 = (2.0 * t3) - (3.0 * t2) + 1.0;
    const double h10 = t3 - (2.0 * t2) + t;
    const double h01 = (-2.0 * t3) + (3.0 * t2);
    const double h11 = t3 - t2;
    return (h00 * y0) + (h10 * h * m0) + (h01 * y1) + (h11 * h * m1);
}

double spline_at(const std::vector<double>& zs, const std::vector<double>& ys, const std::vector<double>& m,
                 bool use_spline, double t, const nimbus_interp& how)
{
    const std::size_t n = zs.size();
    const Bracket br = bracket(zs, t);
    if (br.kind == Bracket::Below) return how.extrapolate_below ? lerp(zs[0], ys[0], zs[1], ys[1], t) : ys[0];
    if (br.kind == Bracket::Above)
        return how.extrapolate_above ? lerp(zs[n - 2], ys[n - 2], zs[n - 1], ys[n - 1], t) : ys[n - 1];
    const std::size_t i2 = br.i2, i1 = i2 - 1;
    if (zs[i2] == t) return ys[i2];
    if (use_spline) return hermite(zs[i1], zs[i2], ys[i1], ys[i2], m[i1], m[i2], t);
    return lerp(zs[i1], ys[i1], zs[i2], ys[i2], t);
}

std::vector<double> spline_series(const std::vector<double>& zs, const std::vector<double>& ys,
                                  const std::vector<double>& zt, const nimbus_interp& how)
{
    bool ok = zs.size() >= 2;
    for (std::size_t i = 0; ok && i < zs.size(); ++i)
        ok = std::isfinite(zs[i]) && std::isfinite(ys[i]) && (i == 0 || zs[i] > zs[i - 1]);
    const std::vector<double> m = ok ? pchip_slopes(zs, ys) : std::vector<double>(zs.size(), 0.0);
    std::vector<double> out(zt.size());
    for (std::size_t i = 0; i < zt.size(); ++i) out[i] = spline_at(zs, ys, m, ok, zt[i], how);
    return out;
}

// The sounding interpolated to the model heights: p (hPa), T, theta, qv, u, v.
Column interpolate(const Column& s, const std::vector<double>& zt, const nimbus_interp& how)
{
    Column r;
    r.z = zt;
    const std::vector<const std::vector<double>*> src = {&s.p_hpa, &s.T, &s.theta, &s.qv, &s.u, &s.v};
    std::vector<std::vector<double>*> dst = {&r.p_hpa, &r.T, &r.theta, &r.qv, &r.u, &r.v};
    for (std::size_t f = 0; f < src.size(); ++f)
    {
        if (how.method == NIMBUS_INTERP_SPLINE)
        {
            *dst[f] = spline_series(s.z, *src[f], zt, how);
            continue;
        }
        dst[f]->resize(zt.size());
        for (std::size_t i = 0; i < zt.size(); ++i) (*dst[f])[i] = linear_at(s.z, *src[f], zt[i], how);
    }
    if (how.method == NIMBUS_INTERP_LOG_PRESSURE)
        for (std::size_t i = 0; i < zt.size(); ++i) r.p_hpa[i] = linear_at(s.z, s.p_hpa, zt[i], how, true);
    return r;
}

struct Outputs
{
    double *rho0, *p0, *theta0, *qv0, *u0, *v0, *T0;
};

// FileSoundingSource::build + sounding_from_data for a theta-primary column.
void column(const Column& s, const nimbus_interp& how, const std::vector<double>& zt, double dz, const Outputs& o)
{
    const std::size_t nz = zt.size();
    if (nz == 0) bad("target model-height column is empty");
    if (!std::isfinite(dz) || dz <= 0.0) bad("target vertical spacing must be positive and finite");
    if (!std::isfinite(zt[0]) || zt[0] < 0.0) bad("target model heights must be finite and non-negative");
    for (std::size_t k = 1; k < nz; ++k)
        if (!std::isfinite(zt[k]) || zt[k] <= zt[k - 1]) bad("target model heights must increase strictly");
    if (how.method < NIMBUS_INTERP_LINEAR || how.method > NIMBUS_INTERP_LOG_PRESSURE)
        bad("unsupported sounding interpolation method");
    // the model's top interface must lie inside the sounding
    const double half = nz > 1 ? 0.5 * (zt[nz - 1] - zt[nz - 2]) : 0.5 * dz;
    const double top = zt[nz - 1] + half;
    const double tol = 1.0e-10 * std::max({1.0, std::abs(top), std::abs(s.z.back())});
    if (top > s.z.back() + tol && !how.extrapolate_above)
        bad("[top-coverage] the model's top interface at " + std::to_string(top) + " m (highest centre " +
            std::to_string(zt[nz - 1]) + " m) lies above the sounding's top at " + std::to_string(s.z.back()) +
            " m, and extrapolation above is off");
    const Column r = interpolate(s, zt, how);
    const double kappa = kRd / kCp;
    // theta stays exact; d(pi)/dz = -g / (cp theta) by the trapezoid rule from
    // the interpolated pressure at the first centre
    if (!std::isfinite(r.p_hpa[0]) || r.p_hpa[0] <= 0.0) bad("first-center pressure must be positive and finite");
    for (std::size_t k = 0; k < nz; ++k)
        if (!std::isfinite(r.theta[k]) || r.theta[k] <= 0.0) bad("interpolated theta column is non-physical");
    double exner = std::pow(r.p_hpa[0] * 100.0 / kP00, kappa);
    if (!std::isfinite(exner) || exner <= 0.0) bad("surface Exner pressure is invalid");
    for (std::size_t k = 0; k < nz; ++k)
    {
        if (k > 0)
        {
            const double dzl = zt[k] - zt[k - 1];
            exner -= (kG / kCp) * dzl * 0.5 * ((1.0 / r.theta[k - 1]) + (1.0 / r.theta[k]));
            if (!std::isfinite(exner) || exner <= 0.0) bad("hydrostatic Exner pressure became non-positive");
        }
        const double th = r.theta[k];
        const double T = th * exner;
        const double p = kP00 * std::pow(exner, 1.0 / kappa);
        const double rho = p / (kRd * T);
        const double q = r.qv[k];
        if (!std::isfinite(q) || q < 0.0) bad("every level requires a finite non-negative mixing ratio");
        if (!std::isfinite(r.u[k]) || !std::isfinite(r.v[k])) bad("winds must be finite at every level");
        // the reference's self-consistency check (init/sounding/sounding_source.cpp:34-70)
        const double th_chk = T * std::pow(kP00 / p, kappa);
        if (!(std::abs(th - th_chk) / std::max(th_chk, 1.0) < 5.0e-3) || !(rho > 0.0) || !std::isfinite(rho))
            bad("column inconsistent at level " + std::to_string(k));
        o.theta0[k] = th;
        o.p0[k] = p;
        o.rho0[k] = rho;
        o.qv0[k] = q;
        o.u0[k] = r.u[k];
        o.v0[k] = r.v[k];
        if (o.T0) o.T0[k] = T;
    }
}

// Round-trip text for a double (shortest %.17g is always exact).
std::string num(double x)
{
    char b[40];
    std::snprintf(b, sizeof b, "%.17g", x);
    return b;
}

template <class F>
int guarded(F&& f)
{
    try
    {
        f();
        return NIMBUS_OK;
    }
    catch (const Bad& e)
    {
        return set_error(NIMBUS_EINVAL, e.what());
    }
    catch (const std::exception& e)
    {
        return set_error(NIMBUS_EINVAL, std::string("CM1 input_sounding: ") + e.what());
    }
}

}  // namespace
}  // namespace nimbus

extern "C" {

int nimbus_cm1_parse(const char* text, size_t len, nimbus_cm1_sounding* out)
{
    using namespace nimbus;
    if (!text || !out) return set_error(NIMBUS_EINVAL, "CM1 input_sounding: null argument");
    return guarded([&] {
        double hdr[3];
        const Column c = parse(text, len, hdr);
        store(c, hdr, *out);
    });
}

int nimbus_cm1_read(const char* path, nimbus_cm1_sounding* out)
{
    using namespace nimbus;
    if (!path || !out) return set_error(NIMBUS_EINVAL, "CM1 input_sounding: null argument");
    struct stat st;
    if (::stat(path, &st) != 0 || !S_ISREG(st.st_mode))
        return set_error(NIMBUS_EINVAL, std::string("CM1 input_sounding is not a regular file: ") + path);
    if (std::uint64_t(st.st_size) > NIMBUS_CM1_MAX_BYTES)
        return set_error(NIMBUS_EINVAL, "CM1 input_sounding exceeds the 8 MiB safety limit");
    std::ifstream f(path, std::ios::binary);
    if (!f) return set_error(NIMBUS_EINVAL, std::string("cannot open CM1 input_sounding: ") + path);
    const std::string text((std::istreambuf_iterator<char>(f)), std::istreambuf_iterator<char>());
    return nimbus_cm1_parse(text.data(), text.size(), out);
}

int nimbus_cm1_format(const nimbus_cm1_sounding* s, char* buf, size_t cap, size_t* len)
{
    using namespace nimbus;
    if (!s || !len || s->n < 1 || !s->z || !s->theta || !s->qv || !s->u || !s->v)
        return set_error(NIMBUS_EINVAL, "CM1 format: null argument or empty column");
    std::string t = num(s->p_sfc / 100.0) + " " + num(s->theta_sfc) + " " + num(s->qv_sfc * 1000.0) + "\n";
    for (int k = 0; k < s->n; ++k)
        t += num(s->z[k]) + " " + num(s->theta[k]) + " " + num(s->qv[k] * 1000.0) + " " + num(s->u[k]) + " " +
             num(s->v[k]) + "\n";
    *len = t.size();
    if (!buf || cap < t.size() + 1)
        return set_error(NIMBUS_EINVAL, "CM1 format: buffer of " + std::to_string(cap) + " bytes < " +
                                            std::to_string(t.size() + 1) + " needed");
    std::memcpy(buf, t.c_str(), t.size() + 1);
    return NIMBUS_OK;
}

int nimbus_cm1_column(const nimbus_cm1_sounding* s, const nimbus_interp* how, int nz, const double* z, double dz,
                      double* rho0, double* p0, double* theta0, double* qv0, double* u0, double* v0, double* T0)
{
    using namespace nimbus;
    if (!s || nz < 0 || (nz > 0 && !z) || !rho0 || !p0 || !theta0 || !qv0 || !u0 || !v0)
        return set_error(NIMBUS_EINVAL, "CM1 column: null argument");
    const nimbus_interp dflt{NIMBUS_INTERP_LINEAR, 0, 0};
    return guarded([&] {
        const Column c = load(*s);
        column(c, how ? *how : dflt, std::vector<double>(z, z + nz), dz, {rho0, p0, theta0, qv0, u0, v0, T0});
    });
}

int nimbus_cm1_base_state(const nimbus_grid* g, const nimbus_cm1_sounding* s, const nimbus_interp* how, double* rho0,
                          double* p0, double* theta0, double* qv0, double* u0, double* v0)
{
    using namespace nimbus;
    if (!g || g->nz < 1 || !(g->dz > 0.0)) return set_error(NIMBUS_EINVAL, "CM1 base state: invalid grid");
    std::vector<double> z(std::size_t(g->nz));
    for (int k = 0; k < g->nz; ++k) z[k] = (k + 0.5) * g->dz;
    return nimbus_cm1_column(s, how, g->nz, z.data(), g->dz, rho0, p0, theta0, qv0, u0, v0, nullptr);
}

}  // extern "C"
