// CPU oracle for the Klemp-Wilhelmson split-explicit large step.
//
// TEST INFRASTRUCTURE ONLY. This file is a plain C++ restatement of the
// Nimbus reference CPU path for one large step of the Cartesian C-grid
// dynamical core. It is used by tests/, by __graft_entry__.smoke() and by the
// cpu_baseline leg of bench.py as the checker and as the CPU baseline; the
// shipped HIP backend (nimbus_amd/) never links, loads or calls it.
//
// Layout mirrors the reference host storage: every 3-D field is row-major
// [nx][ny][nz] with k (vertical) contiguous, index (i*ny + j)*nz + k
// (include/core/field/field3d.hpp:543-574). The C-grid placement follows
// src/boundary_conditions/boundary_conditions_cartesian_cgrid.cpp:11-24:
// u[i] at x_{i+1/2}, v[j] at y_{j+1/2}, w[k] at z_{k+1/2}, scalars at cell
// centres, one ghost column on each lateral side.
//
// The template parameter S is the storage type. S = float reproduces the
// reference CPU path (FP32 storage, FP64 arithmetic, FP32 rounding at every
// store). S = double is the all-FP64 oracle. Arithmetic is FP64 in both.
//
// Parity pins (see tests/): every function below is checked against the
// reference's own known-answer tests, against compiled reference
// translation units (oracle/_ref, built by oracle/Makefile from the
// reference sources where they lie) and against closed forms / invariants.
// The HE-VI vertically implicit acoustic solve (acoustic_substep_hevi) has
// no reference counterpart; it is pinned by closed forms (rest state, beta=0
// structure, dense-solve brute force on tiny columns, conservation).
#pragma once

#include <algorithm>
#include <cmath>
#include <cstddef>
#include <cstdint>
#include <limits>
#include <stdexcept>
#include <vector>

namespace oracle
{

// ---- constants: include/core/infra/physical_constants.hpp:14-34,
// include/microphysics/microphysics_base.hpp:1071-1081,
// include/core/field/field_bounds.hpp:10-20 ----
constexpr double kG = 9.81;
constexpr double kRd = 287.0;
constexpr double kRv = 461.5;
constexpr double kCp = 1004.0;
constexpr double kLv = 2.5e6;
constexpr double kP00 = 100000.0;
constexpr double kT0 = 273.15;
constexpr double kGamma = kCp / (kCp - kRd);
constexpr double kKappa = kRd / kCp;
constexpr float kThetaMin = 200.0f, kThetaMax = 700.0f;
constexpr float kPMin = 1000.0f, kPMax = 200000.0f;
constexpr float kRhoMin = 0.05f, kRhoMax = 2.0f;
constexpr float kWindH = 180.0f, kWindV = 120.0f;
constexpr float kQvMax = 0.05f, kQhMax = 0.05f;
constexpr double kCStar = 30.0;  // boundary_conditions_cartesian_cgrid.cpp:293

enum LateralBC : int { kPeriodic = 0, kRadiative = 1, kZeroGradient = 2, kRigid = 3, kRigidSlip = 4 };
enum BottomBC : int { kBottomFreeSlip = 0, kBottomNoSlip = 1 };
enum TopBC : int { kTopRigid = 0, kTopOpen = 1 };
enum AcousticMode : int { kAcousticExplicit = 0, kAcousticHEVI = 1 };

struct Grid
{
    int nx = 0, ny = 0, nz = 0;
    double dx = 0, dy = 0, dz = 0;
    std::size_t cells() const { return std::size_t(nx) * ny * nz; }
    std::size_t at(int i, int j, int k) const
    {
        return (std::size_t(i) * std::size_t(ny) + std::size_t(j)) * std::size_t(nz) + std::size_t(k);
    }
};

// FP64 base state at cell centres (include/core/runtime/simulation_state.hpp:104-111).
struct Base
{
    std::vector<double> rho0, p0, theta0, qv0, u0, v0;
};

enum SgsMode
{
    kSgs3d = 0,
    kSgsHorizontal = 1,  // "2d"
    kSgsVertical = 2
};
enum SgsFilter
{
    kFilterCubicRoot = 0,  // (dx dy dz)^(1/3)
    kFilterDx = 1          // sqrt(dx dy)
};

struct Config
{
    double dt = 1.0;
    int substeps = 6;           // requested minimum (split_explicit.cpp:127-137)
    int acoustic_mode = kAcousticExplicit;
    double hevi_beta = 0.1;     // off-centering of the implicit vertical step
    double coriolis_f = 0.0;
    int bc_west = kRadiative, bc_east = kRadiative, bc_south = kRadiative, bc_north = kRadiative;
    int bc_bottom = kBottomFreeSlip, bc_top = kTopRigid;
    // Kessler parameters (kessler.hpp:40-46)
    double qc0 = 1.0e-3, c_auto = 1.0e-3, c_accr = 2.2, evap_eff = 1.0, vt_max = 20.0;
    // Rayleigh sponges (runtime_config.hpp:114-122; kessler presets' damping block)
    int sponge = 0;
    double sponge_z_d = 12000.0, sponge_alpha = 0.05;
    int sponge_w_cells = 8;
    double sponge_w_alpha = 0.2;
    int sponge_lateral_cells = 0;
    double sponge_lateral_alpha = 0.05;
    // Smagorinsky SGS (include/turbulence/turbulence_base.hpp:50-70; presets'
    // turbulence block: smagorinsky, mode 2d = horizontal_only, Cs 0.18, Pr_t = Sc_t = 1/3)
    int sgs = 0;
    int sgs_mode = kSgs3d;
    int sgs_filter = kFilterCubicRoot;
    double sgs_cs = 0.18, sgs_pr_t = 1.0 / 3.0, sgs_sc_t = 1.0 / 3.0;
    double sgs_nu_max = 1000.0, sgs_k_max = 1000.0, sgs_dt = 1.0;
    // explicit Laplacian diffusion after the SGS (diffusion_step.cpp:679-1030;
    // DiffusionConfig, include/numerics/diffusion/diffusion_base.hpp:23-40):
    // 0 off, 1 scalars (theta, qv), 2 momentum, 3 both; K_h, K_v in m^2/s
    int diffusion = 0;
    double diffusion_kh = 0.0, diffusion_kv = 0.0;
};

template <class S>
struct State
{
    std::vector<S> u, v, w, rho, p, theta, qv, qc, qr;
    std::vector<double> rain_mm;  // accumulated surface rain per column [nx*ny]
    // KW78 radiative-face stashes (boundary_conditions_cartesian_cgrid.cpp:798-823)
    std::vector<float> rad_w, rad_e, rad_s, rad_n;
    void allocate(const Grid& g)
    {
        for (auto* f : {&u, &v, &w, &rho, &p, &theta, &qv, &qc, &qr}) f->assign(g.cells(), S(0));
        rain_mm.assign(std::size_t(g.nx) * g.ny, 0.0);
    }
};

// ======================================================================
// Thermodynamics (src/microphysics/base/thermodynamics.hpp:29-470)
// ======================================================================

inline double theta_to_temperature(double theta, double p)
{
    // thermodynamics.hpp:29-45
    if (!std::isfinite(theta) || !std::isfinite(p)) return kT0;
    const double th = std::clamp(theta, 150.0, 700.0);
    const double pp = std::clamp(p, 100.0, 200000.0);
    const double T = th * std::pow(pp / kP00, kKappa);
    return std::isfinite(T) ? T : kT0;
}

inline double theta_tendency_from_temperature(double dT, double p)
{
    // thermodynamics.hpp:225-238
    if (!std::isfinite(dT) || !std::isfinite(p)) return 0.0;
    const double pp = std::clamp(p, 100.0, 200000.0);
    const double r = dT * std::pow(kP00 / pp, kKappa);
    return std::isfinite(r) ? r : 0.0;
}

inline double buck_es(double T)
{
    // thermodynamics.hpp:67-88 (Buck water equation at all temperatures)
    if (!std::isfinite(T)) return 0.0;
    const double Tc = std::clamp(T, 150.0, 400.0) - kT0;
    const double es = 611.21 * std::exp((18.678 - Tc / 234.5) * Tc / (257.14 + Tc));
    return std::isfinite(es) ? std::max(0.0, es) : 0.0;
}

inline double qvs_water(double T, double p)
{
    // thermodynamics.hpp:104-118
    const double es = buck_es(T);
    if (!std::isfinite(es) || !std::isfinite(p)) return 0.0;
    const double pp = std::clamp(p, 100.0, 200000.0);
    const double q = 0.62198 * es / std::max(pp - es, 1.0);
    return std::clamp(q, 0.0, 0.2);
}

// d(qvs)/dT along the Buck curve, evaluated from (T, qvs) as the reference's
// Newton slope (thermodynamics.hpp:300-325).
inline double qvs_slope(double T, double qs, double p)
{
    if (!(T > 150.0 && T < 400.0) || !(qs > 0.0 && qs < 0.2)) return 0.0;
    const double pp = std::clamp(p, 100.0, 200000.0);
    const double e = qs * pp / (0.62198 + qs);
    const double Tc = T - kT0;
    const double den = 257.14 + Tc;
    const double num = (18.678 - Tc / 234.5) * Tc;
    const double dnum = 18.678 - 2.0 * Tc / 234.5;
    const double dlog = (dnum * den - num) / (den * den);
    return qs * pp / std::max(pp - e, 1.0) * dlog;
}

// Saturation adjustment: qv' = qvs(T'), T' = T + (Lv/cp)(qv - qv'), solved by a
// bracketed Newton iteration of at most 10 steps (thermodynamics.hpp:282-470).
inline double saturation_adjust(double T_in, double p, double& qv, double& qc)
{
    constexpr double L = kLv / kCp;
    constexpr int iters = 10;
    double T = std::isfinite(T_in) ? std::clamp(T_in, 150.0, 400.0) : kT0;
    qv = std::isfinite(qv) ? std::max(0.0, qv) : 0.0;
    qc = std::isfinite(qc) ? std::max(0.0, qc) : 0.0;
    double qs = qvs_water(T, p);
    if (!std::isfinite(qs) || qs < 0.0) qs = 0.0;
    const double tol_eps = 8.0 * std::numeric_limits<double>::epsilon();
    if (qv > qs)
    {
        const double cmax = std::min(qv, std::max(0.0, (400.0 - T) / L));
        double c = cmax;
        const double r_at_max = qv - cmax - qvs_water(T + L * cmax, p);
        if (r_at_max <= 0.0)
        {
            double lo = 0.0, hi = cmax;
            c = std::clamp((qv - qs) / (1.0 + L * qvs_slope(T, qs, p)), lo, hi);
            for (int it = 0; it < iters; ++it)
            {
                const double Ta = T + L * c;
                const double qa = qvs_water(Ta, p);
                const double f = qv - c - qa;
                if (std::abs(f) <= tol_eps * std::max({1.0, std::abs(qv), std::abs(qa)})) break;
                if (f > 0.0) lo = c; else hi = c;
                const double cand = c - f / (-1.0 - L * qvs_slope(Ta, qa, p));
                c = (cand > lo && cand < hi) ? cand : 0.5 * (lo + hi);
            }
        }
        qv -= c;
        qc += c;
        T += L * c;
    }
    else if (qv < qs && qc > 0.0)
    {
        const double emax = std::min(qc, std::max(0.0, (T - 150.0) / L));
        double e = emax;
        const double r_at_max = qv + emax - qvs_water(T - L * emax, p);
        if (r_at_max >= 0.0)
        {
            double lo = 0.0, hi = emax;
            e = std::clamp((qs - qv) / (1.0 + L * qvs_slope(T, qs, p)), lo, hi);
            for (int it = 0; it < iters; ++it)
            {
                const double Ta = T - L * e;
                const double qa = qvs_water(Ta, p);
                const double f = qv + e - qa;
                if (std::abs(f) <= tol_eps * std::max({1.0, std::abs(qv), std::abs(qa)})) break;
                if (f < 0.0) lo = e; else hi = e;
                const double cand = e - f / (1.0 + L * qvs_slope(Ta, qa, p));
                e = (cand > lo && cand < hi) ? cand : 0.5 * (lo + hi);
            }
        }
        qc -= e;
        qv += e;
        T -= L * e;
    }
    qc = std::max(0.0, qc);
    qv = std::max(0.0, qv);
    return T;
}

// Kessler rain speed without the density correction (fall_speeds.hpp:125-130).
inline double kessler_rain_speed(double q, double rho)
{
    if (q < 1.0e-9) return 0.0;
    return 36.34 * std::pow(1.0e-3 * rho * q, 0.1346);
}

// Dry rho-theta EOS evaluated in FP64 (include/dynamics/eos_pressure.hpp:30-38).
inline double eos_pressure(double rho, double theta)
{
    return kP00 * std::pow(rho * kRd * theta / kP00, kGamma);
}

inline float sanitize(float v, float fallback, float lo, float hi)
{
    // field_bounds.hpp:41-53 (repairs enabled)
    if (!std::isfinite(v)) return fallback;
    return std::clamp(v, lo, hi);
}

// ======================================================================
// WENO5-JS face value (include/numerics/advection/weno5_face.hpp:43-86)
// and the MC-limited slope used near walls (advection_cartesian.cpp:135-162)
// ======================================================================

inline double weno5_face(double a, double b, double c, double d, double e)
{
    // candidate polynomials at the downwind face of cell c (flow c -> d)
    const double q0 = (1.0 / 3.0) * a - (7.0 / 6.0) * b + (11.0 / 6.0) * c;
    const double q1 = -(1.0 / 6.0) * b + (5.0 / 6.0) * c + (1.0 / 3.0) * d;
    const double q2 = (1.0 / 3.0) * c + (5.0 / 6.0) * d - (1.0 / 6.0) * e;
    // Jiang-Shu smoothness indicators
    const double t0a = a - 2.0 * b + c, t0b = a - 4.0 * b + 3.0 * c;
    const double t1a = b - 2.0 * c + d, t1b = b - d;
    const double t2a = c - 2.0 * d + e, t2b = 3.0 * c - 4.0 * d + e;
    const double s0 = 13.0 / 12.0 * t0a * t0a + 0.25 * t0b * t0b;
    const double s1 = 13.0 / 12.0 * t1a * t1a + 0.25 * t1b * t1b;
    const double s2 = 13.0 / 12.0 * t2a * t2a + 0.25 * t2b * t2b;
    // epsilon relative to the stencil magnitude squared plus a tiny floor
    const double m2 = std::max(std::max(std::max(a * a, b * b), std::max(c * c, d * d)), e * e);
    const double eps = 1.0e-6 * m2 + 1.0e-40;
    const double w0 = 0.1 / ((eps + s0) * (eps + s0));
    const double w1 = 0.6 / ((eps + s1) * (eps + s1));
    const double w2 = 0.3 / ((eps + s2) * (eps + s2));
    const double ws = w0 + w1 + w2;
    if (!(ws > 0.0) || !std::isfinite(ws)) return c;
    const double f = (w0 * q0 + w1 * q1 + w2 * q2) / ws;
    return std::isfinite(f) ? f : c;
}

inline double minmod3(double a, double b, double c)
{
    if (a > 0.0 && b > 0.0 && c > 0.0) return std::min(a, std::min(b, c));
    if (a < 0.0 && b < 0.0 && c < 0.0) return std::max(a, std::max(b, c));
    return 0.0;
}

inline double mc_slope(double qm, double q0, double qp)
{
    const double r = qp - q0, l = q0 - qm;
    return minmod3(0.5 * (r + l), 2.0 * r, 2.0 * l);
}

// Conservative PD-limiter storage bound (include/numerics/advection/advection_cartesian.hpp:101-108).
inline float pd_limiter_store(double ratio, bool active_at_one)
{
    if (!(ratio > 0.0 && (ratio < 1.0 || (active_at_one && ratio == 1.0)))) return float(ratio);
    constexpr float shrink = 1.0f - 32.0f * std::numeric_limits<float>::epsilon();
    return std::nextafter(float(ratio) * shrink, 0.0f);
}

// ======================================================================
// UP5/UP3/UP1 upwind-biased first derivatives
// (include/numerics/derivatives/derivative_operators.hpp:324-401)
// ======================================================================

inline double upwind_fd1(double m1, double c, double p1, double a, double inv)
{
    return a >= 0.0 ? (c - m1) * inv : (p1 - c) * inv;
}
inline double upwind_fd3(double m2, double m1, double c, double p1, double p2, double a, double inv)
{
    return a >= 0.0 ? (2.0 * p1 + 3.0 * c - 6.0 * m1 + m2) * (inv / 6.0)
                    : (-p2 + 6.0 * p1 - 3.0 * c - 2.0 * m1) * (inv / 6.0);
}
inline double upwind_fd5(double m3, double m2, double m1, double c, double p1, double p2, double p3,
                         double a, double inv)
{
    return a >= 0.0 ? (-2.0 * m3 + 15.0 * m2 - 60.0 * m1 + 20.0 * c + 30.0 * p1 - 3.0 * p2) * (inv / 60.0)
                    : (2.0 * p3 - 15.0 * p2 + 60.0 * p1 - 20.0 * c - 30.0 * m1 + 3.0 * m2) * (inv / 60.0);
}

// Upwind derivative along one axis: `at(o)` returns f at offset o from the
// centre, `pos` the index along the axis and `n` its extent.
template <class F>
inline double upwind_axis(F at, int pos, int n, double a, double inv)
{
    const int half = std::min(pos, n - 1 - pos);
    if (half <= 0) return 0.0;
    const double c = at(0), m1 = at(-1), p1 = at(1);
    if (half == 1) return upwind_fd1(m1, c, p1, a, inv);
    const double m2 = at(-2), p2 = at(2);
    if (half == 2) return upwind_fd3(m2, m1, c, p1, p2, a, inv);
    return upwind_fd5(at(-3), m2, m1, c, p1, p2, at(3), a, inv);
}

// ======================================================================
// Large-step pieces
// ======================================================================

inline int resolve_substeps(const Grid& g, const Base& b, const Config& cfg)
{
    // split_explicit.cpp:67-156 with c_max from the base state (runtime.cpp:860-878)
    double cmax = 0.0;
    for (int k = 0; k < g.nz; ++k) cmax = std::max(cmax, std::sqrt(kGamma * b.p0[k] / b.rho0[k]));
    double dmin = g.dz;
    if (cfg.acoustic_mode == kAcousticHEVI)
        dmin = std::min(g.dx, g.dy);  // the vertical acoustic mode is implicit
    else
        dmin = std::min({g.dx, g.dy, g.dz});
    const int required = std::max(int(std::ceil(cfg.dt * cmax / (0.5 * dmin))) + 1, 2);
    return cfg.substeps > 0 ? std::max(cfg.substeps, required) : required;
}

// C-grid lateral + vertical boundary closure with optional KW78 radiation
// (boundary_conditions_cartesian_cgrid.cpp:798-1175).
template <class S>
void seed_radiation(const Grid& g, State<S>& s)
{
    const int NX = g.nx, NY = g.ny, NZ = g.nz;
    s.rad_w.resize(std::size_t(NY) * NZ);
    s.rad_e.resize(std::size_t(NY) * NZ);
    s.rad_s.resize(std::size_t(NX) * NZ);
    s.rad_n.resize(std::size_t(NX) * NZ);
    for (int j = 0; j < NY; ++j)
        for (int k = 0; k < NZ; ++k)
        {
            s.rad_w[std::size_t(j) * NZ + k] = float(s.u[g.at(0, j, k)]);
            s.rad_e[std::size_t(j) * NZ + k] = float(s.u[g.at(NX - 2, j, k)]);
        }
    for (int i = 0; i < NX; ++i)
        for (int k = 0; k < NZ; ++k)
        {
            s.rad_s[std::size_t(i) * NZ + k] = float(s.v[g.at(i, 0, k)]);
            s.rad_n[std::size_t(i) * NZ + k] = float(s.v[g.at(i, NY - 2, k)]);
        }
}

inline float kw78_face(float prev, float phase, float grad, float dt, float inv_d, bool west_like)
{
    // include/dynamics/radiative_face.hpp:5-15 (four FP32 operations)
    if (west_like ? phase >= 0.0f : phase <= 0.0f) return prev;
    volatile float a = dt * phase;
    volatile float b = a * grad;
    volatile float c = b * inv_d;
    return prev - c;
}

template <class S>
void apply_boundaries(const Grid& g, const Config& cfg, State<S>& s, double dt_rad)
{
    const int NX = g.nx, NY = g.ny, NZ = g.nz;
    if (s.rad_w.size() != std::size_t(NY) * NZ || s.rad_s.size() != std::size_t(NX) * NZ) seed_radiation(g, s);
    const bool radiate = dt_rad > 0.0;
    const float dtf = float(dt_rad);
    const float inv_dx = 1.0f / float(g.dx), inv_dy = 1.0f / float(g.dy);
    const float cstar = float(kCStar);
    auto& u = s.u; auto& v = s.v; auto& w = s.w;
    auto copy_rp = [&](std::size_t d, std::size_t src) { s.rho[d] = s.rho[src]; s.p[d] = s.p[src]; };
    // west (i = 0 ghost, wall face u[0])
    for (int j = 0; j < NY; ++j)
        for (int k = 0; k < NZ; ++k)
        {
            const std::size_t c = g.at(0, j, k);
            const int bc = cfg.bc_west;
            if (bc == kPeriodic)
            {
                const std::size_t sidx = g.at(NX - 2, j, k);
                u[c] = u[sidx]; v[c] = v[sidx]; w[c] = w[sidx]; copy_rp(c, sidx);
                continue;
            }
            const std::size_t in = g.at(1, j, k);
            if (bc == kRadiative)
            {
                if (radiate)
                {
                    float& up = s.rad_w[std::size_t(j) * NZ + k];
                    const float un = kw78_face(up, up - 0.0f - cstar, float(u[in]) - up, dtf, inv_dx, true);
                    u[c] = S(un);
                    up = un;
                }
            }
            else if (bc == kZeroGradient) u[c] = u[in];
            else u[c] = S(0);
            if (bc == kRigid) { v[c] = S(0); w[c] = S(0); }
            else { v[c] = v[in]; w[c] = w[in]; }
            copy_rp(c, in);
        }
    // east (i = NX-1 ghost, wall face u[NX-2])
    for (int j = 0; j < NY; ++j)
        for (int k = 0; k < NZ; ++k)
        {
            const int bc = cfg.bc_east;
            const std::size_t gi = g.at(NX - 1, j, k);
            if (bc == kPeriodic)
            {
                const std::size_t sidx = g.at(1, j, k);
                u[gi] = u[sidx]; v[gi] = v[sidx]; w[gi] = w[sidx]; copy_rp(gi, sidx);
                continue;
            }
            const std::size_t a = g.at(NX - 2, j, k), am = g.at(NX - 3, j, k);
            if (bc == kRigid || bc == kRigidSlip) u[a] = S(0);
            if (bc == kRadiative && radiate)
            {
                float& up = s.rad_e[std::size_t(j) * NZ + k];
                const float un = kw78_face(up, up - 0.0f + cstar, up - float(u[am]), dtf, inv_dx, false);
                u[a] = S(un);
                up = un;
            }
            else if (bc == kZeroGradient) u[a] = u[am];
            u[gi] = (bc == kRadiative || bc == kZeroGradient) ? u[a] : S(0);
            if (bc == kRigid) { v[gi] = S(0); w[gi] = S(0); }
            else { v[gi] = v[a]; w[gi] = w[a]; }
            copy_rp(gi, a);
        }
    // south (j = 0 ghost, wall face v[i][0])
    for (int i = 0; i < NX; ++i)
        for (int k = 0; k < NZ; ++k)
        {
            const std::size_t c = g.at(i, 0, k);
            const int bc = cfg.bc_south;
            if (bc == kPeriodic)
            {
                const std::size_t sidx = g.at(i, NY - 2, k);
                v[c] = v[sidx]; u[c] = u[sidx]; w[c] = w[sidx]; copy_rp(c, sidx);
                continue;
            }
            const std::size_t in = g.at(i, 1, k);
            if (bc == kRadiative)
            {
                if (radiate)
                {
                    float& vp = s.rad_s[std::size_t(i) * NZ + k];
                    const float vn = kw78_face(vp, vp - 0.0f - cstar, float(v[in]) - vp, dtf, inv_dy, true);
                    v[c] = S(vn);
                    vp = vn;
                }
            }
            else if (bc == kZeroGradient) v[c] = v[in];
            else v[c] = S(0);
            if (bc == kRigid) { u[c] = S(0); w[c] = S(0); }
            else { u[c] = u[in]; w[c] = w[in]; }
            copy_rp(c, in);
        }
    // north (j = NY-1 ghost, wall face v[i][NY-2])
    for (int i = 0; i < NX; ++i)
        for (int k = 0; k < NZ; ++k)
        {
            const int bc = cfg.bc_north;
            const std::size_t gj = g.at(i, NY - 1, k);
            if (bc == kPeriodic)
            {
                const std::size_t sidx = g.at(i, 1, k);
                v[gj] = v[sidx]; u[gj] = u[sidx]; w[gj] = w[sidx]; copy_rp(gj, sidx);
                continue;
            }
            const std::size_t a = g.at(i, NY - 2, k), am = g.at(i, NY - 3, k);
            if (bc == kRigid || bc == kRigidSlip) v[a] = S(0);
            if (bc == kRadiative && radiate)
            {
                float& vp = s.rad_n[std::size_t(i) * NZ + k];
                const float vn = kw78_face(vp, vp - 0.0f + cstar, vp - float(v[am]), dtf, inv_dy, false);
                v[a] = S(vn);
                vp = vn;
            }
            else if (bc == kZeroGradient) v[a] = v[am];
            v[gj] = (bc == kRadiative || bc == kZeroGradient) ? v[a] : S(0);
            if (bc == kRigid) { u[gj] = S(0); w[gj] = S(0); }
            else { u[gj] = u[a]; w[gj] = w[a]; }
            copy_rp(gj, a);
        }
    // vertical: bottom no-slip pins u,v; top rigid lid w = 0, u,v zero-gradient
    for (int i = 0; i < NX; ++i)
        for (int j = 0; j < NY; ++j)
        {
            if (cfg.bc_bottom == kBottomNoSlip) { u[g.at(i, j, 0)] = S(0); v[g.at(i, j, 0)] = S(0); }
            const std::size_t t = g.at(i, j, NZ - 1), tm = g.at(i, j, NZ - 2);
            w[t] = cfg.bc_top == kTopOpen ? w[tm] : S(0);
            u[t] = u[tm];
            v[t] = v[tm];
        }
}

// ----------------------------------------------------------------------
// Slow tendencies (cartesian_cgrid_3d.cpp:1038-1356). dp/dt is not formed:
// the first acoustic substep re-diagnoses p from the EOS in every cell
// before any operator reads it, so the slow pressure tendency never reaches
// an output of the large step.
// ----------------------------------------------------------------------
template <class S>
void slow_tendencies(const Grid& g, const Base& b, const Config& cfg, const State<S>& s,
                     std::vector<S>& du, std::vector<S>& dv, std::vector<S>& dw)
{
    const int NX = g.nx, NY = g.ny, NZ = g.nz;
    const double inv_dx = 1.0 / g.dx, inv_dy = 1.0 / g.dy, inv_dz = 1.0 / g.dz;
    const double f = cfg.coriolis_f;
    du.assign(g.cells(), S(0)); dv.assign(g.cells(), S(0)); dw.assign(g.cells(), S(0));
    const auto& u = s.u; const auto& v = s.v; const auto& w = s.w;
    auto F = [&](const std::vector<S>& fld, int i, int j, int k) { return double(fld[g.at(i, j, k)]); };
    auto finite0 = [](double x) { return std::isfinite(x) ? x : 0.0; };
    auto safe_rho = [](double r) { return std::isfinite(r) && r > 1.0e-6 ? r : 1.0; };
    auto stored = [](double x) { return double(float(x)); };
#pragma omp parallel for collapse(2) schedule(static)
    for (int i = 1; i < NX - 1; ++i)
        for (int j = 1; j < NY - 1; ++j)
            for (int k = 0; k < NZ - 1; ++k)
            {
                const std::size_t c = g.at(i, j, k);
                // ---- u at x-face ----
                {
                    const double ux = F(u, i, j, k);
                    const double vy = 0.25 * (F(v, i, j - 1, k) + F(v, i, j, k) + F(v, i + 1, j - 1, k) + F(v, i + 1, j, k));
                    const double wz = k == 0 ? 0.25 * (F(w, i, j, 0) + F(w, i + 1, j, 0))
                                             : 0.25 * (F(w, i, j, k - 1) + F(w, i, j, k) + F(w, i + 1, j, k - 1) + F(w, i + 1, j, k));
                    const double ddx = upwind_axis([&](int o) { return F(u, i + o, j, k); }, i, NX, ux, inv_dx);
                    const double ddy = upwind_axis([&](int o) { return F(u, i, j + o, k); }, j, NY, vy, inv_dy);
                    const double ddz = k == 0 ? (wz >= 0.0 ? 0.0 : (F(u, i, j, 1) - F(u, i, j, 0)) * inv_dz)
                                              : upwind_axis([&](int o) { return F(u, i, j, k + o); }, k, NZ, wz, inv_dz);
                    const double v0k = (int(b.v0.size()) > k) ? b.v0[k] : 0.0;
                    const double cor = f * (vy - v0k);
                    du[c] = S(finite0(-ux * ddx - vy * ddy - wz * ddz + cor));
                }
                // ---- v at y-face ----
                {
                    const double vy = F(v, i, j, k);
                    const double ux = 0.25 * (F(u, i - 1, j, k) + F(u, i, j, k) + F(u, i - 1, j + 1, k) + F(u, i, j + 1, k));
                    const double wz = k == 0 ? 0.25 * (F(w, i, j, 0) + F(w, i, j + 1, 0))
                                             : 0.25 * (F(w, i, j, k - 1) + F(w, i, j, k) + F(w, i, j + 1, k - 1) + F(w, i, j + 1, k));
                    const double ddx = upwind_axis([&](int o) { return F(v, i + o, j, k); }, i, NX, ux, inv_dx);
                    const double ddy = upwind_axis([&](int o) { return F(v, i, j + o, k); }, j, NY, vy, inv_dy);
                    const double ddz = k == 0 ? (wz >= 0.0 ? 0.0 : (F(v, i, j, 1) - F(v, i, j, 0)) * inv_dz)
                                              : upwind_axis([&](int o) { return F(v, i, j, k + o); }, k, NZ, wz, inv_dz);
                    const double u0k = (int(b.u0.size()) > k) ? b.u0[k] : 0.0;
                    const double cor = f * (ux - u0k);
                    dv[c] = S(finite0(-ux * ddx - vy * ddy - wz * ddz - cor));
                }
                // ---- w at z-face ----
                {
                    const double wz = F(w, i, j, k);
                    const double ux = 0.25 * (F(u, i - 1, j, k) + F(u, i, j, k) + F(u, i - 1, j, k + 1) + F(u, i, j, k + 1));
                    const double vy = 0.25 * (F(v, i, j - 1, k) + F(v, i, j, k) + F(v, i, j - 1, k + 1) + F(v, i, j, k + 1));
                    const double ddx = upwind_axis([&](int o) { return F(w, i + o, j, k); }, i, NX, ux, inv_dx);
                    const double ddy = upwind_axis([&](int o) { return F(w, i, j + o, k); }, j, NY, vy, inv_dy);
                    const double ddz = k == 0 ? (wz >= 0.0 ? F(w, i, j, 0) * inv_dz : (F(w, i, j, 1) - F(w, i, j, 0)) * inv_dz)
                                              : upwind_axis([&](int o) { return F(w, i, j, k + o); }, k, NZ, wz, inv_dz);
                    const double rf = 0.5 * (F(s.rho, i, j, k) + F(s.rho, i, j, k + 1));
                    const double r0f = 0.5 * (b.rho0[k] + b.rho0[k + 1]);
                    const double buoy = -kG * (rf - r0f) / safe_rho(rf);
                    const double qvf = 0.5 * (F(s.qv, i, j, k) + F(s.qv, i, j, k + 1));
                    const double qv0f = 0.5 * (stored(b.qv0[k]) + stored(b.qv0[k + 1]));
                    const double mbuoy = kG * 0.608 * (qvf - qv0f);
                    const double ql_lo = F(s.qc, i, j, k) + F(s.qr, i, j, k);
                    const double ql_hi = F(s.qc, i, j, k + 1) + F(s.qr, i, j, k + 1);
                    const double load = -kG * (0.5 * (ql_lo + ql_hi));
                    dw[c] = S(finite0(-ux * ddx - vy * ddy - wz * ddz + buoy + mbuoy + load));
                }
            }
}

// Slow apply u += dt*du (dynamics.cpp:2055-2095); non-finite tendencies count as 0.
template <class S>
void slow_apply(State<S>& s, const std::vector<S>& du, const std::vector<S>& dv, const std::vector<S>& dw, double dt)
{
    const std::size_t n = s.u.size();
#pragma omp parallel for schedule(static)
    for (std::size_t c = 0; c < n; ++c)
    {
        double a = double(du[c]), b = double(dv[c]), d = double(dw[c]);
        if (!std::isfinite(a)) a = 0.0;
        if (!std::isfinite(b)) b = 0.0;
        if (!std::isfinite(d)) d = 0.0;
        s.u[c] = S(double(s.u[c]) + a * dt);
        s.v[c] = S(double(s.v[c]) + b * dt);
        s.w[c] = S(double(s.w[c]) + d * dt);
    }
}

// Heun slow RK2 (dynamics.cpp:5123-5292): k1 at u_n, predictor u_n + dt*k1,
// k2 at the predictor, state restored, tendency = 0.5*(k1 + k2) in storage
// precision, then the slow apply with dt.
template <class S>
void slow_rk2_and_apply(const Grid& g, const Base& b, const Config& cfg, State<S>& s)
{
    std::vector<S> k1u, k1v, k1w, k2u, k2v, k2w;
    slow_tendencies(g, b, cfg, s, k1u, k1v, k1w);
    const std::vector<S> us = s.u, vs = s.v, ws = s.w;
    slow_apply(s, k1u, k1v, k1w, cfg.dt);
    slow_tendencies(g, b, cfg, s, k2u, k2v, k2w);
    s.u = us; s.v = vs; s.w = ws;
    const std::size_t n = s.u.size();
    for (std::size_t c = 0; c < n; ++c)
    {
        k2u[c] = S(S(0.5) * (k2u[c] + k1u[c]));
        k2v[c] = S(S(0.5) * (k2v[c] + k1v[c]));
        k2w[c] = S(S(0.5) * (k2w[c] + k1w[c]));
    }
    slow_apply(s, k2u, k2v, k2w, cfg.dt);
}

// Time-averaged face mass flux, sampled from the pre-continuity state of each
// substep (advection_cartesian.cpp:610-830, Host precision: FP64 add, storage round).
template <class S>
void accumulate_mass_flux(const Grid& g, const Config& cfg, const State<S>& s, double weight,
                          std::vector<S>& mfx, std::vector<S>& mfy, std::vector<S>& mfz)
{
    const int NX = g.nx, NY = g.ny, NZ = g.nz;
#pragma omp parallel for collapse(2) schedule(static)
    for (int i = 0; i < NX; ++i)
        for (int j = 0; j < NY; ++j)
            for (int k = 0; k < NZ; ++k)
            {
                const std::size_t c = g.at(i, j, k);
                const double rc = double(s.rho[c]);
                if (i < NX - 1)
                {
                    const double fl = 0.5 * (rc + double(s.rho[g.at(i + 1, j, k)])) * double(s.u[c]);
                    mfx[c] = S(double(mfx[c]) + weight * fl);
                }
                if (j < NY - 1)
                {
                    const double fl = 0.5 * (rc + double(s.rho[g.at(i, j + 1, k)])) * double(s.v[c]);
                    mfy[c] = S(double(mfy[c]) + weight * fl);
                }
                double fl = 0.0;
                if (k < NZ - 1) fl = 0.5 * (rc + double(s.rho[g.at(i, j, k + 1)])) * double(s.w[c]);
                else fl = cfg.bc_top == kTopOpen ? rc * double(s.w[c]) : 0.0;
                mfz[c] = S(double(mfz[c]) + weight * fl);
            }
}


// The reference clamps in FP32 on the stored value (field_bounds.hpp:41-53);
// the FP64 oracle applies the same bounds to the FP64 value.
template <class S>
inline double clamp_store(double x, double fallback, float lo, float hi)
{
    if constexpr (sizeof(S) == sizeof(float))
    {
        float f = float(x);
        if (!std::isfinite(f)) f = float(fallback);
        return double(std::clamp(f, lo, hi));
    }
    else
    {
        if (!std::isfinite(x)) x = fallback;
        return std::clamp(x, double(lo), double(hi));
    }
}

inline double finite_or_zero(double x) { return std::isfinite(x) ? x : 0.0; }
inline double safe_rho(double r) { return std::isfinite(r) && r > 1.0e-6 ? r : 1.0; }

// Backward momentum update of the acoustic substep with the new pressure
// (acoustic_cgrid_fast.cpp:159-215). w is skipped when `skip_w` (HE-VI solves it).
template <class S>
void momentum_backward(const Grid& g, const Base& b, State<S>& s, double dts, bool skip_w)
{
    const int NX = g.nx, NY = g.ny, NZ = g.nz;
    const double inv_dx = 1.0 / g.dx, inv_dy = 1.0 / g.dy, inv_dz = 1.0 / g.dz;
#pragma omp parallel for collapse(2) schedule(static)
    for (int i = 0; i < NX; ++i)
        for (int j = 0; j < NY; ++j)
            for (int k = 0; k < NZ; ++k)
            {
                const std::size_t c = g.at(i, j, k);
                double du = 0.0, dv = 0.0, dw = 0.0;
                if (i >= 1 && i < NX - 1 && j >= 1 && j < NY - 1 && k < NZ - 1)
                {
                    const std::size_t ip = g.at(i + 1, j, k), jp = g.at(i, j + 1, k), kp = g.at(i, j, k + 1);
                    const double rc = double(s.rho[c]);
                    const double rx = 0.5 * (rc + double(s.rho[ip]));
                    const double ry = 0.5 * (rc + double(s.rho[jp]));
                    const double rz = 0.5 * (rc + double(s.rho[kp]));
                    const double dp0 = (b.p0[k + 1] - b.p0[k]) * inv_dz;
                    // tendencies are rounded to FP32 before the update in the reference
                    du = double(S(finite_or_zero(-(double(s.p[ip]) - double(s.p[c])) * inv_dx / safe_rho(rx))));
                    dv = double(S(finite_or_zero(-(double(s.p[jp]) - double(s.p[c])) * inv_dy / safe_rho(ry))));
                    dw = double(S(finite_or_zero(-(((double(s.p[kp]) - double(s.p[c])) * inv_dz) - dp0) / safe_rho(rz))));
                }
                s.u[c] = S(clamp_store<S>(double(s.u[c]) + du * dts, 0.0, -kWindH, kWindH));
                s.v[c] = S(clamp_store<S>(double(s.v[c]) + dv * dts, 0.0, -kWindH, kWindH));
                if (!skip_w) s.w[c] = S(clamp_store<S>(double(s.w[c]) + dw * dts, 0.0, -kWindV, kWindV));
            }
}

// Continuity flux divergence of the explicit substep (acoustic_cgrid_fast.cpp:49-116);
// `horizontal_only` drops the vertical terms (HE-VI treats them implicitly).
template <class S>
double continuity_divergence(const Grid& g, const State<S>& s, int i, int j, int k, bool horizontal_only)
{
    const int NZ = g.nz;
    const double inv_dx = 1.0 / g.dx, inv_dy = 1.0 / g.dy, inv_dz = 1.0 / g.dz;
    auto R = [&](std::size_t c) { return double(s.rho[c]); };
    const std::size_t c = g.at(i, j, k);
    const std::size_t im = g.at(i - 1, j, k), jm = g.at(i, j - 1, k);
    const double fxp = 0.5 * (R(c) + R(g.at(i + 1, j, k))) * double(s.u[c]);
    const double fxm = 0.5 * (R(im) + R(c)) * double(s.u[im]);
    const double fyp = 0.5 * (R(c) + R(g.at(i, j + 1, k))) * double(s.v[c]);
    const double fym = 0.5 * (R(jm) + R(c)) * double(s.v[jm]);
    double d = (fxp - fxm) * inv_dx + (fyp - fym) * inv_dy;
    if (!horizontal_only)
    {
        const double fzp = k + 1 < NZ ? 0.5 * (R(c) + R(g.at(i, j, k + 1))) * double(s.w[c]) : R(c) * double(s.w[c]);
        const double fzm = k > 0 ? 0.5 * (R(g.at(i, j, k - 1)) + R(c)) * double(s.w[g.at(i, j, k - 1)]) : 0.0;
        d += (fzp - fzm) * inv_dz;
    }
    return -d;
}

// Density update + EOS pressure of pass 2 (acoustic_cgrid_fast.cpp:118-157).
template <class S>
void density_eos_update(const Base& b, State<S>& s, std::size_t c, int k, double rho_new)
{
    double rn = rho_new;
    if (!std::isfinite(rn) || rn <= 0.0) rn = std::max(0.1, b.rho0[k]);
    const double rs = clamp_store<S>(rn, 1.0, kRhoMin, kRhoMax);
    s.rho[c] = S(rs);
    double pn = kP00;
    const double th = double(s.theta[c]);
    if (th > 0.0) pn = eos_pressure(double(s.rho[c]), th);
    if (!std::isfinite(pn) || pn <= 0.0) pn = kP00;
    s.p[c] = S(clamp_store<S>(pn, 101325.0, kPMin, kPMax));
}

// Explicit forward-backward acoustic substep (src/compute/cpu/acoustic_cgrid_fast.cpp:12-219).
template <class S>
void acoustic_substep_explicit(const Grid& g, const Base& b, State<S>& s, double dts)
{
    const int NX = g.nx, NY = g.ny, NZ = g.nz;
    std::vector<S> drho(g.cells(), S(0));
#pragma omp parallel for collapse(2) schedule(static)
    for (int i = 1; i < NX - 1; ++i)
        for (int j = 1; j < NY - 1; ++j)
            for (int k = 0; k < NZ; ++k)
                drho[g.at(i, j, k)] = S(finite_or_zero(continuity_divergence(g, s, i, j, k, false)));
#pragma omp parallel for collapse(2) schedule(static)
    for (int i = 0; i < NX; ++i)
        for (int j = 0; j < NY; ++j)
            for (int k = 0; k < NZ; ++k)
            {
                const std::size_t c = g.at(i, j, k);
                density_eos_update(b, s, c, k, double(s.rho[c]) + double(drho[c]) * dts);
            }
    momentum_backward(g, b, s, dts, false);
}

// Thomas algorithm for a tridiagonal system a_k x_{k-1} + d_k x_k + c_k x_{k+1} = r_k
// in FP64 with a positive-pivot guard (analog of implicit.cpp:465-500).
inline bool thomas_solve(int n, const double* a, const double* d, const double* c, const double* r, double* x)
{
    std::vector<double> cp(n), rp(n);
    double piv = d[0];
    if (!(piv > 0.0) || !std::isfinite(piv)) return false;
    cp[0] = c[0] / piv;
    rp[0] = r[0] / piv;
    for (int k = 1; k < n; ++k)
    {
        piv = d[k] - a[k] * cp[k - 1];
        if (!(piv > 0.0) || !std::isfinite(piv)) return false;
        cp[k] = c[k] / piv;
        rp[k] = (r[k] - a[k] * rp[k - 1]) / piv;
    }
    x[n - 1] = rp[n - 1];
    for (int k = n - 2; k >= 0; --k) x[k] = rp[k] - cp[k] * x[k + 1];
    return true;
}

// Coefficients of the HE-VI w system for one column. Inputs per level k:
// rho_bar[k] (face density k+1/2, k = 0..nz-2), C[k] = dp/drho at fixed theta
// (cell k), G[k] explicit vertical PGF at face k+1/2 from alpha p* + (1 - alpha) p^n minus
// the base-state gradient, e[k] explicit
// part of the vertical density change of cell k, w_old[k]. Unknowns are the
// interior faces k = 0..nz-2 (ground and lid fluxes are zero).
inline void hevi_column_system(int nz, double dt, double dz, double alpha, const double* rho_bar,
                               const double* C, const double* G, const double* e, const double* w_old,
                               double* A, double* D, double* U, double* Rhs)
{
    const int n = nz - 1;
    const double Bc = alpha * dt / dz;
    for (int k = 0; k < n; ++k)
    {
        const double Ak = dt * alpha / (rho_bar[k] * dz);
        const double rk = rho_bar[k];
        A[k] = k > 0 ? -Ak * Bc * C[k] * rho_bar[k - 1] : 0.0;
        U[k] = k < n - 1 ? -Ak * Bc * C[k + 1] * rho_bar[k + 1] : 0.0;
        D[k] = 1.0 + Ak * Bc * (C[k + 1] + C[k]) * rk;
        Rhs[k] = w_old[k] - dt / rk * G[k] - Ak * (C[k + 1] * e[k + 1] - C[k] * e[k]);
    }
}

// Vertically implicit (HE-VI) acoustic substep. Horizontal continuity and
// momentum are explicit forward-backward as in the reference substep; the
// vertical w-rho-p coupling is off-centred implicit (alpha = (1+beta)/2):
// the vertical pressure gradient acts on alpha p^{n+1} + (1 - alpha) p^n,
// where p^n is the pressure at the start of the substep and p^{n+1} is
// linearised about p* = EOS(rho*), rho* = rho^n plus the explicit horizontal
// divergence. (Off-centring against p* instead of p^n makes the horizontal
// divergence enter the old-level part and is unstable: von Neumann
// amplification up to 1.24 at horizontal Courant 0.7, alpha = 0.55.) The w
// system is solved per column by an FP64 Thomas sweep, after which rho is
// updated in flux form and p re-diagnosed from the EOS. The vertical mass
// flux alpha*w_new + (1-alpha)*w_old is returned in `wz_eff` for the
// transport mass-flux accumulation.
template <class S>
void acoustic_substep_hevi(const Grid& g, const Base& b, State<S>& s, double dts, double beta,
                           std::vector<double>& wz_eff)
{
    const int NX = g.nx, NY = g.ny, NZ = g.nz;
    const double alpha = 0.5 * (1.0 + beta);
    const double inv_dz = 1.0 / g.dz;
    wz_eff.assign(g.cells(), 0.0);
    // columns are independent: horizontal divergence uses old u, v and old rho
    std::vector<S> rho_old = s.rho;
    State<S> snap;  // read-only view of the substep-start state for neighbours
    snap.u = s.u; snap.v = s.v; snap.w = s.w; snap.rho = rho_old;
#pragma omp parallel for collapse(2) schedule(static)
    for (int i = 1; i < NX - 1; ++i)
        for (int j = 1; j < NY - 1; ++j)
        {
            std::vector<double> rstar(NZ), pstar(NZ), C(NZ + 1, 0.0), G(NZ), e(NZ + 1, 0.0), rb(NZ),
                wo(NZ), A(NZ), D(NZ), U(NZ), Rh(NZ), x(NZ, 0.0);
            for (int k = 0; k < NZ; ++k)
            {
                const std::size_t c = g.at(i, j, k);
                rstar[k] = double(rho_old[c]) + dts * finite_or_zero(continuity_divergence(g, snap, i, j, k, true));
                const double th = double(s.theta[c]);
                pstar[k] = eos_pressure(rstar[k], th);
                C[k] = kGamma * pstar[k] / rstar[k];
                wo[k] = double(s.w[c]);
            }
            for (int k = 0; k < NZ - 1; ++k)
            {
                rb[k] = 0.5 * (double(rho_old[g.at(i, j, k)]) + double(rho_old[g.at(i, j, k + 1)]));
                const double pn0 = double(s.p[g.at(i, j, k)]), pn1 = double(s.p[g.at(i, j, k + 1)]);
                G[k] = (alpha * (pstar[k + 1] - pstar[k]) + (1.0 - alpha) * (pn1 - pn0)) * inv_dz -
                       (b.p0[k + 1] - b.p0[k]) * inv_dz;
            }
            rb[NZ - 1] = 0.0;
            for (int k = 0; k < NZ; ++k)
            {
                const double ftop = k < NZ - 1 ? rb[k] * wo[k] : 0.0;
                const double fbot = k > 0 ? rb[k - 1] * wo[k - 1] : 0.0;
                e[k] = -(1.0 - alpha) * dts * inv_dz * (ftop - fbot);
            }
            hevi_column_system(NZ, dts, g.dz, alpha, rb.data(), C.data(), G.data(), e.data(), wo.data(),
                               A.data(), D.data(), U.data(), Rh.data());
            const bool ok = thomas_solve(NZ - 1, A.data(), D.data(), U.data(), Rh.data(), x.data());
            if (!ok)
                for (int k = 0; k < NZ - 1; ++k) x[k] = wo[k];
            // clamp and commit w, then flux-form density update with the weighted flux
            for (int k = 0; k < NZ - 1; ++k)
            {
                const std::size_t c = g.at(i, j, k);
                s.w[c] = S(clamp_store<S>(x[k], 0.0, -kWindV, kWindV));
                x[k] = double(s.w[c]);
            }
            for (int k = 0; k < NZ; ++k)
            {
                const double wt = k < NZ - 1 ? alpha * x[k] + (1.0 - alpha) * wo[k] : 0.0;
                const double wb = k > 0 ? alpha * x[k - 1] + (1.0 - alpha) * wo[k - 1] : 0.0;
                const double ft = k < NZ - 1 ? rb[k] * wt : 0.0;
                const double fb = k > 0 ? rb[k - 1] * wb : 0.0;
                wz_eff[g.at(i, j, k)] = ft;
                const double rn = rstar[k] - dts * inv_dz * (ft - fb);
                density_eos_update(b, s, g.at(i, j, k), k, rn);
            }
        }
    // ghost columns: density unchanged, clamped, pressure re-diagnosed (pass 2 of the reference)
    for (int i = 0; i < NX; ++i)
        for (int j = 0; j < NY; ++j)
        {
            if (i >= 1 && i < NX - 1 && j >= 1 && j < NY - 1) continue;
            for (int k = 0; k < NZ; ++k) density_eos_update(b, s, g.at(i, j, k), k, double(s.rho[g.at(i, j, k)]));
        }
    momentum_backward(g, b, s, dts, true);
    // ghost-column and lid w follow the reference clamp-only rule
    for (int i = 0; i < NX; ++i)
        for (int j = 0; j < NY; ++j)
        {
            const bool interior = i >= 1 && i < NX - 1 && j >= 1 && j < NY - 1;
            for (int k = 0; k < NZ; ++k)
                if (!interior || k == NZ - 1)
                {
                    const std::size_t c = g.at(i, j, k);
                    s.w[c] = S(clamp_store<S>(double(s.w[c]), 0.0, -kWindV, kWindV));
                }
        }
}

// ----------------------------------------------------------------------
// WENO5 conservative scalar transport with retained mass fluxes and the
// positive-definite donor limiter (advection_cartesian.cpp:1941-2650).
// ----------------------------------------------------------------------
template <class S>
void transport_scalar(const Grid& g, const Config& cfg, const std::vector<S>& q, std::vector<S>& out,
                      const std::vector<S>& mfx, const std::vector<S>& mfy, const std::vector<S>& mfz,
                      const std::vector<S>& rho_begin, const std::vector<S>& rho_end, double dt,
                      bool positive_definite)
{
    const int NX = g.nx, NY = g.ny, NZ = g.nz;
    const double inv_dx = 1.0 / g.dx, inv_dy = 1.0 / g.dy, inv_dz = 1.0 / g.dz;
    const bool open_top = cfg.bc_top == kTopOpen;
    std::vector<S> sfx(g.cells(), S(0)), sfy(g.cells(), S(0)), sfz(g.cells(), S(0)), rl(g.cells(), S(1));
    auto Q = [&](int i, int j, int k) { return double(q[g.at(i, j, k)]); };
    // face value along one axis: `at(o)` reads q at offset o from the lower cell
    auto face = [&](auto at, int pos, int n, double mf) {
        if (mf >= 0.0)
        {
            const double q0 = at(0);
            if (pos >= 2 && pos + 2 <= n - 1) return weno5_face(at(-2), at(-1), q0, at(1), at(2));
            const double sl = pos >= 1 ? mc_slope(at(-1), q0, at(1)) : 0.0;
            return q0 + 0.5 * sl;
        }
        const double q0 = at(1);
        if (pos - 1 >= 0 && pos + 3 <= n - 1) return weno5_face(at(3), at(2), q0, at(0), at(-1));
        const double sl = pos + 1 <= n - 2 ? mc_slope(at(0), q0, at(2)) : 0.0;
        return q0 - 0.5 * sl;
    };
#pragma omp parallel for collapse(2) schedule(static)
    for (int i = 0; i < NX; ++i)
        for (int j = 0; j < NY; ++j)
            for (int k = 0; k < NZ; ++k)
            {
                const std::size_t f = g.at(i, j, k);
                if (i < NX - 1)
                {
                    const double m = double(mfx[f]);
                    sfx[f] = S(m * face([&](int o) { return Q(i + o, j, k); }, i, NX, m));
                }
                if (j < NY - 1)
                {
                    const double m = double(mfy[f]);
                    sfy[f] = S(m * face([&](int o) { return Q(i, j + o, k); }, j, NY, m));
                }
                if (k < NZ - 1)
                {
                    const double m = double(mfz[f]);
                    sfz[f] = S(m * face([&](int o) { return Q(i, j, k + o); }, k, NZ, m));
                }
                else if (open_top)
                {
                    sfz[f] = S(double(mfz[f]) * Q(i, j, k));
                }
            }
    if (positive_definite)
    {
#pragma omp parallel for collapse(2) schedule(static)
        for (int i = 0; i < NX; ++i)
            for (int j = 0; j < NY; ++j)
                for (int k = 0; k < NZ; ++k)
                {
                    const std::size_t c = g.at(i, j, k);
                    double out_rate = 0.0;
                    if (i <= NX - 2 && sfx[c] > S(0)) out_rate += double(sfx[c]) * inv_dx;
                    if (i >= 1 && sfx[g.at(i - 1, j, k)] < S(0)) out_rate -= double(sfx[g.at(i - 1, j, k)]) * inv_dx;
                    if (j <= NY - 2 && sfy[c] > S(0)) out_rate += double(sfy[c]) * inv_dy;
                    if (j >= 1 && sfy[g.at(i, j - 1, k)] < S(0)) out_rate -= double(sfy[g.at(i, j - 1, k)]) * inv_dy;
                    if ((k <= NZ - 2 || (open_top && k == NZ - 1)) && sfz[c] > S(0)) out_rate += double(sfz[c]) * inv_dz;
                    if (k >= 1 && sfz[g.at(i, j, k - 1)] < S(0)) out_rate -= double(sfz[g.at(i, j, k - 1)]) * inv_dz;
                    double r = 1.0;
                    if (out_rate > 0.0)
                    {
                        const double avail = double(rho_begin[c]) * std::max(double(q[c]), 0.0) / dt;
                        r = std::min(1.0, avail / out_rate);
                    }
                    float rs = float(r);
                    if (r > 0.0 && r < 1.0) rs = pd_limiter_store(r, false);
                    rl[c] = S(rs);
                }
    }
    auto limited = [&](S raw, std::size_t lo, std::size_t hi) -> double {
        if (!positive_definite || raw == S(0)) return double(raw);
        const S r = raw > S(0) ? rl[lo] : rl[hi];
        return double(raw) * double(r);
    };
    out = q;
    const double rho_floor = double(kRhoMin);
#pragma omp parallel for collapse(2) schedule(static)
    for (int i = 1; i < NX - 1; ++i)
        for (int j = 1; j < NY - 1; ++j)
            for (int k = 0; k < NZ; ++k)
            {
                const std::size_t c = g.at(i, j, k);
                const std::size_t xm = g.at(i - 1, j, k), xp = g.at(i + 1, j, k);
                const std::size_t ym = g.at(i, j - 1, k), yp = g.at(i, j + 1, k);
                const bool has_top = k <= NZ - 2 || (open_top && k == NZ - 1);
                const double sxp = limited(sfx[c], c, xp);
                const double sxm = limited(sfx[xm], xm, c);
                const double syp = limited(sfy[c], c, yp);
                const double sym = limited(sfy[ym], ym, c);
                const double szp = has_top ? limited(sfz[c], c, k < NZ - 1 ? g.at(i, j, k + 1) : c) : 0.0;
                const double szm = k >= 1 ? limited(sfz[g.at(i, j, k - 1)], g.at(i, j, k - 1), c) : 0.0;
                const double rb = double(rho_begin[c]);
                double rho_phi;
                if (positive_definite)
                {
                    double outgoing = 0.0, incoming = 0.0;
                    auto plus = [&](double fl, double inv) { if (fl >= 0.0) outgoing += fl * inv; else incoming -= fl * inv; };
                    auto minus = [&](double fl, double inv) { if (fl <= 0.0) outgoing -= fl * inv; else incoming += fl * inv; };
                    plus(sxp, inv_dx); minus(sxm, inv_dx);
                    plus(syp, inv_dy); minus(sym, inv_dy);
                    if (has_top) plus(szp, inv_dz);
                    if (k >= 1) minus(szm, inv_dz);
                    rho_phi = (rb * double(q[c]) - dt * outgoing) + dt * incoming;
                }
                else
                {
                    const double div = (sxp - sxm) * inv_dx + (syp - sym) * inv_dy + (szp - szm) * inv_dz;
                    rho_phi = rb * double(q[c]) - dt * div;
                }
                const double re = double(rho_end[c]);
                const double den = re > rho_floor ? re : rho_floor;
                out[c] = S(rho_phi / den);
            }
}

// ----------------------------------------------------------------------
// Kessler warm rain (src/microphysics/schemes/kessler/kessler.cpp:103-531).
// Returns per-cell tendencies for theta, qv, qc, qr on the physical domain
// i in [1, nx-1), j in [1, ny-1), all k; and the bottom sink rate per column
// (kessler.cpp:533-561).
// ----------------------------------------------------------------------
struct KesslerParams
{
    double qc0 = 1.0e-3, c_auto = 1.0e-3, c_accr = 2.2, evap_eff = 1.0, vt_max = 20.0;
};

template <class S>
void kessler_tendencies(const Grid& g, const KesslerParams& kp, const std::vector<S>& p, const std::vector<S>& theta,
                        const std::vector<S>& rho, const std::vector<S>& qv, const std::vector<S>& qc,
                        const std::vector<S>& qr, double dt, std::vector<S>& dth, std::vector<S>& dqv,
                        std::vector<S>& dqc, std::vector<S>& dqr)
{
    const int NX = g.nx, NY = g.ny, NZ = g.nz;
    for (auto* f : {&dth, &dqv, &dqc, &dqr}) f->assign(g.cells(), S(0));
    if (!(dt > 0.0) || !std::isfinite(dt)) return;
    const double inv_dt = 1.0 / dt;
    const double L = kLv / kCp;
#pragma omp parallel for collapse(2) schedule(static)
    for (int i = 1; i < NX - 1; ++i)
        for (int j = 1; j < NY - 1; ++j)
            for (int k = 0; k < NZ; ++k)
            {
                const std::size_t c = g.at(i, j, k);
                // saturation adjustment on copies (kessler.cpp:169-208)
                const S T0s = S(theta_to_temperature(double(theta[c]), double(p[c])));
                double T = double(T0s), qva = double(qv[c]), qca = double(qc[c]);
                T = saturation_adjust(T, double(p[c]), qva, qca);
                const S Ts = S(T), qvs_ = S(qva), qcs_ = S(qca);
                const double dqc_adj = double(qcs_) - double(qc[c]);
                if (dqc_adj != 0.0)
                {
                    const double rate = dqc_adj * inv_dt;
                    dqc[c] = S(double(dqc[c]) + double(S(rate)));
                    dqv[c] = S(double(dqv[c]) - double(S(rate)));
                    dth[c] = S(double(dth[c]) + double(S(L * rate)));
                }
                // warm rain on the adjusted state (kessler.cpp:346-460)
                const S qcv = qcs_, qrv = qr[c], qvv = qvs_;
                S prod = S(0);
                if (qcv > S(0))
                {
                    const double factor = 1.0 / (1.0 + kp.c_accr * dt * std::pow(double(std::max(S(0), qrv)), 0.875));
                    const double amount = double(qcv) * (1.0 - factor)
                                        + factor * kp.c_auto * dt * std::max(double(qcv) - kp.qc0, 0.0);
                    prod = S(std::clamp(amount, 0.0, double(qcv)) / dt);
                    dqc[c] = dqc[c] - prod;
                    dqr[c] = dqr[c] + prod;
                }
                if (qrv > S(0) || prod > S(0))
                {
                    const double pv = double(p[c]), rv = double(rho[c]);
                    const double qsat = qvs_water(double(Ts), pv);
                    if (qsat > 0.0 && double(qvv) < qsat && rv > 0.0 && pv > 0.0)
                    {
                        const double qr_av = double(qrv) + double(prod) * dt;
                        const double rcgs = 1.0e-3 * rv;
                        const double rc = rcgs * std::max(qr_av, 0.0);
                        const double vent = (1.6 + 124.9 * std::pow(rc, 0.2046)) * std::pow(rc, 0.525);
                        const double resist = 2.55e8 / (pv * qsat) + 5.4e5;
                        const double raw = kp.evap_eff * (vent / resist) * ((qsat - double(qvv)) / (rcgs * qsat));
                        const double gsat = (kLv * kLv * qsat) / (kCp * kRv * std::max(double(Ts) * double(Ts), 1.0));
                        const double bound = (qsat - double(qvv)) / (1.0 + gsat);
                        const double ev = std::clamp(raw, 0.0, std::min(bound / dt, qr_av / dt));
                        dqv[c] = S(double(dqv[c]) + ev);
                        dqr[c] = S(double(dqr[c]) - ev);
                        dth[c] = S(double(dth[c]) - L * ev);
                    }
                }
            }
    // sedimentation of the pre-process rain, donor then receiver (kessler.cpp:465-531)
#pragma omp parallel for collapse(2) schedule(static)
    for (int i = 1; i < NX - 1; ++i)
        for (int j = 1; j < NY - 1; ++j)
            for (int k = 0; k < NZ; ++k)
            {
                const std::size_t c = g.at(i, j, k);
                const double rk = double(rho[c]);
                const double rkm = k > 0 ? double(rho[g.at(i, j, k - 1)]) : rk;
                const double qv_ = std::max(0.0, double(qr[c]));
                if (qv_ > 0.0)
                {
                    const double Vt = std::min(kessler_rain_speed(qv_, rk) * std::sqrt(1.2 / rk), kp.vt_max);
                    const S flux = S(Vt * qv_);
                    dqr[c] = dqr[c] - S(double(flux) / g.dz);
                    if (k > 0)
                    {
                        const std::size_t cm = g.at(i, j, k - 1);
                        dqr[cm] = dqr[cm] + S(rk / rkm * double(flux) / g.dz);
                    }
                }
            }
    // temperature tendency to theta tendency (kessler.cpp:326-338)
#pragma omp parallel for collapse(2) schedule(static)
    for (int i = 1; i < NX - 1; ++i)
        for (int j = 1; j < NY - 1; ++j)
            for (int k = 0; k < NZ; ++k)
            {
                const std::size_t c = g.at(i, j, k);
                dth[c] = S(theta_tendency_from_temperature(double(dth[c]), double(p[c])));
            }
}

template <class S>
double kessler_bottom_sink(const Grid& g, const KesslerParams& kp, const std::vector<S>& rho,
                           const std::vector<S>& qr, int i, int j)
{
    // kessler.cpp:533-561
    const std::size_t c = g.at(i, j, 0);
    const double r0 = double(rho[c]);
    const S q = qr[c];
    if (!(q > S(0))) return 0.0;
    const double Vt = std::min(kessler_rain_speed(double(q), r0) * std::sqrt(1.2 / r0), kp.vt_max);
    return double(S(S(Vt) * q)) / g.dz;
}

// Longest stable microphysics step of the state (sedimentation_step_limit,
// microphysics_step.cpp:3501-3572): the minimum over the physical cells of
// 0.8 * layer thickness / speed, where speed is the scheme's largest explicit
// terminal velocity for the cell's total water qv + qc + qr at its density
// (KesslerScheme::explicit_terminal_velocity_mps, kessler.cpp:39-44:
// min(Vt_max, kessler_rain(q, rho) sqrt(1.2 / rho))); infinity when no cell
// has a positive speed.
template <class S>
double sedimentation_limit(const Grid& g, const Config& cfg, const State<S>& s)
{
    double limit = std::numeric_limits<double>::infinity();
    for (int i = 1; i < g.nx - 1; ++i)
        for (int j = 1; j < g.ny - 1; ++j)
            for (int k = 0; k < g.nz; ++k)
            {
                const std::size_t c = g.at(i, j, k);
                const double r = double(s.rho[c]);
                const double qt = double(s.qv[c]) + double(s.qc[c]) + double(s.qr[c]);
                const double speed = std::min(cfg.vt_max, kessler_rain_speed(qt, r) * std::sqrt(1.2 / r));
                if (speed > 0.0) limit = std::min(limit, 0.8 * g.dz / speed);
            }
    return limit;
}

// One microphysics step of length h: Kessler tendencies (saturation
// adjustment, warm rain, sedimentation), the bounded commit
// (microphysics_step.cpp:1090-1140: theta step limited to +-50 K and clamped
// to [200, 700] K, water species clamped to [0, 0.05]) and the surface rain
// ledger (microphysics_step.cpp:1917-2024: precipitation flux = sink*rho*dz,
// 1 kg m^-2 = 1 mm).
template <class S>
void microphysics_once(const Grid& g, const Config& cfg, State<S>& s, double h)
{
    const KesslerParams kp{cfg.qc0, cfg.c_auto, cfg.c_accr, cfg.evap_eff, cfg.vt_max};
    std::vector<S> dth, dqv, dqc, dqr;
    kessler_tendencies(g, kp, s.p, s.theta, s.rho, s.qv, s.qc, s.qr, h, dth, dqv, dqc, dqr);
    for (int i = 1; i < g.nx - 1; ++i)
        for (int j = 1; j < g.ny - 1; ++j)
        {
            const double sink = kessler_bottom_sink(g, kp, s.rho, s.qr, i, j);
            s.rain_mm[std::size_t(i) * g.ny + j] += sink * double(s.rho[g.at(i, j, 0)]) * g.dz * h;
        }
#pragma omp parallel for collapse(2) schedule(static)
    for (int i = 1; i < g.nx - 1; ++i)
        for (int j = 1; j < g.ny - 1; ++j)
            for (int k = 0; k < g.nz; ++k)
            {
                const std::size_t c = g.at(i, j, k);
                const double step = std::clamp(double(dth[c]) * h, -50.0, 50.0);
                s.theta[c] = S(clamp_store<S>(double(s.theta[c]) + step, 300.0, kThetaMin, kThetaMax));
                s.qv[c] = S(clamp_store<S>(double(s.qv[c]) + double(dqv[c]) * h, 0.0, 0.0f, kQvMax));
                s.qc[c] = S(clamp_store<S>(double(s.qc[c]) + double(dqc[c]) * h, 0.0, 0.0f, kQhMax));
                s.qr[c] = S(clamp_store<S>(double(s.qr[c]) + double(dqr[c]) * h, 0.0, 0.0f, kQhMax));
            }
}

// Microphysics over an interval with the reference's state-dependent
// subcycling (step_microphysics, microphysics_step.cpp:3619,3700-3750): one
// step when the interval fits under the state's limit; otherwise steps of
// min(remaining, limit), the limit recomputed from the state after each step
// that leaves time remaining. The step lengths are appended to `lengths`
// when it is given.
template <class S>
void microphysics_step(const Grid& g, const Config& cfg, State<S>& s, double dt,
                       std::vector<double>* lengths = nullptr)
{
    double limit = sedimentation_limit(g, cfg, s);
    if (dt <= limit)
    {
        microphysics_once(g, cfg, s, dt);
        if (lengths) lengths->push_back(dt);
        return;
    }
    double elapsed = 0.0;
    int n = 0;
    while (elapsed < dt)
    {
        if (++n > 100000) throw std::runtime_error("microphysics interval requires too many substeps");
        const double h = std::min(dt - elapsed, limit);
        if (!(h > 0.0) || elapsed + h <= elapsed) throw std::runtime_error("microphysics substep cannot advance");
        microphysics_once(g, cfg, s, h);
        if (lengths) lengths->push_back(h);
        elapsed += h;
        if (elapsed < dt) limit = sedimentation_limit(g, cfg, s);
    }
}

// ----------------------------------------------------------------------
// Rayleigh sponges after the EOS projection (dynamics.cpp:7325-7340 order):
// the upper sponge relaxes u, v to the base wind, w to 0 and theta to
// theta0 with alpha(z) = alpha_max sin^2[(pi/2) min(1, (z - z_d)/(z_top - z_d))]
// (w at its face height z_w = (k+1) dz) (rayleigh_damping.cpp:425-465,
// 985-1040); the w-only absorber relaxes w in the top N cells with
// alpha = alpha_w sin^2[(pi/2)(1 - d/(N-1))], d = distance from the lid
// (:467-486, 1055-1080); the lateral band relaxes u, v, w, theta with
// alpha = alpha_l sin^2[(pi/2)(1 - d/W)], d = distance to the nearest
// lateral face in cells (:1174-1183, 1360-1440). Each pass uses the exact
// decay factor 1 - exp(-alpha dt) (:145-163) and rounds to storage.
// ----------------------------------------------------------------------
struct SpongeProfiles
{
    std::vector<double> upper, upper_w, w_only, lateral;
};

inline SpongeProfiles sponge_profiles(const Grid& g, const Config& cfg, double dt)
{
    const double pi_half = 1.5707963267948966;
    SpongeProfiles sp;
    sp.upper.assign(g.nz, 0.0);
    sp.upper_w.assign(g.nz, 0.0);
    sp.w_only.assign(g.nz, 0.0);
    auto factor = [dt](double alpha) { return alpha > 0.0 ? -std::expm1(-alpha * dt) : 0.0; };
    const double z_top = g.nz * g.dz, depth = z_top - cfg.sponge_z_d;
    auto alpha_at = [&](double z) {
        if (depth <= 0.0 || z <= cfg.sponge_z_d) return 0.0;
        const double s = std::sin(pi_half * std::min(1.0, (z - cfg.sponge_z_d) / depth));
        return cfg.sponge_alpha * s * s;
    };
    for (int k = 0; k < g.nz; ++k)
    {
        sp.upper[k] = factor(alpha_at((k + 0.5) * g.dz));
        sp.upper_w[k] = factor(alpha_at((k + 1.0) * g.dz));
    }
    const int n = std::min(std::max(cfg.sponge_w_cells, 0), g.nz);
    for (int d = 0; d < n && cfg.sponge_w_alpha > 0.0; ++d)
    {
        const double frac = n > 1 ? 1.0 - double(d) / double(n - 1) : 1.0;
        const double s = std::sin(pi_half * frac);
        sp.w_only[g.nz - 1 - d] = factor(cfg.sponge_w_alpha * s * s);
    }
    const int W = cfg.sponge_lateral_alpha > 0.0 ? std::max(cfg.sponge_lateral_cells, 0) : 0;
    sp.lateral.assign(W, 0.0);
    for (int d = 0; d < W; ++d)
    {
        const double s = std::sin(pi_half * (1.0 - double(d) / double(W)));
        sp.lateral[d] = factor(cfg.sponge_lateral_alpha * s * s);
    }
    return sp;
}

template <class S>
void apply_sponges(const Grid& g, const Base& b, const Config& cfg, State<S>& s, double dt)
{
    if (!cfg.sponge) return;
    const SpongeProfiles sp = sponge_profiles(g, cfg, dt);
    const int W = int(sp.lateral.size());
    auto relax = [](S& x, double target, double f) {
        if (f > 0.0) x = S(double(x) + f * (target - double(x)));
    };
#pragma omp parallel for collapse(2) schedule(static)
    for (int i = 0; i < g.nx; ++i)
        for (int j = 0; j < g.ny; ++j)
        {
            const int dist = std::min(std::min(i, g.nx - 1 - i), std::min(j, g.ny - 1 - j));
            for (int k = 0; k < g.nz; ++k)
            {
                const std::size_t c = g.at(i, j, k);
                relax(s.u[c], b.u0[k], sp.upper[k]);
                relax(s.v[c], b.v0[k], sp.upper[k]);
                relax(s.w[c], 0.0, sp.upper_w[k]);
                relax(s.theta[c], b.theta0[k], sp.upper[k]);
                relax(s.w[c], 0.0, sp.w_only[k]);
                if (dist < W)
                {
                    const double f = sp.lateral[dist];
                    relax(s.u[c], b.u0[k], f);
                    relax(s.v[c], b.v0[k], f);
                    relax(s.w[c], 0.0, f);
                    relax(s.theta[c], b.theta0[k], f);
                }
            }
        }
}

// ----------------------------------------------------------------------
// Smagorinsky SGS on the C-grid, flat terrain, stability correction "none"
// (src/turbulence/schemes/smagorinsky/smagorinsky.cpp:495-580,
// src/turbulence/base/eddy_viscosity.cpp:1253-1527 closure, 1532-1948
// momentum stress divergence, 1953-2090 scalar flux divergence;
// src/core/orchestration/dynamics/dynamics.cpp:7160-7262 split application,
// sgs_apply.cpp:610-790 commit). Each SGS substep (min(remaining, dt_sgs))
// recomputes nu_t = min(Cs^2 Delta^2 |S|, nu_max) at every cell centre from the
// staggered strain (squares of the edge strains averaged to the centre),
// K_theta = min(nu_t / Pr_t, K_max), K_q = min(nu_t / Sc_t, K_max), then the
// conservative live-density tendencies: momentum from the stress divergence
// (1/rho_face) d_j(2 rho nu S_ij) with edge viscosities, zero shear stress at
// the ground and lid; theta and qv from (1/rho) d_j(rho K d_j phi) with
// zero-flux ground/lid and closed lateral walls; and commits
// x += dt_sub * dx/dt with the storage clamps. The coefficient and tendency
// fields are stored in the storage type, as the reference's FP32 fields.
// ----------------------------------------------------------------------
struct SgsNeighbours
{
    int nx, ny;
    bool wx, wy;
    int ip(int i) const { return wx ? (i <= 1 ? nx - 2 : std::min(i - 1, nx - 2)) : std::max(i - 1, 0); }
    int in(int i) const { return wx ? (i >= nx - 2 ? 1 : std::max(i + 1, 1)) : std::min(i + 1, nx - 1); }
    int jp(int j) const { return wy ? (j <= 1 ? ny - 2 : std::min(j - 1, ny - 2)) : std::max(j - 1, 0); }
    int jn(int j) const { return wy ? (j >= ny - 2 ? 1 : std::max(j + 1, 1)) : std::min(j + 1, ny - 1); }
};

inline SgsNeighbours sgs_neighbours(const Grid& g, const Config& cfg)
{
    return SgsNeighbours{g.nx, g.ny, cfg.bc_west == kPeriodic && cfg.bc_east == kPeriodic,
                         cfg.bc_south == kPeriodic && cfg.bc_north == kPeriodic};
}

inline bool open_lateral(int bc) { return bc == kRadiative || bc == kZeroGradient; }

// |S|^2 at the centre of cell (i, j, k) (eddy_viscosity.cpp:1253-1400).
template <class S>
double sgs_strain_sq(const Grid& g, const SgsNeighbours& nb, int mode, const State<S>& s, int i, int j, int k)
{
    const int NZ = g.nz;
    const double dx = g.dx, dy = g.dy, dz = g.dz;
    auto U = [&](int a, int b, int c) { return double(s.u[g.at(a, b, c)]); };
    auto V = [&](int a, int b, int c) { return double(s.v[g.at(a, b, c)]); };
    auto W = [&](int a, int b, int c) { return c < 0 ? 0.0 : double(s.w[g.at(a, b, c)]); };
    const int im = nb.ip(i), jm = nb.jp(j);
    const double S11 = (U(i, j, k) - U(im, j, k)) / dx;
    const double S22 = (V(i, j, k) - V(i, jm, k)) / dy;
    const double S33 = (W(i, j, k) - W(i, j, k - 1)) / dz;
    auto s12 = [&](int ii, int jj) {
        const int jj1 = nb.jn(jj), ii1 = nb.in(ii);
        return 0.5 * ((U(ii, jj1, k) - U(ii, jj, k)) / dy + (V(ii1, jj, k) - V(ii, jj, k)) / dx);
    };
    struct G2 { double a, b; };  // (d(u|v)/dz, dw/d(x|y)) at an edge
    auto xz = [&](int ii, int kk) {
        if (kk < 0 || kk >= NZ - 1) return G2{0.0, 0.0};
        const int ii1 = nb.in(ii);
        return G2{(U(ii, j, kk + 1) - U(ii, j, kk)) / dz, (W(ii1, j, kk) - W(ii, j, kk)) / dx};
    };
    auto yz = [&](int jj, int kk) {
        if (kk < 0 || kk >= NZ - 1) return G2{0.0, 0.0};
        const int jj1 = nb.jn(jj);
        return G2{(V(i, jj, kk + 1) - V(i, jj, kk)) / dz, (W(i, jj1, kk) - W(i, jj, kk)) / dy};
    };
    const double a = s12(im, jm), b = s12(im, j), c = s12(i, jm), d = s12(i, j);
    const double S12_sq = 0.25 * (a * a + b * b + c * c + d * d);
    const G2 xa = xz(im, k - 1), xb = xz(im, k), xc = xz(i, k - 1), xd = xz(i, k);
    const G2 ya = yz(jm, k - 1), yb = yz(jm, k), yc = yz(j, k - 1), yd = yz(j, k);
    auto avg_b = [](G2 p, G2 q, G2 r, G2 t) { return 0.25 * (p.b * p.b + q.b * q.b + r.b * r.b + t.b * t.b); };
    auto avg_a = [](G2 p, G2 q, G2 r, G2 t) { return 0.25 * (p.a * p.a + q.a * q.a + r.a * r.a + t.a * t.a); };
    if (mode == kSgsHorizontal)
        return 2.0 * (S11 * S11 + S22 * S22) + 4.0 * S12_sq + avg_b(xa, xb, xc, xd) + avg_b(ya, yb, yc, yd);
    if (mode == kSgsVertical) return avg_a(xa, xb, xc, xd) + avg_a(ya, yb, yc, yd) + 2.0 * S33 * S33;
    auto sym = [](G2 p) { const double e = 0.5 * (p.a + p.b); return e * e; };
    const double S13_sq = 0.25 * (sym(xa) + sym(xb) + sym(xc) + sym(xd));
    const double S23_sq = 0.25 * (sym(ya) + sym(yb) + sym(yc) + sym(yd));
    return 2.0 * (S11 * S11 + S22 * S22 + S33 * S33) + 4.0 * (S12_sq + S13_sq + S23_sq);
}

inline double sgs_filter_width(const Grid& g, const Config& cfg)
{
    return cfg.sgs_filter == kFilterDx ? std::sqrt(g.dx * g.dy) : std::cbrt(g.dx * g.dy * g.dz);
}

// nu_t, K_theta, K_q at every cell (smagorinsky.cpp:495-580).
template <class S>
void sgs_closure(const Grid& g, const Config& cfg, const State<S>& s, std::vector<S>& nu, std::vector<S>& kth,
                 std::vector<S>& kq)
{
    const SgsNeighbours nb = sgs_neighbours(g, cfg);
    const double delta = sgs_filter_width(g, cfg);
    const double c2d2 = cfg.sgs_cs * cfg.sgs_cs * delta * delta;
    nu.assign(g.cells(), S(0));
    kth.assign(g.cells(), S(0));
    kq.assign(g.cells(), S(0));
#pragma omp parallel for collapse(2) schedule(static)
    for (int i = 0; i < g.nx; ++i)
        for (int j = 0; j < g.ny; ++j)
            for (int k = 0; k < g.nz; ++k)
            {
                const std::size_t c = g.at(i, j, k);
                const double n = std::min(c2d2 * std::sqrt(sgs_strain_sq(g, nb, cfg.sgs_mode, s, i, j, k)),
                                          cfg.sgs_nu_max);
                nu[c] = S(n);
                kth[c] = S(std::min(n / cfg.sgs_pr_t, cfg.sgs_k_max));
                kq[c] = S(std::min(n / cfg.sgs_sc_t, cfg.sgs_k_max));
            }
}

// du/dt, dv/dt, dw/dt (eddy_viscosity.cpp:1532-1948); zero outside the face ranges.
template <class S>
void sgs_momentum_tendencies(const Grid& g, const Config& cfg, const State<S>& s, const std::vector<S>& nu,
                             std::vector<S>& du, std::vector<S>& dv, std::vector<S>& dw)
{
    const SgsNeighbours nb = sgs_neighbours(g, cfg);
    const bool use_h = cfg.sgs_mode != kSgsVertical, use_v = cfg.sgs_mode != kSgsHorizontal;
    const int NX = g.nx, NY = g.ny, NZ = g.nz;
    const double dx = g.dx, dy = g.dy, dz = g.dz;
    auto U = [&](int a, int b, int c) { return double(s.u[g.at(a, b, c)]); };
    auto V = [&](int a, int b, int c) { return double(s.v[g.at(a, b, c)]); };
    auto W = [&](int a, int b, int c) { return c < 0 ? 0.0 : double(s.w[g.at(a, b, c)]); };
    auto R = [&](int a, int b, int c) { return double(s.rho[g.at(a, b, c)]); };
    auto MU = [&](int a, int b, int c) { return R(a, b, c) * double(nu[g.at(a, b, c)]); };
    auto mu_xy = [&](int a, int b, int c) {
        const int a1 = nb.in(a), b1 = nb.jn(b);
        return 0.25 * (MU(a, b, c) + MU(a1, b, c) + MU(a, b1, c) + MU(a1, b1, c));
    };
    // uniform dz: the thickness-weighted average of the two levels is their mean
    auto mu_xz = [&](int a, int b, int c) {
        const int a1 = nb.in(a);
        const double lower = 0.5 * (MU(a, b, c) + MU(a1, b, c)), upper = 0.5 * (MU(a, b, c + 1) + MU(a1, b, c + 1));
        return (dz * lower + dz * upper) / (dz + dz);
    };
    auto mu_yz = [&](int a, int b, int c) {
        const int b1 = nb.jn(b);
        const double lower = 0.5 * (MU(a, b, c) + MU(a, b1, c)), upper = 0.5 * (MU(a, b, c + 1) + MU(a, b1, c + 1));
        return (dz * lower + dz * upper) / (dz + dz);
    };
    du.assign(g.cells(), S(0));
    dv.assign(g.cells(), S(0));
    dw.assign(g.cells(), S(0));
    const int i_hi = nb.wx ? NX - 2 : NX - 3;
#pragma omp parallel for collapse(2) schedule(static)
    for (int i = 1; i <= i_hi; ++i)
        for (int j = 1; j <= NY - 2; ++j)
        {
            const int jm = nb.jp(j), im = nb.ip(i), ip = nb.in(i);
            double lower_T13 = 0.0;
            for (int k = 0; k < NZ - 1; ++k)
            {
                double div = 0.0;
                if (use_h)
                {
                    const double S11_r = (U(ip, j, k) - U(i, j, k)) / dx, S11_l = (U(i, j, k) - U(im, j, k)) / dx;
                    div += (2.0 * MU(ip, j, k) * S11_r - 2.0 * MU(i, j, k) * S11_l) / dx;
                    auto T12 = [&](int je) {
                        const int i1 = nb.in(i), je1 = nb.jn(je);
                        const double S12 = 0.5 * ((U(i, je1, k) - U(i, je, k)) / dy + (V(i1, je, k) - V(i, je, k)) / dx);
                        return 2.0 * mu_xy(i, je, k) * S12;
                    };
                    div += (T12(j) - T12(jm)) / dy;
                }
                if (use_v)
                {
                    auto T13 = [&](int ke) {
                        if (ke < 0 || ke >= NZ - 1) return 0.0;
                        const double dudz = (U(i, j, ke + 1) - U(i, j, ke)) / dz;
                        const double sg = use_h ? dudz + (W(ip, j, ke) - W(i, j, ke)) / dx : dudz;
                        return mu_xz(i, j, ke) * sg;
                    };
                    const double upper = T13(k);
                    div += (upper - lower_T13) / dz;
                    lower_T13 = upper;
                }
                du[g.at(i, j, k)] = S(div / (0.5 * (R(i, j, k) + R(nb.in(i), j, k))));
            }
        }
#pragma omp parallel for collapse(2) schedule(static)
    for (int i = 1; i <= NX - 2; ++i)
        for (int j = 1; j <= NY - 2; ++j)
        {
            if (!nb.wy && j >= NY - 2) continue;
            const int jm = nb.jp(j), jp = nb.jn(j);
            double lower_T23 = 0.0;
            for (int k = 0; k < NZ - 1; ++k)
            {
                double div = 0.0;
                if (use_h)
                {
                    const double S22_n = (V(i, jp, k) - V(i, j, k)) / dy, S22_s = (V(i, j, k) - V(i, jm, k)) / dy;
                    div += (2.0 * MU(i, jp, k) * S22_n - 2.0 * MU(i, j, k) * S22_s) / dy;
                    auto T21 = [&](int ie) {
                        const int ie1 = nb.in(ie), j1 = nb.jn(j);
                        const double S12 = 0.5 * ((U(ie, j1, k) - U(ie, j, k)) / dy + (V(ie1, j, k) - V(ie, j, k)) / dx);
                        return 2.0 * mu_xy(ie, j, k) * S12;
                    };
                    div += (T21(i) - T21(nb.ip(i))) / dx;
                }
                if (use_v)
                {
                    auto T23 = [&](int ke) {
                        if (ke < 0 || ke >= NZ - 1) return 0.0;
                        const int j1 = nb.jn(j);
                        const double dvdz = (V(i, j, ke + 1) - V(i, j, ke)) / dz;
                        const double sg = use_h ? dvdz + (W(i, j1, ke) - W(i, j, ke)) / dy : dvdz;
                        return mu_yz(i, j, ke) * sg;
                    };
                    const double upper = T23(k);
                    div += (upper - lower_T23) / dz;
                    lower_T23 = upper;
                }
                dv[g.at(i, j, k)] = S(div / (0.5 * (R(i, j, k) + R(i, nb.jn(j), k))));
            }
        }
#pragma omp parallel for collapse(2) schedule(static)
    for (int i = 1; i <= NX - 2; ++i)
        for (int j = 1; j <= NY - 2; ++j)
        {
            double lower_T33 = use_v ? 2.0 * MU(i, j, 0) * ((W(i, j, 0) - W(i, j, -1)) / dz) : 0.0;
            for (int k = 0; k < NZ - 1; ++k)
            {
                const double mass = 0.5 * (R(i, j, k) * dz + R(i, j, k + 1) * dz);
                double force = 0.0;
                if (use_v)
                {
                    const double T33_t = 2.0 * MU(i, j, k + 1) * ((W(i, j, k + 1) - W(i, j, k)) / dz);
                    force += T33_t - lower_T33;
                    lower_T33 = T33_t;
                }
                if (use_h)
                {
                    auto T31 = [&](int ie) {
                        const int ie1 = nb.in(ie);
                        const double dwdx = (W(ie1, j, k) - W(ie, j, k)) / dx;
                        const double sg = use_v ? (U(ie, j, k + 1) - U(ie, j, k)) / dz + dwdx : dwdx;
                        return mu_xz(ie, j, k) * sg;
                    };
                    force += dz * (T31(i) - T31(nb.ip(i))) / dx;
                    auto T32 = [&](int je) {
                        const int je1 = nb.jn(je);
                        const double dwdy = (W(i, je1, k) - W(i, je, k)) / dy;
                        const double sg = use_v ? (V(i, je, k + 1) - V(i, je, k)) / dz + dwdy : dwdy;
                        return mu_yz(i, je, k) * sg;
                    };
                    force += dz * (T32(j) - T32(nb.jp(j))) / dy;
                }
                dw[g.at(i, j, k)] = S(force / mass);
            }
        }
}

// d(phi)/dt of a scalar with diffusivity K (eddy_viscosity.cpp:1953-2090).
template <class S>
void sgs_scalar_tendency(const Grid& g, const Config& cfg, const State<S>& s, const std::vector<S>& K,
                         const std::vector<S>& phi, std::vector<S>& out)
{
    const SgsNeighbours nb = sgs_neighbours(g, cfg);
    const bool use_h = cfg.sgs_mode != kSgsVertical, use_v = cfg.sgs_mode != kSgsHorizontal;
    const int NX = g.nx, NY = g.ny, NZ = g.nz;
    const double dx = g.dx, dy = g.dy, dz = g.dz;
    const bool ow = open_lateral(cfg.bc_west), oe = open_lateral(cfg.bc_east);
    const bool os = open_lateral(cfg.bc_south), on = open_lateral(cfg.bc_north);
    auto T = [&](int a, int b, int c) { return double(s.rho[g.at(a, b, c)]) * double(K[g.at(a, b, c)]); };
    auto P = [&](int a, int b, int c) { return double(phi[g.at(a, b, c)]); };
    out.assign(g.cells(), S(0));
#pragma omp parallel for collapse(2) schedule(static)
    for (int i = 1; i <= NX - 2; ++i)
        for (int j = 1; j <= NY - 2; ++j)
        {
            const int jm = !nb.wy && !os && j == 1 ? j : nb.jp(j);
            const int jp = !nb.wy && !on && j == NY - 2 ? j : nb.jn(j);
            const int im = !nb.wx && !ow && i == 1 ? i : nb.ip(i);
            const int ip = !nb.wx && !oe && i == NX - 2 ? i : nb.in(i);
            for (int k = 0; k < NZ; ++k)
            {
                double div = 0.0;
                if (use_h)
                {
                    const double Fe = 0.5 * (T(i, j, k) + T(ip, j, k)) * (P(ip, j, k) - P(i, j, k)) / dx;
                    const double Fw = 0.5 * (T(im, j, k) + T(i, j, k)) * (P(i, j, k) - P(im, j, k)) / dx;
                    const double Fn = 0.5 * (T(i, j, k) + T(i, jp, k)) * (P(i, jp, k) - P(i, j, k)) / dy;
                    const double Fs = 0.5 * (T(i, jm, k) + T(i, j, k)) * (P(i, j, k) - P(i, jm, k)) / dy;
                    div += (Fe - Fw) / dx + (Fn - Fs) / dy;
                }
                if (use_v)
                {
                    const double Ft = k < NZ - 1 ? 0.5 * (T(i, j, k) + T(i, j, k + 1)) * (P(i, j, k + 1) - P(i, j, k)) / dz
                                                 : 0.0;
                    const double Fb = k > 0 ? 0.5 * (T(i, j, k - 1) + T(i, j, k)) * (P(i, j, k) - P(i, j, k - 1)) / dz
                                            : 0.0;
                    div += (Ft - Fb) / dz;
                }
                out[g.at(i, j, k)] = S(div / double(s.rho[g.at(i, j, k)]));
            }
        }
}

// One SGS substep: closure, tendencies, commit with the storage clamps
// (sgs_apply.cpp:610-790).
template <class S>
void sgs_substep(const Grid& g, const Config& cfg, State<S>& s, double h)
{
    std::vector<S> nu, kth, kq, du, dv, dw, dth, dqv;
    sgs_closure(g, cfg, s, nu, kth, kq);
    sgs_momentum_tendencies(g, cfg, s, nu, du, dv, dw);
    sgs_scalar_tendency(g, cfg, s, kth, s.theta, dth);
    sgs_scalar_tendency(g, cfg, s, kq, s.qv, dqv);
    auto fin = [](double x) { return std::isfinite(x) ? x : 0.0; };
#pragma omp parallel for schedule(static)
    for (std::size_t c = 0; c < g.cells(); ++c)
    {
        s.u[c] = S(clamp_store<S>(double(s.u[c]) + fin(double(du[c])) * h, 0.0, -kWindH, kWindH));
        s.v[c] = S(clamp_store<S>(double(s.v[c]) + fin(double(dv[c])) * h, 0.0, -kWindH, kWindH));
        s.w[c] = S(clamp_store<S>(double(s.w[c]) + fin(double(dw[c])) * h, 0.0, -kWindV, kWindV));
        s.theta[c] = S(clamp_store<S>(double(s.theta[c]) + fin(double(dth[c])) * h, 300.0, kThetaMin, kThetaMax));
        s.qv[c] = S(clamp_store<S>(double(s.qv[c]) + fin(double(dqv[c])) * h, 0.0, 0.0f, kQvMax));
    }
}

// Substep lengths of the split SGS integration: dt_sgs each, the remainder last
// (dynamics.cpp:7160-7220; the explicit stability limit is far above dt_sgs for
// nu_t, K <= 1000 m^2/s at the grid spacings used here and is not applied).
inline std::vector<double> sgs_substeps(const Config& cfg)
{
    std::vector<double> out;
    double done = 0.0;
    while (done < cfg.dt && out.size() < 1000)
    {
        const double h = std::min(cfg.dt - done, cfg.sgs_dt);
        if (!(h > 64.0 * std::numeric_limits<double>::epsilon() * std::max(1.0, cfg.dt))) break;
        out.push_back(h);
        done += h;
    }
    return out;
}

template <class S>
void apply_sgs(const Grid& g, const Config& cfg, State<S>& s)
{
    if (!cfg.sgs) return;
    for (double h : sgs_substeps(cfg)) sgs_substep(g, cfg, s, h);
}

// EOS pressure projection after physics (src/core/orchestration/dynamics/eos_apply.cpp:50-110).
template <class S>
void eos_projection(State<S>& s)
{
    const std::size_t n = s.p.size();
#pragma omp parallel for schedule(static)
    for (std::size_t c = 0; c < n; ++c)
    {
        const double r = double(s.rho[c]), th = double(s.theta[c]);
        if (!(r > 0.0) || !(th > 0.0)) continue;
        const double pd = eos_pressure(r, th);
        if (!std::isfinite(pd) || pd <= 0.0) continue;
        s.p[c] = S(clamp_store<S>(pd, 101325.0, kPMin, kPMax));
    }
}

// ----------------------------------------------------------------------
// One large step in the reference order (dynamics.cpp:6585-7420):
// slow Heun RK2 + apply, zero-time BC closure and KW78 seed, N acoustic
// substeps with mass-flux accumulation and per-substep BCs, WENO5 transport
// of theta, qv, qc, qr with the retained fluxes, Kessler, EOS projection and
// the final BC closure.
// ----------------------------------------------------------------------
template <class S>
void apply_diffusion(const Grid& g, const Config& cfg, State<S>& s);

template <class S>
int large_step(const Grid& g, const Base& b, const Config& cfg, State<S>& s)
{
    const int N = resolve_substeps(g, b, cfg);
    const double dts = cfg.dt / N;
    const double weight = dts / cfg.dt;
    slow_rk2_and_apply(g, b, cfg, s);
    apply_boundaries(g, cfg, s, 0.0);
    seed_radiation(g, s);
    std::vector<S> mfx(g.cells(), S(0)), mfy(g.cells(), S(0)), mfz(g.cells(), S(0));
    const std::vector<S> rho_begin = s.rho;
    std::vector<double> wz;
    for (int n = 0; n < N; ++n)
    {
        if (cfg.acoustic_mode == kAcousticHEVI)
        {
            std::vector<S> zero(g.cells(), S(0));
            accumulate_mass_flux(g, cfg, s, weight, mfx, mfy, zero);
            acoustic_substep_hevi(g, b, s, dts, cfg.hevi_beta, wz);
            for (std::size_t c = 0; c < g.cells(); ++c) mfz[c] = S(double(mfz[c]) + weight * wz[c]);
        }
        else
        {
            accumulate_mass_flux(g, cfg, s, weight, mfx, mfy, mfz);
            acoustic_substep_explicit(g, b, s, dts);
        }
        apply_boundaries(g, cfg, s, dts);
    }
    std::vector<S> tmp;
    transport_scalar(g, cfg, s.theta, tmp, mfx, mfy, mfz, rho_begin, s.rho, cfg.dt, false);
    s.theta.swap(tmp);
    transport_scalar(g, cfg, s.qv, tmp, mfx, mfy, mfz, rho_begin, s.rho, cfg.dt, true);
    s.qv.swap(tmp);
    transport_scalar(g, cfg, s.qc, tmp, mfx, mfy, mfz, rho_begin, s.rho, cfg.dt, true);
    s.qc.swap(tmp);
    transport_scalar(g, cfg, s.qr, tmp, mfx, mfy, mfz, rho_begin, s.rho, cfg.dt, true);
    s.qr.swap(tmp);
    microphysics_step(g, cfg, s, cfg.dt);
    apply_sgs(g, cfg, s);
    apply_diffusion(g, cfg, s);
    eos_projection(s);
    apply_sponges(g, b, cfg, s, cfg.dt);
    apply_boundaries(g, cfg, s, 0.0);
    return N;
}

// ---- explicit Laplacian diffusion tendency of one field
// (src/numerics/diffusion/schemes/explicit/explicit.cpp: control-volume
// domains :52-87, horizontal component :119-194, vertical component :263-339,
// scalar and momentum wrappers :1183-1433; the device contract
// include/compute/dispatch/diffusion.hpp:29-69). loc: 0 scalar, 1 u face,
// 2 v face, 3 w face. Every cell of `tend` is written: zero outside the
// field's control-volume domain; inside, the vertical part (when K_v > 0 or
// kv is given) rounded to the storage type, then the horizontal part (K_h > 0)
// added in FP64 and rounded again, as the reference adds the two passes. ----
struct DiffusionDomain
{
    int i0, i1, j0, j1, k0, k1;
};

inline DiffusionDomain diffusion_domain(int nx, int ny, int nz, int loc, bool c_grid, bool px, bool py)
{
    if (!c_grid || loc == 0) return {1, nx - 1, 1, ny - 1, c_grid ? 0 : 1, c_grid ? nz : nz - 1};
    if (loc == 1) return {1, px ? nx - 1 : nx - 2, 1, ny - 1, 0, nz - 1};
    if (loc == 2) return {1, nx - 1, 1, py ? ny - 1 : ny - 2, 0, nz - 1};
    return {1, nx - 1, 1, ny - 1, 0, nz - 1};
}

template <class S>
void diffusion_tendency(int nx, int ny, int nz, double dx, double dy, const double* dz, const S* field,
                        const S* rho, const S* kv, double K_h, double K_v, int loc, bool c_grid, bool px, bool py,
                        S* tend)
{
    const auto at = [=](int i, int j, int k) { return (std::size_t(i) * ny + j) * nz + k; };
    const std::size_t n = std::size_t(nx) * ny * nz;
    for (std::size_t c = 0; c < n; ++c) tend[c] = S(0);
    const DiffusionDomain d = diffusion_domain(nx, ny, nz, loc, c_grid, px, py);
    if (d.i0 >= d.i1 || d.j0 >= d.j1 || d.k0 >= d.k1) return;
    const bool stag = c_grid && loc != 0;
    const auto rho_at = [&](int i, int j, int k) { return double(rho[at(i, j, k)]); };
    const auto kv_at = [&](int i, int j, int k) { return kv ? double(kv[at(i, j, k)]) : K_v; };
    // density of the field's control volume (u and v: the two cells it joins,
    // the partner wrapping to column / row 1 on a periodic seam; w: the two
    // cells weighted by their thicknesses)
    const auto comp_rho = [&](int i, int j, int k) {
        if (!stag) return rho_at(i, j, k);
        if (loc == 1) return 0.5 * (rho_at(i, j, k) + rho_at(px && i == nx - 2 ? 1 : i + 1, j, k));
        if (loc == 2) return 0.5 * (rho_at(i, j, k) + rho_at(i, py && j == ny - 2 ? 1 : j + 1, k));
        return (rho_at(i, j, k) * dz[k] + rho_at(i, j, k + 1) * dz[k + 1]) / (dz[k] + dz[k + 1]);
    };
    // vertical component
    if (K_v > 0.0 || kv != nullptr)
    {
        const auto cell_tr = [&](int i, int j, int k) { return rho_at(i, j, k) * kv_at(i, j, k); };
        const auto comp_tr = [&](int i, int j, int k) {
            if (!c_grid) return cell_tr(i, j, k);
            if (loc == 1) return 0.5 * (cell_tr(i, j, k) + cell_tr(px && i == nx - 2 ? 1 : i + 1, j, k));
            if (loc == 2) return 0.5 * (cell_tr(i, j, k) + cell_tr(i, py && j == ny - 2 ? 1 : j + 1, k));
            if (loc == 3) return 0.5 * (cell_tr(i, j, k) + cell_tr(i, j, k + 1));
            return cell_tr(i, j, k);
        };
        const bool wc = c_grid && loc == 3;
        // transport through the face above level lo, the distance between the
        // two points it joins and the control mass of level k
        const auto face_tr = [&](int i, int j, int lo) {
            return wc ? cell_tr(i, j, lo + 1) : 0.5 * (comp_tr(i, j, lo) + comp_tr(i, j, lo + 1));
        };
        const auto dist = [&](int lo) { return wc ? dz[lo + 1] : 0.5 * (dz[lo] + dz[lo + 1]); };
        const auto mass = [&](int i, int j, int k) {
            if (wc) return 0.5 * (rho_at(i, j, k) * dz[k] + rho_at(i, j, k + 1) * dz[k + 1]);
            const double r = c_grid ? comp_rho(i, j, k) : rho_at(i, j, k);
            return r * dz[k];
        };
        for (int i = d.i0; i < d.i1; ++i)
            for (int j = d.j0; j < d.j1; ++j)
                for (int k = d.k0; k < d.k1; ++k)
                {
                    const double q = double(field[at(i, j, k)]);
                    double down = 0.0, up = 0.0;
                    if (k > d.k0) down = face_tr(i, j, k - 1) * (q - double(field[at(i, j, k - 1)])) / dist(k - 1);
                    if (k < d.k1 - 1) up = face_tr(i, j, k) * (double(field[at(i, j, k + 1)]) - q) / dist(k);
                    tend[at(i, j, k)] = S(double(tend[at(i, j, k)]) + (up - down) / mass(i, j, k));
                }
    }
    // horizontal component, closed (zero flux) at the domain's lateral faces
    // unless periodic, where the neighbour wraps inside the domain
    if (K_h > 0.0)
    {
        const double idx2 = 1.0 / (dx * dx), idy2 = 1.0 / (dy * dy);
        for (int i = d.i0; i < d.i1; ++i)
            for (int j = d.j0; j < d.j1; ++j)
                for (int k = d.k0; k < d.k1; ++k)
                {
                    const bool wcl = i == d.i0 && !px, ecl = i == d.i1 - 1 && !px;
                    const bool scl = j == d.j0 && !py, ncl = j == d.j1 - 1 && !py;
                    const int iw = i == d.i0 && px ? d.i1 - 1 : i - 1, ie = i == d.i1 - 1 && px ? d.i0 : i + 1;
                    const int js = j == d.j0 && py ? d.j1 - 1 : j - 1, jn = j == d.j1 - 1 && py ? d.j0 : j + 1;
                    const double rc = comp_rho(i, j, k);
                    const double rw = wcl ? 0.0 : 0.5 * (rc + comp_rho(iw, j, k));
                    const double re = ecl ? 0.0 : 0.5 * (rc + comp_rho(ie, j, k));
                    const double rs = scl ? 0.0 : 0.5 * (rc + comp_rho(i, js, k));
                    const double rn = ncl ? 0.0 : 0.5 * (rc + comp_rho(i, jn, k));
                    const double q = double(field[at(i, j, k)]);
                    const double xd = ((ecl ? 0.0 : re * (double(field[at(ie, j, k)]) - q)) -
                                       (wcl ? 0.0 : rw * (q - double(field[at(iw, j, k)])))) *
                                      idx2;
                    const double yd = ((ncl ? 0.0 : rn * (double(field[at(i, jn, k)]) - q)) -
                                       (scl ? 0.0 : rs * (q - double(field[at(i, js, k)])))) *
                                      idy2;
                    tend[at(i, j, k)] = S(double(tend[at(i, j, k)]) + K_h * (xd + yd) / rc);
                }
    }
}

// ---- explicit diffusion in the large step (diffusion_step.cpp:679-1030):
// the tendencies of the state after the SGS, then every cell committed as
// field + dt tendency rounded to the storage type and clamped like the
// reference's commit (winds, theta, qv bounds of field_bounds.hpp). Sides are
// periodic when both faces of an axis are. ----
template <class S>
void apply_diffusion(const Grid& g, const Config& cfg, State<S>& s)
{
    if (cfg.diffusion == 0 || (!(cfg.diffusion_kh > 0.0) && !(cfg.diffusion_kv > 0.0))) return;
    const bool px = cfg.bc_west == kPeriodic && cfg.bc_east == kPeriodic;
    const bool py = cfg.bc_south == kPeriodic && cfg.bc_north == kPeriodic;
    const std::vector<double> dz(std::size_t(g.nz), g.dz);
    const std::size_t n = g.cells();
    auto tend = [&](const std::vector<S>& f, int loc) {
        std::vector<S> t(n);
        diffusion_tendency<S>(g.nx, g.ny, g.nz, g.dx, g.dy, dz.data(), f.data(), s.rho.data(), nullptr,
                              cfg.diffusion_kh, cfg.diffusion_kv, loc, true, px, py, t.data());
        return t;
    };
    if (cfg.diffusion & 2)
    {
        const std::vector<S> du = tend(s.u, 1), dv = tend(s.v, 2), dw = tend(s.w, 3);
        for (std::size_t c = 0; c < n; ++c)
        {
            s.u[c] = S(clamp_store<S>(double(s.u[c]) + double(du[c]) * cfg.dt, 0.0, -kWindH, kWindH));
            s.v[c] = S(clamp_store<S>(double(s.v[c]) + double(dv[c]) * cfg.dt, 0.0, -kWindH, kWindH));
            s.w[c] = S(clamp_store<S>(double(s.w[c]) + double(dw[c]) * cfg.dt, 0.0, -kWindV, kWindV));
        }
    }
    if (cfg.diffusion & 1)
    {
        const std::vector<S> dth = tend(s.theta, 0), dqv = tend(s.qv, 0);
        for (std::size_t c = 0; c < n; ++c)
        {
            s.theta[c] =
                S(clamp_store<S>(double(s.theta[c]) + double(dth[c]) * cfg.dt, 300.0, kThetaMin, kThetaMax));
            s.qv[c] = S(clamp_store<S>(double(s.qv[c]) + double(dqv[c]) * cfg.dt, 0.0, 0.0f, kQvMax));
        }
    }
}

}  // namespace oracle
