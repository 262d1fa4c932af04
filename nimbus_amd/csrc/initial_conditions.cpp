// Seeded analytic initial conditions: Weisman-Klemp-type sounding integrated
// hydrostatically in FP64, a quarter-circle hodograph, cosine-squared warm
// bubbles with relative-humidity-preserving moistening and counter-hashed
// uniform theta noise. Host code only; bit-reproducible on any host.
//
// The reference builds its storm ICs from a parametric sounding and a
// spherical Gaussian bubble (src/core/orchestration/dynamics/initial_conditions_cartesian.cpp:126-233,
// src/init/hodograph/sources/wk_param.cpp:23-61); this generator supplies the
// analytic sounding and ellipsoidal bubbles that the ensemble workloads specify,
// with the reference's conventions: base profiles in FP64 at cell centres, p left
// at the base pressure and rho diagnosed from the dry EOS, qv scaled by the
// saturation ratio inside the bubble.
#include "nimbus_hip.h"
#include "nimbus_internal.hpp"

#include <algorithm>
#include <cmath>
#include <vector>

namespace nimbus
{

namespace
{
constexpr double kPi = 3.14159265358979323846;

double wk_theta(const nimbus_sounding& s, double z)
{
    if (z <= s.z_tropo) return s.theta_sfc + (s.theta_tropo - s.theta_sfc) * std::pow(z / s.z_tropo, 1.25);
    return s.theta_tropo * std::exp(kG / (kCp * s.T_tropo) * (z - s.z_tropo));
}

double wk_rh(const nimbus_sounding& s, double z)
{
    if (z <= s.z_tropo) return s.rh_sfc - (s.rh_sfc - s.rh_top) * std::pow(z / s.z_tropo, 1.25);
    return s.rh_top;
}

double es_buck(double T)
{
    const double Tc = std::clamp(T, 150.0, 400.0) - kT0;
    return 611.21 * std::exp((18.678 - Tc / 234.5) * Tc / (257.14 + Tc));
}

double qvs(double T, double p)
{
    const double e = es_buck(T);
    return std::clamp(0.62198 * e / std::max(p - e, 1.0), 0.0, 0.2);
}

double sounding_qv(const nimbus_sounding& s, double z, double T, double p)
{
    return std::min(wk_rh(s, z) * qvs(T, p), s.qv_cap);
}

// Quarter-circle hodograph of constant shear magnitude along the curve:
// an arc of radius R over [0, curve_depth] followed by a straight tangent
// segment to shear_depth, |V(shear_depth) - V(0)| = shear. `curvature`
// blends toward a straight hodograph with the same endpoint.
void hodograph(const nimbus_sounding& s, double z, double& u, double& v)
{
    const double zc = std::clamp(s.curve_depth, 1.0, s.shear_depth);
    const double H = s.shear_depth;
    const double straight_len_ratio = (kPi / 2.0) * (H - zc) / zc;  // L / R at constant shear
    const double R = s.shear / std::sqrt((1.0 + straight_len_ratio) * (1.0 + straight_len_ratio) + 1.0);
    const double L = straight_len_ratio * R;
    const double zz = std::clamp(z, 0.0, H);
    double uq, vq;
    if (zz <= zc)
    {
        const double ang = kPi - (kPi / 2.0) * (zz / zc);
        uq = R + R * std::cos(ang);
        vq = R * std::sin(ang);
    }
    else
    {
        uq = R + L * (zz - zc) / (H - zc);
        vq = R;
    }
    const double uH = R + L, vH = R;
    const double f = zz / H;
    const double c = std::clamp(s.curvature, 0.0, 1.0);
    u = c * uq + (1.0 - c) * f * uH;
    v = c * vq + (1.0 - c) * f * vH;
    // Galilean shift by half the shear-depth wind keeps the storm near the centre.
    u -= 0.5 * uH;
    v -= 0.5 * vH;
}

std::uint64_t splitmix64(std::uint64_t x)
{
    x += 0x9E3779B97F4A7C15ull;
    x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
    x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
    return x ^ (x >> 31);
}
}  // namespace

double ic_uniform(std::uint64_t seed, std::uint32_t member, int i, int j, int k)
{
    std::uint64_t h = splitmix64(seed);
    h = splitmix64(h ^ (std::uint64_t(member) + 0x1000193ull));
    h = splitmix64(h ^ (std::uint64_t(std::uint32_t(i)) << 1));
    h = splitmix64(h ^ (std::uint64_t(std::uint32_t(j)) << 2));
    h = splitmix64(h ^ (std::uint64_t(std::uint32_t(k)) << 3));
    return double(h >> 11) * (1.0 / 9007199254740992.0);
}

int base_state(const nimbus_grid& g, const nimbus_sounding& s, double* rho0, double* p0, double* th0, double* qv0,
               double* u0, double* v0)
{
    if (g.nz < 2 || !(g.dz > 0.0)) return NIMBUS_EINVAL;
    // Integrate d(pi)/dz = -g / (cp theta_v), theta_v = theta (1 + 0.608 qv),
    // with qv from the RH law at the current (T, p), by a midpoint rule on a
    // fine sub-grid; record at z_k = (k + 1/2) dz.
    const double kappa = kRd / kCp;
    double z = 0.0;
    double pi = std::pow(s.p_sfc / kP00, kappa);
    auto thv_at = [&](double zq, double piq) {
        const double th = wk_theta(s, zq);
        const double p = kP00 * std::pow(piq, 1.0 / kappa);
        const double q = sounding_qv(s, zq, th * piq, p);
        return th * (1.0 + 0.608 * q);
    };
    for (int k = 0; k < g.nz; ++k)
    {
        const double zk = (k + 0.5) * g.dz;
        const int nsub = std::max(1, int(std::ceil((zk - z) / 10.0)));
        const double h = (zk - z) / nsub;
        for (int n = 0; n < nsub; ++n)
        {
            const double half = pi - 0.5 * h * kG / (kCp * thv_at(z, pi));
            pi = pi - h * kG / (kCp * thv_at(z + 0.5 * h, half));
            z += h;
        }
        const double th = wk_theta(s, zk);
        const double p = kP00 * std::pow(pi, 1.0 / kappa);
        const double T = th * pi;
        th0[k] = th;
        p0[k] = p;
        qv0[k] = sounding_qv(s, zk, T, p);
        rho0[k] = p / (kRd * T);
        double u, v;
        hodograph(s, zk, u, v);
        u0[k] = u;
        v0[k] = v;
    }
    return NIMBUS_OK;
}

template <class S>
int fill_from_base(const nimbus_grid& g, const nimbus_base_state& b, const nimbus_ic_params& prm,
                   const nimbus_host_state& out)
{
    if (g.nx < 5 || g.ny < 5 || g.nz < 4) return NIMBUS_EINVAL;
    if (!b.rho0 || !b.p0 || !b.theta0 || !b.qv0 || !b.u0 || !b.v0) return NIMBUS_EINVAL;
    for (int k = 0; k < g.nz; ++k)
        if (!(b.p0[k] > 0.0) || !(b.theta0[k] > 0.0) || !std::isfinite(b.p0[k]) || !std::isfinite(b.theta0[k]))
            return NIMBUS_EINVAL;
    const double *p0 = b.p0, *th0 = b.theta0, *qv0 = b.qv0, *u0 = b.u0, *v0 = b.v0;
    S* u = static_cast<S*>(out.u);
    S* v = static_cast<S*>(out.v);
    S* w = static_cast<S*>(out.w);
    S* rho = static_cast<S*>(out.rho);
    S* p = static_cast<S*>(out.p);
    S* th = static_cast<S*>(out.theta);
    S* qv = static_cast<S*>(out.qv);
    S* qc = static_cast<S*>(out.qc);
    S* qr = static_cast<S*>(out.qr);
    const double kappa = kRd / kCp;
    for (int i = 0; i < g.nx; ++i)
        for (int j = 0; j < g.ny; ++j)
            for (int k = 0; k < g.nz; ++k)
            {
                const std::size_t c = (std::size_t(i) * g.ny + j) * g.nz + k;
                const bool physical = i >= 1 && i < g.nx - 1 && j >= 1 && j < g.ny - 1;
                const double x = (i - 0.5) * g.dx, y = (j - 0.5) * g.dy, z = (k + 0.5) * g.dz;
                double dth_bubble = 0.0;
                for (int b = 0; b < prm.n_bubbles && b < 4; ++b)
                {
                    const nimbus_bubble& bb = prm.bubbles[b];
                    const double ax = (x - bb.xc) / bb.rh, ay = (y - bb.yc) / bb.rh, az = (z - bb.zc) / bb.rz;
                    const double beta = std::sqrt(ax * ax + ay * ay + az * az);
                    if (beta < 1.0)
                    {
                        const double cs = std::cos(0.5 * kPi * beta);
                        dth_bubble += bb.dtheta * cs * cs;
                    }
                }
                double noise = 0.0;
                if (physical && prm.noise_amp != 0.0)
                    noise = prm.noise_amp * (2.0 * ic_uniform(prm.seed, prm.member, i, j, k) - 1.0);
                if (!physical) dth_bubble = 0.0;
                const double pi_k = std::pow(p0[k] / kP00, kappa);
                const double theta = th0[k] + dth_bubble + noise;
                double q = qv0[k];
                if (dth_bubble > 0.0)
                {
                    const double s_old = qvs(th0[k] * pi_k, p0[k]);
                    const double s_new = qvs((th0[k] + dth_bubble) * pi_k, p0[k]);
                    if (s_old > 1.0e-12) q = std::min(q * (s_new / s_old), prm.sounding.qv_cap);
                }
                u[c] = S(u0[k]);
                v[c] = S(v0[k]);
                w[c] = S(0);
                p[c] = S(p0[k]);
                th[c] = S(theta);
                rho[c] = S(p0[k] / (kRd * theta * pi_k));
                qv[c] = S(q);
                qc[c] = S(0);
                qr[c] = S(0);
            }
    if (out.rain_mm) std::fill(out.rain_mm, out.rain_mm + std::size_t(g.nx) * g.ny, 0.0);
    return NIMBUS_OK;
}

template <class S>
int fill_state(const nimbus_grid& g, const nimbus_ic_params& prm, const nimbus_host_state& out)
{
    if (g.nx < 5 || g.ny < 5 || g.nz < 4) return NIMBUS_EINVAL;
    std::vector<double> rho0(g.nz), p0(g.nz), th0(g.nz), qv0(g.nz), u0(g.nz), v0(g.nz);
    const int rc = base_state(g, prm.sounding, rho0.data(), p0.data(), th0.data(), qv0.data(), u0.data(), v0.data());
    if (rc != NIMBUS_OK) return rc;
    const nimbus_base_state b{rho0.data(), p0.data(), th0.data(), qv0.data(), u0.data(), v0.data()};
    return fill_from_base<S>(g, b, prm, out);
}

template int fill_state<float>(const nimbus_grid&, const nimbus_ic_params&, const nimbus_host_state&);
template int fill_state<double>(const nimbus_grid&, const nimbus_ic_params&, const nimbus_host_state&);
template int fill_from_base<float>(const nimbus_grid&, const nimbus_base_state&, const nimbus_ic_params&,
                                   const nimbus_host_state&);
template int fill_from_base<double>(const nimbus_grid&, const nimbus_base_state&, const nimbus_ic_params&,
                                    const nimbus_host_state&);

// The analytic sounding at heights z (z[0] = 0) in CM1 column form: the
// same midpoint integration of d(pi)/dz = -g / (cp theta_v) as base_state,
// carried from one requested height to the next.
int sounding_levels(const nimbus_sounding& s, int n, const double* z, nimbus_cm1_sounding& out)
{
    if (n < 1 || !z || !(z[0] == 0.0)) return NIMBUS_EINVAL;
    if (out.capacity < n || !out.z || !out.theta || !out.qv || !out.u || !out.v) return NIMBUS_EINVAL;
    for (int k = 1; k < n; ++k)
        if (!(z[k] > z[k - 1]) || !std::isfinite(z[k])) return NIMBUS_EINVAL;
    const double kappa = kRd / kCp;
    auto thv_at = [&](double zq, double piq) {
        const double th = wk_theta(s, zq);
        const double p = kP00 * std::pow(piq, 1.0 / kappa);
        return th * (1.0 + 0.608 * sounding_qv(s, zq, th * piq, p));
    };
    double zc = 0.0;
    double pi = std::pow(s.p_sfc / kP00, kappa);
    for (int k = 0; k < n; ++k)
    {
        const int nsub = std::max(1, int(std::ceil((z[k] - zc) / 10.0)));
        const double h = (z[k] - zc) / nsub;
        for (int q = 0; q < nsub && z[k] > zc; ++q)
        {
            const double half = pi - 0.5 * h * kG / (kCp * thv_at(zc, pi));
            pi = pi - h * kG / (kCp * thv_at(zc + 0.5 * h, half));
            zc += h;
        }
        zc = z[k];
        const double th = wk_theta(s, z[k]);
        const double p = kP00 * std::pow(pi, 1.0 / kappa);
        const double T = th * pi;
        out.z[k] = z[k];
        out.theta[k] = th;
        out.qv[k] = sounding_qv(s, z[k], T, p);
        hodograph(s, z[k], out.u[k], out.v[k]);
        if (out.p) out.p[k] = p;
        if (out.T) out.T[k] = T;
    }
    out.n = n;
    out.p_sfc = s.p_sfc;
    out.theta_sfc = out.theta[0];
    out.qv_sfc = out.qv[0];
    return NIMBUS_OK;
}

int resolve_substeps(const nimbus_grid& g, const double* rho0, const double* p0, const nimbus_step_config& c)
{
    double cmax = 0.0;
    for (int k = 0; k < g.nz; ++k)
        if (rho0[k] > 0.0 && p0[k] > 0.0) cmax = std::max(cmax, std::sqrt(kGamma * p0[k] / rho0[k]));
    const double dmin = c.acoustic_mode == NIMBUS_ACOUSTIC_HEVI ? std::min(g.dx, g.dy)
                                                                : std::min({g.dx, g.dy, g.dz});
    const double need = c.dt * cmax / (0.5 * dmin);
    if (!std::isfinite(need) || need > 1.0e6) return -1;
    const int required = std::max(int(std::ceil(need)) + 1, 2);
    return c.substeps > 0 ? std::max(c.substeps, required) : required;
}

}  // namespace nimbus

extern "C" {

void nimbus_default_sounding(nimbus_sounding* s)
{
    s->theta_sfc = 300.0;
    s->theta_tropo = 343.0;
    s->z_tropo = 12000.0;
    s->T_tropo = 213.0;
    s->rh_sfc = 0.95;
    s->rh_top = 0.25;
    s->qv_cap = 0.014;
    s->p_sfc = 100000.0;
    s->shear = 30.0;
    s->shear_depth = 6000.0;
    s->curve_depth = 4000.0;
    s->curvature = 1.0;
}

int nimbus_ic_base_state(const nimbus_grid* g, const nimbus_sounding* s, double* rho0, double* p0, double* theta0,
                         double* qv0, double* u0, double* v0)
{
    if (!g || !s || !rho0 || !p0 || !theta0 || !qv0 || !u0 || !v0) return NIMBUS_EINVAL;
    return nimbus::base_state(*g, *s, rho0, p0, theta0, qv0, u0, v0);
}

int nimbus_ic_fill_f32(const nimbus_grid* g, const nimbus_ic_params* prm, const nimbus_host_state* out)
{
    if (!g || !prm || !out) return NIMBUS_EINVAL;
    return nimbus::fill_state<float>(*g, *prm, *out);
}

int nimbus_ic_fill_f64(const nimbus_grid* g, const nimbus_ic_params* prm, const nimbus_host_state* out)
{
    if (!g || !prm || !out) return NIMBUS_EINVAL;
    return nimbus::fill_state<double>(*g, *prm, *out);
}

int nimbus_ic_fill_base_f32(const nimbus_grid* g, const nimbus_base_state* base, const nimbus_ic_params* prm,
                            const nimbus_host_state* out)
{
    if (!g || !base || !prm || !out) return NIMBUS_EINVAL;
    return nimbus::fill_from_base<float>(*g, *base, *prm, *out);
}

int nimbus_ic_fill_base_f64(const nimbus_grid* g, const nimbus_base_state* base, const nimbus_ic_params* prm,
                            const nimbus_host_state* out)
{
    if (!g || !base || !prm || !out) return NIMBUS_EINVAL;
    return nimbus::fill_from_base<double>(*g, *base, *prm, *out);
}

int nimbus_ic_sounding_levels(const nimbus_sounding* s, int n, const double* z, nimbus_cm1_sounding* out)
{
    if (!s || !out) return NIMBUS_EINVAL;
    return nimbus::sounding_levels(*s, n, z, *out);
}

double nimbus_ic_uniform(uint64_t seed, uint32_t member, int i, int j, int k)
{
    return nimbus::ic_uniform(seed, member, i, j, k);
}

int nimbus_resolve_substeps(const nimbus_grid* g, const nimbus_base_state* b, const nimbus_step_config* c)
{
    if (!g || !b || !c || !b->rho0 || !b->p0) return NIMBUS_EINVAL;
    return nimbus::resolve_substeps(*g, b->rho0, b->p0, *c);
}

}  // extern "C"
