// Entry points with the reference dispatch contracts as they are: the field
// lists, array layout ([nx][ny][nz], k contiguous, caller-owned host float
// arrays), index ranges and output semantics of the reference's device
// dispatches, so that an adapter implementing the reference's
// ComputeBackend families is a field copy per call. Each call stages its
// arrays to the device, runs kernels in that layout and copies the outputs
// back only after every kernel succeeded (a refused or failed call leaves the
// host arrays untouched, include/compute/dispatch/dynamics.hpp:682-683).
//
// Arithmetic is FP64 on FP32 storage, the reference CPU backend's precision,
// so these calls agree with the reference CPU path at its parity tiers.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "nimbus_internal.hpp"

namespace nimbus
{
namespace
{

constexpr int kThreads = 256;

#define REF_TRY(expr)                                                                          \
    do                                                                                         \
    {                                                                                          \
        const hipError_t e_ = (expr);                                                          \
        if (e_ != hipSuccess) return set_error(NIMBUS_EHIP, std::string(#expr) + ": " + hipGetErrorString(e_)); \
    } while (0)

// Device copies of the call's arrays, released when the call returns.
struct DeviceArrays
{
    std::vector<void*> ptrs;
    ~DeviceArrays()
    {
        for (void* p : ptrs) hipFree(p);
    }
    // Device array of n floats, a copy of host (or zero when host is null).
    int take(const float* host, std::size_t n, hipStream_t s, float** out)
    {
        void* p = nullptr;
        const hipError_t e = hipMalloc(&p, std::max<std::size_t>(n, 1) * sizeof(float));
        if (e != hipSuccess) return set_error(NIMBUS_ENOMEM, std::string("dispatch staging: ") + hipGetErrorString(e));
        ptrs.push_back(p);
        const hipError_t c = host ? hipMemcpyAsync(p, host, n * sizeof(float), hipMemcpyHostToDevice, s)
                                  : hipMemsetAsync(p, 0, n * sizeof(float), s);
        if (c != hipSuccess) return set_error(NIMBUS_EHIP, std::string("dispatch staging: ") + hipGetErrorString(c));
        *out = static_cast<float*>(p);
        return NIMBUS_OK;
    }
};

int back(float* host, const float* dev, std::size_t n, hipStream_t s)
{
    REF_TRY(hipMemcpyAsync(host, dev, n * sizeof(float), hipMemcpyDeviceToHost, s));
    return NIMBUS_OK;
}

// Four committed fields: copied to host staging, shown to the observer (if
// any) while the caller's arrays still hold the old state, then published.
int commit_four(float* const* host, float* const* dev, std::size_t n, hipStream_t s,
                nimbus_four_field_observer observer, void* context)
{
    std::vector<float> stage(4 * n);
    for (int q = 0; q < 4; ++q)
        REF_TRY(hipMemcpyAsync(stage.data() + q * n, dev[q], n * sizeof(float), hipMemcpyDeviceToHost, s));
    REF_TRY(hipStreamSynchronize(s));
    if (observer) observer(context, stage.data(), stage.data() + n, stage.data() + 2 * n, stage.data() + 3 * n);
    for (int q = 0; q < 4; ++q) std::memcpy(host[q], stage.data() + q * n, n * sizeof(float));
    return NIMBUS_OK;
}

unsigned blocks_for(std::size_t n) { return unsigned((n + kThreads - 1) / kThreads); }

__device__ __forceinline__ double dfinite0(double x) { return isfinite(x) ? x : 0.0; }

// ---------------------------------------------------------------------
// Slow tendencies (CartesianCgridSlowTendenciesDispatch,
// include/compute/dispatch/dynamics.hpp:374-420; the operator of
// cartesian_cgrid_3d.cpp:1038-1399 on flat terrain).
// ---------------------------------------------------------------------
struct SlowView
{
    const float *u, *v, *w, *rho, *p, *rho0, *qv0, *u0, *v0, *qv, *load;
    float *du, *dv, *dw, *drho, *dp, *adu, *adv, *adw;
    int nx, ny, nz, wind_len;
    double idx, idy, idz, g, f, frame_u, frame_v;
    int moisture, loading;
};

__device__ __forceinline__ std::size_t at(const SlowView& s, int i, int j, int k)
{
    return (std::size_t(i) * s.ny + j) * s.nz + k;
}

// Fifth-order upwind-biased derivative with the reduced orders near the
// axis ends (derivative_operators.hpp:324-401); `stride` steps along the axis.
__device__ double upwind(const float* f, std::size_t c, std::ptrdiff_t stride, int pos, int n, double a, double inv)
{
    const int half = min(pos, n - 1 - pos);
    if (half <= 0) return 0.0;
    auto F = [&](int o) { return double(f[std::ptrdiff_t(c) + o * stride]); };
    const double m1 = F(-1), c0 = F(0), p1 = F(1);
    if (half == 1) return a >= 0.0 ? (c0 - m1) * inv : (p1 - c0) * inv;
    const double m2 = F(-2), p2 = F(2);
    if (half == 2)
        return a >= 0.0 ? (2.0 * p1 + 3.0 * c0 - 6.0 * m1 + m2) * (inv / 6.0)
                        : (-p2 + 6.0 * p1 - 3.0 * c0 - 2.0 * m1) * (inv / 6.0);
    const double m3 = F(-3), p3 = F(3);
    return a >= 0.0 ? (-2.0 * m3 + 15.0 * m2 - 60.0 * m1 + 20.0 * c0 + 30.0 * p1 - 3.0 * p2) * (inv / 60.0)
                    : (2.0 * p3 - 15.0 * p2 + 60.0 * p1 - 20.0 * c0 - 30.0 * m1 + 3.0 * m2) * (inv / 60.0);
}

__global__ __launch_bounds__(kThreads) void k_ref_slow(SlowView s)
{
    const std::size_t n = std::size_t(s.nx) * s.ny * s.nz;
    const std::size_t c = std::size_t(blockIdx.x) * kThreads + threadIdx.x;
    if (c >= n) return;
    const int k = int(c % s.nz), j = int((c / s.nz) % s.ny), i = int(c / (std::size_t(s.nz) * s.ny));
    const std::ptrdiff_t sx = std::ptrdiff_t(s.ny) * s.nz, sy = s.nz, sz = 1;
    double du = 0.0, dv = 0.0, dw = 0.0, dp = 0.0, au = 0.0, av = 0.0, aw = 0.0;
    const bool inner = i >= 1 && i < s.nx - 1 && j >= 1 && j < s.ny - 1;
    if (inner && k < s.nz - 1)
    {
        const float *u = s.u, *v = s.v, *w = s.w;
        auto U = [&](int a, int b, int d) { return double(u[at(s, a, b, d)]); };
        auto V = [&](int a, int b, int d) { return double(v[at(s, a, b, d)]); };
        auto W = [&](int a, int b, int d) { return double(w[at(s, a, b, d)]); };
        const double v0k = k < s.wind_len ? double(s.v0[k]) : 0.0;
        const double u0k = k < s.wind_len ? double(s.u0[k]) : 0.0;
        {  // u at the x-face (cartesian_cgrid_3d.cpp:1064-1165)
            const double ux = U(i, j, k);
            const double vy = 0.25 * (V(i, j - 1, k) + V(i, j, k) + V(i + 1, j - 1, k) + V(i + 1, j, k));
            const double tux = ux - s.frame_u, tvy = vy - s.frame_v;
            const double wz = k == 0 ? 0.25 * (W(i, j, 0) + W(i + 1, j, 0))
                                     : 0.25 * (W(i, j, k - 1) + W(i, j, k) + W(i + 1, j, k - 1) + W(i + 1, j, k));
            const double ddx = upwind(u, c, sx, i, s.nx, tux, s.idx);
            const double ddy = upwind(u, c, sy, j, s.ny, tvy, s.idy);
            const double ddz = k == 0 ? (wz >= 0.0 ? 0.0 : (U(i, j, 1) - U(i, j, 0)) * s.idz)
                                      : upwind(u, c, sz, k, s.nz, wz, s.idz);
            au = -tux * ddx - tvy * ddy - wz * ddz;
            du = au + s.f * (vy - v0k);
        }
        {  // v at the y-face (:1168-1258)
            const double vy = V(i, j, k);
            const double ux = 0.25 * (U(i - 1, j, k) + U(i, j, k) + U(i - 1, j + 1, k) + U(i, j + 1, k));
            const double tux = ux - s.frame_u, tvy = vy - s.frame_v;
            const double wz = k == 0 ? 0.25 * (W(i, j, 0) + W(i, j + 1, 0))
                                     : 0.25 * (W(i, j, k - 1) + W(i, j, k) + W(i, j + 1, k - 1) + W(i, j + 1, k));
            const double ddx = upwind(v, c, sx, i, s.nx, tux, s.idx);
            const double ddy = upwind(v, c, sy, j, s.ny, tvy, s.idy);
            const double ddz = k == 0 ? (wz >= 0.0 ? 0.0 : (V(i, j, 1) - V(i, j, 0)) * s.idz)
                                      : upwind(v, c, sz, k, s.nz, wz, s.idz);
            av = -tux * ddx - tvy * ddy - wz * ddz;
            dv = av - s.f * (ux - u0k);
        }
        {  // w at the z-face with buoyancy, moisture buoyancy and loading (:1260-1356)
            const double wz = W(i, j, k);
            const double ux = 0.25 * (U(i - 1, j, k) + U(i, j, k) + U(i - 1, j, k + 1) + U(i, j, k + 1));
            const double vy = 0.25 * (V(i, j - 1, k) + V(i, j, k) + V(i, j - 1, k + 1) + V(i, j, k + 1));
            const double tux = ux - s.frame_u, tvy = vy - s.frame_v;
            const double ddx = upwind(w, c, sx, i, s.nx, tux, s.idx);
            const double ddy = upwind(w, c, sy, j, s.ny, tvy, s.idy);
            const double ddz = k == 0 ? (wz >= 0.0 ? W(i, j, 0) * s.idz : (W(i, j, 1) - W(i, j, 0)) * s.idz)
                                      : upwind(w, c, sz, k, s.nz, wz, s.idz);
            aw = -tux * ddx - tvy * ddy - wz * ddz;
            const double rf = 0.5 * (double(s.rho[c]) + double(s.rho[c + 1]));
            const double r0f = 0.5 * (double(s.rho0[k]) + double(s.rho0[k + 1]));
            const double rsafe = isfinite(rf) && rf > 1.0e-6 ? rf : 1.0;
            double b = -s.g * (rf - r0f) / rsafe;
            if (s.moisture)
                b += s.g * 0.608 * (0.5 * (double(s.qv[c]) + double(s.qv[c + 1])) -
                                    0.5 * (double(s.qv0[k]) + double(s.qv0[k + 1])));
            if (s.loading) b += -s.g * 0.5 * (double(s.load[c]) + double(s.load[c + 1]));
            dw = aw + b;
        }
    }
    if (inner && k >= 1 && k < s.nz - 1)
    {  // dp/dt = -u.grad p, centred (:1358-1399)
        const double uc = 0.5 * (double(s.u[c]) + double(s.u[c - sx])) - s.frame_u;
        const double vc = 0.5 * (double(s.v[c]) + double(s.v[c - sy])) - s.frame_v;
        const double wc = 0.5 * (double(s.w[c]) + double(s.w[c - 1]));
        const double px = (double(s.p[c + sx]) - double(s.p[c - sx])) * (0.5 * s.idx);
        const double py = (double(s.p[c + sy]) - double(s.p[c - sy])) * (0.5 * s.idy);
        const double pz = (double(s.p[c + 1]) - double(s.p[c - 1])) * (0.5 * s.idz);
        dp = -uc * px - vc * py - wc * pz;
    }
    s.du[c] = float(dfinite0(du));
    s.dv[c] = float(dfinite0(dv));
    s.dw[c] = float(dfinite0(dw));
    s.dp[c] = float(dfinite0(dp));
    s.drho[c] = 0.0f;
    if (s.adu)
    {
        s.adu[c] = float(dfinite0(au));
        s.adv[c] = float(dfinite0(av));
        s.adw[c] = float(dfinite0(aw));
    }
}

// Slow apply (SlowApplyDispatch, dynamics.hpp:328-351; the shader contract of
// slow_apply.comp): x += dt * (dx + pbl), a non-finite sum counting as 0 when
// sanitizing.
__global__ __launch_bounds__(kThreads) void k_ref_slow_apply(std::size_t n, float* u, float* v, float* w, float* p,
                                                             const float* du, const float* dv, const float* dw,
                                                             const float* dp, const float* dup, const float* dvp,
                                                             double dt, int sanitize)
{
    const std::size_t c = std::size_t(blockIdx.x) * kThreads + threadIdx.x;
    if (c >= n) return;
    auto z = [&](double x) { return sanitize && !isfinite(x) ? 0.0 : x; };
    const double a = z(double(du[c]) + double(dup[c])), b = z(double(dv[c]) + double(dvp[c]));
    const double d = z(double(dw[c])), e = z(double(dp[c]));
    u[c] = float(double(u[c]) + a * dt);
    v[c] = float(double(v[c]) + b * dt);
    w[c] = float(double(w[c]) + d * dt);
    p[c] = float(double(p[c]) + e * dt);
}

// Rayleigh state relaxation (StateRelaxationDispatch, dynamics.hpp:354-372;
// state_relaxation_cartesian.comp): upper sponge, w-only absorber and lateral
// band in that order, each a relaxation by 1 - exp(-alpha dt). Profile rows
// (nz each): upper alpha, w-only alpha, u_ref, v_ref, theta0, lateral alpha
// by distance to the nearest lateral face, upper alpha at w faces.
__global__ __launch_bounds__(kThreads) void k_ref_relax(int nx, int ny, int nz, float* u, float* v, float* w,
                                                        float* th, const float* prof, int width, int up, int wonly,
                                                        int lat, double dt)
{
    const std::size_t n = std::size_t(nx) * ny * nz;
    const std::size_t c = std::size_t(blockIdx.x) * kThreads + threadIdx.x;
    if (c >= n) return;
    const int k = int(c % nz), j = int((c / nz) % ny), i = int(c / (std::size_t(nz) * ny));
    const double h = fmax(dt, 1.0e-12);
    const double uref = prof[2 * nz + k], vref = prof[3 * nz + k], th0 = prof[4 * nz + k];
    double uu = u[c], vv = v[c], ww = w[c], tt = th[c];
    if (up)
    {
        const double a = prof[k], aw = prof[6 * nz + k];
        if (a > 0.0 || aw > 0.0)
        {
            const double fa = -expm1(-a * h), fw = -expm1(-aw * h);
            uu += fa * (uref - uu);
            vv += fa * (vref - vv);
            ww += fw * (0.0 - ww);
            tt += fa * (th0 - tt);
        }
    }
    if (wonly)
    {
        const double aw = prof[nz + k];
        if (aw > 0.0) ww += -expm1(-aw * h) * (0.0 - ww);
    }
    if (lat && width > 0)
    {
        const int d = min(min(i, nx - 1 - i), min(j, ny - 1 - j));
        if (d < width)
        {
            const double fa = -expm1(-double(prof[5 * nz + d]) * h);
            uu += fa * (uref - uu);
            vv += fa * (vref - vv);
            ww += fa * (0.0 - ww);
            tt += fa * (th0 - tt);
        }
    }
    u[c] = float(uu);
    v[c] = float(vv);
    w[c] = float(ww);
    th[c] = float(tt);
}

// Microphysics apply (MicrophysicsApplyDispatch, microphysics.hpp:326-363;
// microphysics_apply.comp): bounded theta step, clamped species, counters in
// stats[0..9] and per-level / per-i clamp histograms from stats[10].
struct ApplyView
{
    float* q[8];        // theta, qv, qc, qr, qi, qs, qg, qh
    const float* d[8];  // their tendencies
    const float *drad, *dpbl, *dqvpbl;
    unsigned* stats;
    int stats_len;
    int nx, ny, nz, i0, i1, j0, j1;
    double dt, th_fb, th_min, th_max, max_step, qv_max, hydro_max;
};

__device__ void bump(const ApplyView& a, int slot, unsigned n)
{
    if (a.stats && slot < a.stats_len && n) atomicAdd(a.stats + slot, n);
}

__global__ __launch_bounds__(kThreads) void k_ref_micro_apply(ApplyView a)
{
    const std::size_t n = std::size_t(a.nx) * a.ny * a.nz;
    const std::size_t c = std::size_t(blockIdx.x) * kThreads + threadIdx.x;
    if (c >= n) return;
    const int k = int(c % a.nz), j = int((c / a.nz) % a.ny), i = int(c / (std::size_t(a.nz) * a.ny));
    if (i < a.i0 || i >= a.i1 || j < a.j0 || j >= a.j1) return;
    unsigned bad = 0;
    double t[8];
    for (int q = 0; q < 8; ++q)
    {
        const double x = double(a.d[q][c]);
        t[q] = isfinite(x) ? x : 0.0;
        bad += isfinite(x) ? 0u : 1u;
    }
    double total = (t[0] + double(a.drad[c]) + double(a.dpbl[c])) * a.dt;
    const bool nonfinite_total = !isfinite(total);
    if (nonfinite_total) total = 0.0;
    const double lim = fmin(fmax(total, -a.max_step), a.max_step);
    const double th = double(a.q[0][c]);
    const double thn = th + lim;
    const double safe = isfinite(thn) ? thn : (isfinite(th) ? th : a.th_fb);
    const double bounded = fmin(fmax(safe, a.th_min), a.th_max);
    const bool clamped = bounded != thn || !isfinite(thn);
    a.q[0][c] = float(bounded);
    for (int q = 1; q < 8; ++q)
    {
        const double add = q == 1 ? t[1] + double(a.dqvpbl[c]) : t[q];
        const double x = double(a.q[q][c]) + add * a.dt;
        const double top = q == 1 ? a.qv_max : a.hydro_max;
        a.q[q][c] = float(isfinite(x) ? fmin(fmax(x, 0.0), top) : 0.0);
    }
    bump(a, 0, bad);
    bump(a, 1, nonfinite_total ? 1u : 0u);
    bump(a, 2, lim != total ? 1u : 0u);
    if (clamped)
    {
        bump(a, 3, 1u);
        bump(a, 10 + k, 1u);
        bump(a, 10 + a.nz + i, 1u);
        if (a.stats && a.stats_len >= 10 && atomicExch(a.stats + 4, 1u) == 0u)
        {
            const float thn_f = float(thn), b_f = float(bounded), tot_f = float(total), lim_f = float(lim);
            a.stats[5] = unsigned(c);
            std::memcpy(a.stats + 6, &thn_f, 4);
            std::memcpy(a.stats + 7, &b_f, 4);
            std::memcpy(a.stats + 8, &tot_f, 4);
            std::memcpy(a.stats + 9, &lim_f, 4);
        }
    }
}

// Kessler warm rain on adjusted inputs (KesslerPointwiseDispatch,
// microphysics.hpp:436-467; kessler.cpp:346-460 and the pointwise shader
// contract of kessler.cpp:213-230): autoconversion, accretion and ventilated
// evaporation written as tendencies (temperature tendency in dtheta_dt);
// zero outside the ranges and for the graupel/hail slots.
struct KesslerPtView
{
    const float *T, *p, *qv, *qc, *qr, *rho;
    float *dth, *dqv, *dqc, *dqr, *dqg, *dqh;
    int nx, ny, nz, i0, i1, j0, j1, k0, k1;
    double qc0, c_auto, c_accr, evap, lv_cp, T0, dt;
};

// Buck saturation mixing ratio over water (thermodynamics.hpp:67-118).
__device__ double qvs_ref(double T, double p, double T0)
{
    if (!isfinite(T) || !isfinite(p)) return 0.0;
    const double Tc = fmin(fmax(T, 150.0), 400.0) - T0;
    double es = 611.21 * exp((18.678 - Tc / 234.5) * Tc / (257.14 + Tc));
    es = isfinite(es) ? fmax(0.0, es) : 0.0;
    const double pp = fmin(fmax(p, 100.0), 200000.0);
    return fmin(fmax(0.62198 * es / fmax(pp - es, 1.0), 0.0), 0.2);
}

__global__ __launch_bounds__(kThreads) void k_ref_kessler_pointwise(KesslerPtView v)
{
    const std::size_t n = std::size_t(v.nx) * v.ny * v.nz;
    const std::size_t c = std::size_t(blockIdx.x) * kThreads + threadIdx.x;
    if (c >= n) return;
    const int k = int(c % v.nz), j = int((c / v.nz) % v.ny), i = int(c / (std::size_t(v.nz) * v.ny));
    double dth = 0.0, dqv = 0.0, dqc = 0.0, dqr = 0.0;
    if (i >= v.i0 && i < v.i1 && j >= v.j0 && j < v.j1 && k >= v.k0 && k < v.k1)
    {
        const double T = v.T[c], p = v.p[c], qv = v.qv[c], qc = v.qc[c], qr = v.qr[c], rho = v.rho[c];
        double prod = 0.0;
        if (v.dt > 0.0 && qc > 0.0)
        {
            const double factor = 1.0 / (1.0 + v.c_accr * v.dt * pow(fmax(qr, 0.0), 0.875));
            const double amount = qc * (1.0 - factor) + factor * v.c_auto * v.dt * fmax(qc - v.qc0, 0.0);
            prod = fmin(fmax(amount, 0.0), qc) / v.dt;
            dqc -= prod;
            dqr += prod;
        }
        if (qr > 0.0 || prod > 0.0)
        {
            const double qs = qvs_ref(T, p, v.T0);
            if (qs > 0.0 && qv < qs && rho > 0.0 && p > 0.0)
            {
                const double avail = qr + prod * v.dt;
                const double rcgs = 1.0e-3 * rho;
                const double rc = rcgs * fmax(avail, 0.0);
                const double vent = (1.6 + 124.9 * pow(rc, 0.2046)) * pow(rc, 0.525);
                const double resist = 2.55e8 / (p * qs) + 5.4e5;
                const double raw = v.evap * (vent / resist) * ((qs - qv) / (rcgs * qs));
                const double gsat = v.lv_cp * (kLv / kRv) * qs / fmax(T * T, 1.0);
                const double bound = (qs - qv) / (1.0 + gsat);
                const double ev = fmin(fmax(raw, 0.0), fmin(bound / v.dt, avail / v.dt));
                dqv += ev;
                dqr -= ev;
                dth -= v.lv_cp * ev;
            }
        }
    }
    v.dth[c] = float(dth);
    v.dqv[c] = float(dqv);
    v.dqc[c] = float(dqc);
    v.dqr[c] = float(dqr);
    if (v.dqg) v.dqg[c] = 0.0f;
    if (v.dqh) v.dqh[c] = 0.0f;
}

// Kessler rain sedimentation (KesslerSedimentationDispatch,
// microphysics.hpp:469-492; kessler.cpp:465-531 and kessler_sedimentation
// contract): per column, in k order, the donor loses Vt q / dz_k and the
// level below gains (rho_k / rho_{k-1}) Vt q / dz_{k-1}, added to dqr_dt.
struct KesslerSedView
{
    const float *qr, *rho, *layer;
    float* dqr;
    int nx, ny, nz, i0, i1, j0, j1, k0, k1;
    double dz, vt_max, vt_cap;
};

__global__ __launch_bounds__(kThreads) void k_ref_kessler_sed(KesslerSedView v)
{
    const std::size_t cols = std::size_t(v.nx) * v.ny;
    const std::size_t col = std::size_t(blockIdx.x) * kThreads + threadIdx.x;
    if (col >= cols) return;
    const int j = int(col % v.ny), i = int(col / v.ny);
    if (i < v.i0 || i >= v.i1 || j < v.j0 || j >= v.j1) return;
    const std::size_t base = col * v.nz;
    for (int k = v.k0; k < v.k1; ++k)
    {
        const std::size_t c = base + k;
        const double ddz = v.layer ? double(v.layer[c]) : v.dz;
        const double rdz = v.layer ? (k > v.k0 ? double(v.layer[c - 1]) : ddz) : v.dz;
        double fall = 1.0, recv = 1.0, rho = 1.2;
        if (v.rho)
        {
            const double rk = v.rho[c], rkm = k > v.k0 ? double(v.rho[c - 1]) : rk;
            fall = sqrt(1.2 / rk);
            recv = rk / rkm;
            rho = rk;
        }
        const double q = fmax(double(v.qr[c]), 0.0);
        if (q > 1.0e-9)
        {
            const double vt = fmin(fmin(36.34 * pow(1.0e-3 * rho * q, 0.1346) * fall, v.vt_max), v.vt_cap);
            const double flux = vt * q;
            v.dqr[c] = float(double(v.dqr[c]) - flux / ddz);
            if (k > v.k0) v.dqr[c - 1] = float(double(v.dqr[c - 1]) + recv * flux / rdz);
        }
    }
}

// Conservative WENO5 transport (ConservativeCartesianCgridTransportDispatch,
// advection.hpp:46-103; advection_cartesian.cpp:1941-2650): face fluxes F =
// mf * face value (upwind WENO5 where its cells exist, else the MC-limited
// upwind slope; the support-preserving face for positive scalars when asked),
// the positive-definite donor limiter, and (rho_b q - dt div F) / max(rho_e,
// floor) with outgoing-then-incoming order for positive scalars.
struct TransportView
{
    const float *mfx, *mfy, *mfz, *rb, *re, *q;
    float *fx, *fy, *fz, *rl, *out;
    int nx, ny, nz, pd, support, open_top;
    double idx, idy, idz, dt, floor;
};

__device__ double weno5(double a, double b, double c, double d, double e)
{
    const double q0 = (1.0 / 3.0) * a - (7.0 / 6.0) * b + (11.0 / 6.0) * c;
    const double q1 = -(1.0 / 6.0) * b + (5.0 / 6.0) * c + (1.0 / 3.0) * d;
    const double q2 = (1.0 / 3.0) * c + (5.0 / 6.0) * d - (1.0 / 6.0) * e;
    const double t0a = a - 2.0 * b + c, t0b = a - 4.0 * b + 3.0 * c;
    const double t1a = b - 2.0 * c + d, t1b = b - d;
    const double t2a = c - 2.0 * d + e, t2b = 3.0 * c - 4.0 * d + e;
    const double s0 = 13.0 / 12.0 * t0a * t0a + 0.25 * t0b * t0b;
    const double s1 = 13.0 / 12.0 * t1a * t1a + 0.25 * t1b * t1b;
    const double s2 = 13.0 / 12.0 * t2a * t2a + 0.25 * t2b * t2b;
    const double m2 = fmax(fmax(fmax(a * a, b * b), fmax(c * c, d * d)), e * e);
    const double eps = 1.0e-6 * m2 + 1.0e-40;
    const double w0 = 0.1 / ((eps + s0) * (eps + s0));
    const double w1 = 0.6 / ((eps + s1) * (eps + s1));
    const double w2 = 0.3 / ((eps + s2) * (eps + s2));
    const double ws = w0 + w1 + w2;
    if (!(ws > 0.0) || !isfinite(ws)) return c;
    const double f = (w0 * q0 + w1 * q1 + w2 * q2) / ws;
    return isfinite(f) ? f : c;
}

__device__ double mc(double a, double b, double c)
{
    const double r = c - b, l = b - a, x = 0.5 * (r + l), y = 2.0 * r, z = 2.0 * l;
    if (x > 0.0 && y > 0.0 && z > 0.0) return fmin(x, fmin(y, z));
    if (x < 0.0 && y < 0.0 && z < 0.0) return fmax(x, fmax(y, z));
    return 0.0;
}

// Face value between cells pos and pos+1 of an axis of n cells.
__device__ double face_value(const float* q, std::size_t c, std::ptrdiff_t s, int pos, int n, double m, int pd,
                             int support)
{
    auto Q = [&](int o) { return double(q[std::ptrdiff_t(c) + o * s]); };
    double f, donor;
    if (m >= 0.0)
    {
        donor = Q(0);
        if (pos >= 2 && pos + 2 <= n - 1) f = weno5(Q(-2), Q(-1), donor, Q(1), Q(2));
        else f = donor + 0.5 * (pos >= 1 ? mc(Q(-1), donor, Q(1)) : 0.0);
    }
    else
    {
        donor = Q(1);
        if (pos - 1 >= 0 && pos + 3 <= n - 1) f = weno5(Q(3), Q(2), donor, Q(0), Q(-1));
        else f = donor - 0.5 * (pos + 1 <= n - 2 ? mc(Q(0), donor, Q(2)) : 0.0);
    }
    if (pd && support) f = donor <= 0.0 ? 0.0 : (f <= 0.0 ? donor : f);
    return f;
}

__global__ __launch_bounds__(kThreads) void k_ref_faces(TransportView t)
{
    const std::size_t n = std::size_t(t.nx) * t.ny * t.nz;
    const std::size_t c = std::size_t(blockIdx.x) * kThreads + threadIdx.x;
    if (c >= n) return;
    const int k = int(c % t.nz), j = int((c / t.nz) % t.ny), i = int(c / (std::size_t(t.nz) * t.ny));
    const std::ptrdiff_t sx = std::ptrdiff_t(t.ny) * t.nz, sy = t.nz;
    double fx = 0.0, fy = 0.0, fz = 0.0;
    if (i < t.nx - 1)
    {
        const double m = t.mfx[c];
        fx = m * face_value(t.q, c, sx, i, t.nx, m, t.pd, t.support);
    }
    if (j < t.ny - 1)
    {
        const double m = t.mfy[c];
        fy = m * face_value(t.q, c, sy, j, t.ny, m, t.pd, t.support);
    }
    if (k < t.nz - 1)
    {
        const double m = t.mfz[c];
        fz = m * face_value(t.q, c, 1, k, t.nz, m, t.pd, t.support);
    }
    else if (t.open_top)
        fz = double(t.mfz[c]) * double(t.q[c]);
    t.fx[c] = float(fx);
    t.fy[c] = float(fy);
    t.fz[c] = float(fz);
}

__device__ float limiter_store(double r)
{
    if (!(r > 0.0 && r < 1.0)) return float(r);
    const float shrink = 1.0f - 32.0f * 1.1920928955078125e-7f;
    return nextafterf(float(r) * shrink, 0.0f);
}

__global__ __launch_bounds__(kThreads) void k_ref_limiter(TransportView t)
{
    const std::size_t n = std::size_t(t.nx) * t.ny * t.nz;
    const std::size_t c = std::size_t(blockIdx.x) * kThreads + threadIdx.x;
    if (c >= n) return;
    const int k = int(c % t.nz), j = int((c / t.nz) % t.ny), i = int(c / (std::size_t(t.nz) * t.ny));
    const std::ptrdiff_t sx = std::ptrdiff_t(t.ny) * t.nz, sy = t.nz;
    double out = 0.0;
    if (i <= t.nx - 2 && t.fx[c] > 0.0f) out += double(t.fx[c]) * t.idx;
    if (i >= 1 && t.fx[c - sx] < 0.0f) out -= double(t.fx[c - sx]) * t.idx;
    if (j <= t.ny - 2 && t.fy[c] > 0.0f) out += double(t.fy[c]) * t.idy;
    if (j >= 1 && t.fy[c - sy] < 0.0f) out -= double(t.fy[c - sy]) * t.idy;
    if ((k <= t.nz - 2 || (t.open_top && k == t.nz - 1)) && t.fz[c] > 0.0f) out += double(t.fz[c]) * t.idz;
    if (k >= 1 && t.fz[c - 1] < 0.0f) out -= double(t.fz[c - 1]) * t.idz;
    double r = 1.0;
    if (out > 0.0) r = fmin(1.0, double(t.rb[c]) * fmax(double(t.q[c]), 0.0) / t.dt / out);
    t.rl[c] = (r > 0.0 && r < 1.0) ? limiter_store(r) : float(r);
}

__global__ __launch_bounds__(kThreads) void k_ref_update(TransportView t)
{
    const std::size_t n = std::size_t(t.nx) * t.ny * t.nz;
    const std::size_t c = std::size_t(blockIdx.x) * kThreads + threadIdx.x;
    if (c >= n) return;
    const int k = int(c % t.nz), j = int((c / t.nz) % t.ny), i = int(c / (std::size_t(t.nz) * t.ny));
    if (!(i >= 1 && i < t.nx - 1 && j >= 1 && j < t.ny - 1))
    {
        t.out[c] = t.q[c];
        return;
    }
    const std::ptrdiff_t sx = std::ptrdiff_t(t.ny) * t.nz, sy = t.nz;
    auto lim = [&](float raw, std::size_t lo, std::size_t hi) -> double {
        if (!t.pd || raw == 0.0f) return double(raw);
        return double(raw) * double(raw > 0.0f ? t.rl[lo] : t.rl[hi]);
    };
    const bool top = k <= t.nz - 2 || (t.open_top && k == t.nz - 1);
    const double xp = lim(t.fx[c], c, c + sx), xm = lim(t.fx[c - sx], c - sx, c);
    const double yp = lim(t.fy[c], c, c + sy), ym = lim(t.fy[c - sy], c - sy, c);
    const double zp = top ? lim(t.fz[c], c, k < t.nz - 1 ? c + 1 : c) : 0.0;
    const double zm = k >= 1 ? lim(t.fz[c - 1], c - 1, c) : 0.0;
    const double rb = t.rb[c];
    double rq;
    if (t.pd)
    {
        double og = 0.0, ic = 0.0;
        auto plus = [&](double f, double inv) { if (f >= 0.0) og += f * inv; else ic -= f * inv; };
        auto minus = [&](double f, double inv) { if (f <= 0.0) og -= f * inv; else ic += f * inv; };
        plus(xp, t.idx);
        minus(xm, t.idx);
        plus(yp, t.idy);
        minus(ym, t.idy);
        if (top) plus(zp, t.idz);
        if (k >= 1) minus(zm, t.idz);
        rq = (rb * double(t.q[c]) - t.dt * og) + t.dt * ic;
    }
    else
        rq = rb * double(t.q[c]) - t.dt * ((xp - xm) * t.idx + (yp - ym) * t.idy + (zp - zm) * t.idz);
    const double re = t.re[c];
    t.out[c] = float(rq / (re > t.floor ? re : t.floor));
}

// Standalone mode: mass fluxes from the live winds and density
// (rho averaged to the face times the grid-relative velocity) and the end
// density rho - dt div(MF) (advect_conservative_mass_flux / update contracts).
__global__ __launch_bounds__(kThreads) void k_ref_mass_flux(int nx, int ny, int nz, const float* u, const float* v,
                                                            const float* w, const float* rho, double fu, double fv,
                                                            float* mfx, float* mfy, float* mfz)
{
    const std::size_t n = std::size_t(nx) * ny * nz;
    const std::size_t c = std::size_t(blockIdx.x) * kThreads + threadIdx.x;
    if (c >= n) return;
    const int k = int(c % nz), j = int((c / nz) % ny), i = int(c / (std::size_t(nz) * ny));
    const std::size_t sx = std::size_t(ny) * nz, sy = nz;
    mfx[c] = i < nx - 1 ? float(0.5 * (double(rho[c]) + double(rho[c + sx])) * (double(u[c]) - fu)) : 0.0f;
    mfy[c] = j < ny - 1 ? float(0.5 * (double(rho[c]) + double(rho[c + sy])) * (double(v[c]) - fv)) : 0.0f;
    mfz[c] = k < nz - 1 ? float(0.5 * (double(rho[c]) + double(rho[c + 1])) * double(w[c])) : 0.0f;
}

__global__ __launch_bounds__(kThreads) void k_ref_rho_check(int nx, int ny, int nz, const float* rho,
                                                            const float* mfx, const float* mfy, const float* mfz,
                                                            double idx, double idy, double idz, double dt, float* re)
{
    const std::size_t n = std::size_t(nx) * ny * nz;
    const std::size_t c = std::size_t(blockIdx.x) * kThreads + threadIdx.x;
    if (c >= n) return;
    const int k = int(c % nz), j = int((c / nz) % ny), i = int(c / (std::size_t(nz) * ny));
    const std::size_t sx = std::size_t(ny) * nz, sy = nz;
    const double dx = (double(mfx[c]) - (i >= 1 ? double(mfx[c - sx]) : 0.0)) * idx;
    const double dy = (double(mfy[c]) - (j >= 1 ? double(mfy[c - sy]) : 0.0)) * idy;
    const double dz = (double(mfz[c]) - (k >= 1 ? double(mfz[c - 1]) : 0.0)) * idz;
    re[c] = float(double(rho[c]) - dt * (dx + dy + dz));
}

bool shape_ok(int nx, int ny, int nz) { return nx >= 1 && ny >= 1 && nz >= 1 && double(nx) * ny * nz < 2.1e9; }

// Scratch of one transport call, in units of the cell count: face fluxes x,
// y, z and the limiter ratios, then (standalone mode) the mass fluxes x, y, z
// and the end density. Every element is written before it is read.
constexpr std::size_t kTransportScratchFields = 8;

int validate_transport(const nimbus_conservative_cartesian_cgrid_transport_dispatch& d)
{
    if (d.n_scalars < 1 || d.n_scalars > 16 || !d.q_in || !d.q_out || !d.dt_kappas || !d.rho_data ||
        !shape_ok(d.nx, d.ny, d.nz) || !(d.inv_dx > 0.0f) || !(d.inv_dy > 0.0f) || !(d.inv_dz > 0.0f) ||
        !(d.dt_step > 0.0f))
        return set_error(NIMBUS_EINVAL, "conservative transport: invalid dispatch");
    if (!d.weno) return set_error(NIMBUS_EINVAL, "conservative transport: only the WENO5 reconstruction is built");
    if (d.use_supplied_mass_flux && (!d.density_begin || !d.mass_flux_x || !d.mass_flux_y || !d.mass_flux_z))
        return set_error(NIMBUS_EINVAL, "conservative transport: supplied mass fluxes are incomplete");
    if (!d.use_supplied_mass_flux && (!d.u || !d.v || !d.w))
        return set_error(NIMBUS_EINVAL, "conservative transport: winds are needed without supplied mass fluxes");
    for (int q = 0; q < d.n_scalars; ++q)
    {
        if (!d.q_in[q] || !d.q_out[q]) return set_error(NIMBUS_EINVAL, "conservative transport: null scalar");
        if (!std::isfinite(d.dt_kappas[q]) || d.dt_kappas[q] != 0.0f)
            return set_error(NIMBUS_EINVAL, "conservative transport: chained point diffusion is not supported");
        if (d.preserve_positive_support && (!d.positive_definite || !d.positive_definite[q]))
            return set_error(NIMBUS_EINVAL, "conservative transport: support preservation needs positive scalars");
    }
    return NIMBUS_OK;
}

// Device arrays of one transport call (supplied mass fluxes or winds).
struct TransportArrays
{
    const float *rho, *rho_begin, *mfx, *mfy, *mfz, *u, *v, *w;
    const float* const* q_in;
    float* const* q_out;
};

// Enqueues the transport of every scalar on stream s; `scratch` holds
// kTransportScratchFields * cells floats.
void enqueue_transport(const nimbus_conservative_cartesian_cgrid_transport_dispatch& d, const TransportArrays& a,
                       float* scratch, hipStream_t s)
{
    const std::size_t n = std::size_t(d.nx) * d.ny * d.nz;
    float *fx = scratch, *fy = scratch + n, *fz = scratch + 2 * n, *rl = scratch + 3 * n;
    const float *mfx = a.mfx, *mfy = a.mfy, *mfz = a.mfz, *rb = a.rho_begin, *re = a.rho;
    if (!d.use_supplied_mass_flux)
    {
        // standalone: fluxes from the winds and the live (begin) density; the
        // end density carries the same fluxes
        float *sx = scratch + 4 * n, *sy = scratch + 5 * n, *sz = scratch + 6 * n, *se = scratch + 7 * n;
        hipLaunchKernelGGL(k_ref_mass_flux, dim3(blocks_for(n)), dim3(kThreads), 0, s, d.nx, d.ny, d.nz, a.u, a.v, a.w,
                           a.rho, double(d.frame_velocity_u_ms), double(d.frame_velocity_v_ms), sx, sy, sz);
        hipLaunchKernelGGL(k_ref_rho_check, dim3(blocks_for(n)), dim3(kThreads), 0, s, d.nx, d.ny, d.nz, a.rho, sx,
                           sy, sz, double(d.inv_dx), double(d.inv_dy), double(d.inv_dz), double(d.dt_step), se);
        mfx = sx; mfy = sy; mfz = sz; rb = a.rho; re = se;
    }
    for (int q = 0; q < d.n_scalars; ++q)
    {
        TransportView t{mfx, mfy, mfz, rb, re, a.q_in[q], fx, fy, fz, rl, a.q_out[q], d.nx, d.ny, d.nz,
                        d.positive_definite && d.positive_definite[q] ? 1 : 0, d.preserve_positive_support ? 1 : 0,
                        0, double(d.inv_dx), double(d.inv_dy), double(d.inv_dz), double(d.dt_step),
                        double(d.rho_floor)};
        hipLaunchKernelGGL(k_ref_faces, dim3(blocks_for(n)), dim3(kThreads), 0, s, t);
        if (t.pd) hipLaunchKernelGGL(k_ref_limiter, dim3(blocks_for(n)), dim3(kThreads), 0, s, t);
        hipLaunchKernelGGL(k_ref_update, dim3(blocks_for(n)), dim3(kThreads), 0, s, t);
    }
}

// Explicit Laplacian diffusion tendency of one field (diffusion batch
// contract, include/compute/dispatch/diffusion.hpp:29-69; equations of
// explicit.cpp:52-87,119-194,263-339,1183-1433). One thread per cell writes
// the cell's tendency: zero outside the field's control-volume domain,
// otherwise the vertical part (rounded to FP32), then the horizontal part
// added in FP64 and rounded again. Contraction is off so the FP64 operation
// sequence is the reference's.
struct DiffusionView
{
    const float *f, *rho, *dz, *kv;
    float* out;
    int nx, ny, nz, loc, c_grid, px, py;
    double dx, dy, kh, kvc;
};

struct DiffDomain
{
    int i0, i1, j0, j1, k0, k1;
};

__host__ __device__ inline DiffDomain diff_domain(int nx, int ny, int nz, int loc, bool c_grid, bool px, bool py)
{
    if (!c_grid || loc == 0) return {1, nx - 1, 1, ny - 1, c_grid ? 0 : 1, c_grid ? nz : nz - 1};
    if (loc == 1) return {1, px ? nx - 1 : nx - 2, 1, ny - 1, 0, nz - 1};
    if (loc == 2) return {1, nx - 1, 1, py ? ny - 1 : ny - 2, 0, nz - 1};
    return {1, nx - 1, 1, ny - 1, 0, nz - 1};
}

__global__ __launch_bounds__(kThreads) void k_ref_diffusion(DiffusionView v)
{
#pragma clang fp contract(off)
    const std::size_t n = std::size_t(v.nx) * v.ny * v.nz;
    const std::size_t c = std::size_t(blockIdx.x) * kThreads + threadIdx.x;
    if (c >= n) return;
    const int nx = v.nx, ny = v.ny, nz = v.nz, loc = v.loc;
    const bool cg = v.c_grid != 0, px = v.px != 0, py = v.py != 0;
    const int k = int(c % nz), j = int((c / nz) % ny), i = int(c / (std::size_t(nz) * ny));
    const DiffDomain d = diff_domain(nx, ny, nz, loc, cg, px, py);
    if (i < d.i0 || i >= d.i1 || j < d.j0 || j >= d.j1 || k < d.k0 || k >= d.k1)
    {
        v.out[c] = 0.0f;
        return;
    }
    auto at = [=](int a, int b, int e) { return (std::size_t(a) * ny + b) * nz + e; };
    auto rho = [&](int a, int b, int e) { return double(v.rho[at(a, b, e)]); };
    auto dzk = [&](int e) { return double(v.dz[e]); };
    auto fld = [&](int a, int b, int e) { return double(v.f[at(a, b, e)]); };
    const bool stag = cg && loc != 0;
    auto comp_rho = [&](int a, int b, int e) {
        if (!stag) return rho(a, b, e);
        if (loc == 1) return 0.5 * (rho(a, b, e) + rho(px && a == nx - 2 ? 1 : a + 1, b, e));
        if (loc == 2) return 0.5 * (rho(a, b, e) + rho(a, py && b == ny - 2 ? 1 : b + 1, e));
        return (rho(a, b, e) * dzk(e) + rho(a, b, e + 1) * dzk(e + 1)) / (dzk(e) + dzk(e + 1));
    };
    const double q = fld(i, j, k);
    float t = 0.0f;
    if (v.kvc > 0.0 || v.kv != nullptr)
    {
        auto kvat = [&](int a, int b, int e) { return v.kv ? double(v.kv[at(a, b, e)]) : v.kvc; };
        auto cell_tr = [&](int a, int b, int e) { return rho(a, b, e) * kvat(a, b, e); };
        auto comp_tr = [&](int a, int b, int e) {
            if (!cg) return cell_tr(a, b, e);
            if (loc == 1) return 0.5 * (cell_tr(a, b, e) + cell_tr(px && a == nx - 2 ? 1 : a + 1, b, e));
            if (loc == 2) return 0.5 * (cell_tr(a, b, e) + cell_tr(a, py && b == ny - 2 ? 1 : b + 1, e));
            if (loc == 3) return 0.5 * (cell_tr(a, b, e) + cell_tr(a, b, e + 1));
            return cell_tr(a, b, e);
        };
        const bool wc = cg && loc == 3;
        auto face_tr = [&](int lo) { return wc ? cell_tr(i, j, lo + 1) : 0.5 * (comp_tr(i, j, lo) + comp_tr(i, j, lo + 1)); };
        auto dist = [&](int lo) { return wc ? dzk(lo + 1) : 0.5 * (dzk(lo) + dzk(lo + 1)); };
        const double mass = wc ? 0.5 * (rho(i, j, k) * dzk(k) + rho(i, j, k + 1) * dzk(k + 1))
                               : (cg ? comp_rho(i, j, k) : rho(i, j, k)) * dzk(k);
        double down = 0.0, up = 0.0;
        if (k > d.k0) down = face_tr(k - 1) * (q - fld(i, j, k - 1)) / dist(k - 1);
        if (k < d.k1 - 1) up = face_tr(k) * (fld(i, j, k + 1) - q) / dist(k);
        t = float(double(t) + (up - down) / mass);
    }
    if (v.kh > 0.0)
    {
        const double idx2 = 1.0 / (v.dx * v.dx), idy2 = 1.0 / (v.dy * v.dy);
        const bool wcl = i == d.i0 && !px, ecl = i == d.i1 - 1 && !px;
        const bool scl = j == d.j0 && !py, ncl = j == d.j1 - 1 && !py;
        const int iw = i == d.i0 && px ? d.i1 - 1 : i - 1, ie = i == d.i1 - 1 && px ? d.i0 : i + 1;
        const int js = j == d.j0 && py ? d.j1 - 1 : j - 1, jn = j == d.j1 - 1 && py ? d.j0 : j + 1;
        const double rc = comp_rho(i, j, k);
        const double rw = wcl ? 0.0 : 0.5 * (rc + comp_rho(iw, j, k));
        const double re = ecl ? 0.0 : 0.5 * (rc + comp_rho(ie, j, k));
        const double rs = scl ? 0.0 : 0.5 * (rc + comp_rho(i, js, k));
        const double rn = ncl ? 0.0 : 0.5 * (rc + comp_rho(i, jn, k));
        const double xd = ((ecl ? 0.0 : re * (fld(ie, j, k) - q)) - (wcl ? 0.0 : rw * (q - fld(iw, j, k)))) * idx2;
        const double yd = ((ncl ? 0.0 : rn * (fld(i, jn, k) - q)) - (scl ? 0.0 : rs * (q - fld(i, js, k)))) * idy2;
        t = float(double(t) + v.kh * (xd + yd) / rc);
    }
    v.out[c] = t;
}

constexpr int kMaxDiffusionTargets = 16;

// Host-side refusals of a diffusion batch (the reference validates the same
// inputs before its device dispatch, explicit.cpp:89-116,859-1181).
int validate_diffusion(const nimbus_diffusion_batch_dispatch& d)
{
    auto fin_pos = [](float x) { return std::isfinite(x) && x > 0.0f; };
    if (!d.fields || !d.tendencies || !d.locations || !d.rho || !d.dz || d.count < 1 ||
        d.count > kMaxDiffusionTargets || d.nx < 3 || d.ny < 3 || d.nz < 2 || !shape_ok(d.nx, d.ny, d.nz) ||
        !fin_pos(d.dx) || !fin_pos(d.dy) || !std::isfinite(d.K_h) || d.K_h < 0.0f || !std::isfinite(d.K_v) ||
        d.K_v < 0.0f)
        return set_error(NIMBUS_EINVAL, "diffusion batch: invalid dispatch");
    for (int k = 0; k < d.nz; ++k)
        if (!fin_pos(d.dz[k])) return set_error(NIMBUS_EINVAL, "diffusion batch: layer thickness not positive");
    const std::size_t ny = std::size_t(d.ny), nz = std::size_t(d.nz);
    for (int i = 1; i < d.nx - 1; ++i)
        for (int j = 1; j < d.ny - 1; ++j)
            for (int k = 0; k < d.nz; ++k)
                if (!fin_pos(d.rho[(i * ny + j) * nz + k]))
                    return set_error(NIMBUS_EINVAL, "diffusion batch: density not positive and finite");
    for (int t = 0; t < d.count; ++t)
    {
        if (!d.fields[t] || !d.tendencies[t] || d.locations[t] < 0 || d.locations[t] > 3)
            return set_error(NIMBUS_EINVAL, "diffusion batch: invalid target");
        const DiffDomain dm = diff_domain(d.nx, d.ny, d.nz, d.locations[t], d.c_grid != 0, d.periodic_x != 0,
                                          d.periodic_y != 0);
        const float* kv = d.vertical_K ? d.vertical_K[t] : nullptr;
        for (int i = dm.i0; i < dm.i1; ++i)
            for (int j = dm.j0; j < dm.j1; ++j)
                for (int k = dm.k0; k < dm.k1; ++k)
                {
                    const std::size_t c = (i * ny + j) * nz + k;
                    if (!std::isfinite(d.fields[t][c]))
                        return set_error(NIMBUS_EINVAL, "diffusion batch: non-finite field value");
                    if (kv && (!std::isfinite(kv[c]) || kv[c] < 0.0f))
                        return set_error(NIMBUS_EINVAL, "diffusion batch: vertical diffusivity negative or non-finite");
                }
    }
    return NIMBUS_OK;
}

// Device-resident scratch of the allocation-free transport entry point: one
// array per GPU, held for the life of the process. The host-side name is a
// stand-in, not a device address; kernels get the address that
// hipGetSymbolAddress returns for the current device.
constexpr std::size_t kNoAllocMaxCells = std::size_t(1) << 20;
static __device__ __align__(256) float scratch[kTransportScratchFields * kNoAllocMaxCells];

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mutex;
float* g_scratch_ptr[kMaxDevices] = {};

// The scratch address on `device` (the current device), looked up once.
int scratch_on(int device, float** out)
{
    if (device < 0 || device >= kMaxDevices) return set_error(NIMBUS_ENODEV, "no-alloc transport: device index too large");
    std::lock_guard<std::mutex> lock(g_scratch_mutex);
    if (!g_scratch_ptr[device])
    {
        void* ptr = nullptr;
        const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !ptr)
            return set_error(NIMBUS_EHIP, std::string("no-alloc transport: scratch address: ") + hipGetErrorString(e));
        g_scratch_ptr[device] = static_cast<float*>(ptr);
    }
    *out = g_scratch_ptr[device];
    return NIMBUS_OK;
}

}  // namespace
}  // namespace nimbus

using namespace nimbus;

extern "C" {

int nimbus_hip_dispatch_cartesian_cgrid_slow_tendencies(nimbus_hip_ctx* ctx,
                                                         const nimbus_cartesian_cgrid_slow_tendencies_dispatch* d)
{
    if (!ctx || !d) return set_error(NIMBUS_EINVAL, "cgrid slow tendencies: null argument");
    if (!d->u_data || !d->v_data || !d->w_data || !d->rho_data || !d->p_data || !d->rho0_base_data ||
        !d->du_dt_data || !d->dv_dt_data || !d->dw_dt_data || !d->drho_dt_data || !d->dp_dt_data)
        return set_error(NIMBUS_EINVAL, "cgrid slow tendencies: null field");
    if (!shape_ok(d->nr, d->nth, d->nz) || d->nz < 2 || d->rho0_base_len < d->nz || !(d->dx > 0) || !(d->dy > 0) ||
        !(d->dz > 0))
        return set_error(NIMBUS_EINVAL, "cgrid slow tendencies: invalid shape or spacing");
    if (d->coriolis_beta != 0.0f)
        return set_error(NIMBUS_EINVAL, "cgrid slow tendencies: beta-plane Coriolis is not supported (f-plane only)");
    const bool moist = d->has_moisture && d->qv_data && d->qv0_base_data && d->qv0_base_len >= d->nz;
    const bool load = d->has_loading && d->loading_data;
    const bool adv = d->du_advection_dt_data || d->dv_advection_dt_data || d->dw_advection_dt_data;
    if (adv && !(d->du_advection_dt_data && d->dv_advection_dt_data && d->dw_advection_dt_data))
        return set_error(NIMBUS_EINVAL, "cgrid slow tendencies: advection outputs come as a set of three");
    const int wind_len = d->u0_base_data && d->v0_base_data ? std::min(d->wind0_base_len, d->nz) : 0;
    hipSetDevice(context_device(ctx));
    hipStream_t s = context_stream(ctx);
    const std::size_t n = std::size_t(d->nr) * d->nth * d->nz, nz = std::size_t(d->nz);
    DeviceArrays A;
    SlowView v{};
    float* w_[19] = {};
    int rc = NIMBUS_OK;
    const float* hin[11] = {d->u_data, d->v_data, d->w_data, d->rho_data, d->p_data, d->rho0_base_data,
                            moist ? d->qv0_base_data : nullptr, wind_len ? d->u0_base_data : nullptr,
                            wind_len ? d->v0_base_data : nullptr, moist ? d->qv_data : nullptr,
                            load ? d->loading_data : nullptr};
    // host lengths: fields n; profiles nz (the winds their own length, at most nz)
    const std::size_t wl = std::size_t(wind_len);
    const std::size_t len[11] = {n, n, n, n, n, nz, nz, wl, wl, n, n};
    for (int q = 0; q < 11 && rc == NIMBUS_OK; ++q)
        rc = A.take(hin[q], hin[q] ? len[q] : std::max<std::size_t>(len[q], 1), s, &w_[q]);
    for (int q = 11; q < 19 && rc == NIMBUS_OK; ++q)
        if (q < 16 || adv) rc = A.take(nullptr, n, s, &w_[q]);
    if (rc != NIMBUS_OK) return rc;
    v.u = w_[0]; v.v = w_[1]; v.w = w_[2]; v.rho = w_[3]; v.p = w_[4]; v.rho0 = w_[5]; v.qv0 = w_[6];
    v.u0 = w_[7]; v.v0 = w_[8]; v.qv = w_[9]; v.load = w_[10];
    v.du = w_[11]; v.dv = w_[12]; v.dw = w_[13]; v.drho = w_[14]; v.dp = w_[15];
    v.adu = adv ? w_[16] : nullptr; v.adv = adv ? w_[17] : nullptr; v.adw = adv ? w_[18] : nullptr;
    v.nx = d->nr; v.ny = d->nth; v.nz = d->nz; v.wind_len = wind_len;
    v.idx = 1.0 / double(d->dx); v.idy = 1.0 / double(d->dy); v.idz = 1.0 / double(d->dz);
    v.g = d->g; v.f = d->coriolis_f_val; v.frame_u = d->frame_velocity_u_ms; v.frame_v = d->frame_velocity_v_ms;
    v.moisture = moist; v.loading = load;
    hipLaunchKernelGGL(k_ref_slow, dim3(blocks_for(n)), dim3(kThreads), 0, s, v);
    REF_TRY(hipGetLastError());
    REF_TRY(hipStreamSynchronize(s));
    float* outs[8] = {d->du_dt_data, d->dv_dt_data, d->dw_dt_data, d->drho_dt_data, d->dp_dt_data,
                      d->du_advection_dt_data, d->dv_advection_dt_data, d->dw_advection_dt_data};
    for (int q = 0; q < (adv ? 8 : 5) && rc == NIMBUS_OK; ++q) rc = back(outs[q], w_[11 + q], n, s);
    if (rc == NIMBUS_OK) REF_TRY(hipStreamSynchronize(s));
    if (rc == NIMBUS_OK) count_dispatch(ctx, "cartesian_cgrid_slow_tendencies");
    return rc;
}

int nimbus_hip_dispatch_slow_apply(nimbus_hip_ctx* ctx, const nimbus_slow_apply_dispatch* d)
{
    if (!ctx || !d) return set_error(NIMBUS_EINVAL, "slow apply: null argument");
    if (!d->u || !d->v || !d->w || !d->p_data || !d->du_dt || !d->dv_dt || !d->dw_dt || !d->dp_dt ||
        !shape_ok(d->nx, d->ny, d->nz) || !std::isfinite(d->dt_large))
        return set_error(NIMBUS_EINVAL, "slow apply: invalid dispatch");
    hipSetDevice(context_device(ctx));
    hipStream_t s = context_stream(ctx);
    const std::size_t n = std::size_t(d->nx) * d->ny * d->nz;
    DeviceArrays A;
    float* a[10] = {};
    const float* hin[10] = {d->u, d->v, d->w, d->p_data, d->du_dt, d->dv_dt, d->dw_dt, d->dp_dt, d->du_dt_pbl,
                            d->dv_dt_pbl};
    int rc = NIMBUS_OK;
    for (int q = 0; q < 10 && rc == NIMBUS_OK; ++q) rc = A.take(hin[q], n, s, &a[q]);
    if (rc != NIMBUS_OK) return rc;
    hipLaunchKernelGGL(k_ref_slow_apply, dim3(blocks_for(n)), dim3(kThreads), 0, s, n, a[0], a[1], a[2], a[3], a[4],
                       a[5], a[6], a[7], a[8], a[9], double(d->dt_large), d->sanitize ? 1 : 0);
    REF_TRY(hipGetLastError());
    REF_TRY(hipStreamSynchronize(s));
    float* outs[4] = {d->u, d->v, d->w, d->p_data};
    rc = commit_four(outs, a, n, s, d->completed_observer, d->completed_observer_context);
    if (rc == NIMBUS_OK) count_dispatch(ctx, "slow_apply");
    return rc;
}

int nimbus_hip_dispatch_state_relaxation_cartesian(nimbus_hip_ctx* ctx, const nimbus_state_relaxation_dispatch* d)
{
    if (!ctx || !d) return set_error(NIMBUS_EINVAL, "state relaxation: null argument");
    if (!d->u || !d->v || !d->w || !d->theta_data || !d->profiles || !shape_ok(d->nx, d->ny, d->nz) ||
        d->profiles_len < 7 * d->nz || d->lateral_width < 0 || d->lateral_width > d->nz ||
        !std::isfinite(d->dt_step))
        return set_error(NIMBUS_EINVAL, "state relaxation: invalid dispatch (profiles need 7 nz floats, "
                                        "lateral width at most nz)");
    hipSetDevice(context_device(ctx));
    hipStream_t s = context_stream(ctx);
    const std::size_t n = std::size_t(d->nx) * d->ny * d->nz;
    DeviceArrays A;
    float* a[5] = {};
    const float* hin[4] = {d->u, d->v, d->w, d->theta_data};
    int rc = NIMBUS_OK;
    for (int q = 0; q < 4 && rc == NIMBUS_OK; ++q) rc = A.take(hin[q], n, s, &a[q]);
    if (rc == NIMBUS_OK) rc = A.take(d->profiles, std::size_t(d->profiles_len), s, &a[4]);
    if (rc != NIMBUS_OK) return rc;
    hipLaunchKernelGGL(k_ref_relax, dim3(blocks_for(n)), dim3(kThreads), 0, s, d->nx, d->ny, d->nz, a[0], a[1], a[2],
                       a[3], a[4], d->lateral_width, d->do_upper ? 1 : 0, d->do_w_only ? 1 : 0,
                       d->do_lateral ? 1 : 0, double(d->dt_step));
    REF_TRY(hipGetLastError());
    REF_TRY(hipStreamSynchronize(s));
    float* outs[4] = {d->u, d->v, d->w, d->theta_data};
    rc = commit_four(outs, a, n, s, d->completed_observer, d->completed_observer_context);
    if (rc == NIMBUS_OK) count_dispatch(ctx, "state_relaxation_cartesian");
    return rc;
}

int nimbus_hip_dispatch_microphysics_apply(nimbus_hip_ctx* ctx, const nimbus_microphysics_apply_dispatch* d)
{
    if (!ctx || !d) return set_error(NIMBUS_EINVAL, "microphysics apply: null argument");
    float* fields[8] = {d->theta_data, d->qv_data, d->qc_data, d->qr_data, d->qi_data, d->qs_data, d->qg_data,
                        d->qh_data};
    const float* tend[8] = {d->dtheta, d->dqv, d->dqc, d->dqr, d->dqi, d->dqs, d->dqg, d->dqh};
    if (!shape_ok(d->nx, d->ny, d->nz) || !d->theta_data || !d->qv_data || !d->qc_data || !d->qr_data ||
        !d->dtheta || !d->dqv || !d->dqc || !d->dqr || d->i_begin < 0 || d->i_end > d->nx || d->j_begin < 0 ||
        d->j_end > d->ny || !std::isfinite(d->dt_step) || (d->stats_out && d->stats_len < 10))
        return set_error(NIMBUS_EINVAL, "microphysics apply: invalid dispatch");
    hipSetDevice(context_device(ctx));
    hipStream_t s = context_stream(ctx);
    const std::size_t n = std::size_t(d->nx) * d->ny * d->nz;
    DeviceArrays A;
    ApplyView v{};
    int rc = NIMBUS_OK;
    for (int q = 0; q < 8 && rc == NIMBUS_OK; ++q)
    {
        float* p = nullptr;
        rc = A.take(fields[q], n, s, &p);
        v.q[q] = p;
        if (rc == NIMBUS_OK) rc = A.take(tend[q], n, s, &p);
        v.d[q] = p;
    }
    float *drad = nullptr, *dpbl = nullptr, *dqvp = nullptr, *st = nullptr;
    if (rc == NIMBUS_OK) rc = A.take(d->dtheta_rad, n, s, &drad);
    if (rc == NIMBUS_OK) rc = A.take(d->dtheta_pbl, n, s, &dpbl);
    if (rc == NIMBUS_OK) rc = A.take(d->dqv_pbl, n, s, &dqvp);
    if (rc == NIMBUS_OK && d->stats_out)
        rc = A.take(reinterpret_cast<const float*>(d->stats_out), std::size_t(d->stats_len), s, &st);
    if (rc != NIMBUS_OK) return rc;
    v.drad = drad; v.dpbl = dpbl; v.dqvpbl = dqvp;
    v.stats = reinterpret_cast<unsigned*>(st);
    v.stats_len = d->stats_out ? d->stats_len : 0;
    v.nx = d->nx; v.ny = d->ny; v.nz = d->nz;
    v.i0 = d->i_begin; v.i1 = d->i_end; v.j0 = d->j_begin; v.j1 = d->j_end;
    v.dt = d->dt_step; v.th_fb = d->theta_fallback; v.th_min = d->theta_min; v.th_max = d->theta_max;
    v.max_step = d->max_theta_step; v.qv_max = d->qv_max_val; v.hydro_max = d->hydro_max_val;
    hipLaunchKernelGGL(k_ref_micro_apply, dim3(blocks_for(n)), dim3(kThreads), 0, s, v);
    REF_TRY(hipGetLastError());
    REF_TRY(hipStreamSynchronize(s));
    for (int q = 0; q < 8 && rc == NIMBUS_OK; ++q)
        if (fields[q]) rc = back(fields[q], v.q[q], n, s);
    if (rc == NIMBUS_OK && d->stats_out) rc = back(reinterpret_cast<float*>(d->stats_out), st, d->stats_len, s);
    if (rc == NIMBUS_OK) REF_TRY(hipStreamSynchronize(s));
    if (rc == NIMBUS_OK) count_dispatch(ctx, "microphysics_apply");
    return rc;
}

int nimbus_hip_dispatch_kessler_pointwise(nimbus_hip_ctx* ctx, const nimbus_kessler_pointwise_dispatch* d)
{
    if (!ctx || !d) return set_error(NIMBUS_EINVAL, "kessler pointwise: null argument");
    if (!d->temperature_data || !d->p_data || !d->qv_data || !d->qc_data || !d->qr_data || !d->rho_d_data ||
        !d->dtheta_dt_data || !d->dqv_dt_data || !d->dqc_dt_data || !d->dqr_dt_data || !shape_ok(d->nr, d->nth, d->nz) ||
        d->i_begin < 0 || d->i_end > d->nr || d->j_begin < 0 || d->j_end > d->nth || d->k_begin < 0 ||
        d->k_end > d->nz || !(d->dt_step > 0.0f))
        return set_error(NIMBUS_EINVAL, "kessler pointwise: invalid dispatch");
    hipSetDevice(context_device(ctx));
    hipStream_t s = context_stream(ctx);
    const std::size_t n = std::size_t(d->nr) * d->nth * d->nz;
    DeviceArrays A;
    float* a[12] = {};
    const float* hin[6] = {d->temperature_data, d->p_data, d->qv_data, d->qc_data, d->qr_data, d->rho_d_data};
    int rc = NIMBUS_OK;
    for (int q = 0; q < 12 && rc == NIMBUS_OK; ++q) rc = A.take(q < 6 ? hin[q] : nullptr, n, s, &a[q]);
    if (rc != NIMBUS_OK) return rc;
    KesslerPtView v{a[0], a[1], a[2], a[3], a[4], a[5], a[6], a[7], a[8], a[9], a[10], a[11],
                    d->nr, d->nth, d->nz, d->i_begin, d->i_end, d->j_begin, d->j_end, d->k_begin, d->k_end,
                    d->qc0, d->c_auto, d->c_accr, d->evaporation_efficiency, d->Lv_cp, d->T0, d->dt_step};
    hipLaunchKernelGGL(k_ref_kessler_pointwise, dim3(blocks_for(n)), dim3(kThreads), 0, s, v);
    REF_TRY(hipGetLastError());
    REF_TRY(hipStreamSynchronize(s));
    float* outs[6] = {d->dtheta_dt_data, d->dqv_dt_data, d->dqc_dt_data, d->dqr_dt_data, d->dqg_dt_data,
                      d->dqh_dt_data};
    for (int q = 0; q < 6 && rc == NIMBUS_OK; ++q)
        if (outs[q]) rc = back(outs[q], a[6 + q], n, s);
    if (rc == NIMBUS_OK) REF_TRY(hipStreamSynchronize(s));
    if (rc == NIMBUS_OK) count_dispatch(ctx, "kessler_pointwise");
    return rc;
}

int nimbus_hip_dispatch_kessler_sedimentation(nimbus_hip_ctx* ctx, const nimbus_kessler_sedimentation_dispatch* d)
{
    if (!ctx || !d) return set_error(NIMBUS_EINVAL, "kessler sedimentation: null argument");
    if (!d->qr_data || !d->dqr_dt_data || !shape_ok(d->nr, d->nth, d->nz) || d->i_begin < 0 || d->i_end > d->nr ||
        d->j_begin < 0 || d->j_end > d->nth || d->k_begin < 0 || d->k_end > d->nz ||
        (!d->layer_thickness_m && !(d->dz_val > 0.0f)) || !(d->Vt_max_rain >= 0.0f))
        return set_error(NIMBUS_EINVAL, "kessler sedimentation: invalid dispatch");
    hipSetDevice(context_device(ctx));
    hipStream_t s = context_stream(ctx);
    const std::size_t n = std::size_t(d->nr) * d->nth * d->nz;
    DeviceArrays A;
    float *qr = nullptr, *rho = nullptr, *layer = nullptr, *dqr = nullptr;
    int rc = A.take(d->qr_data, n, s, &qr);
    if (rc == NIMBUS_OK && d->rho_d_data) rc = A.take(d->rho_d_data, n, s, &rho);
    if (rc == NIMBUS_OK && d->layer_thickness_m) rc = A.take(d->layer_thickness_m, n, s, &layer);
    if (rc == NIMBUS_OK) rc = A.take(d->dqr_dt_data, n, s, &dqr);
    if (rc != NIMBUS_OK) return rc;
    KesslerSedView v{qr, rho, layer, dqr, d->nr, d->nth, d->nz, d->i_begin, d->i_end, d->j_begin, d->j_end,
                     d->k_begin, d->k_end, d->dz_val, d->Vt_max_rain, d->vt_cap};
    hipLaunchKernelGGL(k_ref_kessler_sed, dim3(blocks_for(std::size_t(d->nr) * d->nth)), dim3(kThreads), 0, s, v);
    REF_TRY(hipGetLastError());
    REF_TRY(hipStreamSynchronize(s));
    rc = back(d->dqr_dt_data, dqr, n, s);
    if (rc == NIMBUS_OK) REF_TRY(hipStreamSynchronize(s));
    if (rc == NIMBUS_OK) count_dispatch(ctx, "kessler_sedimentation");
    return rc;
}

int nimbus_hip_dispatch_advect_scalars_conservative_cartesian_cgrid(
    nimbus_hip_ctx* ctx, const nimbus_conservative_cartesian_cgrid_transport_dispatch* d)
{
    if (!ctx || !d) return set_error(NIMBUS_EINVAL, "conservative transport: null argument");
    int rc = validate_transport(*d);
    if (rc != NIMBUS_OK) return rc;
    hipSetDevice(context_device(ctx));
    hipStream_t s = context_stream(ctx);
    const std::size_t n = std::size_t(d->nx) * d->ny * d->nz;
    const bool sup = d->use_supplied_mass_flux != 0;
    // device copies of the inputs, one output per scalar, and the scratch
    DeviceArrays A;
    TransportArrays t{};
    float *rho = nullptr, *a = nullptr, *b = nullptr, *c = nullptr, *e = nullptr;
    rc = A.take(d->rho_data, n, s, &rho);
    if (rc == NIMBUS_OK) rc = A.take(sup ? d->mass_flux_x : d->u, n, s, &a);
    if (rc == NIMBUS_OK) rc = A.take(sup ? d->mass_flux_y : d->v, n, s, &b);
    if (rc == NIMBUS_OK) rc = A.take(sup ? d->mass_flux_z : d->w, n, s, &c);
    if (rc == NIMBUS_OK && sup) rc = A.take(d->density_begin, n, s, &e);
    std::vector<float*> qin(d->n_scalars), qout(d->n_scalars);
    for (int q = 0; q < d->n_scalars && rc == NIMBUS_OK; ++q)
    {
        rc = A.take(d->q_in[q], n, s, &qin[q]);
        if (rc == NIMBUS_OK) rc = A.take(nullptr, n, s, &qout[q]);
    }
    float* scratch = nullptr;
    if (rc == NIMBUS_OK) rc = A.take(nullptr, kTransportScratchFields * n, s, &scratch);
    if (rc != NIMBUS_OK) return rc;
    t.rho = rho;
    if (sup) { t.mfx = a; t.mfy = b; t.mfz = c; t.rho_begin = e; }
    else { t.u = a; t.v = b; t.w = c; }
    t.q_in = qin.data();
    t.q_out = qout.data();
    enqueue_transport(*d, t, scratch, s);
    REF_TRY(hipGetLastError());
    REF_TRY(hipStreamSynchronize(s));
    for (int q = 0; q < d->n_scalars && rc == NIMBUS_OK; ++q) rc = back(d->q_out[q], qout[q], n, s);
    if (rc == NIMBUS_OK) REF_TRY(hipStreamSynchronize(s));
    if (rc == NIMBUS_OK) count_dispatch(ctx, "advect_scalars_conservative_cartesian_cgrid");
    return rc;
}

int nimbus_hip_dispatch_diffusion_batch(nimbus_hip_ctx* ctx, const nimbus_diffusion_batch_dispatch* d)
{
    if (!ctx || !d) return set_error(NIMBUS_EINVAL, "diffusion batch: null argument");
    int rc = validate_diffusion(*d);
    if (rc != NIMBUS_OK) return rc;
    hipSetDevice(context_device(ctx));
    hipStream_t s = context_stream(ctx);
    const std::size_t n = std::size_t(d->nx) * d->ny * d->nz;
    DeviceArrays A;
    float *rho = nullptr, *dz = nullptr;
    rc = A.take(d->rho, n, s, &rho);
    if (rc == NIMBUS_OK) rc = A.take(d->dz, std::size_t(d->nz), s, &dz);
    std::vector<float*> out(d->count, nullptr);
    for (int t = 0; t < d->count && rc == NIMBUS_OK; ++t)
    {
        float *f = nullptr, *kv = nullptr;
        rc = A.take(d->fields[t], n, s, &f);
        const float* kvh = d->vertical_K ? d->vertical_K[t] : nullptr;
        if (rc == NIMBUS_OK && kvh) rc = A.take(kvh, n, s, &kv);
        if (rc == NIMBUS_OK) rc = A.take(nullptr, n, s, &out[t]);
        if (rc != NIMBUS_OK) break;
        DiffusionView v{f, rho, dz, kv, out[t], d->nx, d->ny, d->nz, d->locations[t], d->c_grid != 0 ? 1 : 0,
                        d->periodic_x != 0 ? 1 : 0, d->periodic_y != 0 ? 1 : 0, double(d->dx), double(d->dy),
                        double(d->K_h), double(d->K_v)};
        hipLaunchKernelGGL(k_ref_diffusion, dim3(blocks_for(n)), dim3(kThreads), 0, s, v);
    }
    if (rc != NIMBUS_OK) return rc;
    REF_TRY(hipGetLastError());
    REF_TRY(hipStreamSynchronize(s));
    // published only once every target has been computed
    for (int t = 0; t < d->count && rc == NIMBUS_OK; ++t) rc = back(d->tendencies[t], out[t], n, s);
    if (rc == NIMBUS_OK) REF_TRY(hipStreamSynchronize(s));
    if (rc == NIMBUS_OK) count_dispatch(ctx, "diffusion_batch");
    return rc;
}

int nimbus_hip_advect_scalars_conservative_noalloc(nimbus_hip_ctx* ctx,
                                                   const nimbus_conservative_cartesian_cgrid_transport_dispatch* d)
{
    if (!ctx || !d) return set_error(NIMBUS_EINVAL, "no-alloc transport: null argument");
    int rc = validate_transport(*d);
    if (rc != NIMBUS_OK) return rc;
    const std::size_t n = std::size_t(d->nx) * d->ny * d->nz;
    if (n > kNoAllocMaxCells)
        return set_error(NIMBUS_EINVAL, "no-alloc transport: " + std::to_string(n) + " cells exceed the scratch maximum of " +
                                            std::to_string(kNoAllocMaxCells) + " cells");
    const int device = context_device(ctx);
    REF_TRY(hipSetDevice(device));
    float* scr = nullptr;
    rc = scratch_on(device, &scr);
    if (rc != NIMBUS_OK) return rc;
    hipStream_t s = context_stream(ctx);
    const bool sup = d->use_supplied_mass_flux != 0;
    TransportArrays t{};
    t.rho = d->rho_data;
    if (sup) { t.mfx = d->mass_flux_x; t.mfy = d->mass_flux_y; t.mfz = d->mass_flux_z; t.rho_begin = d->density_begin; }
    else { t.u = d->u; t.v = d->v; t.w = d->w; }
    t.q_in = d->q_in;
    t.q_out = d->q_out;
    enqueue_transport(*d, t, scr, s);
    REF_TRY(hipGetLastError());
    REF_TRY(hipStreamSynchronize(s));
    count_dispatch(ctx, "advect_scalars_conservative_noalloc");
    return NIMBUS_OK;
}

}  // extern "C"
