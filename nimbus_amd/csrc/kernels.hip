// CDNA4 (gfx950) kernels of the Klemp-Wilhelmson split-explicit large step.
//
// Device layout: every 3-D field is [k][j][i] with i contiguous and a
// 64-byte-padded row pitch, so the lanes of a wavefront read consecutive i of
// two (j, k) rows (coalesced) and vertical neighbours are one plane apart.
// Cell kernels give each 256-thread block a 32 x 8 (i, j) tile of one k
// plane, ordered so that each XCD sweeps a contiguous slab of planes; column
// kernels give each thread one (i, j) column and march k. rho and p are
// stored as perturbations from the FP64 base state; theta, winds and water
// species are stored whole.
//
// Precision: S is the storage type. Continuity, EOS, pressure gradient,
// base-state terms, slow tendencies, mass-flux accumulation and the HE-VI
// Thomas solve compute in FP64 for both storage types; WENO5 transport and
// Kessler microphysics compute in the storage type (FP32 bulk in the mixed
// mode).
#include "device_math.hpp"
#include "kernels.hpp"

#include <algorithm>

namespace nimbus
{

namespace
{

constexpr int kBlock = 256;

template <class S>
struct AccT;
template <>
struct AccT<float>
{
    using type = float;
};
template <>
struct AccT<double>
{
    using type = double;
};

// Workgroups are dispatched round-robin over the 8 XCDs (workgroup w runs on
// XCD w % 8, each XCD with its own L2). Remapping w so that XCD x receives a
// contiguous range of logical tiles keeps the j +- 1 rows and k +- 1 planes a
// stencil reads in the same L2.
__device__ __forceinline__ int xcd_logical(int w, int n)
{
    constexpr int kXcd = 8;
    const int x = w % kXcd, local = w / kXcd;
    const int q = n / kXcd, r = n % kXcd;
    return x < r ? x * (q + 1) + local : r * (q + 1) + (x - r) * q + local;
}

// Cell kernels: one tile_w x (256 / tile_w) (i, j) tile of one k plane per
// 256-thread block; logical tile order is k-major.
__device__ __forceinline__ bool cell_index(const DevGrid& g, int& i, int& j, int& k)
{
    const int n = g.tiles_plane * g.nz;
    const int logical = g.xcd_remap ? xcd_logical(int(blockIdx.x), n) : int(blockIdx.x);
    k = logical / g.tiles_plane;
    const int t = logical - k * g.tiles_plane;
    const int by = t / g.tiles_x, bx = t - by * g.tiles_x;
    i = (bx << g.tile_shift) + int(threadIdx.x & unsigned(g.tile_w - 1));
    j = (by << (8 - g.tile_shift)) + int(threadIdx.x >> g.tile_shift);
    return i < g.nx && j < g.ny;
}

// Column kernels: one thread per (i, j) column, same tile shape.
__device__ __forceinline__ bool column_index(const DevGrid& g, int& i, int& j)
{
    const int t = g.xcd_remap ? xcd_logical(int(blockIdx.x), g.tiles_plane) : int(blockIdx.x);
    const int by = t / g.tiles_x, bx = t - by * g.tiles_x;
    i = (bx << g.tile_shift) + int(threadIdx.x & unsigned(g.tile_w - 1));
    j = (by << (8 - g.tile_shift)) + int(threadIdx.x >> g.tile_shift);
    return i < g.nx && j < g.ny;
}

__device__ __forceinline__ double finite0(double x) { return isfinite(x) ? x : 0.0; }
__device__ __forceinline__ double safe_rho(double r) { return isfinite(r) && r > 1.0e-6 ? r : 1.0; }

// The reference sanitises the FP32 stored value (field_bounds.hpp:41-53).
template <class S>
__device__ __forceinline__ double clamp_store(double x, double fallback, float lo, float hi)
{
    if constexpr (sizeof(S) == sizeof(float))
    {
        float f = float(x);
        if (!isfinite(f)) f = float(fallback);
        return double(fminf(fmaxf(f, lo), hi));
    }
    else
    {
        if (!isfinite(x)) x = fallback;
        return fmin(fmax(x, double(lo)), double(hi));
    }
}

__device__ __forceinline__ const double* slot(const double* base, int s, int nz) { return base + std::size_t(s) * nz; }

// ---------------------------------------------------------------------
// Slow tendencies at one interior cell (cartesian_cgrid_3d.cpp:1064-1356)
// ---------------------------------------------------------------------
template <class S>
__device__ void slow_tendency_cell(const DevGrid& g, const S* __restrict__ u, const S* __restrict__ v,
                                   const S* __restrict__ w, const S* __restrict__ rp, const S* __restrict__ qv,
                                   const S* __restrict__ qc, const S* __restrict__ qr, const double* __restrict__ base,
                                   double f, int i, int j, int k, double& du, double& dv, double& dw)
{
    const int NX = g.nx, NY = g.ny, NZ = g.nz;
    const double ix = g.idx, iy = g.idy, iz = g.idz;
    const std::size_t c = g.at(i, j, k);
    const std::size_t sx = 1, sy = std::size_t(g.px), sz = std::size_t(g.px) * g.ny;
    auto F = [&](const S* a, std::size_t idx) { return double(a[idx]); };
    // u at x-face
    {
        const double ux = F(u, c);
        const double vy = 0.25 * (F(v, c - sy) + F(v, c) + F(v, c + sx - sy) + F(v, c + sx));
        const double wz = k == 0 ? 0.25 * (F(w, c) + F(w, c + sx))
                                 : 0.25 * (F(w, c - sz) + F(w, c) + F(w, c + sx - sz) + F(w, c + sx));
        const double ddx = dev::upwind_axis<double>([&](int o) { return F(u, c + std::ptrdiff_t(o)); }, i, NX, ux, ix);
        const double ddy = dev::upwind_axis<double>([&](int o) { return F(u, c + std::ptrdiff_t(o) * std::ptrdiff_t(sy)); }, j, NY, vy, iy);
        const double ddz = k == 0 ? (wz >= 0.0 ? 0.0 : (F(u, c + sz) - F(u, c)) * iz)
                                  : dev::upwind_axis<double>([&](int o) { return F(u, c + std::ptrdiff_t(o) * std::ptrdiff_t(sz)); }, k, NZ, wz, iz);
        const double cor = f * (vy - slot(base, kBaseV0, NZ)[k]);
        du = finite0(-ux * ddx - vy * ddy - wz * ddz + cor);
    }
    // v at y-face
    {
        const double vy = F(v, c);
        const double ux = 0.25 * (F(u, c - sx) + F(u, c) + F(u, c - sx + sy) + F(u, c + sy));
        const double wz = k == 0 ? 0.25 * (F(w, c) + F(w, c + sy))
                                 : 0.25 * (F(w, c - sz) + F(w, c) + F(w, c + sy - sz) + F(w, c + sy));
        const double ddx = dev::upwind_axis<double>([&](int o) { return F(v, c + std::ptrdiff_t(o)); }, i, NX, ux, ix);
        const double ddy = dev::upwind_axis<double>([&](int o) { return F(v, c + std::ptrdiff_t(o) * std::ptrdiff_t(sy)); }, j, NY, vy, iy);
        const double ddz = k == 0 ? (wz >= 0.0 ? 0.0 : (F(v, c + sz) - F(v, c)) * iz)
                                  : dev::upwind_axis<double>([&](int o) { return F(v, c + std::ptrdiff_t(o) * std::ptrdiff_t(sz)); }, k, NZ, wz, iz);
        const double cor = f * (ux - slot(base, kBaseU0, NZ)[k]);
        dv = finite0(-ux * ddx - vy * ddy - wz * ddz - cor);
    }
    // w at z-face with buoyancy, moisture buoyancy and condensate loading
    {
        const double wz = F(w, c);
        const double ux = 0.25 * (F(u, c - sx) + F(u, c) + F(u, c - sx + sz) + F(u, c + sz));
        const double vy = 0.25 * (F(v, c - sy) + F(v, c) + F(v, c - sy + sz) + F(v, c + sz));
        const double ddx = dev::upwind_axis<double>([&](int o) { return F(w, c + std::ptrdiff_t(o)); }, i, NX, ux, ix);
        const double ddy = dev::upwind_axis<double>([&](int o) { return F(w, c + std::ptrdiff_t(o) * std::ptrdiff_t(sy)); }, j, NY, vy, iy);
        const double ddz = k == 0 ? (wz >= 0.0 ? F(w, c) * iz : (F(w, c + sz) - F(w, c)) * iz)
                                  : dev::upwind_axis<double>([&](int o) { return F(w, c + std::ptrdiff_t(o) * std::ptrdiff_t(sz)); }, k, NZ, wz, iz);
        const double* rho0 = slot(base, kBaseRho0, NZ);
        const double* qv0s = slot(base, kBaseQv0Stored, NZ);
        const double rpf = 0.5 * (F(rp, c) + F(rp, c + sz));
        const double rf = 0.5 * (rho0[k] + rho0[k + 1]) + rpf;
        const double buoy = -kG * rpf / safe_rho(rf);
        const double qvf = 0.5 * (F(qv, c) + F(qv, c + sz));
        const double mbuoy = kG * 0.608 * (qvf - 0.5 * (qv0s[k] + qv0s[k + 1]));
        const double load = -kG * (0.5 * ((F(qc, c) + F(qr, c)) + (F(qc, c + sz) + F(qr, c + sz))));
        dw = finite0(-ux * ddx - vy * ddy - wz * ddz + buoy + mbuoy + load);
    }
}

template <class S>
__device__ __forceinline__ bool slow_interior(const DevGrid& g, int i, int j, int k)
{
    return i >= 1 && i < g.nx - 1 && j >= 1 && j < g.ny - 1 && k < g.nz - 1;
}

// Heun stage 1: k1 at u_n and the predictor u_n + dt k1 (dynamics.cpp:5178-5232).
template <class S>
__global__ __launch_bounds__(kBlock) void k_slow_stage1(DevGrid g, const S* u, const S* v, const S* w, const S* rp,
                                                        const S* qv, const S* qc, const S* qr, const double* base,
                                                        double f, double dt, S* k1u, S* k1v, S* k1w, S* su, S* sv,
                                                        S* sw)
{
    int i, j, k;
    if (!cell_index(g, i, j, k)) return;
    const std::size_t c = g.at(i, j, k);
    double du = 0.0, dv = 0.0, dw = 0.0;
    if (slow_interior<S>(g, i, j, k)) slow_tendency_cell(g, u, v, w, rp, qv, qc, qr, base, f, i, j, k, du, dv, dw);
    const S a = S(du), b = S(dv), d = S(dw);
    k1u[c] = a;
    k1v[c] = b;
    k1w[c] = d;
    su[c] = S(double(u[c]) + double(a) * dt);
    sv[c] = S(double(v[c]) + double(b) * dt);
    sw[c] = S(double(w[c]) + double(d) * dt);
}

// Heun stage 2 + average in storage precision + slow apply (dynamics.cpp:5233-5292, 2055-2095).
template <class S>
__global__ __launch_bounds__(kBlock) void k_slow_stage2(DevGrid g, S* u, S* v, S* w, const S* rp, const S* qv,
                                                        const S* qc, const S* qr, const double* base, double f,
                                                        double dt, const S* k1u, const S* k1v, const S* k1w,
                                                        const S* su, const S* sv, const S* sw)
{
    int i, j, k;
    if (!cell_index(g, i, j, k)) return;
    const std::size_t c = g.at(i, j, k);
    double du = 0.0, dv = 0.0, dw = 0.0;
    if (slow_interior<S>(g, i, j, k)) slow_tendency_cell(g, su, sv, sw, rp, qv, qc, qr, base, f, i, j, k, du, dv, dw);
    const S au = S(S(0.5) * (S(du) + k1u[c]));
    const S av = S(S(0.5) * (S(dv) + k1v[c]));
    const S aw = S(S(0.5) * (S(dw) + k1w[c]));
    u[c] = S(double(u[c]) + finite0(double(au)) * dt);
    v[c] = S(double(v[c]) + finite0(double(av)) * dt);
    w[c] = S(double(w[c]) + finite0(double(aw)) * dt);
}

template <class S>
__global__ __launch_bounds__(kBlock) void k_slow_tendencies(DevGrid g, const S* u, const S* v, const S* w,
                                                            const S* rp, const S* qv, const S* qc, const S* qr,
                                                            const double* base, double f, S* du_o, S* dv_o, S* dw_o)
{
    int i, j, k;
    if (!cell_index(g, i, j, k)) return;
    const std::size_t c = g.at(i, j, k);
    double du = 0.0, dv = 0.0, dw = 0.0;
    if (slow_interior<S>(g, i, j, k)) slow_tendency_cell(g, u, v, w, rp, qv, qc, qr, base, f, i, j, k, du, dv, dw);
    du_o[c] = S(du);
    dv_o[c] = S(dv);
    dw_o[c] = S(dw);
}

// ---------------------------------------------------------------------
// Acoustic substep, explicit forward step: retained mass-flux
// accumulation from the pre-continuity state (advection_cartesian.cpp:610-830),
// C-grid continuity, density clamp and EOS pressure
// (acoustic_cgrid_fast.cpp:49-157).
// ---------------------------------------------------------------------
// Clamp in FP64 without rounding the whole value to the storage type, so the
// stored perturbation keeps its own precision.
__device__ __forceinline__ double clamp_full(double x, double fallback, float lo, float hi)
{
    if (!isfinite(x)) x = fallback;
    return fmin(fmax(x, double(lo)), double(hi));
}

// Pressure perturbation p - p0 from the dry EOS. FP64 storage evaluates
// p00 (rho R theta / p00)^gamma in FP64; FP32 storage evaluates the same
// expression relative to the base state, p = pe0 (1 + x)^gamma with
// 1 + x = rho theta / (rho0 theta0) and pe0 = EOS(rho0, theta0) in FP64, so
// p - p0 = pe0 expm1(gamma log1p(x)) + (pe0 - p0) keeps FP32 relative
// accuracy on the perturbation rather than on the full pressure.
template <class S>
__device__ __forceinline__ double eos_full_pressure(const double* base, int nz, int k, double rho, double theta)
{
    if constexpr (sizeof(S) == sizeof(float))
    {
        const double x = rho * theta * base[std::size_t(kBaseInvRT0) * nz + k] - 1.0;
        const float e = expm1f(float(kGamma) * log1pf(float(x)));
        return base[std::size_t(kBaseP0) * nz + k] + (double(e) * base[std::size_t(kBasePe0) * nz + k] +
                                                      base[std::size_t(kBasePeDiff) * nz + k]);
    }
    else
    {
        return dev::eos_pressure(rho, theta);
    }
}

template <class S>
__device__ __forceinline__ void density_eos_store(const Bounds& bd, const double* base, int nz, int k, double rn,
                                                  double theta, S* rp_out, S* pp, std::size_t c)
{
    const double rho0k = base[std::size_t(kBaseRho0) * nz + k];
    const double p0k = base[std::size_t(kBaseP0) * nz + k];
    if (!isfinite(rn) || rn <= 0.0) rn = fmax(0.1, rho0k);
    const double rs = clamp_full(rn, 1.0, bd.rho_min, bd.rho_max);
    const S rps = S(rs - rho0k);
    rp_out[c] = rps;
    const double rho_used = rho0k + double(rps);
    double pn = double(bd.p_fallback);
    if (theta > 0.0) pn = eos_full_pressure<S>(base, nz, k, rho_used, theta);
    if (!isfinite(pn) || pn <= 0.0) pn = double(bd.p_fallback);
    pp[c] = S(clamp_full(pn, 101325.0, bd.p_min, bd.p_max) - p0k);
}

template <class S>
__global__ __launch_bounds__(kBlock) void k_acoustic_forward(DevGrid g, const S* __restrict__ u,
                                                             const S* __restrict__ v, const S* __restrict__ w,
                                                             const S* __restrict__ rp_in, S* __restrict__ rp_out,
                                                             S* __restrict__ pp, const S* __restrict__ th,
                                                             S* __restrict__ mfx, S* __restrict__ mfy,
                                                             S* __restrict__ mfz, S* __restrict__ rb,
                                                             const double* __restrict__ base, Bounds bd, double dts,
                                                             double weight, int first, int accumulate, int top_open)
{
    int i, j, k;
    if (!cell_index(g, i, j, k)) return;
    const int NX = g.nx, NY = g.ny, NZ = g.nz;
    const std::size_t c = g.at(i, j, k);
    const std::size_t sy = std::size_t(g.px), sz = std::size_t(g.px) * g.ny;
    const double* rho0 = slot(base, kBaseRho0, NZ);
    const double* p0 = slot(base, kBaseP0, NZ);
    const double r0 = rho0[k];
    const double rc = r0 + double(rp_in[c]);
    if (accumulate)
    {
        if (i < NX - 1)
        {
            const double fl = 0.5 * (rc + (r0 + double(rp_in[c + 1]))) * double(u[c]);
            mfx[c] = S((first ? 0.0 : double(mfx[c])) + weight * fl);
        }
        if (j < NY - 1)
        {
            const double fl = 0.5 * (rc + (r0 + double(rp_in[c + sy]))) * double(v[c]);
            mfy[c] = S((first ? 0.0 : double(mfy[c])) + weight * fl);
        }
        double fl = 0.0;
        if (k < NZ - 1) fl = 0.5 * (rc + (rho0[k + 1] + double(rp_in[c + sz]))) * double(w[c]);
        else if (top_open) fl = rc * double(w[c]);
        mfz[c] = S((first ? 0.0 : double(mfz[c])) + weight * fl);
    }
    if (first) rb[c] = rp_in[c];
    double d = 0.0;
    if (i >= 1 && i < NX - 1 && j >= 1 && j < NY - 1)
    {
        const double rxp = r0 + double(rp_in[c + 1]), rxm = r0 + double(rp_in[c - 1]);
        const double ryp = r0 + double(rp_in[c + sy]), rym = r0 + double(rp_in[c - sy]);
        const double fxp = 0.5 * (rc + rxp) * double(u[c]);
        const double fxm = 0.5 * (rxm + rc) * double(u[c - 1]);
        const double fyp = 0.5 * (rc + ryp) * double(v[c]);
        const double fym = 0.5 * (rym + rc) * double(v[c - sy]);
        const double fzp = k + 1 < NZ ? 0.5 * (rc + (rho0[k + 1] + double(rp_in[c + sz]))) * double(w[c])
                                      : rc * double(w[c]);
        const double fzm = k > 0 ? 0.5 * ((rho0[k - 1] + double(rp_in[c - sz])) + rc) * double(w[c - sz]) : 0.0;
        d = -((fxp - fxm) * g.idx + (fyp - fym) * g.idy + (fzp - fzm) * g.idz);
    }
    const S drho = S(finite0(d));
    density_eos_store<S>(bd, base, NZ, k, rc + double(drho) * dts, double(th[c]), rp_out, pp, c);
}

// Backward momentum update with the new pressure (acoustic_cgrid_fast.cpp:159-215);
// also commits the new density to the home buffer. With skip_w the interior
// w (solved by the HE-VI column kernel) is left untouched.
template <class S>
__global__ __launch_bounds__(kBlock) void k_acoustic_backward(DevGrid g, S* __restrict__ u, S* __restrict__ v,
                                                              S* __restrict__ w, const S* __restrict__ rp_new,
                                                              S* __restrict__ rp_home, const S* __restrict__ pp,
                                                              const double* __restrict__ base, Bounds bd,
                                                              double dts, int skip_w)
{
    int i, j, k;
    if (!cell_index(g, i, j, k)) return;
    const int NX = g.nx, NY = g.ny, NZ = g.nz;
    const std::size_t c = g.at(i, j, k);
    const std::size_t sy = std::size_t(g.px), sz = std::size_t(g.px) * g.ny;
    const double* rho0 = slot(base, kBaseRho0, NZ);
    double du = 0.0, dv = 0.0, dw = 0.0;
    const bool interior = i >= 1 && i < NX - 1 && j >= 1 && j < NY - 1 && k < NZ - 1;
    if (interior)
    {
        const double r0 = rho0[k];
        const double rc = r0 + double(rp_new[c]);
        const double rx = 0.5 * (rc + (r0 + double(rp_new[c + 1])));
        const double ry = 0.5 * (rc + (r0 + double(rp_new[c + sy])));
        const double rz = 0.5 * (rc + (rho0[k + 1] + double(rp_new[c + sz])));
        if constexpr (sizeof(S) == sizeof(float))
        {
            // FP32 tendency, as the reference rounds it to FP32 before the update
            auto safe_f = [](float r) { return isfinite(r) && r > 1.0e-6f ? r : 1.0f; };
            auto fin = [](float x) { return isfinite(x) ? x : 0.0f; };
            const float pc = float(pp[c]);
            du = double(fin(-(float(pp[c + 1]) - pc) * float(g.idx) / safe_f(float(rx))));
            dv = double(fin(-(float(pp[c + sy]) - pc) * float(g.idy) / safe_f(float(ry))));
            dw = double(fin(-((float(pp[c + sz]) - pc) * float(g.idz)) / safe_f(float(rz))));
        }
        else
        {
            const double pc = double(pp[c]);
            du = finite0(-(double(pp[c + 1]) - pc) * g.idx / safe_rho(rx));
            dv = finite0(-(double(pp[c + sy]) - pc) * g.idy / safe_rho(ry));
            dw = finite0(-((double(pp[c + sz]) - pc) * g.idz) / safe_rho(rz));
        }
    }
    u[c] = S(clamp_store<S>(double(u[c]) + du * dts, 0.0, -bd.wind_h, bd.wind_h));
    v[c] = S(clamp_store<S>(double(v[c]) + dv * dts, 0.0, -bd.wind_h, bd.wind_h));
    if (!(skip_w && interior)) w[c] = S(clamp_store<S>(double(w[c]) + dw * dts, 0.0, -bd.wind_v, bd.wind_v));
    if (rp_home != rp_new) rp_home[c] = rp_new[c];
}

// ---------------------------------------------------------------------
// HE-VI column kernel: horizontal continuity explicit, vertical w-rho-p
// coupling off-centred implicit (alpha = (1+beta)/2), p linearised about
// p* = EOS(rho*), FP64 Thomas sweep per column, flux-form density update and
// EOS re-diagnosis. Ghost columns get the reference's pass-2 clamp/EOS only.
// ---------------------------------------------------------------------
template <class S>
__global__ __launch_bounds__(kBlock) void k_hevi_column(DevGrid g, const S* __restrict__ u,
                                                        const S* __restrict__ v, S* __restrict__ w,
                                                        const S* __restrict__ rp_in, S* __restrict__ rp_out,
                                                        S* __restrict__ pp, const S* __restrict__ th,
                                                        S* __restrict__ mfx, S* __restrict__ mfy,
                                                        S* __restrict__ mfz, S* __restrict__ rb,
                                                        double* __restrict__ tc, double* __restrict__ tr,
                                                        const double* __restrict__ base, Bounds bd, double dts,
                                                        double alpha, double weight, int first, int accumulate)
{
    int i, j;
    if (!column_index(g, i, j)) return;
    const int NX = g.nx, NY = g.ny, NZ = g.nz;
    const std::size_t sy = std::size_t(g.px), sz = std::size_t(g.px) * g.ny;
    const double* rho0 = slot(base, kBaseRho0, NZ);
    const double* p0 = slot(base, kBaseP0, NZ);
    const bool interior = i >= 1 && i < NX - 1 && j >= 1 && j < NY - 1;
    const double idz = g.idz;
    const double a1 = 1.0 - alpha;
    // ---- upward sweep: mass fluxes, rho*, p*, Thomas forward elimination ----
    double rho_prev = 0.0, rb_prev = 0.0, C_prev = 0.0, ps_prev = 0.0, wo_prev = 0.0, e_prev = 0.0;
    double rb_prev2 = 0.0, cp_prev = 0.0, rr_prev = 0.0;
    bool ok = true;
    for (int k = 0; k < NZ; ++k)
    {
        const std::size_t c = g.at(i, j, k);
        const double r0 = rho0[k];
        const double rc = r0 + double(rp_in[c]);
        if (accumulate)
        {
            if (i < NX - 1)
            {
                const double fl = 0.5 * (rc + (r0 + double(rp_in[c + 1]))) * double(u[c]);
                mfx[c] = S((first ? 0.0 : double(mfx[c])) + weight * fl);
            }
            if (j < NY - 1)
            {
                const double fl = 0.5 * (rc + (r0 + double(rp_in[c + sy]))) * double(v[c]);
                mfy[c] = S((first ? 0.0 : double(mfy[c])) + weight * fl);
            }
            if (first && !interior) mfz[c] = S(0);
        }
        if (first) rb[c] = rp_in[c];
        if (!interior)
        {
            density_eos_store<S>(bd, base, NZ, k, rc, double(th[c]), rp_out, pp, c);
            continue;
        }
        const double rxp = r0 + double(rp_in[c + 1]), rxm = r0 + double(rp_in[c - 1]);
        const double ryp = r0 + double(rp_in[c + sy]), rym = r0 + double(rp_in[c - sy]);
        const double dh = -((0.5 * (rc + rxp) * double(u[c]) - 0.5 * (rxm + rc) * double(u[c - 1])) * g.idx +
                            (0.5 * (rc + ryp) * double(v[c]) - 0.5 * (rym + rc) * double(v[c - sy])) * g.idy);
        const double rstar = rc + dts * finite0(dh);
        const double ps = dev::eos_pressure(rstar, double(th[c]));
        const double C = kGamma * ps / rstar;
        const double wo = double(w[c]);
        // face density above level k (zero at the lid)
        const double rbk = k < NZ - 1 ? 0.5 * (rc + (rho0[k + 1] + double(rp_in[c + sz]))) : 0.0;
        const double e = -a1 * dts * idz * (rbk * wo - rb_prev * wo_prev);
        rp_out[c] = S(rstar - r0);
        if (k >= 1)
        {
            // row r = k-1 (face k-1/2): unknowns w_{r-1}, w_r, w_{r+1}
            const double rbr = rb_prev;
            const double Ak = dts * alpha / (rbr * g.dz);
            const double Bc = alpha * dts * idz;
            const double A = (k - 1) > 0 ? -Ak * Bc * C_prev * rb_prev2 : 0.0;
            const double U = (k - 1) < NZ - 2 ? -Ak * Bc * C * rbk : 0.0;
            const double D = 1.0 + Ak * Bc * (C + C_prev) * rbr;
            const double G = (ps - ps_prev) * idz - (p0[k] - p0[k - 1]) * idz;
            const double R = wo_prev - dts / rbr * G - Ak * (C * e - C_prev * e_prev);
            const double piv = D - A * cp_prev;
            if (!(piv > 0.0) || !isfinite(piv)) ok = false;
            const double cpr = U / piv;
            const double rrr = (R - A * rr_prev) / piv;
            const std::size_t cr = g.at(i, j, k - 1);
            tc[cr] = cpr;
            tr[cr] = rrr;
            cp_prev = cpr;
            rr_prev = rrr;
        }
        rho_prev = rc;
        rb_prev2 = rb_prev;
        rb_prev = rbk;
        C_prev = C;
        ps_prev = ps;
        wo_prev = wo;
        e_prev = e;
    }
    if (!interior) return;
    // ---- downward sweep: back substitution, density update, weighted z mass flux ----
    double w_above = 0.0, wo_above = 0.0, rb_above = 0.0;  // face r+1 quantities (stored w)
    double x_above = 0.0;                                  // face r+1 Thomas solution (unclamped)
    for (int r = NZ - 2; r >= -1; --r)
    {
        double wn = 0.0, wo = 0.0, rbr = 0.0;
        if (r >= 0)
        {
            const std::size_t cr = g.at(i, j, r);
            wo = double(w[cr]);
            const double x = ok ? tr[cr] - tc[cr] * x_above : wo;
            x_above = x;
            wn = clamp_store<S>(x, 0.0, -bd.wind_v, bd.wind_v);
            w[cr] = S(wn);
            wn = double(S(wn));
            rbr = 0.5 * ((rho0[r] + double(rp_in[cr])) + (rho0[r + 1] + double(rp_in[cr + sz])));
        }
        // finalise cell r+1: top face r+1 (zero at the lid), bottom face r
        const int kc = r + 1;
        const std::size_t cc = g.at(i, j, kc);
        const double ft = kc < NZ - 1 ? rb_above * (alpha * w_above + a1 * wo_above) : 0.0;
        const double fb = r >= 0 ? rbr * (alpha * wn + a1 * wo) : 0.0;
        if (accumulate) mfz[cc] = S((first ? 0.0 : double(mfz[cc])) + weight * ft);
        const double rstar = rho0[kc] + double(rp_out[cc]);
        density_eos_store<S>(bd, base, NZ, kc, rstar - dts * idz * (ft - fb), double(th[cc]), rp_out, pp, cc);
        w_above = wn;
        wo_above = wo;
        rb_above = rbr;
    }
}

// ---------------------------------------------------------------------
// Boundary closure (boundary_conditions_cartesian_cgrid.cpp:798-1175):
// west/east faces, then south/north, then the vertical closure; KW78
// radiative faces integrate their FP32 stash when dt_rad > 0.
// ---------------------------------------------------------------------
__device__ __forceinline__ float kw78_face(float prev, float phase, float grad, float dt, float inv_d, bool west_like)
{
    if (west_like ? phase >= 0.0f : phase <= 0.0f) return prev;
    const float a = __fmul_rn(dt, phase);
    const float b = __fmul_rn(a, grad);
    const float c = __fmul_rn(b, inv_d);
    return __fsub_rn(prev, c);
}

template <class S>
__global__ __launch_bounds__(kBlock) void k_bc_west_east(DevGrid g, S* u, S* v, S* w, S* rp, S* pp, float* rad_w,
                                                         float* rad_e, int bcw, int bce, float dt_rad, float inv_dx)
{
    const int t = blockIdx.x * kBlock + threadIdx.x;
    if (t >= g.ny * g.nz) return;
    const int k = t / g.ny, j = t - k * g.ny;
    const int NX = g.nx, NZ = g.nz;
    const bool radiate = dt_rad > 0.0f;
    const float cstar = 30.0f;
    const std::size_t row = g.at(0, j, k);
    auto cp_rp = [&](std::size_t d, std::size_t s) { rp[d] = rp[s]; pp[d] = pp[s]; };
    {
        const std::size_t c = row, in = row + 1;
        if (bcw == NIMBUS_BC_PERIODIC)
        {
            const std::size_t s = row + (NX - 2);
            u[c] = u[s]; v[c] = v[s]; w[c] = w[s]; cp_rp(c, s);
        }
        else
        {
            if (bcw == NIMBUS_BC_RADIATIVE)
            {
                if (radiate)
                {
                    float& up = rad_w[std::size_t(j) * NZ + k];
                    const float un = kw78_face(up, up - 0.0f - cstar, float(u[in]) - up, dt_rad, inv_dx, true);
                    u[c] = S(un);
                    up = un;
                }
            }
            else if (bcw == NIMBUS_BC_ZERO_GRADIENT) u[c] = u[in];
            else u[c] = S(0);
            if (bcw == NIMBUS_BC_RIGID) { v[c] = S(0); w[c] = S(0); }
            else { v[c] = v[in]; w[c] = w[in]; }
            cp_rp(c, in);
        }
    }
    {
        const std::size_t gi = row + (NX - 1), a = row + (NX - 2), am = row + (NX - 3);
        if (bce == NIMBUS_BC_PERIODIC)
        {
            const std::size_t s = row + 1;
            u[gi] = u[s]; v[gi] = v[s]; w[gi] = w[s]; cp_rp(gi, s);
        }
        else
        {
            if (bce == NIMBUS_BC_RIGID || bce == NIMBUS_BC_RIGID_SLIP) u[a] = S(0);
            if (bce == NIMBUS_BC_RADIATIVE && radiate)
            {
                float& up = rad_e[std::size_t(j) * NZ + k];
                const float un = kw78_face(up, up - 0.0f + cstar, up - float(u[am]), dt_rad, inv_dx, false);
                u[a] = S(un);
                up = un;
            }
            else if (bce == NIMBUS_BC_ZERO_GRADIENT) u[a] = u[am];
            u[gi] = (bce == NIMBUS_BC_RADIATIVE || bce == NIMBUS_BC_ZERO_GRADIENT) ? u[a] : S(0);
            if (bce == NIMBUS_BC_RIGID) { v[gi] = S(0); w[gi] = S(0); }
            else { v[gi] = v[a]; w[gi] = w[a]; }
            cp_rp(gi, a);
        }
    }
}

template <class S>
__global__ __launch_bounds__(kBlock) void k_bc_south_north(DevGrid g, S* u, S* v, S* w, S* rp, S* pp, float* rad_s,
                                                           float* rad_n, int bcs, int bcn, float dt_rad, float inv_dy)
{
    const int t = blockIdx.x * kBlock + threadIdx.x;
    if (t >= g.nx * g.nz) return;
    const int k = t / g.nx, i = t - k * g.nx;
    const int NY = g.ny, NZ = g.nz;
    const bool radiate = dt_rad > 0.0f;
    const float cstar = 30.0f;
    const std::size_t sy = std::size_t(g.px);
    auto cp_rp = [&](std::size_t d, std::size_t s) { rp[d] = rp[s]; pp[d] = pp[s]; };
    {
        const std::size_t c = g.at(i, 0, k), in = c + sy;
        if (bcs == NIMBUS_BC_PERIODIC)
        {
            const std::size_t s = g.at(i, NY - 2, k);
            v[c] = v[s]; u[c] = u[s]; w[c] = w[s]; cp_rp(c, s);
        }
        else
        {
            if (bcs == NIMBUS_BC_RADIATIVE)
            {
                if (radiate)
                {
                    float& vp = rad_s[std::size_t(i) * NZ + k];
                    const float vn = kw78_face(vp, vp - 0.0f - cstar, float(v[in]) - vp, dt_rad, inv_dy, true);
                    v[c] = S(vn);
                    vp = vn;
                }
            }
            else if (bcs == NIMBUS_BC_ZERO_GRADIENT) v[c] = v[in];
            else v[c] = S(0);
            if (bcs == NIMBUS_BC_RIGID) { u[c] = S(0); w[c] = S(0); }
            else { u[c] = u[in]; w[c] = w[in]; }
            cp_rp(c, in);
        }
    }
    {
        const std::size_t gj = g.at(i, NY - 1, k), a = g.at(i, NY - 2, k), am = g.at(i, NY - 3, k);
        if (bcn == NIMBUS_BC_PERIODIC)
        {
            const std::size_t s = g.at(i, 1, k);
            v[gj] = v[s]; u[gj] = u[s]; w[gj] = w[s]; cp_rp(gj, s);
        }
        else
        {
            if (bcn == NIMBUS_BC_RIGID || bcn == NIMBUS_BC_RIGID_SLIP) v[a] = S(0);
            if (bcn == NIMBUS_BC_RADIATIVE && radiate)
            {
                float& vp = rad_n[std::size_t(i) * NZ + k];
                const float vn = kw78_face(vp, vp - 0.0f + cstar, vp - float(v[am]), dt_rad, inv_dy, false);
                v[a] = S(vn);
                vp = vn;
            }
            else if (bcn == NIMBUS_BC_ZERO_GRADIENT) v[a] = v[am];
            v[gj] = (bcn == NIMBUS_BC_RADIATIVE || bcn == NIMBUS_BC_ZERO_GRADIENT) ? v[a] : S(0);
            if (bcn == NIMBUS_BC_RIGID) { u[gj] = S(0); w[gj] = S(0); }
            else { u[gj] = u[a]; w[gj] = w[a]; }
            cp_rp(gj, a);
        }
    }
}

template <class S>
__global__ __launch_bounds__(kBlock) void k_bc_vertical(DevGrid g, S* u, S* v, S* w, int bottom, int top)
{
    int i, j;
    if (!column_index(g, i, j)) return;
    const int NZ = g.nz;
    if (bottom == NIMBUS_BOTTOM_NO_SLIP)
    {
        u[g.at(i, j, 0)] = S(0);
        v[g.at(i, j, 0)] = S(0);
    }
    const std::size_t t = g.at(i, j, NZ - 1), tm = g.at(i, j, NZ - 2);
    w[t] = top == NIMBUS_TOP_OPEN ? w[tm] : S(0);
    u[t] = u[tm];
    v[t] = v[tm];
}

template <class S>
__global__ __launch_bounds__(kBlock) void k_seed_radiation(DevGrid g, const S* u, const S* v, float* rad_w,
                                                           float* rad_e, float* rad_s, float* rad_n)
{
    const int t = blockIdx.x * kBlock + threadIdx.x;
    const int NZ = g.nz;
    if (t < g.ny * NZ)
    {
        const int j = t / NZ, k = t - j * NZ;
        rad_w[t] = float(u[g.at(0, j, k)]);
        rad_e[t] = float(u[g.at(g.nx - 2, j, k)]);
    }
    if (t < g.nx * NZ)
    {
        const int i = t / NZ, k = t - i * NZ;
        rad_s[t] = float(v[g.at(i, 0, k)]);
        rad_n[t] = float(v[g.at(i, g.ny - 2, k)]);
    }
}

// ---------------------------------------------------------------------
// WENO5 retained-flux transport with the PD donor limiter
// (advection_cartesian.cpp:1941-2650), all scalars of the step in one pass.
// ---------------------------------------------------------------------
template <class S>
struct ScalarSet
{
    const S* q[4];
    S* out[4];
    float* rl[4];
    int pd[4];
    int n;
};

// Face flux pass (advect_conservative_scalar_flux): each cell writes the raw
// retained-flux scalar flux through its own upper x, y and z faces, stored in
// the storage type like the reference's SFX/SFY/SFZ scratch.
template <class S>
struct FluxSet
{
    S* fx[4];
    S* fy[4];
    S* fz[4];
};

template <class S>
__global__ __launch_bounds__(kBlock) void k_transport_flux(DevGrid g, ScalarSet<S> set, FluxSet<S> fl,
                                                           const S* __restrict__ mfx, const S* __restrict__ mfy,
                                                           const S* __restrict__ mfz, int open_top)
{
    using A = typename AccT<S>::type;
    int i, j, k;
    if (!cell_index(g, i, j, k)) return;
    const int NX = g.nx, NY = g.ny, NZ = g.nz;
    const std::size_t c = g.at(i, j, k);
    const std::ptrdiff_t sy = g.px, sz = std::ptrdiff_t(g.px) * g.ny;
    const A mx = A(mfx[c]), my = A(mfy[c]), mz = A(mfz[c]);
    for (int s = 0; s < set.n; ++s)
    {
        const S* __restrict__ q = set.q[s];
        S fx = S(0), fy = S(0), fz = S(0);
        if (i <= NX - 2) fx = S(dev::face_flux<A>([=](int o) { return A(q[c + std::ptrdiff_t(o)]); }, i, NX, mx));
        if (j <= NY - 2) fy = S(dev::face_flux<A>([=](int o) { return A(q[c + std::ptrdiff_t(o) * sy]); }, j, NY, my));
        if (k <= NZ - 2) fz = S(dev::face_flux<A>([=](int o) { return A(q[c + std::ptrdiff_t(o) * sz]); }, k, NZ, mz));
        else if (open_top) fz = S(mz * A(q[c]));
        fl.fx[s][c] = fx;
        fl.fy[s][c] = fy;
        fl.fz[s][c] = fz;
    }
}

__device__ __forceinline__ float pd_store(double r)
{
    if (!(r > 0.0 && r < 1.0)) return float(r);
    const float shrink = 1.0f - 32.0f * FLT_EPSILON;
    return nextafterf(float(r) * shrink, 0.0f);
}

// Positive-definite donor limiter (advect_pd_limiter): r = min(1, (rho_begin q / dt) / outgoing).
template <class S>
__global__ __launch_bounds__(kBlock) void k_transport_limiter(DevGrid g, ScalarSet<S> set, FluxSet<S> fl,
                                                              const S* rbp, const double* base, double dt,
                                                              int open_top)
{
    int i, j, k;
    if (!cell_index(g, i, j, k)) return;
    const int NX = g.nx, NY = g.ny, NZ = g.nz;
    const std::size_t c = g.at(i, j, k);
    const std::size_t sy = std::size_t(g.px), sz = std::size_t(g.px) * g.ny;
    const double rho_b = slot(base, kBaseRho0, NZ)[k] + double(rbp[c]);
    for (int s = 0; s < set.n; ++s)
    {
        if (!set.pd[s]) continue;
        const S* fx = fl.fx[s];
        const S* fy = fl.fy[s];
        const S* fz = fl.fz[s];
        double out_rate = 0.0;
        if (i <= NX - 2 && fx[c] > S(0)) out_rate += double(fx[c]) * g.idx;
        if (i >= 1 && fx[c - 1] < S(0)) out_rate -= double(fx[c - 1]) * g.idx;
        if (j <= NY - 2 && fy[c] > S(0)) out_rate += double(fy[c]) * g.idy;
        if (j >= 1 && fy[c - sy] < S(0)) out_rate -= double(fy[c - sy]) * g.idy;
        if ((k <= NZ - 2 || (open_top && k == NZ - 1)) && fz[c] > S(0)) out_rate += double(fz[c]) * g.idz;
        if (k >= 1 && fz[c - sz] < S(0)) out_rate -= double(fz[c - sz]) * g.idz;
        double r = 1.0;
        if (out_rate > 0.0) r = fmin(1.0, rho_b * fmax(double(set.q[s][c]), 0.0) / dt / out_rate);
        set.rl[s][c] = pd_store(r);
    }
}

// Conservative update (advect_conservative_update): q = (rho_begin q - dt div F_lim) / rho_end,
// outgoing-then-incoming for positive-definite scalars.
template <class S>
__global__ __launch_bounds__(kBlock) void k_transport_update(DevGrid g, ScalarSet<S> set, FluxSet<S> fl,
                                                             const S* rbp, const S* rep, const double* base,
                                                             double dt, int open_top)
{
    int i, j, k;
    if (!cell_index(g, i, j, k)) return;
    const int NX = g.nx, NY = g.ny, NZ = g.nz;
    const std::size_t c = g.at(i, j, k);
    if (!(i >= 1 && i < NX - 1 && j >= 1 && j < NY - 1))
    {
        for (int s = 0; s < set.n; ++s) set.out[s][c] = set.q[s][c];
        return;
    }
    const std::size_t sy = std::size_t(g.px), sz = std::size_t(g.px) * g.ny;
    const double r0 = slot(base, kBaseRho0, NZ)[k];
    const double rho_b = r0 + double(rbp[c]);
    const double rho_e = r0 + double(rep[c]);
    const double den = rho_e > 0.05 ? rho_e : 0.05;
    const bool has_top = k <= NZ - 2 || (open_top && k == NZ - 1);
    for (int s = 0; s < set.n; ++s)
    {
        const S* fx = fl.fx[s];
        const S* fy = fl.fy[s];
        const S* fz = fl.fz[s];
        const float* rl = set.rl[s];
        const bool pd = set.pd[s] != 0;
        auto lim = [&](S raw, std::size_t lo, std::size_t hi) -> double {
            if (!pd || raw == S(0)) return double(raw);
            return double(raw) * double(S(raw > S(0) ? rl[lo] : rl[hi]));
        };
        const double sxp = lim(fx[c], c, c + 1);
        const double sxm = lim(fx[c - 1], c - 1, c);
        const double syp = lim(fy[c], c, c + sy);
        const double sym = lim(fy[c - sy], c - sy, c);
        const double szp = has_top ? lim(fz[c], c, k < NZ - 1 ? c + sz : c) : 0.0;
        const double szm = k >= 1 ? lim(fz[c - sz], c - sz, c) : 0.0;
        const double q0 = double(set.q[s][c]);
        double rho_phi;
        if (pd)
        {
            double outg = 0.0, inc = 0.0;
            const double ix = g.idx, iy = g.idy, iz = g.idz;
            if (sxp >= 0.0) outg += sxp * ix; else inc -= sxp * ix;
            if (sxm <= 0.0) outg -= sxm * ix; else inc += sxm * ix;
            if (syp >= 0.0) outg += syp * iy; else inc -= syp * iy;
            if (sym <= 0.0) outg -= sym * iy; else inc += sym * iy;
            if (has_top) { if (szp >= 0.0) outg += szp * iz; else inc -= szp * iz; }
            if (k >= 1) { if (szm <= 0.0) outg -= szm * iz; else inc += szm * iz; }
            rho_phi = (rho_b * q0 - dt * outg) + dt * inc;
        }
        else
        {
            const double div = (sxp - sxm) * g.idx + (syp - sym) * g.idy + (szp - szm) * g.idz;
            rho_phi = rho_b * q0 - dt * div;
        }
        set.out[s][c] = S(rho_phi / den);
    }
}

// ---------------------------------------------------------------------
// Kessler warm rain at one physical cell (kessler.cpp:103-531): saturation
// adjustment on copies booked as rates, finite-step autoconversion and
// accretion, ventilated rain evaporation bounded by the saturation deficit,
// then sedimentation of the pre-process rain (donor, then the receiver
// share from the level above), and the dT -> dtheta conversion.
// ---------------------------------------------------------------------
template <class S>
struct KesslerOut
{
    S dth, dqv, dqc, dqr;
};

template <class S>
__device__ KesslerOut<S> kessler_cell(const DevGrid& g, const KesslerConst& kc, double dt, double p_full,
                                      double rho_c, double rho_up, S theta, S qv, S qc, S qr, S qr_up, bool has_up)
{
    using A = typename AccT<S>::type;
    const A L = A(kLv / kCp);
    const A dtA = A(dt);
    const A p = A(p_full), rho = A(rho_c);
    KesslerOut<S> o{S(0), S(0), S(0), S(0)};
    const S Tf = S(dev::theta_to_temperature<A>(A(theta), p));
    A T = A(Tf), qva = A(qv), qca = A(qc);
    T = dev::saturation_adjust<A>(T, p, qva, qca);
    const S Ts = S(T), qvs = S(qva), qcs = S(qca);
    const A dqc_adj = A(qcs) - A(qc);
    if (dqc_adj != A(0))
    {
        const A rate = dqc_adj * A(1.0 / dt);
        o.dqc = S(o.dqc + S(rate));
        o.dqv = S(o.dqv - S(rate));
        o.dth = S(o.dth + S(L * rate));
    }
    S prod = S(0);
    if (qcs > S(0))
    {
        const A factor = A(1) / (A(1) + A(kc.c_accr) * dtA * pow(fmax(A(0), A(qr)), A(0.875)));
        const A amount = A(qcs) * (A(1) - factor) + factor * A(kc.c_auto) * dtA * fmax(A(qcs) - A(kc.qc0), A(0));
        prod = S(dev::clampv(amount, A(0), A(qcs)) / dtA);
        o.dqc = S(o.dqc - prod);
        o.dqr = S(o.dqr + prod);
    }
    if (qr > S(0) || prod > S(0))
    {
        const A qsat = dev::qvs_water<A>(A(Ts), p);
        if (qsat > A(0) && A(qvs) < qsat && rho > A(0) && p > A(0))
        {
            const A qr_av = A(qr) + A(prod) * dtA;
            const A rcgs = A(1.0e-3) * rho;
            const A rc = rcgs * fmax(qr_av, A(0));
            const A vent = (A(1.6) + A(124.9) * pow(rc, A(0.2046))) * pow(rc, A(0.525));
            const A resist = A(2.55e8) / (p * qsat) + A(5.4e5);
            const A raw = A(kc.evap_eff) * (vent / resist) * ((qsat - A(qvs)) / (rcgs * qsat));
            const A gsat = (A(kLv) * A(kLv) * qsat) / (A(kCp) * A(kRv) * fmax(A(Ts) * A(Ts), A(1)));
            const A bound = (qsat - A(qvs)) / (A(1) + gsat);
            const A ev = dev::clampv(raw, A(0), fmin(bound / dtA, qr_av / dtA));
            o.dqv = S(A(o.dqv) + ev);
            o.dqr = S(A(o.dqr) - ev);
            o.dth = S(A(o.dth) - L * ev);
        }
    }
    const A dzA = A(g.dz);
    const A qr_own = fmax(A(0), A(qr));
    if (qr_own > A(0))
    {
        const A Vt = fmin(dev::kessler_rain_speed<A>(qr_own, rho) * sqrt(A(1.2) / rho), A(kc.vt_max));
        const S flux = S(Vt * qr_own);
        o.dqr = S(o.dqr - S(A(flux) / dzA));
    }
    if (has_up)
    {
        const A qa = fmax(A(0), A(qr_up));
        if (qa > A(0))
        {
            const A ra = A(rho_up);
            const A Vt = fmin(dev::kessler_rain_speed<A>(qa, ra) * sqrt(A(1.2) / ra), A(kc.vt_max));
            const S flux = S(Vt * qa);
            o.dqr = S(o.dqr + S(ra / rho * A(flux) / dzA));
        }
    }
    o.dth = S(dev::theta_tendency_from_temperature<A>(A(o.dth), p));
    return o;
}

template <class S>
__global__ __launch_bounds__(kBlock) void k_kessler_tendencies(DevGrid g, const S* th, const S* qv, const S* qc,
                                                               const S* qr, const S* rp, const S* pp,
                                                               const double* base, KesslerConst kc, double dt,
                                                               S* dth, S* dqv, S* dqc, S* dqr)
{
    int i, j, k;
    if (!cell_index(g, i, j, k)) return;
    const std::size_t c = g.at(i, j, k);
    const int NZ = g.nz;
    if (!(i >= 1 && i < g.nx - 1 && j >= 1 && j < g.ny - 1))
    {
        dth[c] = S(0); dqv[c] = S(0); dqc[c] = S(0); dqr[c] = S(0);
        return;
    }
    const double* rho0 = slot(base, kBaseRho0, NZ);
    const double* p0 = slot(base, kBaseP0, NZ);
    const std::size_t sz = std::size_t(g.px) * g.ny;
    const bool up = k + 1 < NZ;
    const KesslerOut<S> o = kessler_cell<S>(g, kc, dt, p0[k] + double(pp[c]), rho0[k] + double(rp[c]),
                                            up ? rho0[k + 1] + double(rp[c + sz]) : 0.0, th[c], qv[c], qc[c], qr[c],
                                            up ? qr[c + sz] : S(0), up);
    dth[c] = o.dth; dqv[c] = o.dqv; dqc[c] = o.dqc; dqr[c] = o.dqr;
}

// Fused microphysics commit: Kessler tendencies, bounded apply
// (microphysics_step.cpp:1090-1140), surface rain ledger
// (microphysics_step.cpp:1917-2024; kessler.cpp:533-561) and the EOS
// pressure projection of every cell (eos_apply.cpp:50-110).
template <class S>
__global__ __launch_bounds__(kBlock) void k_kessler_apply(DevGrid g, const S* th_in, const S* qv_in, const S* qc_in,
                                                          const S* qr_in, S* th, S* qv, S* qc, S* qr, const S* rp,
                                                          S* pp, const double* base, KesslerConst kc, double dt,
                                                          double* rain, Bounds bd, int do_eos)
{
    int i, j, k;
    if (!cell_index(g, i, j, k)) return;
    const std::size_t c = g.at(i, j, k);
    const int NZ = g.nz;
    const double* rho0 = slot(base, kBaseRho0, NZ);
    const double* p0 = slot(base, kBaseP0, NZ);
    const std::size_t sz = std::size_t(g.px) * g.ny;
    const double rho = rho0[k] + double(rp[c]);
    S thn = th_in[c], qvn = qv_in[c], qcn = qc_in[c], qrn = qr_in[c];
    if (i >= 1 && i < g.nx - 1 && j >= 1 && j < g.ny - 1)
    {
        const bool up = k + 1 < NZ;
        const KesslerOut<S> o = kessler_cell<S>(g, kc, dt, p0[k] + double(pp[c]), rho,
                                                up ? rho0[k + 1] + double(rp[c + sz]) : 0.0, thn, qvn, qcn, qrn,
                                                up ? qr_in[c + sz] : S(0), up);
        if (k == 0)
        {
            const double q0 = double(qrn);
            if (q0 > 0.0)
            {
                double Vt = fmin(double(dev::kessler_rain_speed<double>(q0, rho)) * sqrt(1.2 / rho), kc.vt_max);
                const double sink = double(S(S(Vt) * qrn)) / g.dz;
                rain[std::size_t(j) * g.nx + i] += sink * rho * g.dz * dt;
            }
        }
        const double step = fmin(fmax(double(o.dth) * dt, -50.0), 50.0);
        thn = S(clamp_store<S>(double(thn) + step, 300.0, 200.0f, 700.0f));
        qvn = S(clamp_store<S>(double(qvn) + double(o.dqv) * dt, 0.0, 0.0f, 0.05f));
        qcn = S(clamp_store<S>(double(qcn) + double(o.dqc) * dt, 0.0, 0.0f, 0.05f));
        qrn = S(clamp_store<S>(double(qrn) + double(o.dqr) * dt, 0.0, 0.0f, 0.05f));
    }
    th[c] = thn; qv[c] = qvn; qc[c] = qcn; qr[c] = qrn;
    const double thv = double(thn);
    if (do_eos && rho > 0.0 && thv > 0.0)
    {
        const double pd = dev::eos_pressure(rho, thv);
        if (isfinite(pd) && pd > 0.0) pp[c] = S(clamp_store<S>(pd, 101325.0, bd.p_min, bd.p_max) - p0[k]);
    }
}

template <class S>
__global__ __launch_bounds__(kBlock) void k_copy(std::size_t n, const S* a, S* b)
{
    const std::size_t t = std::size_t(blockIdx.x) * kBlock + threadIdx.x;
    if (t < n) b[t] = a[t];
}

// ---------------------------------------------------------------------
// Deterministic diagnostics: fixed-shape block partials, then one block
// reduces them in a fixed order (so a member's numbers do not depend on
// what else runs on the device).
// ---------------------------------------------------------------------
template <class S>
__global__ __launch_bounds__(kBlock) void k_diag_partials(DevGrid g, const S* w, const S* th, const S* qc,
                                                          const S* qr, const double* rain, const double* base,
                                                          double* partials)
{
    __shared__ double sm[kDiagValues][kBlock];
    double v[kDiagValues] = {-1e300, 1e300, 0.0, 0.0, 0.0, 0.0, -1e300, 0.0};
    const std::size_t ncell = std::size_t(g.nx) * g.ny * g.nz;
    const std::size_t stride = std::size_t(gridDim.x) * kBlock;
    for (std::size_t t = std::size_t(blockIdx.x) * kBlock + threadIdx.x; t < ncell; t += stride)
    {
        const int k = int(t / (std::size_t(g.nx) * g.ny));
        const int rem = int(t - std::size_t(k) * g.nx * g.ny);
        const int j = rem / g.nx, i = rem - j * g.nx;
        if (!(i >= 1 && i < g.nx - 1 && j >= 1 && j < g.ny - 1)) continue;
        const std::size_t c = g.at(i, j, k);
        const double wv = double(w[c]);
        v[0] = fmax(v[0], wv);
        v[1] = fmin(v[1], wv);
        const double tp = double(th[c]) - slot(base, kBaseTh0, g.nz)[k];
        v[2] += tp * tp;
        v[4] += double(qc[c]);
        v[5] += double(qr[c]);
        if (k == 0)
        {
            const double r = rain[std::size_t(j) * g.nx + i];
            v[3] += r;
            v[6] = fmax(v[6], r);
        }
    }
    for (int q = 0; q < kDiagValues; ++q) sm[q][threadIdx.x] = v[q];
    __syncthreads();
    for (int s = kBlock / 2; s > 0; s >>= 1)
    {
        if (int(threadIdx.x) < s)
        {
            sm[0][threadIdx.x] = fmax(sm[0][threadIdx.x], sm[0][threadIdx.x + s]);
            sm[1][threadIdx.x] = fmin(sm[1][threadIdx.x], sm[1][threadIdx.x + s]);
            sm[6][threadIdx.x] = fmax(sm[6][threadIdx.x], sm[6][threadIdx.x + s]);
            sm[2][threadIdx.x] += sm[2][threadIdx.x + s];
            sm[3][threadIdx.x] += sm[3][threadIdx.x + s];
            sm[4][threadIdx.x] += sm[4][threadIdx.x + s];
            sm[5][threadIdx.x] += sm[5][threadIdx.x + s];
            sm[7][threadIdx.x] += sm[7][threadIdx.x + s];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0)
        for (int q = 0; q < kDiagValues; ++q) partials[std::size_t(blockIdx.x) * kDiagValues + q] = sm[q][0];
}

__global__ void k_diag_final(const double* partials, int n, double* out)
{
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    double v[kDiagValues] = {-1e300, 1e300, 0.0, 0.0, 0.0, 0.0, -1e300, 0.0};
    for (int b = 0; b < n; ++b)
    {
        const double* p = partials + std::size_t(b) * kDiagValues;
        v[0] = fmax(v[0], p[0]);
        v[1] = fmin(v[1], p[1]);
        v[6] = fmax(v[6], p[6]);
        v[2] += p[2];
        v[3] += p[3];
        v[4] += p[4];
        v[5] += p[5];
        v[7] += p[7];
    }
    for (int q = 0; q < kDiagValues; ++q) out[q] = v[q];
}

// Host layout [i][j][k] <-> device [k][j][i], subtracting / adding a base slot.
template <class H, class S>
__global__ __launch_bounds__(kBlock) void k_host_to_dev(DevGrid g, const H* src, S* dst, const double* base_slot)
{
    int i, j, k;
    if (!cell_index(g, i, j, k)) return;
    const std::size_t h = (std::size_t(i) * g.ny + j) * g.nz + k;
    dst[g.at(i, j, k)] = S(double(src[h]) - (base_slot ? base_slot[k] : 0.0));
}

template <class S, class H>
__global__ __launch_bounds__(kBlock) void k_dev_to_host(DevGrid g, const S* src, H* dst, const double* base_slot)
{
    int i, j, k;
    if (!cell_index(g, i, j, k)) return;
    const std::size_t h = (std::size_t(i) * g.ny + j) * g.nz + k;
    dst[h] = H(double(src[g.at(i, j, k)]) + (base_slot ? base_slot[k] : 0.0));
}

dim3 cell_grid(const DevGrid& g) { return dim3(unsigned(g.tiles_plane * g.nz)); }
dim3 column_grid(const DevGrid& g) { return dim3(unsigned(g.tiles_plane)); }
dim3 linear_grid(std::size_t n) { return dim3(unsigned((n + kBlock - 1) / kBlock)); }

template <class S>
S* P(void* p) { return static_cast<S*>(p); }

}  // namespace

hipEvent_t Launcher::take_event()
{
    if (event_pool && !event_pool->empty())
    {
        hipEvent_t e = event_pool->back();
        event_pool->pop_back();
        return e;
    }
    hipEvent_t e = nullptr;
    if (hipEventCreate(&e) != hipSuccess) error = hipErrorOutOfMemory;
    return e;
}

void Launcher::begin(const char* name, hipEvent_t& a)
{
    (void)name;
    a = nullptr;
    if (!timing) return;
    a = take_event();
    if (a) hipEventRecord(a, stream);
}

void Launcher::end(const char* name, hipEvent_t a)
{
    if (launches) ++*launches;
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess && error == hipSuccess) error = e;
    if (!timing || !a) return;
    hipEvent_t b = take_event();
    if (!b) return;
    hipEventRecord(b, stream);
    if (pending) pending->push_back(TimedLaunch{name, a, b});
}

#define NIMBUS_LAUNCH(L, name, kernel, grid, ...)                               \
    do                                                                          \
    {                                                                           \
        hipEvent_t ev_;                                                         \
        (L).begin(name, ev_);                                                   \
        hipLaunchKernelGGL(kernel, grid, dim3(kBlock), 0, (L).stream, __VA_ARGS__); \
        (L).end(name, ev_);                                                     \
    } while (0)

template <class S>
void launch_slow_rk2(Launcher& L, const DevGrid& g, const DevFields& f, const StepParams& p)
{
    NIMBUS_LAUNCH(L, "slow_stage1", (k_slow_stage1<S>), cell_grid(g), g, P<S>(f.u), P<S>(f.v), P<S>(f.w),
                  P<S>(f.rp), P<S>(f.qv), P<S>(f.qc), P<S>(f.qr), (const double*)f.base, p.coriolis_f, p.dt,
                  P<S>(f.k1u), P<S>(f.k1v), P<S>(f.k1w), P<S>(f.su), P<S>(f.sv), P<S>(f.sw));
    NIMBUS_LAUNCH(L, "slow_stage2", (k_slow_stage2<S>), cell_grid(g), g, P<S>(f.u), P<S>(f.v), P<S>(f.w),
                  P<S>(f.rp), P<S>(f.qv), P<S>(f.qc), P<S>(f.qr), (const double*)f.base, p.coriolis_f, p.dt,
                  P<S>(f.k1u), P<S>(f.k1v), P<S>(f.k1w), P<S>(f.su), P<S>(f.sv), P<S>(f.sw));
}

template <class S>
void launch_slow_tendencies(Launcher& L, const DevGrid& g, const DevFields& f, const StepParams& p, void* du,
                            void* dv, void* dw)
{
    NIMBUS_LAUNCH(L, "slow_tendencies", (k_slow_tendencies<S>), cell_grid(g), g, P<S>(f.u), P<S>(f.v), P<S>(f.w),
                  P<S>(f.rp), P<S>(f.qv), P<S>(f.qc), P<S>(f.qr), (const double*)f.base, p.coriolis_f, P<S>(du),
                  P<S>(dv), P<S>(dw));
}

template <class S>
void launch_boundaries(Launcher& L, const DevGrid& g, const DevFields& f, const StepParams& p, double dt_rad)
{
    const float dtf = float(dt_rad);
    NIMBUS_LAUNCH(L, "bc_west_east", (k_bc_west_east<S>), linear_grid(std::size_t(g.ny) * g.nz), g, P<S>(f.u),
                  P<S>(f.v), P<S>(f.w), P<S>(f.rp), P<S>(f.pp), f.rad_w, f.rad_e, p.bc_west, p.bc_east, dtf,
                  1.0f / float(g.dx));
    NIMBUS_LAUNCH(L, "bc_south_north", (k_bc_south_north<S>), linear_grid(std::size_t(g.nx) * g.nz), g, P<S>(f.u),
                  P<S>(f.v), P<S>(f.w), P<S>(f.rp), P<S>(f.pp), f.rad_s, f.rad_n, p.bc_south, p.bc_north, dtf,
                  1.0f / float(g.dy));
    NIMBUS_LAUNCH(L, "bc_vertical", (k_bc_vertical<S>), column_grid(g), g, P<S>(f.u), P<S>(f.v), P<S>(f.w),
                  p.bc_bottom, p.bc_top);
}

template <class S>
void launch_seed_radiation(Launcher& L, const DevGrid& g, const DevFields& f)
{
    NIMBUS_LAUNCH(L, "seed_radiation", (k_seed_radiation<S>), linear_grid(std::size_t(std::max(g.nx, g.ny)) * g.nz),
                  g, P<S>(f.u), P<S>(f.v), f.rad_w, f.rad_e, f.rad_s, f.rad_n);
}

template <class S>
void launch_acoustic_substep(Launcher& L, const DevGrid& g, const DevFields& f, const StepParams& p, double dts,
                             double weight, bool first, bool accumulate)
{
    if (p.acoustic_mode == NIMBUS_ACOUSTIC_HEVI)
    {
        const double alpha = 0.5 * (1.0 + p.hevi_beta);
        NIMBUS_LAUNCH(L, "acoustic_hevi_column", (k_hevi_column<S>), column_grid(g), g, P<S>(f.u), P<S>(f.v),
                      P<S>(f.w), P<S>(f.rp), P<S>(f.ralt), P<S>(f.pp), P<S>(f.th), P<S>(f.mfx), P<S>(f.mfy),
                      P<S>(f.mfz), P<S>(f.rb), f.tri_c, f.tri_r, (const double*)f.base, p.bounds, dts, alpha, weight,
                      int(first), int(accumulate));
        NIMBUS_LAUNCH(L, "acoustic_backward", (k_acoustic_backward<S>), cell_grid(g), g, P<S>(f.u), P<S>(f.v),
                      P<S>(f.w), P<S>(f.ralt), P<S>(f.rp), P<S>(f.pp), (const double*)f.base, p.bounds, dts, 1);
        return;
    }
    NIMBUS_LAUNCH(L, "acoustic_forward", (k_acoustic_forward<S>), cell_grid(g), g, P<S>(f.u), P<S>(f.v), P<S>(f.w),
                  P<S>(f.rp), P<S>(f.ralt), P<S>(f.pp), P<S>(f.th), P<S>(f.mfx), P<S>(f.mfy), P<S>(f.mfz),
                  P<S>(f.rb), (const double*)f.base, p.bounds, dts, weight, int(first), int(accumulate),
                  int(p.bc_top == NIMBUS_TOP_OPEN));
    NIMBUS_LAUNCH(L, "acoustic_backward", (k_acoustic_backward<S>), cell_grid(g), g, P<S>(f.u), P<S>(f.v),
                  P<S>(f.w), P<S>(f.ralt), P<S>(f.rp), P<S>(f.pp), (const double*)f.base, p.bounds, dts, 0);
}

template <class S>
void launch_transport(Launcher& L, const DevGrid& g, const DevFields& f, const StepParams& p, int n_scalars,
                      void* const* q_in, void* const* q_out, const int* pd)
{
    ScalarSet<S> set{};
    FluxSet<S> fl{};
    float* rls[3] = {f.rl_qv, f.rl_qc, f.rl_qr};
    int npd = 0;
    set.n = n_scalars;
    for (int s = 0; s < n_scalars; ++s)
    {
        set.q[s] = P<S>(q_in[s]);
        set.out[s] = P<S>(q_out[s]);
        set.pd[s] = pd[s];
        set.rl[s] = pd[s] ? rls[npd++] : nullptr;
        fl.fx[s] = P<S>(f.flux[3 * s]);
        fl.fy[s] = P<S>(f.flux[3 * s + 1]);
        fl.fz[s] = P<S>(f.flux[3 * s + 2]);
    }
    const int open_top = int(p.bc_top == NIMBUS_TOP_OPEN);
    NIMBUS_LAUNCH(L, "transport_flux", (k_transport_flux<S>), cell_grid(g), g, set, fl, P<S>(f.mfx), P<S>(f.mfy),
                  P<S>(f.mfz), open_top);
    if (npd > 0)
        NIMBUS_LAUNCH(L, "transport_limiter", (k_transport_limiter<S>), cell_grid(g), g, set, fl, P<S>(f.rb),
                      (const double*)f.base, p.dt, open_top);
    NIMBUS_LAUNCH(L, "transport_update", (k_transport_update<S>), cell_grid(g), g, set, fl, P<S>(f.rb), P<S>(f.rp),
                  (const double*)f.base, p.dt, open_top);
}

template <class S>
void launch_kessler_tendencies(Launcher& L, const DevGrid& g, const DevFields& f, const StepParams& p, double dt,
                               void* dth, void* dqv, void* dqc, void* dqr)
{
    NIMBUS_LAUNCH(L, "kessler_tendencies", (k_kessler_tendencies<S>), cell_grid(g), g, P<S>(f.th), P<S>(f.qv),
                  P<S>(f.qc), P<S>(f.qr), P<S>(f.rp), P<S>(f.pp), (const double*)f.base, p.kessler, dt, P<S>(dth),
                  P<S>(dqv), P<S>(dqc), P<S>(dqr));
}

template <class S>
void launch_large_step(Launcher& L, const DevGrid& g, const DevFields& f, const StepParams& p)
{
    const int N = p.substeps;
    const double dts = p.dt / N;
    const double weight = dts / p.dt;
    launch_slow_rk2<S>(L, g, f, p);
    launch_boundaries<S>(L, g, f, p, 0.0);
    launch_seed_radiation<S>(L, g, f);
    for (int n = 0; n < N; ++n)
    {
        launch_acoustic_substep<S>(L, g, f, p, dts, weight, n == 0, true);
        launch_boundaries<S>(L, g, f, p, dts);
    }
    void* qin[4] = {f.th, f.qv, f.qc, f.qr};
    void* qout[4] = {f.tth, f.tqv, f.tqc, f.tqr};
    const int pd[4] = {0, 1, 1, 1};
    launch_transport<S>(L, g, f, p, 4, qin, qout, pd);
    // sedimentation CFL bound 0.8 dz / Vt_max (microphysics_step.cpp:3501-3560), state independent
    const double limit = 0.8 * g.dz / p.kessler.vt_max;
    const int nsub = p.dt <= limit ? 1 : int(ceil(p.dt / limit));
    const double h = p.dt / nsub;
    for (int n = 0; n < nsub; ++n)
    {
        const bool from_temp = (n % 2) == 0;
        void* const* src = from_temp ? qout : qin;
        void* const* dst = from_temp ? qin : qout;
        NIMBUS_LAUNCH(L, "kessler_apply", (k_kessler_apply<S>), cell_grid(g), g, P<S>(src[0]), P<S>(src[1]),
                      P<S>(src[2]), P<S>(src[3]), P<S>(dst[0]), P<S>(dst[1]), P<S>(dst[2]), P<S>(dst[3]), P<S>(f.rp),
                      P<S>(f.pp), (const double*)f.base, p.kessler, h, f.rain, p.bounds, int(n == nsub - 1));
    }
    if (nsub % 2 == 0)
        for (int s = 0; s < 4; ++s)
            NIMBUS_LAUNCH(L, "copy", (k_copy<S>), linear_grid(g.cells()), g.cells(), P<S>(qout[s]), P<S>(qin[s]));
    launch_boundaries<S>(L, g, f, p, 0.0);
}

template <class S>
void launch_diagnostics(Launcher& L, const DevGrid& g, const DevFields& f, double* partials, int n_partials,
                        double* out)
{
    NIMBUS_LAUNCH(L, "diag_partials", (k_diag_partials<S>), dim3(n_partials), g, P<S>(f.w), P<S>(f.th), P<S>(f.qc),
                  P<S>(f.qr), (const double*)f.rain, (const double*)f.base, partials);
    hipEvent_t ev;
    L.begin("diag_final", ev);
    hipLaunchKernelGGL(k_diag_final, dim3(1), dim3(64), 0, L.stream, (const double*)partials, n_partials, out);
    L.end("diag_final", ev);
}

template <class H, class S>
void launch_host_to_dev(Launcher& L, const DevGrid& g, const void* src, void* dst, const double* base_slot)
{
    NIMBUS_LAUNCH(L, "host_to_dev", (k_host_to_dev<H, S>), cell_grid(g), g, static_cast<const H*>(src), P<S>(dst),
                  base_slot);
}

template <class S, class H>
void launch_dev_to_host(Launcher& L, const DevGrid& g, const void* src, void* dst, const double* base_slot)
{
    NIMBUS_LAUNCH(L, "dev_to_host", (k_dev_to_host<S, H>), cell_grid(g), g, static_cast<const S*>(src),
                  static_cast<H*>(dst), base_slot);
}

#define NIMBUS_INSTANTIATE(S)                                                                                        \
    template void launch_large_step<S>(Launcher&, const DevGrid&, const DevFields&, const StepParams&);             \
    template void launch_slow_rk2<S>(Launcher&, const DevGrid&, const DevFields&, const StepParams&);               \
    template void launch_slow_tendencies<S>(Launcher&, const DevGrid&, const DevFields&, const StepParams&, void*,   \
                                            void*, void*);                                                           \
    template void launch_acoustic_substep<S>(Launcher&, const DevGrid&, const DevFields&, const StepParams&, double, \
                                             double, bool, bool);                                                    \
    template void launch_boundaries<S>(Launcher&, const DevGrid&, const DevFields&, const StepParams&, double);     \
    template void launch_seed_radiation<S>(Launcher&, const DevGrid&, const DevFields&);                            \
    template void launch_transport<S>(Launcher&, const DevGrid&, const DevFields&, const StepParams&, int,          \
                                      void* const*, void* const*, const int*);                                       \
    template void launch_kessler_tendencies<S>(Launcher&, const DevGrid&, const DevFields&, const StepParams&,      \
                                               double, void*, void*, void*, void*);                                  \
    template void launch_diagnostics<S>(Launcher&, const DevGrid&, const DevFields&, double*, int, double*);        \
    template void launch_host_to_dev<float, S>(Launcher&, const DevGrid&, const void*, void*, const double*);       \
    template void launch_host_to_dev<double, S>(Launcher&, const DevGrid&, const void*, void*, const double*);      \
    template void launch_dev_to_host<S, float>(Launcher&, const DevGrid&, const void*, void*, const double*);       \
    template void launch_dev_to_host<S, double>(Launcher&, const DevGrid&, const void*, void*, const double*);

NIMBUS_INSTANTIATE(float)
NIMBUS_INSTANTIATE(double)

}  // namespace nimbus
