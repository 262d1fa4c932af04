// CDNA4 (gfx950) kernels of the Klemp-Wilhelmson split-explicit large step.
//
// Device layout: every 3-D field is [k][j][i] with i contiguous and a
// 64-byte-padded row pitch, so the 64 lanes of a wavefront read consecutive
// i of one (j, k) row and vertical neighbours are one plane apart. rho and p
// are stored as perturbations from the FP64 base state; theta, winds and
// water species are stored whole.
//
// Launch shapes. Every launch covers all members of a batch: workgroup b
// reads entry b of a block map {member, local work item}. Stencil and
// pointwise kernels give each thread one cell and each 256-thread workgroup
// a tile of 64 x 4 (or 32 x 8) columns at one level; the work items of a
// column tile are numbered level-fastest within vertical chunks, so the
// workgroups an XCD runs back to back are vertical neighbours and a
// stencil's planes above and below are still in that XCD's L2 while every
// thread stays independent (full occupancy, no serial march). The HE-VI
// column solve gives each thread one whole column.
//
// Precision: S is the storage type. Continuity, EOS, base-state terms,
// mass-flux accumulation and the HE-VI column system with its Thomas solve
// compute in FP64 for both storage types. The slow tendencies (UP5
// advection, buoyancy, Coriolis), the acoustic pressure-gradient tendency,
// WENO5 transport, Kessler microphysics and the SGS operators compute in the
// storage type's arithmetic (FP32 bulk in the mixed mode).
#include "device_math.hpp"
#include "kernels.hpp"

#include <algorithm>

namespace nimbus
{

namespace
{

constexpr int kBlock = 256;

template <class S>
struct AccT
{
    using type = S;
};

// Device arrays are addressed as global memory (address space 1). The
// pointers live in the member argument table, so without the explicit
// address space the compiler emits flat loads and stores, which also count
// against the LDS/scalar counter and force full waits inside pipelined loops.
template <class T>
using gptr = __attribute__((address_space(1))) T*;

template <class T>
__device__ __forceinline__ gptr<T> G(T* p)
{
    return (gptr<T>)(p);
}

template <class S>
__device__ __forceinline__ gptr<S> P(void* p)
{
    return (gptr<S>)(static_cast<S*>(p));
}

__device__ __forceinline__ double finite0(double x) { return isfinite(x) ? x : 0.0; }

// 1/x in FP64: the hardware reciprocal estimate refined by two Newton-Raphson
// steps r <- r + r (1 - x r), each of which squares the relative error, so the
// result is within about an ulp of the quotient for finite nonzero x of
// normal range (zero and non-finite x give a non-finite result).
__device__ __forceinline__ double rcp_nr(double x)
{
    double r = __builtin_amdgcn_rcp(x);
    r = fma(r, fma(-x, r, 1.0), r);
    return fma(r, fma(-x, r, 1.0), r);
}
__device__ __forceinline__ float finite0(float x) { return isfinite(x) ? x : 0.0f; }
__device__ __forceinline__ double safe_rho(double r) { return isfinite(r) && r > 1.0e-6 ? r : 1.0; }
__device__ __forceinline__ float safe_rho(float r) { return isfinite(r) && r > 1.0e-6f ? r : 1.0f; }

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

// Clamp in FP64 without rounding the whole value to the storage type, so the
// stored perturbation keeps its own precision.
__device__ __forceinline__ double clamp_full(double x, double fallback, float lo, float hi)
{
    if (!isfinite(x)) x = fallback;
    return fmin(fmax(x, double(lo)), double(hi));
}

__device__ __forceinline__ gptr<const double> slot(const double* base, int s, int nz)
{
    return G(base) + std::size_t(s) * nz;
}

// One field as a raw buffer: the base address in a scalar buffer resource
// and 32-bit byte offsets, the lane-dependent part (position in a plane) in
// a VGPR and the plane part in an SGPR. A stencil's loads then cost no
// 64-bit address arithmetic per operand. Offsets stay below 2^31: a member
// field holds fewer than 2^28 padded cells (runtime grid check).
template <class S>
struct BufField
{
    __amdgpu_buffer_rsrc_t r;
    __device__ __forceinline__ S operator()(int voff, int soff) const
    {
        if constexpr (sizeof(S) == sizeof(float))
            return __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(r, voff, soff, 0));
        else
            return __builtin_bit_cast(double, __builtin_amdgcn_raw_buffer_load_b64(r, voff, soff, 0));
    }
};

// Buffer resource word 3 for gfx9-family parts: 32-bit elements, raw (stride 0) addressing.
constexpr int kBufferWord3 = 0x00020000;

template <class S>
__device__ __forceinline__ BufField<S> buf(const void* p, std::size_t bytes)
{
    return BufField<S>{__builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(p), 0, int(bytes), kBufferWord3)};
}

// ---------------------------------------------------------------------
// Work decoding
// ---------------------------------------------------------------------
struct Work
{
    int m, local;
};

__device__ __forceinline__ Work work_of(const int2* __restrict__ map)
{
    const int2 e = map[blockIdx.x];
    return Work{e.x, e.y};
}

// One thread per (i, j) column: of a tile_w x tile_h tile of columns, or of
// a run of kBlock consecutive columns in the linear geometry.
struct Column
{
    int i, j;
};

__device__ __forceinline__ Column column_of(const MemberArgs& A, int local)
{
    const MemberGeom& geo = A.geo;
    if (geo.linear)
    {
        // run `local` of kBlock consecutive columns in row-major order: a
        // level's loads of a workgroup are one or two contiguous segments
        const unsigned col = unsigned(local) * unsigned(kBlock) + threadIdx.x;
        const unsigned nx = unsigned(A.g.nx);
        const unsigned j = col / nx;
        return Column{int(col - j * nx), int(j)};
    }
    const int ty = local / geo.tiles_x, tx = local - ty * geo.tiles_x;
    const int tid = int(threadIdx.x);
    return Column{tx * geo.tile_w + (tid & (geo.tile_w - 1)), ty * geo.tile_h + tid / geo.tile_w};
}

// One cell per thread. Column tiles (tile_cell_of): tile t of the member at
// level k; the local work index runs k fastest within a vertical chunk
// (k_chunks chunks of kz levels), then over the column tiles, then over the
// chunks, so the workgroups an XCD runs back to back are the vertical
// neighbours of one column tile and a stencil's k-1 .. k+3 planes are still
// in its L2. Linear runs (cell_of with MemberGeom::linear): 256 consecutive
// elements of one level's row-major plane, levels in order, so a launch
// streams through memory; the slow tendencies and the WENO5 face fluxes and
// update (+-3 stencils in every direction) keep the column tiles, which serve
// their j and k neighbours from L2.
struct Cell
{
    int i, j, k;
    bool valid;
};

__device__ __forceinline__ Cell tile_cell_of(const MemberArgs& A, int local)
{
    const MemberGeom& geo = A.geo;
    const int nz = A.g.nz;
    const int kz = (nz + geo.k_chunks - 1) / geo.k_chunks;
    const int tiles = geo.tiles_x * geo.tiles_y;
    const int kk = local % kz, rest = local / kz;
    const int t = rest % tiles, kc = rest / tiles;
    const int ty = t / geo.tiles_x, tx = t - ty * geo.tiles_x;
    const int tid = int(threadIdx.x);
    Cell c;
    c.k = kc * kz + kk;
    c.i = tx * geo.tile_w + (tid & (geo.tile_w - 1));
    c.j = ty * geo.tile_h + tid / geo.tile_w;
    c.valid = c.k < nz && c.i < A.g.nx && c.j < A.g.ny;
    return c;
}

__device__ __forceinline__ Cell cell_of(const MemberArgs& A, int local)
{
    const MemberGeom& geo = A.geo;
    const int nz = A.g.nz;
    if (geo.linear)
    {
        // run r of kBlock consecutive elements of level k's padded plane: the
        // level stays uniform over the workgroup, so per-level data stays
        // scalar
        const int k = local / geo.runs;
        const unsigned idx = unsigned(local - k * geo.runs) * unsigned(kBlock) + threadIdx.x;
        const unsigned px = unsigned(A.g.px);
        const unsigned row = idx / px;
        Cell c;
        c.i = int(idx - row * px);
        c.j = int(row);
        c.k = k;
        c.valid = k < nz && c.j < A.g.ny && c.i < A.g.nx;
        return c;
    }
    return tile_cell_of(A, local);
}

__device__ __forceinline__ int clampi(int x, int lo, int hi) { return x < lo ? lo : (x > hi ? hi : x); }

// Keeps a loaded value where it is: the compiler may otherwise sink a load
// into the branch that uses it, which puts a full memory wait inside that
// branch. Pinning every value of a stencil after its loads are issued means
// one wait for all of them.
template <class T>
__device__ __forceinline__ void pin(T& x)
{
    asm volatile("" : "+v"(x));
}

// ---------------------------------------------------------------------
// Slow tendencies (cartesian_cgrid_3d.cpp:1064-1356) at one cell, in the
// arithmetic type T of the storage (FP32 for the mixed mode, with buoyancy
// in perturbation form; FP64 for FP64 storage). Every input is loaded first
// (SlowStencil, indices clamped to the domain), so a cell's loads are all in
// flight together and the UP5/UP3/UP1 choice near the walls only selects
// among registers.
// ---------------------------------------------------------------------
template <class T>
struct SlowStencil
{
    // 7-point lines through the cell, index 3 = the cell: along x, y and z
    T ux[7], uy[7], uz[7], vx[7], vy[7], vz[7], wx[7], wy[7], wz[7];
    // cross-face neighbours for the interpolated advecting velocities
    T v_xp_ym, w_xp_zm, u_xm_yp, w_yp_zm, u_xm_zp, v_ym_zp;
    T rp0, rp1, qv0, qv1, l0, l1;  // cell and the one above
    // base state: v0, u0 at k; rho0 and the stored qv0 averaged to the face above
    T v0, u0, rho0f, qv0f;
};

template <class S, class T>
__device__ __forceinline__ void load_slow_stencil(const DevGrid& g, gptr<const S> u, gptr<const S> v,
                                                  gptr<const S> w, gptr<const S> rp, gptr<const S> qv,
                                                  gptr<const S> qc, gptr<const S> qr, const double* base, int i,
                                                  int j, int k, SlowStencil<T>& L)
{
    const int NX = g.nx, NY = g.ny, NZ = g.nz;
    const std::size_t c = g.at(i, j, k);
    const std::ptrdiff_t sy = g.px, sz = std::ptrdiff_t(g.plane());
#pragma unroll
    for (int o = 0; o < 7; ++o)
    {
        const std::ptrdiff_t ox = clampi(i + o - 3, 0, NX - 1) - i;
        const std::ptrdiff_t oy = std::ptrdiff_t(clampi(j + o - 3, 0, NY - 1) - j) * sy;
        const std::ptrdiff_t oz = std::ptrdiff_t(clampi(k + o - 3, 0, NZ - 1) - k) * sz;
        L.ux[o] = T(u[c + ox]); L.vx[o] = T(v[c + ox]); L.wx[o] = T(w[c + ox]);
        L.uy[o] = T(u[c + oy]); L.vy[o] = T(v[c + oy]); L.wy[o] = T(w[c + oy]);
        L.uz[o] = T(u[c + oz]); L.vz[o] = T(v[c + oz]); L.wz[o] = T(w[c + oz]);
    }
    const std::ptrdiff_t xm = i >= 1 ? -1 : 0, xp = i < NX - 1 ? 1 : 0;
    const std::ptrdiff_t ym = j >= 1 ? -sy : 0, yp = j < NY - 1 ? sy : 0;
    const std::ptrdiff_t zm = k >= 1 ? -sz : 0, zp = k < NZ - 1 ? sz : 0;
    L.v_xp_ym = T(v[c + xp + ym]);
    L.w_xp_zm = T(w[c + xp + zm]);
    L.u_xm_yp = T(u[c + xm + yp]);
    L.w_yp_zm = T(w[c + yp + zm]);
    L.u_xm_zp = T(u[c + xm + zp]);
    L.v_ym_zp = T(v[c + ym + zp]);
    L.rp0 = T(rp[c]); L.rp1 = T(rp[c + zp]);
    L.qv0 = T(qv[c]); L.qv1 = T(qv[c + zp]);
    L.l0 = T(qc[c]) + T(qr[c]);
    L.l1 = T(qc[c + zp]) + T(qr[c + zp]);
    const int k1 = k < NZ - 1 ? k + 1 : k;
    const auto rho0 = slot(base, kBaseRho0, NZ);
    const auto qv0s = slot(base, kBaseQv0Stored, NZ);
    L.v0 = T(slot(base, kBaseV0, NZ)[k]);
    L.u0 = T(slot(base, kBaseU0, NZ)[k]);
    L.rho0f = T(0.5 * (rho0[k] + rho0[k1]));
    L.qv0f = T(0.5 * (qv0s[k] + qv0s[k1]));
#pragma unroll
    for (int o = 0; o < 7; ++o)
    {
        pin(L.ux[o]); pin(L.uy[o]); pin(L.uz[o]); pin(L.vx[o]); pin(L.vy[o]); pin(L.vz[o]);
        pin(L.wx[o]); pin(L.wy[o]); pin(L.wz[o]);
    }
}

template <class T>
__device__ __forceinline__ T up7(const T* s, int pos, int n, T a, T inv)
{
    return dev::upwind_axis<T>([&](int o) { return s[3 + o]; }, pos, n, a, inv);
}

// Tendencies of an interior cell (1 <= i <= NX-2, 1 <= j <= NY-2, k <= NZ-2).
template <class T>
__device__ __forceinline__ void slow_tendency_cell(const DevGrid& g, const SlowStencil<T>& L, double fcor_d,
                                                   int i, int j, int k, T& du, T& dv, T& dw)
{
    const int NX = g.nx, NY = g.ny, NZ = g.nz;
    const T ix = T(g.idx), iy = T(g.idy), iz = T(g.idz), fcor = T(fcor_d);
    const T q = T(0.25), h = T(0.5);
    // u at x-face
    {
        const T ux = L.ux[3];
        const T vy = q * (L.vy[2] + L.vx[3] + L.v_xp_ym + L.vx[4]);
        const T wz = k == 0 ? q * (L.wz[3] + L.wx[4]) : q * (L.wz[2] + L.wz[3] + L.w_xp_zm + L.wx[4]);
        const T ddx = up7(L.ux, i, NX, ux, ix);
        const T ddy = up7(L.uy, j, NY, vy, iy);
        const T ddz = k == 0 ? (wz >= T(0) ? T(0) : (L.uz[4] - L.uz[3]) * iz) : up7(L.uz, k, NZ, wz, iz);
        const T cor = fcor * (vy - L.v0);
        du = finite0(-ux * ddx - vy * ddy - wz * ddz + cor);
    }
    // v at y-face
    {
        const T vy = L.vy[3];
        const T ux = q * (L.ux[2] + L.uy[3] + L.u_xm_yp + L.uy[4]);
        const T wz = k == 0 ? q * (L.wz[3] + L.wy[4]) : q * (L.wz[2] + L.wz[3] + L.w_yp_zm + L.wy[4]);
        const T ddx = up7(L.vx, i, NX, ux, ix);
        const T ddy = up7(L.vy, j, NY, vy, iy);
        const T ddz = k == 0 ? (wz >= T(0) ? T(0) : (L.vz[4] - L.vz[3]) * iz) : up7(L.vz, k, NZ, wz, iz);
        const T cor = fcor * (ux - L.u0);
        dv = finite0(-ux * ddx - vy * ddy - wz * ddz - cor);
    }
    // w at z-face with buoyancy -g rho'/rho, moisture buoyancy and condensate loading
    {
        const T wz = L.wz[3];
        const T ux = q * (L.ux[2] + L.uz[3] + L.u_xm_zp + L.uz[4]);
        const T vy = q * (L.vy[2] + L.vz[3] + L.v_ym_zp + L.vz[4]);
        const T ddx = up7(L.wx, i, NX, ux, ix);
        const T ddy = up7(L.wy, j, NY, vy, iy);
        const T ddz = k == 0 ? (wz >= T(0) ? L.wz[3] * iz : (L.wz[4] - L.wz[3]) * iz) : up7(L.wz, k, NZ, wz, iz);
        const T rpf = h * (L.rp0 + L.rp1);
        const T rf = L.rho0f + rpf;
        const T buoy = -T(kG) * rpf / safe_rho(rf);
        const T qvf = h * (L.qv0 + L.qv1);
        const T mbuoy = T(kG * 0.608) * (qvf - L.qv0f);
        const T load = -T(kG) * (h * (L.l0 + L.l1));
        dw = finite0(-ux * ddx - vy * ddy - wz * ddz + buoy + mbuoy + load);
    }
}

// Heun slow RK2 (dynamics.cpp:5178-5292, 2055-2095), one cell per thread.
// MODE 0: stage 1 at u_n -> k1 and predictor u_n + dt k1 (into u2, v2, w2).
// MODE 1: stage 2 at the predictor, average in storage precision, apply to u_n.
// MODE 2: tendencies only (per-operation dispatch) into out_u/v/w.
template <class S, int MODE>
__global__ __launch_bounds__(kBlock) void k_slow(const MemberArgs* __restrict__ ma, const int2* __restrict__ map,
                                                 S* out_u, S* out_v, S* out_w)
{
    const Work wk = work_of(map);
    const MemberArgs& A = ma[wk.m];
    const DevGrid& g = A.g;
    const Cell cl = tile_cell_of(A, wk.local);
    if (!cl.valid) return;
    const int i = cl.i, j = cl.j, k = cl.k;
    const int NX = g.nx, NY = g.ny, NZ = g.nz;
    const DevFields& f = A.f;
    const double dt = A.p.dt;
    const std::size_t c = g.at(i, j, k);
    using T = typename AccT<S>::type;
    SlowStencil<T> L;
    load_slow_stencil<S, T>(g, P<S>(MODE == 1 ? f.u2 : f.u), P<S>(MODE == 1 ? f.v2 : f.v),
                            P<S>(MODE == 1 ? f.w2 : f.w), P<S>(f.rp), P<S>(f.qv), P<S>(f.qc), P<S>(f.qr), f.base, i, j,
                            k, L);
    // the stage-2 operands, loaded with the stencil
    S k1u = S(0), k1v = S(0), k1w = S(0), un = S(0), vn = S(0), wn = S(0);
    if constexpr (MODE == 1)
    {
        k1u = P<S>(f.k1u)[c]; k1v = P<S>(f.k1v)[c]; k1w = P<S>(f.k1w)[c];
        un = P<S>(f.u)[c]; vn = P<S>(f.v)[c]; wn = P<S>(f.w)[c];
    }
    T du = T(0), dv = T(0), dw = T(0);
    if (i >= 1 && i < NX - 1 && j >= 1 && j < NY - 1 && k < NZ - 1)
        slow_tendency_cell<T>(g, L, A.p.coriolis_f, i, j, k, du, dv, dw);
    if constexpr (MODE == 0)
    {
        const S a = S(du), b = S(dv), d = S(dw);
        P<S>(f.k1u)[c] = a;
        P<S>(f.k1v)[c] = b;
        P<S>(f.k1w)[c] = d;
        // the window centres are the cell's own u, v, w (exact in T)
        P<S>(f.u2)[c] = S(double(S(L.uz[3])) + double(a) * dt);
        P<S>(f.v2)[c] = S(double(S(L.vz[3])) + double(b) * dt);
        P<S>(f.w2)[c] = S(double(S(L.wz[3])) + double(d) * dt);
    }
    else if constexpr (MODE == 1)
    {
        const S au = S(S(0.5) * (S(du) + k1u));
        const S av = S(S(0.5) * (S(dv) + k1v));
        const S aw = S(S(0.5) * (S(dw) + k1w));
        P<S>(f.u)[c] = S(double(un) + finite0(double(au)) * dt);
        P<S>(f.v)[c] = S(double(vn) + finite0(double(av)) * dt);
        P<S>(f.w)[c] = S(double(wn) + finite0(double(aw)) * dt);
    }
    else
    {
        out_u[c] = S(du);
        out_v[c] = S(dv);
        out_w[c] = S(dw);
    }
}

// ---------------------------------------------------------------------
// Density clamp and EOS (acoustic_cgrid_fast.cpp:118-157).
// Pressure perturbation p - p0 from the dry EOS. FP64 storage evaluates
// p00 (rho R theta / p00)^gamma in FP64; FP32 storage evaluates the same
// expression relative to the base state, p = pe0 (1 + x)^gamma with
// 1 + x = rho theta / (rho0 theta0) and pe0 = EOS(rho0, theta0) in FP64, so
// p - p0 = pe0 expm1(gamma log1p(x)) + (pe0 - p0) keeps FP32 relative
// accuracy on the perturbation rather than on the full pressure.
// ---------------------------------------------------------------------
// (1 + x)^gamma - 1 in FP32. For |x| < 1/8 (every realistic perturbation)
// it is the binomial series sum_{n=1..8} C(gamma, n) x^n, whose truncation
// error is below 2e-10 relative, evaluated by Horner's rule (about 1 ulp);
// larger |x| take expm1(gamma log1p(x)).
__device__ __forceinline__ float pow1p_gamma_m1(float x)
{
    if (fabsf(x) < 0.125f)
    {
        // C(gamma, n) for gamma = cp / (cp - Rd) = 1004 / 717
        float p = 3.214725001e-03f;
        p = fmaf(x, p, -4.592693052e-03f);
        p = fmaf(x, p, 6.989304557e-03f);
        p = fmaf(x, p, -1.164974359e-02f);
        p = fmaf(x, p, 2.240575685e-02f);
        p = fmaf(x, p, -5.602415924e-02f);
        p = fmaf(x, p, 2.802510849e-01f);
        p = fmaf(x, p, 1.400278940e+00f);
        return x * p;
    }
    return expm1f(float(kGamma) * log1pf(x));
}

// Base-state profiles used by the EOS: rho0, p0, 1 / (rho0 theta0),
// pe0 = EOS(rho0, theta0) and pe0 - p0 at level k, read from the member's
// FP64 base table in global memory (GProf) or from a copy staged in LDS
// (LProf, layout [rho0 | p0 | 1/(rho0 theta0) | pe0 | pe0 - p0] x nz).
struct GProf
{
    gptr<const double> b;
    int nz;
    __device__ double rho0(int k) const { return b[std::size_t(kBaseRho0) * nz + k]; }
    __device__ double p0(int k) const { return b[std::size_t(kBaseP0) * nz + k]; }
    __device__ double inv_rt0(int k) const { return b[std::size_t(kBaseInvRT0) * nz + k]; }
    __device__ double pe0(int k) const { return b[std::size_t(kBasePe0) * nz + k]; }
    __device__ double pe_diff(int k) const { return b[std::size_t(kBasePeDiff) * nz + k]; }
};

constexpr int kLdsProfiles = 5;

struct LProf
{
    const double* s;
    int nz;
    __device__ double rho0(int k) const { return s[k]; }
    __device__ double p0(int k) const { return s[nz + k]; }
    __device__ double inv_rt0(int k) const { return s[2 * nz + k]; }
    __device__ double pe0(int k) const { return s[3 * nz + k]; }
    __device__ double pe_diff(int k) const { return s[4 * nz + k]; }
};

__device__ __forceinline__ GProf gprof(const double* base, int nz) { return GProf{G(base), nz}; }

// The same profile values of one level, loaded once into registers.
struct VProf
{
    double r0, p0_, irt, pe0_, ped;
    __device__ double rho0(int) const { return r0; }
    __device__ double p0(int) const { return p0_; }
    __device__ double inv_rt0(int) const { return irt; }
    __device__ double pe0(int) const { return pe0_; }
    __device__ double pe_diff(int) const { return ped; }
};

__device__ __forceinline__ VProf vprof(const double* base, int nz, int k)
{
    const GProf g = gprof(base, nz);
    return VProf{g.rho0(k), g.p0(k), g.inv_rt0(k), g.pe0(k), g.pe_diff(k)};
}

// Copies the LProf profiles of a member into LDS (all threads of the block take part).
__device__ __forceinline__ LProf stage_profiles(double* lds, const double* base, int nz)
{
    const gptr<const double> b = G(base);
    constexpr int slots[kLdsProfiles] = {kBaseRho0, kBaseP0, kBaseInvRT0, kBasePe0, kBasePeDiff};
    for (int t = int(threadIdx.x); t < kLdsProfiles * nz; t += int(blockDim.x))
    {
        const int q = t / nz, k = t - q * nz;
        lds[t] = b[std::size_t(slots[q]) * nz + k];
    }
    __syncthreads();
    return LProf{lds, nz};
}

template <class S, class Prof>
__device__ __forceinline__ double eos_full_pressure(const Prof& pr, int k, double rho, double theta)
{
    if constexpr (sizeof(S) == sizeof(float))
    {
        const double x = rho * theta * pr.inv_rt0(k) - 1.0;
        const float e = pow1p_gamma_m1(float(x));
        return pr.p0(k) + (double(e) * pr.pe0(k) + pr.pe_diff(k));
    }
    else
    {
        return dev::eos_pressure(rho, theta);
    }
}

template <class S>
struct RhoP
{
    S r, p;  // stored rho' and p'
};

template <class S, class Prof>
__device__ __forceinline__ RhoP<S> density_eos(const Bounds& bd, const Prof& pr, int k, double rn, double theta)
{
    const double rho0k = pr.rho0(k);
    const double p0k = pr.p0(k);
    if (!isfinite(rn) || rn <= 0.0) rn = fmax(0.1, rho0k);
    const double rs = clamp_full(rn, 1.0, bd.rho_min, bd.rho_max);
    RhoP<S> o;
    o.r = S(rs - rho0k);
    const double rho_used = rho0k + double(o.r);
    double pn = double(bd.p_fallback);
    if (theta > 0.0) pn = eos_full_pressure<S>(pr, k, rho_used, theta);
    if (!isfinite(pn) || pn <= 0.0) pn = double(bd.p_fallback);
    o.p = S(clamp_full(pn, 101325.0, bd.p_min, bd.p_max) - p0k);
    return o;
}

// ---------------------------------------------------------------------
// Explicit acoustic substep (acoustic_cgrid_fast.cpp:12-219) as a forward
// pass (retained mass-flux accumulation of the pre-continuity state,
// advection_cartesian.cpp:610-830; continuity; density clamp; EOS) and a
// backward pass (momentum with the new pressure). Substep s reads rho' from
// buffer s & 1 and writes buffer (s + 1) & 1, so no workgroup reads a
// neighbour's rho' after it changed; u, v, w are updated in place by the
// backward pass, which reads only its own cell's velocities.
// ---------------------------------------------------------------------
template <class S>
struct AcousticBufs
{
    gptr<const S> uo, vo, wo, ro, th;
    gptr<S> rn, pp, mfx, mfy, mfz, rb;
};

template <class S>
__device__ __forceinline__ RhoP<S> acoustic_forward(const DevGrid& g, const AcousticBufs<S>& B,
                                                    const double* __restrict__ base, const Bounds& bd, double dts,
                                                    double weight, int i, int j, int k, bool store, bool accumulate,
                                                    bool first, bool top_open)
{
    const int NX = g.nx, NY = g.ny, NZ = g.nz;
    const std::size_t c = g.at(i, j, k);
    const std::size_t sy = std::size_t(g.px), sz = g.plane();
    // all loads first, neighbours clamped to the domain (masked below)
    const std::size_t cxp = i < NX - 1 ? c + 1 : c, cyp = j < NY - 1 ? c + sy : c, czp = k < NZ - 1 ? c + sz : c;
    const std::size_t cxm = i >= 1 ? c - 1 : c, cym = j >= 1 ? c - sy : c, czm = k >= 1 ? c - sz : c;
    const int kp = k < NZ - 1 ? k + 1 : k, km = k >= 1 ? k - 1 : k;
    S ro_c = B.ro[c], ro_xp = B.ro[cxp], ro_yp = B.ro[cyp], ro_zp = B.ro[czp];
    S ro_xm = B.ro[cxm], ro_ym = B.ro[cym], ro_zm = B.ro[czm];
    S u_c = B.uo[c], v_c = B.vo[c], w_c = B.wo[c], u_xm = B.uo[cxm], v_ym = B.vo[cym], w_zm = B.wo[czm];
    S th_c = B.th[c];
    const bool acc = store && accumulate;
    S mfx0 = B.mfx[c], mfy0 = B.mfy[c], mfz0 = B.mfz[c];  // used only when accumulating onto them
    const auto rho0 = slot(base, kBaseRho0, NZ);
    VProf pv = vprof(base, NZ, k);
    const double r0 = pv.r0, r0p = rho0[kp], r0m = rho0[km];
    pin(ro_c); pin(ro_xp); pin(ro_yp); pin(ro_zp); pin(ro_xm); pin(ro_ym); pin(ro_zm);
    pin(u_c); pin(v_c); pin(w_c); pin(u_xm); pin(v_ym); pin(w_zm); pin(th_c);
    pin(mfx0); pin(mfy0); pin(mfz0);
    const double rc = r0 + double(ro_c);
    const double rxp = i < NX - 1 ? r0 + double(ro_xp) : 0.0;
    const double ryp = j < NY - 1 ? r0 + double(ro_yp) : 0.0;
    const double rzp = k < NZ - 1 ? r0p + double(ro_zp) : 0.0;
    const double uc = double(u_c), vc = double(v_c), wc = double(w_c);
    if (acc)
    {
        if (i < NX - 1) B.mfx[c] = S((first ? 0.0 : double(mfx0)) + weight * (0.5 * (rc + rxp) * uc));
        if (j < NY - 1) B.mfy[c] = S((first ? 0.0 : double(mfy0)) + weight * (0.5 * (rc + ryp) * vc));
        double fl = 0.0;
        if (k < NZ - 1) fl = 0.5 * (rc + rzp) * wc;
        else if (top_open) fl = rc * wc;
        B.mfz[c] = S((first ? 0.0 : double(mfz0)) + weight * fl);
    }
    if (store && first) B.rb[c] = ro_c;
    double d = 0.0;
    if (i >= 1 && i < NX - 1 && j >= 1 && j < NY - 1)
    {
        const double rxm = r0 + double(ro_xm);
        const double rym = r0 + double(ro_ym);
        const double fxp = 0.5 * (rc + rxp) * uc;
        const double fxm = 0.5 * (rxm + rc) * double(u_xm);
        const double fyp = 0.5 * (rc + ryp) * vc;
        const double fym = 0.5 * (rym + rc) * double(v_ym);
        const double fzp = k + 1 < NZ ? 0.5 * (rc + rzp) * wc : rc * wc;
        const double fzm = k > 0 ? 0.5 * ((r0m + double(ro_zm)) + rc) * double(w_zm) : 0.0;
        d = -((fxp - fxm) * g.idx + (fyp - fym) * g.idy + (fzp - fzm) * g.idz);
    }
    const S drho = S(finite0(d));
    const RhoP<S> o = density_eos<S>(bd, pv, k, rc + double(drho) * dts, double(th_c));
    if (store)
    {
        B.rn[c] = o.r;
        B.pp[c] = o.p;
    }
    return o;
}

// Backward momentum update of one cell with the new rho' and p' of the cell
// and its +x, +y, +z neighbours (acoustic_cgrid_fast.cpp:159-215).
template <class S>
__device__ __forceinline__ void momentum_update(const DevGrid& g, double r0, double r0p, const Bounds& bd,
                                                double dts, int i, int j, int k, RhoP<S> cc, RhoP<S> xp, RhoP<S> yp,
                                                RhoP<S> zp, double u, double v, double w, bool skip_w, gptr<S> un,
                                                gptr<S> vn, gptr<S> wn, std::size_t c)
{
    const int NX = g.nx, NY = g.ny, NZ = g.nz;
    double du = 0.0, dv = 0.0, dw = 0.0;
    const bool interior = i >= 1 && i < NX - 1 && j >= 1 && j < NY - 1 && k < NZ - 1;
    if (interior)
    {
        const double rc = r0 + double(cc.r);
        const double rx = 0.5 * (rc + (r0 + double(xp.r)));
        const double ry = 0.5 * (rc + (r0 + double(yp.r)));
        const double rz = 0.5 * (rc + (r0p + double(zp.r)));
        if constexpr (sizeof(S) == sizeof(float))
        {
            // FP32 tendency, as the reference rounds it to FP32 before the update
            auto safe_f = [](float r) { return isfinite(r) && r > 1.0e-6f ? r : 1.0f; };
            auto fin = [](float x) { return isfinite(x) ? x : 0.0f; };
            const float pc = float(cc.p);
            du = double(fin(-(float(xp.p) - pc) * float(g.idx) / safe_f(float(rx))));
            dv = double(fin(-(float(yp.p) - pc) * float(g.idy) / safe_f(float(ry))));
            dw = double(fin(-((float(zp.p) - pc) * float(g.idz)) / safe_f(float(rz))));
        }
        else
        {
            const double pc = double(cc.p);
            du = finite0(-(double(xp.p) - pc) * g.idx / safe_rho(rx));
            dv = finite0(-(double(yp.p) - pc) * g.idy / safe_rho(ry));
            dw = finite0(-((double(zp.p) - pc) * g.idz) / safe_rho(rz));
        }
    }
    un[c] = S(clamp_store<S>(u + du * dts, 0.0, -bd.wind_h, bd.wind_h));
    vn[c] = S(clamp_store<S>(v + dv * dts, 0.0, -bd.wind_h, bd.wind_h));
    if (!(skip_w && interior)) wn[c] = S(clamp_store<S>(w + dw * dts, 0.0, -bd.wind_v, bd.wind_v));
}

template <class S>
__global__ __launch_bounds__(kBlock) void k_acoustic_forward(const MemberArgs* __restrict__ ma,
                                                             const int2* __restrict__ map, int s, int accumulate,
                                                             int init)
{
    const Work wk = work_of(map);
    const MemberArgs& A = ma[wk.m];
    if (s >= A.p.substeps) return;
    const Cell cl = cell_of(A, wk.local);
    if (!cl.valid) return;
    const DevFields& f = A.f;
    const int par = s & 1;
    AcousticBufs<S> B;
    B.uo = P<S>(f.u);
    B.vo = P<S>(f.v);
    B.wo = P<S>(f.w);
    B.ro = P<S>(par ? f.rp2 : f.rp);
    B.rn = P<S>(par ? f.rp : f.rp2);
    B.th = P<S>(f.th);
    B.pp = P<S>(f.pp);
    B.mfx = P<S>(f.mfx);
    B.mfy = P<S>(f.mfy);
    B.mfz = P<S>(f.mfz);
    B.rb = P<S>(f.rb);
    acoustic_forward<S>(A.g, B, f.base, A.p.bounds, A.p.dts, A.p.mf_weight, cl.i, cl.j, cl.k, true,
                        accumulate != 0, init != 0 && s == 0, A.p.bc_top == NIMBUS_TOP_OPEN);
}

// ---------------------------------------------------------------------
// HE-VI substep: horizontal continuity explicit, vertical w-rho-p coupling
// off-centred implicit (alpha = (1+beta)/2): the vertical pressure gradient
// acts on alpha p^{n+1} + (1 - alpha) p^n with p^{n+1} linearised about
// p* = EOS(rho*); FP64 Thomas sweep per column, flux-form density update and
// EOS re-diagnosis. Two kernels: k_hevi_explicit (one thread per cell) does
// the horizontal part, k_hevi_column (one thread per column) the vertical
// solve. rho' is read from buffer s & 1 and written to (s + 1) & 1; u, v, w
// stay in their home buffers (w is only touched by its own column).
// ---------------------------------------------------------------------

// Part 1, per cell: retained horizontal mass fluxes, the density after the
// explicit horizontal divergence rho* and its EOS pressure p*, stored as
// perturbations (rho* - rho0 in the new-density buffer, p* - p0 in the u2
// scratch field, which is dead between the slow stage 2 and transport).
// Ghost columns get the reference's pass-2 clamp/EOS here and are not
// touched by the column solve.
template <class S>
__global__ __launch_bounds__(kBlock) void k_hevi_explicit(const MemberArgs* __restrict__ ma,
                                                          const int2* __restrict__ map, int s, int accumulate,
                                                          int init)
{
    const Work wk = work_of(map);
    const MemberArgs& A = ma[wk.m];
    if (s >= A.p.substeps) return;
    const Cell cl = cell_of(A, wk.local);
    if (!cl.valid) return;
    const DevGrid& g = A.g;
    const int NX = g.nx, NY = g.ny, NZ = g.nz;
    const int i = cl.i, j = cl.j, k = cl.k;
    const DevFields& f = A.f;
    const int par = s & 1;
    const auto u = P<S>(f.u), v = P<S>(f.v), th = P<S>(f.th);
    const auto rp_in = P<S>(par ? f.rp2 : f.rp);
    const auto rp_out = P<S>(par ? f.rp : f.rp2);
    const auto mfx = P<S>(f.mfx), mfy = P<S>(f.mfy);
    const bool acc = accumulate != 0, first = init != 0 && s == 0;
    const double dts = A.p.dts, weight = A.p.mf_weight;
    const std::size_t sy = std::size_t(g.px);
    const std::size_t c = g.at(i, j, k);
    const std::size_t cxp = i < NX - 1 ? c + 1 : c, cyp = j < NY - 1 ? c + sy : c;
    const std::size_t cxm = i >= 1 ? c - 1 : c, cym = j >= 1 ? c - sy : c;
    S rpc = rp_in[c], rpxp = rp_in[cxp], rpyp = rp_in[cyp], rpxm = rp_in[cxm], rpym = rp_in[cym];
    S uc = u[c], um = u[cxm], vc = v[c], vm = v[cym], thc = th[c];
    S mfx0 = mfx[c], mfy0 = mfy[c];  // used only when accumulating onto them
    const VProf pv = vprof(f.base, NZ, k);
    pin(rpc); pin(rpxp); pin(rpyp); pin(rpxm); pin(rpym); pin(uc); pin(um); pin(vc); pin(vm); pin(thc);
    pin(mfx0); pin(mfy0);
    const double r0 = pv.r0;
    const double rc = r0 + double(rpc);
    const double rxp = r0 + double(rpxp), ryp = r0 + double(rpyp);
    if (acc)
    {
        if (i < NX - 1) mfx[c] = S((first ? 0.0 : double(mfx0)) + weight * (0.5 * (rc + rxp) * double(uc)));
        if (j < NY - 1) mfy[c] = S((first ? 0.0 : double(mfy0)) + weight * (0.5 * (rc + ryp) * double(vc)));
    }
    if (first) P<S>(f.rb)[c] = rpc;
    if (!(i >= 1 && i < NX - 1 && j >= 1 && j < NY - 1))
    {
        // ghost column: density unchanged (clamped), EOS pressure, no vertical flux
        if (acc && first) P<S>(f.mfz)[c] = S(0);
        const RhoP<S> o = density_eos<S>(A.p.bounds, pv, k, rc, double(thc));
        rp_out[c] = o.r;
        P<S>(f.pp)[c] = o.p;
        return;
    }
    const double rxm = r0 + double(rpxm), rym = r0 + double(rpym);
    const double dh = -((0.5 * (rc + rxp) * double(uc) - 0.5 * (rxm + rc) * double(um)) * g.idx +
                        (0.5 * (rc + ryp) * double(vc) - 0.5 * (rym + rc) * double(vm)) * g.idy);
    const double rstar = rc + dts * finite0(dh);
    // linearisation point p* = EOS(rho*): FP64 power law in FP64 storage, the
    // perturbation form (FP32 relative accuracy on p* - p0) in the mixed mode
    const double ps = eos_full_pressure<S>(pv, k, rstar, double(thc));
    rp_out[c] = S(rstar - r0);
    P<S>(f.u2)[c] = S(ps - pv.p0_);
}

// Part 2, one thread per interior column. Up-sweep: each row (face) of the
// w system from its two cells' rho*, p*, substep-start p and face
// densities, eliminated at once (FP64 c', d' kept in tri_c / tri_r).
// Down-sweep: back substitution, off-centred vertical mass flux, final
// density and EOS pressure. The inputs of the next kRing - 1 levels are in
// flight while a level is processed: a ring of kRing register sets, the loop
// unrolled by kRing so the ring rotates by renaming, every load unconditional
// (levels clamped to the column), the base-state profiles in LDS.
constexpr int kRing = 4;

template <class S>
struct HeviUp
{
    S rs, ps, pn, w, rp;  // rho* - rho0, p* - p0, substep-start p', w, substep-start rho'
};

template <class S>
struct HeviDown
{
    double tr, tc;
    S w, rp, rs, th, mfz;
};

template <class S>
__global__ __launch_bounds__(kBlock) void k_hevi_column(const MemberArgs* __restrict__ ma,
                                                        const int2* __restrict__ map, int s, int accumulate, int init)
{
    const Work wk = work_of(map);
    const MemberArgs& A = ma[wk.m];
    const DevGrid& g = A.g;
    // uniform over the block (one member per block); members the batch solves
    // in the partitioned form are not this kernel's
    if (s >= A.p.substeps || A.p.hevi_seg) return;
    const Column t = column_of(A, wk.local);
    const int NX = g.nx, NY = g.ny, NZ = g.nz;
    extern __shared__ double s_prof[];
    const LProf pr = stage_profiles(s_prof, A.f.base, NZ);
    const int i = t.i, j = t.j;
    if (!(i >= 1 && i < NX - 1 && j >= 1 && j < NY - 1)) return;  // no barrier after this point
    const DevFields& f = A.f;
    const int par = s & 1;
    const gptr<S> __restrict__ w = P<S>(f.w);
    const gptr<const S> __restrict__ rp_in = P<S>(par ? f.rp2 : f.rp);
    const gptr<S> __restrict__ rp_out = P<S>(par ? f.rp : f.rp2);
    const gptr<S> __restrict__ pp = P<S>(f.pp);
    const gptr<const S> __restrict__ psb = P<S>(f.u2);
    const gptr<const S> __restrict__ th = P<S>(f.th);
    const gptr<S> __restrict__ mfz = P<S>(f.mfz);
    const gptr<double> __restrict__ tc = G(f.tri_c);
    const gptr<double> __restrict__ tr = G(f.tri_r);
    const Bounds bd = A.p.bounds;
    const double dts = A.p.dts, alpha = A.p.hevi_alpha, weight = A.p.mf_weight;
    const bool acc = accumulate != 0, first = init != 0 && s == 0;
    const std::size_t c0 = g.at(i, j, 0), sz = g.plane();
    const double idz = g.idz;
    const double a1 = 1.0 - alpha;

    // ---- up-sweep
    auto load_up = [&](int k) {
        const std::size_t c = c0 + std::size_t(k) * sz;
        HeviUp<S> L;
        L.rs = rp_out[c]; L.ps = psb[c]; L.pn = pp[c]; L.w = w[c]; L.rp = rp_in[c];
        return L;
    };
    HeviUp<S> ring[kRing];
#pragma unroll
    for (int q = 0; q < kRing; ++q) ring[q] = load_up(min(q, NZ - 1));
    double rb_prev = 0.0, C_prev = 0.0, ps_prev = 0.0, pn_prev = 0.0, wo_prev = 0.0, e_prev = 0.0;
    double rb_prev2 = 0.0, cp_prev = 0.0, rr_prev = 0.0;
    bool ok = true;
    for (int k0 = 0; k0 < NZ; k0 += kRing)
    {
#pragma unroll
        for (int q = 0; q < kRing; ++q)
        {
            const int k = k0 + q;
            const HeviUp<S> cur = ring[q];
            const HeviUp<S> nxt = ring[(q + 1) % kRing];  // level k + 1
            if (k < NZ)  // not a loop exit, and the refill below reuses cur's registers
            {
                const double r0 = pr.rho0(k);
                const double rc = r0 + double(cur.rp);
                const double rstar = r0 + double(cur.rs);
                const double ps = pr.p0(k) + double(cur.ps);
                const double C = kGamma * ps * (1.0 / rstar);
                const double wo = double(cur.w);
                // face density above level k (zero at the lid)
                const double rbk = k < NZ - 1 ? 0.5 * (rc + (pr.rho0(k + 1) + double(nxt.rp))) : 0.0;
                const double e = -a1 * dts * idz * (rbk * wo - rb_prev * wo_prev);
                const double pnk = double(cur.pn);
                if (k >= 1)
                {
                    // row r = k-1 (face k-1/2): unknowns w_{r-1}, w_r, w_{r+1}
                    const double rbr = rb_prev;
                    const double inv_rbr = 1.0 / rbr;
                    const double Bc = alpha * dts * idz;
                    const double Ak = Bc * inv_rbr;
                    const double Aa = (k - 1) > 0 ? -Ak * Bc * C_prev * rb_prev2 : 0.0;
                    const double U = (k - 1) < NZ - 2 ? -Ak * Bc * C * rbk : 0.0;
                    const double D = 1.0 + Ak * Bc * (C + C_prev) * rbr;
                    // PGF on alpha p* + (1 - alpha) p^n (p^n: substep-start pressure), base gradient removed
                    const double G_ = (alpha * ((ps - ps_prev) - (pr.p0(k) - pr.p0(k - 1))) + a1 * (pnk - pn_prev)) * idz;
                    const double R = wo_prev - dts * inv_rbr * G_ - Ak * (C * e - C_prev * e_prev);
                    const double piv = D - Aa * cp_prev;
                    if (!(piv > 0.0) || !isfinite(piv)) ok = false;
                    const double inv_piv = 1.0 / piv;
                    const double cpr = U * inv_piv;
                    const double rrr = (R - Aa * rr_prev) * inv_piv;
                    const std::size_t cr = c0 + std::size_t(k - 1) * sz;
                    tc[cr] = cpr;
                    tr[cr] = rrr;
                    cp_prev = cpr;
                    rr_prev = rrr;
                }
                rb_prev2 = rb_prev;
                rb_prev = rbk;
                C_prev = C;
                ps_prev = ps;
                pn_prev = pnk;
                wo_prev = wo;
                e_prev = e;
            }
            ring[q] = load_up(min(k + kRing, NZ - 1));
        }
    }

    // ---- down-sweep: iteration t handles face r = NZ-2-t (none at r = -1)
    // and finalises cell r+1 (top face r+1, zero at the lid; bottom face r).
    auto load_down = [&](int r) {
        const std::size_t cr = c0 + std::size_t(r) * sz;
        HeviDown<S> d;
        d.w = w[cr]; d.tr = tr[cr]; d.tc = tc[cr]; d.rp = rp_in[cr]; d.rs = rp_out[cr]; d.th = th[cr];
        d.mfz = mfz[cr];
        return d;
    };
    double w_above = 0.0, wo_above = 0.0, rb_above = 0.0;  // face r+1 quantities (stored w)
    double x_above = 0.0;                                  // face r+1 Thomas solution (unclamped)
    HeviDown<S> cell = load_down(NZ - 1);                  // cell r+1 (its w, tr, tc are not used)
    HeviDown<S> dring[kRing];
#pragma unroll
    for (int q = 0; q < kRing; ++q) dring[q] = load_down(max(NZ - 2 - q, 0));
    for (int t0 = 0; t0 < NZ; t0 += kRing)
    {
#pragma unroll
        for (int q = 0; q < kRing; ++q)
        {
            const int r = NZ - 2 - (t0 + q);
            const HeviDown<S> dn = dring[q];
            if (r >= -1)
            {
                double wn = 0.0, wo = 0.0, rbr = 0.0;
                if (r >= 0)
                {
                    const std::size_t cr = c0 + std::size_t(r) * sz;
                    wo = double(dn.w);
                    const double x = ok ? dn.tr - dn.tc * x_above : wo;
                    x_above = x;
                    wn = clamp_store<S>(x, 0.0, -bd.wind_v, bd.wind_v);
                    w[cr] = S(wn);
                    wn = double(S(wn));
                    rbr = 0.5 * ((pr.rho0(r) + double(dn.rp)) + (pr.rho0(r + 1) + double(cell.rp)));
                }
                const int kc = r + 1;
                const std::size_t cc = c0 + std::size_t(kc) * sz;
                const double ft = kc < NZ - 1 ? rb_above * (alpha * w_above + a1 * wo_above) : 0.0;
                const double fb = r >= 0 ? rbr * (alpha * wn + a1 * wo) : 0.0;
                if (acc) mfz[cc] = S((first ? 0.0 : double(cell.mfz)) + weight * ft);
                const double rstar = pr.rho0(kc) + double(cell.rs);
                const RhoP<S> o = density_eos<S>(bd, pr, kc, rstar - dts * idz * (ft - fb), double(cell.th));
                rp_out[cc] = o.r;
                pp[cc] = o.p;
                w_above = wn;
                wo_above = wo;
                rb_above = rbr;
                cell = dn;
            }
            dring[q] = load_down(max(r - kRing, 0));
        }
    }
}

// Part 2, partitioned form (columns of at most kSegMaxNz levels, chosen per
// member by StepParams::hevi_seg): the same rows as the sequential sweep, in
// FP64, solved without coefficient arrays in memory. Each column's faces
// 0 .. nz-2 are split into runs of ls consecutive faces, one thread per run
// (kSegCols consecutive columns of one row x kSegments runs per workgroup, so
// every load is 16 contiguous lanes of one level). A thread
//  1. loads its run's levels once (one level below and two above for the
//     face densities);
//  2. forward-eliminates its rows to x_r + c1_r x_{r+1} + a1_r L' = d1_r,
//     where L' is the last unknown of the run below (a1 carries that
//     coupling; the first run has none, and its c1, d1 are the sequential
//     sweep's c', d'), and back-eliminates its first row to
//     x_{s0} = dt0 - at0 L' - ct0 L, with L its own last unknown;
//  3. forms its row of the system of the runs' last unknowns (row p:
//     a1 L_{p-1} + (1 - c1 at') L_p - c1 ct' L_{p+1} = d1 - c1 dt', primes for
//     run p + 1's first row), solved by cyclic reduction over the column's
//     run threads in LDS;
//  4. substitutes back down its run and finalises the cells whose lower face
//     is in it (the face above the run is run p + 1's first unknown,
//     evaluated from the same LDS values by the same expression).
// Quotients use rcp_nr (about an ulp) rather than correctly rounded
// division. Every pivot of the runs and of the reduction must be positive
// and finite, otherwise the whole column keeps its substep-start w, as in the
// sequential sweep. A run's neighbours overwrite its halo levels only after
// the barriers, by which time every halo value has been consumed.
template <class S>
__global__ __launch_bounds__(kBlock) void k_hevi_column_seg(const MemberArgs* __restrict__ ma,
                                                            const int2* __restrict__ map, int s, int accumulate,
                                                            int init)
{
    constexpr int NQ = kSegRows + 3;  // levels s0 - 1 .. s0 + ls + 1
    const Work wk = work_of(map);
    const MemberArgs& A = ma[wk.m];
    const DevGrid& g = A.g;
    const int NX = g.nx, NY = g.ny, NZ = g.nz;
    // uniform over the block (one member per block): no barrier is skipped by part of it
    if (s >= A.p.substeps || !A.p.hevi_seg || NZ > kSegMaxNz) return;
    extern __shared__ double s_prof[];
    const LProf pr = stage_profiles(s_prof, A.f.base, NZ);
    double* const ex = s_prof + kLdsProfiles * NZ;         // [11][kBlock]: run ends and reduced rows
    double* const lv = ex + 11 * kBlock;                   // [kBlock]: last unknown of each run
    int* const okv = reinterpret_cast<int*>(lv + kBlock);  // [2][kBlock]: run and reduction verdicts

    const int tid = int(threadIdx.x);
    const int col = tid & (kSegCols - 1), p = tid / kSegCols;
    const int row_blocks = (NX + kSegCols - 1) / kSegCols;
    const int j = wk.local / row_blocks;
    const int i = (wk.local - j * row_blocks) * kSegCols + col;
    const bool interior = i >= 1 && i < NX - 1 && j >= 1 && j < NY - 1;
    const int n = NZ - 1;  // unknown faces 0 .. NZ-2
    const int ls = (n + kSegments - 1) / kSegments;
    const int runs = (n + ls - 1) / ls;
    const int s0 = p * ls;
    const int m = interior && p < runs ? min(ls, n - s0) : 0;

    const DevFields& f = A.f;
    const int par = s & 1;
    const gptr<S> __restrict__ w = P<S>(f.w);
    const gptr<const S> __restrict__ rp_in = P<S>(par ? f.rp2 : f.rp);
    const gptr<S> __restrict__ rp_out = P<S>(par ? f.rp : f.rp2);
    const gptr<S> __restrict__ pp = P<S>(f.pp);
    const gptr<const S> __restrict__ psb = P<S>(f.u2);
    const gptr<const S> __restrict__ th = P<S>(f.th);
    const gptr<S> __restrict__ mfz = P<S>(f.mfz);
    const Bounds bd = A.p.bounds;
    const double dts = A.p.dts, alpha = A.p.hevi_alpha, weight = A.p.mf_weight;
    const bool acc = accumulate != 0, first = init != 0 && s == 0;
    const double idz = g.idz;
    const double a1 = 1.0 - alpha;
    const std::size_t c0 = m > 0 ? g.at(i, j, 0) : 0, sz = g.plane();

    // 1. level q is level s0 - 1 + q (clamped to the column for the load)
    S rp[NQ], wo_s[NQ], rs[NQ], ps[NQ], pn[NQ], tq[NQ], mq[NQ];
#pragma unroll
    for (int q = 0; q < NQ; ++q)
    {
        rp[q] = wo_s[q] = rs[q] = ps[q] = pn[q] = tq[q] = mq[q] = S(0);
        if (q <= m + 2 && m > 0)
        {
            const std::size_t c = c0 + std::size_t(clampi(s0 - 1 + q, 0, NZ - 1)) * sz;
            rp[q] = rp_in[c];
            wo_s[q] = w[c];
            if (q >= 1 && q <= m + 1)
            {
                rs[q] = rp_out[c];
                ps[q] = psb[c];
                pn[q] = pp[c];
            }
        }
    }
    auto lev = [&](int q) { return clampi(s0 - 1 + q, 0, NZ - 1); };
    // old w of face s0 - 1 + q (none below the ground face)
    auto wo = [&](int q) { return s0 - 1 + q >= 0 ? double(wo_s[q]) : 0.0; };
    // full density of cell s0 - 1 + q
    auto rc_at = [&](int q) { return pr.rho0(lev(q)) + double(rp[q]); };
    // density of face s0 - 1 + q between cells q and q + 1 (zero below the ground face and at the lid)
    auto rb_of = [&](int q, double rc_lo, double rc_hi) {
        const int fq = s0 - 1 + q;
        return fq >= 0 && fq < NZ - 1 ? 0.5 * (rc_lo + rc_hi) : 0.0;
    };
    // cell s0 - 1 + q: C = gamma p* / rho* and the full p*
    auto cell_c = [&](int q, double& Cq, double& psq) {
        const int k = lev(q);
        const double rstar = pr.rho0(k) + double(rs[q]);
        psq = pr.p0(k) + double(ps[q]);
        Cq = kGamma * psq * rcp_nr(rstar);
    };

    // 2. forward elimination of rows r = s0 + t (cells r, r + 1; faces r - 1, r, r + 1),
    // C, the old-w divergence e and the face densities rolled along as in the sequential sweep
    double c1[kSegRows], a1s[kSegRows], d1[kSegRows];
    bool ok = true;
    double rb_lo = 0.0, rb_mid = 0.0, rc_nx = 0.0, C_lo = 0.0, ps_lo = 0.0, e_lo = 0.0;
    if (m > 0)
    {
        const double r0 = rc_at(0), r1 = rc_at(1);
        rc_nx = rc_at(2);
        rb_lo = rb_of(0, r0, r1);
        rb_mid = rb_of(1, r1, rc_nx);
        cell_c(1, C_lo, ps_lo);
        e_lo = -a1 * dts * idz * (rb_mid * wo(1) - rb_lo * wo(0));
    }
    const double Bc = alpha * dts * idz;
    double cp = 0.0, ap = 0.0, dp = 0.0;  // the last eliminated row
#pragma unroll
    for (int t = 0; t < kSegRows; ++t)
    {
        if (t < m)
        {
            const int r = s0 + t;
            const double rc_hi = rc_at(t + 3);
            const double rb_hi = rb_of(t + 2, rc_nx, rc_hi);
            rc_nx = rc_hi;
            double C_hi, ps_hi;
            cell_c(t + 2, C_hi, ps_hi);
            const double e_hi = -a1 * dts * idz * (rb_hi * wo(t + 2) - rb_mid * wo(t + 1));
            const double rbr = rb_mid;
            const double inv_rbr = rcp_nr(rbr);
            const double Ak = Bc * inv_rbr;
            const double Aa = r > 0 ? -Ak * Bc * C_lo * rb_lo : 0.0;
            const double U = r < NZ - 2 ? -Ak * Bc * C_hi * rb_hi : 0.0;
            const double D = 1.0 + Ak * Bc * (C_hi + C_lo) * rbr;
            const double G_ = (alpha * ((ps_hi - ps_lo) - (pr.p0(r + 1) - pr.p0(r))) +
                               a1 * (double(pn[t + 2]) - double(pn[t + 1]))) *
                              idz;
            const double R = wo(t + 1) - dts * inv_rbr * G_ - Ak * (C_hi * e_hi - C_lo * e_lo);
            const double piv = D - Aa * cp;
            if (!(piv > 0.0) || !isfinite(piv)) ok = false;
            const double inv_piv = rcp_nr(piv);
            c1[t] = U * inv_piv;
            a1s[t] = (t == 0 ? Aa : -Aa * ap) * inv_piv;
            d1[t] = (R - Aa * dp) * inv_piv;
            cp = c1[t];
            ap = a1s[t];
            dp = d1[t];
            rb_lo = rb_mid;
            rb_mid = rb_hi;
            C_lo = C_hi;
            ps_lo = ps_hi;
            e_lo = e_hi;
        }
        else
        {
            c1[t] = a1s[t] = d1[t] = 0.0;
        }
    }
    // the first row back-eliminated: x_{s0} = dt0 - at0 L' - ct0 L (the last row is L itself)
    double dt0 = 0.0, at0 = 0.0, ct0 = -1.0;
#pragma unroll
    for (int t = kSegRows - 2; t >= 0; --t)
    {
        if (t < m - 1)
        {
            dt0 = d1[t] - c1[t] * dt0;
            at0 = a1s[t] - c1[t] * at0;
            ct0 = -c1[t] * ct0;
        }
    }

    // 3. the system of the runs' last unknowns L_p, one row per run thread:
    // a1 L_{p-1} + (1 - c1 at') L_p - c1 ct' L_{p+1} = d1 - c1 dt' (primes:
    // run p + 1's first row), solved by cyclic reduction over the kSegments
    // run threads of the column in LDS (log2(kSegments) steps of stride 1, 2,
    // 4, ..., double buffered so one barrier per step suffices). Runs past the
    // column and non-interior columns carry identity rows.
    double* const nb = ex;                 // [3][kBlock]: at0, ct0, dt0 of each run
    double* const rbuf = ex + 3 * kBlock;  // [2][4][kBlock]: reduced rows (sub, dia, sup, rhs)
    nb[0 * kBlock + tid] = at0;
    nb[1 * kBlock + tid] = ct0;
    nb[2 * kBlock + tid] = dt0;
    okv[tid] = ok ? 1 : 0;
    __syncthreads();
    double sub = 0.0, dia = 1.0, sup = 0.0, rhs = 0.0;
    if (m > 0)
    {
        const int un = tid + kSegCols;
        const bool more = p + 1 < runs;
        const double aN = more ? nb[0 * kBlock + un] : 0.0;
        const double cN = more ? nb[1 * kBlock + un] : 0.0;
        const double dN = more ? nb[2 * kBlock + un] : 0.0;
        sub = ap;
        dia = 1.0 - cp * aN;
        sup = -cp * cN;
        rhs = dp - cp * dN;
    }
    bool col_ok = true;
    for (int q = 0; q < runs; ++q) col_ok = col_ok && okv[q * kSegCols + col] != 0;
#pragma unroll
    for (int st = 0; st < kSegSteps; ++st)
    {
        const int stride = 1 << st;
        double* const b = rbuf + (st & 1) * 4 * kBlock;
        b[0 * kBlock + tid] = sub;
        b[1 * kBlock + tid] = dia;
        b[2 * kBlock + tid] = sup;
        b[3 * kBlock + tid] = rhs;
        __syncthreads();
        double k1 = 0.0, k2 = 0.0, nsub = 0.0, nsup = 0.0, ndia = dia, nrhs = rhs;
        if (p - stride >= 0)
        {
            const int u = tid - stride * kSegCols;
            k1 = sub * rcp_nr(b[1 * kBlock + u]);
            nsub = -k1 * b[0 * kBlock + u];
            ndia -= k1 * b[2 * kBlock + u];
            nrhs -= k1 * b[3 * kBlock + u];
        }
        if (p + stride < kSegments)
        {
            const int u = tid + stride * kSegCols;
            k2 = sup * rcp_nr(b[1 * kBlock + u]);
            nsup = -k2 * b[2 * kBlock + u];
            ndia -= k2 * b[0 * kBlock + u];
            nrhs -= k2 * b[3 * kBlock + u];
        }
        sub = nsub;
        sup = nsup;
        dia = ndia;
        rhs = nrhs;
        if (m > 0 && (!(dia > 0.0) || !isfinite(dia))) col_ok = false;
    }
    const double Lp = rhs * rcp_nr(dia);
    lv[tid] = Lp;
    okv[kBlock + tid] = col_ok ? 1 : 0;
    __syncthreads();
    if (m <= 0) return;  // no barrier after this point
    for (int q = 0; q < runs; ++q) col_ok = col_ok && okv[kBlock + q * kSegCols + col] != 0;

    // 4. back substitution from the top of the run down; each cell s0 + t + 1 is
    // finalised once both of its faces are known (the face above the run is
    // run p + 1's first unknown, the same expression and operands as there)
    // theta and the mass flux of the cells this run finalises (no other run
    // writes them) are loaded only now to keep them out of the elimination's
    // registers; rho' and rho* of levels s0 .. s0 + m + 1 are read again here rather than
    // kept: rho' is not written by this kernel, and rho* of a cell is
    // overwritten only by the run that finalises it, after this read
#pragma unroll
    for (int q = 1; q < NQ; ++q)
    {
        if (q <= m + 2)
        {
            const std::size_t c = c0 + std::size_t(lev(q)) * sz;
            rp[q] = rp_in[c];
            if (q >= (p == 0 ? 1 : 2) && q <= m + 1)
            {
                rs[q] = rp_out[c];
                tq[q] = th[c];
                if (acc && !first) mq[q] = mfz[c];
            }
        }
    }
    const double Lm = p > 0 ? lv[tid - kSegCols] : 0.0;
    // old w, density and stored new w of the face above the run (face s0 + m)
    double wo_e = 0.0, rc_e = 0.0, rc_e1 = 0.0;
#pragma unroll
    for (int q = 1; q < NQ - 1; ++q)
        if (q == m + 1)
        {
            wo_e = wo(q);
            rc_e = rc_at(q);
            rc_e1 = rc_at(q + 1);
        }
    double rb_up = rb_of(m + 1, rc_e, rc_e1);
    double rc_up = rc_e;  // density of the cell above the current face
    double w_up = 0.0;
    if (s0 + m < n)
    {
        const int u = tid + kSegCols;
        const double x = col_ok ? nb[2 * kBlock + u] - nb[0 * kBlock + u] * Lp - nb[1 * kBlock + u] * lv[u] : wo_e;
        w_up = double(S(clamp_store<S>(x, 0.0, -bd.wind_v, bd.wind_v)));
    }
    auto finalise = [&](int q, double rb_top, double w_top, double rb_bot, double w_bot) {
        // cell s0 - 1 + q: upper face q, lower face q - 1
        const int kc = s0 - 1 + q;
        const std::size_t cc = c0 + std::size_t(kc) * sz;
        const double ft = kc < NZ - 1 ? rb_top * (alpha * w_top + a1 * wo(q)) : 0.0;
        const double fb = kc >= 1 ? rb_bot * (alpha * w_bot + a1 * wo(q - 1)) : 0.0;
        if (acc) mfz[cc] = S((first ? 0.0 : double(mq[q])) + weight * ft);
        const double rstar = pr.rho0(kc) + double(rs[q]);
        const RhoP<S> o = density_eos<S>(bd, pr, kc, rstar - dts * idz * (ft - fb), double(tq[q]));
        rp_out[cc] = o.r;
        pp[cc] = o.p;
    };
    double x_up = Lp;  // unclamped solution of the face above
#pragma unroll
    for (int t = kSegRows - 1; t >= 0; --t)
    {
        if (t < m)
        {
            double x = t == m - 1 ? Lp : t == 0 ? dt0 - at0 * Lm - ct0 * Lp : d1[t] - c1[t] * x_up - a1s[t] * Lm;
            x_up = x;
            if (!col_ok) x = wo(t + 1);
            const double c = clamp_store<S>(x, 0.0, -bd.wind_v, bd.wind_v);
            w[c0 + std::size_t(s0 + t) * sz] = S(c);
            const double wn = double(S(c));
            const double rc_lo = rc_at(t + 1);
            const double rb = rb_of(t + 1, rc_lo, rc_up);
            finalise(t + 2, rb_up, w_up, rb, wn);
            w_up = wn;
            rb_up = rb;
            rc_up = rc_lo;
        }
    }
    if (p == 0) finalise(1, rb_up, w_up, 0.0, 0.0);  // cell 0 above the ground face
}

// Backward momentum update of substep s with the new rho' (buffer
// (s + 1) & 1) and pressure, in place. HE-VI members keep the interior w the
// column solve set; their ghost-column and lid w get the clamp only.
template <class S>
__global__ __launch_bounds__(kBlock) void k_momentum(const MemberArgs* __restrict__ ma,
                                                     const int2* __restrict__ map, int s)
{
    const Work wk = work_of(map);
    const MemberArgs& A = ma[wk.m];
    if (s >= A.p.substeps) return;
    const Cell cl = cell_of(A, wk.local);
    if (!cl.valid) return;
    const DevGrid& g = A.g;
    const int NX = g.nx, NY = g.ny, NZ = g.nz;
    const int i = cl.i, j = cl.j, k = cl.k;
    const DevFields& f = A.f;
    const gptr<const S> __restrict__ rn = P<S>((s & 1) ? f.rp : f.rp2);
    const gptr<const S> __restrict__ pp = P<S>(f.pp);
    const auto u = P<S>(f.u);
    const auto v = P<S>(f.v);
    const auto w = P<S>(f.w);
    const std::size_t sy = std::size_t(g.px), sz = g.plane();
    const std::size_t c = g.at(i, j, k);
    // neighbours clamped to the domain (the cell itself stands in for a missing one)
    const std::size_t cxp = i < NX - 1 ? c + 1 : c, cyp = j < NY - 1 ? c + sy : c, czp = k < NZ - 1 ? c + sz : c;
    RhoP<S> cc{rn[c], pp[c]}, xp{rn[cxp], pp[cxp]}, yp{rn[cyp], pp[cyp]}, zp{rn[czp], pp[czp]};
    S uc = u[c], vc = v[c], wc = w[c];
    const auto rho0 = slot(f.base, kBaseRho0, NZ);
    double r0 = rho0[k], r0p = rho0[k < NZ - 1 ? k + 1 : k];
    pin(cc.r); pin(cc.p); pin(xp.r); pin(xp.p); pin(yp.r); pin(yp.p); pin(zp.r); pin(zp.p);
    pin(uc); pin(vc); pin(wc); pin(r0); pin(r0p);
    momentum_update<S>(g, r0, r0p, A.p.bounds, A.p.dts, i, j, k, cc, xp, yp, zp, double(uc), double(vc), double(wc),
                       A.p.acoustic_mode == NIMBUS_ACOUSTIC_HEVI, u, v, w, c);
}

// ---------------------------------------------------------------------
// Boundary closure (boundary_conditions_cartesian_cgrid.cpp:798-1175):
// west/east faces, then south/north, then the vertical closure; KW78
// radiative faces integrate their FP32 stash when radiating.
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
struct BcBufs
{
    gptr<S> u, v, w, rp, pp;
    float dt_rad;
};

// Buffers current after substep s (s < 0: home buffers, no radiation):
// u, v, w live in their home buffers, rho' in buffer (s + 1) & 1.
template <class S>
__device__ __forceinline__ bool bc_buffers(const MemberArgs& A, int s, BcBufs<S>& b)
{
    const DevFields& f = A.f;
    if (s >= 0 && s >= A.p.substeps) return false;
    const int par = s >= 0 ? ((s + 1) & 1) : 0;
    b.u = P<S>(f.u);
    b.v = P<S>(f.v);
    b.w = P<S>(f.w);
    b.rp = P<S>(par ? f.rp2 : f.rp);
    b.pp = P<S>(f.pp);
    b.dt_rad = s >= 0 ? float(A.p.dts) : 0.0f;
    return true;
}

// One lateral boundary pair of one row: the low face (ghost cell c, first
// interior cell in) and the high face (interior face cell a, the one inside
// it am, ghost cell gi) for the face-normal velocity n and the tangential
// t1, t2. Every value is loaded before anything is stored, so the row costs
// one round trip (periodic copies read a and in, which the other face does
// not modify before, in the reference's order, they are read).
template <class S>
__device__ __forceinline__ void lateral_pair(gptr<S> n, gptr<S> t1, gptr<S> t2, gptr<S> rp, gptr<S> pp,
                                             gptr<float> stash_lo, gptr<float> stash_hi, int bc_lo, int bc_hi,
                                             bool radiate, float dt_rad, float inv_d, std::size_t c,
                                             std::size_t in, std::size_t am, std::size_t a, std::size_t gi)
{
    const float cstar = 30.0f;
    S n_in = n[in], t1_in = t1[in], t2_in = t2[in], rp_in = rp[in], pp_in = pp[in];
    S n_a = n[a], t1_a = t1[a], t2_a = t2[a], rp_a = rp[a], pp_a = pp[a], n_am = n[am];
    float s_lo = *stash_lo, s_hi = *stash_hi;
    pin(n_in); pin(t1_in); pin(t2_in); pin(rp_in); pin(pp_in);
    pin(n_a); pin(t1_a); pin(t2_a); pin(rp_a); pin(pp_a); pin(n_am); pin(s_lo); pin(s_hi);
    // low face
    if (bc_lo == NIMBUS_BC_PERIODIC)
    {
        n[c] = n_a; t1[c] = t1_a; t2[c] = t2_a; rp[c] = rp_a; pp[c] = pp_a;
    }
    else
    {
        if (bc_lo == NIMBUS_BC_RADIATIVE)
        {
            if (radiate)
            {
                const float nn = kw78_face(s_lo, s_lo - 0.0f - cstar, float(n_in) - s_lo, dt_rad, inv_d, true);
                n[c] = S(nn);
                *stash_lo = nn;
            }
        }
        else n[c] = bc_lo == NIMBUS_BC_ZERO_GRADIENT ? n_in : S(0);
        const bool rigid = bc_lo == NIMBUS_BC_RIGID;
        t1[c] = rigid ? S(0) : t1_in;
        t2[c] = rigid ? S(0) : t2_in;
        rp[c] = rp_in;
        pp[c] = pp_in;
    }
    // high face
    if (bc_hi == NIMBUS_BC_PERIODIC)
    {
        n[gi] = n_in; t1[gi] = t1_in; t2[gi] = t2_in; rp[gi] = rp_in; pp[gi] = pp_in;
    }
    else
    {
        S na = n_a;
        if (bc_hi == NIMBUS_BC_RIGID || bc_hi == NIMBUS_BC_RIGID_SLIP) na = S(0);
        if (bc_hi == NIMBUS_BC_RADIATIVE && radiate)
        {
            const float nn = kw78_face(s_hi, s_hi - 0.0f + cstar, s_hi - float(n_am), dt_rad, inv_d, false);
            na = S(nn);
            *stash_hi = nn;
        }
        else if (bc_hi == NIMBUS_BC_ZERO_GRADIENT) na = n_am;
        if (bc_hi == NIMBUS_BC_RIGID || bc_hi == NIMBUS_BC_RIGID_SLIP || (bc_hi == NIMBUS_BC_RADIATIVE && radiate) ||
            bc_hi == NIMBUS_BC_ZERO_GRADIENT)
            n[a] = na;
        n[gi] = (bc_hi == NIMBUS_BC_RADIATIVE || bc_hi == NIMBUS_BC_ZERO_GRADIENT) ? na : S(0);
        const bool rigid = bc_hi == NIMBUS_BC_RIGID;
        t1[gi] = rigid ? S(0) : t1_a;
        t2[gi] = rigid ? S(0) : t2_a;
        rp[gi] = rp_a;
        pp[gi] = pp_a;
    }
}

template <class S>
__global__ __launch_bounds__(kBlock) void k_bc_west_east(const MemberArgs* __restrict__ ma,
                                                         const int2* __restrict__ map, int s)
{
    const Work wk = work_of(map);
    const MemberArgs& A = ma[wk.m];
    BcBufs<S> b;
    if (!bc_buffers<S>(A, s, b)) return;
    const DevGrid& g = A.g;
    const int t = wk.local * kBlock + int(threadIdx.x);
    if (t >= g.ny * g.nz) return;
    const int k = t / g.ny, j = t - k * g.ny;
    const int NX = g.nx, NZ = g.nz;
    const std::size_t row = g.at(0, j, k);
    const std::size_t st = std::size_t(j) * NZ + k;
    lateral_pair<S>(b.u, b.v, b.w, b.rp, b.pp, G(A.f.rad_w) + st, G(A.f.rad_e) + st, A.p.bc_west, A.p.bc_east,
                    b.dt_rad > 0.0f, b.dt_rad, 1.0f / float(g.dx), row, row + 1, row + (NX - 3), row + (NX - 2),
                    row + (NX - 1));
}

template <class S>
__global__ __launch_bounds__(kBlock) void k_bc_south_north(const MemberArgs* __restrict__ ma,
                                                           const int2* __restrict__ map, int s)
{
    const Work wk = work_of(map);
    const MemberArgs& A = ma[wk.m];
    BcBufs<S> b;
    if (!bc_buffers<S>(A, s, b)) return;
    const DevGrid& g = A.g;
    const int t = wk.local * kBlock + int(threadIdx.x);
    if (t >= g.nx * g.nz) return;
    const int k = t / g.nx, i = t - k * g.nx;
    const int NY = g.ny, NZ = g.nz;
    const std::size_t sy = std::size_t(g.px);
    const std::size_t c = g.at(i, 0, k);
    const std::size_t st = std::size_t(i) * NZ + k;
    lateral_pair<S>(b.v, b.u, b.w, b.rp, b.pp, G(A.f.rad_s) + st, G(A.f.rad_n) + st, A.p.bc_south, A.p.bc_north,
                    b.dt_rad > 0.0f, b.dt_rad, 1.0f / float(g.dy), c, c + sy, c + (NY - 3) * sy, c + (NY - 2) * sy,
                    c + (NY - 1) * sy);
}

template <class S>
__global__ __launch_bounds__(kBlock) void k_bc_vertical(const MemberArgs* __restrict__ ma,
                                                        const int2* __restrict__ map, int s)
{
    const Work wk = work_of(map);
    const MemberArgs& A = ma[wk.m];
    BcBufs<S> b;
    if (!bc_buffers<S>(A, s, b)) return;
    const DevGrid& g = A.g;
    const Column t = column_of(A, wk.local);
    if (t.i >= g.nx || t.j >= g.ny) return;
    const int NZ = g.nz;
    const std::size_t top = g.at(t.i, t.j, NZ - 1), tm = g.at(t.i, t.j, NZ - 2);
    const S wt = b.w[tm], ut = b.u[tm], vt = b.v[tm];  // loads before the stores
    if (A.p.bc_bottom == NIMBUS_BOTTOM_NO_SLIP)
    {
        b.u[g.at(t.i, t.j, 0)] = S(0);
        b.v[g.at(t.i, t.j, 0)] = S(0);
    }
    b.w[top] = A.p.bc_top == NIMBUS_TOP_OPEN ? wt : S(0);
    b.u[top] = ut;
    b.v[top] = vt;
}

template <class S>
__global__ __launch_bounds__(kBlock) void k_seed_radiation(const MemberArgs* __restrict__ ma,
                                                           const int2* __restrict__ map)
{
    const Work wk = work_of(map);
    const MemberArgs& A = ma[wk.m];
    const DevGrid& g = A.g;
    const int t = wk.local * kBlock + int(threadIdx.x);
    const int NZ = g.nz;
    const auto u = P<S>(A.f.u);
    const auto v = P<S>(A.f.v);
    if (t < g.ny * NZ)
    {
        const int j = t / NZ, k = t - j * NZ;
        G(A.f.rad_w)[t] = float(u[g.at(0, j, k)]);
        G(A.f.rad_e)[t] = float(u[g.at(g.nx - 2, j, k)]);
    }
    if (t < g.nx * NZ)
    {
        const int i = t / NZ, k = t - i * NZ;
        G(A.f.rad_s)[t] = float(v[g.at(i, 0, k)]);
        G(A.f.rad_n)[t] = float(v[g.at(i, g.ny - 2, k)]);
    }
}

// Moves rho' from buffer 1 to buffer 0 when the member's substep count
// leaves it there (n_done odd; n_done < 0 means the member's own count).
template <class S>
__global__ __launch_bounds__(kBlock) void k_commit(const MemberArgs* __restrict__ ma, const int2* __restrict__ map,
                                                   int n_done)
{
    const Work wk = work_of(map);
    const MemberArgs& A = ma[wk.m];
    const int n = n_done < 0 ? A.p.substeps : min(n_done, A.p.substeps);
    if ((n & 1) == 0) return;
    const Cell cl = cell_of(A, wk.local);
    if (!cl.valid) return;
    const std::size_t c = A.g.at(cl.i, cl.j, cl.k);
    P<S>(A.f.rp)[c] = P<S>(A.f.rp2)[c];
}

// ---------------------------------------------------------------------
// WENO5 retained-flux transport with the PD donor limiter
// (advection_cartesian.cpp:1941-2650): face flux pass, limiter pass,
// conservative update pass, each a column march. Scalars are the member's
// theta, qv, qc, qr slots; the result goes to the tth .. tqr slots.
// ---------------------------------------------------------------------
constexpr int kMaxScalars = 4;

template <class S>
__device__ __forceinline__ gptr<const S> scalar_in(const DevFields& f, int q)
{
    return P<S>(q == 0 ? f.th : q == 1 ? f.qv : q == 2 ? f.qc : f.qr);
}
template <class S>
__device__ __forceinline__ gptr<S> scalar_out(const DevFields& f, int q)
{
    return P<S>(q == 0 ? f.tth : q == 1 ? f.tqv : q == 2 ? f.tqc : f.tqr);
}
// Limiter factor slot of scalar q: the PD scalars take rl_qv, rl_qc, rl_qr in order.
__device__ __forceinline__ gptr<float> limiter_slot(const DevFields& f, int pd_mask, int q)
{
    int n = 0;
    for (int r = 0; r < q; ++r) n += (pd_mask >> r) & 1;
    return G(n == 0 ? f.rl_qv : n == 1 ? f.rl_qc : f.rl_qr);
}

// Face flux pass (advect_conservative_scalar_flux): each cell writes the raw
// flux through its own upper x, y and z faces, stored in the storage type
// like the reference's SFX/SFY/SFZ scratch.
template <class S>
__global__ __launch_bounds__(kBlock) void k_transport_flux(const MemberArgs* __restrict__ ma,
                                                           const int2* __restrict__ map, int n_scalars)
{
    using A_ = typename AccT<S>::type;
    const Work wk = work_of(map);
    const MemberArgs& A = ma[wk.m];
    const Cell cl = tile_cell_of(A, wk.local);
    if (!cl.valid) return;
    const DevGrid& g = A.g;
    const int NX = g.nx, NY = g.ny, NZ = g.nz;
    const int i = cl.i, j = cl.j, k = cl.k;
    const DevFields& f = A.f;
    const std::ptrdiff_t c = std::ptrdiff_t(g.at(i, j, k));
    const bool open_top = A.p.bc_top == NIMBUS_TOP_OPEN;
    const A_ mx = A_(P<S>(f.mfx)[c]), my = A_(P<S>(f.mfy)[c]), mz = A_(P<S>(f.mfz)[c]);
    std::ptrdiff_t xo[6], yo[6], zo[6];
#pragma unroll
    for (int o = 0; o < 6; ++o)
    {
        xo[o] = clampi(i + o - 2, 0, NX - 1) - i;
        yo[o] = std::ptrdiff_t(clampi(j + o - 2, 0, NY - 1) - j) * g.px;
        zo[o] = std::ptrdiff_t(clampi(k + o - 2, 0, NZ - 1) - k) * std::ptrdiff_t(g.plane());
    }
    // Phase 1: the 16 stencil values of every scalar (offsets clamped to the
    // domain, so every load is valid and none waits behind a branch).
    A_ v[kMaxScalars][16];
#pragma unroll
    for (int q = 0; q < kMaxScalars; ++q)
    {
        if (q >= n_scalars) break;
        const gptr<const S> qs = scalar_in<S>(f, q);
        v[q][0] = A_(qs[c]);
#pragma unroll
        for (int o = 0; o < 5; ++o)
        {
            const int oo = o < 2 ? o : o + 1;  // stencil points -2, -1, +1, +2, +3
            v[q][1 + o] = A_(qs[c + xo[oo]]);
            v[q][6 + o] = A_(qs[c + yo[oo]]);
            v[q][11 + o] = A_(qs[c + zo[oo]]);
        }
    }
    // Phase 2: one WENO5 face per direction and scalar (-2 -1 0 +1 +2 +3 about the face).
#pragma unroll
    for (int q = 0; q < kMaxScalars; ++q)
    {
        if (q >= n_scalars) break;
        const A_* x = &v[q][1];
        const A_* y = &v[q][6];
        const A_* z = &v[q][11];
        const A_ q0 = v[q][0];
        S sx = S(0), sy_ = S(0), sz_ = S(0);
        if (i <= NX - 2) sx = S(dev::face_flux6<A_>(x[0], x[1], q0, x[2], x[3], x[4], i, NX, mx));
        if (j <= NY - 2) sy_ = S(dev::face_flux6<A_>(y[0], y[1], q0, y[2], y[3], y[4], j, NY, my));
        if (k <= NZ - 2)
            sz_ = S(dev::face_flux6<A_>(z[0], z[1], q0, z[2], z[3], z[4], k, NZ, mz));
        else if (open_top)
            sz_ = S(mz * q0);
        P<S>(f.flux[3 * q])[c] = sx;
        P<S>(f.flux[3 * q + 1])[c] = sy_;
        P<S>(f.flux[3 * q + 2])[c] = sz_;
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
__global__ __launch_bounds__(kBlock) void k_transport_limiter(const MemberArgs* __restrict__ ma,
                                                              const int2* __restrict__ map, int n_scalars,
                                                              int pd_mask)
{
    const Work wk = work_of(map);
    const MemberArgs& A = ma[wk.m];
    const Cell cl = cell_of(A, wk.local);
    if (!cl.valid) return;
    const DevGrid& g = A.g;
    const int NX = g.nx, NY = g.ny, NZ = g.nz;
    const int i = cl.i, j = cl.j, k = cl.k;
    const DevFields& f = A.f;
    const std::size_t sy = std::size_t(g.px), sz = g.plane();
    const std::size_t c = g.at(i, j, k);
    const bool open_top = A.p.bc_top == NIMBUS_TOP_OPEN;
    const double rho_b = slot(f.base, kBaseRho0, NZ)[k] + double(P<S>(f.rb)[c]);
    // Every load is unconditional (a missing lower neighbour reads the cell
    // itself and is masked), so the loads of a cell issue back to back and no
    // branch on a loaded value sits between them.
    const std::size_t cxm = i >= 1 ? c - 1 : c, cym = j >= 1 ? c - sy : c, czm = k >= 1 ? c - sz : c;
    const bool hx = i <= NX - 2, hy = j <= NY - 2, hz = k <= NZ - 2 || (open_top && k == NZ - 1);
    const bool lx = i >= 1, ly = j >= 1, lz = k >= 1;
    S fl[kMaxScalars][6], qs[kMaxScalars];
#pragma unroll
    for (int q = 0; q < kMaxScalars; ++q)
    {
        if (q >= n_scalars) break;
        if (!((pd_mask >> q) & 1)) continue;
        const gptr<const S> fx = P<S>(f.flux[3 * q]);
        const gptr<const S> fy = P<S>(f.flux[3 * q + 1]);
        const gptr<const S> fz = P<S>(f.flux[3 * q + 2]);
        fl[q][0] = fx[c]; fl[q][1] = fx[cxm]; fl[q][2] = fy[c]; fl[q][3] = fy[cym];
        fl[q][4] = fz[c]; fl[q][5] = fz[czm];
        qs[q] = scalar_in<S>(f, q)[c];
    }
#pragma unroll
    for (int q = 0; q < kMaxScalars; ++q)
    {
        if (q >= n_scalars) break;
        if (!((pd_mask >> q) & 1)) continue;
        double out_rate = 0.0;
        out_rate += hx && fl[q][0] > S(0) ? double(fl[q][0]) * g.idx : 0.0;
        out_rate -= lx && fl[q][1] < S(0) ? double(fl[q][1]) * g.idx : 0.0;
        out_rate += hy && fl[q][2] > S(0) ? double(fl[q][2]) * g.idy : 0.0;
        out_rate -= ly && fl[q][3] < S(0) ? double(fl[q][3]) * g.idy : 0.0;
        out_rate += hz && fl[q][4] > S(0) ? double(fl[q][4]) * g.idz : 0.0;
        out_rate -= lz && fl[q][5] < S(0) ? double(fl[q][5]) * g.idz : 0.0;
        double r = 1.0;
        if (out_rate > 0.0) r = fmin(1.0, rho_b * fmax(double(qs[q]), 0.0) / A.p.dt / out_rate);
        limiter_slot(f, pd_mask, q)[c] = pd_store(r);
    }
}

// Conservative update (advect_conservative_update): q = (rho_begin q - dt div F_lim) / rho_end,
// outgoing-then-incoming for positive-definite scalars; ghost columns copy through.
template <class S>
__global__ __launch_bounds__(kBlock) void k_transport_update(const MemberArgs* __restrict__ ma,
                                                             const int2* __restrict__ map, int n_scalars,
                                                             int pd_mask)
{
    const Work wk = work_of(map);
    const MemberArgs& A = ma[wk.m];
    const Cell cl = tile_cell_of(A, wk.local);
    if (!cl.valid) return;
    const DevGrid& g = A.g;
    const int NX = g.nx, NY = g.ny, NZ = g.nz;
    const int i = cl.i, j = cl.j, k = cl.k;
    const DevFields& f = A.f;
    const std::size_t sy = std::size_t(g.px), sz = g.plane();
    const std::size_t c = g.at(i, j, k);
    if (!(i >= 1 && i < NX - 1 && j >= 1 && j < NY - 1))
    {
#pragma unroll
        for (int q = 0; q < kMaxScalars; ++q)
            if (q < n_scalars) scalar_out<S>(f, q)[c] = scalar_in<S>(f, q)[c];
        return;
    }
    const bool open_top = A.p.bc_top == NIMBUS_TOP_OPEN;
    const double dt = A.p.dt;
    const double r0 = slot(f.base, kBaseRho0, NZ)[k];
    const double rho_b = r0 + double(P<S>(f.rb)[c]);
    const double rho_e = r0 + double(P<S>(f.rp)[c]);
    const double den = rho_e > 0.05 ? rho_e : 0.05;
    const bool has_top = k <= NZ - 2 || (open_top && k == NZ - 1);
    const std::size_t czp = k < NZ - 1 ? c + sz : c, czm = k >= 1 ? c - sz : c;
    const double ix = g.idx, iy = g.idy, iz = g.idz;
    // Phase 1: every load of every scalar, unconditionally (interior column:
    // c +- 1 and c +- sy exist; vertical neighbours clamped; a scalar without
    // the limiter reads some limiter field and ignores it). Doing them before
    // any store lets them all be in flight at once.
    S fl[kMaxScalars][6], qs[kMaxScalars];
    float rr[kMaxScalars][7];
#pragma unroll
    for (int q = 0; q < kMaxScalars; ++q)
    {
        if (q >= n_scalars) break;
        const gptr<const S> fx = P<S>(f.flux[3 * q]);
        const gptr<const S> fy = P<S>(f.flux[3 * q + 1]);
        const gptr<const S> fz = P<S>(f.flux[3 * q + 2]);
        fl[q][0] = fx[c]; fl[q][1] = fx[c - 1]; fl[q][2] = fy[c]; fl[q][3] = fy[c - sy];
        fl[q][4] = fz[c]; fl[q][5] = fz[czm];
        qs[q] = scalar_in<S>(f, q)[c];
        const gptr<const float> rl = limiter_slot(f, pd_mask, q);
        rr[q][0] = rl[c]; rr[q][1] = rl[c + 1]; rr[q][2] = rl[c - 1]; rr[q][3] = rl[c + sy];
        rr[q][4] = rl[c - sy]; rr[q][5] = rl[czp]; rr[q][6] = rl[czm];
    }
    // Phase 2: limited face fluxes and the conservative update.
#pragma unroll
    for (int q = 0; q < kMaxScalars; ++q)
    {
        if (q >= n_scalars) break;
        const bool pd = (pd_mask >> q) & 1;
        // a face flux scaled by its donor cell's ratio (lo: the cell below/left
        // of the face, hi: the one above/right)
        auto lim = [&](S raw, float r_lo, float r_hi) -> double {
            const float r = raw > S(0) ? r_lo : r_hi;
            return pd && raw != S(0) ? double(raw) * double(S(r)) : double(raw);
        };
        const double sxp = lim(fl[q][0], rr[q][0], rr[q][1]);
        const double sxm = lim(fl[q][1], rr[q][2], rr[q][0]);
        const double syp = lim(fl[q][2], rr[q][0], rr[q][3]);
        const double sym = lim(fl[q][3], rr[q][4], rr[q][0]);
        const double szp = has_top ? lim(fl[q][4], rr[q][0], rr[q][5]) : 0.0;
        const double szm = k >= 1 ? lim(fl[q][5], rr[q][6], rr[q][0]) : 0.0;
        const double q0 = double(qs[q]);
        double rho_phi;
        if (pd)
        {
            // outgoing, then incoming (adding a signed zero leaves a sum unchanged)
            double outg = 0.0, inc = 0.0;
            outg += sxp >= 0.0 ? sxp * ix : 0.0;
            inc -= sxp >= 0.0 ? 0.0 : sxp * ix;
            outg -= sxm <= 0.0 ? sxm * ix : 0.0;
            inc += sxm <= 0.0 ? 0.0 : sxm * ix;
            outg += syp >= 0.0 ? syp * iy : 0.0;
            inc -= syp >= 0.0 ? 0.0 : syp * iy;
            outg -= sym <= 0.0 ? sym * iy : 0.0;
            inc += sym <= 0.0 ? 0.0 : sym * iy;
            outg += has_top && szp >= 0.0 ? szp * iz : 0.0;
            inc -= has_top && szp < 0.0 ? szp * iz : 0.0;
            outg -= k >= 1 && szm <= 0.0 ? szm * iz : 0.0;
            inc += k >= 1 && szm > 0.0 ? szm * iz : 0.0;
            rho_phi = (rho_b * q0 - dt * outg) + dt * inc;
        }
        else
        {
            const double div = (sxp - sxm) * ix + (syp - sym) * iy + (szp - szm) * iz;
            rho_phi = rho_b * q0 - dt * div;
        }
        scalar_out<S>(f, q)[c] = S(rho_phi / den);
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
__device__ KesslerOut<S> kessler_cell(double dz, const KesslerConst& kc, double dt, double p_full, double rho_c,
                                      double rho_up, S theta, S qv, S qc, S qr, S qr_up, bool has_up)
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
        const A factor = A(1) / (A(1) + A(kc.c_accr) * dtA * dev::fpow(fmax(A(0), A(qr)), A(0.875)));
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
            const A vent = (A(1.6) + A(124.9) * dev::fpow(rc, A(0.2046))) * dev::fpow(rc, A(0.525));
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
    const A dzA = A(dz);
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
__global__ __launch_bounds__(kBlock) void k_kessler_tendencies(const MemberArgs* __restrict__ ma,
                                                               const int2* __restrict__ map, double dt, S* dth,
                                                               S* dqv, S* dqc, S* dqr)
{
    const Work wk = work_of(map);
    const MemberArgs& A = ma[wk.m];
    const Cell cl = cell_of(A, wk.local);
    if (!cl.valid) return;
    const DevGrid& g = A.g;
    const int NX = g.nx, NY = g.ny, NZ = g.nz;
    const int i = cl.i, j = cl.j, k = cl.k;
    const DevFields& f = A.f;
    const std::size_t c = g.at(i, j, k);
    if (!(i >= 1 && i < NX - 1 && j >= 1 && j < NY - 1))
    {
        dth[c] = S(0); dqv[c] = S(0); dqc[c] = S(0); dqr[c] = S(0);
        return;
    }
    const auto rho0 = slot(f.base, kBaseRho0, NZ);
    const auto p0 = slot(f.base, kBaseP0, NZ);
    const auto th = P<S>(f.th), qv = P<S>(f.qv), qc = P<S>(f.qc), qr = P<S>(f.qr), rp = P<S>(f.rp),
            pp = P<S>(f.pp);
    const std::size_t sz = g.plane();
    const bool up = k + 1 < NZ;
    const KesslerOut<S> o =
        kessler_cell<S>(g.dz, A.p.kessler, dt, p0[k] + double(pp[c]), rho0[k] + double(rp[c]),
                        up ? rho0[k + 1] + double(rp[c + sz]) : 0.0, th[c], qv[c], qc[c], qr[c],
                        up ? qr[c + sz] : S(0), up);
    dth[c] = o.dth; dqv[c] = o.dqv; dqc[c] = o.dqc; dqr[c] = o.dqr;
}

// Fused microphysics commit: Kessler tendencies, bounded apply
// (microphysics_step.cpp:1090-1140), surface rain ledger
// (microphysics_step.cpp:1917-2024; kessler.cpp:533-561) and the EOS
// pressure projection of every cell (eos_apply.cpp:50-110). Subcycle n reads
// the transported slots (n even) or the home slots (n odd) and writes the
// other set, so the level above is still the subcycle's input.
// Sedimentation limit of a member's state before microphysics subcycle n
// (sedimentation_step_limit, microphysics_step.cpp:3501-3572): the minimum
// over physical cells of 0.8 dz / speed, speed = min(Vt_max,
// kessler_rain(qv + qc + qr, rho) sqrt(1.2 / rho)), folded into the member's
// clock by an unsigned 64-bit minimum of the (positive) FP64 bit patterns.
// Reads the set subcycle n reads: the transported slots for even n.
template <class S>
__global__ __launch_bounds__(kBlock) void k_micro_limit(const MemberArgs* __restrict__ ma,
                                                        const int2* __restrict__ map, int n)
{
    __shared__ double red[kBlock];
    const Work wk = work_of(map);
    const MemberArgs& A = ma[wk.m];
    if (!A.p.kessler_adaptive || n >= A.p.kessler_sub) return;
    const Cell cl = cell_of(A, wk.local);
    const DevGrid& g = A.g;
    const DevFields& f = A.f;
    double lim = __builtin_inf();
    if (cl.valid && cl.i >= 1 && cl.i < g.nx - 1 && cl.j >= 1 && cl.j < g.ny - 1)
    {
        const bool from_temp = (n & 1) == 0;
        const std::size_t c = g.at(cl.i, cl.j, cl.k);
        const double qt = double(P<S>(from_temp ? f.tqv : f.qv)[c]) + double(P<S>(from_temp ? f.tqc : f.qc)[c]) +
                          double(P<S>(from_temp ? f.tqr : f.qr)[c]);
        const double rho = slot(f.base, kBaseRho0, g.nz)[cl.k] + double(P<S>(f.rp)[c]);
        const double speed = fmin(A.p.kessler.vt_max, dev::kessler_rain_speed<double>(qt, rho) * sqrt(1.2 / rho));
        if (speed > 0.0) lim = 0.8 * g.dz / speed;
    }
    red[threadIdx.x] = lim;
    __syncthreads();
    for (int h = kBlock / 2; h > 0; h >>= 1)
    {
        if (int(threadIdx.x) < h) red[threadIdx.x] = fmin(red[threadIdx.x], red[threadIdx.x + h]);
        __syncthreads();
    }
    if (threadIdx.x == 0 && red[0] < __builtin_inf())
        atomicMin(reinterpret_cast<unsigned long long*>(f.clock + 2),
                  static_cast<unsigned long long>(__double_as_longlong(red[0])));
}

// Microphysics clock of each adaptive member before subcycle n
// (step_microphysics, microphysics_step.cpp:3700-3750): the step is
// min(remaining, limit), the whole interval when it fits under the limit, and
// 0 once the interval is done (the subcycle then passes the state through);
// the last launch of the budget takes whatever remains. Resets the limit
// slot for the next reduction. One thread per member.
__global__ __launch_bounds__(64) void k_micro_clock(const MemberArgs* __restrict__ ma, int n_members, int n)
{
    const int m = int(blockIdx.x) * 64 + int(threadIdx.x);
    if (m >= n_members) return;
    const MemberArgs& A = ma[m];
    if (!A.p.kessler_adaptive || n >= A.p.kessler_sub) return;
    double* clk = A.f.clock;
    const double dt = A.p.dt;
    const double elapsed = n == 0 ? 0.0 : clk[0];
    const double limit = clk[2];
    double h = 0.0;
    if (elapsed < dt) h = n == A.p.kessler_sub - 1 ? dt - elapsed : fmin(dt - elapsed, limit);
    clk[0] = elapsed + h;
    clk[1] = h;
    clk[2] = __builtin_inf();
}

template <class S>
__global__ __launch_bounds__(kBlock) void k_kessler_apply(const MemberArgs* __restrict__ ma,
                                                          const int2* __restrict__ map, int n)
{
    const Work wk = work_of(map);
    const MemberArgs& A = ma[wk.m];
    if (n >= A.p.kessler_sub) return;
    const Cell cl = cell_of(A, wk.local);
    if (!cl.valid) return;
    const DevGrid& g = A.g;
    const int NX = g.nx, NY = g.ny, NZ = g.nz;
    const int i = cl.i, j = cl.j, k = cl.k;
    const DevFields& f = A.f;
    const bool from_temp = (n & 1) == 0;
    const auto th_in = P<S>(from_temp ? f.tth : f.th);
    const auto qv_in = P<S>(from_temp ? f.tqv : f.qv);
    const auto qc_in = P<S>(from_temp ? f.tqc : f.qc);
    const auto qr_in = P<S>(from_temp ? f.tqr : f.qr);
    const auto th = P<S>(from_temp ? f.th : f.tth);
    const auto qv = P<S>(from_temp ? f.qv : f.tqv);
    const auto qc = P<S>(from_temp ? f.qc : f.tqc);
    const auto qr = P<S>(from_temp ? f.qr : f.tqr);
    const auto rp = P<S>(f.rp);
    const auto pp = P<S>(f.pp);
    const auto rho0 = slot(f.base, kBaseRho0, NZ);
    const auto p0 = slot(f.base, kBaseP0, NZ);
    const KesslerConst kc = A.p.kessler;
    const Bounds bd = A.p.bounds;
    // subcycle length: from the member's clock when it subcycles by its
    // state, else the whole step (one subcycle)
    const double h = A.p.kessler_adaptive ? G(f.clock)[1] : A.p.dt / A.p.kessler_sub;
    // with the SGS or the diffusion on, the projection follows them instead
    const bool do_eos = n == A.p.kessler_sub - 1 && A.p.sgs_n == 0 && A.p.dif_apply == 0;
    const std::size_t c = g.at(i, j, k), sz = g.plane();
    // all loads first (the level above clamped at the lid, unused there)
    const bool up = k + 1 < NZ;
    const std::size_t cu = up ? c + sz : c;
    S thn = th_in[c], qvn = qv_in[c], qcn = qc_in[c], qrn = qr_in[c];
    S rpc = rp[c], ppc = pp[c], rpu = rp[cu], qru = qr_in[cu];
    const double r0 = rho0[k], r0u = rho0[up ? k + 1 : k], p0k = p0[k];
    pin(thn); pin(qvn); pin(qcn); pin(qrn); pin(rpc); pin(ppc); pin(rpu); pin(qru);
    const double rho = r0 + double(rpc);
    if (h > 0.0 && i >= 1 && i < NX - 1 && j >= 1 && j < NY - 1)
    {
        const KesslerOut<S> o = kessler_cell<S>(g.dz, kc, h, p0k + double(ppc), rho, up ? r0u + double(rpu) : 0.0,
                                                thn, qvn, qcn, qrn, up ? qru : S(0), up);
        if (k == 0)
        {
            const double q0 = double(qrn);
            if (q0 > 0.0)
            {
                double Vt = fmin(double(dev::kessler_rain_speed<double>(q0, rho)) * sqrt(1.2 / rho), kc.vt_max);
                const double sink = double(S(S(Vt) * qrn)) / g.dz;
                G(f.rain)[std::size_t(j) * NX + i] += sink * rho * g.dz * h;
            }
        }
        const double step = fmin(fmax(double(o.dth) * h, -50.0), 50.0);
        thn = S(clamp_store<S>(double(thn) + step, 300.0, 200.0f, 700.0f));
        qvn = S(clamp_store<S>(double(qvn) + double(o.dqv) * h, 0.0, 0.0f, 0.05f));
        qcn = S(clamp_store<S>(double(qcn) + double(o.dqc) * h, 0.0, 0.0f, 0.05f));
        qrn = S(clamp_store<S>(double(qrn) + double(o.dqr) * h, 0.0, 0.0f, 0.05f));
    }
    th[c] = thn; qv[c] = qvn; qc[c] = qcn; qr[c] = qrn;
    const double thv = double(thn);
    if (do_eos && rho > 0.0 && thv > 0.0)
    {
        const double pd = dev::eos_pressure(rho, thv);
        if (isfinite(pd) && pd > 0.0) pp[c] = S(clamp_store<S>(pd, 101325.0, bd.p_min, bd.p_max) - p0k);
    }
}

// Copies the transported-slot result home after an even number of subcycles.
template <class S>
__global__ __launch_bounds__(kBlock) void k_scalars_home(const MemberArgs* __restrict__ ma,
                                                         const int2* __restrict__ map)
{
    const Work wk = work_of(map);
    const MemberArgs& A = ma[wk.m];
    if (A.p.kessler_sub & 1) return;
    const Cell cl = cell_of(A, wk.local);
    if (!cl.valid) return;
    const DevFields& f = A.f;
    const std::size_t c = A.g.at(cl.i, cl.j, cl.k);
    P<S>(f.th)[c] = P<S>(f.tth)[c];
    P<S>(f.qv)[c] = P<S>(f.tqv)[c];
    P<S>(f.qc)[c] = P<S>(f.tqc)[c];
    P<S>(f.qr)[c] = P<S>(f.tqr)[c];
}

// ---------------------------------------------------------------------
// Smagorinsky SGS (smagorinsky.cpp:495-580; eddy_viscosity.cpp:1253-1527
// closure, 1532-1948 momentum, 1953-2090 scalars; sgs_apply.cpp:610-790
// commit), applied after the microphysics in sgs_n substeps, each three
// per-cell passes: closure (nu, K_theta, K_q into the dead k1 fields), then
// momentum and scalar tendencies, each committed with the storage clamps
// into the other of two buffer sets (home / scratch ping-pong), with the EOS
// projection after the last. FP64 arithmetic for FP64 storage, as the
// reference's CPU path; FP32 for the mixed mode. The per-edge stresses a
// reference column march carries from level to level are recomputed here
// from the same operands, so the values are identical.
// ---------------------------------------------------------------------
// EOS projection of one cell after the SGS commit (eos_apply.cpp:50-110).
// ---------------------------------------------------------------------
// Rayleigh sponges after the EOS projection (rayleigh_damping.cpp:985-1080,
// 1360-1440): upper sponge on u, v (to the base wind), w (to 0) and theta
// (to theta0), the w-only absorber under the lid, then the lateral band;
// each relaxation x += f (target - x) with the exact decay factor
// f = 1 - exp(-alpha dt) precomputed per level, rounded to storage. Members
// with SGS apply them in the pass that commits their last SGS substep;
// others in k_sponge.
// ---------------------------------------------------------------------
struct SpongeAt
{
    double fu, fw, fwo, fl;  // upper sponge at centres and at w faces, w-only absorber, lateral band
    double u0, v0, th0;      // targets
};

__device__ __forceinline__ SpongeAt sponge_at(const MemberArgs& A, int i, int j, int k)
{
    const DevGrid& g = A.g;
    const int NZ = g.nz;
    const double* base = A.f.base;
    SpongeAt z;
    z.fu = slot(base, kBaseSpongeUp, NZ)[k];
    z.fw = slot(base, kBaseSpongeUpW, NZ)[k];
    z.fwo = slot(base, kBaseSpongeW, NZ)[k];
    const int dist = min(min(i, g.nx - 1 - i), min(j, g.ny - 1 - j));
    z.fl = dist < A.p.sponge_lateral ? G(A.f.sponge_lat)[dist] : 0.0;
    z.u0 = slot(base, kBaseU0, NZ)[k];
    z.v0 = slot(base, kBaseV0, NZ)[k];
    z.th0 = slot(base, kBaseTh0, NZ)[k];
    return z;
}

template <class S>
__device__ __forceinline__ void sponge_relax(S& x, double target, double fac)
{
    if (fac > 0.0) x = S(double(x) + fac * (target - double(x)));
}

// per field in the reference's order: upper sponge, w-only absorber, lateral band
template <class S>
__device__ __forceinline__ void sponge_wind(const SpongeAt& z, S& u, S& v, S& w)
{
    sponge_relax(u, z.u0, z.fu);
    sponge_relax(v, z.v0, z.fu);
    sponge_relax(w, 0.0, z.fw);
    sponge_relax(w, 0.0, z.fwo);
    sponge_relax(u, z.u0, z.fl);
    sponge_relax(v, z.v0, z.fl);
    sponge_relax(w, 0.0, z.fl);
}

template <class S>
__device__ __forceinline__ void sponge_theta(const SpongeAt& z, S& th)
{
    sponge_relax(th, z.th0, z.fu);
    sponge_relax(th, z.th0, z.fl);
}

template <class S>
__device__ __forceinline__ void sgs_eos(const MemberArgs& A, int c, int k, S theta)
{
    const DevFields& f = A.f;
    const double r0 = slot(f.base, kBaseRho0, A.g.nz)[k], p0k = slot(f.base, kBaseP0, A.g.nz)[k];
    const double rho = r0 + double(P<S>(f.rp)[c]);
    const double thv = double(theta);
    if (rho > 0.0 && thv > 0.0)
    {
        const double pd = dev::eos_pressure(rho, thv);
        if (isfinite(pd) && pd > 0.0)
            P<S>(f.pp)[c] = S(clamp_store<S>(pd, 101325.0, A.p.bounds.p_min, A.p.bounds.p_max) - p0k);
    }
}

template <class S>
struct SgsView
{
    using T = typename AccT<S>::type;  // FP32 arithmetic for FP32 storage, FP64 for FP64
    static constexpr int es = int(sizeof(S));
    gptr<const S> u, v, w, rho_p, th, qv;
    BufField<S> bu, bv, bw;            // the same u, v, w as buffers
    gptr<const double> rho0;
    int nx, ny, nz, px, plane;
    int wx, wy;
    // byte offsets of a column position (lane part) and of a level (uniform part)
    __device__ int xo(int i) const { return i * es; }
    __device__ int yo(int j) const { return j * px * es; }
    __device__ int zo(int k) const { return k * plane * es; }
    __device__ T Ub(int i, int j, int k) const { return T(bu(xo(i) + yo(j), zo(k))); }
    __device__ T Vb(int i, int j, int k) const { return T(bv(xo(i) + yo(j), zo(k))); }
    __device__ T Wb(int i, int j, int k) const { return T(bw(xo(i) + yo(j), zo(k))); }
    __device__ int at(int i, int j, int k) const { return k * plane + j * px + i; }
    __device__ int ip(int i) const { return wx ? (i <= 1 ? nx - 2 : min(i - 1, nx - 2)) : max(i - 1, 0); }
    __device__ int in(int i) const { return wx ? (i >= nx - 2 ? 1 : max(i + 1, 1)) : min(i + 1, nx - 1); }
    __device__ int jp(int j) const { return wy ? (j <= 1 ? ny - 2 : min(j - 1, ny - 2)) : max(j - 1, 0); }
    __device__ int jn(int j) const { return wy ? (j >= ny - 2 ? 1 : max(j + 1, 1)) : min(j + 1, ny - 1); }
    __device__ T U(int i, int j, int k) const { return T(u[at(i, j, k)]); }
    __device__ T V(int i, int j, int k) const { return T(v[at(i, j, k)]); }
    __device__ T W(int i, int j, int k) const { return k < 0 ? T(0) : T(w[at(i, j, k)]); }
    // live density: rho0 + rho', rounded to storage as the reference's whole-value field
    __device__ T R(int i, int j, int k) const { return T(S(rho0[k] + double(rho_p[at(i, j, k)]))); }
};

// SGS substep n reads u, v, w, theta, qv from the home fields when n is even
// and from the scratch set (u2, v2, w2, rp2, mfx: dead after transport) when
// n is odd, and writes the other set, so the commit is part of the tendency
// passes; an odd count leaves the state in the scratch set for k_sgs_home.
struct SgsSet
{
    void *u, *v, *w, *th, *qv;
};

__device__ __forceinline__ SgsSet sgs_set(const DevFields& f, bool scratch)
{
    return scratch ? SgsSet{f.u2, f.v2, f.w2, f.rp2, f.mfx} : SgsSet{f.u, f.v, f.w, f.th, f.qv};
}

template <class S>
__device__ SgsView<S> sgs_view(const MemberArgs& A, int n)
{
    const DevFields& f = A.f;
    const SgsSet cur = sgs_set(f, n & 1);
    SgsView<S> s;
    s.u = P<S>(cur.u); s.v = P<S>(cur.v); s.w = P<S>(cur.w); s.rho_p = P<S>(f.rp); s.th = P<S>(cur.th);
    s.qv = P<S>(cur.qv);
    const std::size_t fbytes = std::size_t(A.g.plane()) * A.g.nz * sizeof(S);
    s.bu = buf<S>(cur.u, fbytes);
    s.bv = buf<S>(cur.v, fbytes);
    s.bw = buf<S>(cur.w, fbytes);
    s.rho0 = slot(f.base, kBaseRho0, A.g.nz);
    s.nx = A.g.nx; s.ny = A.g.ny; s.nz = A.g.nz; s.px = A.g.px; s.plane = int(A.g.plane());
    s.wx = A.p.sgs_wrap_x;
    s.wy = A.p.sgs_wrap_y;
    return s;
}

// |S|^2 at the centre of cell (i, j, k) (eddy_viscosity.cpp:1253-1400). The
// neighbour indices are the reference's (clamped, or wrapped on periodic
// axes) and every value is loaded before the arithmetic; edge strains of
// edges on or above the lid and below the ground are zero.
template <class S, int Mode>
__device__ __forceinline__ typename AccT<S>::type sgs_strain_sq(const SgsView<S>& s, typename AccT<S>::type idx,
                                                                typename AccT<S>::type idy,
                                                                typename AccT<S>::type idz, int i, int j, int k)
{
    using T = typename AccT<S>::type;
    const int NZ = s.nz;
    const T h = T(0.5), q = T(0.25), two = T(2);
    const int im = s.ip(i), jm = s.jp(j);
    const int ii[2] = {im, i}, jj[2] = {jm, j};
    const bool kv[2] = {k >= 1, k <= NZ - 2};  // edges k-1/2 and k+1/2 exist
    const int kl[2] = {k >= 1 ? k - 1 : 0, k <= NZ - 2 ? k : NZ - 2};
    // Only the operands the mode's formula uses are loaded: horizontal
    // derivatives of u, v and w unless it is the vertical-only mode, vertical
    // derivatives of u, v, w unless it is the horizontal-only mode.
    constexpr bool hz = Mode != NIMBUS_SGS_VERTICAL, vt = Mode != NIMBUS_SGS_HORIZONTAL;
    T u0 = T(0), uw = T(0), v0 = T(0), vs = T(0), w0 = T(0), wb = T(0);
    T a12[2][2][4] = {};  // s12 operands: u(ii, jn(jj)), u(ii, jj), v(in(ii), jj), v(ii, jj)
    T xzv[2][2][4] = {};  // xz operands: u(ii, kk+1), u(ii, kk), w(in(ii), kk), w(ii, kk)
    T yzv[2][2][4] = {};  // yz operands: v(jj, kk+1), v(jj, kk), w(jn(jj), kk), w(jj, kk)
    if (hz)
    {
        u0 = s.Ub(i, j, k);
        uw = s.Ub(im, j, k);
        v0 = s.Vb(i, j, k);
        vs = s.Vb(i, jm, k);
#pragma unroll
        for (int x = 0; x < 2; ++x)
#pragma unroll
            for (int y = 0; y < 2; ++y)
            {
                a12[x][y][0] = s.Ub(ii[x], s.jn(jj[y]), k);
                a12[x][y][1] = s.Ub(ii[x], jj[y], k);
                a12[x][y][2] = s.Vb(s.in(ii[x]), jj[y], k);
                a12[x][y][3] = s.Vb(ii[x], jj[y], k);
            }
#pragma unroll
        for (int x = 0; x < 2; ++x)
#pragma unroll
            for (int z = 0; z < 2; ++z)
            {
                const int kk = kl[z];
                xzv[x][z][2] = s.Wb(s.in(ii[x]), j, kk);
                xzv[x][z][3] = s.Wb(ii[x], j, kk);
                yzv[x][z][2] = s.Wb(i, s.jn(jj[x]), kk);
                yzv[x][z][3] = s.Wb(i, jj[x], kk);
            }
    }
    if (vt)
    {
        w0 = s.Wb(i, j, k);
        wb = s.Wb(i, j, k >= 1 ? k - 1 : 0);
#pragma unroll
        for (int x = 0; x < 2; ++x)
#pragma unroll
            for (int z = 0; z < 2; ++z)
            {
                const int kk = kl[z];
                xzv[x][z][0] = s.Ub(ii[x], j, kk + 1);
                xzv[x][z][1] = s.Ub(ii[x], j, kk);
                yzv[x][z][0] = s.Vb(i, jj[x], kk + 1);
                yzv[x][z][1] = s.Vb(i, jj[x], kk);
            }
    }
    pin(u0); pin(uw); pin(v0); pin(vs); pin(w0); pin(wb);
#pragma unroll
    for (int x = 0; x < 2; ++x)
#pragma unroll
        for (int y = 0; y < 2; ++y)
#pragma unroll
            for (int o = 0; o < 4; ++o)
            {
                pin(a12[x][y][o]);
                pin(xzv[x][y][o]);
                pin(yzv[x][y][o]);
            }
    const T S11 = (u0 - uw) * idx;
    const T S22 = (v0 - vs) * idy;
    const T S33 = (w0 - (k >= 1 ? wb : T(0))) * idz;
    auto s12 = [&](int x, int y) {
        return h * ((a12[x][y][0] - a12[x][y][1]) * idy + (a12[x][y][2] - a12[x][y][3]) * idx);
    };
    struct G2
    {
        T a, b;
    };
    auto xz = [&](int x, int z) {
        if (!kv[z]) return G2{T(0), T(0)};
        return G2{(xzv[x][z][0] - xzv[x][z][1]) * idz, (xzv[x][z][2] - xzv[x][z][3]) * idx};
    };
    auto yz = [&](int y, int z) {
        if (!kv[z]) return G2{T(0), T(0)};
        return G2{(yzv[y][z][0] - yzv[y][z][1]) * idz, (yzv[y][z][2] - yzv[y][z][3]) * idy};
    };
    const T a = s12(0, 0), b = s12(0, 1), c = s12(1, 0), d = s12(1, 1);
    const T S12_sq = q * (a * a + b * b + c * c + d * d);
    // edges: (west or centre column) x (below k: z = 0, above: z = 1)
    const G2 xa = xz(0, 0), xb = xz(0, 1), xc = xz(1, 0), xd = xz(1, 1);
    const G2 ya = yz(0, 0), yb = yz(0, 1), yc = yz(1, 0), yd = yz(1, 1);
    auto avg_b = [&](G2 p, G2 r, G2 t, G2 z) { return q * (p.b * p.b + r.b * r.b + t.b * t.b + z.b * z.b); };
    auto avg_a = [&](G2 p, G2 r, G2 t, G2 z) { return q * (p.a * p.a + r.a * r.a + t.a * t.a + z.a * z.a); };
    if constexpr (Mode == NIMBUS_SGS_HORIZONTAL)
        return two * (S11 * S11 + S22 * S22) + T(4) * S12_sq + avg_b(xa, xb, xc, xd) + avg_b(ya, yb, yc, yd);
    if constexpr (Mode == NIMBUS_SGS_VERTICAL) return avg_a(xa, xb, xc, xd) + avg_a(ya, yb, yc, yd) + two * S33 * S33;
    auto sym = [&](G2 p) {
        const T e = h * (p.a + p.b);
        return e * e;
    };
    const T S13_sq = q * (sym(xa) + sym(xb) + sym(xc) + sym(xd));
    const T S23_sq = q * (sym(ya) + sym(yb) + sym(yc) + sym(yd));
    return two * (S11 * S11 + S22 * S22 + S33 * S33) + T(4) * (S12_sq + S13_sq + S23_sq);
}

// 1 / spacing in the arithmetic type: FP64 divides by the spacing as the
// reference does; FP32 multiplies by the rounded reciprocal.
template <class T>
struct Inv
{
    T x, y, z;
};

template <class S>
__global__ __launch_bounds__(kBlock) void k_sgs_closure(const MemberArgs* __restrict__ ma,
                                                        const int2* __restrict__ map, int n)
{
    using T = typename AccT<S>::type;
    const Work wk = work_of(map);
    const MemberArgs& A = ma[wk.m];
    if (n >= A.p.sgs_n) return;
    const Cell cl = cell_of(A, wk.local);
    if (!cl.valid) return;
    const SgsView<S> sv = sgs_view<S>(A, n);
    const int c = sv.at(cl.i, cl.j, cl.k);
    const T idx = T(A.g.idx), idy = T(A.g.idy), idz = T(A.g.idz);
    T str;
    // one code path per mode (uniform over the workgroup)
    if (A.p.sgs_mode == NIMBUS_SGS_HORIZONTAL)
        str = sgs_strain_sq<S, NIMBUS_SGS_HORIZONTAL>(sv, idx, idy, idz, cl.i, cl.j, cl.k);
    else if (A.p.sgs_mode == NIMBUS_SGS_VERTICAL)
        str = sgs_strain_sq<S, NIMBUS_SGS_VERTICAL>(sv, idx, idy, idz, cl.i, cl.j, cl.k);
    else
        str = sgs_strain_sq<S, NIMBUS_SGS_3D>(sv, idx, idy, idz, cl.i, cl.j, cl.k);
    const double nu = fmin(A.p.sgs_c2d2 * sqrt(double(str)), A.p.sgs_nu_max);
    P<S>(A.f.k1u)[c] = S(nu);
    P<S>(A.f.k1v)[c] = S(fmin(nu * A.p.sgs_inv_pr, A.p.sgs_k_max));
    P<S>(A.f.k1w)[c] = S(fmin(nu * A.p.sgs_inv_sc, A.p.sgs_k_max));
}

// Momentum tendencies (eddy_viscosity.cpp:1532-1948), committed into the
// other buffer set (sgs_set) with the storage clamps. Every
// operand is one of the 27 cells around (i, j, k) reached by the reference's
// neighbour rules (west iw = prev(i), east ie = next(i), likewise in y; next
// of prev is the cell itself on every face these loops visit); the 59 the
// formulas use are loaded before any arithmetic.
template <class S, int Mode>
__device__ __forceinline__ void sgs_momentum_cell(const MemberArgs& A, const Cell& cl, int n)
{
    using T = typename AccT<S>::type;
    const SgsView<S> s = sgs_view<S>(A, n);
    const int NX = s.nx, NY = s.ny, NZ = s.nz;
    const int i = cl.i, j = cl.j, k = cl.k;
    const int c = s.at(i, j, k);
    const bool has_u = i >= 1 && i <= (s.wx ? NX - 2 : NX - 3) && j >= 1 && j <= NY - 2 && k < NZ - 1;
    const bool has_v = i >= 1 && i <= NX - 2 && j >= 1 && j <= NY - 2 && (s.wy || j < NY - 2) && k < NZ - 1;
    const bool has_w = i >= 1 && i <= NX - 2 && j >= 1 && j <= NY - 2 && k < NZ - 1;
    const SgsSet nxt = sgs_set(A.f, !(n & 1));
    const double hstep = n == A.p.sgs_n - 1 ? A.p.sgs_h_last : A.p.sgs_h;
    const Bounds& bd = A.p.bounds;
    // commit x + h dx/dt with the storage clamps (sgs_apply.cpp:610-660); the
    // tendency is rounded to storage first, as the reference's tendency field
    // the last substep of an even count lands in the home fields: the sponges follow it here
    const bool sponge_here = A.p.sponge && n == A.p.sgs_n - 1 && (A.p.sgs_n & 1) == 0 && A.p.dif_apply == 0;
    auto commit = [&](T du, T dv, T dw) {
        S un = S(clamp_store<S>(double(s.u[c]) + finite0(double(S(du))) * hstep, 0.0, -bd.wind_h, bd.wind_h));
        S vn = S(clamp_store<S>(double(s.v[c]) + finite0(double(S(dv))) * hstep, 0.0, -bd.wind_h, bd.wind_h));
        S wn = S(clamp_store<S>(double(s.w[c]) + finite0(double(S(dw))) * hstep, 0.0, -bd.wind_v, bd.wind_v));
        if (sponge_here) sponge_wind<S>(sponge_at(A, i, j, k), un, vn, wn);
        P<S>(nxt.u)[c] = un;
        P<S>(nxt.v)[c] = vn;
        P<S>(nxt.w)[c] = wn;
    };
    if (!(has_u || has_v || has_w))
    {
        commit(T(0), T(0), T(0));
        return;
    }
    const int iw = s.ip(i), ie = s.in(i), js = s.jp(j), jn = s.jn(j);
    const int km = k > 0 ? k - 1 : k, kp = k + 1;  // k <= NZ - 2 here
    const bool lo = k > 0;                          // a face below exists
    // cell offsets: X[0..2] = iw, i, ie; Y[0..2] = js, j, jn; Z[0..2] = km, k, kp
    // (bytes; the level is uniform over the workgroup, so the plane part is scalar)
    constexpr int es = int(sizeof(S));
    const int xs[3] = {iw * es, i * es, ie * es}, ys[3] = {js * s.px * es, j * s.px * es, jn * s.px * es};
    const int zs[3] = {km * s.plane * es, k * s.plane * es, kp * s.plane * es};
    const std::size_t fbytes = std::size_t(s.plane) * NZ * es;
    const BufField<S> bu = s.bu, bv = s.bv, bw = s.bw;
    const BufField<S> bn = buf<S>(A.f.k1u, fbytes), br = buf<S>(A.f.rp, fbytes);
    auto ld = [&](const BufField<S>& f, int a, int b, int cc) { return T(f(xs[a] + ys[b], zs[cc])); };
    // U, V, W at the nine cells each equation uses; nu and rho' at sixteen
    T u_ = ld(bu, 1, 1, 1), u_e = ld(bu, 2, 1, 1), u_w = ld(bu, 0, 1, 1);
    T u_n = ld(bu, 1, 2, 1), u_s = ld(bu, 1, 0, 1), u_wn = ld(bu, 0, 2, 1);
    // Mode < 0: the member's mode at run time (one code path for all modes)
    const int mode = Mode < 0 ? A.p.sgs_mode : Mode;
    const bool use_h = mode != NIMBUS_SGS_VERTICAL, use_v = mode != NIMBUS_SGS_HORIZONTAL;
    // operands of vertical derivatives only: not loaded in the horizontal-only mode
    T u_t = T(0), u_b = T(0), u_wt = T(0), v_t = T(0), v_b = T(0), v_st = T(0);
    T w_t = T(0), w_b = T(0), w_eb = T(0), w_nb = T(0);
    if (use_v)
    {
        u_t = ld(bu, 1, 1, 2);
        u_b = ld(bu, 1, 1, 0);
        u_wt = ld(bu, 0, 1, 2);
        v_t = ld(bv, 1, 1, 2);
        v_b = ld(bv, 1, 1, 0);
        v_st = ld(bv, 1, 0, 2);
        w_t = ld(bw, 1, 1, 2);
        w_b = ld(bw, 1, 1, 0);
        w_eb = ld(bw, 2, 1, 0);
        w_nb = ld(bw, 1, 2, 0);
    }
    T v_ = ld(bv, 1, 1, 1), v_n = ld(bv, 1, 2, 1), v_s = ld(bv, 1, 0, 1);
    T v_e = ld(bv, 2, 1, 1), v_w = ld(bv, 0, 1, 1), v_es = ld(bv, 2, 0, 1);
    T w_ = ld(bw, 1, 1, 1), w_e = ld(bw, 2, 1, 1), w_w = ld(bw, 0, 1, 1);
    T w_n = ld(bw, 1, 2, 1), w_s = ld(bw, 1, 0, 1);
    constexpr int kMu = 16;
    const int mu_pts[kMu][3] = {{1, 1, 1}, {2, 1, 1}, {0, 1, 1}, {1, 2, 1}, {1, 0, 1}, {2, 2, 1}, {2, 0, 1}, {0, 2, 1},
                                {1, 1, 2}, {1, 1, 0}, {2, 1, 2}, {2, 1, 0}, {0, 1, 2}, {1, 2, 2}, {1, 2, 0}, {1, 0, 2}};
    T nu_[kMu], rp_[kMu];
#pragma unroll
    for (int q = 0; q < kMu; ++q)
    {
        nu_[q] = T(0);
        rp_[q] = T(0);
        // the three cells below (b, eb, nb) enter only vertical-derivative terms
        if (use_v || mu_pts[q][2] != 0)
        {
            nu_[q] = ld(bn, mu_pts[q][0], mu_pts[q][1], mu_pts[q][2]);
            rp_[q] = ld(br, mu_pts[q][0], mu_pts[q][1], mu_pts[q][2]);
        }
    }
    const double r0k = s.rho0[k], r0m = s.rho0[km], r0p = s.rho0[kp];
    pin(u_); pin(u_e); pin(u_w); pin(u_n); pin(u_s); pin(u_wn); pin(u_t); pin(u_b); pin(u_wt);
    pin(v_); pin(v_n); pin(v_s); pin(v_e); pin(v_w); pin(v_es); pin(v_t); pin(v_b); pin(v_st);
    pin(w_); pin(w_e); pin(w_w); pin(w_n); pin(w_s); pin(w_t); pin(w_b); pin(w_eb); pin(w_nb);
    // live density and rho nu at the sixteen cells (density rounded to storage as the reference's field)
    T R_[kMu], M_[kMu];
#pragma unroll
    for (int q = 0; q < kMu; ++q)
    {
        const double r0 = mu_pts[q][2] == 1 ? r0k : (mu_pts[q][2] == 2 ? r0p : r0m);
        R_[q] = T(S(r0 + double(rp_[q])));
        M_[q] = R_[q] * nu_[q];
    }
    // named mu = rho nu: c, e, w, n, s, en, es, wn, t, b, et, eb, wt, nt, nb, st
    const T Mc = M_[0], Me = M_[1], Mw = M_[2], Mn = M_[3], Ms = M_[4], Men = M_[5], Mes = M_[6], Mwn = M_[7];
    const T Mt = M_[8], Mb = M_[9], Met = M_[10], Meb = M_[11], Mwt = M_[12], Mnt = M_[13], Mnb = M_[14], Mst = M_[15];
    const T dx = T(A.g.dx), dy = T(A.g.dy), dz = T(A.g.dz);
    const T h = T(0.5), q4 = T(0.25), two = T(2);
    // x / d: a division in FP64 (as the reference), a multiply by 1/d in FP32
    const T rdx = T(1) / dx, rdy = T(1) / dy, rdz = T(1) / dz;
    auto ovx = [&](T x) { return sizeof(T) == sizeof(double) ? x / dx : x * rdx; };
    auto ovy = [&](T x) { return sizeof(T) == sizeof(double) ? x / dy : x * rdy; };
    auto ovz = [&](T x) { return sizeof(T) == sizeof(double) ? x / dz : x * rdz; };
    const T wb = lo ? w_b : T(0);   // w below the cell (zero at the ground)
    const T web = lo ? w_eb : T(0);
    const T wnb = lo ? w_nb : T(0);
    // the reference's thickness-weighted mean of two levels; FP32 multiplies by 1 / (2 dz)
    const T r2dz = T(1) / (dz + dz);
    auto vert_mean = [&](T lower, T upper) {
        return sizeof(T) == sizeof(double) ? (dz * lower + dz * upper) / (dz + dz) : (dz * lower + dz * upper) * r2dz;
    };
    // quotient: a division in FP64, the hardware reciprocal and a multiply in FP32
    auto quot = [](T a, T b) {
        if constexpr (sizeof(T) == sizeof(double))
            return a / b;
        else
            return a * __builtin_amdgcn_rcpf(b);
    };
    T du = T(0);
    if (has_u)
    {
        T div = T(0);
        if (use_h)
        {
            const T S11_r = ovx((u_e - u_)), S11_l = ovx((u_ - u_w));
            div += ovx((two * Me * S11_r - two * Mc * S11_l));
            const T T12n = two * (q4 * (Mc + Me + Mn + Men)) * (h * (ovy((u_n - u_)) + ovx((v_e - v_))));
            const T T12s = two * (q4 * (Ms + Mes + Mc + Me)) * (h * (ovy((u_ - u_s)) + ovx((v_es - v_s))));
            div += ovy((T12n - T12s));
        }
        if (use_v)
        {
            const T T13t = vert_mean(h * (Mc + Me), h * (Mt + Met)) *
                           (ovz((u_t - u_)) + (use_h ? ovx((w_e - w_)) : T(0)));
            const T T13b = lo ? vert_mean(h * (Mb + Meb), h * (Mc + Me)) *
                                    (ovz((u_ - u_b)) + (use_h ? ovx((web - wb)) : T(0)))
                              : T(0);
            div += ovz((T13t - T13b));
        }
        du = quot(div, h * (R_[0] + R_[1]));
    }
    T dv = T(0);
    if (has_v)
    {
        T div = T(0);
        if (use_h)
        {
            const T S22_n = ovy((v_n - v_)), S22_s = ovy((v_ - v_s));
            div += ovy((two * Mn * S22_n - two * Mc * S22_s));
            const T T21e = two * (q4 * (Mc + Me + Mn + Men)) * (h * (ovy((u_n - u_)) + ovx((v_e - v_))));
            const T T21w = two * (q4 * (Mw + Mc + Mwn + Mn)) * (h * (ovy((u_wn - u_w)) + ovx((v_ - v_w))));
            div += ovx((T21e - T21w));
        }
        if (use_v)
        {
            const T T23t = vert_mean(h * (Mc + Mn), h * (Mt + Mnt)) *
                           (ovz((v_t - v_)) + (use_h ? ovy((w_n - w_)) : T(0)));
            const T T23b = lo ? vert_mean(h * (Mb + Mnb), h * (Mc + Mn)) *
                                    (ovz((v_ - v_b)) + (use_h ? ovy((wnb - wb)) : T(0)))
                              : T(0);
            div += ovz((T23t - T23b));
        }
        dv = quot(div, h * (R_[0] + R_[3]));
    }
    T dw = T(0);
    if (has_w)
    {
        const T mass = h * (R_[0] * dz + R_[8] * dz);
        T force = T(0);
        if (use_v)
        {
            const T T33_t = two * Mt * (ovz((w_t - w_)));
            const T T33_b = two * Mc * (ovz((w_ - wb)));
            force += T33_t - T33_b;
        }
        if (use_h)
        {
            const T T31e = vert_mean(h * (Mc + Me), h * (Mt + Met)) *
                           ((use_v ? ovz((u_t - u_)) : T(0)) + ovx((w_e - w_)));
            const T T31w = vert_mean(h * (Mw + Mc), h * (Mwt + Mt)) *
                           ((use_v ? ovz((u_wt - u_w)) : T(0)) + ovx((w_ - w_w)));
            force += ovx(dz * (T31e - T31w));
            const T T32n = vert_mean(h * (Mc + Mn), h * (Mt + Mnt)) *
                           ((use_v ? ovz((v_t - v_)) : T(0)) + ovy((w_n - w_)));
            const T T32s = vert_mean(h * (Ms + Mc), h * (Mst + Mt)) *
                           ((use_v ? ovz((v_st - v_s)) : T(0)) + ovy((w_ - w_s)));
            force += ovy(dz * (T32n - T32s));
        }
        dw = quot(force, mass);
    }
    commit(du, dv, dw);
}

template <class S>
__global__ __launch_bounds__(kBlock) void k_sgs_momentum(const MemberArgs* __restrict__ ma,
                                                         const int2* __restrict__ map, int n)
{
    const Work wk = work_of(map);
    const MemberArgs& A = ma[wk.m];
    if (n >= A.p.sgs_n) return;
    const Cell cl = cell_of(A, wk.local);
    if (!cl.valid) return;
    // a single code path with the mode read at run time: per-mode copies of
    // this kernel raise its register count and measured slower
    sgs_momentum_cell<S, -1>(A, cl, n);
}

// Scalar tendencies of theta and qv (eddy_viscosity.cpp:1953-2090),
// committed into the other buffer set; the EOS projection after the last
// substep when it lands in the home fields.
template <class S, int Mode>
__device__ __forceinline__ void sgs_scalars_cell(const MemberArgs& A, const Cell& cl, int n)
{
    using T = typename AccT<S>::type;
    const SgsView<S> s = sgs_view<S>(A, n);
    const int NX = s.nx, NY = s.ny, NZ = s.nz;
    const int i = cl.i, j = cl.j, k = cl.k;
    const T dx = T(A.g.dx), dy = T(A.g.dy), dz = T(A.g.dz), h = T(0.5);
    const T rdx = T(1) / dx, rdy = T(1) / dy, rdz = T(1) / dz;
    auto ovx = [&](T x) { return sizeof(T) == sizeof(double) ? x / dx : x * rdx; };
    auto ovy = [&](T x) { return sizeof(T) == sizeof(double) ? x / dy : x * rdy; };
    auto ovz = [&](T x) { return sizeof(T) == sizeof(double) ? x / dz : x * rdz; };
    constexpr bool use_h = Mode != NIMBUS_SGS_VERTICAL, use_v = Mode != NIMBUS_SGS_HORIZONTAL;
    const int c = s.at(i, j, k);
    T dth = T(0), dqv = T(0);
    if (i >= 1 && i <= NX - 2 && j >= 1 && j <= NY - 2)
    {
        const bool ow = A.p.sgs_open[0], oe = A.p.sgs_open[1], os = A.p.sgs_open[2], on = A.p.sgs_open[3];
        const int jm = !s.wy && !os && j == 1 ? j : s.jp(j);
        const int jpp = !s.wy && !on && j == NY - 2 ? j : s.jn(j);
        const int im = !s.wx && !ow && i == 1 ? i : s.ip(i);
        const int ipp = !s.wx && !oe && i == NX - 2 ? i : s.in(i);
        // the seven-point star, loaded before any arithmetic (vertical neighbours clamped, masked below)
        const int km = k > 0 ? k - 1 : k, kp = k < NZ - 1 ? k + 1 : k;
        // lane part and level part of each point's byte offset
        const int vo[7] = {s.xo(i) + s.yo(j),   s.xo(ipp) + s.yo(j), s.xo(im) + s.yo(j), s.xo(i) + s.yo(jpp),
                           s.xo(i) + s.yo(jm),  s.xo(i) + s.yo(j),   s.xo(i) + s.yo(j)};
        const int kk[7] = {k, k, k, k, k, kp, km};
        const std::size_t fbytes = std::size_t(s.plane) * NZ * sizeof(S);
        const SgsSet cur = sgs_set(A.f, n & 1);
        const BufField<S> bR = buf<S>(A.f.rp, fbytes), bK1 = buf<S>(A.f.k1v, fbytes), bK2 = buf<S>(A.f.k1w, fbytes);
        const BufField<S> bT = buf<S>(cur.th, fbytes), bQ = buf<S>(cur.qv, fbytes);
        T r[7], k1[7], k2[7], p1[7], p2[7];
#pragma unroll
        for (int o = 0; o < 7; ++o)
        {
            r[o] = k1[o] = k2[o] = p1[o] = p2[o] = T(0);
            // lateral points for the horizontal terms, vertical ones for the vertical terms
            if (o == 0 || (o <= 4 ? use_h : use_v))
            {
                const int so = s.zo(kk[o]);
                r[o] = T(S(s.rho0[kk[o]] + double(bR(vo[o], so))));
                k1[o] = T(bK1(vo[o], so));
                k2[o] = T(bK2(vo[o], so));
                p1[o] = T(bT(vo[o], so));
                p2[o] = T(bQ(vo[o], so));
            }
        }
        auto scalar = [&](const T* K, const T* P_) {
            T tr[7];
#pragma unroll
            for (int o = 0; o < 7; ++o) tr[o] = r[o] * K[o];
            T div = T(0);
            if (use_h)
            {
                const T Fe = ovx(h * (tr[0] + tr[1]) * (P_[1] - P_[0]));
                const T Fw = ovx(h * (tr[2] + tr[0]) * (P_[0] - P_[2]));
                const T Fn = ovy(h * (tr[0] + tr[3]) * (P_[3] - P_[0]));
                const T Fs = ovy(h * (tr[4] + tr[0]) * (P_[0] - P_[4]));
                div += ovx((Fe - Fw)) + ovy((Fn - Fs));
            }
            if (use_v)
            {
                const T Ft = k < NZ - 1 ? ovz(h * (tr[0] + tr[5]) * (P_[5] - P_[0])) : T(0);
                const T Fb = k > 0 ? ovz(h * (tr[6] + tr[0]) * (P_[0] - P_[6])) : T(0);
                div += ovz((Ft - Fb));
            }
            if constexpr (sizeof(T) == sizeof(double))
                return div / r[0];
            else
                return div * __builtin_amdgcn_rcpf(r[0]);
        };
        dth = scalar(k1, p1);
        dqv = scalar(k2, p2);
    }
    const SgsSet nxt = sgs_set(A.f, !(n & 1));
    const double hstep = n == A.p.sgs_n - 1 ? A.p.sgs_h_last : A.p.sgs_h;
    S thn = S(clamp_store<S>(double(s.th[c]) + finite0(double(S(dth))) * hstep, 300.0, 200.0f, 700.0f));
    P<S>(nxt.qv)[c] = S(clamp_store<S>(double(s.qv[c]) + finite0(double(S(dqv))) * hstep, 0.0, 0.0f, 0.05f));
    // an even count ends in the home fields: the EOS projection, then the
    // sponges, follow the last commit here
    if (n == A.p.sgs_n - 1 && (A.p.sgs_n & 1) == 0 && A.p.dif_apply == 0)
    {
        sgs_eos<S>(A, c, k, thn);
        if (A.p.sponge) sponge_theta<S>(sponge_at(A, i, j, k), thn);
    }
    P<S>(nxt.th)[c] = thn;
}

template <class S>
__global__ __launch_bounds__(kBlock) void k_sgs_scalars(const MemberArgs* __restrict__ ma,
                                                        const int2* __restrict__ map, int n)
{
    const Work wk = work_of(map);
    const MemberArgs& A = ma[wk.m];
    if (n >= A.p.sgs_n) return;
    const Cell cl = cell_of(A, wk.local);
    if (!cl.valid) return;
    // one code path per mode (uniform over the workgroup)
    if (A.p.sgs_mode == NIMBUS_SGS_HORIZONTAL)
        sgs_scalars_cell<S, NIMBUS_SGS_HORIZONTAL>(A, cl, n);
    else if (A.p.sgs_mode == NIMBUS_SGS_VERTICAL)
        sgs_scalars_cell<S, NIMBUS_SGS_VERTICAL>(A, cl, n);
    else
        sgs_scalars_cell<S, NIMBUS_SGS_3D>(A, cl, n);
}

// Odd SGS counts: the state of the last substep back to the home fields,
// then the EOS projection and the sponges (unless the diffusion follows).
template <class S>
__global__ __launch_bounds__(kBlock) void k_sgs_home(const MemberArgs* __restrict__ ma, const int2* __restrict__ map)
{
    const Work wk = work_of(map);
    const MemberArgs& A = ma[wk.m];
    if ((A.p.sgs_n & 1) == 0) return;
    const Cell cl = cell_of(A, wk.local);
    if (!cl.valid) return;
    const DevFields& f = A.f;
    const int c = int(A.g.at(cl.i, cl.j, cl.k));
    S u = P<S>(f.u2)[c], v = P<S>(f.v2)[c], w = P<S>(f.w2)[c], th = P<S>(f.rp2)[c];
    const S qv = P<S>(f.mfx)[c];
    if (A.p.dif_apply == 0) sgs_eos<S>(A, c, cl.k, th);
    if (A.p.sponge && A.p.dif_apply == 0)
    {
        const SpongeAt z = sponge_at(A, cl.i, cl.j, cl.k);
        sponge_wind<S>(z, u, v, w);
        sponge_theta<S>(z, th);
    }
    P<S>(f.u)[c] = u;
    P<S>(f.v)[c] = v;
    P<S>(f.w)[c] = w;
    P<S>(f.th)[c] = th;
    P<S>(f.qv)[c] = qv;
}

// Sponges of the members without SGS (the others apply them with their last SGS substep).
template <class S>
__global__ __launch_bounds__(kBlock) void k_sponge(const MemberArgs* __restrict__ ma, const int2* __restrict__ map)
{
    const Work wk = work_of(map);
    const MemberArgs& A = ma[wk.m];
    if (!A.p.sponge || A.p.sgs_n > 0 || A.p.dif_apply != 0) return;
    const Cell cl = cell_of(A, wk.local);
    if (!cl.valid) return;
    const SpongeAt z = sponge_at(A, cl.i, cl.j, cl.k);
    if (z.fu == 0.0 && z.fw == 0.0 && z.fwo == 0.0 && z.fl == 0.0) return;
    const DevFields& f = A.f;
    const std::size_t c = A.g.at(cl.i, cl.j, cl.k);
    const auto u = P<S>(f.u);
    const auto v = P<S>(f.v);
    const auto w = P<S>(f.w);
    const auto th = P<S>(f.th);
    S uc = u[c], vc = v[c], wc = w[c], tc = th[c];
    sponge_wind<S>(z, uc, vc, wc);
    sponge_theta<S>(z, tc);
    u[c] = uc;
    v[c] = vc;
    w[c] = wc;
    th[c] = tc;
}

// ---------------------------------------------------------------------
// Explicit Laplacian diffusion after the SGS (diffusion_step.cpp:679-1030;
// equations of explicit.cpp:52-87,119-194,263-339,1183-1433 on the C-grid
// control volumes): one pass writes the tendencies of the state (u, v, w into
// the dead Heun stage-1 slots, theta and qv into two dead face-flux slots), a
// second commits field + dt tendency with the storage clamps, then the EOS
// projection and the sponges, which the microphysics or SGS pass leaves to it.
// FP64 arithmetic; the vertical part is rounded to storage before the
// horizontal part is added, as the reference adds its two passes.
// ---------------------------------------------------------------------
// Addresses are offsets from the cell's own element (neighbour columns and
// rows precomputed once, with the periodic wrap), and every operand is loaded
// before the arithmetic: neighbours that a closed edge or the lid leaves out
// are read at clamped, in-domain positions and their terms selected away.
// The arithmetic is the reference's expression for expression.
template <class S>
__device__ __forceinline__ double dif_tendency(const MemberArgs& A, gptr<const S> fld, int loc, int i, int j,
                                               int k)
{
#pragma clang fp contract(off)
    const DevGrid& g = A.g;
    const int nx = g.nx, ny = g.ny, nz = g.nz;
    const bool px = A.p.dif_wrap_x != 0, py = A.p.dif_wrap_y != 0;
    int i0 = 1, i1 = nx - 1, j0 = 1, j1 = ny - 1, k0 = 0, k1 = loc == 0 ? nz : nz - 1;
    if (loc == 1) i1 = px ? nx - 1 : nx - 2;
    if (loc == 2) j1 = py ? ny - 1 : ny - 2;
    if (i < i0 || i >= i1 || j < j0 || j >= j1 || k < k0 || k >= k1) return 0.0;
    const std::ptrdiff_t sy = g.px, sz = std::ptrdiff_t(g.plane());
    const std::size_t c = g.at(i, j, k);
    const gptr<const double> rho0 = slot(A.f.base, kBaseRho0, nz);
    const gptr<const S> rp = P<S>(A.f.rp) + c;
    const gptr<const S> fq = fld + c;
    // neighbour columns (wrapped on periodic sides) and clamped levels
    const int iw = i == i0 && px ? i1 - 1 : i - 1, ie = i == i1 - 1 && px ? i0 : i + 1;
    const int js = j == j0 && py ? j1 - 1 : j - 1, jn = j == j1 - 1 && py ? j0 : j + 1;
    const std::ptrdiff_t ow = iw - i, oe = ie - i, os = std::ptrdiff_t(js - j) * sy, on = std::ptrdiff_t(jn - j) * sy;
    const int km = k > k0 ? k - 1 : k, kp = k < nz - 1 ? k + 1 : k;
    const std::ptrdiff_t dm = std::ptrdiff_t(km - k) * sz, dp = std::ptrdiff_t(kp - k) * sz;
    // the partner column of a u (v) control volume: the next one, wrapped
    auto xpart = [&](int a) -> std::ptrdiff_t { return px && a == nx - 2 ? std::ptrdiff_t(1 - a) : 1; };
    auto ypart = [&](int b) -> std::ptrdiff_t { return (py && b == ny - 2 ? std::ptrdiff_t(1 - b) : 1) * sy; };
    const double dz = g.dz;
    // live density at offset o (same level e as the offset's level part)
    auto rho = [&](std::ptrdiff_t o, int e) { return rho0[e] + double(rp[o]); };
    // control-volume density at column/row offset o of column a, row b, level e
    auto comp_rho = [&](std::ptrdiff_t o, int a, int b, int e) {
        const std::ptrdiff_t d = std::ptrdiff_t(e - k) * sz;
        if (loc == 0) return rho(o + d, e);
        if (loc == 1) return 0.5 * (rho(o + d, e) + rho(o + xpart(a) + d, e));
        if (loc == 2) return 0.5 * (rho(o + d, e) + rho(o + ypart(b) + d, e));
        return (rho(o + d, e) * dz + rho(o + d + sz, e + 1) * dz) / (dz + dz);
    };
    const double qc = double(fq[0]);
    S t = S(0);
    const double kv = A.p.dif_kv;
    if (kv > 0.0)
    {
        const double q_m = double(fq[dm]), q_p = double(fq[dp]);
        auto cell_tr = [&](std::ptrdiff_t o, int e) { return rho(o, e) * kv; };
        // vertical transmissibility of the control volume at level e (offset d)
        auto comp_tr = [&](std::ptrdiff_t d, int e) {
            if (loc == 1) return 0.5 * (cell_tr(d, e) + cell_tr(d + xpart(i), e));
            if (loc == 2) return 0.5 * (cell_tr(d, e) + cell_tr(d + ypart(j), e));
            return cell_tr(d, e);
        };
        const bool wc = loc == 3;
        const double dist = wc ? dz : 0.5 * (dz + dz);
        double tr_down, tr_up, mass;
        if (wc)
        {
            // faces of a w control volume are the cells k and k + 1
            tr_down = cell_tr(0, k);
            tr_up = cell_tr(sz, k + 1);
            mass = 0.5 * (rho(0, k) * dz + rho(sz, k + 1) * dz);
        }
        else
        {
            const double tm = comp_tr(dm, km), t0 = comp_tr(0, k), tp = comp_tr(dp, kp);
            tr_down = 0.5 * (tm + t0);
            tr_up = 0.5 * (t0 + tp);
            mass = comp_rho(0, i, j, k) * dz;
        }
        const double down = k > k0 ? tr_down * (qc - q_m) / dist : 0.0;
        const double up = k < k1 - 1 ? tr_up * (q_p - qc) / dist : 0.0;
        t = S(double(t) + (up - down) / mass);
    }
    const double kh = A.p.dif_kh;
    if (kh > 0.0)
    {
        const double q_w = double(fq[ow]), q_e = double(fq[oe]), q_s = double(fq[os]), q_n = double(fq[on]);
        const bool wcl = i == i0 && !px, ecl = i == i1 - 1 && !px;
        const bool scl = j == j0 && !py, ncl = j == j1 - 1 && !py;
        const double rc = comp_rho(0, i, j, k);
        const double cw = comp_rho(ow, iw, j, k), ce = comp_rho(oe, ie, j, k);
        const double cs = comp_rho(os, i, js, k), cn = comp_rho(on, i, jn, k);
        const double rw = wcl ? 0.0 : 0.5 * (rc + cw);
        const double re = ecl ? 0.0 : 0.5 * (rc + ce);
        const double rs = scl ? 0.0 : 0.5 * (rc + cs);
        const double rn = ncl ? 0.0 : 0.5 * (rc + cn);
        const double xd = ((ecl ? 0.0 : re * (q_e - qc)) - (wcl ? 0.0 : rw * (qc - q_w))) * A.p.dif_idx2;
        const double yd = ((ncl ? 0.0 : rn * (q_n - qc)) - (scl ? 0.0 : rs * (qc - q_s))) * A.p.dif_idy2;
        t = S(double(t) + kh * (xd + yd) / rc);
    }
    return double(t);
}

template <class S>
__global__ __launch_bounds__(kBlock) void k_diffusion_tend(const MemberArgs* __restrict__ ma,
                                                           const int2* __restrict__ map)
{
#pragma clang fp contract(off)
    const Work wk = work_of(map);
    const MemberArgs& A = ma[wk.m];
    if (A.p.dif_apply == 0) return;
    const Cell cl = cell_of(A, wk.local);
    if (!cl.valid) return;
    const DevFields& f = A.f;
    const std::size_t c = A.g.at(cl.i, cl.j, cl.k);
    if (A.p.dif_apply & NIMBUS_DIFFUSE_MOMENTUM)
    {
        P<S>(f.k1u)[c] = S(dif_tendency<S>(A, P<S>(f.u), 1, cl.i, cl.j, cl.k));
        P<S>(f.k1v)[c] = S(dif_tendency<S>(A, P<S>(f.v), 2, cl.i, cl.j, cl.k));
        P<S>(f.k1w)[c] = S(dif_tendency<S>(A, P<S>(f.w), 3, cl.i, cl.j, cl.k));
    }
    if (A.p.dif_apply & NIMBUS_DIFFUSE_SCALARS)
    {
        P<S>(f.flux[7])[c] = S(dif_tendency<S>(A, P<S>(f.th), 0, cl.i, cl.j, cl.k));
        P<S>(f.flux[8])[c] = S(dif_tendency<S>(A, P<S>(f.qv), 0, cl.i, cl.j, cl.k));
    }
}

template <class S>
__global__ __launch_bounds__(kBlock) void k_diffusion_commit(const MemberArgs* __restrict__ ma,
                                                             const int2* __restrict__ map)
{
#pragma clang fp contract(off)
    const Work wk = work_of(map);
    const MemberArgs& A = ma[wk.m];
    if (A.p.dif_apply == 0) return;
    const Cell cl = cell_of(A, wk.local);
    if (!cl.valid) return;
    const DevFields& f = A.f;
    const std::size_t c = A.g.at(cl.i, cl.j, cl.k);
    const Bounds& bd = A.p.bounds;
    const double dt = A.p.dt;
    S u = P<S>(f.u)[c], v = P<S>(f.v)[c], w = P<S>(f.w)[c], th = P<S>(f.th)[c];
    if (A.p.dif_apply & NIMBUS_DIFFUSE_MOMENTUM)
    {
        u = S(clamp_store<S>(double(u) + double(P<S>(f.k1u)[c]) * dt, 0.0, -bd.wind_h, bd.wind_h));
        v = S(clamp_store<S>(double(v) + double(P<S>(f.k1v)[c]) * dt, 0.0, -bd.wind_h, bd.wind_h));
        w = S(clamp_store<S>(double(w) + double(P<S>(f.k1w)[c]) * dt, 0.0, -bd.wind_v, bd.wind_v));
    }
    if (A.p.dif_apply & NIMBUS_DIFFUSE_SCALARS)
    {
        th = S(clamp_store<S>(double(th) + double(P<S>(f.flux[7])[c]) * dt, 300.0, 200.0f, 700.0f));
        P<S>(f.qv)[c] = S(clamp_store<S>(double(P<S>(f.qv)[c]) + double(P<S>(f.flux[8])[c]) * dt, 0.0, 0.0f, 0.05f));
    }
    sgs_eos<S>(A, int(c), cl.k, th);
    if (A.p.sponge)
    {
        const SpongeAt z = sponge_at(A, cl.i, cl.j, cl.k);
        sponge_wind<S>(z, u, v, w);
        sponge_theta<S>(z, th);
    }
    P<S>(f.u)[c] = u;
    P<S>(f.v)[c] = v;
    P<S>(f.w)[c] = w;
    P<S>(f.th)[c] = th;
}

// ---------------------------------------------------------------------
// Deterministic diagnostics: kDiagBlocksPerMember fixed-shape partials per
// member, then one workgroup per member reduces them in a fixed order, so a
// member's numbers do not depend on what else is in the batch.
// ---------------------------------------------------------------------
template <class S>
__global__ __launch_bounds__(kBlock) void k_diag_partials(const MemberArgs* __restrict__ ma,
                                                          const int2* __restrict__ map)
{
    __shared__ double sm[kDiagValues][kBlock];
    const Work wk = work_of(map);
    const MemberArgs& A = ma[wk.m];
    const DevGrid& g = A.g;
    const DevFields& f = A.f;
    const auto w = P<S>(f.w);
    const auto th = P<S>(f.th);
    const auto qc = P<S>(f.qc);
    const auto qr = P<S>(f.qr);
    double v[kDiagValues] = {-1e300, 1e300, 0.0, 0.0, 0.0, 0.0, -1e300, 0.0};
    const std::size_t ncell = std::size_t(g.nx) * g.ny * g.nz;
    const std::size_t stride = std::size_t(kDiagBlocksPerMember) * kBlock;
    for (std::size_t t = std::size_t(wk.local) * kBlock + threadIdx.x; t < ncell; t += stride)
    {
        const int k = int(t / (std::size_t(g.nx) * g.ny));
        const int rem = int(t - std::size_t(k) * g.nx * g.ny);
        const int j = rem / g.nx, i = rem - j * g.nx;
        if (!(i >= 1 && i < g.nx - 1 && j >= 1 && j < g.ny - 1)) continue;
        const std::size_t c = g.at(i, j, k);
        const double wv = double(w[c]);
        v[0] = fmax(v[0], wv);
        v[1] = fmin(v[1], wv);
        const double tp = double(th[c]) - slot(f.base, kBaseTh0, g.nz)[k];
        v[2] += tp * tp;
        v[4] += double(qc[c]);
        v[5] += double(qr[c]);
        if (k == 0)
        {
            const double r = G(f.rain)[std::size_t(j) * g.nx + i];
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
        for (int q = 0; q < kDiagValues; ++q) G(f.diag)[std::size_t(wk.local) * kDiagValues + q] = sm[q][0];
}

// One workgroup per member (blockIdx.x = member index).
__global__ void k_diag_final(const MemberArgs* __restrict__ ma, int n_members)
{
    if (threadIdx.x != 0 || int(blockIdx.x) >= n_members) return;
    const auto d = G(ma[blockIdx.x].f.diag);
    double v[kDiagValues] = {-1e300, 1e300, 0.0, 0.0, 0.0, 0.0, -1e300, 0.0};
    for (int b = 0; b < kDiagBlocksPerMember; ++b)
    {
        const auto p = d + std::size_t(b) * kDiagValues;
        v[0] = fmax(v[0], p[0]);
        v[1] = fmin(v[1], p[1]);
        v[6] = fmax(v[6], p[6]);
        v[2] += p[2];
        v[3] += p[3];
        v[4] += p[4];
        v[5] += p[5];
        v[7] += p[7];
    }
    for (int q = 0; q < kDiagValues; ++q) d[std::size_t(kDiagBlocksPerMember) * kDiagValues + q] = v[q];
}

// Host layout [i][j][k] <-> device [k][j][i], subtracting / adding a base slot.
// Grid: (ceil(nx / 64), ny, nz) workgroups of 64 threads.
template <class H, class S>
__global__ __launch_bounds__(64) void k_host_to_dev(DevGrid g, const H* src, S* dst, const double* base_slot)
{
    const int i = int(blockIdx.x) * 64 + int(threadIdx.x), j = int(blockIdx.y), k = int(blockIdx.z);
    if (i >= g.nx) return;
    const std::size_t h = (std::size_t(i) * g.ny + j) * g.nz + k;
    dst[g.at(i, j, k)] = S(double(src[h]) - (base_slot ? base_slot[k] : 0.0));
}

template <class S, class H>
__global__ __launch_bounds__(64) void k_dev_to_host(DevGrid g, const S* src, H* dst, const double* base_slot)
{
    const int i = int(blockIdx.x) * 64 + int(threadIdx.x), j = int(blockIdx.y), k = int(blockIdx.z);
    if (i >= g.nx) return;
    const std::size_t h = (std::size_t(i) * g.ny + j) * g.nz + k;
    dst[h] = H(double(src[g.at(i, j, k)]) + (base_slot ? base_slot[k] : 0.0));
}

}  // namespace

// ---------------------------------------------------------------------
// Host side: geometry, launch helpers and launchers
// ---------------------------------------------------------------------
void fill_geometry(MemberArgs& a, int tile_w, int k_chunks, int xcd_remap)
{
    MemberGeom& G = a.geo;
    G.linear = tile_w == 0 ? 1 : 0;
    G.runs = int((a.g.plane() + kBlock - 1) / kBlock);
    G.tile_w = tile_w == 32 ? 32 : 64;
    G.tile_h = kBlock / G.tile_w;
    G.tiles_x = (a.g.nx + G.tile_w - 1) / G.tile_w;
    G.tiles_y = (a.g.ny + G.tile_h - 1) / G.tile_h;
    G.k_chunks = std::max(1, std::min(k_chunks, a.g.nz));
    G.xcd_remap = xcd_remap ? 1 : 0;
}

int work_items(const MemberArgs& a, WorkKind kind)
{
    const MemberGeom& G = a.geo;
    const int tiles = G.tiles_x * G.tiles_y;
    const auto lin = [](std::size_t n) { return int((n + kBlock - 1) / kBlock); };
    switch (kind)
    {
    case W_CELL:
        return G.linear ? G.runs * a.g.nz : tiles * G.k_chunks * ((a.g.nz + G.k_chunks - 1) / G.k_chunks);
    case W_COLUMN: return G.linear ? int((std::size_t(a.g.nx) * a.g.ny + kBlock - 1) / kBlock) : tiles;
    case W_WE: return lin(std::size_t(a.g.ny) * a.g.nz);
    case W_SN: return lin(std::size_t(a.g.nx) * a.g.nz);
    case W_SEED: return lin(std::size_t(std::max(a.g.nx, a.g.ny)) * a.g.nz);
    case W_DIAG: return kDiagBlocksPerMember;
    case W_SEGCOL: return ((a.g.nx + kSegCols - 1) / kSegCols) * a.g.ny;
    case W_TILECELL: return tiles * G.k_chunks * ((a.g.nz + G.k_chunks - 1) / G.k_chunks);
    default: return 0;
    }
}

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
    if (a && hipEventRecord(a, stream) != hipSuccess && error == hipSuccess) error = hipErrorUnknown;
}

void Launcher::end(const char* name, hipEvent_t a)
{
    if (launches) ++*launches;
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess && error == hipSuccess) error = e;
    if (!timing || !a) return;
    hipEvent_t b = take_event();
    if (!b) return;
    if (hipEventRecord(b, stream) != hipSuccess && error == hipSuccess) error = hipErrorUnknown;
    if (pending) pending->push_back(TimedLaunch{name, a, b});
}

namespace
{

template <class K, class... Args>
void launch_map_lds(Launcher& L, const char* name, K kernel, const BatchView& b, int subset, WorkKind kind,
                    std::size_t lds_bytes, Args... args)
{
    const BlockMap& m = b.maps[subset][kind];
    if (m.blocks <= 0) return;
    hipEvent_t ev;
    L.begin(name, ev);
    hipLaunchKernelGGL(kernel, dim3(unsigned(m.blocks)), dim3(kBlock), lds_bytes, L.stream, b.args, m.entries,
                       args...);
    L.end(name, ev);
}

template <class K, class... Args>
void launch_map(Launcher& L, const char* name, K kernel, const BatchView& b, int subset, WorkKind kind,
                Args... args)
{
    launch_map_lds(L, name, kernel, b, subset, kind, 0, args...);
}

// LDS bytes of the staged base-state profiles for the deepest member of a batch.
std::size_t profile_lds_bytes(const BatchView& b) { return std::size_t(kLdsProfiles) * b.max_nz * sizeof(double); }

// LDS of the partitioned column solve: the profiles, 12 doubles per thread
// (run ends, two buffers of reduced rows, last unknowns) and 2 pivot verdicts.
std::size_t segment_lds_bytes(const BatchView& b)
{
    return profile_lds_bytes(b) + std::size_t(12) * kBlock * sizeof(double) + std::size_t(2) * kBlock * sizeof(int);
}

template <class S>
S* PS(void* p)
{
    return static_cast<S*>(p);
}

}  // namespace

template <class S>
void launch_slow_rk2(Launcher& L, const BatchView& b)
{
    launch_map(L, "slow_stage1", k_slow<S, 0>, b, S_ALL, W_TILECELL, (S*)nullptr, (S*)nullptr, (S*)nullptr);
    launch_map(L, "slow_stage2", k_slow<S, 1>, b, S_ALL, W_TILECELL, (S*)nullptr, (S*)nullptr, (S*)nullptr);
}

template <class S>
void launch_slow_tendencies(Launcher& L, const BatchView& b, void* du, void* dv, void* dw)
{
    launch_map(L, "slow_tendencies", k_slow<S, 2>, b, S_ALL, W_TILECELL, PS<S>(du), PS<S>(dv), PS<S>(dw));
}

template <class S>
void launch_acoustic_substep(Launcher& L, const BatchView& b, int s, bool accumulate, bool init_at_first)
{
    launch_map(L, "acoustic_forward", k_acoustic_forward<S>, b, S_EXPLICIT, W_CELL, s, int(accumulate),
               int(init_at_first));
    launch_map(L, "acoustic_hevi_explicit", k_hevi_explicit<S>, b, S_HEVI, W_CELL, s, int(accumulate),
               int(init_at_first));
    if (b.any_hevi_seg)
        launch_map_lds(L, "acoustic_hevi_column_seg", k_hevi_column_seg<S>, b, S_HEVI, W_SEGCOL,
                       segment_lds_bytes(b), s, int(accumulate), int(init_at_first));
    if (b.any_hevi_serial)
        launch_map_lds(L, "acoustic_hevi_column", k_hevi_column<S>, b, S_HEVI, W_COLUMN, profile_lds_bytes(b), s,
                       int(accumulate), int(init_at_first));
    launch_map(L, "acoustic_momentum", k_momentum<S>, b, S_ALL, W_CELL, s);
}

template <class S>
void launch_boundaries(Launcher& L, const BatchView& b, int s)
{
    launch_map(L, "bc_west_east", k_bc_west_east<S>, b, S_ALL, W_WE, s);
    launch_map(L, "bc_south_north", k_bc_south_north<S>, b, S_ALL, W_SN, s);
    launch_map(L, "bc_vertical", k_bc_vertical<S>, b, S_ALL, W_COLUMN, s);
}

template <class S>
void launch_seed_radiation(Launcher& L, const BatchView& b)
{
    launch_map(L, "seed_radiation", k_seed_radiation<S>, b, S_ALL, W_SEED);
}

template <class S>
void launch_commit(Launcher& L, const BatchView& b, int n_done)
{
    // a whole step (n_done < 0) has nothing to commit when every member's substep count is even
    if (n_done >= 0 || b.any_odd_substeps) launch_map(L, "acoustic_commit", k_commit<S>, b, S_ALL, W_CELL, n_done);
}

template <class S>
void launch_transport(Launcher& L, const BatchView& b, int n_scalars, const int* pd, bool into_temp)
{
    (void)into_temp;
    int mask = 0;
    for (int q = 0; q < n_scalars; ++q) mask |= (pd[q] ? 1 : 0) << q;
    launch_map(L, "transport_flux", k_transport_flux<S>, b, S_ALL, W_TILECELL, n_scalars);
    if (mask) launch_map(L, "transport_limiter", k_transport_limiter<S>, b, S_ALL, W_CELL, n_scalars, mask);
    launch_map(L, "transport_update", k_transport_update<S>, b, S_ALL, W_TILECELL, n_scalars, mask);
}

template <class S>
void launch_kessler_tendencies(Launcher& L, const BatchView& b, double dt, void* dth, void* dqv, void* dqc,
                               void* dqr)
{
    launch_map(L, "kessler_tendencies", k_kessler_tendencies<S>, b, S_ALL, W_CELL, dt, PS<S>(dth), PS<S>(dqv),
               PS<S>(dqc), PS<S>(dqr));
}

template <class S>
void launch_large_step(Launcher& L, const BatchView& b)
{
    launch_slow_rk2<S>(L, b);
    launch_boundaries<S>(L, b, -1);
    launch_seed_radiation<S>(L, b);
    for (int s = 0; s < b.max_substeps; ++s)
    {
        launch_acoustic_substep<S>(L, b, s, true, true);
        launch_boundaries<S>(L, b, s);
    }
    launch_commit<S>(L, b, -1);
    const int pd[4] = {0, 1, 1, 1};
    launch_transport<S>(L, b, 4, pd, true);
    for (int n = 0; n < b.max_kessler_sub; ++n)
    {
        if (b.any_kessler_adaptive)
        {
            launch_map(L, "micro_limit", k_micro_limit<S>, b, S_ALL, W_CELL, n);
            hipEvent_t ev;
            L.begin("micro_clock", ev);
            hipLaunchKernelGGL(k_micro_clock, dim3(unsigned((b.n + 63) / 64)), dim3(64), 0, L.stream, b.args, b.n,
                               n);
            L.end("micro_clock", ev);
        }
        launch_map(L, "kessler_apply", k_kessler_apply<S>, b, S_ALL, W_CELL, n);
    }
    if (b.any_even_kessler_sub) launch_map(L, "scalars_home", k_scalars_home<S>, b, S_ALL, W_CELL);
    for (int n = 0; n < b.max_sgs_n; ++n)
    {
        launch_map(L, "sgs_closure", k_sgs_closure<S>, b, S_ALL, W_CELL, n);
        launch_map(L, "sgs_momentum", k_sgs_momentum<S>, b, S_ALL, W_CELL, n);
        launch_map(L, "sgs_scalars", k_sgs_scalars<S>, b, S_ALL, W_CELL, n);
    }
    if (b.any_odd_sgs_n) launch_map(L, "sgs_home", k_sgs_home<S>, b, S_ALL, W_CELL);
    if (b.any_diffusion)
    {
        launch_map(L, "diffusion_tendencies", k_diffusion_tend<S>, b, S_ALL, W_CELL);
        launch_map(L, "diffusion_commit", k_diffusion_commit<S>, b, S_ALL, W_CELL);
    }
    if (b.any_sponge_unfused) launch_map(L, "sponge", k_sponge<S>, b, S_ALL, W_CELL);
    launch_boundaries<S>(L, b, -1);
}

template <class S>
void launch_diagnostics(Launcher& L, const BatchView& b)
{
    launch_map(L, "diag_partials", k_diag_partials<S>, b, S_ALL, W_DIAG);
    hipEvent_t ev;
    L.begin("diag_final", ev);
    hipLaunchKernelGGL(k_diag_final, dim3(unsigned(b.n)), dim3(64), 0, L.stream, b.args, b.n);
    L.end("diag_final", ev);
}

template <class H, class S>
void launch_host_to_dev(Launcher& L, const DevGrid& g, const void* src, void* dst, const double* base_slot)
{
    hipEvent_t ev;
    L.begin("host_to_dev", ev);
    hipLaunchKernelGGL((k_host_to_dev<H, S>), dim3(unsigned((g.nx + 63) / 64), unsigned(g.ny), unsigned(g.nz)),
                       dim3(64), 0, L.stream, g, static_cast<const H*>(src), PS<S>(dst), base_slot);
    L.end("host_to_dev", ev);
}

template <class S, class H>
void launch_dev_to_host(Launcher& L, const DevGrid& g, const void* src, void* dst, const double* base_slot)
{
    hipEvent_t ev;
    L.begin("dev_to_host", ev);
    hipLaunchKernelGGL((k_dev_to_host<S, H>), dim3(unsigned((g.nx + 63) / 64), unsigned(g.ny), unsigned(g.nz)),
                       dim3(64), 0, L.stream, g, static_cast<const S*>(src), static_cast<H*>(dst), base_slot);
    L.end("dev_to_host", ev);
}

#define NIMBUS_INSTANTIATE(S)                                                                                    \
    template void launch_large_step<S>(Launcher&, const BatchView&);                                            \
    template void launch_slow_rk2<S>(Launcher&, const BatchView&);                                              \
    template void launch_slow_tendencies<S>(Launcher&, const BatchView&, void*, void*, void*);                  \
    template void launch_acoustic_substep<S>(Launcher&, const BatchView&, int, bool, bool);                     \
    template void launch_boundaries<S>(Launcher&, const BatchView&, int);                                       \
    template void launch_seed_radiation<S>(Launcher&, const BatchView&);                                        \
    template void launch_commit<S>(Launcher&, const BatchView&, int);                                           \
    template void launch_transport<S>(Launcher&, const BatchView&, int, const int*, bool);                      \
    template void launch_kessler_tendencies<S>(Launcher&, const BatchView&, double, void*, void*, void*, void*); \
    template void launch_diagnostics<S>(Launcher&, const BatchView&);                                           \
    template void launch_host_to_dev<float, S>(Launcher&, const DevGrid&, const void*, void*, const double*);   \
    template void launch_host_to_dev<double, S>(Launcher&, const DevGrid&, const void*, void*, const double*);  \
    template void launch_dev_to_host<S, float>(Launcher&, const DevGrid&, const void*, void*, const double*);   \
    template void launch_dev_to_host<S, double>(Launcher&, const DevGrid&, const void*, void*, const double*);

NIMBUS_INSTANTIATE(float)
NIMBUS_INSTANTIATE(double)

}  // namespace nimbus
