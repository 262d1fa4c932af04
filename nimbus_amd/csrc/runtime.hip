// Host runtime of the HIP backend: device context, resident ensemble members
// (device-owned state in [k][j][i] layout), batches of members that share
// every kernel launch (device member table + block maps), hipGraph capture of
// a batch's whole large step, per-kernel HIP-event timing and the C ABI
// declared in nimbus_hip.h.
#include "kernels.hpp"
#include "nimbus_internal.hpp"

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <chrono>
#include <limits>
#include <map>
#include <memory>
#include <sys/stat.h>
#include <string>
#include <vector>

namespace nimbus
{
namespace
{

thread_local std::string g_last_error;

int fail(int code, const std::string& msg)
{
    g_last_error = msg;
    return code;
}

int hip_fail(hipError_t e, const char* what)
{
    return fail(NIMBUS_EHIP, std::string(what) + ": " + hipGetErrorString(e));
}

#define NIMBUS_HIP_TRY(expr)                                    \
    do                                                          \
    {                                                           \
        const hipError_t e_ = (expr);                           \
        if (e_ != hipSuccess) return hip_fail(e_, #expr);       \
    } while (0)

int pitch_for(int nx, int elem)
{
    const int per = 64 / elem;  // 64-byte rows
    return (nx + per - 1) / per * per;
}

// Field slots of a member allocation, in order. The face-flux scratch of the
// scalar transport reuses seven slots that are dead while it runs (Heun
// stage-1 tendencies and the second acoustic buffers) plus five of its own.
enum FieldSlot
{
    F_U, F_V, F_W, F_RP, F_PP, F_TH, F_QV, F_QC, F_QR,
    F_U2, F_V2, F_W2, F_RP2,
    F_MFX, F_MFY, F_MFZ, F_RB,
    F_K1U, F_K1V, F_K1W,
    F_TTH, F_TQV, F_TQC, F_TQR,
    F_FLUX7, F_FLUX_END = F_FLUX7 + 5,
    F_COUNT = F_FLUX_END
};

std::size_t align256(std::size_t x) { return (x + 255) / 256 * 256; }

struct Layout
{
    std::size_t field, rl, tri, rad_w, rad_s, rain, base, diag, lat, clock, total;
};

// Lateral sponge band width in cells (0 when the sponge or its band is off).
int lateral_cells(const nimbus_step_config& c, const nimbus_grid& g)
{
    if (!c.sponge || !(c.sponge_lateral_alpha > 0.0)) return 0;
    return std::max(0, std::min(c.sponge_lateral_cells, std::max(g.nx, g.ny)));
}

Layout layout_for(const DevGrid& g, int elem, bool hevi, int lateral = 0)
{
    Layout L{};
    L.field = align256(g.cells() * std::size_t(elem));
    L.rl = align256(g.cells() * sizeof(float));
    // FP64 Thomas coefficients: only the sequential sweep keeps them in memory
    L.tri = hevi && !hevi_partitioned_always(elem, g.nz) ? align256(g.cells() * sizeof(double)) : 0;
    L.rad_w = align256(std::size_t(g.ny) * g.nz * sizeof(float));
    L.rad_s = align256(std::size_t(g.nx) * g.nz * sizeof(float));
    L.rain = align256(std::size_t(g.nx) * g.ny * sizeof(double));
    L.base = align256(std::size_t(kBaseCount) * g.nz * sizeof(double));
    L.diag = align256(std::size_t(kDiagBlocksPerMember + 1) * kDiagValues * sizeof(double));
    L.lat = align256(std::size_t(std::max(lateral, 1)) * sizeof(double));
    L.clock = align256(4 * sizeof(double));
    L.total = L.field * F_COUNT + 3 * L.rl + 2 * L.tri + 2 * L.rad_w + 2 * L.rad_s + L.rain + L.base + L.diag +
              L.lat + L.clock;
    return L;
}

DevGrid dev_grid(const nimbus_grid& g, int elem)
{
    DevGrid d{};
    d.nx = g.nx;
    d.ny = g.ny;
    d.nz = g.nz;
    d.px = pitch_for(g.nx, elem);
    d.dx = g.dx;
    d.dy = g.dy;
    d.dz = g.dz;
    d.idx = 1.0 / g.dx;
    d.idy = 1.0 / g.dy;
    d.idz = 1.0 / g.dz;
    return d;
}

// Default ordering of the per-cell work: vertical chunks of about 20 levels,
// so a column tile's levels run back to back while the neighbouring tiles
// in y (one tile row later) are still close enough to share the L2.
int default_k_chunks(const DevGrid& g, int tile_w)
{
    (void)tile_w;
    return std::max(1, (g.nz + 10) / 20);
}

bool valid_grid(const nimbus_grid& g)
{
    return g.nx >= 5 && g.ny >= 5 && g.nz >= 4 && g.nz <= kMaxLevels && std::isfinite(g.dx) && g.dx > 0 && std::isfinite(g.dy) && g.dy > 0 &&
           std::isfinite(g.dz) && g.dz > 0 && dev_grid(g, 4).cells() < (std::size_t(1) << 28);
}

bool valid_config(const nimbus_step_config& c)
{
    auto bc_ok = [](int b) { return b >= NIMBUS_BC_PERIODIC && b <= NIMBUS_BC_RIGID_SLIP; };
    return std::isfinite(c.dt) && c.dt > 0 && c.substeps >= 0 && c.substeps < 10000 &&
           (c.acoustic_mode == NIMBUS_ACOUSTIC_EXPLICIT || c.acoustic_mode == NIMBUS_ACOUSTIC_HEVI) &&
           std::isfinite(c.hevi_beta) && c.hevi_beta >= 0.0 && c.hevi_beta <= 1.0 && std::isfinite(c.coriolis_f) &&
           bc_ok(c.bc_west) && bc_ok(c.bc_east) && bc_ok(c.bc_south) && bc_ok(c.bc_north) &&
           (c.bc_bottom == NIMBUS_BOTTOM_FREE_SLIP || c.bc_bottom == NIMBUS_BOTTOM_NO_SLIP) &&
           (c.bc_top == NIMBUS_TOP_RIGID || c.bc_top == NIMBUS_TOP_OPEN) &&
           !(c.acoustic_mode == NIMBUS_ACOUSTIC_HEVI && c.bc_top == NIMBUS_TOP_OPEN) && c.vt_max > 0 &&
           std::isfinite(c.vt_max) && c.qc0 >= 0 && c.c_auto >= 0 && c.c_accr >= 0 && c.evap_eff >= 0 &&
           (!c.sponge || (std::isfinite(c.sponge_z_d) && std::isfinite(c.sponge_alpha) && c.sponge_alpha >= 0 &&
                          c.sponge_w_cells >= 0 && std::isfinite(c.sponge_w_alpha) && c.sponge_w_alpha >= 0 &&
                          c.sponge_lateral_cells >= 0 && std::isfinite(c.sponge_lateral_alpha) &&
                          c.sponge_lateral_alpha >= 0)) &&
           (!c.sgs || (c.sgs_mode >= NIMBUS_SGS_3D && c.sgs_mode <= NIMBUS_SGS_VERTICAL &&
                       (c.sgs_filter == NIMBUS_SGS_FILTER_CUBIC_ROOT || c.sgs_filter == NIMBUS_SGS_FILTER_DX) &&
                       std::isfinite(c.sgs_cs) && c.sgs_cs >= 0 && std::isfinite(c.sgs_pr_t) && c.sgs_pr_t > 0 &&
                       std::isfinite(c.sgs_sc_t) && c.sgs_sc_t > 0 && std::isfinite(c.sgs_nu_max) &&
                       c.sgs_nu_max >= 0 && std::isfinite(c.sgs_k_max) && c.sgs_k_max >= 0 &&
                       std::isfinite(c.sgs_dt) && c.sgs_dt > 0 && c.dt / c.sgs_dt <= 1000.0)) &&
           c.diffusion >= 0 && c.diffusion <= (NIMBUS_DIFFUSE_SCALARS | NIMBUS_DIFFUSE_MOMENTUM) &&
           (c.diffusion == 0 || (std::isfinite(c.diffusion_kh) && c.diffusion_kh >= 0.0 &&
                                 std::isfinite(c.diffusion_kv) && c.diffusion_kv >= 0.0));
}

struct Member
{
    nimbus_member_desc desc{};
    std::vector<double> host_base;  // kBaseCount * nz
    int elem = 4;
    MemberArgs args{};              // host copy of the device member-table entry
    void* block = nullptr;          // one device allocation for everything
    std::size_t bytes = 0;
    std::int64_t steps = 0;
    int tile_w = 64, k_chunks = 1, xcd_remap = 1;
    const DevGrid& g() const { return args.g; }
    const DevFields& f() const { return args.f; }
};

// A batch: members of one precision that share every launch.
struct Batch
{
    std::vector<std::uint32_t> ids;
    int elem = 4;
    MemberArgs* d_args = nullptr;
    int2* d_maps = nullptr;
    BatchView view{};
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    std::uint64_t last_use = 0;
    ~Batch()
    {
        if (exec) hipGraphExecDestroy(exec);
        if (graph) hipGraphDestroy(graph);
        if (d_args) hipFree(d_args);
        if (d_maps) hipFree(d_maps);
    }
};

}  // namespace

int set_error(int code, const std::string& msg) { return fail(code, msg); }

int context_device(nimbus_hip_ctx* ctx);
hipStream_t context_stream(nimbus_hip_ctx* ctx);
void count_dispatch(nimbus_hip_ctx* ctx, const char* op);

}  // namespace nimbus

struct nimbus_hip_ctx
{
    int device = 0;
    hipStream_t stream = nullptr;
    std::map<std::uint32_t, std::unique_ptr<nimbus::Member>> members;
    std::uint32_t next_id = 1;
    bool graphs = false;
    bool timing = false;
    std::vector<nimbus::TimedLaunch> pending;
    std::vector<hipEvent_t> event_pool;
    struct Row
    {
        std::uint64_t launches = 0;
        double ms = 0.0;
    };
    std::map<std::string, Row> table;
    std::uint64_t launches = 0;
    hipEvent_t step_a = nullptr, step_b = nullptr;
    double last_step_ms = 0.0;
    std::map<std::pair<int, std::vector<std::uint32_t>>, std::unique_ptr<nimbus::Batch>> batches;
    std::uint64_t batch_clock = 0;
    std::map<std::string, std::uint64_t> op_counts;  // successful per-operation dispatches by entry point
    // membership-change costs: batch table builds and whole-step graph
    // captures (capture + instantiate), host wall clock
    std::uint64_t batch_builds = 0, graph_captures = 0;
    double batch_build_ms = 0.0, graph_capture_ms = 0.0;
};

namespace nimbus
{
int context_device(nimbus_hip_ctx* ctx) { return ctx->device; }
hipStream_t context_stream(nimbus_hip_ctx* ctx) { return ctx->stream; }
void count_dispatch(nimbus_hip_ctx* ctx, const char* op) { ++ctx->op_counts[op]; }
}  // namespace nimbus

namespace nimbus
{
namespace
{

constexpr std::size_t kMaxCachedBatches = 8;

Launcher make_launcher(nimbus_hip_ctx* ctx)
{
    Launcher L;
    L.stream = ctx->stream;
    L.timing = ctx->timing;
    L.pending = &ctx->pending;
    L.event_pool = &ctx->event_pool;
    L.launches = &ctx->launches;
    return L;
}

// Moves completed per-kernel event pairs into the timing table (after a sync).
void harvest(nimbus_hip_ctx* ctx)
{
    for (auto& t : ctx->pending)
    {
        float ms = 0.0f;
        if (hipEventElapsedTime(&ms, t.start, t.stop) == hipSuccess)
        {
            auto& row = ctx->table[t.name];
            row.launches += 1;
            row.ms += ms;
        }
        ctx->event_pool.push_back(t.start);
        ctx->event_pool.push_back(t.stop);
    }
    ctx->pending.clear();
}

Member* find(nimbus_hip_ctx* ctx, std::uint32_t id)
{
    auto it = ctx->members.find(id);
    return it == ctx->members.end() ? nullptr : it->second.get();
}

void drop_batches(nimbus_hip_ctx* ctx)
{
    hipStreamSynchronize(ctx->stream);
    ctx->batches.clear();
}

void set_geometry(Member& m, int tile_w, int k_chunks, int remap)
{
    m.tile_w = tile_w == 0 ? 0 : (tile_w == 32 ? 32 : 64);
    m.k_chunks = k_chunks > 0 ? std::min(k_chunks, m.args.g.nz) : default_k_chunks(m.args.g, m.tile_w);
    m.xcd_remap = remap ? 1 : 0;
    fill_geometry(m.args, m.tile_w, m.k_chunks, m.xcd_remap);
}

// Fixes the HE-VI column-solve form of a member (NIMBUS_COLUMN_SOLVE_*).
int set_column_solve(Member& m, int form)
{
    if (m.desc.config.acoustic_mode != NIMBUS_ACOUSTIC_HEVI)
        return fail(NIMBUS_EINVAL, "set_column_solve: the member does not use the HE-VI acoustic mode");
    const int nz = m.args.g.nz;
    int seg = 0;
    switch (form)
    {
    case NIMBUS_COLUMN_SOLVE_AUTO: seg = hevi_default_seg(m.elem, m.args.g.nx, m.args.g.ny, nz); break;
    case NIMBUS_COLUMN_SOLVE_PARTITIONED:
        if (nz > kSegMaxNz) return fail(NIMBUS_EINVAL, "set_column_solve: the partitioned solve takes at most 129 levels");
        seg = 1;
        break;
    case NIMBUS_COLUMN_SOLVE_SEQUENTIAL:
        if (!m.args.f.tri_c)
            return fail(NIMBUS_EINVAL, "set_column_solve: FP64 members of at most 129 levels hold no scratch for the "
                                       "sequential sweep");
        seg = 0;
        break;
    default: return fail(NIMBUS_EINVAL, "set_column_solve: unknown form");
    }
    m.args.p.hevi_seg = seg;
    return NIMBUS_OK;
}

int allocate_member(const nimbus_member_desc& desc, Member& m)
{
    m.desc = desc;
    m.elem = desc.precision == NIMBUS_PRECISION_MIXED ? 4 : 8;
    m.args.g = dev_grid(desc.grid, m.elem);
    const int nz = desc.grid.nz;
    const bool hevi = desc.config.acoustic_mode == NIMBUS_ACOUSTIC_HEVI;
    m.host_base.assign(std::size_t(kBaseCount) * nz, 0.0);
    const double* src[6] = {desc.base.rho0, desc.base.p0, desc.base.theta0, desc.base.qv0, desc.base.u0,
                            desc.base.v0};
    for (int s = 0; s < 6; ++s)
        for (int k = 0; k < nz; ++k)
        {
            double v = src[s] ? src[s][k] : 0.0;
            if (s == kBaseQv0Stored) v = double(float(v));
            m.host_base[std::size_t(s) * nz + k] = v;
        }
    // Reference point of the perturbation EOS. Without a theta0 profile the
    // EOS-consistent theta of (rho0, p0) is used, so EOS(rho0, theta0) = p0.
    for (int k = 0; k < nz; ++k)
    {
        double* hb = m.host_base.data();
        const double r0 = hb[std::size_t(kBaseRho0) * nz + k], p0 = hb[std::size_t(kBaseP0) * nz + k];
        double& th0 = hb[std::size_t(kBaseTh0) * nz + k];
        if (!(th0 > 0.0) && r0 > 0.0 && p0 > 0.0) th0 = p0 / (r0 * kRd * std::pow(p0 / kP00, kKappa));
        double pe0 = 0.0, inv = 0.0;
        if (r0 > 0.0 && th0 > 0.0)
        {
            pe0 = kP00 * std::pow(r0 * kRd * th0 / kP00, kGamma);
            inv = 1.0 / (r0 * th0);
        }
        hb[std::size_t(kBasePe0) * nz + k] = pe0;
        hb[std::size_t(kBasePeDiff) * nz + k] = pe0 - p0;
        hb[std::size_t(kBaseInvRT0) * nz + k] = inv;
    }
    const int substeps = resolve_substeps(desc.grid, m.host_base.data() + std::size_t(kBaseRho0) * nz,
                                          m.host_base.data() + std::size_t(kBaseP0) * nz, desc.config);
    if (substeps < 1)
        return fail(NIMBUS_EINVAL, "acoustic substep count unresolvable (non-positive base density or "
                                   "more than 1e6 substeps required)");
    const int W = lateral_cells(desc.config, desc.grid);
    // Rayleigh sponge decay factors 1 - exp(-alpha dt) (rayleigh_damping.cpp:145-163,425-486,1174-1183)
    std::vector<double> lat(std::max(W, 1), 0.0);
    if (desc.config.sponge)
    {
        const nimbus_step_config& c = desc.config;
        const double pi_half = 1.5707963267948966, dtl = c.dt;
        auto factor = [dtl](double alpha) { return alpha > 0.0 ? -std::expm1(-alpha * dtl) : 0.0; };
        const double z_top = nz * desc.grid.dz, depth = z_top - c.sponge_z_d;
        auto alpha_at = [&](double z) {
            if (depth <= 0.0 || z <= c.sponge_z_d) return 0.0;
            const double sn = std::sin(pi_half * std::min(1.0, (z - c.sponge_z_d) / depth));
            return c.sponge_alpha * sn * sn;
        };
        double* hb = m.host_base.data();
        for (int k = 0; k < nz; ++k)
        {
            hb[std::size_t(kBaseSpongeUp) * nz + k] = factor(alpha_at((k + 0.5) * desc.grid.dz));
            hb[std::size_t(kBaseSpongeUpW) * nz + k] = factor(alpha_at((k + 1.0) * desc.grid.dz));
        }
        const int n = std::min(std::max(c.sponge_w_cells, 0), nz);
        for (int d = 0; d < n && c.sponge_w_alpha > 0.0; ++d)
        {
            const double frac = n > 1 ? 1.0 - double(d) / double(n - 1) : 1.0;
            const double sn = std::sin(pi_half * frac);
            hb[std::size_t(kBaseSpongeW) * nz + (nz - 1 - d)] = factor(c.sponge_w_alpha * sn * sn);
        }
        for (int d = 0; d < W; ++d)
        {
            const double sn = std::sin(pi_half * (1.0 - double(d) / double(W)));
            lat[d] = factor(c.sponge_lateral_alpha * sn * sn);
        }
    }
    const Layout L = layout_for(m.args.g, m.elem, hevi, W);
    m.bytes = L.total;
    hipError_t e = hipMalloc(&m.block, L.total);
    if (e != hipSuccess)
    {
        m.block = nullptr;
        return fail(NIMBUS_ENOMEM, std::string("member allocation failed: ") + hipGetErrorString(e));
    }
    char* p = static_cast<char*>(m.block);
    void* fields[F_COUNT];
    for (int s = 0; s < F_COUNT; ++s)
    {
        fields[s] = p;
        p += L.field;
    }
    DevFields& f = m.args.f;
    f.u = fields[F_U]; f.v = fields[F_V]; f.w = fields[F_W]; f.rp = fields[F_RP]; f.pp = fields[F_PP];
    f.th = fields[F_TH]; f.qv = fields[F_QV]; f.qc = fields[F_QC]; f.qr = fields[F_QR];
    f.u2 = fields[F_U2]; f.v2 = fields[F_V2]; f.w2 = fields[F_W2]; f.rp2 = fields[F_RP2];
    f.mfx = fields[F_MFX]; f.mfy = fields[F_MFY]; f.mfz = fields[F_MFZ]; f.rb = fields[F_RB];
    f.k1u = fields[F_K1U]; f.k1v = fields[F_K1V]; f.k1w = fields[F_K1W];
    f.tth = fields[F_TTH]; f.tqv = fields[F_TQV]; f.tqc = fields[F_TQC]; f.tqr = fields[F_TQR];
    void* flux_slots[12] = {f.k1u, f.k1v, f.k1w, f.u2, f.v2, f.w2, f.rp2, fields[F_FLUX7], fields[F_FLUX7 + 1],
                            fields[F_FLUX7 + 2], fields[F_FLUX7 + 3], fields[F_FLUX7 + 4]};
    for (int q = 0; q < 12; ++q) f.flux[q] = flux_slots[q];
    f.rl_qv = reinterpret_cast<float*>(p); p += L.rl;
    f.rl_qc = reinterpret_cast<float*>(p); p += L.rl;
    f.rl_qr = reinterpret_cast<float*>(p); p += L.rl;
    f.tri_c = nullptr;
    f.tri_r = nullptr;
    if (L.tri)
    {
        f.tri_c = reinterpret_cast<double*>(p); p += L.tri;
        f.tri_r = reinterpret_cast<double*>(p); p += L.tri;
    }
    f.rad_w = reinterpret_cast<float*>(p); p += L.rad_w;
    f.rad_e = reinterpret_cast<float*>(p); p += L.rad_w;
    f.rad_s = reinterpret_cast<float*>(p); p += L.rad_s;
    f.rad_n = reinterpret_cast<float*>(p); p += L.rad_s;
    f.rain = reinterpret_cast<double*>(p); p += L.rain;
    f.base = reinterpret_cast<double*>(p); p += L.base;
    f.diag = reinterpret_cast<double*>(p); p += L.diag;
    f.sponge_lat = reinterpret_cast<double*>(p); p += L.lat;
    f.clock = reinterpret_cast<double*>(p); p += L.clock;
    // Zero everything once (not inside any graph), then upload the base state
    // and the microphysics clock (its limit slot starts at +infinity).
    e = hipMemset(m.block, 0, L.total);
    if (e == hipSuccess)
    {
        const double clock0[4] = {0.0, 0.0, std::numeric_limits<double>::infinity(), 0.0};
        e = hipMemcpy(f.clock, clock0, sizeof clock0, hipMemcpyHostToDevice);
    }
    if (e == hipSuccess)
        e = hipMemcpy(f.base, m.host_base.data(), m.host_base.size() * sizeof(double), hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(f.sponge_lat, lat.data(), lat.size() * sizeof(double), hipMemcpyHostToDevice);
    if (e != hipSuccess) return hip_fail(e, "member initialisation");
    StepParams& sp = m.args.p;
    const nimbus_step_config& c = desc.config;
    sp.dt = c.dt;
    sp.substeps = substeps;
    sp.dts = c.dt / substeps;
    sp.mf_weight = sp.dts / c.dt;
    sp.acoustic_mode = c.acoustic_mode;
    sp.hevi_seg = hevi ? hevi_default_seg(m.elem, desc.grid.nx, desc.grid.ny, nz) : 0;
    sp.hevi_alpha = 0.5 * (1.0 + c.hevi_beta);
    sp.coriolis_f = c.coriolis_f;
    sp.bc_west = c.bc_west; sp.bc_east = c.bc_east; sp.bc_south = c.bc_south; sp.bc_north = c.bc_north;
    sp.bc_bottom = c.bc_bottom; sp.bc_top = c.bc_top;
    // Microphysics subcycling (microphysics_step.cpp:3501-3572,3700-3750): the
    // reference's limit is the minimum over cells of 0.8 dz / speed with
    // speed <= Vt_max, so it is never below 0.8 dz / Vt_max. A step that fits
    // under that bound is one subcycle whatever the state; a longer one takes
    // the state-dependent lengths min(remaining, limit) from a device-side
    // clock, within a launch budget that covers the shortest possible limit
    // (plus one launch for a rounding remainder).
    const double limit = 0.8 * desc.grid.dz / c.vt_max;
    sp.kessler_adaptive = c.dt <= limit ? 0 : 1;
    sp.kessler_sub = c.dt <= limit ? 1 : int(std::ceil(c.dt / limit)) + 1;
    sp.kessler = KesslerConst{c.qc0, c.c_auto, c.c_accr, c.evap_eff, c.vt_max};
    sp.sponge = c.sponge ? 1 : 0;
    sp.sponge_lateral = W;
    // Smagorinsky SGS substeps: dt_sgs each, the remainder last (the same
    // accumulation as the CPU reference's loop)
    sp.sgs_n = 0;
    if (c.sgs)
    {
        double done = 0.0, h = 0.0, first = 0.0;
        while (done < c.dt && sp.sgs_n < 1000)
        {
            h = std::min(c.dt - done, c.sgs_dt);
            if (!(h > 64.0 * std::numeric_limits<double>::epsilon() * std::max(1.0, c.dt))) break;
            if (sp.sgs_n == 0) first = h;
            ++sp.sgs_n;
            done += h;
        }
        sp.sgs_h = first;
        sp.sgs_h_last = h;
        const double delta = c.sgs_filter == NIMBUS_SGS_FILTER_DX ? std::sqrt(desc.grid.dx * desc.grid.dy)
                                                                  : std::cbrt(desc.grid.dx * desc.grid.dy * desc.grid.dz);
        sp.sgs_c2d2 = c.sgs_cs * c.sgs_cs * delta * delta;
        sp.sgs_inv_pr = 1.0 / c.sgs_pr_t;
        sp.sgs_inv_sc = 1.0 / c.sgs_sc_t;
        sp.sgs_mode = c.sgs_mode;
        sp.sgs_nu_max = c.sgs_nu_max;
        sp.sgs_k_max = c.sgs_k_max;
        sp.sgs_wrap_x = c.bc_west == NIMBUS_BC_PERIODIC && c.bc_east == NIMBUS_BC_PERIODIC;
        sp.sgs_wrap_y = c.bc_south == NIMBUS_BC_PERIODIC && c.bc_north == NIMBUS_BC_PERIODIC;
        const int lat[4] = {c.bc_west, c.bc_east, c.bc_south, c.bc_north};
        for (int q = 0; q < 4; ++q)
            sp.sgs_open[q] = lat[q] == NIMBUS_BC_RADIATIVE || lat[q] == NIMBUS_BC_ZERO_GRADIENT;
    }
    // explicit diffusion after the SGS: off when neither coefficient is positive
    sp.dif_apply = c.diffusion != 0 && (c.diffusion_kh > 0.0 || c.diffusion_kv > 0.0) ? c.diffusion : 0;
    sp.dif_kh = c.diffusion_kh;
    sp.dif_kv = c.diffusion_kv;
    sp.dif_idx2 = 1.0 / (desc.grid.dx * desc.grid.dx);
    sp.dif_idy2 = 1.0 / (desc.grid.dy * desc.grid.dy);
    sp.dif_wrap_x = c.bc_west == NIMBUS_BC_PERIODIC && c.bc_east == NIMBUS_BC_PERIODIC;
    sp.dif_wrap_y = c.bc_south == NIMBUS_BC_PERIODIC && c.bc_north == NIMBUS_BC_PERIODIC;
    sp.bounds = Bounds{};
    // linear runs for the per-cell kernels (a launch streams through memory:
    // 3-17 % faster per kernel than 64 x 4 column tiles at 400 x 400 x 100),
    // column tiles with the default vertical chunks for the rest
    set_geometry(m, 0, 0, 1);
    return NIMBUS_OK;
}

void free_member(Member& m)
{
    if (m.block) hipFree(m.block);
    m.block = nullptr;
}

const double* base_slot(const Member& m, int s) { return m.f().base + std::size_t(s) * m.g().nz; }

// Block b of a launch over n logical work items runs on XCD b % 8; the
// permutation hands each XCD a contiguous range of logical items, so the
// horizontally neighbouring tiles a stencil reads share that XCD's L2.
int xcd_logical(int b, int n)
{
    constexpr int kXcd = 8;
    const int x = b % kXcd, local = b / kXcd;
    const int q = n / kXcd, r = n % kXcd;
    return x < r ? x * (q + 1) + local : r * (q + 1) + (x - r) * q + local;
}

// Builds the device member table and block maps of a batch.
int build_batch(Batch& b, const std::vector<Member*>& ms)
{
    std::vector<MemberArgs> args;
    args.reserve(ms.size());
    b.view.max_substeps = 0;
    b.view.max_kessler_sub = 0;
    b.view.any_sponge = 0;
    b.view.max_nz = 0;
    b.view.max_sgs_n = 0;
    b.view.any_odd_substeps = b.view.any_even_kessler_sub = b.view.any_odd_sgs_n = 0;
    b.view.any_sponge_unfused = 0;
    b.view.any_diffusion = 0;
    bool remap = true;
    for (Member* m : ms)
    {
        args.push_back(m->args);
        b.view.max_substeps = std::max(b.view.max_substeps, m->args.p.substeps);
        b.view.max_kessler_sub = std::max(b.view.max_kessler_sub, m->args.p.kessler_sub);
        b.view.any_sponge = b.view.any_sponge || m->args.p.sponge;
        b.view.max_nz = std::max(b.view.max_nz, m->args.g.nz);
        b.view.max_sgs_n = std::max(b.view.max_sgs_n, m->args.p.sgs_n);
        b.view.any_odd_substeps |= m->args.p.substeps & 1;
        b.view.any_even_kessler_sub |= (m->args.p.kessler_sub & 1) == 0;
        b.view.any_odd_sgs_n |= m->args.p.sgs_n & 1;
        b.view.any_sponge_unfused |= m->args.p.sponge && m->args.p.sgs_n == 0 && m->args.p.dif_apply == 0;
        b.view.any_diffusion |= m->args.p.dif_apply != 0;
        b.view.any_kessler_adaptive |= m->args.p.kessler_adaptive;
        if (m->args.p.acoustic_mode == NIMBUS_ACOUSTIC_HEVI)
        {
            const bool part = m->args.p.hevi_seg != 0;
            b.view.any_hevi_seg |= part;
            b.view.any_hevi_serial |= !part;
        }
        remap = remap && m->xcd_remap;
    }
    std::vector<int2> all;
    std::size_t offsets[S_SUBSETS][W_KINDS];
    int counts[S_SUBSETS][W_KINDS];
    for (int s = 0; s < S_SUBSETS; ++s)
        for (int k = 0; k < W_KINDS; ++k)
        {
            std::vector<int2> logical;
            for (std::size_t q = 0; q < ms.size(); ++q)
            {
                const int mode = ms[q]->args.p.acoustic_mode;
                if (s == S_EXPLICIT && mode != NIMBUS_ACOUSTIC_EXPLICIT) continue;
                if (s == S_HEVI && mode != NIMBUS_ACOUSTIC_HEVI) continue;
                const int n = work_items(ms[q]->args, WorkKind(k));
                for (int l = 0; l < n; ++l) logical.push_back(int2{int(q), l});
            }
            const int n = int(logical.size());
            offsets[s][k] = all.size();
            counts[s][k] = n;
            for (int blk = 0; blk < n; ++blk) all.push_back(logical[remap ? xcd_logical(blk, n) : blk]);
        }
    NIMBUS_HIP_TRY(hipMalloc(&b.d_args, sizeof(MemberArgs) * args.size()));
    NIMBUS_HIP_TRY(hipMemcpy(b.d_args, args.data(), sizeof(MemberArgs) * args.size(), hipMemcpyHostToDevice));
    NIMBUS_HIP_TRY(hipMalloc(&b.d_maps, sizeof(int2) * std::max<std::size_t>(all.size(), 1)));
    if (!all.empty())
        NIMBUS_HIP_TRY(hipMemcpy(b.d_maps, all.data(), sizeof(int2) * all.size(), hipMemcpyHostToDevice));
    b.view.args = b.d_args;
    b.view.n = int(ms.size());
    b.view.elem = b.elem;
    for (int s = 0; s < S_SUBSETS; ++s)
        for (int k = 0; k < W_KINDS; ++k)
            b.view.maps[s][k] = BlockMap{b.d_maps + offsets[s][k], counts[s][k]};
    return NIMBUS_OK;
}

// Cached batch for a set of member ids of one precision.
int get_batch(nimbus_hip_ctx* ctx, const std::vector<Member*>& ms, const std::vector<std::uint32_t>& ids, Batch** out)
{
    std::vector<std::uint32_t> key = ids;
    std::sort(key.begin(), key.end());
    const int elem = ms.front()->elem;
    auto k = std::make_pair(elem, key);
    auto it = ctx->batches.find(k);
    if (it == ctx->batches.end())
    {
        if (ctx->batches.size() >= kMaxCachedBatches)
        {
            auto oldest = ctx->batches.begin();
            for (auto jt = ctx->batches.begin(); jt != ctx->batches.end(); ++jt)
                if (jt->second->last_use < oldest->second->last_use) oldest = jt;
            hipStreamSynchronize(ctx->stream);
            ctx->batches.erase(oldest);
        }
        auto b = std::make_unique<Batch>();
        b->ids = key;
        b->elem = elem;
        std::vector<Member*> sorted;
        for (std::uint32_t id : key) sorted.push_back(find(ctx, id));
        const auto t0 = std::chrono::steady_clock::now();
        const int rc = build_batch(*b, sorted);
        if (rc != NIMBUS_OK) return rc;
        ctx->batch_builds += 1;
        ctx->batch_build_ms += std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
        it = ctx->batches.emplace(k, std::move(b)).first;
    }
    it->second->last_use = ++ctx->batch_clock;
    *out = it->second.get();
    return NIMBUS_OK;
}

void enqueue_step(Launcher& L, const Batch& b)
{
    if (b.elem == 4) launch_large_step<float>(L, b.view);
    else launch_large_step<double>(L, b.view);
}

int run_batch_step(nimbus_hip_ctx* ctx, Batch& b)
{
    if (ctx->graphs && !ctx->timing)
    {
        if (!b.exec)
        {
            const auto t0 = std::chrono::steady_clock::now();
            Launcher L = make_launcher(ctx);
            L.timing = false;
            NIMBUS_HIP_TRY(hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal));
            enqueue_step(L, b);
            hipGraph_t graph = nullptr;
            const hipError_t e = hipStreamEndCapture(ctx->stream, &graph);
            if (L.error != hipSuccess) return hip_fail(L.error, "step capture launch");
            if (e != hipSuccess) return hip_fail(e, "hipStreamEndCapture");
            b.graph = graph;
            NIMBUS_HIP_TRY(hipGraphInstantiate(&b.exec, b.graph, nullptr, nullptr, 0));
            ctx->graph_captures += 1;
            ctx->graph_capture_ms +=
                std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
        }
        NIMBUS_HIP_TRY(hipGraphLaunch(b.exec, ctx->stream));
    }
    else
    {
        Launcher L = make_launcher(ctx);
        enqueue_step(L, b);
        if (L.error != hipSuccess) return hip_fail(L.error, "step launch");
    }
    return NIMBUS_OK;
}

int step_members(nimbus_hip_ctx* ctx, const std::uint32_t* ids, int n_ids, int n_steps, bool wait)
{
    if (!ctx || (!ids && n_ids > 0) || n_ids < 0 || n_steps < 0) return fail(NIMBUS_EINVAL, "step: bad arguments");
    std::vector<Member*> groups[2];
    std::vector<std::uint32_t> gids[2];
    for (int q = 0; q < n_ids; ++q)
    {
        Member* m = find(ctx, ids[q]);
        if (!m) return fail(NIMBUS_ENOENT, "step: unknown member");
        for (int r = 0; r < q; ++r)
            if (ids[r] == ids[q]) return fail(NIMBUS_EINVAL, "step: member listed twice");
        const int gi = m->elem == 4 ? 0 : 1;
        groups[gi].push_back(m);
        gids[gi].push_back(ids[q]);
    }
    Batch* batches[2] = {nullptr, nullptr};
    for (int gi = 0; gi < 2; ++gi)
        if (!groups[gi].empty())
        {
            const int rc = get_batch(ctx, groups[gi], gids[gi], &batches[gi]);
            if (rc != NIMBUS_OK) return rc;
        }
    NIMBUS_HIP_TRY(hipEventRecord(ctx->step_a, ctx->stream));
    for (int s = 0; s < n_steps; ++s)
        for (Batch* b : batches)
            if (b)
            {
                const int rc = run_batch_step(ctx, *b);
                if (rc != NIMBUS_OK) return rc;
            }
    NIMBUS_HIP_TRY(hipEventRecord(ctx->step_b, ctx->stream));
    for (int gi = 0; gi < 2; ++gi)
        for (Member* m : groups[gi]) m->steps += n_steps;
    if (wait)
    {
        NIMBUS_HIP_TRY(hipStreamSynchronize(ctx->stream));
        float ms = 0.0f;
        NIMBUS_HIP_TRY(hipEventElapsedTime(&ms, ctx->step_a, ctx->step_b));
        ctx->last_step_ms = ms;
        harvest(ctx);
    }
    return NIMBUS_OK;
}

// Host field h (reference layout, element H) -> device field d via a staging buffer.
template <class H>
int upload_field(nimbus_hip_ctx* ctx, Member& m, const void* h, void* d, const double* subtract, void* staging)
{
    const std::size_t n = std::size_t(m.g().nx) * m.g().ny * m.g().nz;
    NIMBUS_HIP_TRY(hipMemcpyAsync(staging, h, n * sizeof(H), hipMemcpyHostToDevice, ctx->stream));
    Launcher L = make_launcher(ctx);
    L.timing = false;
    if (m.elem == 4) launch_host_to_dev<H, float>(L, m.g(), staging, d, subtract);
    else launch_host_to_dev<H, double>(L, m.g(), staging, d, subtract);
    if (L.error != hipSuccess) return hip_fail(L.error, "upload transpose");
    return NIMBUS_OK;
}

template <class H>
int download_field(nimbus_hip_ctx* ctx, Member& m, const void* d, void* h, const double* add, void* staging)
{
    const std::size_t n = std::size_t(m.g().nx) * m.g().ny * m.g().nz;
    Launcher L = make_launcher(ctx);
    L.timing = false;
    if (m.elem == 4) launch_dev_to_host<float, H>(L, m.g(), d, staging, add);
    else launch_dev_to_host<double, H>(L, m.g(), d, staging, add);
    if (L.error != hipSuccess) return hip_fail(L.error, "download transpose");
    NIMBUS_HIP_TRY(hipMemcpyAsync(h, staging, n * sizeof(H), hipMemcpyDeviceToHost, ctx->stream));
    NIMBUS_HIP_TRY(hipStreamSynchronize(ctx->stream));
    return NIMBUS_OK;
}

struct Staging
{
    void* p = nullptr;
    ~Staging()
    {
        if (p) hipFree(p);
    }
};

template <class H>
int upload_state(nimbus_hip_ctx* ctx, Member& m, const nimbus_host_state& s)
{
    Staging st;
    const std::size_t n = std::size_t(m.g().nx) * m.g().ny * m.g().nz;
    NIMBUS_HIP_TRY(hipMalloc(&st.p, n * sizeof(H)));
    const void* hs[9] = {s.u, s.v, s.w, s.rho, s.p, s.theta, s.qv, s.qc, s.qr};
    const DevFields& f = m.f();
    void* ds[9] = {f.u, f.v, f.w, f.rp, f.pp, f.th, f.qv, f.qc, f.qr};
    for (int q = 0; q < 9; ++q)
    {
        if (!hs[q]) return fail(NIMBUS_EINVAL, "upload: null field");
        const double* sub = q == 3 ? base_slot(m, kBaseRho0) : (q == 4 ? base_slot(m, kBaseP0) : nullptr);
        const int rc = upload_field<H>(ctx, m, hs[q], ds[q], sub, st.p);
        if (rc != NIMBUS_OK) return rc;
        NIMBUS_HIP_TRY(hipStreamSynchronize(ctx->stream));
    }
    const std::size_t cols = std::size_t(m.g().nx) * m.g().ny;
    std::vector<double> rain(cols, 0.0);
    if (s.rain_mm)
        for (int i = 0; i < m.g().nx; ++i)
            for (int j = 0; j < m.g().ny; ++j)
                rain[std::size_t(j) * m.g().nx + i] = s.rain_mm[std::size_t(i) * m.g().ny + j];
    NIMBUS_HIP_TRY(hipMemcpy(f.rain, rain.data(), cols * sizeof(double), hipMemcpyHostToDevice));
    return NIMBUS_OK;
}

template <class H>
int download_state(nimbus_hip_ctx* ctx, Member& m, const nimbus_host_state& s)
{
    Staging st;
    const std::size_t n = std::size_t(m.g().nx) * m.g().ny * m.g().nz;
    NIMBUS_HIP_TRY(hipMalloc(&st.p, n * sizeof(H)));
    void* hs[9] = {s.u, s.v, s.w, s.rho, s.p, s.theta, s.qv, s.qc, s.qr};
    const DevFields& f = m.f();
    void* ds[9] = {f.u, f.v, f.w, f.rp, f.pp, f.th, f.qv, f.qc, f.qr};
    for (int q = 0; q < 9; ++q)
    {
        if (!hs[q]) continue;
        const double* add = q == 3 ? base_slot(m, kBaseRho0) : (q == 4 ? base_slot(m, kBaseP0) : nullptr);
        const int rc = download_field<H>(ctx, m, ds[q], hs[q], add, st.p);
        if (rc != NIMBUS_OK) return rc;
    }
    if (s.rain_mm)
    {
        const std::size_t cols = std::size_t(m.g().nx) * m.g().ny;
        std::vector<double> rain(cols);
        NIMBUS_HIP_TRY(hipMemcpy(rain.data(), f.rain, cols * sizeof(double), hipMemcpyDeviceToHost));
        for (int i = 0; i < m.g().nx; ++i)
            for (int j = 0; j < m.g().ny; ++j)
                s.rain_mm[std::size_t(i) * m.g().ny + j] = rain[std::size_t(j) * m.g().nx + i];
    }
    return NIMBUS_OK;
}

// A scratch member for one per-operation dispatch (mixed precision, given
// grid and base profiles) with its own one-member batch; both are released
// when the dispatch returns.
struct ScratchMember
{
    Member m;
    Batch batch;
    ~ScratchMember() { free_member(m); }
    int finish_setup()
    {
        batch.elem = m.elem;
        std::vector<Member*> one{&m};
        return build_batch(batch, one);
    }
};

int make_scratch(ScratchMember& sm, int nx, int ny, int nz, double dx, double dy, double dz, const double* rho0,
                 const double* p0, const double* th0, const double* qv0, const double* u0, const double* v0,
                 int acoustic_mode)
{
    nimbus_member_desc d{};
    d.grid = nimbus_grid{nx, ny, nz, dx, dy, dz};
    d.config.dt = 1.0;
    d.config.substeps = 1;
    d.config.acoustic_mode = acoustic_mode;
    d.config.hevi_beta = 0.1;
    d.config.vt_max = 20.0;
    d.precision = NIMBUS_PRECISION_MIXED;
    d.base = nimbus_base_state{rho0, p0, th0, qv0, u0, v0};
    if (!valid_grid(d.grid)) return fail(NIMBUS_EINVAL, "dispatch: invalid grid");
    return allocate_member(d, sm.m);
}

template <class T>
bool all_nonnull(std::initializer_list<T> ps)
{
    for (auto p : ps)
        if (!p) return false;
    return true;
}

}  // namespace
}  // namespace nimbus

using namespace nimbus;

extern "C" {

const char* nimbus_hip_last_error(void) { return g_last_error.c_str(); }

int nimbus_hip_create(int device, nimbus_hip_ctx** out)
{
    if (!out) return fail(NIMBUS_EINVAL, "create: null output");
    *out = nullptr;
    int count = 0;
    if (hipGetDeviceCount(&count) != hipSuccess || count <= 0) return fail(NIMBUS_ENODEV, "no HIP device");
    if (device < 0 || device >= count) return fail(NIMBUS_ENODEV, "device index out of range");
    NIMBUS_HIP_TRY(hipSetDevice(device));
    auto ctx = std::make_unique<nimbus_hip_ctx>();
    ctx->device = device;
    NIMBUS_HIP_TRY(hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking));
    NIMBUS_HIP_TRY(hipEventCreate(&ctx->step_a));
    NIMBUS_HIP_TRY(hipEventCreate(&ctx->step_b));
    *out = ctx.release();
    return NIMBUS_OK;
}

int nimbus_hip_destroy(nimbus_hip_ctx* ctx)
{
    if (!ctx) return NIMBUS_OK;
    hipSetDevice(ctx->device);
    hipStreamSynchronize(ctx->stream);
    ctx->batches.clear();
    for (auto& kv : ctx->members) free_member(*kv.second);
    ctx->members.clear();
    harvest(ctx);
    for (hipEvent_t e : ctx->event_pool) hipEventDestroy(e);
    if (ctx->step_a) hipEventDestroy(ctx->step_a);
    if (ctx->step_b) hipEventDestroy(ctx->step_b);
    if (ctx->stream) hipStreamDestroy(ctx->stream);
    delete ctx;
    return NIMBUS_OK;
}

int nimbus_hip_mem_info(nimbus_hip_ctx* ctx, uint64_t* free_bytes, uint64_t* total_bytes)
{
    if (!ctx || !free_bytes || !total_bytes) return fail(NIMBUS_EINVAL, "mem_info: null argument");
    size_t f = 0, t = 0;
    NIMBUS_HIP_TRY(hipMemGetInfo(&f, &t));
    *free_bytes = f;
    *total_bytes = t;
    return NIMBUS_OK;
}

uint64_t nimbus_hip_dispatch_count(nimbus_hip_ctx* ctx) { return ctx ? ctx->launches : 0; }

int nimbus_hip_device_arch(nimbus_hip_ctx* ctx, char* buf, size_t n)
{
    if (!ctx || !buf || n == 0) return fail(NIMBUS_EINVAL, "device_arch: bad arguments");
    hipDeviceProp_t prop{};
    NIMBUS_HIP_TRY(hipGetDeviceProperties(&prop, ctx->device));
    std::snprintf(buf, n, "%s", prop.gcnArchName);
    return NIMBUS_OK;
}

int nimbus_hip_batch_stats(nimbus_hip_ctx* ctx, uint64_t* builds, double* build_ms, uint64_t* captures,
                           double* capture_ms)
{
    if (!ctx) return fail(NIMBUS_EINVAL, "batch_stats: null context");
    if (builds) *builds = ctx->batch_builds;
    if (build_ms) *build_ms = ctx->batch_build_ms;
    if (captures) *captures = ctx->graph_captures;
    if (capture_ms) *capture_ms = ctx->graph_capture_ms;
    return NIMBUS_OK;
}

uint64_t nimbus_hip_successful_dispatch_count(nimbus_hip_ctx* ctx, const char* op)
{
    if (!ctx || !op) return 0;
    auto it = ctx->op_counts.find(op);
    return it == ctx->op_counts.end() ? 0 : it->second;
}

uint64_t nimbus_hip_member_bytes(const nimbus_member_desc* desc)
{
    if (!desc || !valid_grid(desc->grid)) return 0;
    const int elem = desc->precision == NIMBUS_PRECISION_MIXED ? 4 : 8;
    return layout_for(dev_grid(desc->grid, elem), elem, desc->config.acoustic_mode == NIMBUS_ACOUSTIC_HEVI,
                      lateral_cells(desc->config, desc->grid))
        .total;
}

int nimbus_hip_member_create(nimbus_hip_ctx* ctx, const nimbus_member_desc* desc, uint32_t* id)
{
    if (!ctx || !desc || !id) return fail(NIMBUS_EINVAL, "member_create: null argument");
    if (!valid_grid(desc->grid)) return fail(NIMBUS_EINVAL, "member_create: invalid grid");
    if (!valid_config(desc->config)) return fail(NIMBUS_EINVAL, "member_create: invalid step config");
    if (desc->precision != NIMBUS_PRECISION_FP64 && desc->precision != NIMBUS_PRECISION_MIXED)
        return fail(NIMBUS_EINVAL, "member_create: invalid precision");
    if (!desc->base.rho0 || !desc->base.p0 || !desc->base.theta0)
        return fail(NIMBUS_EINVAL, "member_create: base state needs rho0, p0, theta0");
    for (int k = 0; k < desc->grid.nz; ++k)
        if (!(desc->base.rho0[k] > 0.0) || !(desc->base.p0[k] > 0.0) || !std::isfinite(desc->base.rho0[k]) ||
            !std::isfinite(desc->base.p0[k]))
            return fail(NIMBUS_EINVAL, "member_create: base pressure and density must be positive and finite");
    hipSetDevice(ctx->device);
    auto m = std::make_unique<Member>();
    const int rc = allocate_member(*desc, *m);
    if (rc != NIMBUS_OK)
    {
        free_member(*m);
        return rc;
    }
    *id = ctx->next_id++;
    ctx->members[*id] = std::move(m);
    return NIMBUS_OK;
}

int nimbus_hip_member_destroy(nimbus_hip_ctx* ctx, uint32_t id)
{
    if (!ctx) return fail(NIMBUS_EINVAL, "member_destroy: null context");
    Member* m = find(ctx, id);
    if (!m) return fail(NIMBUS_ENOENT, "member_destroy: unknown member");
    hipStreamSynchronize(ctx->stream);
    for (auto it = ctx->batches.begin(); it != ctx->batches.end();)
    {
        const auto& key = it->first.second;
        if (std::find(key.begin(), key.end(), id) != key.end()) it = ctx->batches.erase(it);
        else ++it;
    }
    free_member(*m);
    ctx->members.erase(id);
    return NIMBUS_OK;
}

static int member_io(nimbus_hip_ctx* ctx, uint32_t id, const nimbus_host_state* s, bool up, int elem_host)
{
    if (!ctx || !s) return fail(NIMBUS_EINVAL, "member io: null argument");
    Member* m = find(ctx, id);
    if (!m) return fail(NIMBUS_ENOENT, "member io: unknown member");
    hipSetDevice(ctx->device);
    hipStreamSynchronize(ctx->stream);
    if (up) return elem_host == 4 ? upload_state<float>(ctx, *m, *s) : upload_state<double>(ctx, *m, *s);
    return elem_host == 4 ? download_state<float>(ctx, *m, *s) : download_state<double>(ctx, *m, *s);
}

// The host element type follows the member precision.
int nimbus_hip_member_upload(nimbus_hip_ctx* ctx, uint32_t id, const nimbus_host_state* s)
{
    Member* m = ctx ? find(ctx, id) : nullptr;
    return member_io(ctx, id, s, true, m ? m->elem : 4);
}

int nimbus_hip_member_download(nimbus_hip_ctx* ctx, uint32_t id, const nimbus_host_state* s)
{
    Member* m = ctx ? find(ctx, id) : nullptr;
    return member_io(ctx, id, s, false, m ? m->elem : 4);
}

int nimbus_hip_member_diag(nimbus_hip_ctx* ctx, uint32_t id, nimbus_diag* d)
{
    if (!ctx || !d) return fail(NIMBUS_EINVAL, "diag: null argument");
    Member* m = find(ctx, id);
    if (!m) return fail(NIMBUS_ENOENT, "diag: unknown member");
    hipSetDevice(ctx->device);
    Batch* b = nullptr;
    std::vector<Member*> one{m};
    std::vector<std::uint32_t> ids{id};
    int rc = get_batch(ctx, one, ids, &b);
    if (rc != NIMBUS_OK) return rc;
    Launcher L = make_launcher(ctx);
    L.timing = false;
    if (m->elem == 4) launch_diagnostics<float>(L, b->view);
    else launch_diagnostics<double>(L, b->view);
    if (L.error != hipSuccess) return hip_fail(L.error, "diag launch");
    double v[kDiagValues];
    NIMBUS_HIP_TRY(hipMemcpyAsync(v, m->f().diag + std::size_t(kDiagBlocksPerMember) * kDiagValues, sizeof(v),
                                  hipMemcpyDeviceToHost, ctx->stream));
    NIMBUS_HIP_TRY(hipStreamSynchronize(ctx->stream));
    const double cells = double(m->g().nx - 2) * double(m->g().ny - 2) * double(m->g().nz);
    d->max_w = v[0];
    d->min_w = v[1];
    d->rms_theta_p = std::sqrt(v[2] / cells);
    d->rain_total_mm = v[3];
    d->max_rain_mm = v[6];
    d->mean_qc = v[4] / cells;
    d->mean_qr = v[5] / cells;
    d->steps = m->steps;
    return NIMBUS_OK;
}

// ---------------------------------------------------------------------
// State I/O: NPY field export and member snapshots.
// ---------------------------------------------------------------------

int nimbus_hip_member_export_npy(nimbus_hip_ctx* ctx, uint32_t id, const char* dir)
{
    if (!ctx || !dir) return fail(NIMBUS_EINVAL, "export_npy: null argument");
    Member* m = find(ctx, id);
    if (!m) return fail(NIMBUS_ENOENT, "export_npy: unknown member");
    struct stat st;
    if (::stat(dir, &st) != 0 || !S_ISDIR(st.st_mode))
        return fail(NIMBUS_EINVAL, std::string("export_npy: not a directory: ") + dir);
    const int nx = m->g().nx, ny = m->g().ny, nz = m->g().nz;
    const std::size_t n = std::size_t(nx) * ny * nz;
    static const char* names[9] = {"u", "v", "w", "rho", "p", "theta", "qv", "qc", "qr"};
    std::vector<std::vector<float>> f32(9, std::vector<float>(n));
    std::vector<double> rain(std::size_t(nx) * ny);
    if (m->elem == 4)
    {
        nimbus_host_state hs{f32[0].data(), f32[1].data(), f32[2].data(), f32[3].data(), f32[4].data(),
                             f32[5].data(), f32[6].data(), f32[7].data(), f32[8].data(), rain.data()};
        const int rc = member_io(ctx, id, &hs, false, 4);
        if (rc != NIMBUS_OK) return rc;
    }
    else
    {
        std::vector<std::vector<double>> f64(9, std::vector<double>(n));
        nimbus_host_state hs{f64[0].data(), f64[1].data(), f64[2].data(), f64[3].data(), f64[4].data(),
                             f64[5].data(), f64[6].data(), f64[7].data(), f64[8].data(), rain.data()};
        const int rc = member_io(ctx, id, &hs, false, 8);
        if (rc != NIMBUS_OK) return rc;
        for (int q = 0; q < 9; ++q)
            for (std::size_t c = 0; c < n; ++c) f32[q][c] = float(f64[q][c]);
    }
    const std::int64_t shape3[3] = {nx, ny, nz}, shape2[2] = {nx, ny};
    for (int q = 0; q < 9; ++q)
    {
        const std::string path = std::string(dir) + "/" + names[q] + ".npy";
        const int rc = write_npy(path.c_str(), f32[q].data(), NIMBUS_NPY_F4, 3, shape3);
        if (rc != NIMBUS_OK) return rc;
    }
    std::vector<float> rain32(rain.begin(), rain.end());
    const std::string path = std::string(dir) + "/accumulated_rainfall.npy";
    return write_npy(path.c_str(), rain32.data(), NIMBUS_NPY_F4, 2, shape2);
}

namespace
{
constexpr char kSnapMagic[8] = {'N', 'I', 'M', 'B', 'U', 'S', 'M', 'S'};
constexpr std::uint32_t kSnapVersion = 4;

struct SnapHead
{
    char magic[8];
    std::uint32_t version, endian, grid_bytes, config_bytes;
    std::int32_t precision, tile_w, k_chunks, xcd_remap;
    std::int32_t hevi_seg, reserved;
    std::int64_t steps;
    std::uint64_t field_bytes, rain_bytes;
};

constexpr std::uint64_t kFnvBasis = 1469598103934665603ULL, kFnvPrime = 1099511628211ULL;

// FNV-1a, 64 bit
std::uint64_t fnv1a(const std::uint8_t* p, std::size_t n, std::uint64_t h = kFnvBasis)
{
    for (std::size_t i = 0; i < n; ++i)
    {
        h ^= p[i];
        h *= kFnvPrime;
    }
    return h;
}

// Snapshot checksum: the little-endian 64-bit words of the buffer go round
// robin to four lanes, each lane h = (h ^ word) * prime, h ^= h >> 29 (both
// steps invertible, so a changed word always changes its lane, and the shift
// carries changes in high bits down); the tail bytes go to lane 0 byte-wise,
// and the four lanes are folded with byte-wise FNV-1a. Four independent
// multiply chains hash several bytes per cycle, where byte-wise FNV-1a over a
// 400x400x100 member (577 MB) takes most of a second.
std::uint64_t snap_checksum(const std::uint8_t* p, std::size_t n)
{
    std::uint64_t h[4] = {kFnvBasis, kFnvBasis ^ 1u, kFnvBasis ^ 2u, kFnvBasis ^ 3u};
    const std::size_t words = n / 8;
    auto mix = [](std::uint64_t v, std::uint64_t x) {
        v = (v ^ x) * kFnvPrime;
        return v ^ (v >> 29);
    };
    std::size_t w = 0;
    for (; w + 4 <= words; w += 4)
    {
        std::uint64_t x[4];
        std::memcpy(x, p + 8 * w, sizeof x);
        h[0] = mix(h[0], x[0]);
        h[1] = mix(h[1], x[1]);
        h[2] = mix(h[2], x[2]);
        h[3] = mix(h[3], x[3]);
    }
    for (; w < words; ++w)
    {
        std::uint64_t x;
        std::memcpy(&x, p + 8 * w, sizeof x);
        h[w & 3] = mix(h[w & 3], x);
    }
    for (std::size_t b = 8 * words; b < n; ++b) h[0] = mix(h[0], p[b]);
    std::uint64_t out = kFnvBasis;
    for (int l = 0; l < 4; ++l)
    {
        std::uint8_t le[8];
        for (int q = 0; q < 8; ++q) le[q] = std::uint8_t(h[l] >> (8 * q));
        out = fnv1a(le, 8, out);
    }
    return out;
}

struct SnapSizes
{
    std::size_t field, rain, base, total;
};

SnapSizes snap_sizes(const Member& m)
{
    const bool hevi = m.desc.config.acoustic_mode == NIMBUS_ACOUSTIC_HEVI;
    const Layout L = layout_for(m.args.g, m.elem, hevi, lateral_cells(m.desc.config, m.desc.grid));
    SnapSizes z{};
    z.field = L.field;
    z.rain = std::size_t(m.g().nx) * m.g().ny * sizeof(double);
    z.base = 6 * std::size_t(m.g().nz) * sizeof(double);
    z.total = sizeof(SnapHead) + sizeof(nimbus_grid) + sizeof(nimbus_step_config) + z.base + 9 * z.field + z.rain +
              sizeof(std::uint64_t);
    return z;
}

void* prognostic(const Member& m, int q)
{
    const DevFields& f = m.f();
    void* ds[9] = {f.u, f.v, f.w, f.rp, f.pp, f.th, f.qv, f.qc, f.qr};
    return ds[q];
}
}  // namespace

int nimbus_hip_member_snapshot_size(nimbus_hip_ctx* ctx, uint32_t id, uint64_t* bytes)
{
    if (!ctx || !bytes) return fail(NIMBUS_EINVAL, "snapshot_size: null argument");
    Member* m = find(ctx, id);
    if (!m) return fail(NIMBUS_ENOENT, "snapshot_size: unknown member");
    *bytes = snap_sizes(*m).total;
    return NIMBUS_OK;
}

int nimbus_hip_member_snapshot(nimbus_hip_ctx* ctx, uint32_t id, void* buf, uint64_t cap)
{
    if (!ctx || !buf) return fail(NIMBUS_EINVAL, "snapshot: null argument");
    Member* m = find(ctx, id);
    if (!m) return fail(NIMBUS_ENOENT, "snapshot: unknown member");
    const SnapSizes z = snap_sizes(*m);
    if (cap < z.total) return fail(NIMBUS_EINVAL, "snapshot: buffer smaller than nimbus_hip_member_snapshot_size");
    hipSetDevice(ctx->device);
    NIMBUS_HIP_TRY(hipStreamSynchronize(ctx->stream));
    std::uint8_t* p = static_cast<std::uint8_t*>(buf);
    SnapHead h{};
    std::memcpy(h.magic, kSnapMagic, 8);
    h.version = kSnapVersion;
    h.endian = 0x01020304u;
    h.grid_bytes = sizeof(nimbus_grid);
    h.config_bytes = sizeof(nimbus_step_config);
    h.precision = m->desc.precision;
    h.tile_w = m->tile_w;
    h.k_chunks = m->k_chunks;
    h.xcd_remap = m->xcd_remap;
    h.hevi_seg = m->args.p.hevi_seg;
    h.steps = m->steps;
    h.field_bytes = z.field;
    h.rain_bytes = z.rain;
    std::memcpy(p, &h, sizeof h);
    p += sizeof h;
    std::memcpy(p, &m->desc.grid, sizeof(nimbus_grid));
    p += sizeof(nimbus_grid);
    std::memcpy(p, &m->desc.config, sizeof(nimbus_step_config));
    p += sizeof(nimbus_step_config);
    std::memcpy(p, m->host_base.data(), z.base);  // slots 0..5: rho0 p0 theta0 qv0 u0 v0
    p += z.base;
    for (int q = 0; q < 9; ++q, p += z.field)
        NIMBUS_HIP_TRY(hipMemcpy(p, prognostic(*m, q), z.field, hipMemcpyDeviceToHost));
    NIMBUS_HIP_TRY(hipMemcpy(p, m->f().rain, z.rain, hipMemcpyDeviceToHost));
    p += z.rain;
    const std::uint64_t sum = snap_checksum(static_cast<const std::uint8_t*>(buf), z.total - sizeof(std::uint64_t));
    std::memcpy(p, &sum, sizeof sum);
    return NIMBUS_OK;
}

int nimbus_hip_member_restore(nimbus_hip_ctx* ctx, const void* buf, uint64_t len, uint32_t* id)
{
    if (!ctx || !buf || !id) return fail(NIMBUS_EINVAL, "restore: null argument");
    const std::uint8_t* p0 = static_cast<const std::uint8_t*>(buf);
    SnapHead h{};
    if (len < sizeof h) return fail(NIMBUS_EINVAL, "restore: truncated snapshot");
    std::memcpy(&h, p0, sizeof h);
    if (std::memcmp(h.magic, kSnapMagic, 8) != 0) return fail(NIMBUS_EINVAL, "restore: not a member snapshot");
    if (h.version != kSnapVersion || h.endian != 0x01020304u || h.grid_bytes != sizeof(nimbus_grid) ||
        h.config_bytes != sizeof(nimbus_step_config))
        return fail(NIMBUS_EINVAL, "restore: snapshot written by an incompatible library version");
    const std::size_t fixed = sizeof h + sizeof(nimbus_grid) + sizeof(nimbus_step_config);
    if (len < fixed) return fail(NIMBUS_EINVAL, "restore: truncated snapshot");
    nimbus_member_desc d{};
    std::memcpy(&d.grid, p0 + sizeof h, sizeof(nimbus_grid));
    std::memcpy(&d.config, p0 + sizeof h + sizeof(nimbus_grid), sizeof(nimbus_step_config));
    d.precision = h.precision;
    if (!valid_grid(d.grid)) return fail(NIMBUS_EINVAL, "restore: invalid grid in snapshot");
    const std::size_t nz = std::size_t(d.grid.nz);
    const std::size_t base_bytes = 6 * nz * sizeof(double);
    const std::size_t total = fixed + base_bytes + 9 * h.field_bytes + h.rain_bytes + sizeof(std::uint64_t);
    if (len != total) return fail(NIMBUS_EINVAL, "restore: snapshot length does not match its header");
    std::uint64_t sum = 0;
    std::memcpy(&sum, p0 + total - sizeof sum, sizeof sum);
    if (sum != snap_checksum(p0, total - sizeof sum)) return fail(NIMBUS_EINVAL, "restore: checksum mismatch");
    std::vector<double> prof(6 * nz);
    std::memcpy(prof.data(), p0 + fixed, base_bytes);
    d.base = nimbus_base_state{prof.data(), prof.data() + nz, prof.data() + 2 * nz, prof.data() + 3 * nz,
                               prof.data() + 4 * nz, prof.data() + 5 * nz};
    const int rc = nimbus_hip_member_create(ctx, &d, id);
    if (rc != NIMBUS_OK) return rc;
    Member* m = find(ctx, *id);
    const SnapSizes z = snap_sizes(*m);
    if (z.field != h.field_bytes || z.rain != h.rain_bytes)
    {
        nimbus_hip_member_destroy(ctx, *id);
        return fail(NIMBUS_EINVAL, "restore: field layout differs from the snapshot's");
    }
    set_geometry(*m, h.tile_w, h.k_chunks, h.xcd_remap);
    if (m->desc.config.acoustic_mode == NIMBUS_ACOUSTIC_HEVI)
    {
        const int rc2 = set_column_solve(*m, h.hevi_seg ? NIMBUS_COLUMN_SOLVE_PARTITIONED : NIMBUS_COLUMN_SOLVE_SEQUENTIAL);
        if (rc2 != NIMBUS_OK)
        {
            nimbus_hip_member_destroy(ctx, *id);
            return rc2;
        }
    }
    m->steps = h.steps;
    const std::uint8_t* p = p0 + fixed + base_bytes;
    for (int q = 0; q < 9; ++q, p += z.field)
    {
        const hipError_t e = hipMemcpy(prognostic(*m, q), p, z.field, hipMemcpyHostToDevice);
        if (e != hipSuccess)
        {
            nimbus_hip_member_destroy(ctx, *id);
            return hip_fail(e, "restore upload");
        }
    }
    const hipError_t e = hipMemcpy(m->f().rain, p, z.rain, hipMemcpyHostToDevice);
    if (e != hipSuccess)
    {
        nimbus_hip_member_destroy(ctx, *id);
        return hip_fail(e, "restore upload");
    }
    return NIMBUS_OK;
}

int nimbus_hip_member_save(nimbus_hip_ctx* ctx, uint32_t id, const char* path)
{
    if (!path) return fail(NIMBUS_EINVAL, "save: null path");
    std::uint64_t n = 0;
    int rc = nimbus_hip_member_snapshot_size(ctx, id, &n);
    if (rc != NIMBUS_OK) return rc;
    std::vector<std::uint8_t> buf(n);
    rc = nimbus_hip_member_snapshot(ctx, id, buf.data(), n);
    if (rc != NIMBUS_OK) return rc;
    std::FILE* f = std::fopen(path, "wb");
    if (!f) return fail(NIMBUS_EINVAL, std::string("save: cannot open ") + path);
    const bool ok = std::fwrite(buf.data(), 1, buf.size(), f) == buf.size();
    if (std::fclose(f) != 0 || !ok) return fail(NIMBUS_EINVAL, std::string("save: write failed: ") + path);
    return NIMBUS_OK;
}

int nimbus_hip_member_load(nimbus_hip_ctx* ctx, const char* path, uint32_t* id)
{
    if (!path) return fail(NIMBUS_EINVAL, "load: null path");
    std::FILE* f = std::fopen(path, "rb");
    if (!f) return fail(NIMBUS_EINVAL, std::string("load: cannot open ") + path);
    std::vector<std::uint8_t> buf;
    std::uint8_t chunk[1 << 16];
    std::size_t got;
    while ((got = std::fread(chunk, 1, sizeof chunk, f)) > 0) buf.insert(buf.end(), chunk, chunk + got);
    std::fclose(f);
    return nimbus_hip_member_restore(ctx, buf.data(), buf.size(), id);
}

int nimbus_hip_step(nimbus_hip_ctx* ctx, const uint32_t* ids, int n_ids, int n_steps)
{
    if (ctx) hipSetDevice(ctx->device);
    return step_members(ctx, ids, n_ids, n_steps, true);
}

int nimbus_hip_step_async(nimbus_hip_ctx* ctx, const uint32_t* ids, int n_ids, int n_steps)
{
    if (ctx) hipSetDevice(ctx->device);
    return step_members(ctx, ids, n_ids, n_steps, false);
}

int nimbus_hip_synchronize(nimbus_hip_ctx* ctx)
{
    if (!ctx) return fail(NIMBUS_EINVAL, "synchronize: null context");
    NIMBUS_HIP_TRY(hipStreamSynchronize(ctx->stream));
    float ms = 0.0f;
    if (hipEventElapsedTime(&ms, ctx->step_a, ctx->step_b) == hipSuccess) ctx->last_step_ms = ms;
    harvest(ctx);
    return NIMBUS_OK;
}

int nimbus_hip_member_set_tiling(nimbus_hip_ctx* ctx, uint32_t id, int tile_w, int k_chunks, int xcd_remap)
{
    if (!ctx) return fail(NIMBUS_EINVAL, "set_tiling: null context");
    Member* m = find(ctx, id);
    if (!m) return fail(NIMBUS_ENOENT, "set_tiling: unknown member");
    if (tile_w != 0 && tile_w != 32 && tile_w != 64)
        return fail(NIMBUS_EINVAL, "set_tiling: tile width must be 0 (linear), 32 or 64");
    if (k_chunks < 1 || k_chunks > m->g().nz) return fail(NIMBUS_EINVAL, "set_tiling: k_chunks must be in 1..nz");
    drop_batches(ctx);
    set_geometry(*m, tile_w, k_chunks, xcd_remap);
    return NIMBUS_OK;
}

int nimbus_hip_member_set_column_solve(nimbus_hip_ctx* ctx, uint32_t id, int form)
{
    if (!ctx) return fail(NIMBUS_EINVAL, "set_column_solve: null context");
    Member* m = find(ctx, id);
    if (!m) return fail(NIMBUS_ENOENT, "set_column_solve: unknown member");
    const int before = m->args.p.hevi_seg;
    const int rc = set_column_solve(*m, form);
    if (rc == NIMBUS_OK && m->args.p.hevi_seg != before) drop_batches(ctx);
    return rc;
}

int nimbus_hip_member_column_solve(nimbus_hip_ctx* ctx, uint32_t id, int* form)
{
    if (!ctx || !form) return fail(NIMBUS_EINVAL, "column_solve: null argument");
    Member* m = find(ctx, id);
    if (!m) return fail(NIMBUS_ENOENT, "column_solve: unknown member");
    *form = m->desc.config.acoustic_mode != NIMBUS_ACOUSTIC_HEVI ? NIMBUS_COLUMN_SOLVE_AUTO
            : m->args.p.hevi_seg                                  ? NIMBUS_COLUMN_SOLVE_PARTITIONED
                                                                  : NIMBUS_COLUMN_SOLVE_SEQUENTIAL;
    return NIMBUS_OK;
}

int nimbus_hip_member_autotune(nimbus_hip_ctx* ctx, uint32_t id, int trials, int* best_tile_w, int* best_k_chunks,
                               int* best_remap, double* best_ms)
{
    if (!ctx || trials < 1) return fail(NIMBUS_EINVAL, "autotune: bad arguments");
    Member* m = find(ctx, id);
    if (!m) return fail(NIMBUS_ENOENT, "autotune: unknown member");
    hipSetDevice(ctx->device);
    void* saved = nullptr;
    NIMBUS_HIP_TRY(hipMalloc(&saved, m->bytes));
    // Every exit restores the member's device block, step count, launch
    // geometry (the best one found on success) and the context's timing
    // switch, so a failed trial leaves the member as it was.
    struct Restore
    {
        nimbus_hip_ctx* ctx;
        Member* m;
        void* saved;
        bool copied = false;
        std::int64_t steps;
        int w, k, r;
        bool timing;
        ~Restore()
        {
            hipStreamSynchronize(ctx->stream);
            if (copied) hipMemcpy(m->block, saved, m->bytes, hipMemcpyDeviceToDevice);
            hipFree(saved);
            m->steps = steps;
            ctx->timing = timing;
            drop_batches(ctx);
            set_geometry(*m, w, k, r);
        }
    } guard{ctx, m, saved, false, m->steps, m->tile_w, m->k_chunks, m->xcd_remap, ctx->timing};
    NIMBUS_HIP_TRY(hipMemcpyAsync(saved, m->block, m->bytes, hipMemcpyDeviceToDevice, ctx->stream));
    NIMBUS_HIP_TRY(hipStreamSynchronize(ctx->stream));
    guard.copied = true;
    ctx->timing = false;
    double best = 1e300;
    int bw = m->tile_w, bk = m->k_chunks, br = m->xcd_remap;
    const std::uint32_t ids[1] = {id};
    const int nz = m->g().nz;
    std::vector<int> chunks;
    for (int c : {1, 2, 3, 5, 8, 12})
        if (c <= std::max(1, nz / 4)) chunks.push_back(c);
    chunks.push_back(nz);  // plane-by-plane order
    for (int w : {64, 32, 0})
        for (int c : chunks)
            for (int r : {1, 0})
            {
                drop_batches(ctx);
                set_geometry(*m, w, c, r);
                int rc = step_members(ctx, ids, 1, 1, true);  // warm-up (and graph capture)
                if (rc == NIMBUS_OK) rc = step_members(ctx, ids, 1, trials, true);
                if (rc != NIMBUS_OK) return rc;
                const double ms = ctx->last_step_ms / trials;
                if (ms < best)
                {
                    best = ms;
                    bw = w;
                    bk = c;
                    br = r;
                }
                NIMBUS_HIP_TRY(hipMemcpyAsync(m->block, saved, m->bytes, hipMemcpyDeviceToDevice, ctx->stream));
            }
    guard.w = bw;
    guard.k = bk;
    guard.r = br;
    if (best_tile_w) *best_tile_w = bw;
    if (best_k_chunks) *best_k_chunks = bk;
    if (best_remap) *best_remap = br;
    if (best_ms) *best_ms = best;
    return NIMBUS_OK;
}

int nimbus_hip_set_graphs(nimbus_hip_ctx* ctx, int enable)
{
    if (!ctx) return fail(NIMBUS_EINVAL, "set_graphs: null context");
    ctx->graphs = enable != 0;
    return NIMBUS_OK;
}

int nimbus_hip_set_kernel_timing(nimbus_hip_ctx* ctx, int enable)
{
    if (!ctx) return fail(NIMBUS_EINVAL, "set_kernel_timing: null context");
    ctx->timing = enable != 0;
    return NIMBUS_OK;
}

int nimbus_hip_kernel_timing(nimbus_hip_ctx* ctx, int max_rows, char (*names)[48], uint64_t* launches,
                             double* total_ms)
{
    if (!ctx || max_rows < 0) return fail(NIMBUS_EINVAL, "kernel_timing: bad arguments");
    int r = 0;
    for (auto& kv : ctx->table)
    {
        if (r >= max_rows) break;
        if (names) std::snprintf(names[r], 48, "%s", kv.first.c_str());
        if (launches) launches[r] = kv.second.launches;
        if (total_ms) total_ms[r] = kv.second.ms;
        ++r;
    }
    return r;
}

int nimbus_hip_reset_kernel_timing(nimbus_hip_ctx* ctx)
{
    if (!ctx) return fail(NIMBUS_EINVAL, "reset_kernel_timing: null context");
    hipStreamSynchronize(ctx->stream);
    harvest(ctx);
    ctx->table.clear();
    return NIMBUS_OK;
}

double nimbus_hip_last_step_ms(nimbus_hip_ctx* ctx) { return ctx ? ctx->last_step_ms : 0.0; }

// ---------------- per-operation dispatch ----------------

int nimbus_hip_dispatch_acoustic_substep_fused(nimbus_hip_ctx* ctx, const nimbus_acoustic_substep_fused_dispatch* d)
{
    // Validation mirrors valid_acoustic_cgrid_dispatch (src/compute/cpu/acoustic_cgrid.cpp:11-46).
    if (!ctx || !d) return fail(NIMBUS_EINVAL, "acoustic: null argument");
    auto pos = [](double x) { return std::isfinite(x) && x > 0.0; };
    auto bounds_ok = [](float lo, float hi) { return std::isfinite(lo) && std::isfinite(hi) && lo > 0.0f && lo <= hi; };
    if (!all_nonnull<const void*>({d->u, d->v, d->w, d->rho, d->p, d->theta_data, d->rho0_base_host, d->p0_base_host}) ||
        d->nx < 5 || d->ny < 5 || d->nz < 4 || d->base_len < d->nz || !pos(d->dx_host) || !pos(d->dy_host) ||
        !pos(d->dz_host) || !pos(d->dt_small_host) || !pos(d->gamma_host) || !bounds_ok(d->rho_min, d->rho_max) ||
        !bounds_ok(d->p_min, d->p_max) || !pos(d->p_ref_fallback) || !(d->wind_clamp_h >= 0.0f) ||
        !(d->wind_clamp_v >= 0.0f) || std::fabs(d->gamma_host - kGamma) > 1e-12)
        return fail(NIMBUS_EINVAL, "acoustic: invalid dispatch");
    hipSetDevice(ctx->device);
    hipStreamSynchronize(ctx->stream);
    ScratchMember sm{};
    int rc = make_scratch(sm, d->nx, d->ny, d->nz, d->dx_host, d->dy_host, d->dz_host, d->rho0_base_host,
                          d->p0_base_host, nullptr, nullptr, nullptr, nullptr, NIMBUS_ACOUSTIC_EXPLICIT);
    if (rc != NIMBUS_OK) return rc;
    Member& m = sm.m;
    StepParams& p = m.args.p;
    p.bounds = Bounds{d->rho_min, d->rho_max, d->p_min, d->p_max, d->p_ref_fallback, d->wind_clamp_h,
                      d->wind_clamp_v};
    p.substeps = 1;
    p.dts = d->dt_small_host;
    p.dt = d->dt_small_host;
    p.mf_weight = 0.0;
    if ((rc = sm.finish_setup()) != NIMBUS_OK) return rc;
    std::vector<float> zero(std::size_t(d->nx) * d->ny * d->nz, 0.0f);
    nimbus_host_state hs{d->u, d->v, d->w, d->rho, d->p, const_cast<float*>(d->theta_data), zero.data(), zero.data(),
                         zero.data(), nullptr};
    rc = upload_state<float>(ctx, m, hs);
    if (rc != NIMBUS_OK) return rc;
    Launcher L = make_launcher(ctx);
    launch_acoustic_substep<float>(L, sm.batch.view, 0, false, false);
    launch_commit<float>(L, sm.batch.view, 1);
    if (L.error != hipSuccess) return hip_fail(L.error, "acoustic launch");
    nimbus_host_state out{d->u, d->v, d->w, d->rho, d->p, nullptr, nullptr, nullptr, nullptr, nullptr};
    rc = download_state<float>(ctx, m, out);
    harvest(ctx);
    return rc;
}

int nimbus_hip_dispatch_acoustic_substeps_batched(nimbus_hip_ctx* ctx,
                                                  const nimbus_acoustic_substeps_batched_dispatch* d)
{
    if (!ctx || !d) return fail(NIMBUS_EINVAL, "batched acoustic: null argument");
    if (!all_nonnull<const void*>({d->u, d->v, d->w, d->rho, d->p, d->theta_data, d->rho0_base_host, d->p0_base_host}) ||
        d->n_substeps < 1 || d->n_substeps > 10000 || !(d->dt_small > 0.0) ||
        (d->accumulate_mass_flux && !all_nonnull<const void*>({d->mass_flux_x, d->mass_flux_y, d->mass_flux_z})))
        return fail(NIMBUS_EINVAL, "batched acoustic: invalid dispatch");
    nimbus_step_config cfg{};
    cfg.dt = d->dt_small * d->n_substeps;
    cfg.acoustic_mode = d->acoustic_mode;
    cfg.hevi_beta = d->hevi_beta;
    cfg.bc_west = d->bc_west; cfg.bc_east = d->bc_east; cfg.bc_south = d->bc_south; cfg.bc_north = d->bc_north;
    cfg.bc_bottom = d->bc_bottom; cfg.bc_top = d->bc_top;
    cfg.vt_max = 20.0;
    if (!valid_config(cfg)) return fail(NIMBUS_EINVAL, "batched acoustic: invalid boundary or mode");
    hipSetDevice(ctx->device);
    hipStreamSynchronize(ctx->stream);
    ScratchMember sm{};
    int rc = make_scratch(sm, d->nx, d->ny, d->nz, d->dx, d->dy, d->dz, d->rho0_base_host, d->p0_base_host, nullptr,
                          nullptr, nullptr, nullptr, d->acoustic_mode);
    if (rc != NIMBUS_OK) return rc;
    Member& m = sm.m;
    StepParams& p = m.args.p;
    p.acoustic_mode = d->acoustic_mode;
    p.hevi_alpha = 0.5 * (1.0 + d->hevi_beta);
    p.bc_west = d->bc_west; p.bc_east = d->bc_east; p.bc_south = d->bc_south; p.bc_north = d->bc_north;
    p.bc_bottom = d->bc_bottom; p.bc_top = d->bc_top;
    p.substeps = d->n_substeps;
    p.dts = d->dt_small;
    p.dt = d->dt_small * d->n_substeps;
    p.mf_weight = d->mass_flux_weight;
    if ((rc = sm.finish_setup()) != NIMBUS_OK) return rc;
    std::vector<float> zero(std::size_t(d->nx) * d->ny * d->nz, 0.0f);
    nimbus_host_state hs{d->u, d->v, d->w, d->rho, d->p, const_cast<float*>(d->theta_data), zero.data(), zero.data(),
                         zero.data(), nullptr};
    rc = upload_state<float>(ctx, m, hs);
    if (rc != NIMBUS_OK) return rc;
    Staging st;
    const std::size_t n = zero.size();
    NIMBUS_HIP_TRY(hipMalloc(&st.p, n * sizeof(float)));
    const DevFields& f = m.f();
    if (d->accumulate_mass_flux)
    {
        const void* src[3] = {d->mass_flux_x, d->mass_flux_y, d->mass_flux_z};
        void* dst[3] = {f.mfx, f.mfy, f.mfz};
        for (int q = 0; q < 3; ++q)
        {
            rc = upload_field<float>(ctx, m, src[q], dst[q], nullptr, st.p);
            if (rc != NIMBUS_OK) return rc;
            NIMBUS_HIP_TRY(hipStreamSynchronize(ctx->stream));
        }
    }
    Launcher L = make_launcher(ctx);
    launch_seed_radiation<float>(L, sm.batch.view);
    for (int s = 0; s < d->n_substeps; ++s)
    {
        launch_acoustic_substep<float>(L, sm.batch.view, s, d->accumulate_mass_flux != 0, false);
        launch_boundaries<float>(L, sm.batch.view, s);
    }
    launch_commit<float>(L, sm.batch.view, d->n_substeps);
    if (L.error != hipSuccess) return hip_fail(L.error, "batched acoustic launch");
    nimbus_host_state out{d->u, d->v, d->w, d->rho, d->p, nullptr, nullptr, nullptr, nullptr, nullptr};
    rc = download_state<float>(ctx, m, out);
    if (rc != NIMBUS_OK) return rc;
    if (d->accumulate_mass_flux)
    {
        void* dst[3] = {d->mass_flux_x, d->mass_flux_y, d->mass_flux_z};
        const void* src[3] = {f.mfx, f.mfy, f.mfz};
        for (int q = 0; q < 3; ++q)
        {
            rc = download_field<float>(ctx, m, src[q], dst[q], nullptr, st.p);
            if (rc != NIMBUS_OK) return rc;
        }
    }
    harvest(ctx);
    return NIMBUS_OK;
}

int nimbus_hip_dispatch_slow_tendencies(nimbus_hip_ctx* ctx, const nimbus_slow_tendencies_dispatch* d)
{
    if (!ctx || !d) return fail(NIMBUS_EINVAL, "slow tendencies: null argument");
    if (!all_nonnull<const void*>({d->u_data, d->v_data, d->w_data, d->rho_data, d->qv_data, d->qc_data, d->qr_data,
                                   d->rho0_base, d->du_dt_data, d->dv_dt_data, d->dw_dt_data}))
        return fail(NIMBUS_EINVAL, "slow tendencies: null field");
    hipSetDevice(ctx->device);
    hipStreamSynchronize(ctx->stream);
    std::vector<double> p0(d->nz, kP00);
    ScratchMember sm{};
    int rc = make_scratch(sm, d->nr, d->nth, d->nz, d->dx, d->dy, d->dz, d->rho0_base, p0.data(), nullptr,
                          d->qv0_base, d->u0_base, d->v0_base, NIMBUS_ACOUSTIC_EXPLICIT);
    if (rc != NIMBUS_OK) return rc;
    Member& m = sm.m;
    m.args.p.coriolis_f = d->coriolis_f;
    if ((rc = sm.finish_setup()) != NIMBUS_OK) return rc;
    const std::size_t n = std::size_t(d->nr) * d->nth * d->nz;
    std::vector<float> zero(n, 0.0f);
    nimbus_host_state hs{const_cast<float*>(d->u_data), const_cast<float*>(d->v_data), const_cast<float*>(d->w_data),
                         const_cast<float*>(d->rho_data), zero.data(), zero.data(), const_cast<float*>(d->qv_data),
                         const_cast<float*>(d->qc_data), const_cast<float*>(d->qr_data), nullptr};
    rc = upload_state<float>(ctx, m, hs);
    if (rc != NIMBUS_OK) return rc;
    const DevFields& f = m.f();
    Launcher L = make_launcher(ctx);
    launch_slow_tendencies<float>(L, sm.batch.view, f.k1u, f.k1v, f.k1w);
    if (L.error != hipSuccess) return hip_fail(L.error, "slow tendencies launch");
    Staging st;
    NIMBUS_HIP_TRY(hipMalloc(&st.p, n * sizeof(float)));
    void* outs[3] = {d->du_dt_data, d->dv_dt_data, d->dw_dt_data};
    void* srcs[3] = {f.k1u, f.k1v, f.k1w};
    for (int q = 0; q < 3; ++q)
    {
        rc = download_field<float>(ctx, m, srcs[q], outs[q], nullptr, st.p);
        if (rc != NIMBUS_OK) return rc;
    }
    harvest(ctx);
    return NIMBUS_OK;
}

int nimbus_hip_dispatch_slow_rk2(nimbus_hip_ctx* ctx, const nimbus_slow_rk2_dispatch* d)
{
    if (!ctx || !d) return fail(NIMBUS_EINVAL, "slow rk2: null argument");
    if (!all_nonnull<const void*>({d->u, d->v, d->w, d->rho, d->qv, d->qc, d->qr, d->rho0_base}) || !(d->dt_large > 0))
        return fail(NIMBUS_EINVAL, "slow rk2: invalid dispatch");
    hipSetDevice(ctx->device);
    hipStreamSynchronize(ctx->stream);
    std::vector<double> p0(d->nz, kP00);
    ScratchMember sm{};
    int rc = make_scratch(sm, d->nx, d->ny, d->nz, d->dx, d->dy, d->dz, d->rho0_base, p0.data(), nullptr,
                          d->qv0_base, d->u0_base, d->v0_base, NIMBUS_ACOUSTIC_EXPLICIT);
    if (rc != NIMBUS_OK) return rc;
    Member& m = sm.m;
    m.args.p.coriolis_f = d->coriolis_f;
    m.args.p.dt = d->dt_large;
    if ((rc = sm.finish_setup()) != NIMBUS_OK) return rc;
    const std::size_t n = std::size_t(d->nx) * d->ny * d->nz;
    std::vector<float> zero(n, 0.0f);
    nimbus_host_state hs{d->u, d->v, d->w, const_cast<float*>(d->rho), zero.data(), zero.data(),
                         const_cast<float*>(d->qv), const_cast<float*>(d->qc), const_cast<float*>(d->qr), nullptr};
    rc = upload_state<float>(ctx, m, hs);
    if (rc != NIMBUS_OK) return rc;
    Launcher L = make_launcher(ctx);
    launch_slow_rk2<float>(L, sm.batch.view);
    if (L.error != hipSuccess) return hip_fail(L.error, "slow rk2 launch");
    nimbus_host_state out{d->u, d->v, d->w, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr};
    rc = download_state<float>(ctx, m, out);
    harvest(ctx);
    return rc;
}

int nimbus_hip_dispatch_advect_scalars(nimbus_hip_ctx* ctx, const nimbus_advect_scalars_dispatch* d)
{
    if (!ctx || !d) return fail(NIMBUS_EINVAL, "advect: null argument");
    if (d->n_scalars < 0 || d->n_scalars > 4 ||
        !all_nonnull<const void*>({d->mass_flux_x, d->mass_flux_y, d->mass_flux_z, d->rho_begin, d->rho_end}) ||
        !(d->dt > 0.0) || (d->bc_top != NIMBUS_TOP_RIGID && d->bc_top != NIMBUS_TOP_OPEN))
        return fail(NIMBUS_EINVAL, "advect: invalid dispatch");
    int npd = 0;
    for (int s = 0; s < d->n_scalars; ++s)
    {
        if (!d->q[s]) return fail(NIMBUS_EINVAL, "advect: null scalar");
        npd += d->positive_definite[s] != 0;
    }
    if (npd > 3) return fail(NIMBUS_EINVAL, "advect: at most three positive-definite scalars");
    if (d->n_scalars == 0) return NIMBUS_OK;
    hipSetDevice(ctx->device);
    hipStreamSynchronize(ctx->stream);
    std::vector<double> rho0(d->nz, 0.0), p0(d->nz, kP00);
    ScratchMember sm{};
    int rc = make_scratch(sm, d->nx, d->ny, d->nz, d->dx, d->dy, d->dz, rho0.data(), p0.data(), nullptr, nullptr,
                          nullptr, nullptr, NIMBUS_ACOUSTIC_EXPLICIT);
    if (rc != NIMBUS_OK) return rc;
    Member& m = sm.m;
    m.args.p.dt = d->dt;
    m.args.p.bc_top = d->bc_top;
    if ((rc = sm.finish_setup()) != NIMBUS_OK) return rc;
    const std::size_t n = std::size_t(d->nx) * d->ny * d->nz;
    Staging st;
    NIMBUS_HIP_TRY(hipMalloc(&st.p, n * sizeof(float)));
    const DevFields& f = m.f();
    // rho0 = 0 in the scratch member, so rho' slots hold the full densities.
    const void* src[5] = {d->mass_flux_x, d->mass_flux_y, d->mass_flux_z, d->rho_begin, d->rho_end};
    void* dst[5] = {f.mfx, f.mfy, f.mfz, f.rb, f.rp};
    for (int q = 0; q < 5; ++q)
    {
        rc = upload_field<float>(ctx, m, src[q], dst[q], nullptr, st.p);
        if (rc != NIMBUS_OK) return rc;
        NIMBUS_HIP_TRY(hipStreamSynchronize(ctx->stream));
    }
    void* qin[4] = {f.th, f.qv, f.qc, f.qr};
    void* qout[4] = {f.tth, f.tqv, f.tqc, f.tqr};
    int pd[4] = {0, 0, 0, 0};
    for (int s = 0; s < d->n_scalars; ++s)
    {
        rc = upload_field<float>(ctx, m, d->q[s], qin[s], nullptr, st.p);
        if (rc != NIMBUS_OK) return rc;
        NIMBUS_HIP_TRY(hipStreamSynchronize(ctx->stream));
        pd[s] = d->positive_definite[s] != 0;
    }
    Launcher L = make_launcher(ctx);
    launch_transport<float>(L, sm.batch.view, d->n_scalars, pd, true);
    if (L.error != hipSuccess) return hip_fail(L.error, "advect launch");
    for (int s = 0; s < d->n_scalars; ++s)
    {
        rc = download_field<float>(ctx, m, qout[s], d->q[s], nullptr, st.p);
        if (rc != NIMBUS_OK) return rc;
    }
    harvest(ctx);
    return NIMBUS_OK;
}

int nimbus_hip_dispatch_kessler(nimbus_hip_ctx* ctx, const nimbus_kessler_dispatch* d)
{
    if (!ctx || !d) return fail(NIMBUS_EINVAL, "kessler: null argument");
    if (!all_nonnull<const void*>({d->p, d->theta, d->rho, d->qv, d->qc, d->qr, d->dtheta_dt, d->dqv_dt, d->dqc_dt,
                                   d->dqr_dt}) ||
        !(d->dt > 0.0) || !(d->dz > 0.0) || !(d->vt_max > 0.0))
        return fail(NIMBUS_EINVAL, "kessler: invalid dispatch");
    hipSetDevice(ctx->device);
    hipStreamSynchronize(ctx->stream);
    ScratchMember sm{};
    // p0 = rho0 = 0: the perturbation slots carry the full fields.
    nimbus_member_desc desc{};
    desc.grid = nimbus_grid{d->nx, d->ny, d->nz, 1.0, 1.0, d->dz};
    if (!valid_grid(desc.grid)) return fail(NIMBUS_EINVAL, "kessler: invalid grid");
    desc.config.dt = d->dt;
    desc.config.vt_max = d->vt_max;
    desc.precision = NIMBUS_PRECISION_MIXED;
    std::vector<double> one(d->nz, 1.0);
    desc.base = nimbus_base_state{one.data(), one.data(), nullptr, nullptr, nullptr, nullptr};
    int rc = allocate_member(desc, sm.m);
    if (rc != NIMBUS_OK) return rc;
    Member& m = sm.m;
    const DevFields& f = m.f();
    NIMBUS_HIP_TRY(hipMemset(f.base, 0, sizeof(double) * kBaseCount * d->nz));
    m.args.p.kessler = KesslerConst{d->qc0, d->c_auto, d->c_accr, d->evap_eff, d->vt_max};
    if ((rc = sm.finish_setup()) != NIMBUS_OK) return rc;
    const std::size_t n = std::size_t(d->nx) * d->ny * d->nz;
    std::vector<float> zero(n, 0.0f);
    nimbus_host_state hs{zero.data(), zero.data(), zero.data(), const_cast<float*>(d->rho), const_cast<float*>(d->p),
                         const_cast<float*>(d->theta), const_cast<float*>(d->qv), const_cast<float*>(d->qc),
                         const_cast<float*>(d->qr), nullptr};
    // upload_state subtracts base slots that are zero here
    rc = upload_state<float>(ctx, m, hs);
    if (rc != NIMBUS_OK) return rc;
    Launcher L = make_launcher(ctx);
    launch_kessler_tendencies<float>(L, sm.batch.view, d->dt, f.k1u, f.k1v, f.k1w, f.u2);
    if (L.error != hipSuccess) return hip_fail(L.error, "kessler launch");
    Staging st;
    NIMBUS_HIP_TRY(hipMalloc(&st.p, n * sizeof(float)));
    void* outs[4] = {d->dtheta_dt, d->dqv_dt, d->dqc_dt, d->dqr_dt};
    void* srcs[4] = {f.k1u, f.k1v, f.k1w, f.u2};
    for (int q = 0; q < 4; ++q)
    {
        rc = download_field<float>(ctx, m, srcs[q], outs[q], nullptr, st.p);
        if (rc != NIMBUS_OK) return rc;
    }
    harvest(ctx);
    return NIMBUS_OK;
}

}  // extern "C"
