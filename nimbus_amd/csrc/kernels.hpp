// Launch interface of the large-step kernels.
//
// Every kernel works on a batch of members (independent domains of possibly
// different shapes, one precision per batch). The batch keeps a device array
// of MemberArgs (grid, field pointers, step parameters) and, per launch
// shape, a block map: entry b of the map names the member and the member-local
// work tile of workgroup b. One launch therefore covers every member of the
// batch, and the map order decides which XCD runs which tiles.
#pragma once

#include "nimbus_internal.hpp"

#include <string>
#include <vector>

namespace nimbus
{

struct TimedLaunch
{
    const char* name;
    hipEvent_t start, stop;
};

struct Launcher
{
    hipStream_t stream = nullptr;
    bool timing = false;
    std::vector<TimedLaunch>* pending = nullptr;
    std::vector<hipEvent_t>* event_pool = nullptr;
    std::uint64_t* launches = nullptr;
    hipError_t error = hipSuccess;

    hipEvent_t take_event();
    void begin(const char* name, hipEvent_t& a);
    void end(const char* name, hipEvent_t a);
};

// Device field set of one member (precision-erased pointers). rho' has a
// second buffer: acoustic substep s reads buffer s & 1 and writes (s + 1) & 1.
struct DevFields
{
    void *u, *v, *w, *rp, *pp, *th, *qv, *qc, *qr;       // prognostic state (rho, p as perturbations)
    void *u2, *v2, *w2, *rp2;                            // Heun predictor u, v, w; second rho' buffer
    void *mfx, *mfy, *mfz, *rb;                          // retained mass fluxes, rho' at step begin
    void *k1u, *k1v, *k1w;                               // Heun stage-1 tendencies
    void *tth, *tqv, *tqc, *tqr;                         // transported scalars before microphysics
    void* flux[12];                                      // per scalar: x, y, z face fluxes (7 alias dead slots)
    float *rl_qv, *rl_qc, *rl_qr;                        // PD limiter factors
    double *tri_c, *tri_r;                               // HE-VI Thomas coefficients (FP64)
    float *rad_w, *rad_e, *rad_s, *rad_n;                // KW78 radiative stashes
    double* rain;                                        // accumulated surface rain [ny][nx]
    double* base;                                        // kBaseCount * nz FP64 profiles
    double* diag;                                        // kDiagBlocksPerMember * kDiagValues partials + result
    double* sponge_lat;                                  // lateral sponge factor by distance to the nearest face
    double* clock;                                       // microphysics subcycle clock: elapsed, step, limit bits
};

struct StepParams
{
    double dt;          // large step
    double dts;         // acoustic substep
    double mf_weight;   // retained mass-flux weight per substep (dts / dt in the large step)
    int substeps;
    int acoustic_mode;
    int hevi_seg;       // the member's HE-VI columns are solved in the partitioned form (fixed per member)
    double hevi_alpha;  // (1 + beta) / 2
    double coriolis_f;
    int bc_west, bc_east, bc_south, bc_north, bc_bottom, bc_top;
    int kessler_sub;    // microphysics subcycle launches per large step
    int kessler_adaptive; // subcycle lengths follow the state's sedimentation limit (device clock)
    int sponge;         // Rayleigh sponges on
    int sponge_lateral; // lateral band width (cells)
    // Smagorinsky SGS: sgs_n substeps of sgs_h seconds, the last sgs_h_last;
    // sgs_n = 0 when off
    int sgs_n;
    double sgs_h, sgs_h_last;
    int sgs_mode;                 // NIMBUS_SGS_*
    double sgs_c2d2;              // Cs^2 Delta^2
    double sgs_inv_pr, sgs_inv_sc; // 1 / Pr_t, 1 / Sc_t
    double sgs_nu_max, sgs_k_max;
    int sgs_wrap_x, sgs_wrap_y;   // periodic topology
    int sgs_open[4];              // west, east, south, north lateral faces open (radiative or zero-gradient)
    // explicit diffusion after the SGS: NIMBUS_DIFFUSE_* bits (0 off); when on,
    // the EOS projection and the sponges follow its commit instead of the
    // microphysics or SGS pass
    int dif_apply;
    double dif_kh, dif_kv;
    double dif_idx2, dif_idy2;  // 1 / (dx dx), 1 / (dy dy)
    int dif_wrap_x, dif_wrap_y;
    KesslerConst kessler;
    Bounds bounds;
};

// Launch geometry of one member: workgroups of tile_w x (256 / tile_w)
// columns (four waves of 64 lanes). Per-cell kernels visit the levels of a
// column tile consecutively within each of k_chunks vertical chunks
// (k_chunks = nz reproduces plane-by-plane order).
struct MemberGeom
{
    int tile_w, tile_h, tiles_x, tiles_y;
    int k_chunks;
    int xcd_remap;
    // per-cell work in runs of kBlock consecutive elements of one level's
    // padded [j][i] plane (`runs` per level) instead of column tiles at one
    // level, per-column work in runs of kBlock consecutive columns
    int linear, runs;
};

struct MemberArgs
{
    DevGrid g;
    MemberGeom geo;
    DevFields f;
    StepParams p;
};

// Work shapes of a launch: what one map entry (one workgroup) covers.
enum WorkKind
{
    W_CELL = 0,    // one tile of columns at one level, or one run of a level (one cell per thread)
    W_COLUMN = 1,  // one tile, whole columns (one column per thread)
    W_WE = 2,      // 256 (j, k) rows of the west/east faces
    W_SN = 3,      // 256 (i, k) rows of the south/north faces
    W_SEED = 4,    // 256 entries of the radiative stashes
    W_DIAG = 5,    // one of kDiagBlocksPerMember reduction blocks
    W_SEGCOL = 6,  // kSegCols consecutive columns of one row, kSegments runs of faces each
    W_TILECELL = 7,  // one tile of columns at one level whatever the geometry (wide stencils)
    W_KINDS = 8
};

// Partitioned HE-VI column solve: a column of nz - 1 unknown faces is split
// into kSegments runs of at most kSegRows faces, one thread per run, for
// columns of at most kSegMaxNz levels. The form is a property of the member,
// fixed at creation (StepParams::hevi_seg, default from the member's own
// shape, nimbus_hip_member_set_column_solve to override), so a member's
// arithmetic never depends on the batch it runs in. Default: FP64 members
// always (measured faster than the sequential sweep, which is memory-bound
// on its FP64 coefficient arrays); FP32-storage members with fewer than
// kSegSmallMemberColumns columns, which alone cannot fill the GPU with one
// thread per column (measured: 64x64x40 0.49 -> 0.36 ms per large step,
// 160x160x80 1.29 -> 1.23; by 255x255 columns the partitioned form's
// latency-bound FP64 elimination is the slower one, and at 400x400x100
// 9.06 against 8.04). Deeper columns take the sequential sweep, which keeps
// its FP64 coefficients in two cell-sized scratch arrays (allocated for every
// member that may be solved by it: FP32 storage or more than kSegMaxNz levels).
#ifndef NIMBUS_SEG_ROWS
#define NIMBUS_SEG_ROWS 8
#endif
constexpr int kSegRows = NIMBUS_SEG_ROWS;        // faces per run (one thread)
constexpr int kSegments = 128 / kSegRows;        // runs per column: columns of up to 129 levels
constexpr int kSegCols = 256 / kSegments;        // columns per 256-thread workgroup
constexpr int kSegSteps = kSegments == 64 ? 6 : kSegments == 32 ? 5 : kSegments == 16 ? 4 : kSegments == 8 ? 3 : 0;
static_assert(kSegSteps > 0 && (1 << kSegSteps) == kSegments, "runs per column must be 8, 16, 32 or 64");
constexpr int kSegMaxNz = kSegments * kSegRows + 1;
constexpr int kSegSmallMemberColumns = 32768;
constexpr bool hevi_partitioned_always(int elem, int nz) { return elem == 8 && nz <= kSegMaxNz; }
// Default column-solve form of a member: 1 = partitioned, 0 = sequential sweep.
constexpr int hevi_default_seg(int elem, int nx, int ny, int nz)
{
    return nz <= kSegMaxNz && (elem == 8 || long(nx) * long(ny) < long(kSegSmallMemberColumns)) ? 1 : 0;
}

// Member subsets a launch can be restricted to.
enum MemberSubset
{
    S_ALL = 0,
    S_EXPLICIT = 1,
    S_HEVI = 2,
    S_SUBSETS = 3
};

struct BlockMap
{
    const int2* entries = nullptr;  // device: {member index, local work index}
    int blocks = 0;
};

struct BatchView
{
    const MemberArgs* args = nullptr;  // device array
    int n = 0;
    int max_substeps = 0;
    int max_kessler_sub = 0;
    int any_sponge = 0;
    int max_sgs_n = 0;
    // parities that decide whether a copy-home kernel has work: some member
    // with an odd substep count (acoustic commit), an even microphysics
    // subcycle count (scalars home), an odd SGS substep count (SGS home)
    int any_odd_substeps = 0, any_even_kessler_sub = 0, any_odd_sgs_n = 0;
    int any_sponge_unfused = 0;  // a member with sponges and no SGS or diffusion pass to apply them in
    int any_diffusion = 0;       // a member with explicit diffusion
    int any_kessler_adaptive = 0; // a member whose microphysics subcycles follow its sedimentation limit
    int any_hevi_seg = 0;        // HE-VI members solved in the partitioned form
    int any_hevi_serial = 0;     // HE-VI members solved by the sequential sweep
    int max_nz = 0;
    int elem = 4;
    BlockMap maps[S_SUBSETS][W_KINDS];
};

// Number of local work items of `kind` for one member.
int work_items(const MemberArgs& a, WorkKind kind);
void fill_geometry(MemberArgs& a, int tile_w, int k_chunks, int xcd_remap);

// Whole large step of every member of the batch.
template <class S>
void launch_large_step(Launcher& L, const BatchView& b);

// Individual phases (per-operation dispatch entry points).
template <class S>
void launch_slow_rk2(Launcher& L, const BatchView& b);
template <class S>
void launch_slow_tendencies(Launcher& L, const BatchView& b, void* du, void* dv, void* dw);
// Substep s of the acoustic loop (0-based): reads buffer s & 1, writes (s + 1) & 1.
template <class S>
void launch_acoustic_substep(Launcher& L, const BatchView& b, int s, bool accumulate, bool init_at_first);
// Boundary closure; s >= 0 selects the buffers after substep s and radiates
// with the member's dts; s < 0 closes the home buffers without radiation.
template <class S>
void launch_boundaries(Launcher& L, const BatchView& b, int s);
template <class S>
void launch_seed_radiation(Launcher& L, const BatchView& b);
// Moves rho' from buffer 1 to buffer 0 for members whose last substep (n_done - 1) left it there.
template <class S>
void launch_commit(Launcher& L, const BatchView& b, int n_done);
template <class S>
void launch_transport(Launcher& L, const BatchView& b, int n_scalars, const int* pd, bool into_temp);
template <class S>
void launch_kessler_tendencies(Launcher& L, const BatchView& b, double dt, void* dth, void* dqv, void* dqc,
                               void* dqr);
template <class S>
void launch_diagnostics(Launcher& L, const BatchView& b);
template <class H, class S>
void launch_host_to_dev(Launcher& L, const DevGrid& g, const void* src, void* dst, const double* base_slot);
template <class S, class H>
void launch_dev_to_host(Launcher& L, const DevGrid& g, const void* src, void* dst, const double* base_slot);

constexpr int kDiagValues = 8;
constexpr int kDiagBlocksPerMember = 64;

}  // namespace nimbus
