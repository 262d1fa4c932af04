/*
 * nimbus_hip.h - C ABI of the MI355X (gfx950) HIP compute backend for the
 * Nimbus Klemp-Wilhelmson split-explicit large step.
 *
 * Conventions
 *  - Every entry point returns 0 on success and a negative NIMBUS_E* code on
 *    failure; nimbus_hip_last_error() returns the thread's last message. A
 *    failed dispatch leaves caller-owned host fields untouched, as the
 *    reference dispatch contract requires (include/compute/dispatch/dynamics.hpp:682-683).
 *  - Host 3-D arrays use the reference layout [nx][ny][nz] with k contiguous,
 *    index (i*ny + j)*nz + k (include/core/field/field3d.hpp:543-574). nx and ny
 *    include the one-cell lateral ghost layer on each side.
 *  - Device memory is owned by the context; host memory by the caller. Calls
 *    on one context are serialised by the caller (one context per host thread),
 *    as the reference backend's single orchestration thread does.
 */
#ifndef NIMBUS_HIP_H
#define NIMBUS_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum
{
    NIMBUS_OK = 0,
    NIMBUS_EINVAL = -1,    /* invalid argument or shape: nothing was touched */
    NIMBUS_ENODEV = -2,    /* no HIP device / device index out of range      */
    NIMBUS_ENOMEM = -3,    /* device allocation or admission budget refused  */
    NIMBUS_EHIP = -4,      /* a HIP runtime call failed                      */
    NIMBUS_ENOENT = -5     /* unknown member id                              */
};

enum
{
    NIMBUS_PRECISION_FP64 = 0,  /* FP64 storage, FP64 arithmetic                      */
    NIMBUS_PRECISION_MIXED = 1  /* FP32 storage, FP64 acoustic/EOS/base/Thomas terms  */
};

enum
{
    NIMBUS_ACOUSTIC_EXPLICIT = 0, /* reference forward-backward substep               */
    NIMBUS_ACOUSTIC_HEVI = 1      /* vertically implicit w-p' with FP64 Thomas solve  */
};

/* Lateral / vertical boundary codes (boundary_conditions_cartesian_cgrid.cpp:26-60). */
enum
{
    NIMBUS_BC_PERIODIC = 0,
    NIMBUS_BC_RADIATIVE = 1,
    NIMBUS_BC_ZERO_GRADIENT = 2,
    NIMBUS_BC_RIGID = 3,
    NIMBUS_BC_RIGID_SLIP = 4
};
enum { NIMBUS_BOTTOM_FREE_SLIP = 0, NIMBUS_BOTTOM_NO_SLIP = 1 };
enum { NIMBUS_TOP_RIGID = 0, NIMBUS_TOP_OPEN = 1 };
/* Smagorinsky modes and filter widths (turbulence/runtime.cpp:49-72). */
enum { NIMBUS_SGS_3D = 0, NIMBUS_SGS_HORIZONTAL = 1, NIMBUS_SGS_VERTICAL = 2 };
enum { NIMBUS_SGS_FILTER_CUBIC_ROOT = 0, NIMBUS_SGS_FILTER_DX = 1 };

typedef struct nimbus_hip_ctx nimbus_hip_ctx;

typedef struct
{
    int nx, ny, nz;       /* cells including lateral ghosts; nx, ny >= 5, 4 <= nz <= 1024, fewer than 2^28 cells (rows padded to 64 B) */
    double dx, dy, dz;    /* metres */
} nimbus_grid;

/* Large-step configuration (split_explicit.cpp:67-258; kessler.hpp:40-46). */
typedef struct
{
    double dt;            /* large step (s) */
    int substeps;         /* requested acoustic substeps; promoted to the CFL minimum */
    int acoustic_mode;    /* NIMBUS_ACOUSTIC_* */
    double hevi_beta;     /* off-centering of the implicit vertical step (HE-VI) */
    double coriolis_f;    /* f-plane parameter (1/s) */
    int bc_west, bc_east, bc_south, bc_north, bc_bottom, bc_top;
    double qc0, c_auto, c_accr, evap_eff, vt_max;
    /* Rayleigh sponges applied after the EOS projection
     * (src/core/infra/rayleigh_damping.cpp:403-470,985-1080,1160-1440): u, v
     * relax to the base wind, w to 0 and theta to theta0 with rate
     * alpha sin^2 ramps; rho and p are not relaxed. sponge = 0 disables all. */
    int sponge;
    double sponge_z_d;            /* upper sponge bottom (m); alpha rises to sponge_alpha at the lid */
    double sponge_alpha;          /* 1/s (0.05) */
    int sponge_w_cells;           /* w-only absorber depth under the lid (cells, 8) */
    double sponge_w_alpha;        /* 1/s (0.2) */
    int sponge_lateral_cells;     /* lateral band width (cells; 10 km in the presets) */
    double sponge_lateral_alpha;  /* 1/s (0.05) */
    /* Smagorinsky SGS after the microphysics and before the EOS projection
     * (turbulence/schemes/smagorinsky/smagorinsky.cpp:495-580,
     * turbulence/base/eddy_viscosity.cpp:1253-2090, core/orchestration/dynamics/
     * dynamics.cpp:7160-7262, sgs_apply.cpp:610-790): closure and conservative
     * momentum/theta/qv tendencies recomputed every SGS substep of sgs_dt (the
     * remainder last). sgs = 0 disables it. */
    int sgs;
    int sgs_mode;                 /* NIMBUS_SGS_3D, NIMBUS_SGS_HORIZONTAL ("2d" in the presets), NIMBUS_SGS_VERTICAL */
    int sgs_filter;               /* NIMBUS_SGS_FILTER_CUBIC_ROOT (default), NIMBUS_SGS_FILTER_DX */
    double sgs_cs;                /* 0.18 */
    double sgs_pr_t, sgs_sc_t;    /* 1/3, 1/3 */
    double sgs_nu_max, sgs_k_max; /* caps, 1000 m^2/s */
    double sgs_dt;                /* longest SGS substep (1 s) */
    /* Explicit Laplacian diffusion after the SGS and before the EOS
     * projection (core/orchestration/physics/diffusion_step.cpp:679-1030,
     * numerics/diffusion/schemes/explicit/explicit.cpp): the live-density
     * finite-volume tendencies of the state after the SGS, committed as
     * field + dt tendency with the storage clamps. diffusion: 0 off,
     * NIMBUS_DIFFUSE_SCALARS (theta, qv), NIMBUS_DIFFUSE_MOMENTUM (u, v, w),
     * or both OR-ed (the reference's apply_to "scalars", "momentum", "all");
     * lateral sides diffuse periodically when both faces of an axis are
     * periodic and are closed (zero flux) otherwise. */
    int diffusion;
    double diffusion_kh, diffusion_kv; /* m^2/s */
} nimbus_step_config;
enum { NIMBUS_DIFFUSE_SCALARS = 1, NIMBUS_DIFFUSE_MOMENTUM = 2 };

/* FP64 base-state profiles at cell centres, each nz long
 * (include/core/runtime/simulation_state.hpp:104-111). */
typedef struct
{
    const double *rho0, *p0, *theta0, *qv0, *u0, *v0;
} nimbus_base_state;

/* Host views of the nine prognostic fields, reference layout. The element type
 * is float for the *_f32 entry points and double for *_f64. rain_mm (nx*ny
 * doubles, accumulated surface rain) may be null. */
typedef struct
{
    void *u, *v, *w, *rho, *p, *theta, *qv, *qc, *qr;
    double* rain_mm;
} nimbus_host_state;

/* Analytic Weisman-Klemp-type sounding with a quarter-circle hodograph. */
typedef struct
{
    double theta_sfc;        /* K (300) */
    double theta_tropo;      /* K (343) */
    double z_tropo;          /* m (12000) */
    double T_tropo;          /* K (213) */
    double rh_sfc, rh_top;   /* 0.95, 0.25 */
    double qv_cap;           /* kg/kg (0.014) */
    double p_sfc;            /* Pa (100000) */
    double shear;            /* |V(6 km) - V(0)| m/s (30) */
    double shear_depth;      /* m (6000) */
    double curve_depth;      /* m (4000): depth of the quarter circle */
    double curvature;        /* 1 = quarter circle, 0 = straight line */
} nimbus_sounding;

typedef struct
{
    double xc, yc, zc;       /* centre (m); x, y measured from the west/south wall */
    double rh, rz;           /* horizontal / vertical radius (m) */
    double dtheta;           /* K */
} nimbus_bubble;

typedef struct
{
    nimbus_sounding sounding;
    int n_bubbles;
    nimbus_bubble bubbles[4];
    double noise_amp;        /* K, uniform in [-amp, amp) on physical cells */
    uint64_t seed;
    uint32_t member;
} nimbus_ic_params;

typedef struct
{
    nimbus_grid grid;
    nimbus_step_config config;
    int precision;           /* NIMBUS_PRECISION_* */
    nimbus_base_state base;  /* copied at creation */
} nimbus_member_desc;

typedef struct
{
    double max_w;            /* max w over the physical domain (m/s) */
    double min_w;
    double rms_theta_p;      /* RMS of theta - theta0 over physical cells (K) */
    double rain_total_mm;    /* sum of accumulated surface rain over columns (mm) */
    double max_rain_mm;
    double mean_qc, mean_qr; /* domain means over physical cells (kg/kg) */
    int64_t steps;           /* large steps completed */
} nimbus_diag;

/* ---------------- context ---------------- */
const char* nimbus_hip_last_error(void);
int nimbus_hip_create(int device, nimbus_hip_ctx** out);
int nimbus_hip_destroy(nimbus_hip_ctx* ctx);
int nimbus_hip_mem_info(nimbus_hip_ctx* ctx, uint64_t* free_bytes, uint64_t* total_bytes);
/* The device's architecture name (e.g. "gfx950:sramecc+:xnack-"), NUL terminated in buf[n]. */
int nimbus_hip_device_arch(nimbus_hip_ctx* ctx, char* buf, size_t n);
/* Kernel launches made by the context (ComputeBackend::successful_dispatch_count,
 * backend_interface.hpp:47). */
uint64_t nimbus_hip_dispatch_count(nimbus_hip_ctx* ctx);
/* Successful calls of one reference-contract entry point, named by the part
 * after "nimbus_hip_dispatch_" (e.g. "kessler_pointwise"); the per-operation
 * count behind ComputeBackend::successful_kernel_dispatch_count(KernelId)
 * (backend_interface.hpp:60). */
uint64_t nimbus_hip_successful_dispatch_count(nimbus_hip_ctx* ctx, const char* op);

/* ---------------- initial conditions (host only, no device needed) ---------------- */
void nimbus_default_sounding(nimbus_sounding* s);
/* Hydrostatic FP64 base state of the sounding at cell centres z_k = (k+1/2) dz. */
int nimbus_ic_base_state(const nimbus_grid* g, const nimbus_sounding* s, double* rho0, double* p0,
                         double* theta0, double* qv0, double* u0, double* v0);
/* Full initial state: base + bubbles (RH preserving) + seeded noise; p = p0, rho from the EOS. */
int nimbus_ic_fill_f32(const nimbus_grid* g, const nimbus_ic_params* prm, const nimbus_host_state* out);
int nimbus_ic_fill_f64(const nimbus_grid* g, const nimbus_ic_params* prm, const nimbus_host_state* out);
/* The same fill over caller-supplied base profiles (nz each; a CM1 sounding's
 * nimbus_cm1_base_state output, say): bubbles, noise and the RH-preserving
 * moistening as above, with prm->sounding used only for its qv_cap. */
int nimbus_ic_fill_base_f32(const nimbus_grid* g, const nimbus_base_state* base, const nimbus_ic_params* prm,
                            const nimbus_host_state* out);
int nimbus_ic_fill_base_f64(const nimbus_grid* g, const nimbus_base_state* base, const nimbus_ic_params* prm,
                            const nimbus_host_state* out);
/* Counter-based uniform [0,1) hash of (seed, member, i, j, k) used for the noise. */
double nimbus_ic_uniform(uint64_t seed, uint32_t member, int i, int j, int k);
/* Effective acoustic substep count (split_explicit.cpp:67-156; HE-VI checks only the horizontal CFL). */
int nimbus_resolve_substeps(const nimbus_grid* g, const nimbus_base_state* b, const nimbus_step_config* c);

/* ---------------- CM1 input_sounding (host only, no device needed) ----------------
 * CM1's native text sounding: a header "p_sfc[hPa] theta_sfc[K] qv_sfc[g/kg]",
 * then one row "z[m] theta[K] qv[g/kg] u[m/s] v[m/s]" per level, the first at
 * z = 0; '!' and '#' start comments, Fortran 'd'/'D' exponents are accepted
 * (src/soundings/schemes/cm1/cm1_sounding.hpp:5-15). In memory the column is
 * in SI units (Pa, K, kg/kg, m, m/s) with the pressure and temperature the
 * reader reconstructs hydrostatically from the header. Arrays are
 * caller-owned with `capacity` entries each. */
#define NIMBUS_CM1_MAX_LEVELS 100000           /* soundings/sounding_input_limits.hpp:10 */
#define NIMBUS_CM1_MAX_BYTES (8u * 1024u * 1024u) /* soundings/sounding_input_limits.hpp:9 */

typedef struct
{
    int n;                             /* levels after the header (output of parse/read) */
    int capacity;                      /* entries each array below holds */
    double p_sfc, theta_sfc, qv_sfc;   /* header: Pa, K, kg/kg */
    double *z, *theta, *qv, *u, *v;    /* the file's columns */
    double *p, *T;                     /* hydrostatic reconstruction (Pa, K); may be null for format */
} nimbus_cm1_sounding;

enum
{
    NIMBUS_INTERP_LINEAR = 0,          /* soundings_base.hpp:95 interpolation_method 0 */
    NIMBUS_INTERP_SPLINE = 1,          /* monotone PCHIP */
    NIMBUS_INTERP_LOG_PRESSURE = 2     /* linear, with log-linear pressure */
};

typedef struct
{
    int method;                        /* NIMBUS_INTERP_* (default linear) */
    int extrapolate_below, extrapolate_above; /* default 0, 0 */
} nimbus_interp;

/* Parse and validate CM1 text, reconstruct p and T (Exner integration of
 * theta_v by the trapezoid rule from the header pressure) and apply the
 * reader's quality control, refusing rather than shortening a profile
 * (Cm1SoundingScheme::load_sounding, cm1_sounding.cpp:159-281). Fails with
 * NIMBUS_EINVAL and a line-numbered message; when capacity is too small,
 * out->n holds the level count needed. */
int nimbus_cm1_parse(const char* text, size_t len, nimbus_cm1_sounding* out);
/* The same from a regular file of at most NIMBUS_CM1_MAX_BYTES
 * (Cm1SoundingScheme::initialize, cm1_sounding.cpp:138-157). */
int nimbus_cm1_read(const char* path, nimbus_cm1_sounding* out);
/* Write the column as CM1 text (header from p_sfc, theta_sfc, qv_sfc; rows
 * from z, theta, qv, u, v) with round-trip precision into buf (cap bytes, NUL
 * terminated); *len receives the length without the NUL, also when cap is too
 * small (then NIMBUS_EINVAL). No reference counterpart: the writer the
 * exchange of soundings with the reference needs. */
int nimbus_cm1_format(const nimbus_cm1_sounding* s, char* buf, size_t cap, size_t* len);
/* Model column at heights z[0..nz) (strictly increasing, >= 0; dz the grid
 * spacing): interpolation to z (Cm1SoundingScheme::interpolate_to_heights,
 * cm1_sounding.cpp:283-312; sounding_algorithms.cpp:427-641;
 * sharpy_interpolation.cpp:13-407), the coverage check of the model top
 * interface and the dry Exner re-integration that keeps theta exact
 * (FileSoundingSource::build and sounding_from_data,
 * init/sounding/sources/file_sounding.cpp:63-327,344-438). Outputs the base
 * profiles rho0, p0, theta0, qv0, u0, v0 and temperature T0 (nz each; T0 may
 * be null). */
int nimbus_cm1_column(const nimbus_cm1_sounding* s, const nimbus_interp* how, int nz, const double* z, double dz,
                      double* rho0, double* p0, double* theta0, double* qv0, double* u0, double* v0, double* T0);
/* nimbus_cm1_column at the cell centres z_k = (k + 1/2) dz of grid g: the
 * base state the reference assigns from a sounding
 * (core/orchestration/dynamics/equations.cpp:288-305). */
int nimbus_cm1_base_state(const nimbus_grid* g, const nimbus_cm1_sounding* s, const nimbus_interp* how,
                          double* rho0, double* p0, double* theta0, double* qv0, double* u0, double* v0);
/* The analytic sounding sampled at heights z[0..n) (z[0] = 0) as a CM1
 * column: theta, qv and the hodograph wind at each height, the header from
 * the surface values; p and T (when non-null) from the same hydrostatic
 * integration as nimbus_ic_base_state. No reference counterpart: it exports
 * the seeded workloads' sounding in the reference's input format. */
int nimbus_ic_sounding_levels(const nimbus_sounding* s, int n, const double* z, nimbus_cm1_sounding* out);

/* ---------------- state I/O (host only) ---------------- */
enum
{
    NIMBUS_NPY_F4 = 0,   /* '<f4' */
    NIMBUS_NPY_F8 = 1    /* '<f8' */
};
/* Write a C-order array of 1 to 3 dimensions as a NumPy v1.0 .npy file, byte
 * for byte what the reference's writer produces (npy::write_2d / write_3d,
 * src/core/output/npy_writer.cpp:33-182). */
int nimbus_npy_write(const char* path, const void* data, int dtype, int ndim, const int64_t* shape);

/* ---------------- resident ensemble members ---------------- */
int nimbus_hip_member_create(nimbus_hip_ctx* ctx, const nimbus_member_desc* desc, uint32_t* id);
int nimbus_hip_member_destroy(nimbus_hip_ctx* ctx, uint32_t id);
/* Bytes the member would hold on the device (admission planning). */
uint64_t nimbus_hip_member_bytes(const nimbus_member_desc* desc);
int nimbus_hip_member_upload(nimbus_hip_ctx* ctx, uint32_t id, const nimbus_host_state* s);
int nimbus_hip_member_download(nimbus_hip_ctx* ctx, uint32_t id, const nimbus_host_state* s);
int nimbus_hip_member_diag(nimbus_hip_ctx* ctx, uint32_t id, nimbus_diag* d);
/* Field export of one member: u, v, w, theta, p, rho, qv, qc, qr as <name>.npy
 * of shape (nx, ny, nz) '<f4' in the reference layout (full rho and p), and
 * accumulated_rainfall.npy of shape (nx, ny), into the existing directory
 * `dir` -- the files the reference's NPY field export writes for these
 * fields (AsyncOutputWriter::write_npy_field, src/core/output/output_writer.cpp:
 * 1331-1350; names src/core/output/output_config.cpp:216-253). FP64 members
 * are rounded to FP32, as the reference's fields are stored. */
int nimbus_hip_member_export_npy(nimbus_hip_ctx* ctx, uint32_t id, const char* dir);
/* Member snapshot for eviction and resume: the description (grid, step
 * configuration, precision, launch tiling), the six base profiles, the step
 * count, and the prognostic fields and surface rain in their device storage
 * form (perturbation rho' and p', storage precision), with a 64-bit
 * checksum (FNV-1a steps over 64-bit words in four lanes). Restoring it and stepping on gives the results of never having
 * stopped, bit for bit. The role of the reference's simulation checkpoint
 * (src/core/runtime/simulation_checkpoint.cpp, --resume / --checkpoint-every;
 * wire helpers include/core/runtime/checkpoint_format.hpp:22-60), not its file
 * format: that one records the whole run configuration. */
int nimbus_hip_member_snapshot_size(nimbus_hip_ctx* ctx, uint32_t id, uint64_t* bytes);
int nimbus_hip_member_snapshot(nimbus_hip_ctx* ctx, uint32_t id, void* buf, uint64_t cap);
int nimbus_hip_member_restore(nimbus_hip_ctx* ctx, const void* buf, uint64_t len, uint32_t* id);
/* The same through a file. */
int nimbus_hip_member_save(nimbus_hip_ctx* ctx, uint32_t id, const char* path);
int nimbus_hip_member_load(nimbus_hip_ctx* ctx, const char* path, uint32_t* id);
/* Advance each listed member by n_steps large steps; returns after the device finished.
 * The listed members of one precision run as one batch: every kernel launch of
 * the step covers all of them (uneven grids included), and the batch's whole
 * large step is one captured hipGraph when graphs are enabled. Members with
 * different substep or microphysics subcycle counts share the batch. */
int nimbus_hip_step(nimbus_hip_ctx* ctx, const uint32_t* ids, int n_ids, int n_steps);
/* Same, enqueued on the context stream without waiting. */
int nimbus_hip_step_async(nimbus_hip_ctx* ctx, const uint32_t* ids, int n_ids, int n_steps);
int nimbus_hip_synchronize(nimbus_hip_ctx* ctx);
/* Launch geometry of a member: column tiles of tile_w x (256 / tile_w) columns
 * (tile_w 32 or 64), column marches split into k_chunks vertical chunks
 * (1 .. nz), and XCD-contiguous workgroup ordering on/off. tile_w 0 (the
 * default): per-cell work in runs of 256 consecutive cells of one level's
 * row-major plane, per-column work in runs of 256 consecutive columns, the
 * wide-stencil passes (slow tendencies, WENO5 fluxes and update) in 64 x 4
 * tiles with k_chunks. */
int nimbus_hip_member_set_tiling(nimbus_hip_ctx* ctx, uint32_t id, int tile_w, int k_chunks, int xcd_remap);
/* HE-VI column-solve form of a member. It is fixed per member (never chosen
 * by the batch), so a member's results are the same alone and in any batch.
 * AUTO: partitioned for FP64 members of at most 129 levels and for FP32
 * members of fewer than 32768 columns and at most 129 levels, the sequential
 * sweep otherwise. PARTITIONED is refused above 129 levels, SEQUENTIAL for FP64
 * members of at most 129 levels (they hold no coefficient scratch). No
 * reference counterpart: the reference has no vertically implicit solve. */
enum
{
    NIMBUS_COLUMN_SOLVE_AUTO = 0,
    NIMBUS_COLUMN_SOLVE_SEQUENTIAL = 1,  /* one thread per column, FP64 coefficients in scratch */
    NIMBUS_COLUMN_SOLVE_PARTITIONED = 2  /* 16 runs of faces per column, reduced system in LDS */
};
int nimbus_hip_member_set_column_solve(nimbus_hip_ctx* ctx, uint32_t id, int form);
/* The member's form (SEQUENTIAL or PARTITIONED; AUTO for explicit-acoustic members). */
int nimbus_hip_member_column_solve(nimbus_hip_ctx* ctx, uint32_t id, int* form);
/* Autotuner: times `trials` large steps per launch-geometry candidate on a saved
 * copy of the member state, keeps the fastest and reports it. Every exit,
 * failures included, restores the member's state and step count. */
int nimbus_hip_member_autotune(nimbus_hip_ctx* ctx, uint32_t id, int trials, int* best_tile_w, int* best_k_chunks,
                               int* best_remap, double* best_ms);
/* Enable hipGraph capture of whole large steps (one graph per batch of members). */
int nimbus_hip_set_graphs(nimbus_hip_ctx* ctx, int enable);
/* Per-kernel timing with HIP events on the launch stream (0 disables). */
int nimbus_hip_set_kernel_timing(nimbus_hip_ctx* ctx, int enable);
/* Kernel timing table: name, launches, total milliseconds. Returns the number of rows. */
int nimbus_hip_kernel_timing(nimbus_hip_ctx* ctx, int max_rows, char (*names)[48], uint64_t* launches,
                             double* total_ms);
int nimbus_hip_reset_kernel_timing(nimbus_hip_ctx* ctx);
/* Cost of membership changes so far: batch member-table builds (device
 * allocation and upload of the table and block maps) and whole-step graph
 * captures (capture and instantiation), counts and host milliseconds. A
 * batch is built and captured the first time a member set steps together;
 * the 8 most recent sets stay cached. */
int nimbus_hip_batch_stats(nimbus_hip_ctx* ctx, uint64_t* builds, double* build_ms, uint64_t* captures,
                           double* capture_ms);
/* Elapsed device time (ms) of the last nimbus_hip_step call measured with HIP events. */
double nimbus_hip_last_step_ms(nimbus_hip_ctx* ctx);

/* ---------------- per-operation dispatch (reference dispatch contracts) ----------------
 * Each takes caller-owned host float arrays in the reference layout, runs the
 * same device kernels the resident step uses (FP32 storage, mixed arithmetic)
 * and writes the outputs back. */

/* Replaces CpuComputeBackend::dispatch_acoustic_substep_fused_cartesian_cgrid
 * (src/compute/cpu/acoustic_cgrid.cpp:50-76; struct include/compute/dispatch/dynamics.hpp:220-277). */
typedef struct
{
    float *u, *v, *w, *rho, *p;
    const float* theta_data;
    const double* rho0_base_host;
    const double* p0_base_host;
    int base_len;
    int nx, ny, nz;
    double dx_host, dy_host, dz_host, dt_small_host, gamma_host;
    float rho_min, rho_max, p_min, p_max, p_ref_fallback, wind_clamp_h, wind_clamp_v;
} nimbus_acoustic_substep_fused_dispatch;
int nimbus_hip_dispatch_acoustic_substep_fused(nimbus_hip_ctx* ctx, const nimbus_acoustic_substep_fused_dispatch* d);

/* Replaces DynamicsComputeKernels::dispatch_acoustic_substeps_batched_cartesian_cgrid
 * (include/compute/dispatch/dynamics.hpp:279-325,726-731): n_substeps substeps with
 * per-substep BCs and time-averaged mass-flux export (weight per substep). The
 * KW78 radiative stash is seeded from the input faces. */
typedef struct
{
    float *u, *v, *w, *rho, *p;
    const float* theta_data;
    const double* rho0_base_host;
    const double* p0_base_host;
    int nx, ny, nz;
    double dx, dy, dz, dt_small;
    int n_substeps;
    int acoustic_mode;
    double hevi_beta;
    int bc_west, bc_east, bc_south, bc_north, bc_bottom, bc_top;
    float *mass_flux_x, *mass_flux_y, *mass_flux_z;
    int accumulate_mass_flux;
    double mass_flux_weight;
} nimbus_acoustic_substeps_batched_dispatch;
int nimbus_hip_dispatch_acoustic_substeps_batched(nimbus_hip_ctx* ctx,
                                                  const nimbus_acoustic_substeps_batched_dispatch* d);

/* Replaces dispatch_cartesian_cgrid_slow_tendencies (include/compute/dispatch/dynamics.hpp:374-420,780):
 * UP5 momentum advection, buoyancy, moisture buoyancy, loading, f-plane Coriolis. */
typedef struct
{
    const float *u_data, *v_data, *w_data, *rho_data, *qv_data, *qc_data, *qr_data;
    const double *rho0_base, *qv0_base, *u0_base, *v0_base;
    float *du_dt_data, *dv_dt_data, *dw_dt_data;
    int nr, nth, nz;
    double dx, dy, dz, coriolis_f;
} nimbus_slow_tendencies_dispatch;
int nimbus_hip_dispatch_slow_tendencies(nimbus_hip_ctx* ctx, const nimbus_slow_tendencies_dispatch* d);

/* Heun slow RK2 + slow apply (dynamics.cpp:5123-5292,1955-2095) on u, v, w. */
typedef struct
{
    float *u, *v, *w;
    const float *rho, *qv, *qc, *qr;
    const double *rho0_base, *qv0_base, *u0_base, *v0_base;
    int nx, ny, nz;
    double dx, dy, dz, dt_large, coriolis_f;
} nimbus_slow_rk2_dispatch;
int nimbus_hip_dispatch_slow_rk2(nimbus_hip_ctx* ctx, const nimbus_slow_rk2_dispatch* d);

/* Replaces dispatch_advect_scalars_conservative_cartesian_cgrid
 * (include/compute/dispatch/advection.hpp:46-103,226-231): WENO5 retained-flux
 * transport of up to 4 scalars with the positive-definite donor limiter. */
typedef struct
{
    int n_scalars;
    float* q[4];              /* in/out */
    int positive_definite[4];
    const float *mass_flux_x, *mass_flux_y, *mass_flux_z;
    const float *rho_begin, *rho_end;
    int nx, ny, nz;
    double dx, dy, dz, dt;
    int bc_top;
} nimbus_advect_scalars_dispatch;
int nimbus_hip_dispatch_advect_scalars(nimbus_hip_ctx* ctx, const nimbus_advect_scalars_dispatch* d);

/* Replaces dispatch_kessler_pointwise_backend + dispatch_kessler_sedimentation_backend
 * (include/compute/dispatch/microphysics.hpp:436,726,752; kessler.cpp:103-531):
 * saturation adjustment, warm rain and sedimentation tendencies on the physical domain. */
typedef struct
{
    const float *p, *theta, *rho, *qv, *qc, *qr;
    float *dtheta_dt, *dqv_dt, *dqc_dt, *dqr_dt;
    int nx, ny, nz;
    double dz, dt;
    double qc0, c_auto, c_accr, evap_eff, vt_max;
} nimbus_kessler_dispatch;
int nimbus_hip_dispatch_kessler(nimbus_hip_ctx* ctx, const nimbus_kessler_dispatch* d);

/* ---------------- reference dispatch contracts as they are ----------------
 * The structs below carry the field lists of the reference's device dispatch
 * structs (include/compute/dispatch/{dynamics,microphysics,advection}.hpp)
 * with C types (bool -> int), so an adapter implementing the reference's
 * ComputeBackend families copies fields one for one. Arrays are caller-owned
 * host float arrays in the reference layout; outputs are written back only
 * when the whole call succeeded. Arithmetic: FP64 on FP32 storage (the
 * reference CPU backend's precision). */

/* Replaces DynamicsComputeKernels::dispatch_cartesian_cgrid_slow_tendencies
 * (include/compute/dispatch/dynamics.hpp:374-420,780): UP5 momentum advection
 * with 4-point advecting velocities, buoyancy -g (rho_f - rho0_f) / rho_f,
 * moisture buoyancy 0.608 g (qv_f - qv0_f) when has_moisture, hydrometeor
 * loading -g loading_f (host pre-summed) when has_loading, f-plane Coriolis on
 * the perturbation wind, dp/dt = -u.grad p (centred) and drho/dt = 0; every
 * output zero outside the reference's update ranges. Frame velocities
 * subtract from the horizontal advecting velocities. Refused: coriolis_beta
 * != 0 (beta plane). rho0/qv0/u0/v0 profiles are nz long (FP32 as the
 * reference stores them). */
typedef struct
{
    const float *u_data, *v_data, *w_data, *rho_data, *p_data;
    const float* rho0_base_data;
    int rho0_base_len;
    const float* qv0_base_data;
    int qv0_base_len;
    const float *u0_base_data, *v0_base_data;
    int wind0_base_len;
    const float *qv_data, *loading_data;
    float *du_dt_data, *dv_dt_data, *dw_dt_data, *drho_dt_data, *dp_dt_data;
    int nr, nth, nz;
    float dx, dy, dz, g, coriolis_f_val;
    int has_moisture, has_loading, slow_rk2_invariant_input_residency;
    float *du_advection_dt_data, *dv_advection_dt_data, *dw_advection_dt_data; /* optional, all or none */
    float frame_velocity_u_ms, frame_velocity_v_ms;
    float coriolis_beta;
} nimbus_cartesian_cgrid_slow_tendencies_dispatch;
int nimbus_hip_dispatch_cartesian_cgrid_slow_tendencies(nimbus_hip_ctx* ctx,
                                                         const nimbus_cartesian_cgrid_slow_tendencies_dispatch* d);

/* Observer called after the device finished and before the host arrays are
 * written, with host copies of the committed fields (not to be retained);
 * include/compute/dispatch/common.hpp:22-27. */
typedef void (*nimbus_four_field_observer)(void* context, const float* field0, const float* field1,
                                           const float* field2, const float* field3);

/* Replaces DynamicsComputeKernels::dispatch_slow_apply (dynamics.hpp:328-351,760;
 * the CPU callback dynamics.cpp:1955-2095): u += dt (du + du_pbl), v += dt (dv +
 * dv_pbl), w += dt dw, p += dt dp; with sanitize a non-finite sum counts as
 * zero. Null pbl arrays count as zero. */
typedef struct
{
    float *u, *v, *w, *p_data;
    const float *du_dt, *dv_dt, *dw_dt, *dp_dt;
    const float *du_dt_pbl, *dv_dt_pbl;
    int nx, ny, nz;
    float dt_large;
    nimbus_four_field_observer completed_observer;
    void* completed_observer_context;
    int sanitize;
} nimbus_slow_apply_dispatch;
int nimbus_hip_dispatch_slow_apply(nimbus_hip_ctx* ctx, const nimbus_slow_apply_dispatch* d);

/* Replaces DynamicsComputeKernels::dispatch_state_relaxation_cartesian
 * (dynamics.hpp:354-372,747): upper Rayleigh sponge, additive w-only lid
 * absorber and lateral band on u, v, w, theta, each x += (1 - exp(-alpha dt))
 * (x_ref - x). profiles: 7 rows of nz floats (upper alpha, w-only alpha,
 * u_ref, v_ref, theta0, lateral alpha by distance to the nearest lateral face,
 * upper alpha at w faces); lateral_width <= nz. */
typedef struct
{
    float *u, *v, *w, *theta_data;
    const float* profiles;
    int profiles_len;
    int nx, ny, nz;
    int lateral_width;
    int do_upper, do_w_only, do_lateral;
    float dt_step;
    nimbus_four_field_observer completed_observer;
    void* completed_observer_context;
} nimbus_state_relaxation_dispatch;
int nimbus_hip_dispatch_state_relaxation_cartesian(nimbus_hip_ctx* ctx, const nimbus_state_relaxation_dispatch* d);

/* Replaces MicrophysicsComputeKernels::dispatch_microphysics_apply
 * (include/compute/dispatch/microphysics.hpp:326-363,688): on columns
 * i in [i_begin, i_end), j in [j_begin, j_end): theta += clamp((dtheta + rad +
 * pbl) dt, +-max_theta_step), bounded to [theta_min, theta_max] (fallback for
 * non-finite values); species += tendency dt clamped to [0, qv_max_val] /
 * [0, hydro_max_val] (0 when non-finite). Optional stats_out (stats_len >= 10,
 * and 10 + nz + nx for the histograms): counts of non-finite tendencies,
 * non-finite totals, limited steps, clamps, and the first clamp's record. The
 * ice and graupel/hail species and the rad/pbl terms may be null. */
typedef struct
{
    float *theta_data, *qv_data, *qc_data, *qr_data, *qi_data, *qs_data, *qg_data, *qh_data;
    const float *dtheta, *dqv, *dqc, *dqr, *dqi, *dqs, *dqg, *dqh;
    const float *dtheta_rad, *dtheta_pbl, *dqv_pbl;
    int nx, ny, nz;
    int i_begin, i_end, j_begin, j_end;
    float dt_step, theta_fallback, theta_min, theta_max, max_theta_step, qv_max_val, hydro_max_val;
    unsigned int* stats_out;
    int stats_len;
} nimbus_microphysics_apply_dispatch;
int nimbus_hip_dispatch_microphysics_apply(nimbus_hip_ctx* ctx, const nimbus_microphysics_apply_dispatch* d);

/* Replaces MicrophysicsComputeKernels::dispatch_kessler_pointwise
 * (microphysics.hpp:436-467,726; caller kessler.cpp:213-230): warm rain on the
 * host-adjusted temperature, qv and qc (the caller books the saturation
 * adjustment itself): autoconversion + accretion, ventilated evaporation
 * bounded by the saturation deficit; dtheta_dt receives the temperature
 * tendency (the caller converts it to theta). Writes every output: zero
 * outside the index ranges and for the qg/qh slots (which may be null). */
typedef struct
{
    const float *temperature_data, *p_data, *qv_data, *qc_data, *qr_data, *rho_d_data, *qh_data;
    float *dtheta_dt_data, *dqv_dt_data, *dqc_dt_data, *dqr_dt_data, *dqg_dt_data, *dqh_dt_data;
    int nr, nth, nz;
    int i_begin, i_end, j_begin, j_end, k_begin, k_end;
    float qc0, c_auto, c_accr, evaporation_efficiency, Lv_cp, T0, dt_step;
} nimbus_kessler_pointwise_dispatch;
int nimbus_hip_dispatch_kessler_pointwise(nimbus_hip_ctx* ctx, const nimbus_kessler_pointwise_dispatch* d);

/* Replaces MicrophysicsComputeKernels::dispatch_kessler_sedimentation
 * (microphysics.hpp:469-492,752; kessler.cpp:465-531): adds the rain
 * sedimentation divergence to dqr_dt, column by column in k order: the donor
 * loses Vt qr / dz_k, the level below gains (rho_k / rho_{k-1}) Vt qr / dz_{k-1},
 * Vt = min(36.34 (1e-3 rho qr)^0.1346 sqrt(1.2 / rho), Vt_max_rain, vt_cap).
 * layer_thickness_m (optional) replaces dz_val per cell; rho_d_data may be
 * null (rho = 1.2). The qg/qh slots are inert. */
typedef struct
{
    const float *qr_data, *qg_data, *qh_data, *rho_d_data;
    float *dqr_dt_data, *dqg_dt_data, *dqh_dt_data;
    int nr, nth, nz;
    int i_begin, i_end, j_begin, j_end, k_begin, k_end;
    float dz_val, Vt_max_rain, vt_cap, dt_step;
    const float* layer_thickness_m;
} nimbus_kessler_sedimentation_dispatch;
int nimbus_hip_dispatch_kessler_sedimentation(nimbus_hip_ctx* ctx, const nimbus_kessler_sedimentation_dispatch* d);

/* Replaces AdvectionComputeKernels::dispatch_advect_scalars_conservative_cartesian_cgrid
 * (include/compute/dispatch/advection.hpp:46-103,226-231;
 * advection_cartesian.cpp:1652-2749): WENO5 flux-form transport of n_scalars
 * (<= 16) with the positive-definite donor limiter. With
 * use_supplied_mass_flux the retained time-averaged fluxes and density_begin
 * are used and rho_data is the end density; without, the fluxes are formed
 * from u, v, w (frame-relative) and rho_data, which is then the begin
 * density, and the end density carries the same fluxes. q_out[s] may equal
 * q_in[s]. Refused: weno == 0 (the TVD reconstruction is not built) and
 * nonzero dt_kappas (chained point diffusion, as the reference's device
 * backends refuse it). The scheme's own fields inv_dx2.. are unused. */
typedef struct
{
    const float *u, *v, *w, *rho_data;
    const float* const* q_in;
    float* const* q_out;
    const float* dt_kappas;
    const int* positive_definite;
    int n_scalars;
    int nx, ny, nz;
    float inv_dx, inv_dy, inv_dz, inv_dx2, inv_dy2, inv_dz2;
    float dt_step, rho_floor;
    const float *density_begin, *mass_flux_x, *mass_flux_y, *mass_flux_z;
    int use_supplied_mass_flux;
    float frame_velocity_u_ms, frame_velocity_v_ms;
    int weno, preserve_positive_support;
} nimbus_conservative_cartesian_cgrid_transport_dispatch;
int nimbus_hip_dispatch_advect_scalars_conservative_cartesian_cgrid(
    nimbus_hip_ctx* ctx, const nimbus_conservative_cartesian_cgrid_transport_dispatch* d);

/* Control-volume kind of a diffused field (DiffusionFieldLocation,
 * include/compute/dispatch/diffusion.hpp:22-28). */
enum
{
    NIMBUS_DIFFUSION_SCALAR = 0,
    NIMBUS_DIFFUSION_U_FACE = 1,
    NIMBUS_DIFFUSION_V_FACE = 2,
    NIMBUS_DIFFUSION_W_FACE = 3
};

/* Replaces DiffusionComputeKernels::dispatch_diffusion_batch
 * (include/compute/dispatch/diffusion.hpp:52-69,149-209; its single-field
 * dispatch_diffusion, :29-50, is a batch of one; equations of
 * ExplicitDiffusionScheme, src/numerics/diffusion/schemes/explicit/
 * explicit.cpp:119-339,1183-1433): for each of `count` fields the live-density
 * finite-volume Laplacian tendency d(field)/dt = (1/rho_cv) div(rho K
 * grad(field)) on the field's control volumes (scalar cells or C-grid u, v,
 * w faces; A-grid cells when c_grid is 0), K_h horizontally, K_v (or the
 * field's vertical_K array) vertically with layer thicknesses dz[nz]; zero
 * flux at closed control-volume edges, wrap on periodic sides. Every
 * tendency cell is written (zero outside the control volumes). The vertical
 * part is rounded to FP32 before the horizontal part is added, as the
 * reference adds its two passes; FP64 arithmetic. Refused (NIMBUS_EINVAL,
 * tendencies untouched): nx or ny < 3, nz < 2, count outside 1..16, a
 * location outside 0..3, non-positive or non-finite dx, dy or dz, negative or
 * non-finite K, a non-finite field value or a non-positive density inside a
 * control-volume domain, a negative or non-finite vertical_K there. */
typedef struct
{
    const float* const* fields;
    float* const* tendencies;
    const int* locations;
    int count;
    const float* rho;
    const float* dz;
    int nx, ny, nz;
    float dx, dy, K_h, K_v;
    int c_grid, periodic_x, periodic_y;
    const float* const* vertical_K; /* null, or count arrays (each may be null) */
} nimbus_diffusion_batch_dispatch;
int nimbus_hip_dispatch_diffusion_batch(nimbus_hip_ctx* ctx, const nimbus_diffusion_batch_dispatch* d);

/* Allocation-free variant of nimbus_hip_dispatch_advect_scalars_conservative_cartesian_cgrid.
 * No counterpart in the reference: it is for callers that cannot allocate
 * device memory (code under HIP graph capture, tight serving loops, harnesses
 * that forbid hipMalloc). Every array in *d (winds, densities, mass fluxes,
 * q_in[s], q_out[s]) is a DEVICE pointer owned by the caller; the pointer
 * arrays q_in/q_out, dt_kappas and positive_definite stay in host memory.
 * The face fluxes, limiter ratios and (standalone mode) the formed mass
 * fluxes live in one static device array per GPU of 8 floats per cell for at
 * most NIMBUS_NOALLOC_MAX_CELLS cells (32 MiB), held until the process ends;
 * larger grids are refused with NIMBUS_EINVAL. The call makes no hipMalloc,
 * hipFree or other allocation, runs on the context stream and returns when
 * the device finished. Calls that overlap on one GPU share that array and
 * race; successive calls need no clearing (every scratch element is written
 * before it is read). */
#define NIMBUS_NOALLOC_MAX_CELLS (1u << 20)
int nimbus_hip_advect_scalars_conservative_noalloc(nimbus_hip_ctx* ctx,
                                                   const nimbus_conservative_cartesian_cgrid_transport_dispatch* d);

#ifdef __cplusplus
}
#endif

#endif /* NIMBUS_HIP_H */
