// Launch interface of the large-step kernels. Every launcher enqueues on the
// stream held by the Launcher, which optionally brackets each launch with HIP
// events for per-kernel timing.
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

// Device field set of one member. Precision-erased pointers; the kernels are
// instantiated for float and double storage.
struct DevFields
{
    void *u, *v, *w, *rp, *pp, *th, *qv, *qc, *qr;       // prognostic state (rho, p as perturbations)
    void *mfx, *mfy, *mfz, *rb;                          // retained mass fluxes, rho' at step begin
    void *k1u, *k1v, *k1w, *su, *sv, *sw;                // Heun stage-1 tendencies and predictor
    void* ralt;                                          // rho' ping-pong of the acoustic substep
    void *tth, *tqv, *tqc, *tqr;                         // transported scalars before microphysics
    void* flux[12];                                      // per scalar: x, y, z face fluxes
    float *rl_qv, *rl_qc, *rl_qr;                        // PD limiter factors
    double *tri_c, *tri_r;                               // HE-VI Thomas coefficients (FP64)
    float *rad_w, *rad_e, *rad_s, *rad_n;                // KW78 radiative stashes
    double* rain;                                        // accumulated surface rain [ny][nx]
    double* base;                                        // kBaseCount * nz FP64 profiles
};

struct StepParams
{
    double dt;
    int substeps;
    int acoustic_mode;
    double hevi_beta;
    double coriolis_f;
    int bc_west, bc_east, bc_south, bc_north, bc_bottom, bc_top;
    KesslerConst kessler;
    Bounds bounds;
};

template <class S>
void launch_large_step(Launcher& L, const DevGrid& g, const DevFields& f, const StepParams& p);

// Individual phases (used by the per-operation dispatch entry points).
template <class S>
void launch_slow_rk2(Launcher& L, const DevGrid& g, const DevFields& f, const StepParams& p);
template <class S>
void launch_slow_tendencies(Launcher& L, const DevGrid& g, const DevFields& f, const StepParams& p, void* du,
                            void* dv, void* dw);
template <class S>
void launch_acoustic_substep(Launcher& L, const DevGrid& g, const DevFields& f, const StepParams& p, double dts,
                             double weight, bool first, bool accumulate);
template <class S>
void launch_boundaries(Launcher& L, const DevGrid& g, const DevFields& f, const StepParams& p, double dt_rad);
template <class S>
void launch_seed_radiation(Launcher& L, const DevGrid& g, const DevFields& f);
template <class S>
void launch_transport(Launcher& L, const DevGrid& g, const DevFields& f, const StepParams& p, int n_scalars,
                      void* const* q_in, void* const* q_out, const int* pd);
template <class S>
void launch_kessler_tendencies(Launcher& L, const DevGrid& g, const DevFields& f, const StepParams& p, double dt,
                               void* dth, void* dqv, void* dqc, void* dqr);
template <class S>
void launch_diagnostics(Launcher& L, const DevGrid& g, const DevFields& f, double* partials, int n_partials,
                        double* out);
template <class H, class S>
void launch_host_to_dev(Launcher& L, const DevGrid& g, const void* src, void* dst, const double* base_slot);
template <class S, class H>
void launch_dev_to_host(Launcher& L, const DevGrid& g, const void* src, void* dst, const double* base_slot);

constexpr int kDiagValues = 8;
constexpr int kDiagBlocks = 512;

}  // namespace nimbus
