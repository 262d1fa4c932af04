// HipComputeBackend: the reference's ComputeBackend plugin contract
// (include/compute/backend_interface.hpp:19-255) implemented over the C ABI of
// libnimbus_hip.so (include/nimbus_hip.h). This file is compiled inside the
// reference tree, next to its CPU/Vulkan/Metal backends (see the integration
// guide for the factory and enum lines); it includes the reference's headers
// and nimbus_hip.h only, and links against libnimbus_hip.so.
//
// Every dispatch copies the reference struct's fields into the C ABI struct
// of the same contract and calls it; a refused call (nonzero status) returns
// false with the host fields untouched, so the caller's fallback policy
// (compute_backend.cpp:151-186) applies unchanged. Boundary-condition codes
// are translated from the reference's LateralBC/BottomBC/TopBC encodings
// (include/init/config/config_v1.hpp:40-67). Exceptions never leave the
// backend: the C ABI reports errors by status, and allocation failures in the
// adapter itself are caught and turned into a refusal.
#include "compute/backend_interface.hpp"
#include "compute/backend_types.hpp"
#include "compute/dispatch/advection.hpp"
#include "compute/dispatch/diffusion.hpp"
#include "compute/dispatch/dynamics.hpp"
#include "compute/dispatch/microphysics.hpp"
#include "compute/kernel_manifest.hpp"

#include "nimbus_hip.h"

#include <memory>
#include <string>
#include <vector>

namespace nimbus_hip_adapter
{

int lateral_code(int reference_code)
{
    switch (reference_code)  // LateralBC: Periodic 0, Radiative 1, Rigid 2, RigidSlip 3, ZeroGradient 4, LegacyOpen 5
    {
    case 0: return NIMBUS_BC_PERIODIC;
    case 1: return NIMBUS_BC_RADIATIVE;
    case 2: return NIMBUS_BC_RIGID;
    case 3: return NIMBUS_BC_RIGID_SLIP;
    case 4: return NIMBUS_BC_ZERO_GRADIENT;
    default: return -1;  // LegacyOpen and unknown codes: refused
    }
}

int bottom_code(int reference_code)  // BottomBC: RigidNoSlip 0, RigidSlip 1
{
    return reference_code == 0 ? NIMBUS_BOTTOM_NO_SLIP : reference_code == 1 ? NIMBUS_BOTTOM_FREE_SLIP : -1;
}

int top_code(int reference_code)  // TopBC: Rigid 0, RigidWithSponge 1, Open 2
{
    return reference_code == 0 || reference_code == 1 ? NIMBUS_TOP_RIGID : reference_code == 2 ? NIMBUS_TOP_OPEN : -1;
}

// FP64 copy of an FP32 base profile when the dispatch carries no host view.
const double* profile64(const double* host, const float* data, int n, std::vector<double>& keep)
{
    if (host) return host;
    if (!data || n <= 0) return nullptr;
    keep.assign(data, data + n);
    return keep.data();
}

class Dynamics final : public DynamicsComputeKernels
{
public:
    explicit Dynamics(nimbus_hip_ctx*& ctx) : ctx_(ctx) {}

    bool supports(DynamicsKernelOp op) const override
    {
        return ctx_ && (op == DynamicsKernelOp::AcousticSubstepFusedCartesianCgrid ||
                        op == DynamicsKernelOp::AcousticSubstepsBatchedCartesianCgrid ||
                        op == DynamicsKernelOp::StateRelaxationCartesian || op == DynamicsKernelOp::SlowApply ||
                        op == DynamicsKernelOp::CartesianCgridSlowTendencies);
    }
    bool supports_batched_cartesian_cgrid_mass_flux_export() const override { return ctx_ != nullptr; }

    bool dispatch_acoustic_substep_fused_cartesian_cgrid(const AcousticSubstepFusedCartesianCgridDispatch& d) override
    {
        if (!ctx_ || d.frame_velocity_u_ms != 0.0f || d.frame_velocity_v_ms != 0.0f || d.energy_report_host ||
            d.dp0_dz_low_data)
            return false;
        try
        {
            std::vector<double> r0, p0;
            nimbus_acoustic_substep_fused_dispatch c{};
            c.u = d.u; c.v = d.v; c.w = d.w; c.rho = d.rho; c.p = d.p;
            c.theta_data = d.theta_data;
            c.rho0_base_host = profile64(d.rho0_base_host, d.rho0_base_data, d.base_len, r0);
            c.p0_base_host = profile64(d.p0_base_host, d.p0_base_data, d.base_len, p0);
            c.base_len = d.base_len;
            c.nx = d.nx; c.ny = d.ny; c.nz = d.nz;
            c.dx_host = d.dx_host > 0.0 ? d.dx_host : d.dx;
            c.dy_host = d.dy_host > 0.0 ? d.dy_host : d.dy;
            c.dz_host = d.dz_host > 0.0 ? d.dz_host : d.dz;
            c.dt_small_host = d.dt_small_host > 0.0 ? d.dt_small_host : d.dt_small;
            c.gamma_host = d.gamma_host > 0.0 ? d.gamma_host : d.gamma_val;
            c.rho_min = d.rho_min; c.rho_max = d.rho_max; c.p_min = d.p_min; c.p_max = d.p_max;
            c.p_ref_fallback = d.p_ref_fallback; c.wind_clamp_h = d.wind_clamp_h; c.wind_clamp_v = d.wind_clamp_v;
            return nimbus_hip_dispatch_acoustic_substep_fused(ctx_, &c) == NIMBUS_OK;
        }
        catch (...)
        {
            return false;
        }
    }

    bool dispatch_acoustic_substeps_batched_cartesian_cgrid(
        const AcousticSubstepsBatchedCartesianCgridDispatch& d) override
    {
        const int w = lateral_code(d.bc_west), e = lateral_code(d.bc_east), s = lateral_code(d.bc_south),
                  n = lateral_code(d.bc_north), b = bottom_code(d.bc_bottom), t = top_code(d.bc_top);
        if (!ctx_ || w < 0 || e < 0 || s < 0 || n < 0 || b < 0 || t < 0 || d.frame_velocity_u_ms != 0.0f ||
            d.frame_velocity_v_ms != 0.0f || d.dp0_dz_low_data || d.wave_history || d.wave_history_output)
            return false;
        try
        {
            std::vector<double> r0, p0;
            nimbus_acoustic_substeps_batched_dispatch c{};
            c.u = d.u; c.v = d.v; c.w = d.w; c.rho = d.rho; c.p = d.p;
            c.theta_data = d.theta_data;
            c.rho0_base_host = profile64(nullptr, d.rho0_base_data, d.base_len, r0);
            c.p0_base_host = profile64(nullptr, d.p0_base_data, d.base_len, p0);
            c.nx = d.nx; c.ny = d.ny; c.nz = d.nz;
            c.dx = d.dx; c.dy = d.dy; c.dz = d.dz; c.dt_small = d.dt_small;
            c.n_substeps = d.n_substeps;
            c.acoustic_mode = NIMBUS_ACOUSTIC_EXPLICIT;
            c.hevi_beta = 0.0;
            c.bc_west = w; c.bc_east = e; c.bc_south = s; c.bc_north = n; c.bc_bottom = b; c.bc_top = t;
            c.mass_flux_x = d.mass_flux_x; c.mass_flux_y = d.mass_flux_y; c.mass_flux_z = d.mass_flux_z;
            c.accumulate_mass_flux = d.accumulate_mass_flux ? 1 : 0;
            c.mass_flux_weight = d.mass_flux_weight;
            return nimbus_hip_dispatch_acoustic_substeps_batched(ctx_, &c) == NIMBUS_OK;
        }
        catch (...)
        {
            return false;
        }
    }

    bool dispatch_state_relaxation_cartesian(const StateRelaxationDispatch& d) override
    {
        if (!ctx_) return false;
        nimbus_state_relaxation_dispatch c{};
        c.u = d.u; c.v = d.v; c.w = d.w; c.theta_data = d.theta_data;
        c.profiles = d.profiles; c.profiles_len = d.profiles_len;
        c.nx = d.nx; c.ny = d.ny; c.nz = d.nz; c.lateral_width = d.lateral_width;
        c.do_upper = d.do_upper; c.do_w_only = d.do_w_only; c.do_lateral = d.do_lateral;
        c.dt_step = d.dt_step;
        c.completed_observer = d.completed_observer;
        c.completed_observer_context = d.completed_observer_context;
        return nimbus_hip_dispatch_state_relaxation_cartesian(ctx_, &c) == NIMBUS_OK;
    }

    bool dispatch_slow_apply(const SlowApplyDispatch& d) override
    {
        if (!ctx_) return false;
        nimbus_slow_apply_dispatch c{};
        c.u = d.u; c.v = d.v; c.w = d.w; c.p_data = d.p_data;
        c.du_dt = d.du_dt; c.dv_dt = d.dv_dt; c.dw_dt = d.dw_dt; c.dp_dt = d.dp_dt;
        c.du_dt_pbl = d.du_dt_pbl; c.dv_dt_pbl = d.dv_dt_pbl;
        c.nx = d.nx; c.ny = d.ny; c.nz = d.nz; c.dt_large = d.dt_large;
        c.completed_observer = d.completed_observer;
        c.completed_observer_context = d.completed_observer_context;
        c.sanitize = d.sanitize ? 1 : 0;
        return nimbus_hip_dispatch_slow_apply(ctx_, &c) == NIMBUS_OK;
    }

    bool dispatch_cartesian_cgrid_slow_tendencies(const CartesianCgridSlowTendenciesDispatch& d) override
    {
        if (!ctx_) return false;
        nimbus_cartesian_cgrid_slow_tendencies_dispatch c{};
        c.u_data = d.u_data; c.v_data = d.v_data; c.w_data = d.w_data; c.rho_data = d.rho_data; c.p_data = d.p_data;
        c.rho0_base_data = d.rho0_base_data; c.rho0_base_len = d.rho0_base_len;
        c.qv0_base_data = d.qv0_base_data; c.qv0_base_len = d.qv0_base_len;
        c.u0_base_data = d.u0_base_data; c.v0_base_data = d.v0_base_data; c.wind0_base_len = d.wind0_base_len;
        c.qv_data = d.qv_data; c.loading_data = d.loading_data;
        c.du_dt_data = d.du_dt_data; c.dv_dt_data = d.dv_dt_data; c.dw_dt_data = d.dw_dt_data;
        c.drho_dt_data = d.drho_dt_data; c.dp_dt_data = d.dp_dt_data;
        c.nr = d.nr; c.nth = d.nth; c.nz = d.nz;
        c.dx = d.dx; c.dy = d.dy; c.dz = d.dz; c.g = d.g; c.coriolis_f_val = d.coriolis_f_val;
        c.has_moisture = d.has_moisture; c.has_loading = d.has_loading;
        c.slow_rk2_invariant_input_residency = d.slow_rk2_invariant_input_residency;
        c.du_advection_dt_data = d.du_advection_dt_data;
        c.dv_advection_dt_data = d.dv_advection_dt_data;
        c.dw_advection_dt_data = d.dw_advection_dt_data;
        c.frame_velocity_u_ms = d.frame_velocity_u_ms; c.frame_velocity_v_ms = d.frame_velocity_v_ms;
        c.coriolis_beta = d.coriolis_beta;
        return nimbus_hip_dispatch_cartesian_cgrid_slow_tendencies(ctx_, &c) == NIMBUS_OK;
    }

private:
    nimbus_hip_ctx*& ctx_;
};

class Advection final : public AdvectionComputeKernels
{
public:
    explicit Advection(nimbus_hip_ctx*& ctx) : ctx_(ctx) {}

    bool supports(AdvectionKernelOp op) const override
    {
        return ctx_ && op == AdvectionKernelOp::ConservativeCartesianCgridTransport;
    }
    bool supports_supplied_cartesian_cgrid_mass_flux() const override { return ctx_ != nullptr; }

    bool dispatch_advect_scalars_conservative_cartesian_cgrid(
        const ConservativeCartesianCgridTransportDispatch& d) override
    {
        // operator-time science reports and interface registers are the CPU
        // family's; a device dispatch asked for them refuses
        if (!ctx_ || d.host_boundary_reports || d.host_energy_reports || d.host_interface_registers) return false;
        nimbus_conservative_cartesian_cgrid_transport_dispatch c{};
        c.u = d.u; c.v = d.v; c.w = d.w; c.rho_data = d.rho_data;
        c.q_in = d.q_in; c.q_out = d.q_out; c.dt_kappas = d.dt_kappas; c.positive_definite = d.positive_definite;
        c.n_scalars = d.n_scalars;
        c.nx = d.nx; c.ny = d.ny; c.nz = d.nz;
        c.inv_dx = d.inv_dx; c.inv_dy = d.inv_dy; c.inv_dz = d.inv_dz;
        c.inv_dx2 = d.inv_dx2; c.inv_dy2 = d.inv_dy2; c.inv_dz2 = d.inv_dz2;
        c.dt_step = d.dt_step; c.rho_floor = d.rho_floor;
        c.density_begin = d.density_begin;
        c.mass_flux_x = d.mass_flux_x; c.mass_flux_y = d.mass_flux_y; c.mass_flux_z = d.mass_flux_z;
        c.use_supplied_mass_flux = d.use_supplied_mass_flux ? 1 : 0;
        c.frame_velocity_u_ms = d.frame_velocity_u_ms; c.frame_velocity_v_ms = d.frame_velocity_v_ms;
        c.weno = d.weno ? 1 : 0;
        c.preserve_positive_support = d.preserve_positive_support ? 1 : 0;
        return nimbus_hip_dispatch_advect_scalars_conservative_cartesian_cgrid(ctx_, &c) == NIMBUS_OK;
    }

private:
    nimbus_hip_ctx*& ctx_;
};

class Microphysics final : public MicrophysicsComputeKernels
{
public:
    explicit Microphysics(nimbus_hip_ctx*& ctx) : ctx_(ctx) {}

    bool supports(MicrophysicsKernelOp op) const override
    {
        return ctx_ && (op == MicrophysicsKernelOp::KesslerPointwise ||
                        op == MicrophysicsKernelOp::KesslerSedimentation ||
                        op == MicrophysicsKernelOp::MicrophysicsApply);
    }

    bool dispatch_microphysics_apply(const MicrophysicsApplyDispatch& d) override
    {
        if (!ctx_) return false;
        nimbus_microphysics_apply_dispatch c{};
        c.theta_data = d.theta_data; c.qv_data = d.qv_data; c.qc_data = d.qc_data; c.qr_data = d.qr_data;
        c.qi_data = d.qi_data; c.qs_data = d.qs_data; c.qg_data = d.qg_data; c.qh_data = d.qh_data;
        c.dtheta = d.dtheta; c.dqv = d.dqv; c.dqc = d.dqc; c.dqr = d.dqr;
        c.dqi = d.dqi; c.dqs = d.dqs; c.dqg = d.dqg; c.dqh = d.dqh;
        c.dtheta_rad = d.dtheta_rad; c.dtheta_pbl = d.dtheta_pbl; c.dqv_pbl = d.dqv_pbl;
        c.nx = d.nx; c.ny = d.ny; c.nz = d.nz;
        c.i_begin = d.i_begin; c.i_end = d.i_end; c.j_begin = d.j_begin; c.j_end = d.j_end;
        c.dt_step = d.dt_step; c.theta_fallback = d.theta_fallback; c.theta_min = d.theta_min;
        c.theta_max = d.theta_max; c.max_theta_step = d.max_theta_step; c.qv_max_val = d.qv_max_val;
        c.hydro_max_val = d.hydro_max_val;
        c.stats_out = d.stats_out; c.stats_len = d.stats_len;
        return nimbus_hip_dispatch_microphysics_apply(ctx_, &c) == NIMBUS_OK;
    }

    bool dispatch_kessler_pointwise(const KesslerPointwiseDispatch& d) override
    {
        if (!ctx_) return false;
        nimbus_kessler_pointwise_dispatch c{};
        c.temperature_data = d.temperature_data; c.p_data = d.p_data; c.qv_data = d.qv_data;
        c.qc_data = d.qc_data; c.qr_data = d.qr_data; c.rho_d_data = d.rho_d_data; c.qh_data = d.qh_data;
        c.dtheta_dt_data = d.dtheta_dt_data; c.dqv_dt_data = d.dqv_dt_data; c.dqc_dt_data = d.dqc_dt_data;
        c.dqr_dt_data = d.dqr_dt_data; c.dqg_dt_data = d.dqg_dt_data; c.dqh_dt_data = d.dqh_dt_data;
        c.nr = d.nr; c.nth = d.nth; c.nz = d.nz;
        c.i_begin = d.i_begin; c.i_end = d.i_end; c.j_begin = d.j_begin; c.j_end = d.j_end;
        c.k_begin = d.k_begin; c.k_end = d.k_end;
        c.qc0 = d.qc0; c.c_auto = d.c_auto; c.c_accr = d.c_accr;
        c.evaporation_efficiency = d.evaporation_efficiency; c.Lv_cp = d.Lv_cp; c.T0 = d.T0;
        c.dt_step = d.dt_step;
        return nimbus_hip_dispatch_kessler_pointwise(ctx_, &c) == NIMBUS_OK;
    }

    bool dispatch_kessler_sedimentation(const KesslerSedimentationDispatch& d) override
    {
        if (!ctx_) return false;
        nimbus_kessler_sedimentation_dispatch c{};
        c.qr_data = d.qr_data; c.qg_data = d.qg_data; c.qh_data = d.qh_data; c.rho_d_data = d.rho_d_data;
        c.dqr_dt_data = d.dqr_dt_data; c.dqg_dt_data = d.dqg_dt_data; c.dqh_dt_data = d.dqh_dt_data;
        c.nr = d.nr; c.nth = d.nth; c.nz = d.nz;
        c.i_begin = d.i_begin; c.i_end = d.i_end; c.j_begin = d.j_begin; c.j_end = d.j_end;
        c.k_begin = d.k_begin; c.k_end = d.k_end;
        c.dz_val = d.dz_val; c.Vt_max_rain = d.Vt_max_rain; c.vt_cap = d.vt_cap; c.dt_step = d.dt_step;
        c.layer_thickness_m = d.layer_thickness_m;
        return nimbus_hip_dispatch_kessler_sedimentation(ctx_, &c) == NIMBUS_OK;
    }

private:
    nimbus_hip_ctx*& ctx_;
};

class Diffusion final : public DiffusionComputeKernels
{
public:
    explicit Diffusion(nimbus_hip_ctx*& ctx) : ctx_(ctx) {}

    bool supports(DiffusionKernelOp op) const override { return ctx_ && op == DiffusionKernelOp::Diffusion; }

    bool dispatch_diffusion(const DiffusionDispatch& d) override
    {
        const float* const fields[1] = {d.field};
        float* const tendencies[1] = {d.tendency};
        const DiffusionFieldLocation locations[1] = {d.location};
        const float* const vertical[1] = {d.vertical_K};
        const DiffusionBatchDispatch b{fields, tendencies, locations, 1, d.rho, d.dz, d.nx, d.ny, d.nz, d.dx, d.dy,
                                       d.K_h, d.K_v, d.c_grid, d.periodic_x, d.periodic_y,
                                       d.vertical_K ? vertical : nullptr};
        return dispatch_diffusion_batch(b);
    }

    bool dispatch_diffusion_batch(const DiffusionBatchDispatch& d) override
    {
        if (!ctx_ || d.count < 1 || !d.locations) return false;
        try
        {
            std::vector<int> locations(static_cast<std::size_t>(d.count));
            for (int t = 0; t < d.count; ++t) locations[static_cast<std::size_t>(t)] = static_cast<int>(d.locations[t]);
            nimbus_diffusion_batch_dispatch c{};
            c.fields = d.fields; c.tendencies = d.tendencies; c.locations = locations.data(); c.count = d.count;
            c.rho = d.rho; c.dz = d.dz;
            c.nx = d.nx; c.ny = d.ny; c.nz = d.nz;
            c.dx = d.dx; c.dy = d.dy; c.K_h = d.K_h; c.K_v = d.K_v;
            c.c_grid = d.c_grid ? 1 : 0; c.periodic_x = d.periodic_x ? 1 : 0; c.periodic_y = d.periodic_y ? 1 : 0;
            c.vertical_K = d.vertical_K;
            return nimbus_hip_dispatch_diffusion_batch(ctx_, &c) == NIMBUS_OK;
        }
        catch (...)
        {
            return false;
        }
    }

private:
    nimbus_hip_ctx*& ctx_;
};

class HipComputeBackend final : public ComputeBackend
{
public:
    HipComputeBackend() : dynamics_(ctx_), advection_(ctx_), microphysics_(ctx_), diffusion_(ctx_) {}
    ~HipComputeBackend() override { shutdown(); }

    std::string name() const override { return "hip"; }
    void configure(const ComputeBackendConfig& config) override { device_ = config.device_index < 0 ? 0 : config.device_index; }

    bool initialize(std::string& error) override
    {
        if (ctx_) return true;
        if (nimbus_hip_create(device_, &ctx_) != NIMBUS_OK)
        {
            error = std::string("HIP backend: ") + nimbus_hip_last_error();
            ctx_ = nullptr;
            return false;
        }
        return true;
    }

    void shutdown() override
    {
        if (ctx_) nimbus_hip_destroy(ctx_);
        ctx_ = nullptr;
    }

    ComputeMemoryProperties memory_properties() const override
    {
        ComputeMemoryProperties m{};
        m.accelerator = ctx_ != nullptr;
        std::uint64_t free_bytes = 0, total_bytes = 0;
        if (ctx_ && nimbus_hip_mem_info(ctx_, &free_bytes, &total_bytes) == NIMBUS_OK)
        {
            m.device_local_heap_bytes = total_bytes;
            m.max_storage_buffer_bytes = total_bytes;
            m.device_budget_is_live = true;
            m.device_available_budget_bytes = free_bytes;
        }
        return m;
    }

    std::uint64_t successful_dispatch_count() const override { return ctx_ ? nimbus_hip_dispatch_count(ctx_) : 0; }

    std::uint64_t successful_kernel_dispatch_count(nimbus::compute::KernelId kernel) const override
    {
        using nimbus::compute::KernelId;
        const char* op = nullptr;
        switch (kernel)
        {
        case KernelId::CartesianCgridSlowTendencies: op = "cartesian_cgrid_slow_tendencies"; break;
        case KernelId::SlowApply: op = "slow_apply"; break;
        case KernelId::StateRelaxationCartesian: op = "state_relaxation_cartesian"; break;
        case KernelId::KesslerPointwise: op = "kessler_pointwise"; break;
        case KernelId::KesslerSedimentation: op = "kessler_sedimentation"; break;
        case KernelId::MicrophysicsApply: op = "microphysics_apply"; break;
        case KernelId::AdvectConservativeUpdateCartesianCgrid: op = "advect_scalars_conservative_cartesian_cgrid"; break;
        case KernelId::Diffusion: op = "diffusion_batch"; break;
        default: return 0;
        }
        return ctx_ ? nimbus_hip_successful_dispatch_count(ctx_, op) : 0;
    }

    AdvectionComputeKernels* advection_kernels() override { return ctx_ ? &advection_ : nullptr; }
    const AdvectionComputeKernels* advection_kernels() const override { return ctx_ ? &advection_ : nullptr; }
    DynamicsComputeKernels* dynamics_kernels() override { return ctx_ ? &dynamics_ : nullptr; }
    const DynamicsComputeKernels* dynamics_kernels() const override { return ctx_ ? &dynamics_ : nullptr; }
    MicrophysicsComputeKernels* microphysics_kernels() override { return ctx_ ? &microphysics_ : nullptr; }
    const MicrophysicsComputeKernels* microphysics_kernels() const override { return ctx_ ? &microphysics_ : nullptr; }
    DiffusionComputeKernels* diffusion_kernels() override { return ctx_ ? &diffusion_ : nullptr; }
    const DiffusionComputeKernels* diffusion_kernels() const override { return ctx_ ? &diffusion_ : nullptr; }

    // The per-operation calls stage their arrays each time, so nothing is
    // resident between them: the residency hooks have no state to keep.
    void begin_resident_step() override {}
    void invalidate_resident_field(const char*) noexcept override {}
    void invalidate_all_resident_fields() noexcept override {}

private:
    int device_ = 0;
    nimbus_hip_ctx* ctx_ = nullptr;
    Dynamics dynamics_;
    Advection advection_;
    Microphysics microphysics_;
    Diffusion diffusion_;
};

}  // namespace nimbus_hip_adapter

// Factory entry point beside create_cpu/vulkan/metal_compute_backend
// (include/compute/compute_backend_factory.hpp:8-10).
std::unique_ptr<ComputeBackend> create_hip_compute_backend()
{
    return std::make_unique<nimbus_hip_adapter::HipComputeBackend>();
}
