// Drives the HipComputeBackend adapter through the reference's ComputeBackend
// interface on a host without a GPU: the factory, the name, families that are
// unavailable before initialize(), and an initialize() that fails with an
// error message instead of throwing.
#include "compute/backend_interface.hpp"

#include <cstdio>
#include <memory>
#include <string>

std::unique_ptr<ComputeBackend> create_hip_compute_backend();

int main()
{
    std::unique_ptr<ComputeBackend> b = create_hip_compute_backend();
    if (!b || b->name() != "hip") return 1;
    if (b->dynamics_kernels() || b->advection_kernels() || b->microphysics_kernels() || b->diffusion_kernels())
        return 2;
    ComputeBackendConfig cfg;
    cfg.backend = "hip";
    cfg.device_index = 0;
    b->configure(cfg);
    std::string err;
    const bool ok = b->initialize(err);
    if (ok)
    {
        if (!b->dynamics_kernels() ||
            !b->dynamics_kernels()->supports(DynamicsKernelOp::CartesianCgridSlowTendencies) ||
            !b->diffusion_kernels() || !b->diffusion_kernels()->supports(DiffusionKernelOp::Diffusion))
            return 3;
        std::printf("initialized\n");
    }
    else
    {
        if (err.empty()) return 4;
        std::printf("refused: %s\n", err.c_str());
    }
    b->shutdown();
    std::printf("ok\n");
    return 0;
}
