// Internal declarations shared by the host runtime, the kernels and the
// initial-condition generator of the HIP backend.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>
#include <string>

#include "nimbus_hip.h"

namespace nimbus
{

// include/core/infra/physical_constants.hpp:14-34; microphysics_base.hpp:1071-1081
constexpr double kG = 9.81;
constexpr double kRd = 287.0;
constexpr double kRv = 461.5;
constexpr double kCp = 1004.0;
constexpr double kLv = 2.5e6;
constexpr double kP00 = 100000.0;
constexpr double kT0 = 273.15;
constexpr double kGamma = kCp / (kCp - kRd);
constexpr double kKappa = kRd / kCp;

int base_state(const nimbus_grid& g, const nimbus_sounding& s, double* rho0, double* p0, double* th0, double* qv0,
               double* u0, double* v0);
template <class S>
int fill_state(const nimbus_grid& g, const nimbus_ic_params& prm, const nimbus_host_state& out);
template <class S>
int fill_from_base(const nimbus_grid& g, const nimbus_base_state& b, const nimbus_ic_params& prm,
                   const nimbus_host_state& out);
int sounding_levels(const nimbus_sounding& s, int n, const double* z, nimbus_cm1_sounding& out);
// Records msg as the calling thread's last error (nimbus_hip_last_error) and returns code.
int set_error(int code, const std::string& msg);
int write_npy(const char* path, const void* data, int dtype, int ndim, const std::int64_t* shape);
double ic_uniform(std::uint64_t seed, std::uint32_t member, int i, int j, int k);
int resolve_substeps(const nimbus_grid& g, const double* rho0, const double* p0, const nimbus_step_config& c);
// Device, launch stream and dispatch counter of a context (runtime.hip).
int context_device(nimbus_hip_ctx* ctx);
hipStream_t context_stream(nimbus_hip_ctx* ctx);
void count_dispatch(nimbus_hip_ctx* ctx, const char* op);

// Device grid descriptor. Device fields are stored [k][j][i] with i
// contiguous and the row pitch px (elements) padded to 64 bytes.
struct DevGrid
{
    int nx, ny, nz, px;
    double dx, dy, dz;
    double idx, idy, idz;  // 1/dx, 1/dy, 1/dz
    __host__ __device__ std::size_t at(int i, int j, int k) const
    {
        return (std::size_t(k) * std::size_t(ny) + std::size_t(j)) * std::size_t(px) + std::size_t(i);
    }
    __host__ __device__ std::size_t cells() const { return std::size_t(px) * ny * nz; }
    __host__ __device__ std::size_t plane() const { return std::size_t(px) * ny; }
};

// Base-state profile block in device memory (FP64), each nz long.
enum BaseSlot
{
    kBaseRho0 = 0,
    kBaseP0 = 1,
    kBaseTh0 = 2,
    kBaseQv0Stored = 3,  // qv0 rounded to FP32 as the reference stores it (cartesian_common.hpp:31-35)
    kBaseU0 = 4,
    kBaseV0 = 5,
    kBasePe0 = 6,          // EOS(rho0, theta0): reference point of the perturbation EOS
    kBasePeDiff = 7,       // EOS(rho0, theta0) - p0
    kBaseInvRT0 = 8,       // 1 / (rho0 theta0)
    kBaseSpongeUp = 9,     // upper sponge decay factor 1 - exp(-alpha dt) at cell centres
    kBaseSpongeUpW = 10,   // the same at w faces (z = (k+1) dz)
    kBaseSpongeW = 11,     // w-only absorber factor under the lid
    kBaseCount = 12
};

// Deepest supported column: kernels stage five FP64 base-state profiles of
// the deepest member in LDS (5 * 8 * 1024 B = 40 KiB per workgroup).
constexpr int kMaxLevels = 1024;

// Prognostic bounds (include/core/field/field_bounds.hpp:10-20).
struct Bounds
{
    float rho_min = 0.05f, rho_max = 2.0f;
    float p_min = 1000.0f, p_max = 200000.0f;
    float p_fallback = 100000.0f;
    float wind_h = 180.0f, wind_v = 120.0f;
};

struct KesslerConst
{
    double qc0, c_auto, c_accr, evap_eff, vt_max;
};

}  // namespace nimbus
