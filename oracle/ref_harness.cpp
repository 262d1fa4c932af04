// TEST INFRASTRUCTURE ONLY: a driver that calls the Nimbus reference's own
// CPU code (compiled from the reference sources where they lie, see the
// `ref` target of oracle/Makefile) to produce golden vectors for the oracle
// and the HIP backend. It runs only in the development container; its
// outputs are committed as fixtures under tests/golden/.
//
// Usage: ref_harness <op> <dir>. Inputs are raw little-endian arrays in
// <dir> (named per op below, with a params.txt of scalars), outputs are
// written back into <dir>.
//
//   weno      in5.f64 (N x 5)          -> weno.f64 (N), offset.f32 (N, from in5 rounded to float)
//   thermo    tpqq.f64 (N x 4: T, p, qv, qc) -> satadj.f64 (N x 3), qvs.f64 (N), t_of_theta.f64 (N), vt.f64 (N)
//   upwind    params: nx ny nz dx dy dz; f.f32; pts.i32 (M x 4: i j k axis); a.f64 (M) -> upwind.f64 (M)
//   kessler   params: nx ny nz dz dt; p theta rho qv qc qr .f32 -> dth dqv dqc dqr .f32 and sink.f64 (nx*ny)
//   acoustic  params: nx ny nz dx dy dz dts; rho0 p0 .f64 (nz); u v w rho p theta .f32 -> *_out.f32
#include "compute/cpu/cpu_backend.hpp"
#include "core/field/field3d.hpp"
#include "core/field/field_bounds.hpp"
#include "core/infra/physical_constants.hpp"
#include "core/infra/physical_cell_domain.hpp"
#include "microphysics/base/fall_speeds.hpp"
#include "microphysics/base/thermodynamics.hpp"
#include "microphysics/schemes/kessler/kessler.hpp"
#include "numerics/advection/weno5_face.hpp"
#include "numerics/derivatives/derivative_operators.hpp"

#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <iostream>
#include <string>
#include <vector>

namespace
{

template <class T>
std::vector<T> read_bin(const std::string& path)
{
    std::ifstream f(path, std::ios::binary | std::ios::ate);
    if (!f) throw std::runtime_error("cannot open " + path);
    const std::streamsize n = f.tellg();
    f.seekg(0);
    std::vector<T> v(std::size_t(n) / sizeof(T));
    f.read(reinterpret_cast<char*>(v.data()), n);
    return v;
}

template <class T>
void write_bin(const std::string& path, const std::vector<T>& v)
{
    std::ofstream f(path, std::ios::binary);
    f.write(reinterpret_cast<const char*>(v.data()), std::streamsize(v.size() * sizeof(T)));
}

std::vector<double> read_params(const std::string& dir)
{
    std::ifstream f(dir + "/params.txt");
    std::vector<double> p;
    double x;
    while (f >> x) p.push_back(x);
    return p;
}

Field3D to_field(const std::vector<float>& v, int nx, int ny, int nz)
{
    Field3D f(nx, ny, nz, 0.0f);
    std::copy(v.begin(), v.end(), f.data());
    return f;
}

std::vector<float> from_field(const Field3D& f) { return std::vector<float>(f.data(), f.data() + f.size()); }

int op_weno(const std::string& d)
{
    const auto in = read_bin<double>(d + "/in5.f64");
    const std::size_t n = in.size() / 5;
    std::vector<double> out(n);
    std::vector<float> off(n);
    for (std::size_t q = 0; q < n; ++q)
    {
        const double* a = &in[5 * q];
        out[q] = weno5_face_value(a[0], a[1], a[2], a[3], a[4]);
        off[q] = weno5_face_offset_fp32(float(a[0]), float(a[1]), float(a[2]), float(a[3]), float(a[4]));
    }
    write_bin(d + "/weno.f64", out);
    write_bin(d + "/offset.f32", off);
    return 0;
}

int op_thermo(const std::string& d)
{
    const auto in = read_bin<double>(d + "/tpqq.f64");
    const std::size_t n = in.size() / 4;
    std::vector<double> sat(3 * n), qvs(n), tt(n), vt(n);
    for (std::size_t q = 0; q < n; ++q)
    {
        double T = in[4 * q], p = in[4 * q + 1], qv = in[4 * q + 2], qc = in[4 * q + 3];
        qvs[q] = tmv::thermodynamics::saturation_mixing_ratio_water(T, p);
        tt[q] = tmv::thermodynamics::theta_to_temperature(T + 20.0, p);
        vt[q] = tmv::fall_speeds::kessler_rain(qv, 0.5 + qc * 100.0);
        const double Ta = tmv::thermodynamics::saturation_adjustment(T, p, qv, qc);
        sat[3 * q] = Ta;
        sat[3 * q + 1] = qv;
        sat[3 * q + 2] = qc;
    }
    write_bin(d + "/satadj.f64", sat);
    write_bin(d + "/qvs.f64", qvs);
    write_bin(d + "/t_of_theta.f64", tt);
    write_bin(d + "/vt.f64", vt);
    return 0;
}

int op_upwind(const std::string& d)
{
    const auto p = read_params(d);
    const int nx = int(p[0]), ny = int(p[1]), nz = int(p[2]);
    const Field3D f = to_field(read_bin<float>(d + "/f.f32"), nx, ny, nz);
    const auto pts = read_bin<std::int32_t>(d + "/pts.i32");
    const auto a = read_bin<double>(d + "/a.f64");
    const StaggeredCartesianDerivatives deriv(p[3], p[4], p[5], nx, ny, nz);
    std::vector<double> out(a.size());
    for (std::size_t q = 0; q < a.size(); ++q)
    {
        const int i = pts[4 * q], j = pts[4 * q + 1], k = pts[4 * q + 2], axis = pts[4 * q + 3];
        out[q] = axis == 0 ? deriv.upwind_dx(f, a[q], i, j, k)
                           : (axis == 1 ? deriv.upwind_dy(f, a[q], i, j, k) : deriv.upwind_dz(f, a[q], i, j, k));
    }
    write_bin(d + "/upwind.f64", out);
    return 0;
}

int op_kessler(const std::string& d)
{
    const auto p = read_params(d);
    const int nx = int(p[0]), ny = int(p[1]), nz = int(p[2]);
    const double dz = p[3], dt = p[4];
    auto F = [&](const char* name) { return to_field(read_bin<float>(d + "/" + name + ".f32"), nx, ny, nz); };
    const Field3D pf = F("p"), th = F("theta"), rho = F("rho"), qv = F("qv"), qc = F("qc"), qr = F("qr");
    const Field3D zero(nx, ny, nz, 0.0f);
    Field3D dth, dqv, dqc, dqr, dqi, dqs, dqg, dqh;
    tmv::KesslerScheme k;
    k.set_vertical_spacing_m(dz);
    k.set_physical_cell_domain(nx, ny, nz, PhysicalCellDomain{1, nx - 1, 1, ny - 1, 0, nz});
    k.compute_tendencies(pf, th, rho, qv, qc, qr, zero, zero, zero, zero, dt, dth, dqv, dqc, dqr, dqi, dqs, dqg, dqh);
    write_bin(d + "/dth.f32", from_field(dth));
    write_bin(d + "/dqv.f32", from_field(dqv));
    write_bin(d + "/dqc.f32", from_field(dqc));
    write_bin(d + "/dqr.f32", from_field(dqr));
    std::vector<double> sink(std::size_t(nx) * ny, 0.0);
    for (int i = 1; i < nx - 1; ++i)
        for (int j = 1; j < ny - 1; ++j)
            sink[std::size_t(i) * ny + j] = k.bottom_sedimentation_sink_rate(rho, qr, zero, zero, zero, i, j, dt);
    write_bin(d + "/sink.f64", sink);
    return 0;
}

int op_acoustic(const std::string& d)
{
    const auto p = read_params(d);
    const int nx = int(p[0]), ny = int(p[1]), nz = int(p[2]);
    const double dx = p[3], dy = p[4], dz = p[5], dts = p[6];
    const auto rho0 = read_bin<double>(d + "/rho0.f64");
    const auto p0 = read_bin<double>(d + "/p0.f64");
    auto u = read_bin<float>(d + "/u.f32"), v = read_bin<float>(d + "/v.f32"), w = read_bin<float>(d + "/w.f32");
    auto rho = read_bin<float>(d + "/rho.f32"), pr = read_bin<float>(d + "/p.f32");
    const auto th = read_bin<float>(d + "/theta.f32");
    std::vector<float> scratch(u.size(), 0.0f);
    AcousticSubstepFusedCartesianCgridDispatch disp;
    disp.u = u.data(); disp.v = v.data(); disp.w = w.data(); disp.rho = rho.data(); disp.p = pr.data();
    disp.theta_data = th.data();
    disp.base_len = nz;
    disp.nx = nx; disp.ny = ny; disp.nz = nz;
    disp.dx = float(dx); disp.dy = float(dy); disp.dz = float(dz);
    disp.dt_small = float(dts);
    disp.gamma_val = float(physical_constants::dry_air_gamma);
    disp.rho_min = density_min_kgm3; disp.rho_max = density_max_kgm3;
    disp.p_min = pressure_min_pa; disp.p_max = pressure_max_pa;
    disp.p_ref_fallback = float(physical_constants::reference_pressure_pa);
    disp.wind_clamp_h = wind_horizontal_abs_max_ms; disp.wind_clamp_v = wind_vertical_abs_max_ms;
    disp.rho0_base_host = rho0.data(); disp.p0_base_host = p0.data();
    disp.dx_host = dx; disp.dy_host = dy; disp.dz_host = dz; disp.dt_small_host = dts;
    disp.gamma_host = physical_constants::dry_air_gamma;
    disp.drho_scratch = scratch.data();
    disp.repair_primary_state = true;
    nimbus::compute::cpu_internal::CpuComputeBackend be;
    std::string err;
    be.initialize(err);
    if (!be.dispatch_acoustic_substep_fused_cartesian_cgrid(disp))
    {
        std::cerr << "reference acoustic dispatch refused the inputs\n";
        return 2;
    }
    write_bin(d + "/u_out.f32", u);
    write_bin(d + "/v_out.f32", v);
    write_bin(d + "/w_out.f32", w);
    write_bin(d + "/rho_out.f32", rho);
    write_bin(d + "/p_out.f32", pr);
    return 0;
}

}  // namespace

int main(int argc, char** argv)
{
    if (argc != 3)
    {
        std::cerr << "usage: ref_harness <weno|thermo|upwind|kessler|acoustic> <dir>\n";
        return 1;
    }
    const std::string op = argv[1], dir = argv[2];
    try
    {
        if (op == "weno") return op_weno(dir);
        if (op == "thermo") return op_thermo(dir);
        if (op == "upwind") return op_upwind(dir);
        if (op == "kessler") return op_kessler(dir);
        if (op == "acoustic") return op_acoustic(dir);
    }
    catch (const std::exception& e)
    {
        std::cerr << "error: " << e.what() << "\n";
        return 3;
    }
    std::cerr << "unknown op\n";
    return 1;
}
