// TEST INFRASTRUCTURE ONLY: runs the Nimbus reference's own Smagorinsky
// operators (src/turbulence/base/eddy_viscosity.cpp, compiled where it lies
// by the `ref` target of oracle/Makefile) on a state given as files, to
// produce golden vectors for the SGS oracle. The coefficient loop follows
// SmagorinskyScheme::compute_eddy_coefficients
// (src/turbulence/schemes/smagorinsky/smagorinsky.cpp:495-567, flat C-grid,
// stability correction "none").
//
// Usage: ref_sgs <dir>. <dir>/params.txt holds
//   nx ny nz dx dy dz mode filter Cs Pr_t Sc_t nu_max K_max periodic_x periodic_y open_w open_e open_s open_n
// (mode 0 = 3d, 1 = horizontal_only, 2 = vertical_only; filter 0 = cubic_root, 1 = dx);
// inputs u v w rho theta qv .f32 ([nx][ny][nz]) and rho0.f64 (nz); outputs
// nu kth kq du dv dw dth dqv .f32.
#include "core/field/field3d.hpp"
#include "turbulence/base/eddy_viscosity.hpp"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <fstream>
#include <string>
#include <vector>

using namespace tmv;

namespace
{
template <class T>
std::vector<T> read_bin(const std::string& path, std::size_t n)
{
    std::vector<T> v(n);
    std::ifstream f(path, std::ios::binary);
    f.read(reinterpret_cast<char*>(v.data()), std::streamsize(n * sizeof(T)));
    return v;
}

void load(Field3D& f, const std::string& path, std::size_t n)
{
    const std::vector<float> v = read_bin<float>(path, n);
    std::memcpy(f.data(), v.data(), n * sizeof(float));
}

void save(const Field3D& f, const std::string& path, std::size_t n)
{
    std::ofstream o(path, std::ios::binary);
    o.write(reinterpret_cast<const char*>(f.data()), std::streamsize(n * sizeof(float)));
}
}  // namespace

int main(int argc, char** argv)
{
    if (argc < 2) return 2;
    const std::string d = argv[1];
    std::ifstream pf(d + "/params.txt");
    int nx, ny, nz, mode, filter, px, py, ow, oe, os, on;
    double dx, dy, dz, cs, prt, sct, nu_max, k_max;
    pf >> nx >> ny >> nz >> dx >> dy >> dz >> mode >> filter >> cs >> prt >> sct >> nu_max >> k_max >> px >> py >> ow >>
        oe >> os >> on;
    const std::size_t n = std::size_t(nx) * ny * nz;
    Field3D u(nx, ny, nz, 0.0f), v(nx, ny, nz, 0.0f), w(nx, ny, nz, 0.0f), rho(nx, ny, nz, 0.0f),
        th(nx, ny, nz, 0.0f), qv(nx, ny, nz, 0.0f);
    load(u, d + "/u.f32", n);
    load(v, d + "/v.f32", n);
    load(w, d + "/w.f32", n);
    load(rho, d + "/rho.f32", n);
    load(th, d + "/theta.f32", n);
    load(qv, d + "/qv.f32", n);
    std::vector<double> rho0 = read_bin<double>(d + "/rho0.f64", std::size_t(nz));

    TurbulenceConfig cfg;
    cfg.mode = mode == 1 ? "horizontal_only" : (mode == 2 ? "vertical_only" : "3d");
    cfg.filter_width = filter == 1 ? "dx" : "cubic_root";
    cfg.Cs = cs;
    cfg.Pr_t = prt;
    cfg.Sc_t = sct;
    cfg.nu_t_max = nu_max;
    cfg.K_max = k_max;
    TurbulenceStateView s;
    s.u = &u; s.v = &v; s.w = &w; s.rho = &rho; s.theta = &th; s.qv = &qv; s.rho0 = &rho0;
    s.NR = nx; s.NTH = ny; s.NZ = nz;
    s.periodic_x = px != 0;
    s.periodic_y = py != 0;
    s.open_lateral = {{ow != 0, oe != 0, os != 0, on != 0}};
    GridMetrics g;
    g.dx = dx;
    g.dy = dy;
    g.dz.assign(std::size_t(nz), dz);

    Field3D nu(nx, ny, nz, 0.0f), kth(nx, ny, nz, 0.0f), kq(nx, ny, nz, 0.0f);
    for (int i = 0; i < nx; ++i)
        for (int j = 0; j < ny; ++j)
            for (int k = 0; k < nz; ++k)
            {
                const double strain_sq = eddy_viscosity::compute_directional_strain_sq_centered(s, g, i, j, k, cfg.mode);
                const double delta = eddy_viscosity::compute_filter_width(cfg, g, 0, 0, k);
                double nu_t = eddy_viscosity::compute_smagorinsky_viscosity(cs, delta, std::sqrt(strain_sq), 1.0);
                nu_t = std::min(nu_t, nu_max);
                nu[i][j][k] = static_cast<float>(nu_t);
                double kt, kqq, ktke;
                eddy_viscosity::compute_eddy_diffusivities(nu_t, prt, sct, kt, kqq, ktke);
                kth[i][j][k] = static_cast<float>(std::min(kt, k_max));
                kq[i][j][k] = static_cast<float>(std::min(kqq, k_max));
            }
    Field3D du(nx, ny, nz, 0.0f), dv(nx, ny, nz, 0.0f), dw(nx, ny, nz, 0.0f), dth(nx, ny, nz, 0.0f),
        dqv(nx, ny, nz, 0.0f);
    eddy_viscosity::compute_momentum_diffusion_tendencies(cfg, s, g, nu, du, dv, dw, s.periodic_x, s.periodic_y,
                                                          nullptr);
    eddy_viscosity::compute_scalar_diffusion_tendencies(cfg, s, g, kth, th, dth);
    eddy_viscosity::compute_scalar_diffusion_tendencies(cfg, s, g, kq, qv, dqv);
    save(nu, d + "/nu.f32", n);
    save(kth, d + "/kth.f32", n);
    save(kq, d + "/kq.f32", n);
    save(du, d + "/du.f32", n);
    save(dv, d + "/dv.f32", n);
    save(dw, d + "/dw.f32", n);
    save(dth, d + "/dth.f32", n);
    save(dqv, d + "/dqv.f32", n);
    return 0;
}
