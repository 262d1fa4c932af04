// TEST INFRASTRUCTURE ONLY: runs the Nimbus reference's own explicit
// diffusion scheme (src/numerics/diffusion/schemes/explicit/explicit.cpp,
// compiled where it lies by the `ref` target of oracle/Makefile) through
// ExplicitDiffusionScheme::compute_diffusion_tendencies on the CPU path (no
// compute runtime, flat grid), to produce golden vectors for the diffusion
// oracle.
//
// Usage: ref_diffusion <dir>. <dir>/params.txt holds
//   nx ny nz dx dy K_h K_v c_grid periodic_x periodic_y variable_K apply_to
// (apply_to 0 = scalars, 1 = momentum, 2 = all); inputs u v w theta qv rho
// .f32 ([nx][ny][nz]), kv.f32 when variable_K, dz.f64 (nz); outputs du dv dw
// dtheta dqv .f32.
#include "core/field/field3d.hpp"
#include "numerics/diffusion/schemes/explicit/explicit.hpp"

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
    int nx, ny, nz, c_grid, px, py, variable, apply;
    double dx, dy, kh, kv;
    pf >> nx >> ny >> nz >> dx >> dy >> kh >> kv >> c_grid >> px >> py >> variable >> apply;
    const std::size_t n = std::size_t(nx) * ny * nz;
    Field3D u(nx, ny, nz, 0.0f), v(nx, ny, nz, 0.0f), w(nx, ny, nz, 0.0f), th(nx, ny, nz, 0.0f),
        qv(nx, ny, nz, 0.0f), rho(nx, ny, nz, 0.0f), kvf(nx, ny, nz, 0.0f);
    load(u, d + "/u.f32", n);
    load(v, d + "/v.f32", n);
    load(w, d + "/w.f32", n);
    load(th, d + "/theta.f32", n);
    load(qv, d + "/qv.f32", n);
    load(rho, d + "/rho.f32", n);
    if (variable) load(kvf, d + "/kv.f32", n);

    GridMetrics grid;
    grid.dx = dx;
    grid.dy = dy;
    grid.dz = read_bin<double>(d + "/dz.f64", std::size_t(nz));

    DiffusionConfig cfg;
    cfg.K_h = kh;
    cfg.K_v = kv;
    cfg.apply_to = apply == 0 ? "scalars" : (apply == 1 ? "momentum" : "all");
    cfg.use_variable_K = variable != 0;
    cfg.dt_diffusion = 1.0;
    ExplicitDiffusionScheme scheme;
    scheme.initialize(cfg);

    DiffusionStateView s;
    s.grid = &grid;
    s.rho = &rho;
    s.c_grid = c_grid != 0;
    s.periodic_x = px != 0;
    s.periodic_y = py != 0;
    if (apply != 0)
    {
        s.u = &u;
        s.v = &v;
        s.w = &w;
    }
    if (apply != 1)
    {
        s.theta = &th;
        s.qv = &qv;
    }
    if (variable)
    {
        s.K_scalar = &kvf;
        s.K_momentum = &kvf;
    }
    DiffusionTendencies t;
    scheme.compute_diffusion_tendencies(cfg, s, t, nullptr);
    save(t.dudt_diff, d + "/du.f32", n);
    save(t.dvdt_diff, d + "/dv.f32", n);
    save(t.dwdt_diff, d + "/dw.f32", n);
    save(t.dthetadt_diff, d + "/dtheta.f32", n);
    save(t.dqvdt_diff, d + "/dqv.f32", n);
    return 0;
}
