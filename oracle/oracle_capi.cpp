// C entry points of the CPU oracle (test infrastructure only; see nimbus_oracle.hpp).
// Host arrays use the reference layout [nx][ny][nz], k contiguous.
#include "nimbus_oracle.hpp"

#include <cstring>
#include <omp.h>

extern "C" {

struct oc_grid
{
    int nx, ny, nz;
    double dx, dy, dz;
};

struct oc_config
{
    double dt;
    int substeps;
    int acoustic_mode;
    double hevi_beta;
    double coriolis_f;
    int bc_west, bc_east, bc_south, bc_north, bc_bottom, bc_top;
    double qc0, c_auto, c_accr, evap_eff, vt_max;
    int sponge;
    double sponge_z_d, sponge_alpha;
    int sponge_w_cells;
    double sponge_w_alpha;
    int sponge_lateral_cells;
    double sponge_lateral_alpha;
    int sgs, sgs_mode, sgs_filter;
    double sgs_cs, sgs_pr_t, sgs_sc_t, sgs_nu_max, sgs_k_max, sgs_dt;
    int diffusion;
    double diffusion_kh, diffusion_kv;
};

struct oc_base
{
    const double *rho0, *p0, *theta0, *qv0, *u0, *v0;
};

// Nine prognostic fields (u, v, w, rho, p, theta, qv, qc, qr) plus the
// accumulated surface rain [nx*ny] (may be null).
struct oc_state
{
    void *u, *v, *w, *rho, *p, *theta, *qv, *qc, *qr;
    double* rain_mm;
};
}

namespace
{
using namespace oracle;

Grid to_grid(const oc_grid* g) { return Grid{g->nx, g->ny, g->nz, g->dx, g->dy, g->dz}; }

Config to_config(const oc_config* c)
{
    Config k;
    k.dt = c->dt;
    k.substeps = c->substeps;
    k.acoustic_mode = c->acoustic_mode;
    k.hevi_beta = c->hevi_beta;
    k.coriolis_f = c->coriolis_f;
    k.bc_west = c->bc_west; k.bc_east = c->bc_east; k.bc_south = c->bc_south; k.bc_north = c->bc_north;
    k.bc_bottom = c->bc_bottom; k.bc_top = c->bc_top;
    k.qc0 = c->qc0; k.c_auto = c->c_auto; k.c_accr = c->c_accr; k.evap_eff = c->evap_eff; k.vt_max = c->vt_max;
    k.sponge = c->sponge; k.sponge_z_d = c->sponge_z_d; k.sponge_alpha = c->sponge_alpha;
    k.sponge_w_cells = c->sponge_w_cells; k.sponge_w_alpha = c->sponge_w_alpha;
    k.sponge_lateral_cells = c->sponge_lateral_cells; k.sponge_lateral_alpha = c->sponge_lateral_alpha;
    k.sgs = c->sgs; k.sgs_mode = c->sgs_mode; k.sgs_filter = c->sgs_filter; k.sgs_cs = c->sgs_cs;
    k.sgs_pr_t = c->sgs_pr_t; k.sgs_sc_t = c->sgs_sc_t; k.sgs_nu_max = c->sgs_nu_max; k.sgs_k_max = c->sgs_k_max;
    k.sgs_dt = c->sgs_dt;
    k.diffusion = c->diffusion; k.diffusion_kh = c->diffusion_kh; k.diffusion_kv = c->diffusion_kv;
    return k;
}

Base to_base(const oc_base* b, int nz)
{
    Base r;
    auto cp = [nz](const double* p) { return p ? std::vector<double>(p, p + nz) : std::vector<double>(nz, 0.0); };
    r.rho0 = cp(b->rho0); r.p0 = cp(b->p0); r.theta0 = cp(b->theta0);
    r.qv0 = cp(b->qv0); r.u0 = cp(b->u0); r.v0 = cp(b->v0);
    return r;
}

template <class S>
void load(const Grid& g, const oc_state* st, State<S>& s)
{
    s.allocate(g);
    void* const src[] = {st->u, st->v, st->w, st->rho, st->p, st->theta, st->qv, st->qc, st->qr};
    std::vector<S>* dst[] = {&s.u, &s.v, &s.w, &s.rho, &s.p, &s.theta, &s.qv, &s.qc, &s.qr};
    for (int f = 0; f < 9; ++f) std::memcpy(dst[f]->data(), src[f], g.cells() * sizeof(S));
    if (st->rain_mm) std::memcpy(s.rain_mm.data(), st->rain_mm, sizeof(double) * g.nx * g.ny);
}

template <class S>
void store(const Grid& g, const State<S>& s, const oc_state* st)
{
    void* const dst[] = {st->u, st->v, st->w, st->rho, st->p, st->theta, st->qv, st->qc, st->qr};
    const std::vector<S>* src[] = {&s.u, &s.v, &s.w, &s.rho, &s.p, &s.theta, &s.qv, &s.qc, &s.qr};
    for (int f = 0; f < 9; ++f) std::memcpy(dst[f], src[f]->data(), g.cells() * sizeof(S));
    if (st->rain_mm) std::memcpy(st->rain_mm, s.rain_mm.data(), sizeof(double) * g.nx * g.ny);
}

template <class S>
int run_steps(const oc_grid* gp, const oc_base* bp, const oc_config* cp, const oc_state* st, int nsteps)
{
    const Grid g = to_grid(gp);
    const Base b = to_base(bp, g.nz);
    const Config c = to_config(cp);
    State<S> s;
    load(g, st, s);
    int n = 0;
    for (int t = 0; t < nsteps; ++t) n = large_step(g, b, c, s);
    store(g, s, st);
    return n;
}

template <class S>
std::vector<S> vec(const void* p, std::size_t n)
{
    const S* q = static_cast<const S*>(p);
    return std::vector<S>(q, q + n);
}
template <class S>
int substeps_batched(const oc_grid* gp, const oc_base* bp, const oc_config* cp, const oc_state* st, int n,
                     double dts, double weight, S* mfx, S* mfy, S* mfz)
{
    const Grid g = to_grid(gp);
    const Base b = to_base(bp, g.nz);
    const Config c = to_config(cp);
    State<S> s;
    load(g, st, s);
    std::vector<S> fx(mfx, mfx + g.cells()), fy(mfy, mfy + g.cells()), fz(mfz, mfz + g.cells());
    seed_radiation(g, s);
    std::vector<double> wz;
    for (int q = 0; q < n; ++q)
    {
        if (c.acoustic_mode == kAcousticHEVI)
        {
            std::vector<S> zero(g.cells(), S(0));
            accumulate_mass_flux(g, c, s, weight, fx, fy, zero);
            acoustic_substep_hevi(g, b, s, dts, c.hevi_beta, wz);
            for (std::size_t k = 0; k < g.cells(); ++k) fz[k] = S(double(fz[k]) + weight * wz[k]);
        }
        else
        {
            accumulate_mass_flux(g, c, s, weight, fx, fy, fz);
            acoustic_substep_explicit(g, b, s, dts);
        }
        apply_boundaries(g, c, s, dts);
    }
    store(g, s, st);
    std::memcpy(mfx, fx.data(), sizeof(S) * g.cells());
    std::memcpy(mfy, fy.data(), sizeof(S) * g.cells());
    std::memcpy(mfz, fz.data(), sizeof(S) * g.cells());
    return 0;
}

}  // namespace

extern "C" {

int oracle_resolve_substeps(const oc_grid* g, const oc_base* b, const oc_config* c)
{
    return resolve_substeps(to_grid(g), to_base(b, g->nz), to_config(c));
}

int oracle_large_steps_f32(const oc_grid* g, const oc_base* b, const oc_config* c, const oc_state* s, int n)
{
    return run_steps<float>(g, b, c, s, n);
}

int oracle_large_steps_f64(const oc_grid* g, const oc_base* b, const oc_config* c, const oc_state* s, int n)
{
    return run_steps<double>(g, b, c, s, n);
}

// One explicit fused acoustic substep on (u, v, w, rho, p) given theta, the
// reference's AcousticSubstepFusedCartesianCgridDispatch contract.
int oracle_acoustic_substep_f32(const oc_grid* gp, const oc_base* bp, const oc_state* st, double dts)
{
    const Grid g = to_grid(gp);
    const Base b = to_base(bp, g.nz);
    State<float> s;
    load(g, st, s);
    acoustic_substep_explicit(g, b, s, dts);
    store(g, s, st);
    return 0;
}

int oracle_acoustic_substep_f64(const oc_grid* gp, const oc_base* bp, const oc_state* st, double dts)
{
    const Grid g = to_grid(gp);
    const Base b = to_base(bp, g.nz);
    State<double> s;
    load(g, st, s);
    acoustic_substep_explicit(g, b, s, dts);
    store(g, s, st);
    return 0;
}

// One HE-VI acoustic substep; wz_eff receives the weighted vertical mass flux.
int oracle_hevi_substep_f64(const oc_grid* gp, const oc_base* bp, const oc_state* st, double dts, double beta,
                            double* wz_eff)
{
    const Grid g = to_grid(gp);
    const Base b = to_base(bp, g.nz);
    State<double> s;
    load(g, st, s);
    std::vector<double> wz;
    acoustic_substep_hevi(g, b, s, dts, beta, wz);
    store(g, s, st);
    if (wz_eff) std::memcpy(wz_eff, wz.data(), sizeof(double) * g.cells());
    return 0;
}

int oracle_hevi_substep_f32(const oc_grid* gp, const oc_base* bp, const oc_state* st, double dts, double beta,
                            double* wz_eff)
{
    const Grid g = to_grid(gp);
    const Base b = to_base(bp, g.nz);
    State<float> s;
    load(g, st, s);
    std::vector<double> wz;
    acoustic_substep_hevi(g, b, s, dts, beta, wz);
    store(g, s, st);
    if (wz_eff) std::memcpy(wz_eff, wz.data(), sizeof(double) * g.cells());
    return 0;
}

// n substeps with per-substep boundaries after seeding the KW78 stash from the
// input faces, accumulating weight * face mass flux into mf* (the reference's
// batched substep contract, include/compute/dispatch/dynamics.hpp:279-325,726-731).
int oracle_acoustic_substeps_batched_f32(const oc_grid* gp, const oc_base* bp, const oc_config* cp, const oc_state* st,
                                         int n, double dts, double weight, float* mfx, float* mfy, float* mfz)
{
    return substeps_batched<float>(gp, bp, cp, st, n, dts, weight, mfx, mfy, mfz);
}

int oracle_acoustic_substeps_batched_f64(const oc_grid* gp, const oc_base* bp, const oc_config* cp, const oc_state* st,
                                         int n, double dts, double weight, double* mfx, double* mfy, double* mfz)
{
    return substeps_batched<double>(gp, bp, cp, st, n, dts, weight, mfx, mfy, mfz);
}

int oracle_apply_boundaries_f32(const oc_grid* gp, const oc_config* cp, const oc_state* st, double dt_rad)
{
    const Grid g = to_grid(gp);
    State<float> s;
    load(g, st, s);
    seed_radiation(g, s);
    apply_boundaries(g, to_config(cp), s, dt_rad);
    store(g, s, st);
    return 0;
}

int oracle_apply_boundaries_f64(const oc_grid* gp, const oc_config* cp, const oc_state* st, double dt_rad)
{
    const Grid g = to_grid(gp);
    State<double> s;
    load(g, st, s);
    seed_radiation(g, s);
    apply_boundaries(g, to_config(cp), s, dt_rad);
    store(g, s, st);
    return 0;
}

int oracle_slow_tendencies_f32(const oc_grid* gp, const oc_base* bp, const oc_config* cp, const oc_state* st,
                               float* du, float* dv, float* dw)
{
    const Grid g = to_grid(gp);
    State<float> s;
    load(g, st, s);
    std::vector<float> a, b, c;
    slow_tendencies(g, to_base(bp, g.nz), to_config(cp), s, a, b, c);
    std::memcpy(du, a.data(), sizeof(float) * g.cells());
    std::memcpy(dv, b.data(), sizeof(float) * g.cells());
    std::memcpy(dw, c.data(), sizeof(float) * g.cells());
    return 0;
}

int oracle_slow_tendencies_f64(const oc_grid* gp, const oc_base* bp, const oc_config* cp, const oc_state* st,
                               double* du, double* dv, double* dw)
{
    const Grid g = to_grid(gp);
    State<double> s;
    load(g, st, s);
    std::vector<double> a, b, c;
    slow_tendencies(g, to_base(bp, g.nz), to_config(cp), s, a, b, c);
    std::memcpy(du, a.data(), sizeof(double) * g.cells());
    std::memcpy(dv, b.data(), sizeof(double) * g.cells());
    std::memcpy(dw, c.data(), sizeof(double) * g.cells());
    return 0;
}

int oracle_slow_rk2_f32(const oc_grid* gp, const oc_base* bp, const oc_config* cp, const oc_state* st)
{
    const Grid g = to_grid(gp);
    State<float> s;
    load(g, st, s);
    slow_rk2_and_apply(g, to_base(bp, g.nz), to_config(cp), s);
    store(g, s, st);
    return 0;
}

int oracle_slow_rk2_f64(const oc_grid* gp, const oc_base* bp, const oc_config* cp, const oc_state* st)
{
    const Grid g = to_grid(gp);
    State<double> s;
    load(g, st, s);
    slow_rk2_and_apply(g, to_base(bp, g.nz), to_config(cp), s);
    store(g, s, st);
    return 0;
}

// Retained-flux WENO5 transport of one scalar (float storage, reference mode).
int oracle_transport_f32(const oc_grid* gp, const oc_config* cp, const float* q, float* out, const float* mfx,
                         const float* mfy, const float* mfz, const float* rho_begin, const float* rho_end,
                         double dt, int positive_definite)
{
    const Grid g = to_grid(gp);
    const std::size_t n = g.cells();
    std::vector<float> o;
    transport_scalar(g, to_config(cp), vec<float>(q, n), o, vec<float>(mfx, n), vec<float>(mfy, n),
                     vec<float>(mfz, n), vec<float>(rho_begin, n), vec<float>(rho_end, n), dt,
                     positive_definite != 0);
    std::memcpy(out, o.data(), sizeof(float) * n);
    return 0;
}

int oracle_transport_f64(const oc_grid* gp, const oc_config* cp, const double* q, double* out, const double* mfx,
                         const double* mfy, const double* mfz, const double* rho_begin, const double* rho_end,
                         double dt, int positive_definite)
{
    const Grid g = to_grid(gp);
    const std::size_t n = g.cells();
    std::vector<double> o;
    transport_scalar(g, to_config(cp), vec<double>(q, n), o, vec<double>(mfx, n), vec<double>(mfy, n),
                     vec<double>(mfz, n), vec<double>(rho_begin, n), vec<double>(rho_end, n), dt,
                     positive_definite != 0);
    std::memcpy(out, o.data(), sizeof(double) * n);
    return 0;
}

// Kessler tendencies (float storage, reference mode) on the physical domain.
int oracle_kessler_tendencies_f32(const oc_grid* gp, const oc_config* cp, const float* p, const float* theta,
                                  const float* rho, const float* qv, const float* qc, const float* qr, double dt,
                                  float* dth, float* dqv, float* dqc, float* dqr)
{
    const Grid g = to_grid(gp);
    const std::size_t n = g.cells();
    const Config c = to_config(cp);
    const KesslerParams kp{c.qc0, c.c_auto, c.c_accr, c.evap_eff, c.vt_max};
    std::vector<float> a, b, d, e;
    kessler_tendencies(g, kp, vec<float>(p, n), vec<float>(theta, n), vec<float>(rho, n), vec<float>(qv, n),
                       vec<float>(qc, n), vec<float>(qr, n), dt, a, b, d, e);
    std::memcpy(dth, a.data(), sizeof(float) * n);
    std::memcpy(dqv, b.data(), sizeof(float) * n);
    std::memcpy(dqc, d.data(), sizeof(float) * n);
    std::memcpy(dqr, e.data(), sizeof(float) * n);
    return 0;
}

// The microphysics interval; the lengths of its steps go to `lengths`
// (capacity `cap`), their count is returned.
int oracle_microphysics_step_f32(const oc_grid* gp, const oc_config* cp, const oc_state* st, double dt,
                                 double* lengths, int cap)
{
    const Grid g = to_grid(gp);
    State<float> s;
    load(g, st, s);
    std::vector<double> h;
    microphysics_step(g, to_config(cp), s, dt, &h);
    store(g, s, st);
    for (int q = 0; q < int(h.size()) && q < cap; ++q) lengths[q] = h[q];
    return int(h.size());
}

int oracle_microphysics_step_f64(const oc_grid* gp, const oc_config* cp, const oc_state* st, double dt,
                                 double* lengths, int cap)
{
    const Grid g = to_grid(gp);
    State<double> s;
    load(g, st, s);
    std::vector<double> h;
    microphysics_step(g, to_config(cp), s, dt, &h);
    store(g, s, st);
    for (int q = 0; q < int(h.size()) && q < cap; ++q) lengths[q] = h[q];
    return int(h.size());
}

int oracle_apply_sponges_f64(const oc_grid* gp, const oc_base* bp, const oc_config* cp, const oc_state* st, double dt)
{
    const Grid g = to_grid(gp);
    const Base b = to_base(bp, g.nz);
    State<double> s;
    load(g, st, s);
    apply_sponges(g, b, to_config(cp), s, dt);
    store(g, s, st);
    return 0;
}

// Smagorinsky closure and tendencies of one state, no commit (FP64).
}  // extern "C"

template <class S>
static int sgs_terms(const oc_grid* gp, const oc_config* cp, const oc_state* st, double* nu, double* kth,
                     double* kq, double* du, double* dv, double* dw, double* dth, double* dqv)
{
    const Grid g = to_grid(gp);
    const Config c = to_config(cp);
    State<S> s;
    load(g, st, s);
    std::vector<S> a, b, d, e, f, h, x, y;
    sgs_closure(g, c, s, a, b, d);
    sgs_momentum_tendencies(g, c, s, a, e, f, h);
    sgs_scalar_tendency(g, c, s, b, s.theta, x);
    sgs_scalar_tendency(g, c, s, d, s.qv, y);
    const std::size_t n = g.cells();
    std::copy(a.begin(), a.end(), nu); std::copy(b.begin(), b.end(), kth); std::copy(d.begin(), d.end(), kq);
    std::copy(e.begin(), e.end(), du); std::copy(f.begin(), f.end(), dv); std::copy(h.begin(), h.end(), dw);
    std::copy(x.begin(), x.end(), dth); std::copy(y.begin(), y.end(), dqv);
    (void)n;
    return 0;
}

extern "C" {

int oracle_sgs_terms_f64(const oc_grid* gp, const oc_config* cp, const oc_state* st, double* nu, double* kth,
                         double* kq, double* du, double* dv, double* dw, double* dth, double* dqv)
{
    return sgs_terms<double>(gp, cp, st, nu, kth, kq, du, dv, dw, dth, dqv);
}

// FP32 storage (the reference's fields): coefficients and tendencies rounded to FP32.
int oracle_sgs_terms_f32(const oc_grid* gp, const oc_config* cp, const oc_state* st, double* nu, double* kth,
                         double* kq, double* du, double* dv, double* dw, double* dth, double* dqv)
{
    return sgs_terms<float>(gp, cp, st, nu, kth, kq, du, dv, dw, dth, dqv);
}

// The split SGS integration over one large step (all SGS substeps).
int oracle_apply_sgs_f64(const oc_grid* gp, const oc_config* cp, const oc_state* st)
{
    const Grid g = to_grid(gp);
    State<double> s;
    load(g, st, s);
    apply_sgs(g, to_config(cp), s);
    store(g, s, st);
    return 0;
}

int oracle_apply_sgs_f32(const oc_grid* gp, const oc_config* cp, const oc_state* st)
{
    const Grid g = to_grid(gp);
    State<float> s;
    load(g, st, s);
    apply_sgs(g, to_config(cp), s);
    store(g, s, st);
    return 0;
}

double oracle_weno5_face(double a, double b, double c, double d, double e) { return weno5_face(a, b, c, d, e); }
double oracle_mc_slope(double a, double b, double c) { return mc_slope(a, b, c); }
double oracle_upwind_fd5(double m3, double m2, double m1, double c, double p1, double p2, double p3, double a,
                         double inv)
{
    return upwind_fd5(m3, m2, m1, c, p1, p2, p3, a, inv);
}
double oracle_upwind_fd3(double m2, double m1, double c, double p1, double p2, double a, double inv)
{
    return upwind_fd3(m2, m1, c, p1, p2, a, inv);
}
// UP5/UP3/UP1 derivative of a float field (reference layout) at (i, j, k) along axis 0/1/2.
double oracle_upwind_point(const float* f, int nx, int ny, int nz, int i, int j, int k, int axis, double a,
                           double inv)
{
    const Grid g{nx, ny, nz, 1.0, 1.0, 1.0};
    if (axis == 0) return upwind_axis([&](int o) { return double(f[g.at(i + o, j, k)]); }, i, nx, a, inv);
    if (axis == 1) return upwind_axis([&](int o) { return double(f[g.at(i, j + o, k)]); }, j, ny, a, inv);
    return upwind_axis([&](int o) { return double(f[g.at(i, j, k + o)]); }, k, nz, a, inv);
}

double oracle_qvs_water(double T, double p) { return qvs_water(T, p); }
double oracle_theta_to_temperature(double th, double p) { return theta_to_temperature(th, p); }
double oracle_eos_pressure(double rho, double theta) { return eos_pressure(rho, theta); }
double oracle_saturation_adjust(double T, double p, double* qv, double* qc) { return saturation_adjust(T, p, *qv, *qc); }
double oracle_kessler_rain_speed(double q, double rho) { return kessler_rain_speed(q, rho); }
int oracle_thomas(int n, const double* a, const double* d, const double* c, const double* r, double* x)
{
    return thomas_solve(n, a, d, c, r, x) ? 0 : -1;
}
void oracle_hevi_column_system(int nz, double dt, double dz, double alpha, const double* rho_bar, const double* C,
                               const double* G, const double* e, const double* w_old, double* A, double* D,
                               double* U, double* R)
{
    hevi_column_system(nz, dt, dz, alpha, rho_bar, C, G, e, w_old, A, D, U, R);
}

// The large step's explicit diffusion pass alone on a state (FP64 storage).
int oracle_apply_diffusion_f64(const oc_grid* gp, const oc_config* cp, const oc_state* st)
{
    const Grid g = to_grid(gp);
    State<double> s;
    s.allocate(g);
    load(g, st, s);
    apply_diffusion(g, to_config(cp), s);
    store(g, s, st);
    return 0;
}

// Explicit diffusion tendency of one field (dz: nz layer thicknesses; kv may be null).
void oracle_diffusion_tendency_f32(int nx, int ny, int nz, double dx, double dy, const double* dz, const float* field,
                                   const float* rho, const float* kv, double kh, double kvc, int loc, int c_grid,
                                   int px, int py, float* tend)
{
    diffusion_tendency<float>(nx, ny, nz, dx, dy, dz, field, rho, kv, kh, kvc, loc, c_grid != 0, px != 0, py != 0,
                              tend);
}
void oracle_diffusion_tendency_f64(int nx, int ny, int nz, double dx, double dy, const double* dz, const double* field,
                                   const double* rho, const double* kv, double kh, double kvc, int loc, int c_grid,
                                   int px, int py, double* tend)
{
    diffusion_tendency<double>(nx, ny, nz, dx, dy, dz, field, rho, kv, kh, kvc, loc, c_grid != 0, px != 0, py != 0,
                               tend);
}

// Thread count of the OpenMP loops for the rest of the process (n >= 1).
void oracle_set_threads(int n) { omp_set_num_threads(n > 0 ? n : 1); }
int oracle_max_threads() { return omp_get_max_threads(); }

}  // extern "C"
