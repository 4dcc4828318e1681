// dsh_host.cpp - host-only parts of the C ABI: default parameters, the Init.txt reader
// (kernel.cu:1705-1742) and the seeded shock-tube set-up used by the benchmark
// configurations (the structure of init_Sod_X_Dust, kernel.cu:1391-1693).
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>

#include "../../include/dsh.h"
#include "dsh_ic.h"

namespace {

// Lattice of one tube: x <= 0 at spacing dx (rows anchored at x = 0 going -x, as
// kernel.cu:1421), x > 0 at 2 dx starting at +2dx (kernel.cu:1461); the cross-section
// holds 2 n and n points per axis, so both halves tile the periodic box exactly and
// equal-mass particles give a density ratio of exactly 8.
struct Lattice {
    int nx[2], ny[2], nz[2];
    double dx[2], dy[2], dz[2];
    double m, rho[2];
};

Lattice lattice(const dsh_tube_spec& t) {
    Lattice L{};
    L.dx[0] = t.dx_left;
    L.dx[1] = 2.0 * t.dx_left;
    L.nx[0] = (int)std::floor((0.0 - t.x_min) / L.dx[0] + 0.5);
    L.nx[1] = (int)std::floor((t.x_max - 0.0) / L.dx[1] + 0.5);
    L.ny[1] = std::max(1, (int)std::floor(t.y_width / L.dx[1] + 0.5));
    L.nz[1] = std::max(1, (int)std::floor(t.z_width / L.dx[1] + 0.5));
    L.ny[0] = 2 * L.ny[1];
    L.nz[0] = 2 * L.nz[1];
    for (int s = 0; s < 2; ++s) {
        L.dy[s] = t.y_width / L.ny[s];
        L.dz[s] = t.z_width / L.nz[s];
    }
    L.m = t.rho_left * (L.dx[0] * L.dy[0] * L.dz[0]);
    L.rho[0] = t.rho_left;
    L.rho[1] = L.m / (L.dx[1] * L.dy[1] * L.dz[1]);
    return L;
}

bool spec_ok(const dsh_tube_spec* t) {
    return t && t->dx_left > 0 && t->x_max > 0 && t->x_min < 0 && t->y_width > 0 && t->z_width > 0 &&
           t->rho_left > 0 && t->dust_to_gas > 0 && t->K > 0;
}

}  // namespace

namespace dsh {

int make_tube_ic(const dsh_tube_spec& t, int species, TubeIC& c) {
    if (species < 0 || species > 1) return DSH_ERR_ARG;
    const Lattice L = lattice(t);
    const long long n0 = (long long)L.nx[0] * L.ny[0] * L.nz[0];
    const long long n1 = (long long)L.nx[1] * L.ny[1] * L.nz[1];
    if (n0 + n1 > 2000000000ll) return DSH_ERR_ARG;
    const double ratio = species == 0 ? 1.0 : t.dust_to_gas;
    const double cs = 1.0, gam = 1.4;
    c = TubeIC{};
    c.n = (int)(n0 + n1);
    c.n0 = (int)n0;
    for (int s = 0; s < 2; ++s) {
        c.nx[s] = L.nx[s];
        c.ny[s] = L.ny[s];
        c.nz[s] = L.nz[s];
        c.dx[s] = L.dx[s];
        c.dy[s] = L.dy[s];
        c.dz[s] = L.dz[s];
        c.hloc[s] = 1.2 * std::cbrt(L.dx[s] * L.dy[s] * L.dz[s]);
        c.rho[s] = L.rho[s] * ratio;
        c.p[s] = species == 0 ? cs * cs * c.rho[s] : 0.0;
    }
    c.m = L.m * ratio;
    c.e = species == 0 ? cs * cs / (gam - 1.0) : 0.0;
    c.ts_num_k = t.K;
    c.x_min = t.x_min;
    c.x_max = t.x_max;
    c.y_width = t.y_width;
    c.z_width = t.z_width;
    c.jitter = t.jitter;
    c.frozen_h = t.frozen_h;
    c.seed = t.seed;
    c.tube_index = t.tube_index;
    c.species = species;
    return DSH_OK;
}

}  // namespace dsh

// ---- the reference's own problem: init_Sod_X_Dust (kernel.cu:1391-1575) ----
//
// Gas on two lattices of spacing dlh = (m / rho)^(1/3) = 1 / Particle_on_length, x <= 0
// anchored at 0 going -x and x > 0 starting at +dlh, kept where |x| <= 1, |y|, |z| <= 0.8;
// e = 2.5 left, 2.0 right, rho = 1, p from the equation of state with gamma 1.4 and c = 1.
// The frozen flag is evaluated for every lattice candidate and stored on the last placed
// particle (kernel.cu:1441-1448,1481-1488), with X_cube standing in for Y_cube in the y
// test; before the first placement the reference writes Ind_gas[-1], which is skipped.
// Two dust species copy the gas positions and flags, with mass and density scaled by 0.75
// and 0.5, and t_stop = R / (sqrt(1.4 p / rho) rho) with R = 0.01 and 1.0.
namespace {

struct RefLattice {
    double dlh[2], x0[2], sgn[2];
    int im[2], jm[2], km[2];
    double e[2], p[2], rho, m;
};

RefLattice ref_lattice(const dsh_params& P) {
    RefLattice L{};
    L.m = 1.0 / (P.particles_per_length * P.particles_per_length * P.particles_per_length);
    L.rho = 1.0;
    L.e[0] = 2.5;
    L.e[1] = 2.0;
    for (int s = 0; s < 2; ++s) {
        L.p[s] = P.eos_type == 0 ? L.rho * L.e[s] * (1.4 - 1) : 1.0 * 1.0 * L.rho;
        L.dlh[s] = std::pow(L.m / L.rho, 1 / 3.0);
        const double span = s == 0 ? (0.0 - P.x_min) : (P.x_max - 0.0);
        L.im[s] = (int)(span / L.dlh[s]) + 1;
        L.jm[s] = (int)((P.y_max - P.y_min) / L.dlh[s]) + 1;
        L.km[s] = (int)((P.z_max - P.z_min) / L.dlh[s]) + 1;
    }
    return L;
}

// Visits the lattice candidates in the reference's loop order; place(x, y, z, side) is
// called for placed particles and flag(ind) after every candidate.
template <class Place, class Flag>
void ref_walk(const dsh_params& P, const RefLattice& L, Place&& place, Flag&& flag) {
    const double border = 2.5 * P.h;
    const double X_cube = 1.0, Y_cube = 0.8, Z_cube = 0.8;
    for (int side = 0; side < 2; ++side) {
        const double d = L.dlh[side];
        for (int i = 0; i <= L.im[side]; i++)
            for (int j = 0; j <= L.jm[side]; j++)
                for (int k = 0; k <= L.km[side]; k++) {
                    const double x = side == 0 ? 0.0 - (double)(i * d) : d + (double)(i * d);
                    const double y = P.y_min + (double)(j * d);
                    const double z = P.z_min + (double)(k * d);
                    if ((std::fabs(x) <= X_cube) && (std::fabs(y) <= Y_cube) && (std::fabs(z) <= Z_cube))
                        place(x, y, z, side);
                    flag(((std::fabs(x) <= X_cube - border) && (std::fabs(y) <= X_cube - border) &&
                          (std::fabs(z) <= Z_cube - border))
                             ? 0
                             : 1);
                }
    }
}

}  // namespace

extern "C" {

int dsh_struct_sizes(size_t out[4]) {
    if (!out) return DSH_ERR_ARG;
    out[0] = sizeof(dsh_params);
    out[1] = sizeof(dsh_soa_view);
    out[2] = sizeof(dsh_tube_spec);
    out[3] = sizeof(dsh_server_opts);
    return DSH_OK;
}

int dsh_default_params(dsh_params* p) {
    if (!p) return DSH_ERR_ARG;
    std::memset(p, 0, sizeof(*p));
    // the reference Init.txt (Init.txt:2-18) and constants (data.h:6-7, kernel.cu:1498-1502,1744)
    p->x_min = -2.0;
    p->x_max = 2.0;
    p->y_min = -2.0;
    p->y_max = 2.0;
    p->z_min = -2.0;
    p->z_max = 2.0;
    p->eos_type = 0;
    p->max_particles = 1000000;
    p->particles_per_length = 40;
    p->h = 0.1;
    p->avg_cell_coeff = 0.5;
    p->tau = 5.0e-3;
    p->t_end = 0.2;
    p->t_out = 0.1;
    p->alpha = 1.0;
    p->beta = 2.0;
    p->eps = 0.01;
    p->gamma = 1.4;
    p->cs = 1.0;
    p->h_mode = 0;
    p->eta = 1.2;
    p->n_dust = 2;
    p->drag_mode = 0;
    p->K[0] = p->K[1] = 1.0;
    p->R_dust[0] = 0.01;
    p->R_dust[1] = 1.0;
    p->mass_ratio[0] = 0.75;
    p->mass_ratio[1] = 0.5;
    p->eps_fallback[0] = 0.5;
    p->eps_fallback[1] = 0.33;
    p->tstop_fallback = 1.0;
    p->eps2_uses_count1 = 1;
    p->dt_mode = DSH_DT_FIXED;
    p->cfl = 0.3;
    p->max_steps = 0;
    p->cell_size = 0.0;
    p->cell_slack = 0.05;
    p->avg_factor = 2;
    p->precision = DSH_FP64;
    p->fp64_wall_band = 1;
    p->wall_band_speed = 0.5;
    p->wall_band_horizon = 50;
    return DSH_OK;
}

int dsh_tube_params(const dsh_tube_spec* t, dsh_params* p) {
    if (!spec_ok(t) || !p) return DSH_ERR_ARG;
    dsh_default_params(p);
    p->x_min = t->x_min;
    p->x_max = t->x_max;
    p->y_min = 0.0;
    p->y_max = t->y_width;
    p->z_min = 0.0;
    p->z_max = t->z_width;
    p->eos_type = 1;
    p->max_particles = 0;
    p->particles_per_length = 1.0 / t->dx_left;
    p->h = 1.2 * t->dx_left;
    p->avg_cell_coeff = 0.0;
    p->tau = 0.0;
    p->t_end = t->t_end;
    p->t_out = t->t_end;
    p->periodic_y = 1;
    p->periodic_z = 1;
    p->h_mode = 1;
    p->n_dust = 1;
    p->drag_mode = 1;
    p->K[0] = p->K[1] = t->K;
    p->mass_ratio[0] = p->mass_ratio[1] = t->dust_to_gas;
    p->eps2_uses_count1 = 0;
    p->dt_mode = t->dt_mode;
    p->cfl = t->cfl;
    p->max_steps = t->max_steps;
    // averaging cell = one neighbour cell (2.1 h_min, ~16 particles per species on the dense
    // lattice): close to the ~8 per species of the reference's own run (0.5 h with h = 4
    // spacings), and the width whose shock-tube solution is closest to the isothermal
    // Riemann limits (coupled c_s/sqrt(1+eps), decoupled c_s)
    p->avg_factor = 1;
    p->precision = t->precision;
    return DSH_OK;
}

int dsh_tube_count(const dsh_tube_spec* t, int* n) {
    if (!spec_ok(t) || !n) return DSH_ERR_ARG;
    const Lattice L = lattice(*t);
    const long long c = (long long)L.nx[0] * L.ny[0] * L.nz[0] + (long long)L.nx[1] * L.ny[1] * L.nz[1];
    if (c > 2000000000ll) return DSH_ERR_ARG;
    n[0] = (int)c;
    n[1] = (int)c;
    n[2] = 0;
    return DSH_OK;
}

int dsh_tube_ic(const dsh_tube_spec* t, int species, const dsh_soa_view* v) {
    if (!spec_ok(t) || !v || species < 0 || species > 1) return DSH_ERR_ARG;
    dsh::TubeIC c;
    const int rc = dsh::make_tube_ic(*t, species, c);
    if (rc) return rc;
    for (int id = 0; id < c.n; ++id) {
        const dsh::ICParticle q = dsh::tube_ic_particle(c, id);
        if (v->x) v->x[id] = q.x;
        if (v->y) v->y[id] = q.y;
        if (v->z) v->z[id] = q.z;
        if (v->vx) v->vx[id] = 0.0;
        if (v->vy) v->vy[id] = 0.0;
        if (v->vz) v->vz[id] = 0.0;
        if (v->ax) v->ax[id] = 0.0;
        if (v->ay) v->ay[id] = 0.0;
        if (v->az) v->az[id] = 0.0;
        if (v->m) v->m[id] = q.m;
        if (v->rho) v->rho[id] = q.rho;
        if (v->p) v->p[id] = q.p;
        if (v->e) v->e[id] = q.e;
        if (v->h) v->h[id] = q.h;
        if (v->t_stop) v->t_stop[id] = q.ts;
        if (v->edot) v->edot[id] = 0.0;
        if (v->ind) v->ind[id] = q.ind;
        if (v->nn) v->nn[id] = -1;
        if (v->cell) v->cell[id] = -1;
    }
    return DSH_OK;
}

int dsh_reference_setup(dsh_params* p, int* n_per_species) {
    if (!p || !n_per_species) return DSH_ERR_ARG;
    if (!(p->particles_per_length > 0.0) || !(p->h > 0.0)) return DSH_ERR_ARG;
    // parameter overrides of init_Sod_X_Dust (kernel.cu:1495-1502) and the averaging grid
    // built from them (kernel.cu:1577-1582): the averaging cell is the neighbour cell here
    p->n_dust = 2;
    p->avg_cell_coeff = 0.5;
    p->mass_ratio[0] = 0.75;
    p->mass_ratio[1] = 0.5;
    p->R_dust[0] = 0.01;
    p->R_dust[1] = 1.0;
    p->cell_size = p->avg_cell_coeff * p->h;
    p->avg_factor = 1;
    const RefLattice L = ref_lattice(*p);
    long long n = 0;
    ref_walk(*p, L, [&](double, double, double, int) { ++n; }, [](int) {});
    if (n == 0 || n > 2000000000ll) return DSH_ERR_ARG;
    if (p->max_particles > 0 && n > p->max_particles) return DSH_ERR_ARG;
    for (int s = 0; s < DSH_MAX_SPECIES; ++s) n_per_species[s] = (int)n;
    return DSH_OK;
}

int dsh_reference_ic(const dsh_params* p, int species, const dsh_soa_view* v) {
    if (!p || !v || species < 0 || species > 2) return DSH_ERR_ARG;
    const RefLattice L = ref_lattice(*p);
    const double ratio = species == 0 ? 1.0 : p->mass_ratio[species - 1];
    int last = -1;
    ref_walk(
        *p, L,
        [&](double x, double y, double z, int side) {
            const int q = ++last;
            if (v->x) v->x[q] = x;
            if (v->y) v->y[q] = y;
            if (v->z) v->z[q] = z;
            if (v->vx) v->vx[q] = 0.0;
            if (v->vy) v->vy[q] = 0.0;
            if (v->vz) v->vz[q] = 0.0;
            if (v->ax) v->ax[q] = 0.0;
            if (v->ay) v->ay[q] = 0.0;
            if (v->az) v->az[q] = 0.0;
            if (v->m) v->m[q] = L.m * ratio;
            if (v->rho) v->rho[q] = L.rho * ratio;
            if (v->p) v->p[q] = species == 0 ? L.p[side] : 0.0;
            if (v->e) v->e[q] = species == 0 ? L.e[side] : 0.0;
            if (v->h) v->h[q] = p->h;
            if (v->t_stop) {
                const double R = species == 0 ? 0.0 : p->R_dust[species - 1];
                v->t_stop[q] = species == 0 ? 0.0 : R / (std::sqrt(p->gamma * L.p[side] / L.rho) * L.rho);
            }
            if (v->edot) v->edot[q] = 0.0;
            if (v->nn) v->nn[q] = -1;
            if (v->cell) v->cell[q] = -1;
        },
        [&](int ind) {
            if (last >= 0 && v->ind) v->ind[last] = ind;
        });
    return DSH_OK;
}

// Init.txt: a comment line, then one value per line (kernel.cu:1705-1722).  Lines after
// the 18th, when present, are extension values in the order of dsh_params.
int dsh_read_init(const char* path, dsh_params* p) {
    if (!path || !p) return DSH_ERR_ARG;
    FILE* f = std::fopen(path, "r");
    if (!f) return DSH_ERR_IO;
    dsh_default_params(p);
    char s[256];
    int line = 0, ok = 1;
    double* dbl1[] = {&p->x_min, &p->x_max, &p->y_min, &p->y_max, &p->z_min, &p->z_max};
    if (!std::fgets(s, sizeof s, f)) ok = 0;
    for (int i = 0; ok && i < 6; ++i) ok = std::fgets(s, sizeof s, f) && std::sscanf(s, "%lf", dbl1[i]) == 1;
    if (ok) ok = std::fgets(s, sizeof s, f) && std::sscanf(s, "%d", &p->eos_type) == 1;
    if (ok) ok = std::fgets(s, sizeof s, f) && std::sscanf(s, "%d", &p->max_particles) == 1;
    double* dbl2[] = {&p->particles_per_length, &p->h, &p->avg_cell_coeff, &p->tau, &p->t_end,
                      &p->t_out, &p->alpha, &p->beta, &p->eps};
    for (int i = 0; ok && i < 9; ++i) ok = std::fgets(s, sizeof s, f) && std::sscanf(s, "%lf", dbl2[i]) == 1;
    line = 18;
    if (!ok) {
        std::fclose(f);
        return DSH_ERR_IO;
    }
    // optional extension lines
    struct Ext {
        char kind;
        void* dst;
    } ext[] = {{'i', &p->periodic_y},     {'i', &p->periodic_z},  {'d', &p->gamma},        {'d', &p->cs},
               {'i', &p->h_mode},         {'d', &p->eta},         {'i', &p->n_dust},       {'i', &p->drag_mode},
               {'d', &p->K[0]},           {'d', &p->K[1]},        {'d', &p->mass_ratio[0]}, {'d', &p->mass_ratio[1]},
               {'d', &p->eps_fallback[0]}, {'d', &p->eps_fallback[1]}, {'d', &p->tstop_fallback},
               {'i', &p->eps2_uses_count1}, {'i', &p->dt_mode},   {'d', &p->cfl},          {'i', &p->max_steps},
               {'d', &p->cell_size},      {'d', &p->cell_slack},  {'i', &p->avg_factor},   {'i', &p->precision},
               {'i', &p->fp64_wall_band}, {'d', &p->wall_band_speed}, {'i', &p->wall_band_horizon}};
    for (auto& e : ext) {
        if (!std::fgets(s, sizeof s, f)) break;
        const int r = e.kind == 'i' ? std::sscanf(s, "%d", (int*)e.dst) : std::sscanf(s, "%lf", (double*)e.dst);
        if (r != 1) break;
        ++line;
    }
    std::fclose(f);
    (void)line;
    return DSH_OK;
}

}  // extern "C"
