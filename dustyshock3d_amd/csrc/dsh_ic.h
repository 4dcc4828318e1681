// dsh_ic.h - the seeded shock-tube set-up, one particle at a time, shared by the host
// set-up (dsh_tube_ic) and the device generator (k_tube_ic): both call tube_ic_particle on
// the same TubeIC, so the two paths agree bit for bit (integer hashing, then +, -, * and /
// in FP64 without contraction).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/dsh.h"

namespace dsh {

// Everything the per-particle set-up needs for one (tube, species); built on the host
// (cube roots included) by make_tube_ic.
struct TubeIC {
    int n;                 // particles of this species
    int n0;                // of which on the x <= 0 lattice
    int nx[2], ny[2], nz[2];
    double dx[2], dy[2], dz[2];
    double hloc[2], rho[2], p[2];
    double m, e, ts_num_k;  // mass, gas e, K (dust t_stop = rho / K)
    double x_min, x_max, y_width, z_width, jitter, frozen_h;
    uint64_t seed;
    int tube_index, species;
};

struct ICParticle {
    double x, y, z, m, rho, p, e, h, ts;
    int ind;
};

__host__ __device__ inline uint64_t ic_mix64(uint64_t x) {
    x += 0x9E3779B97F4A7C15ull;
    x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
    x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
    return x ^ (x >> 31);
}

// Counter-based stream keyed by (seed, tube, species, id, axis): the set-up of one
// particle does not depend on how many others were generated before it.
__host__ __device__ inline double ic_jitter_unit(uint64_t seed, int tube, int species, int id, int axis) {
    uint64_t z = ic_mix64(seed ^ ((uint64_t)(uint32_t)tube * 0xD1B54A32D192ED03ull));
    z = ic_mix64(z ^ ((uint64_t)(uint32_t)(species * 4 + axis) * 0x8CB92BA72F3D8DD7ull));
    z = ic_mix64(z ^ ((uint64_t)(uint32_t)id * 0xA0761D6478BD642Full));
    const double u = (double)(z >> 11) * (1.0 / 9007199254740992.0);
    return 2.0 * u - 1.0;
}

__host__ __device__ inline double ic_wrap(double y, double w) {
    if (y < 0.0) y = y + w;
    if (y >= w) y = y - w;
    return y;
}

// Particle id of one species: lattice x <= 0 (rows anchored at x = 0 going -x, as
// kernel.cu:1421), then x > 0 (starting at +dx, kernel.cu:1461); k fastest.
__host__ __device__ inline ICParticle tube_ic_particle(const TubeIC& c, int id) {
    const int side = id < c.n0 ? 0 : 1;
    const int local = side == 0 ? id : id - c.n0;
    const int nyz = c.ny[side] * c.nz[side];
    const int i = local / nyz;
    const int j = (local / c.nz[side]) % c.ny[side];
    const int k = local % c.nz[side];
    const double dx = c.dx[side], dy = c.dy[side], dz = c.dz[side];
    double x = side == 0 ? 0.0 - (double)i * dx : dx + (double)i * dx;
    double y = 0.0 + (double)j * dy;
    double z = 0.0 + (double)k * dz;
    x = x + c.jitter * dx * ic_jitter_unit(c.seed, c.tube_index, c.species, id, 0);
    y = y + c.jitter * dy * ic_jitter_unit(c.seed, c.tube_index, c.species, id, 1);
    z = z + c.jitter * dz * ic_jitter_unit(c.seed, c.tube_index, c.species, id, 2);
    y = ic_wrap(y, c.y_width);
    z = ic_wrap(z, c.z_width);
    ICParticle q;
    q.x = x;
    q.y = y;
    q.z = z;
    q.m = c.m;
    q.rho = c.rho[side];
    q.p = c.p[side];
    q.e = c.e;
    q.h = c.hloc[side];
    q.ts = c.species == 0 ? 0.0 : c.rho[side] / c.ts_num_k;
    const double hl = c.hloc[side];
    q.ind = (x < c.x_min + c.frozen_h * hl || x > c.x_max - c.frozen_h * hl) ? 1 : 0;
    return q;
}

// Host: the TubeIC of one species of a spec (defined in dsh_host.cpp).
int make_tube_ic(const dsh_tube_spec& spec, int species, TubeIC& out);

}  // namespace dsh
