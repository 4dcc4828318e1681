// dsh_types.h - plain structs shared by the host engine and the gfx950 kernels.
//
// A batch holds T independent tubes.  Particles of every (species, tube) pair live
// in one global structure-of-arrays space, species-major then tube-major, so that a
// single counting sort over one global cell-key space sorts every tube and species at
// once: segment (s, t) owns particles [p_off, p_off + n) and neighbour cells
// [cell_off, cell_off + ncell), and the exclusive scan of cell counts reproduces p_off
// at each segment boundary.
#pragma once

#include <stdint.h>

namespace dsh {

constexpr int kMaxSpecies = 3;
constexpr int kMaxDust = 2;
constexpr int kMaxCellReach = 64;  // largest search radius, in cells along one axis
constexpr double kRadiusMargin = 1e-9;

// Error bits of a tube (TubeDyn::err).
constexpr int kErrGrid = 1;        // smoothing length needs a search radius above kMaxCellReach
constexpr int kErrNonFinite = 2;   // NaN/Inf in the conservation sums

struct TubeConst {
    // neighbour grid: cell (kx,ky,kz), key = kx + n0 (ky + n1 kz)
    double o[3], c[3], ic[3], L[3], hi[3];
    double cmin;
    int n[3];
    int per[3];
    int f;             // averaging cell = f^3 neighbour cells
    int na[3];         // averaging cells per axis
    int ncell;         // n0 n1 n2
    int navg;
    int avg_off;       // this tube's averaging cells (and conservation partials) in the batch arrays
    int npart;         // conservation partials: navg, or one per run of cells when f == 1
    int nsp;           // 1 + n_dust
    int cell_off[kMaxSpecies];
    int p_off[kMaxSpecies];
    int np[kMaxSpecies];
    int pseg[kMaxSpecies];   // index of segment (species, tube) in the particle map, -1 if none
    // physics
    double gamma, cs, alpha, beta, eps, eta;
    double K[kMaxDust], mu[kMaxDust], eps_fb[kMaxDust], t_fb;
    double R[kMaxDust];    // drag_mode 0: t_stop = R / (sqrt(gamma p_g / rho_g) rho_g)
    int drag_mode;         // 0: reference law with the gas particle of the same index, 1: rho_d / K
    double tau, cfl, t_end;
    int eos, h_mode, eps2_quirk, dt_mode, max_steps, precision;
    int wall_band;         // mixed precision: FP64 band next to the frozen walls (TubeDyn::band_lo/hi)
    double band_speed;     // band edge speed, units of cs
    int band_horizon;      // steps the band's time term covers at most (0: the whole run)
};

struct TubeDyn {
    double t, dt;
    int steps, done, final_step, err;
    int done_cur;                  // buffer holding a finished tube's final state (not re-sorted)
    int pad_;
    unsigned long long hmax_bits[kMaxSpecies];  // max h per species this step (positive doubles)
    double cons[4];                // momentum (3), energy after the last step
    // frozen particles lie at x <= xw_lo or x >= xw_hi (x-end walls; +-1e300 if a side has none)
    double xw_lo, xw_hi;
    // this step's FP64 band: a 64-particle chunk with a particle at x < band_lo or x > band_hi
    double band_lo, band_hi;
};

// Double-buffered particle fields: buffer b of a field starts at base + b * ntot.
struct PBuf {
    double *x, *y, *z, *vx, *vy, *vz, *m, *h, *e, *rho, *p, *ts;
    double *ax, *ay, *az, *ed;
    int *ind, *id;
    long long ntot;
};

// A contiguous run of items (particles, cells or averaging cells) that belong to one
// (species, tube) pair; blk is where its blocks start in the launch's block map.
struct Seg {
    int off;       // first item
    int n;         // item count
    int tube;
    int species;
};

}  // namespace dsh
