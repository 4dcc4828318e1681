// dsh_kernels.h - host-side launchers of the gfx950 kernels (dsh_kernels.hip).
#pragma once

#include <hip/hip_runtime.h>

#include "dsh_ic.h"
#include "dsh_types.h"

namespace dsh {

// Block map of one launch: segment s owns blocks [blk[s], blk[s+1]).
struct Map {
    const Seg* seg;
    const int* blk;
    int nseg;
    int nblocks;
};

// Packed record of one gas particle in sorted order for the FP32 force pass, written by the
// density pass of a mixed-precision tube: the force pass stages a candidate with four 16-byte
// loads instead of gathering and converting eleven FP64 fields.
struct alignas(16) ForceRec {
    double x;
    int y, z;          // fixed point: units of L / 2^30 from the box origin
    float h, vx, vy, vz, m, rho, pr2, cs;
    unsigned flags;    // bit 0 frozen, bit 1 in an FP64-band chunk, bit 2 y, z inside the fixed-point range
    unsigned pad[3];
};
static_assert(sizeof(ForceRec) == 64, "four 16-byte loads");

// pair-kernel variants: one thread per target with global-memory candidates, or one
// wave per 64-target chunk with LDS-staged candidate rows
constexpr int kPairDirect = 0;
constexpr int kPairTiled = 1;
// Candidates staged per LDS piece by the tiled pair kernels (template instances).
constexpr int kTiles[] = {128, 160, 192, 224, 256};
inline bool valid_tile(int t) {
    for (int k : kTiles)
        if (k == t) return true;
    return false;
}
// Waves (64-target chunks, each with its own LDS tile) per block of the tiled pair kernels.
constexpr int kBlockWaves[] = {1, 2, 4};
inline bool valid_block_waves(int w) {
    for (int k : kBlockWaves)
        if (k == w) return true;
    return false;
}

struct StepArgs {
    const TubeConst* tc;
    TubeDyn* td;
    int* cur;              // device: [0] buffer index of the current state, [1] finalize's block ticket
    PBuf B;
    int ntubes;
    // per particle (global sorted space)
    int *key, *slot, *tsrc, *tkey, *wkey;
    double *pr2, *csav;
    int* ncount;
    unsigned long long* nhash;
    // per neighbour cell
    int *count, *start;
    double* cellh;
    double* reach;
    double* wtmp;          // 2 x ncells: separable window maxima of cell h
    int ncells;
    int* scan_bsum;
    int scan_nb;
    // per averaging cell
    double* part;          // 5 per averaging cell
    // launch maps
    Map pmap, gmap, cmap, amap;
    Map pmap64, gmap64;    // 64 targets per block (tiled pair kernels)
    Map amap_runs;         // idic_block_cells() averaging cells per block, idic_run_cells() per wave (f == 1)
    unsigned idic_blocks;          // bit nd: a tube with averaging cells of f^3 > 1 neighbour cells, nd dust species
    unsigned idic_runs;            // bit nd: a tube with f == 1 and nd dust species
    unsigned long long* masks;   // gas density hit masks reused by the force pass
    ForceRec* frec;        // per gas particle (sorted), mixed precision; null otherwise
    int* ginv;             // gas slot of each gas index (tubes with drag_mode 0), or null
    int variant;
    int tile;              // candidates staged per piece in the tiled pair kernels (kTiles)
    int waves;             // waves per block of the tiled pair kernels (kBlockWaves)
    // mixed precision with an FP64 wall band: per sorted particle, 1 if its 64-particle chunk
    // has a particle in its tube's band (written by k_band each step); null otherwise
    unsigned char* pband;
    // with pband or several tubes (else null): the chunk order of the pair kernels
    // (pmap64.nblocks entries; then the band count, gas band count, count of the other running
    // chunks and of their gas chunks), then per-chunk classes (pmap64.nblocks) and two counts
    // per block of the ordering kernels
    int* bperm;
};
// ints of StepArgs::bperm for nall pmap64 chunks
inline size_t bperm_ints(int nall) { return 2 * (size_t)nall + 4 + 2 * ((size_t)(nall + 1023) / 1024) + 1; }

// scratch: kPrepParts doubles per particle block of pmap
constexpr int kPrepParts = 7;
hipError_t launch_prepare(const StepArgs& a, double* scratch, hipStream_t s);
hipError_t launch_key(const StepArgs& a, hipStream_t s);
hipError_t launch_scan(const StepArgs& a, hipStream_t s);
hipError_t launch_scatter(const StepArgs& a, hipStream_t s);
hipError_t launch_rank_permute(const StepArgs& a, hipStream_t s);
hipError_t launch_radius(const StepArgs& a, hipStream_t s);
hipError_t launch_density(const StepArgs& a, int precision, bool stats, hipStream_t s);
hipError_t launch_force(const StepArgs& a, int precision, hipStream_t s);
hipError_t launch_idic(const StepArgs& a, hipStream_t s);
int idic_run_cells();     // averaging cells per wave of the f == 1 drag kernel
int idic_block_cells();   // averaging cells per block of it (one conservation partial each)
hipError_t launch_finalize(const StepArgs& a, hipStream_t s);
// seeded shock-tube set-up of one (tube, species) straight into buffer slots off..off+n-1
hipError_t launch_tube_ic(const TubeIC& c, const PBuf& B, long long off, hipStream_t s);

int scan_blocks(int n);
size_t mask_words(int gas_chunks);

}  // namespace dsh
