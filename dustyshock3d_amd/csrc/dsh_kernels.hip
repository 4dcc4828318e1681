// dsh_kernels.hip - gfx950 kernels of one dusty-gas SPH step over a batch of tubes.
//
// Step (one stream, no host round trips):
//   key          neighbour-cell key per particle; wave-aggregated atomic slot in its cell
//   scan         exclusive scan of cell counts -> cell starts (global over all segments);
//                clears the counts for the next step
//   scatter      particle index and key -> (cell start + slot)
//   rank_permute rank inside the cell by (x, id) (gathered from the state), gather the state
//                into the work buffer, max h per cell and per (species, tube); the sorted
//                order is (cell key, x, id), and the reference's linked lists (cell, id) are
//                rebuilt at download
//   radius       per-cell search radius in cells, from a window max of cell h (clears it)
//   density      same-species density sums; pressure, sound speed, stopping time
//   force        gas pressure + artificial-viscosity acceleration and de/dt
//   idic         per averaging cell: sums, implicit two-fluid drag solve (FP64), per-particle
//                velocity update, drift, energy, smoothing length, conservation partials
//   finalize     per tube: conservation sums, time, next CFL step, completion; its last
//                block swaps the state and work buffers
#include "dsh_kernels.h"

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <algorithm>
#include <type_traits>

namespace dsh {
namespace {

constexpr int kBS = 256;
constexpr int kScanItems = 8;
constexpr int kScanTile = kBS * kScanItems;
constexpr double kInvPi = 1.0 / 3.14159265;   // kernel.cu:13-15

__device__ __forceinline__ int find_seg(const Map& mp, int b) {
    int lo = 0, hi = mp.nseg - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (mp.blk[mid] <= b) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

__device__ __forceinline__ int floordiv(int a, int n) { return (a >= 0) ? a / n : -((-a + n - 1) / n); }

// Cell index along one axis: int((x - o) / c) as kernel.cu:1949 and IDIC.h:28 compute it
// (FP64 divide, truncation), clamped into [0, n).
__device__ __forceinline__ int axis_cell(double x, double o, double c, int n) {
    const double t = (x - o) / c;
    if (!(t >= 0.0)) return 0;
    if (t >= (double)n) return n - 1;
    return (int)t;
}

__device__ __forceinline__ double wrap(double y, double lo, double L, double hi) {
    if (y < lo) y = y + L;
    if (y >= hi) y = y - L;
    return y;
}

__device__ __forceinline__ uint64_t mix64(uint64_t x) {
    x += 0x9E3779B97F4A7C15ull;
    x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
    x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
    return x ^ (x >> 31);
}

// Order-free fingerprint term of one neighbour: (id, image) packs injectively into 64 bits
// (image offsets are at most kMaxCellReach = 64 < 128 cells away, so sy + 128 and sz + 128 fit
// 8 bits each) and mix64 is a bijection, so distinct neighbours give distinct terms.
__device__ __forceinline__ uint64_t pair_tag(int id, int sy, int sz) {
    return mix64(((uint64_t)(uint32_t)id << 16) | ((uint64_t)(uint32_t)(sy + 128) << 8) | (uint64_t)(uint32_t)(sz + 128));
}

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

__device__ __forceinline__ int wave_sum_i(int v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

__device__ __forceinline__ double wave_min(double v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v = fmin(v, __shfl_xor(v, o, 64));
    return v;
}

__device__ __forceinline__ double wave_max(double v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v = fmax(v, __shfl_xor(v, o, 64));
    return v;
}

// ---------------------------------------------------------------- real-type helpers
// The pair kernels take reciprocals and square roots of positive, normal FP64 values only
// (smoothing lengths, densities, squared distances of distinct particles; a zero distance is
// handled by the caller).  For those the hardware approximations refined by Newton steps are
// within an ulp of the correctly rounded result, without the scaling and special-case
// fix-ups of the IEEE sequences (about a third of their instructions).
// 1/x: v_rcp_f64, then two Newton steps y += y (1 - x y).
template <typename R> __device__ __forceinline__ R rcp(R x);
template <> __device__ __forceinline__ double rcp<double>(double x) {
    double y = __builtin_amdgcn_rcp(x);
    y = __builtin_fma(y, __builtin_fma(-x, y, 1.0), y);
    y = __builtin_fma(y, __builtin_fma(-x, y, 1.0), y);
    return y;
}
template <> __device__ __forceinline__ float rcp<float>(float x) { return __builtin_amdgcn_rcpf(x); }

// sqrt(x): Goldschmidt iteration from v_rsq_f64 (g -> sqrt(x), h -> 1/(2 sqrt(x))), then one
// residual correction g += (x - g^2) h; 0 for x = 0 (the density's self pair).
template <typename R> __device__ __forceinline__ R sqrt_r(R x);
template <> __device__ __forceinline__ double sqrt_r<double>(double x) {
    const double y = __builtin_amdgcn_rsq(x);
    double g = x * y, h = 0.5 * y;
    const double r = __builtin_fma(-h, g, 0.5);
    g = __builtin_fma(g, r, g);
    h = __builtin_fma(h, r, h);
    g = __builtin_fma(__builtin_fma(-g, g, x), h, g);
    return x > 0.0 ? g : 0.0;
}
// v_sqrt_f32 (1 ulp); arguments are squared distances, never denormal-small in practice.
template <> __device__ __forceinline__ float sqrt_r<float>(float x) { return __builtin_amdgcn_sqrtf(x); }

// Neighbour sums accumulate in FP64 in both precisions.  An FP32 pair term is exact in FP64
// and the FP64 sum of a particle's terms rounds far below FP32 resolution, so the exactly
// antisymmetric FP32 force terms of a pair still cancel in the total momentum; a
// Kahan-compensated FP32 sum does not (its per-particle result is rounded once more, and
// those roundings do not cancel between the two particles of a pair).  It also costs fewer
// VALU slots than Kahan's four FP32 operations: one conversion and one FP64 add, both full
// rate on CDNA4 without packed FP32.
template <typename R> struct Acc {
    double s = 0.0;
    __device__ __forceinline__ void add(R v) { s += (double)v; }
    __device__ __forceinline__ double get() const { return s; }
};

// Cubic spline W(r,h) (kernel.cu:11-22) and dW/dr for r >= 0 (kernel.cu:24-38),
// written with one reciprocal of h per pair.
// Both branches are evaluated and one is selected, so a wave with mixed q runs one path.
template <typename R> __device__ __forceinline__ R spline_w(R q) {
    const R inner = R(1) - R(1.5) * q * q + R(0.75) * q * q * q;
    const R u = R(2) - q;
    const R outer = R(0.25) * u * u * u;
    return q < R(1) ? inner : (q <= R(2) ? outer : R(0));
}
template <typename R> __device__ __forceinline__ R spline_dw(R q) {
    const R inner = R(-3) * q + R(2.25) * q * q;
    const R u = R(2) - q;
    const R outer = R(-0.75) * u * u;
    return q <= R(1) ? inner : (q <= R(2) ? outer : R(0));
}

// W polynomial for a neighbour (q = r/h <= 2 up to the last rounding of r): Horner form
// with fused multiply-adds; the q > 2 zero of spline_w is not needed.
__device__ __forceinline__ float fma_r(float a, float b, float c) { return __builtin_fmaf(a, b, c); }
__device__ __forceinline__ double fma_r(double a, double b, double c) { return __builtin_fma(a, b, c); }
template <typename R> __device__ __forceinline__ R spline_w_hit(R q) {
    const R inner = fma_r(q * q, fma_r(R(0.75), q, R(-1.5)), R(1));
    const R u = R(2) - q;
    const R outer = R(0.25) * (u * u) * u;
    return q < R(1) ? inner : outer;
}

// dW polynomial for a neighbour (q <= 2 up to the last rounding), Horner form, no branch.
template <typename R> __device__ __forceinline__ R spline_dw_hit(R q) {
    const R inner = q * fma_r(R(2.25), q, R(-3));
    const R u = R(2) - q;
    const R outer = R(-0.75) * (u * u);
    return q <= R(1) ? inner : outer;
}

// dW of a signed coordinate difference, without the 1/(pi h^4) factor (kernel.cu:24-38: odd in u).
template <typename R> __device__ __forceinline__ R spline_dw_signed(R u) {
    return u < R(0) ? -spline_dw<R>(-u) : spline_dw<R>(u);
}

// de/dt weight of the plain ForceKernel (kernel.cu:182-187): each velocity component times
// dW of the same coordinate difference x_i - x_j = -d (gas-only runs).
template <typename R>
__device__ __forceinline__ R plain_edot_weight(R dvx, R dvy, R dvz, R fx, R fy, R fz, R ih) {
    const R c4 = (R)kInvPi * (ih * ih * ih * ih);
    return (dvx * (c4 * spline_dw_signed<R>(-fx * ih)) + dvy * (c4 * spline_dw_signed<R>(-fy * ih))) +
           dvz * (c4 * spline_dw_signed<R>(-fz * ih));
}

// ---------------------------------------------------------------- neighbour traversal
// Visits every candidate j within ceil(reach / c_a) cells of (kx,ky,kz) along each axis a,
// periodic in y/z (each offset is a distinct (cell, image) pair, so thin tubes get every
// image in reach), clamped in x, and calls f.pair for
// d2 < (h_i + h_j)^2 with d2 = fma(dz,dz,fma(dy,dy,dx*dx)) in FP64.  Rows of cells along
// x are contiguous in the sorted arrays, so each (dy,dz) offset is one index range.
template <class F>
__device__ __forceinline__ void for_pairs(const TubeConst& T, int cell_off, const int* __restrict__ start,
                                          int kx, int ky, int kz, double reach, const double* __restrict__ X,
                                          const double* __restrict__ Y, const double* __restrict__ Z,
                                          const double* __restrict__ H, double xi, double yi, double zi,
                                          double hi, F& f) {
    const int nx = T.n[0], ny = T.n[1], nz = T.n[2];
    const int Rx = (int)ceil(reach * T.ic[0]);
    const int Ry = (int)ceil(reach * T.ic[1]);
    const int Rz = (int)ceil(reach * T.ic[2]);
    const int x0 = max(kx - Rx, 0), x1 = min(kx + Rx, nx - 1);
    for (int oz = -Rz; oz <= Rz; ++oz) {
        int cz = kz + oz, sz = 0;
        if (T.per[2]) {
            sz = floordiv(cz, nz);
            cz -= sz * nz;
        } else if (cz < 0 || cz >= nz) {
            continue;
        }
        const double shz = (double)sz * T.L[2];
        for (int oy = -Ry; oy <= Ry; ++oy) {
            int cy = ky + oy, sy = 0;
            if (T.per[1]) {
                sy = floordiv(cy, ny);
                cy -= sy * ny;
            } else if (cy < 0 || cy >= ny) {
                continue;
            }
            const double shy = (double)sy * T.L[1];
            const int row = cell_off + nx * (cy + ny * cz);
            const int jb = start[row + x0], je = start[row + x1 + 1];
            for (int j = jb; j < je; ++j) {
                const double dx = X[j] - xi;
                const double dy = (Y[j] - yi) + shy;
                const double dz = (Z[j] - zi) + shz;
                const double d2 = __builtin_fma(dz, dz, __builtin_fma(dy, dy, dx * dx));
                const double hs = hi + H[j];
                if (d2 < hs * hs) f.pair(j, dx, dy, dz, d2, sy, sz);
            }
        }
    }
}

struct ItemPos {
    Seg sg;
    int li;
};

__device__ __forceinline__ ItemPos locate(const Map& mp, int per_block) {
    const int s = find_seg(mp, blockIdx.x);
    ItemPos ip;
    ip.sg = mp.seg[s];
    ip.li = (blockIdx.x - mp.blk[s]) * per_block;
    return ip;
}

// ---------------------------------------------------------------- kernels

// Initial time step and conservation sums from the uploaded state (buffer 0): per-block
// partials over every particle, then one block per tube reduces its blocks in order.
__global__ __launch_bounds__(kBS) void k_prepare_part(Map mp, const TubeConst* __restrict__ tc, PBuf B,
                                                      double* __restrict__ part) {
    __shared__ double sh[kPrepParts][kBS / 64];
    const ItemPos ip = locate(mp, kBS);
    const int li = ip.li + threadIdx.x;
    const TubeConst& T = tc[ip.sg.tube];
    double mn = 1e300, px = 0, py = 0, pz = 0, E = 0, flo = -1e300, fhi = 1e300;
    if (li < ip.sg.n) {
        const int p = ip.sg.off + li;
        const double vx = B.vx[p], vy = B.vy[p], vz = B.vz[p], m = B.m[p];
        px = m * vx;
        py = m * vy;
        pz = m * vz;
        E = 0.5 * m * (vx * vx + vy * vy + vz * vz);
        if (ip.sg.species == 0) E += m * B.e[p];
        if (B.ind[p] == 0) {
            mn = B.h[p] / (T.cs + sqrt(vx * vx + vy * vy + vz * vz));
        } else {
            // frozen wall particle: extent of the wall on its side of the tube's middle
            const double x = B.x[p];
            if (x < T.o[0] + 0.5 * T.L[0]) flo = x;
            else fhi = x;
        }
    }
    px = wave_sum(px);
    py = wave_sum(py);
    pz = wave_sum(pz);
    E = wave_sum(E);
    mn = wave_min(mn);
    flo = wave_max(flo);
    fhi = wave_min(fhi);
    const int w = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0) {
        sh[0][w] = px;
        sh[1][w] = py;
        sh[2][w] = pz;
        sh[3][w] = E;
        sh[4][w] = mn;
        sh[5][w] = flo;
        sh[6][w] = fhi;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        double a[4] = {0, 0, 0, 0}, m2 = 1e300, lo = -1e300, hi = 1e300;
        for (int k = 0; k < kBS / 64; ++k) {
            for (int c = 0; c < 4; ++c) a[c] += sh[c][k];
            m2 = fmin(m2, sh[4][k]);
            lo = fmax(lo, sh[5][k]);
            hi = fmin(hi, sh[6][k]);
        }
        double* o = part + (long long)kPrepParts * blockIdx.x;
        for (int c = 0; c < 4; ++c) o[c] = a[c];
        o[4] = m2;
        o[5] = lo;
        o[6] = hi;
    }
}

// FP64 band of the coming step of a mixed-precision tube.  Momentum enters a tube only
// through its frozen walls, so the particles next to them, within f * band_speed * cs * (time
// left) + (3 + f) neighbour cells of the frozen layers (f: neighbour cells per averaging cell
// along an axis), run in FP64: FP32 rounding made outside the band reaches the walls late or
// not at all before the run ends.  The time left counts at most band_horizon steps.  An
// open-ended run keeps the fixed part; no band in FP64 mode or when switched off.
__device__ __forceinline__ void set_band(const TubeConst& T, TubeDyn& D) {
    D.band_lo = -1e300;
    D.band_hi = 1e300;
    if (!T.wall_band || T.precision != DSH_MIXED) return;
    double left = 1e300;
    if (T.t_end < 1e29) left = T.t_end - D.t;
    // a CFL step may grow: twice the current step per step left
    const double step_bound = D.dt * (T.dt_mode == DSH_DT_CFL ? 2.0 : 1.0);
    if (T.max_steps > 0) left = fmin(left, (double)(T.max_steps - D.steps) * step_bound);
    if (!(left < 1e29)) left = 0.0;   // an open-ended run: the fixed part alone
    // at most band_horizon steps ahead: later, averaging-cell membership differences decide
    if (T.band_horizon > 0) left = fmin(left, (double)T.band_horizon * step_bound);
    // an averaging cell of f neighbour cells passes a particle's influence f cells on in one
    // step, so both terms scale with f
    const double w = T.band_speed * (double)T.f * T.cs * fmax(left, 0.0) + (double)(3 + T.f) * T.c[0];
    D.band_lo = D.xw_lo + w;
    D.band_hi = D.xw_hi - w;
}

__global__ __launch_bounds__(kBS) void k_prepare(Map mp, const TubeConst* __restrict__ tc, TubeDyn* __restrict__ td,
                                                 const double* __restrict__ part) {
    __shared__ double sh[kPrepParts][kBS / 64];
    const int t = blockIdx.x;
    const TubeConst& T = tc[t];
    double mn = 1e300, a4[4] = {0, 0, 0, 0}, flo = -1e300, fhi = 1e300;
    for (int s = 0; s < T.nsp; ++s) {
        const int sg = T.pseg[s];
        for (int b = mp.blk[sg] + threadIdx.x; b < mp.blk[sg + 1]; b += kBS) {
            const double* o = part + (long long)kPrepParts * b;
            for (int c = 0; c < 4; ++c) a4[c] += o[c];
            mn = fmin(mn, o[4]);
            flo = fmax(flo, o[5]);
            fhi = fmin(fhi, o[6]);
        }
    }
    for (int c = 0; c < 4; ++c) a4[c] = wave_sum(a4[c]);
    mn = wave_min(mn);
    flo = wave_max(flo);
    fhi = wave_min(fhi);
    const int w = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0) {
        for (int c = 0; c < 4; ++c) sh[c][w] = a4[c];
        sh[4][w] = mn;
        sh[5][w] = flo;
        sh[6][w] = fhi;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        double a[4] = {0, 0, 0, 0}, m2 = 1e300, lo = -1e300, hi = 1e300;
        for (int k = 0; k < kBS / 64; ++k) {
            for (int c = 0; c < 4; ++c) a[c] += sh[c][k];
            m2 = fmin(m2, sh[4][k]);
            lo = fmax(lo, sh[5][k]);
            hi = fmin(hi, sh[6][k]);
        }
        TubeDyn& D = td[t];
        D.xw_lo = lo;
        D.xw_hi = hi;
        D.t = 0.0;
        D.dt = (T.dt_mode == 0) ? T.tau : T.cfl * m2;
        D.steps = 0;
        D.done = 0;
        D.final_step = 0;
        D.err = 0;
        D.done_cur = 0;
        for (int c = 0; c < 4; ++c) D.cons[c] = a[c];
        for (int s = 0; s < kMaxSpecies; ++s) D.hmax_bits[s] = 0ull;
        set_band(T, D);
    }
}

// Cell key and slot.  Consecutive lanes hold consecutive particles of the previous
// step's sorted order, so equal keys come in runs: one atomic per run, found with a
// wave ballot of run heads.  A finished (or failed) tube is not re-sorted: its particles
// are counted in its first cell of the species, which keeps every later tube's cell starts
// (one prefix sum over all cells) in place, and get no key or slot (k_scatter and
// k_rank_permute skip the tube).
__global__ __launch_bounds__(kBS) void k_key(Map mp, const TubeConst* __restrict__ tc,
                                             const TubeDyn* __restrict__ td, const int* __restrict__ cur,
                                             PBuf B, int* __restrict__ key, int* __restrict__ slot,
                                             int* __restrict__ count) {
    const ItemPos ip = locate(mp, kBS);
    const int li = ip.li + threadIdx.x;
    const bool valid = li < ip.sg.n;
    const int p = ip.sg.off + li;
    const TubeConst& T = tc[ip.sg.tube];
    if (td[ip.sg.tube].done || td[ip.sg.tube].err) {
        if (threadIdx.x == 0) atomicAdd(&count[T.cell_off[ip.sg.species]], min(kBS, ip.sg.n - ip.li));
        return;
    }
    const long long so = (long long)(*cur) * B.ntot;
    int k = -1;
    if (valid) {
        const int kx = axis_cell(B.x[so + p], T.o[0], T.c[0], T.n[0]);
        const int ky = axis_cell(B.y[so + p], T.o[1], T.c[1], T.n[1]);
        const int kz = axis_cell(B.z[so + p], T.o[2], T.c[2], T.n[2]);
        k = T.cell_off[ip.sg.species] + kx + T.n[0] * (ky + T.n[1] * kz);
    }
    const int lane = threadIdx.x & 63;
    const int prev = __shfl_up(k, 1, 64);
    const bool head = valid && (lane == 0 || prev != k);
    const unsigned long long heads = __ballot(head);
    const unsigned long long vmask = __ballot(valid);
    const int nvalid = __popcll(vmask);
    const unsigned long long upto = (lane == 63) ? heads : (heads & ((2ull << lane) - 1ull));
    const int h0 = upto ? (63 - __clzll(upto)) : 0;
    const unsigned long long after = (lane == 63) ? 0ull : (heads & ~((2ull << lane) - 1ull));
    int next = after ? (__ffsll((long long)after) - 1) : 64;
    if (next > nvalid) next = nvalid;
    int base = 0;
    if (head) base = atomicAdd(&count[k], next - lane);
    base = __shfl(base, h0, 64);
    if (valid) {
        key[p] = k;
        slot[p] = base + (lane - h0);
    }
}

// ---- exclusive scan of cell counts (three passes) ----
__device__ __forceinline__ int block_exclusive_scan(int v, int* total) {
    __shared__ int ws[kBS / 64];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    int x = v;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const int y = __shfl_up(x, o, 64);
        if (lane >= o) x += y;
    }
    if (lane == 63) ws[w] = x;
    __syncthreads();
    int wo = 0, tot = 0;
    for (int k = 0; k < kBS / 64; ++k) {
        if (k < w) wo += ws[k];
        tot += ws[k];
    }
    __syncthreads();
    *total = tot;
    return wo + x - v;
}

__global__ __launch_bounds__(kBS) void k_scan_reduce(const int* __restrict__ in, int n, int* __restrict__ bsum) {
    const int base = blockIdx.x * kScanTile + threadIdx.x * kScanItems;
    int s = 0;
#pragma unroll
    for (int k = 0; k < kScanItems; ++k)
        if (base + k < n) s += in[base + k];
    int tot;
    block_exclusive_scan(s, &tot);
    if (threadIdx.x == 0) bsum[blockIdx.x] = tot;
}

__global__ __launch_bounds__(kBS) void k_scan_top(int* __restrict__ bsum, int nb) {
    int carry = 0;
    for (int base = 0; base < nb; base += kBS) {
        const int i = base + threadIdx.x;
        const int v = (i < nb) ? bsum[i] : 0;
        int tot;
        const int ex = block_exclusive_scan(v, &tot);
        if (i < nb) bsum[i] = carry + ex;
        carry += tot;
    }
}

// The counts are consumed here: each is cleared after it is read, ready for the next step's
// k_key.
__global__ __launch_bounds__(kBS) void k_scan_down(int* __restrict__ in, int n, const int* __restrict__ bsum,
                                                   int* __restrict__ out) {
    const int base = blockIdx.x * kScanTile + threadIdx.x * kScanItems;
    int v[kScanItems];
    int s = 0;
#pragma unroll
    for (int k = 0; k < kScanItems; ++k) {
        v[k] = (base + k < n) ? in[base + k] : 0;
        if (base + k < n) in[base + k] = 0;
        s += v[k];
    }
    int tot;
    int run = block_exclusive_scan(s, &tot) + bsum[blockIdx.x];
#pragma unroll
    for (int k = 0; k < kScanItems; ++k) {
        if (base + k < n) out[base + k] = run;
        run += v[k];
    }
    if (base + kScanItems >= n && base < n) out[n] = run;  // the thread holding the last item
}

// The bits of a double as an unsigned integer that orders like the value (NaNs included,
// so an order exists even for a failed tube).
__device__ __forceinline__ unsigned long long order_bits(double v) {
    const unsigned long long u = (unsigned long long)__double_as_longlong(v);
    return (u >> 63) ? ~u : (u | 0x8000000000000000ull);
}

__global__ __launch_bounds__(kBS) void k_scatter(Map mp, const TubeDyn* __restrict__ td, const int* __restrict__ key,
                                                 const int* __restrict__ slot, const int* __restrict__ start,
                                                 int* __restrict__ tsrc, int* __restrict__ tkey) {
    const ItemPos ip = locate(mp, kBS);
    const int li = ip.li + threadIdx.x;
    if (li >= ip.sg.n) return;
    if (td[ip.sg.tube].done || td[ip.sg.tube].err) return;   // not re-sorted (k_key)
    const int p = ip.sg.off + li;
    const int k = key[p];
    const int pos = start[k] + slot[p];
    tsrc[pos] = p;
    tkey[pos] = k;
}


// Largest h per neighbour cell and per (tube, species), one thread per position of the
// block (called by every thread): keys ascend along a wave, so a segmented max over runs of
// equal keys leaves each run's maximum on its last lane, which raises the cell's value
// (h > 0, so the FP64 bits order like the values).  Cells without particles keep the 0
// they start the step with (cleared at set-up and by k_radius).
__device__ __forceinline__ void cell_max_h(const ItemPos& ip, bool valid, int k, double h,
                                           double* __restrict__ cellh, TubeDyn* __restrict__ td) {
    __shared__ double wmax[kBS / 64];
    const int lane = threadIdx.x & 63;
    for (int o = 1; o < 64; o <<= 1) {
        const double ho = __shfl_up(h, o, 64);
        const int ko = __shfl_up(k, o, 64);
        if (lane >= o && ko == k) h = fmax(h, ho);
    }
    const int kn = __shfl_down(k, 1, 64);
    if (valid && (lane == 63 || kn != k))
        atomicMax(reinterpret_cast<unsigned long long*>(cellh + k), (unsigned long long)__double_as_longlong(h));
    const double m = wave_max(h);
    if (lane == 0) wmax[threadIdx.x >> 6] = m;
    __syncthreads();
    if (threadIdx.x == 0) {
        double bm = 0.0;
        for (int w = 0; w < kBS / 64; ++w) bm = fmax(bm, wmax[w]);
        unsigned long long* t = &td[ip.sg.tube].hmax_bits[ip.sg.species];
        const unsigned long long bits = (unsigned long long)__double_as_longlong(bm);
        // most blocks do not raise the tube maximum: read before the atomic
        if (bm > 0.0 && bits > __atomic_load_n(t, __ATOMIC_RELAXED)) atomicMax(t, bits);
    }
}

constexpr int kRankRun = 1024;   // cell members staged per block by k_rank_permute

// Final position = cell start + number of cell members before it in (x, id) order: cells
// are sorted by x, so a row of cells is one x-sorted run for the pair kernels' windows.
// A finished tube is still sorted every step (its positions moved on its last drift, so
// the order can change once); its last acceleration and de/dt move with it.
__global__ __launch_bounds__(kBS) void k_rank_permute(Map mp, TubeDyn* __restrict__ td,
                                                      const int* __restrict__ cur, PBuf B,
                                                      const int* __restrict__ start, const int* __restrict__ tsrc,
                                                      const int* __restrict__ tkey, int* __restrict__ wkey, int* __restrict__ ginv,
                                                      double* __restrict__ cellh) {
    // the block's positions lie in one segment and keys ascend along them, so the members of
    // every cell the block touches form one run [w0, w1), staged in LDS when it fits
    __shared__ unsigned long long sx[kRankRun];
    __shared__ int sid[kRankRun];
    const ItemPos ip = locate(mp, kBS);
    // a finished (or failed) tube is sorted no more: its final state stays in buffer done_cur,
    // in the order of its last step (wkey), where downloads read it
    if (td[ip.sg.tube].done || td[ip.sg.tube].err) return;
    const int li = ip.li + threadIdx.x;
    const int qf = ip.sg.off + ip.li;
    const int ql = ip.sg.off + min(ip.li + kBS, ip.sg.n) - 1;
    const int w0 = start[tkey[qf]], w1 = start[tkey[ql] + 1];
    const bool staged = w1 - w0 <= kRankRun;
    const long long so = (long long)(*cur) * B.ntot;
    if (staged) {
        for (int k = threadIdx.x; k < w1 - w0; k += kBS) {
            const long long g = so + tsrc[w0 + k];
            sx[k] = order_bits(B.x[g]);
            sid[k] = B.id[g];
        }
    }
    __syncthreads();
    const bool valid = li < ip.sg.n;
    int c = -1;
    double hq = 0.0;
    if (valid) {
    const int q = ip.sg.off + li;
    c = tkey[q];
    const int lo = start[c], hi = start[c + 1];
    const int src = tsrc[q];
    const int my = B.id[so + src];
    const unsigned long long mx = order_bits(B.x[so + src]);
    int r = 0;
    if (staged) {
        for (int k = lo - w0; k < hi - w0; ++k) r += (sx[k] < mx || (sx[k] == mx && sid[k] < my)) ? 1 : 0;
    } else {
        for (int k = lo; k < hi; ++k) {
            const long long g = so + tsrc[k];
            const unsigned long long kx = order_bits(B.x[g]);
            r += (kx < mx || (kx == mx && B.id[g] < my)) ? 1 : 0;
        }
    }
    const int dst = lo + r;
    const long long wo = (long long)(1 - *cur) * B.ntot;
    const long long gs = so + src, gd = wo + dst;
    // every field is read before the first store: a store could alias any field as far as
    // the compiler knows, so loads left after stores would each wait for a round trip.
    // Density, pressure and stopping time are recomputed this step for active particles;
    // frozen particles carry theirs along.
    const bool gas = ip.sg.species == 0;
    const double x = B.x[gs], y = B.y[gs], z = B.z[gs];
    const double vx = B.vx[gs], vy = B.vy[gs], vz = B.vz[gs];
    const double m = B.m[gs];
    hq = B.h[gs];
    const int ind = B.ind[gs];
    const bool keep = ind != 0;
    const double rho = keep ? B.rho[gs] : 0.0;
    const double e = gas ? B.e[gs] : 0.0;
    const double pr = gas && keep ? B.p[gs] : 0.0;
    const double ts = !gas && keep ? B.ts[gs] : 0.0;
    B.x[gd] = x;
    B.y[gd] = y;
    B.z[gd] = z;
    B.vx[gd] = vx;
    B.vy[gd] = vy;
    B.vz[gd] = vz;
    B.m[gd] = m;
    B.h[gd] = hq;
    B.ind[gd] = ind;
    B.id[gd] = my;
    if (keep) B.rho[gd] = rho;
    if (gas) {
        if (ginv) ginv[ip.sg.off + my] = dst;
        B.e[gd] = e;
        if (keep) B.p[gd] = pr;
    } else if (keep) {
        B.ts[gd] = ts;
    }
    wkey[dst] = c;
    }
    cell_max_h(ip, valid, c, hq, cellh, td);
}

// Reach of a cell: every neighbour j of a particle i of the cell has |x_i - x_j| < h_i + h_j
// <= h_cell + h_j.  Cells more than W_a = ceil(2 h_max (1+m) / c_a) cells away along an
// axis are at least W_a c_a >= 2 h_max apart, so h_j is bounded by the max of cell h over
// that window (a periodic axis shorter than the window is scanned once), and the pair
// kernels search ceil(reach / c_a) cells along each axis.
// Search window of the current step along each axis: W_a = ceil(2 h_max (1+m) / c_a).
__device__ __forceinline__ bool radius_window(const TubeConst& T, const TubeDyn& D, int species, int w[3]) {
    const double hs = __longlong_as_double((long long)D.hmax_bits[species]);
    const double span = 2.0 * hs * (1.0 + kRadiusMargin);
    bool ok = true;
    for (int a = 0; a < 3; ++a) {
        w[a] = (int)ceil(span * T.ic[a]);
        ok = ok && w[a] <= kMaxCellReach;
    }
    return ok;
}

// One axis of the window maximum of cell h, separably: out[c] = max of in over the cells
// within W_a of c along axis a (clamped on a bounded axis, wrapped on a periodic one; a
// periodic axis shorter than the window is taken whole).  Max is exact, so three passes
// give the same value as the full window.
// Max of `in` over the window of w[axis] cells either side of local cell li along one axis.
__device__ __forceinline__ double axis_window_max(const TubeConst& T, const int w[3], const double* __restrict__ in,
                                                  int off, int li, int axis) {
    const int n0 = T.n[0], n1 = T.n[1];
    const int k[3] = {li % n0, (li / n0) % n1, li / (n0 * n1)};
    const int stride = axis == 0 ? 1 : (axis == 1 ? n0 : n0 * n1);
    const int n = T.n[axis], ka = k[axis];
    const int base = off + li - ka * stride;
    double m = 0.0;
    if (T.per[axis] && 2 * w[axis] + 1 >= n) {
        for (int c = 0; c < n; ++c) m = fmax(m, in[base + c * stride]);
    } else if (T.per[axis]) {
        for (int o = -w[axis]; o <= w[axis]; ++o) {
            int c = ka + o;
            c -= floordiv(c, n) * n;
            m = fmax(m, in[base + c * stride]);
        }
    } else {
        const int c0 = max(ka - w[axis], 0), c1 = min(ka + w[axis], n - 1);
        for (int c = c0; c <= c1; ++c) m = fmax(m, in[base + c * stride]);
    }
    return m;
}

// The x and y passes of the separable window max (the z pass is taken inside k_radius).  A
// finished tube's cells are skipped: it is not re-sorted, so its cell maxima stay zero and
// k_radius gives its cells no reach without reading the window.
__global__ __launch_bounds__(kBS) void k_window_max(Map mp, const TubeConst* __restrict__ tc,
                                                    const TubeDyn* __restrict__ td, const double* __restrict__ in,
                                                    double* __restrict__ out, int axis) {
    const ItemPos ip = locate(mp, kBS);
    const int li = ip.li + threadIdx.x;
    if (li >= ip.sg.n) return;
    if (td[ip.sg.tube].done || td[ip.sg.tube].err) return;
    const TubeConst& T = tc[ip.sg.tube];
    if (T.h_mode == 0) return;
    int w[3];
    if (!radius_window(T, td[ip.sg.tube], ip.sg.species, w)) return;
    out[ip.sg.off + li] = axis_window_max(T, w, in, ip.sg.off, li, axis);
}

// Reach of a cell: every neighbour j of a particle i of the cell has |x_i - x_j| < h_i + h_j
// <= h_cell + h_j.  Cells more than W_a = ceil(2 h_max (1+m) / c_a) cells away along an
// axis are at least W_a c_a >= 2 h_max apart, so h_j is bounded by the max of cell h over
// that window (the x and y passes of k_window_max, then the z pass here), and the pair
// kernels search ceil(reach / c_a) cells along each axis.
// The cell maxima of h are consumed here (the window passes have read them): each is cleared
// after it is read, ready for the next step's rank pass.
__global__ __launch_bounds__(kBS) void k_radius(Map mp, const TubeConst* __restrict__ tc, TubeDyn* __restrict__ td,
                                                double* __restrict__ cellh, const double* __restrict__ wxy,
                                                double* __restrict__ reach) {
    const ItemPos ip = locate(mp, kBS);
    const int li = ip.li + threadIdx.x;
    if (li >= ip.sg.n) return;
    const int c = ip.sg.off + li;
    const TubeConst& T = tc[ip.sg.tube];
    TubeDyn& D = td[ip.sg.tube];
    const double hc = cellh[c];
    cellh[c] = 0.0;
    if (hc <= 0.0) {
        reach[c] = 0.0;
        return;
    }
    int w[3];
    if (!radius_window(T, D, ip.sg.species, w)) {
        if (li == 0) atomicOr(&D.err, kErrGrid);
        reach[c] = 0.0;
        return;
    }
    if (T.h_mode == 0) {
        const double hs = __longlong_as_double((long long)D.hmax_bits[ip.sg.species]);
        reach[c] = 2.0 * hs * (1.0 + kRadiusMargin);
        return;
    }
    reach[c] = (hc + axis_window_max(T, w, wxy, ip.sg.off, li, 2)) * (1.0 + kRadiusMargin);
}

// Pressure + viscosity terms of one pair (IDIC.h:225-250): returns the acceleration
// coefficient of (dx,dy,dz) and the de/dt contribution.
// d2 is the squared distance: FP64 takes sqrt and 1/r; FP32 takes 1/r from v_rsq_f32 and
// r = d2/r, one transcendental instead of two.
template <typename T>
__device__ __forceinline__ void force_terms(T d2, T hh, T fx, T fy, T fz, T dvx, T dvy, T dvz, T rhoi, T rhoj,
                                            T pr2i, T pr2j, T csi, T csj, T mi, T mj, T alpha, T beta, T epsv,
                                            bool plain, T& coef, T& edc) {
    T dist, rinv, dist2;
    if constexpr (sizeof(T) == 8) {
        dist = sqrt_r<T>(d2);
        rinv = rcp<T>(dist);
        dist2 = dist * dist;
    } else {
        rinv = __builtin_amdgcn_rsqf(d2);
        dist = d2 * rinv;
        dist2 = d2;
    }
    const T ih = rcp<T>(hh);
    // (v_i - v_j).(x_i - x_j), IDIC.h:226-229.  Explicit FMAs (not compiler contraction): the
    // two orientations of a pair run the same operations on the same values (each product's
    // operands both change sign), so the pair's terms stay mirror images bit for bit.
    const T nu = fma_r(dvx, -fx, fma_r(dvy, -fy, dvz * (-fz)));
    T fnu = T(0);
    if (nu < T(0)) {
        const T cnu = T(0.5) * (csi + csj);
        const T mu = hh * nu * rcp<T>(fma_r(epsv * hh, hh, dist2));
        fnu = fma_r(beta * mu, mu, -alpha * cnu * mu) * rcp<T>(T(0.5) * (rhoi + rhoj));
    }
    const T dw = (T)kInvPi * (ih * ih * ih * ih) * spline_dw_hit<T>(dist * ih);
    const T dwr = dw * rinv;
    const T ei = mi * fma_r(fnu, T(0.5), pr2i);
    if (plain)
        edc = ei * plain_edot_weight<T>(dvx, dvy, dvz, fx, fy, fz, ih);
    else
        edc = ei * nu * dwr;
    coef = mj * (pr2i + pr2j + fnu) * dwr;
}

// Band flag of every sorted particle's 64-particle chunk (one wave per chunk; the chunks of
// the tiled pair kernels): 1 if a particle of the chunk lies in its tube's FP64 band this
// step.  Band chunks run the FP64 pair kernels; FP32 chunks evaluate their pairs with band
// (and frozen) particles in FP64, so both particles of a pair compute its terms alike.
// The chunk's class for the fused kernels' order: 1 band, 0 other, 2 a finished or failed
// tube's (no pair work: left out of the order).
__global__ __launch_bounds__(kBS) void k_band(Map mp, const TubeDyn* __restrict__ td, const int* __restrict__ cur,
                                              PBuf B, unsigned char* __restrict__ pband, const int* __restrict__ blk64,
                                              int* __restrict__ cflag) {
    const ItemPos ip = locate(mp, kBS);
    const int li = ip.li + threadIdx.x;
    const bool valid = li < ip.sg.n;
    const TubeDyn& D = td[ip.sg.tube];
    const long long wo = (long long)(1 - *cur) * B.ntot;
    bool in = false;
    if (valid && pband != nullptr) {
        const double x = B.x[wo + ip.sg.off + li];
        in = x < D.band_lo || x > D.band_hi;
    }
    const bool any = __ballot(in) != 0ull;
    if (valid && pband != nullptr) pband[ip.sg.off + li] = any ? 1 : 0;
    // the chunk's flag by its pmap64 block (pmap and pmap64 list the same segments, and both
    // start each segment on a chunk boundary)
    if (cflag != nullptr && (threadIdx.x & 63) == 0) {
        const int ci = li >> 6;
        const int cls = (D.done || D.err) ? 2 : (any ? 1 : 0);
        if (ci < ((ip.sg.n + 63) >> 6)) cflag[blk64[find_seg(mp, blockIdx.x)] + ci] = cls;
    }
}

// Chunk order of the fused mixed-precision pair kernel in perm[0, nall): the band's chunks
// ascending, then the other chunks of running tubes ascending, then the chunks of finished
// tubes.  perm[nall] = band chunks, perm[nall + 1] = band chunks below gas_blocks (the gas
// chunks), perm[nall + 2] = other running chunks, perm[nall + 3] = those below gas_blocks.
// Two passes over blocks of kPermThreads chunks: counts per block (band, finished), then each
// block writes its chunks at its offsets.
constexpr int kPermThreads = 1024;
__device__ __forceinline__ int block_sum_1024(int v, int* sh) {
    const int w = threadIdx.x >> 6;
    int t = v;
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) t += __shfl_xor(t, o, 64);
    if ((threadIdx.x & 63) == 0) sh[w] = t;
    __syncthreads();
    int r = 0;
#pragma unroll
    for (int k = 0; k < kPermThreads / 64; ++k) r += sh[k];
    __syncthreads();
    return r;
}
__global__ __launch_bounds__(kPermThreads) void k_perm_count(const int* __restrict__ cflag, int nall,
                                                            int* __restrict__ bcnt) {
    __shared__ int sh[kPermThreads / 64];
    const int b = blockIdx.x * kPermThreads + threadIdx.x;
    const int f = b < nall ? cflag[b] : 0;
    const int band = block_sum_1024(f == 1 ? 1 : 0, sh);
    const int gone = block_sum_1024(f == 2 ? 1 : 0, sh);
    if (threadIdx.x == 0) {
        bcnt[2 * blockIdx.x] = band;
        bcnt[2 * blockIdx.x + 1] = gone;
    }
}
// Count of the chunks of class c before this thread's chunk inside the block: wave prefix by
// ballot, then the waves before it.
__device__ __forceinline__ int class_before_in_block(bool is_c, int* wpre) {
    const unsigned long long m = __ballot(is_c);
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int in_wave = __popcll(m & ((1ull << lane) - 1ull));
    if (lane == 0) wpre[w] = __popcll(m);
    __syncthreads();
    int waves_before = 0;
    for (int k = 0; k < w; ++k) waves_before += wpre[k];
    __syncthreads();
    return waves_before + in_wave;
}
__global__ __launch_bounds__(kPermThreads) void k_perm_write(const int* __restrict__ cflag, int nall,
                                                            const int* __restrict__ bcnt, int gas_blocks,
                                                            int* __restrict__ perm) {
    __shared__ int sh[kPermThreads / 64];
    __shared__ int wpre[kPermThreads / 64];
    // band and finished chunks of the blocks before this one, and of all blocks (nbk <= a few
    // hundred)
    int bb = 0, bt = 0, gb = 0, gt = 0;
    for (int k = threadIdx.x; k < (int)gridDim.x; k += kPermThreads) {
        const int vb = bcnt[2 * k], vg = bcnt[2 * k + 1];
        bt += vb;
        gt += vg;
        bb += k < (int)blockIdx.x ? vb : 0;
        gb += k < (int)blockIdx.x ? vg : 0;
    }
    const int band_blocks_before = block_sum_1024(bb, sh);
    const int band_total = block_sum_1024(bt, sh);
    const int gone_blocks_before = block_sum_1024(gb, sh);
    const int gone_total = block_sum_1024(gt, sh);
    const int rest_total = nall - band_total - gone_total;
    const int b = blockIdx.x * kPermThreads + threadIdx.x;
    const int f = b < nall ? cflag[b] : 0;
    const int band_before = band_blocks_before + class_before_in_block(b < nall && f == 1, wpre);
    const int gone_before = gone_blocks_before + class_before_in_block(b < nall && f == 2, wpre);
    const int rest_before = b - band_before - gone_before;
    if (b < nall) {
        if (f == 1) perm[band_before] = b;
        else if (f == 2) perm[band_total + rest_total + gone_before] = b;
        else perm[band_total + rest_before] = b;
        if (b == nall - 1) {
            perm[nall] = band_total;
            perm[nall + 2] = rest_total;
        }
        if (b == gas_blocks - 1) {
            perm[nall + 1] = band_before + (f == 1 ? 1 : 0);
            perm[nall + 3] = rest_before + (f == 0 ? 1 : 0);
        }
    }
}

template <typename R, bool kStats>
struct DensPair {
    const double* __restrict__ M;
    const double* __restrict__ H;
    const int* __restrict__ ID;
    double hi;
    Acc<R> acc;
    int cnt = 0;
    uint64_t hash = 0;
    __device__ __forceinline__ void pair(int j, double, double, double, double d2, int sy, int sz) {
        const R r = sqrt_r<R>((R)d2);
        const R hh = (R)(0.5 * (hi + H[j]));
        const R ih = rcp<R>(hh);
        const R q = r * ih;
        const R w = (R)kInvPi * (ih * ih * ih) * spline_w<R>(q);
        acc.add((R)M[j] * w);
        if (kStats) {
            ++cnt;
            hash += pair_tag(ID[j], sy, sz);
        }
    }
};

template <typename R, bool kStats>
__global__ __launch_bounds__(kBS) void k_density(Map mp, const TubeConst* __restrict__ tc,
                                                 const TubeDyn* __restrict__ td, const int* __restrict__ cur,
                                                 PBuf B, const int* __restrict__ start,
                                                 const int* __restrict__ wkey, const double* __restrict__ reach,
                                                 double* __restrict__ pr2, double* __restrict__ csav,
                                                 int* __restrict__ ncount, unsigned long long* __restrict__ nhash,
                                                 const unsigned char* __restrict__ pband) {
    const ItemPos ip = locate(mp, kBS);
    const int li = ip.li + threadIdx.x;
    if (li >= ip.sg.n) return;
    const TubeDyn& D = td[ip.sg.tube];
    if (D.done || D.err) return;
    const TubeConst& T = tc[ip.sg.tube];
    const int s = ip.sg.species;
    const int q = ip.sg.off + li;
    const long long wo = (long long)(1 - *cur) * B.ntot;
    const double* X = B.x + wo;
    const double* Y = B.y + wo;
    const double* Z = B.z + wo;
    const double* H = B.h + wo;
    double rho = B.rho[wo + q];
    if (B.ind[wo + q] == 0) {
        const int cell_off = T.cell_off[s];
        const int lc = wkey[q] - cell_off;
        const int kx = lc % T.n[0];
        const int ky = (lc / T.n[0]) % T.n[1];
        const int kz = lc / (T.n[0] * T.n[1]);
        // a target of an FP64-band chunk sums in FP64
        auto run = [&](auto& f) {
            for_pairs(T, cell_off, start, kx, ky, kz, reach[wkey[q]], X, Y, Z, H, X[q], Y[q], Z[q], H[q], f);
            rho = f.acc.get();
            if (kStats) {
                ncount[q] = f.cnt;
                nhash[q] = f.hash;
            }
        };
        if (sizeof(R) == 4 && pband != nullptr && pband[q] != 0) {
            DensPair<double, kStats> f{B.m + wo, H, B.id + wo, H[q]};
            run(f);
        } else {
            DensPair<R, kStats> f{B.m + wo, H, B.id + wo, H[q]};
            run(f);
        }
        B.rho[wo + q] = rho;
        if (s == 0) {
            const double p = (T.eos == 0) ? rho * B.e[wo + q] * (T.gamma - 1) : T.cs * T.cs * rho;
            B.p[wo + q] = p;
        } else {
            if (T.drag_mode == 1) B.ts[wo + q] = rho / T.K[s - 1];
        }
    } else if (kStats) {
        ncount[q] = 0;
        nhash[q] = 0ull;
    }
    if (s == 0) {
        const double p = B.p[wo + q];
        pr2[q] = p / (rho * rho);
        csav[q] = sqrt(T.gamma * p / rho);
    }
}

// One target's force sums over the direct traversal.  Mixed precision: a pair with a frozen
// neighbour, or with both particles in FP64-band chunks, is evaluated in FP64 (its partner
// evaluates it so too), every other pair in FP32, all accumulated in FP64.
template <typename R>
struct ForcePair {
    const double* __restrict__ M;
    const double* __restrict__ H;
    const double* __restrict__ VX;
    const double* __restrict__ VY;
    const double* __restrict__ VZ;
    const double* __restrict__ RHO;
    const double* __restrict__ PR2;
    const double* __restrict__ CS;
    const int* __restrict__ IND;
    const unsigned char* __restrict__ PB;
    double hi, vxi, vyi, vzi, rhoi, pr2i, csi, mi, alpha, beta, eps;
    bool plain;   // gas-only run: de/dt of the plain ForceKernel
    bool f64i;    // the target's chunk is in the FP64 band
    Acc<R> ax, ay, az, ed;
    double dx_ = 0.0, dy_ = 0.0, dz_ = 0.0, de_ = 0.0;
    __device__ __forceinline__ void pair(int j, double dx, double dy, double dz, double d2, int, int) {
        if (!(d2 > 0.0)) return;
        const double hh = 0.5 * (hi + H[j]);
        if (sizeof(R) == 8 || IND[j] != 0 || (f64i && PB[j] != 0)) {
            double coef, edc;
            force_terms<double>(d2, hh, dx, dy, dz, vxi - VX[j], vyi - VY[j], vzi - VZ[j], rhoi, RHO[j], pr2i, PR2[j],
                                csi, CS[j], mi, M[j], alpha, beta, eps, plain, coef, edc);
            dx_ += dx * coef;
            dy_ += dy * coef;
            dz_ += dz * coef;
            de_ += edc;
        } else {
            const R fx = (R)dx, fy = (R)dy, fz = (R)dz;
            R coef, edc;
            force_terms<R>((R)d2, (R)hh, fx, fy, fz, (R)(vxi - VX[j]), (R)(vyi - VY[j]), (R)(vzi - VZ[j]), (R)rhoi,
                           (R)RHO[j], (R)pr2i, (R)PR2[j], (R)csi, (R)CS[j], (R)mi, (R)M[j], (R)alpha, (R)beta, (R)eps,
                           plain, coef, edc);
            ax.add(fx * coef);
            ay.add(fy * coef);
            az.add(fz * coef);
            ed.add(edc);
        }
    }
};

template <typename R>
__global__ __launch_bounds__(kBS) void k_force(Map mp, const TubeConst* __restrict__ tc,
                                               const TubeDyn* __restrict__ td, const int* __restrict__ cur, PBuf B,
                                               const int* __restrict__ start, const int* __restrict__ wkey,
                                               const double* __restrict__ reach, const double* __restrict__ pr2,
                                               const double* __restrict__ csav,
                                               const unsigned char* __restrict__ pband) {
    const ItemPos ip = locate(mp, kBS);
    const int li = ip.li + threadIdx.x;
    if (li >= ip.sg.n) return;
    const TubeDyn& D = td[ip.sg.tube];
    if (D.done || D.err) return;
    const TubeConst& T = tc[ip.sg.tube];
    const int q = ip.sg.off + li;
    const long long wo = (long long)(1 - *cur) * B.ntot;
    if (B.ind[wo + q] != 0) {
        B.ax[wo + q] = 0.0;
        B.ay[wo + q] = 0.0;
        B.az[wo + q] = 0.0;
        B.ed[wo + q] = 0.0;
        return;
    }
    const double* X = B.x + wo;
    const double* Y = B.y + wo;
    const double* Z = B.z + wo;
    const double* H = B.h + wo;
    const int cell_off = T.cell_off[0];
    const int lc = wkey[q] - cell_off;
    const int kx = lc % T.n[0];
    const int ky = (lc / T.n[0]) % T.n[1];
    const int kz = lc / (T.n[0] * T.n[1]);
    ForcePair<R> f{B.m + wo, H, B.vx + wo, B.vy + wo, B.vz + wo, B.rho + wo, pr2, csav, B.ind + wo, pband,
                   H[q], B.vx[wo + q], B.vy[wo + q], B.vz[wo + q],
                   B.rho[wo + q], pr2[q], csav[q], B.m[wo + q],
                   T.alpha, T.beta, T.eps, T.nsp == 1, pband != nullptr && pband[q] != 0};
    for_pairs(T, cell_off, start, kx, ky, kz, reach[wkey[q]], X, Y, Z, H, X[q], Y[q], Z[q], H[q], f);
    B.ax[wo + q] = f.ax.get() + f.dx_;
    B.ay[wo + q] = f.ay.get() + f.dy_;
    B.az[wo + q] = f.az.get() + f.dz_;
    B.ed[wo + q] = f.ed.get() + f.de_;
}

// ---------------------------------------------------------------- tiled pair kernels
// One wave per chunk of 64 consecutive sorted targets.  Targets of one x-row of cells
// are handled together: for every (dy,dz) row offset the union of their candidate
// cells is one contiguous sorted range, staged into LDS; each lane then runs the FP64
// predicate over its own cell range only, collecting hits in a 64-bit mask per window of
// 64 candidates, and does the pair arithmetic for the hits alone.  The candidate set of
// a lane is the one of the direct kernels (cells within ceil(reach / c_a) of its own),
// so the neighbour sets are identical.
// candidates staged per piece: a template parameter (128 to 256, kTiles) chosen at run time;
// the pieces split each row's range in order, so every size gives the same sums bit for bit
constexpr int kMaskSlots = 48;  // stored hit-mask words per lane of a gas chunk

__device__ __forceinline__ int xcd_remap(int b, int nb) {
    // consecutive workgroups are dealt round-robin to the 8 XCDs; give each XCD a
    // contiguous run of chunks so neighbouring chunks share that XCD's L2
    const int per = nb >> 3;
    if (b >= per * 8) return b;
    return (b & 7) * per + (b >> 3);
}

// Wave maximum of v in [0, 512) by a bit-by-bit search on ballots: comparisons and scalar
// operations only, no cross-lane data movement through LDS.
__device__ __forceinline__ int wave_max_u9(int v) {
    int r = 0;
#pragma unroll
    for (int b = 8; b >= 0; --b) {
        const int c = r | (1 << b);
        if (__ballot(v >= c) != 0ull) r = c;
    }
    return r;
}
// a value equal on every lane, as a scalar (kept in SGPRs)
__device__ __forceinline__ int uni(int v) { return __builtin_amdgcn_readfirstlane(v); }
__device__ __forceinline__ double uni(double v) {
    const unsigned long long b = (unsigned long long)__double_as_longlong(v);
    const unsigned lo = (unsigned)__builtin_amdgcn_readfirstlane((int)(unsigned)b);
    const unsigned hi = (unsigned)__builtin_amdgcn_readfirstlane((int)(unsigned)(b >> 32));
    return __longlong_as_double((long long)(((unsigned long long)hi << 32) | lo));
}

// One candidate of the FP64 neighbour predicate, as a bit at position t (t < 32, uniform).
// One candidate's position and smoothing length, gathered from the staged arrays.
struct P4 {
    double x, y, z, h;
};
// Pair predicate d2 < (h_i + h_j)^2 with d2 = fma(dz, dz, fma(dy, dy, dx*dx)); a periodic
// image offset is added to the coordinate difference, not to the coordinate.
__device__ __forceinline__ uint32_t pred1(const P4& c, double xi, double yi, double zi, double hi) {
    const double dx = c.x - xi;
    const double dy = c.y - yi;
    const double dz = c.z - zi;
    const double dd = __builtin_fma(dz, dz, __builtin_fma(dy, dy, dx * dx));
    const double hs = hi + c.h;
    return dd < hs * hs ? 1u : 0u;
}
__device__ __forceinline__ uint32_t pred1_shift(const P4& c, double xi, double yi, double zi, double hi, double shy,
                                                double shz) {
    const double dx = c.x - xi;
    const double dy = (c.y - yi) + shy;
    const double dz = (c.z - zi) + shz;
    const double dd = __builtin_fma(dz, dz, __builtin_fma(dy, dy, dx * dx));
    const double hs = hi + c.h;
    return dd < hs * hs ? 1u : 0u;
}

// 32-bit fixed point of a box coordinate: units of L / 2^30 from the box origin, so every
// coordinate from one box length below the box to one above fits and an image shift is 2^30
// units.  Differences are taken modulo 2^32, exact for any pair closer than 2 L.
constexpr double kFixUnits = 1073741824.0;   // 2^30
__device__ __forceinline__ int fix_coord(double v, double o, double units_per_length) {
    return __double2int_rn((v - o) * units_per_length);
}
__device__ __forceinline__ int fix_diff(int a, int b, int images) {
    return (int)((unsigned)a - (unsigned)b + (unsigned)images * 1073741824u);
}

// A density window reads slots [base, base + wn4) with base = min(a, kTileP - wn4), inside the
// piece for every lane, and shifts the lane's bits down by a - base; bits past a lane's own
// range are masked.  The force pass reads the density pass's hit masks; its rare fallback
// (more windows than mask slots) stays inside each lane's range.  Separate arrays (not one
// record per candidate) keep the lanes' 8-byte reads on distinct banks.  Velocities are
// staged in the pair-arithmetic type R: v_i - v_j is formed in R, antisymmetric bit for bit.
// The FP32 force pass stages y and z as 32-bit fixed point and h as FP32 (x stays FP64):
// an integer difference negates exactly, so dy and dz stay antisymmetric, and the smaller
// record lets larger pieces (more mirror-row pairs) keep 4 waves per SIMD.
template <typename R, int KIND>
constexpr bool kFixYZ = KIND == 1 && sizeof(R) == 4;
template <typename R, int KIND, bool kStats, int kTileP, bool kPairs = kFixYZ<R, KIND>>
struct TileLds {
    using YZ = double;
    using HT = double;
    double x[kTileP];
    YZ y[kTileP], z[kTileP];
    HT h[kTileP];
    R vx[KIND ? kTileP : 1], vy[KIND ? kTileP : 1], vz[KIND ? kTileP : 1];
    R m[KIND ? kTileP : 1];
    // density: h / 2 and m / pi side by side, one LDS read per hit at the stride of x
    struct alignas(2 * sizeof(R)) HM {
        R hh, mw;
    };
    HM hm[KIND ? 1 : kTileP];
    R rho[KIND ? kTileP : 1], pr2[KIND ? kTileP : 1], cs[KIND ? kTileP : 1];
    unsigned char ind[KIND ? kTileP : 1];
    int id[kStats ? kTileP : 1];
    __device__ YZ Y(int u) const { return y[u]; }
    __device__ YZ Z(int u) const { return z[u]; }
    __device__ HT H(int u) const { return h[u]; }
    __device__ R M(int u) const { return m[u]; }
    __device__ R VX(int u) const { return vx[u]; }
    __device__ R VY(int u) const { return vy[u]; }
    __device__ R VZ(int u) const { return vz[u]; }
    __device__ R RHO(int u) const { return rho[u]; }
    __device__ R PR2(int u) const { return pr2[u]; }
    __device__ R CS(int u) const { return cs[u]; }
};
// The FP32 force tile: every 4-byte field in 8-byte pairs, so that all of a candidate's
// fields sit at multiples of its index times 8 (one address for x and every pair, no
// address arithmetic per field) and each pair is one LDS access.
template <typename R, int KIND, bool kStats, int kTileP>
struct TileLds<R, KIND, kStats, kTileP, true> {
    using YZ = int;
    using HT = float;
    struct alignas(8) I2 {
        int a, b;
    };
    struct alignas(8) F2 {
        float a, b;
    };
    double x[kTileP];
    I2 yz[kTileP];    // y, z (fixed point)
    F2 hm[kTileP];    // h, m
    F2 vxy[kTileP];   // vx, vy
    F2 vzp[kTileP];   // vz, pr2
    F2 rc[kTileP];    // rho, cs (read by the viscosity term only)
    unsigned char ind[kTileP];
    int id[kStats ? kTileP : 1];
    __device__ int Y(int u) const { return yz[u].a; }
    __device__ int Z(int u) const { return yz[u].b; }
    __device__ float H(int u) const { return hm[u].a; }
    __device__ float M(int u) const { return hm[u].b; }
    __device__ float VX(int u) const { return vxy[u].a; }
    __device__ float VY(int u) const { return vxy[u].b; }
    __device__ float VZ(int u) const { return vzp[u].a; }
    __device__ float PR2(int u) const { return vzp[u].b; }
    __device__ float RHO(int u) const { return rc[u].a; }
    __device__ float CS(int u) const { return rc[u].b; }
    __device__ void set(int k, int y, int z, float h, float m, float vx, float vy, float vz, float rho, float pr2,
                        float cs) {
        yz[k] = I2{y, z};
        hm[k] = F2{h, m};
        vxy[k] = F2{vx, vy};
        vzp[k] = F2{vz, pr2};
        rc[k] = F2{rho, cs};
    }
};

// The LDS tile of a wave is its own (a block of kWaves waves holds kWaves tiles), so the
// staging barriers order one wave's LDS writes and reads only: waves of a block never wait
// for each other, and a wave may leave early.  One wave per block: the block barrier.
template <int kWaves>
__device__ __forceinline__ void tile_sync() {
    if constexpr (kWaves == 1) {
        __syncthreads();
    } else {
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    }
}

// KIND 0: density (all species); KIND 1: gas force and de/dt.
// The gas density pass stores its predicate hit masks (one 64-bit word per lane per
// uniform (row, piece, window) iteration, kMaskSlots per chunk); the force pass replays
// the same iteration order and reads them instead of re-running the predicate.
// Occupancy targets: both passes are latency- as well as VALU-bound, so the register
// budget is capped at 4 (force) / 5 (density) waves per SIMD; with pieces of 192 candidates
// the LDS of a wave (9.4 KB FP32 force, 7.7 KB FP32 density) allows as many.
// One 64-target chunk b (the block index of the 64-per-block map mp) of a tiled pair pass, by
// one wave.  With pband (mixed precision with an FP64 wall band), a chunk whose band flag does
// not match the instance's precision is left to the other instance; the force pair of two
// band particles is FP64, of a frozen and an active particle FP64, of any other two FP32.
template <typename R, int KIND, bool kStats, int kTileP, int kWaves>
__device__ __forceinline__ void pairs_chunk(TileLds<R, KIND, kStats, kTileP>& L, int b, const Map& mp,
                                            const TubeConst* __restrict__ tc,
                                            const TubeDyn* __restrict__ td, const int* __restrict__ cur,
                                            const PBuf& B, const int* __restrict__ start,
                                            const int* __restrict__ wkey, const double* __restrict__ reach,
                                            double* __restrict__ pr2, double* __restrict__ csav,
                                            int* __restrict__ ncount, unsigned long long* __restrict__ nhash,
                                            unsigned long long* __restrict__ masks, int mask_chunks,
                                            const unsigned char* __restrict__ pband, ForceRec* __restrict__ frec) {
    static_assert(kTileP <= 256, "window lengths are reduced with wave_max_u9");
    static_assert(kTileP >= 64, "a density window of 64 slots fits in one piece");
    const int sgi = find_seg(mp, b);
    const Seg sg = mp.seg[sgi];
    const int lane = threadIdx.x & 63;
    const int li = (b - mp.blk[sgi]) * 64 + lane;
    const TubeDyn& D = td[sg.tube];
    if (D.done || D.err) return;
    const TubeConst& T = tc[sg.tube];
    const int s = sg.species;
    const bool valid = li < sg.n;
    const int q = sg.off + (valid ? li : sg.n - 1);
    // mixed precision with a wall band: the chunks of the band run in the FP64 instance of this
    // kernel, the others in the FP32 instance (one flag per chunk, so the exit is wave-uniform)
    if (pband != nullptr && ((uni((int)pband[q]) != 0) != (sizeof(R) == 8))) return;
    const long long wo = (long long)(1 - *cur) * B.ntot;
    const int nx = T.n[0], ny = T.n[1], nz = T.n[2];
    const int cell_off = T.cell_off[s];
    const bool active = valid && B.ind[wo + q] == 0;
    const int lc = wkey[q] - cell_off;
    const int kx = lc % nx;
    const int row = lc / nx;
    const double rch = reach[wkey[q]];
    const double xi = B.x[wo + q], yi = B.y[wo + q], zi = B.z[wo + q], hi = B.h[wo + q];
    const int rxl = (int)ceil(rch * T.ic[0]);
    const int ryl = (int)ceil(rch * T.ic[1]);
    const int rzl = (int)ceil(rch * T.ic[2]);
    // hit masks: written by the gas density pass, read by the force pass
    const bool use_masks = masks != nullptr && s == 0 && b < mask_chunks;
    unsigned long long* mrow = use_masks ? masks + (size_t)b * kMaskSlots * 64 + lane : nullptr;
    // after the masks: each lane's x-window (start << 16 | length) per visited row segment,
    // written by the gas density pass, read by the force pass in place of its own search
    unsigned int* wrow =
        use_masks ? reinterpret_cast<unsigned int*>(masks + (size_t)mask_chunks * kMaskSlots * 64) +
                        (size_t)b * kMaskSlots * 64 + lane
                  : nullptr;
    int segi = 0;
    int slot = 0;
    // force pass: the stored x-windows of the next two segments and the hit masks of the next
    // two slots, loaded together with a piece's candidates (so their memory round trip overlaps
    // the staging instead of following it); pw_at / pm_at: the segment / slot they belong to
    unsigned int pw[2] = {0u, 0u};
    unsigned long long pm[2] = {0ull, 0ull};
    int pw_at = -2, pm_at = -2;
    // once the force pass searches a window itself (past the stored ones) it takes no stored
    // mask either: masks follow the windows they were made for
    bool win_own = false;

    Acc<R> a0, a1, a2, a3;  // density: a0; force: ax, ay, az, de/dt
    double d0 = 0.0, d1 = 0.0, d2s = 0.0, d3 = 0.0;  // force, FP64 terms of frozen neighbours
    int cnt = 0;
    uint64_t hash = 0;
    double vxi = 0, vyi = 0, vzi = 0, rhoi_d = 1, pr2i_d = 0, csi_d = 0, mi_d = 0;
    if (KIND == 1) {
        vxi = B.vx[wo + q];
        vyi = B.vy[wo + q];
        vzi = B.vz[wo + q];
        rhoi_d = B.rho[wo + q];
        pr2i_d = pr2[q];
        csi_d = csav[q];
        mi_d = B.m[wo + q];
    }
    const R rhoi = (R)rhoi_d, pr2i = (R)pr2i_d, csi = (R)csi_d, mi = (R)mi_d;
    const R vxi_r = (R)vxi, vyi_r = (R)vyi, vzi_r = (R)vzi;
    const R alpha = (R)T.alpha, beta = (R)T.beta, epsv = (R)T.eps;
    const bool plain = T.nsp == 1;   // gas-only run: de/dt of the plain ForceKernel
    const R hhi = (R)(0.5 * hi);
    constexpr bool kFix = kFixYZ<R, KIND>;
    const double ups_y = kFixUnits / T.L[1], ups_z = kFixUnits / T.L[2];
    const float qy = (float)(T.L[1] / kFixUnits), qz = (float)(T.L[2] / kFixUnits);
    const int Yi = fix_coord(yi, T.o[1], ups_y), Zi = fix_coord(zi, T.o[2], ups_z);
    const float hif = (float)hi;
    // a neighbour within reach 2 or more box lengths away across periodic images, or a
    // coordinate outside the fixed-point range (one box length either side of the box), would
    // break the fixed-point difference: such waves (or pieces) take dy, dz from FP64 positions
    auto fix_range_ok = [&](double y, double z) {
        const double uy = (y - T.o[1]) * ups_y, uz = (z - T.o[2]) * ups_z;
        return uy > -0.95 * kFixUnits && uy < 1.95 * kFixUnits && uz > -0.95 * kFixUnits && uz < 1.95 * kFixUnits;
    };
    const bool far_images =
        kFix && __ballot(rch >= 1.75 * fmin(T.L[1], T.L[2]) || !fix_range_ok(yi, zi)) != 0ull;
    // an FP64-band wave evaluates its pairs with active neighbours outside the band as their
    // FP32 waves do, taking dy, dz from FP64 positions under the same rule (a wave whose reach
    // spans 1.75 box lengths)
    const bool far_mirror = KIND == 1 && sizeof(R) == 8 && pband != nullptr &&
                            __ballot(rch >= 1.75 * fmin(T.L[1], T.L[2]) || !fix_range_ok(yi, zi)) != 0ull;
    bool far_piece = far_images;

    // One row of cells (a (dy,dz) offset) for the targets of the current x-row: the union
    // range [ub, ue) of all lanes, this lane's own range [jl, jh) (empty if it does not need
    // the row), and the periodic image of the row.
    struct RowSeg {
        int ub, ue, jl, jh, sy, sz;
        double shy, shz;
        bool on, shifted;
    };
    // Stages the sorted candidates [s0, s0+n0) then [s1, s1+n1) into LDS slots [0, n0+n1).
    // All loads of a candidate are issued before its LDS writes (the scheduling barrier
    // keeps the compiler from serialising them under register pressure).
    auto stage = [&](int s0, int n0, int s1, int n1) {
        tile_sync<kWaves>();
        if constexpr (KIND == 1) {
            if (use_masks && segi < kMaskSlots) {
                pw_at = segi;
                pw[0] = wrow[(size_t)segi * 64];
                pw[1] = segi + 1 < kMaskSlots ? wrow[(size_t)(segi + 1) * 64] : 0u;
            }
            if (use_masks && slot < kMaskSlots && !win_own) {
                pm_at = slot;
                pm[0] = mrow[(size_t)slot * 64];
                pm[1] = slot + 1 < kMaskSlots ? mrow[(size_t)(slot + 1) * 64] : 0ull;
            }
        }
        bool out_of_range = false;
        for (int k = lane; k < n0 + n1; k += 64) {
            const int src = k < n0 ? s0 + k : s1 + (k - n0);
            if constexpr (kFix) {
                if (frec != nullptr) {
                    const ForceRec r = frec[src];
                    __builtin_amdgcn_sched_barrier(0);
                    L.x[k] = r.x;
                    L.set(k, r.y, r.z, r.h, r.m, r.vx, r.vy, r.vz, r.rho, r.pr2, r.cs);
                    L.ind[k] = (unsigned char)(r.flags & 3u);
                    out_of_range = out_of_range || (r.flags & 4u) == 0u;
                    continue;
                }
            }
            const long long g = wo + src;
            const double x = B.x[g], y = B.y[g], z = B.z[g], h = B.h[g], m = B.m[g];
            double vx = 0, vy = 0, vz = 0, rho = 0, p2 = 0, cs = 0;
            int ind = 0, id = 0;
            if (KIND == 1) {
                vx = B.vx[g];
                vy = B.vy[g];
                vz = B.vz[g];
                rho = B.rho[g];
                p2 = pr2[src];
                cs = csav[src];
                ind = B.ind[g];
            }
            if (kStats) id = B.id[g];
            __builtin_amdgcn_sched_barrier(0);
            L.x[k] = x;
            if constexpr (kFix) {
                L.set(k, fix_coord(y, T.o[1], ups_y), fix_coord(z, T.o[2], ups_z), (float)h, (float)m, (float)vx,
                      (float)vy, (float)vz, (float)rho, (float)p2, (float)cs);
                out_of_range = out_of_range || !fix_range_ok(y, z);
                // bit 0: frozen (an FP64 pair); bit 1: in an FP64-band chunk
                L.ind[k] = (unsigned char)((ind != 0 ? 1 : 0) | (pband != nullptr && pband[src] != 0 ? 2 : 0));
            } else {
                L.y[k] = y;
                L.z[k] = z;
                L.h[k] = h;
                if (KIND == 0) {
                    L.hm[k] = {(R)(0.5 * h), (R)(m * kInvPi)};
                } else {
                    L.m[k] = (R)m;
                    L.vx[k] = (R)vx;
                    L.vy[k] = (R)vy;
                    L.vz[k] = (R)vz;
                    L.rho[k] = (R)rho;
                    L.pr2[k] = (R)p2;
                    L.cs[k] = (R)cs;
                    // bit 0: frozen (an FP64 pair); bit 1: in an FP64-band chunk
                    L.ind[k] = (unsigned char)((ind != 0 ? 1 : 0) | (pband != nullptr && pband[src] != 0 ? 2 : 0));
                }
            }
            if (kStats) L.id[k] = id;
        }
        if constexpr (kFix) far_piece = far_images || __ballot(out_of_range) != 0ull;
        tile_sync<kWaves>();
    };
    // Hit mask of one window of 64 candidates of this lane's range, starting at staged slot
    // a (>= 0); len = candidates of the lane's range from a on, wn = the wave's maximum of
    // len (<= 64 used).  The force pass reads the density pass's mask.
    auto window_mask = [&](int a, int len, int wn, const RowSeg& g, int gb) -> unsigned long long {
        unsigned long long mask = 0ull;
        if constexpr (KIND == 1) {
            if (use_masks && slot < kMaskSlots && !win_own) {
                mask = slot == pm_at ? pm[0] : (slot == pm_at + 1 ? pm[1] : mrow[(size_t)slot * 64]);
            } else {
                // no stored mask: each lane tests its own candidates only
                const int nt = min(max(len, 0), 64);
                for (int t = 0; t < nt; ++t) {
                    const int u = a + t;
                    P4 c;
                    double ty = yi, tz = zi, th = hi;
                    if constexpr (kFix) {
                        // FP64 y, z, h of the candidate (sorted index gb + u) and of the target
                        // from memory
                        const long long gg = wo + gb + u;
                        c = P4{L.x[u], B.y[gg], B.z[gg], B.h[gg]};
                        int qq = q;
                        asm volatile("" : "+v"(qq));
                        ty = B.y[wo + qq];
                        tz = B.z[wo + qq];
                        th = B.h[wo + qq];
                    } else {
                        c = P4{L.x[u], (double)L.Y(u), (double)L.Z(u), (double)L.H(u)};
                    }
                    const uint32_t hit = g.shifted ? pred1_shift(c, xi, ty, tz, th, g.shy, g.shz)
                                                   : pred1(c, xi, ty, tz, th);
                    mask |= (unsigned long long)hit << t;
                }
            }
        } else {
            // uniform trip count: every lane evaluates wn4 slots from a base that keeps them
            // inside the piece (a lane whose window starts past base finds candidate a + u on
            // bit a - base + u); bits past the lane's own range are cleared once per window
            wn = min(wn, 64);
            const unsigned long long valid = len >= 64 ? ~0ull : (len <= 0 ? 0ull : ((1ull << len) - 1ull));
            const int base = min(a, kTileP - ((wn + 3) & ~3));
            const int shift = a - base;
            // candidates are shifted in from the top, so candidate t lands on bit t
            // z and h are addressed from a register of their own: one ds_read2 offset cannot
            // reach them from x's address, and without it the compiler adds two base offsets
            // per pair of slots (the empty asm keeps it from folding zbase back into base)
            int zbase = base;
            asm volatile("" : "+v"(zbase));
            const double *cx_ = L.x + base, *cy_ = L.y + base, *cz_ = L.z + zbase, *ch_ = L.h + zbase;
            // whole groups of 4 (slots past a lane's window are masked)
            const int wn4 = (wn + 3) & ~3;
            const int n0 = min(wn4, 32);
            uint32_t mlo = 0u, mhi = 0u;
            const double shy = g.shy, shz = g.shz;
            auto group = [&](auto shifted_tag, uint32_t m, int t) {
                constexpr bool kSh = decltype(shifted_tag)::value;
#pragma unroll
                for (int k = 3; k >= 0; --k) {
                    const P4 c{cx_[t + k], cy_[t + k], cz_[t + k], ch_[t + k]};
                    m = (m << 1) | (kSh ? pred1_shift(c, xi, yi, zi, hi, shy, shz) : pred1(c, xi, yi, zi, hi));
                }
                return m;
            };
            if (!g.shifted) {
                for (int t = wn4 - 4; t >= 32; t -= 4) mhi = group(std::false_type{}, mhi, t);
                for (int t = n0 - 4; t >= 0; t -= 4) mlo = group(std::false_type{}, mlo, t);
            } else {
                for (int t = wn4 - 4; t >= 32; t -= 4) mhi = group(std::true_type{}, mhi, t);
                for (int t = n0 - 4; t >= 0; t -= 4) mlo = group(std::true_type{}, mlo, t);
            }
            const unsigned long long raw = ((unsigned long long)mhi << 32) | mlo;
            mask = shift >= 64 ? 0ull : (raw >> shift) & valid;
        }
        if (KIND == 0 && use_masks && slot < kMaskSlots) mrow[(size_t)slot * 64] = mask;
        ++slot;
        return mask;
    };
    // Pair arithmetic of one hit at staged slot u; gi = the candidate's sorted index.
    auto hit = [&](auto shifted_tag, int u, double shy, double shz, int sy, int sz, int gi) {
        constexpr bool kSh = decltype(shifted_tag)::value;
        const double dx = L.x[u] - xi;
        if constexpr (KIND == 1 && sizeof(R) == 8) {
            if (pband != nullptr && L.ind[u] == 0) {
                // FP64-band wave, active neighbour outside the band: the FP32 term exactly as
                // the neighbour's FP32 wave evaluates its mirror (fixed-point dy, dz, FP32
                // fields), so the pair stays antisymmetric bit for bit
                const float fx = (float)dx;
                float fy, fz, dd;
                if (!far_mirror) {
                    fy = (float)fix_diff(fix_coord(L.Y(u), T.o[1], ups_y), Yi, sy) * qy;
                    fz = (float)fix_diff(fix_coord(L.Z(u), T.o[2], ups_z), Zi, sz) * qz;
                    dd = __builtin_fmaf(fz, fz, __builtin_fmaf(fy, fy, fx * fx));
                } else {
                    const double dy = kSh ? (L.Y(u) - yi) + shy : L.Y(u) - yi;
                    const double dz = kSh ? (L.Z(u) - zi) + shz : L.Z(u) - zi;
                    fy = (float)dy;
                    fz = (float)dz;
                    dd = (float)__builtin_fma(dz, dz, __builtin_fma(dy, dy, dx * dx));
                }
                if (dd > 0.0f) {
                    float coef, edc;
                    force_terms<float>(dd, 0.5f * (hif + (float)L.H(u)), fx, fy, fz, (float)vxi - (float)L.VX(u),
                                       (float)vyi - (float)L.VY(u), (float)vzi - (float)L.VZ(u), (float)rhoi_d,
                                       (float)L.RHO(u), (float)pr2i_d, (float)L.PR2(u), (float)csi_d, (float)L.CS(u),
                                       (float)mi_d, (float)L.M(u), (float)T.alpha, (float)T.beta, (float)T.eps, plain,
                                       coef, edc);
                    a0.add(fx * coef);
                    a1.add(fy * coef);
                    a2.add(fz * coef);
                    a3.add(edc);
                }
                return;
            }
        }
        if constexpr (kFix) {
            if ((L.ind[u] & 1) != 0) {
                // a frozen neighbour has no partner term: FP64 keeps the momentum it injects
                // exact to rounding (its y, z, h and the target's values are read from memory;
                // the empty asm keeps their addresses from being hoisted out of the loops)
                int qq = q;
                asm volatile("" : "+v"(qq));
                const long long g = wo + gi, gq = wo + qq;
                const double dy = kSh ? (B.y[g] - B.y[gq]) + shy : B.y[g] - B.y[gq];
                const double dz = kSh ? (B.z[g] - B.z[gq]) + shz : B.z[g] - B.z[gq];
                const double dd = __builtin_fma(dz, dz, __builtin_fma(dy, dy, dx * dx));
                if (dd > 0.0) {
                    double coef, edc;
                    force_terms<double>(dd, 0.5 * (B.h[gq] + B.h[g]), dx, dy, dz, B.vx[gq] - B.vx[g],
                                        B.vy[gq] - B.vy[g], B.vz[gq] - B.vz[g], B.rho[gq], B.rho[g], pr2[qq],
                                        pr2[gi], csav[qq], csav[gi], B.m[gq], B.m[g], T.alpha, T.beta, T.eps,
                                        plain, coef, edc);
                    d0 += dx * coef;
                    d1 += dy * coef;
                    d2s += dz * coef;
                    d3 += edc;
                }
            } else {
                // dx from FP64 positions, dy and dz from fixed-point differences (the periodic
                // image is sy, sz box lengths); every term is antisymmetric in (i, j)
                const float fx = (float)dx;
                float fy, fz, dd;
                if (!far_piece) {
                    fy = (float)fix_diff(L.Y(u), Yi, sy) * qy;
                    fz = (float)fix_diff(L.Z(u), Zi, sz) * qz;
                    dd = __builtin_fmaf(fz, fz, __builtin_fmaf(fy, fy, fx * fx));
                } else {
                    int qq = q;
                    asm volatile("" : "+v"(qq));
                    const long long g = wo + gi, gq = wo + qq;
                    const double dy = kSh ? (B.y[g] - B.y[gq]) + shy : B.y[g] - B.y[gq];
                    const double dz = kSh ? (B.z[g] - B.z[gq]) + shz : B.z[g] - B.z[gq];
                    fy = (float)dy;
                    fz = (float)dz;
                    dd = (float)__builtin_fma(dz, dz, __builtin_fma(dy, dy, dx * dx));
                }
                if (dd > 0.0f) {
                    R coef, edc;
                    force_terms<R>(dd, 0.5f * (hif + L.H(u)), fx, fy, fz, vxi_r - L.VX(u), vyi_r - L.VY(u),
                                   vzi_r - L.VZ(u), rhoi, L.RHO(u), pr2i, L.PR2(u), csi, L.CS(u), mi, L.M(u),
                                   alpha, beta, epsv, plain, coef, edc);
                    a0.add(fx * coef);
                    a1.add(fy * coef);
                    a2.add(fz * coef);
                    a3.add(edc);
                }
            }
            return;
        }
        if constexpr (KIND == 0) {
            // every LDS read of the hit is issued before the arithmetic waits on any
            const auto hm = L.hm[u];
            const double dy = kSh ? (L.Y(u) - yi) + shy : L.Y(u) - yi;
            const double dz = kSh ? (L.Z(u) - zi) + shz : L.Z(u) - zi;
            const double dd = __builtin_fma(dz, dz, __builtin_fma(dy, dy, dx * dx));
            const R r = sqrt_r<R>((R)dd);
            const R ih = rcp<R>(hhi + hm.hh);
            a0.add(hm.mw * (ih * ih * ih) * spline_w_hit<R>(r * ih));
            if (kStats) {
                ++cnt;
                hash += pair_tag(L.id[u], sy, sz);
            }
            return;
        }
        const double dy = kSh ? (L.Y(u) - yi) + shy : L.Y(u) - yi;
        const double dz = kSh ? (L.Z(u) - zi) + shz : L.Z(u) - zi;
        const double dd = __builtin_fma(dz, dz, __builtin_fma(dy, dy, dx * dx));
        if (dd > 0.0) {
            const double hhd = 0.5 * (hi + L.H(u));
            if (sizeof(R) == 4 && (L.ind[u] & 1) != 0) {
                // a frozen neighbour has no partner term: FP64 keeps the momentum it
                // injects exact to rounding
                // (the target's FP64 values are read again here rather than kept in registers;
                // the empty asm keeps their addresses from being hoisted out of the loops)
                int qq = q;
                asm volatile("" : "+v"(qq));
                const long long g = wo + gi, gq = wo + qq;
                double coef, edc;
                force_terms<double>(dd, hhd, dx, dy, dz, B.vx[gq] - B.vx[g], B.vy[gq] - B.vy[g],
                                    B.vz[gq] - B.vz[g], B.rho[gq], B.rho[g], pr2[qq], pr2[gi], csav[qq], csav[gi],
                                    B.m[gq], B.m[g], T.alpha, T.beta, T.eps, plain, coef, edc);
                d0 += dx * coef;
                d1 += dy * coef;
                d2s += dz * coef;
                d3 += edc;
            } else {
                R coef, edc;
                const R fx = (R)dx, fy = (R)dy, fz = (R)dz;
                force_terms<R>((R)dd, (R)hhd, fx, fy, fz, vxi_r - L.VX(u), vyi_r - L.VY(u),
                               vzi_r - L.VZ(u), rhoi, L.RHO(u), pr2i, L.PR2(u), csi, L.CS(u), mi, L.M(u), alpha,
                               beta, epsv, plain, coef, edc);
                a0.add(fx * coef);
                a1.add(fy * coef);
                a2.add(fz * coef);
                a3.add(edc);
            }
        }
    };
    // Hits of one window whose slot 0 is staged slot o and whose candidates start at sorted
    // index gb + o.
    auto hits1 = [&](auto shifted_tag, unsigned long long m, int o, int gb, const RowSeg& g) {
        while (m) {
            const int t = __ffsll((long long)m) - 1;
            m &= m - 1ull;
            hit(shifted_tag, o + t, g.shy, g.shz, g.sy, g.sz, gb + o + t);
        }
    };
    // Hits of two windows: a lane takes all of its hits of the first (candidates from sorted
    // index gA + staged slot), then of the second (gB + slot), so the loop runs max over
    // lanes of the total, not the sum of the two maxima.  Per lane: the first window's hits
    // in candidate order, then the second's.
    auto hits2 = [&](auto shifted_tag, unsigned long long mA, int oA, int gA, const RowSeg& A,
                     unsigned long long mB, int oB, int gB, const RowSeg& B2) {
        bool second = mA == 0ull;
        unsigned long long m = second ? mB : mA;
        int o = second ? oB : oA;
        while (m) {
            const int t = __ffsll((long long)m) - 1;
            m &= m - 1ull;
            const int u = o + t;
            hit(shifted_tag, u, second ? B2.shy : A.shy, second ? B2.shz : A.shz, second ? B2.sy : A.sy,
                second ? B2.sz : A.sz, (second ? gB : gA) + u);
            if (m == 0ull && !second) {
                second = true;
                m = mB;
                o = oB;
            }
        }
    };
    // This lane's x-window inside staged slots [a, e) of one row: rows are sorted by x (cells
    // ascend in x, a cell by x then id), so the candidates with -rch < x_j - x_i < rch (the
    // differences rounded as the predicate rounds them) are one run.  Any other candidate has
    // d2 >= dx^2 >= rch^2 > (h_i + h_j)^2 and cannot be a neighbour.  Returns the run as
    // (start, length); two binary searches on the staged x over [a, e), run side by side so
    // that each step's two LDS reads are in flight together (the searches are chains of
    // dependent reads; run one after the other they cost twice the round trips).
    auto xwindow = [&](int a, int e, int& len) {
        if (e <= a) {
            len = 0;
            return a;
        }
        int lo = a, nlo = e - a, hi = a, nhi = e - a;
        while (nlo > 0 || nhi > 0) {
            const int ml = lo + (nlo >> 1), mh = hi + (nhi >> 1);
            const double vl = L.x[min(ml, e - 1)], vh = L.x[min(mh, e - 1)];
            if (nlo > 0) {
                if (vl - xi <= -rch) {
                    lo = ml + 1;
                    nlo -= (nlo >> 1) + 1;
                } else {
                    nlo >>= 1;
                }
            }
            if (nhi > 0) {
                if (vh - xi < rch) {
                    hi = mh + 1;
                    nhi -= (nhi >> 1) + 1;
                } else {
                    nhi >>= 1;
                }
            }
        }
        len = hi - lo;
        return lo;
    };
    // The x-window of this lane in one row segment of the current piece (staged slots
    // [a0, e0)): searched by the density pass and stored for gas chunks, read back by the
    // force pass.
    auto seg_window_range = [&](int a0, int e0, int& len) {
        int a;
        if (KIND == 1 && use_masks && segi < kMaskSlots) {
            const unsigned int w = segi == pw_at ? pw[0] : (segi == pw_at + 1 ? pw[1] : wrow[(size_t)segi * 64]);
            a = (int)(w >> 16);
            len = (int)(w & 0xffffu);
        } else {
            if (KIND == 1) win_own = true;
            a = xwindow(a0, e0, len);
            if (KIND == 0 && use_masks && segi < kMaskSlots) wrow[(size_t)segi * 64] = ((unsigned)a << 16) | (unsigned)len;
        }
        ++segi;
        return a;
    };
    // One visit: row A alone, in pieces of kTileP, or rows A and B2 staged together as one
    // piece (pair; only when both fit).  Per piece each row is a segment of staged slots; a
    // lane's x-window in a segment is cut into windows of 64 candidates, and window masks
    // take slots in the order A's windows, then B2's.  When each segment has at most one
    // window, one hit loop takes both (a lane runs its A hits, then its B2 hits); otherwise
    // one window per iteration.  Either way a target sums its hits in the order A then B2,
    // candidates in sorted order, so sums do not depend on the piece size or on whether a
    // pair fitted.  The pair kernels' large bodies (predicate windows, force hits) each
    // have one call site here.
    auto visit = [&](const RowSeg& A, const RowSeg& B2, bool pair) {
        const int nA = A.ue - A.ub, nB = pair ? B2.ue - B2.ub : 0;
        // a pair is one piece even when A's range is empty
        const int npieces = pair ? 1 : (nA + kTileP - 1) / kTileP;
        for (int ip = 0; ip < npieces; ++ip) {
            const int pb = A.ub + ip * kTileP;
            const int pnA = pair ? nA : min(kTileP, A.ue - pb);
            stage(pb, pnA, B2.ub, nB);
            int lenA, lenB = 0;
            const int aA = seg_window_range(max(A.jl, pb) - pb, min(A.jh, pb + pnA) - pb, lenA);
            const int aB =
                pair ? seg_window_range(nA + max(B2.jl, B2.ub) - B2.ub, nA + min(B2.jh, B2.ue) - B2.ub, lenB) : 0;
            // lengths are within one piece (<= 256 candidates)
            const int mlA = wave_max_u9(lenA);
            const int mlB = pair ? wave_max_u9(lenB) : 0;
            const int winA = (mlA + 63) >> 6, winB = (mlB + 63) >> 6;
            const bool both = winA <= 1 && winB <= 1;
            const int nit = both ? 1 : winA + winB;
            for (int k = 0; k < nit; ++k) {
                unsigned long long mA = 0ull, mB = 0ull;
                int oA = aA, oB = aB;
                // the window of segment `second` taken at iteration k, if any
                auto seg_window = [&](bool second) {
                    const int w = both ? 0 : (second ? k - winA : k);
                    const bool use = both ? (second ? winB : winA) > 0 : (second ? k >= winA : k < winA);
                    if (!use) return;
                    const int o = (second ? aB : aA) + 64 * w;
                    const unsigned long long m =
                        window_mask(o, (second ? lenB : lenA) - 64 * w, (second ? mlB : mlA) - 64 * w,
                                    second ? B2 : A, second ? B2.ub - nA : pb);
                    if (second) {
                        mB = m;
                        oB = o;
                    } else {
                        mA = m;
                        oA = o;
                    }
                };
                if (KIND == 0) {
                    // small bodies: specialised copies
                    seg_window(false);
                    if (pair) seg_window(true);
                    if (!pair) {
                        if (A.shifted) hits1(std::true_type{}, mA, oA, pb, A);
                        else hits1(std::false_type{}, mA, oA, pb, A);
                    } else if (A.shifted || B2.shifted) {
                        hits2(std::true_type{}, mA, oA, pb, A, mB, oB, B2.ub - nA, B2);
                    } else {
                        hits2(std::false_type{}, mA, oA, pb, A, mB, oB, B2.ub - nA, B2);
                    }
                } else {
                    // one copy of the large bodies
#pragma clang loop unroll(disable)
                    for (int sg2 = 0; sg2 < 2; ++sg2) seg_window(sg2 == 1);
                    hits2(std::true_type{}, mA, oA, pb, A, mB, oB, B2.ub - nA, B2);
                }
            }
        }
    };

    bool pending = active;
    while (true) {
        const unsigned long long pm = __ballot(pending);
        if (pm == 0ull) break;
        const int leader = __ffsll((long long)pm) - 1;
        const int row0 = uni(__shfl(row, leader, 64));
        const bool mine = pending && row == row0;
        // targets are sorted by cell, so along the lanes of one x-row kx does not decrease
        const unsigned long long mm = __ballot(mine);
        const int cxa = __builtin_amdgcn_readlane(kx, __ffsll((long long)mm) - 1);
        const int cxb = __builtin_amdgcn_readlane(kx, 63 - __builtin_clzll(mm));
        const double rmax = wave_max(mine ? rch : 0.0);
        // x extent of this row's targets: a candidate with x_j - xmin <= -rmax or
        // x_j - xmax >= rmax (as the window searches round them) is in no lane's x-window
        const double xmin = uni(wave_min(mine ? xi : __builtin_inf()));
        const double xmax = uni(wave_max(mine ? xi : -__builtin_inf()));
        const int Rx = uni((int)ceil(rmax * T.ic[0]));
        const int Ry = uni((int)ceil(rmax * T.ic[1]));
        const int Rz = uni((int)ceil(rmax * T.ic[2]));
        const int ky0 = row0 % ny, kz0 = row0 / ny;
        const int nry = 2 * Ry + 1, nrows = nry * (2 * Rz + 1);
        auto row_seg = [&](int r) {
            RowSeg g{};
            g.on = false;
            const int oy = r % nry - Ry, oz = r / nry - Rz;
            int cz = kz0 + oz, sz = 0;
            if (T.per[2]) {
                sz = floordiv(cz, nz);
                cz -= sz * nz;
            } else if (cz < 0 || cz >= nz) {
                return g;
            }
            int cy = ky0 + oy, sy = 0;
            if (T.per[1]) {
                sy = floordiv(cy, ny);
                cy -= sy * ny;
            } else if (cy < 0 || cy >= ny) {
                return g;
            }
            const bool need = mine && abs(oz) <= rzl && abs(oy) <= ryl;
            if (__ballot(need) == 0ull) return g;
            const int rowbase = cell_off + nx * (cy + ny * cz);
            g.on = true;
            g.sy = sy;
            g.sz = sz;
            g.shy = (double)sy * T.L[1];
            g.shz = (double)sz * T.L[2];
            g.shifted = (sy != 0) || (sz != 0);
            g.ub = uni(start[rowbase + max(cxa - Rx, 0)]);
            g.ue = uni(start[rowbase + min(cxb + Rx, nx - 1) + 1]);
            if (g.ue > g.ub) {
                // Trim the row's ends (whole cells of the reach) to the candidates some lane
                // can reach, up to 64 at each end: x ascends along the row, so the
                // unreachable candidates at each end are a prefix / suffix, counted by ballot.
                // Smaller rows let more mirror pairs share a piece; the sets and the order of
                // every target's hits are unchanged (only candidates outside every window go).
                const int klo = g.ub + lane, khi = g.ue - 64 + lane;
                const bool vlo = klo < g.ue, vhi = khi >= g.ub;
                const double xl = B.x[wo + (vlo ? klo : g.ub)], xh = B.x[wo + (vhi ? khi : g.ub)];
                const int cut_lo = __popcll(__ballot(vlo && xl - xmin <= -rmax));
                const int cut_hi = __popcll(__ballot(vhi && xh - xmax >= rmax));
                g.ub += cut_lo;
                g.ue -= cut_hi;
            }
            g.jl = 0;
            g.jh = 0;
            if (need) {
                g.jl = start[rowbase + max(kx - rxl, 0)];
                g.jh = start[rowbase + min(kx + rxl, nx - 1) + 1];
            }
            return g;
        };
        // rows in mirror pairs (r, nrows-1-r): the hits of a target in a row and in its
        // mirror row add up to nearly the same count for every lane
        for (int v = 0; v < (nrows + 1) / 2; ++v) {
            const RowSeg A = row_seg(v);
            const int w = nrows - 1 - v;
            const RowSeg B2 = w != v ? row_seg(w) : RowSeg{};
            // one call site of visit: the pair, or A and then B2 alone
            const bool pair = A.on && B2.on && (A.ue - A.ub) + (B2.ue - B2.ub) <= kTileP;
#pragma clang loop unroll(disable)
            for (int k = 0; k < (pair ? 1 : 2); ++k) {
                const bool first = k == 0;
                if (!(first ? A.on : B2.on)) continue;
                visit(first ? A : B2, first ? B2 : A, pair);
            }
        }
        pending = pending && !mine;
    }

    if (!valid) return;
    if (KIND == 0) {
        double rho = B.rho[wo + q];
        if (active) {
            rho = a0.get();
            B.rho[wo + q] = rho;
            if (s == 0) {
                B.p[wo + q] = (T.eos == 0) ? rho * B.e[wo + q] * (T.gamma - 1) : T.cs * T.cs * rho;
            } else {
                if (T.drag_mode == 1) B.ts[wo + q] = rho / T.K[s - 1];
            }
        }
        if (kStats) {
            ncount[q] = active ? cnt : 0;
            nhash[q] = active ? hash : 0ull;
        }
        if (s == 0) {
            const double p = B.p[wo + q];
            const double pr = p / (rho * rho), cs = sqrt(T.gamma * p / rho);
            pr2[q] = pr;
            csav[q] = cs;
            if (frec != nullptr && T.precision == DSH_MIXED) {
                // the FP32 force pass's staged values of this particle, converted once
                ForceRec r;
                r.x = xi;
                r.y = fix_coord(yi, T.o[1], ups_y);
                r.z = fix_coord(zi, T.o[2], ups_z);
                r.h = (float)hi;
                r.vx = (float)B.vx[wo + q];
                r.vy = (float)B.vy[wo + q];
                r.vz = (float)B.vz[wo + q];
                r.m = (float)B.m[wo + q];
                r.rho = (float)rho;
                r.pr2 = (float)pr;
                r.cs = (float)cs;
                r.flags = (B.ind[wo + q] != 0 ? 1u : 0u) | (pband != nullptr && pband[q] != 0 ? 2u : 0u) |
                          (fix_range_ok(yi, zi) ? 4u : 0u);
                r.pad[0] = r.pad[1] = r.pad[2] = 0u;
                frec[q] = r;
            }
        }
    } else {
        B.ax[wo + q] = active ? a0.get() + d0 : 0.0;
        B.ay[wo + q] = active ? a1.get() + d1 : 0.0;
        B.az[wo + q] = active ? a2.get() + d2s : 0.0;
        B.ed[wo + q] = active ? a3.get() + d3 : 0.0;
    }
}

// Block b of the part [p0, p1) of a grid, remapped so that each XCD (which takes the blocks
// b with b % 8 equal to its index) runs one contiguous stretch of the part.
__device__ __forceinline__ int part_remap(int b, int p0, int p1) {
    const int x = b & 7;
    int before = 0;
#pragma unroll
    for (int y = 0; y < 8; ++y) {
        const int first = p0 + ((y - p0) & 7);
        const int cnt = first < p1 ? (p1 - first + 7) >> 3 : 0;
        before += y < x ? cnt : 0;
    }
    const int first = p0 + ((x - p0) & 7);
    return p0 + before + ((b - first) >> 3);
}

// Every chunk of the map, kWaves per block (wave w of block b takes chunk b * kWaves + w;
// blocks remapped so each XCD takes a contiguous run).  With perm (a batch of several tubes):
// the running tubes' chunks only, perm[0, perm[nall + 2]) (gas map: perm[nall + 3]), dealt to
// the XCDs the same way, so they spread over every XCD however many tubes have finished.
template <typename R, int KIND, bool kStats, int kTileP, int kWaves>
__global__ __launch_bounds__(64 * kWaves) __attribute__((amdgpu_waves_per_eu(KIND == 0 ? 5 : 4))) void k_pairs_tiled(
    Map mp, const TubeConst* __restrict__ tc, const TubeDyn* __restrict__ td, const int* __restrict__ cur, PBuf B,
    const int* __restrict__ start, const int* __restrict__ wkey, const double* __restrict__ reach,
    double* __restrict__ pr2, double* __restrict__ csav, int* __restrict__ ncount,
    unsigned long long* __restrict__ nhash, unsigned long long* __restrict__ masks, int mask_chunks,
    const unsigned char* __restrict__ pband, ForceRec* __restrict__ frec, const int* __restrict__ perm, int nall) {
    __shared__ TileLds<R, KIND, kStats, kTileP> L[kWaves];
    const int wave = kWaves == 1 ? 0 : uni((int)(threadIdx.x >> 6));
    int b;
    if (perm != nullptr) {
        const int nrun = KIND == 0 ? uni(perm[nall + 2]) : uni(perm[nall + 3]);
        const int br = (nrun + kWaves - 1) / kWaves;
        if ((int)blockIdx.x >= br) return;
        const int e = part_remap(blockIdx.x, 0, br) * kWaves + wave;
        if (e >= nrun) return;
        b = uni(perm[e]);
    } else {
        b = xcd_remap(blockIdx.x, gridDim.x) * kWaves + wave;
        if (kWaves > 1 && b >= mp.nblocks) return;
    }
    pairs_chunk<R, KIND, kStats, kTileP, kWaves>(L[wave], b, mp, tc, td, cur, B, start, wkey, reach, pr2, csav,
                                                 ncount, nhash, masks, mask_chunks, pband, frec);
}

// Pieces of the FP64 wall-band path inside the fused kernels: the largest of 128, 96, 64
// candidates whose FP64 force tile fits in the FP32 force tile's LDS (its records are about
// twice the FP32 ones).  The density pass uses the same size: the force pass replays the
// density pass's stored hit masks, whose slots follow the piece layout.
template <int kTileP>
constexpr int band_tile() {
    constexpr size_t f32 = sizeof(TileLds<float, 1, false, kTileP>);
    if (sizeof(TileLds<double, 1, false, 128>) <= f32) return 128;
    if (sizeof(TileLds<double, 1, false, 96>) <= f32) return 96;
    return 64;
}

// Mixed precision with an FP64 wall band, both paths in one launch: perm (built by
// k_perm_write this step) lists the band's chunks first, then the other chunks of running
// tubes, each part in ascending chunk order; finished tubes' chunks come last and get no
// block.  Blocks below the band count run the FP64 path on band chunks and are dispatched
// first, so the band's slower waves overlap the FP32 ones instead of trailing them; the next
// blocks run the FP32 path.  For the gas force map (KIND 1) the parts are the gas chunks of
// each (perm[nall + 1] and perm[nall + 3] of them) and the FP32 part starts at perm[nband_all].
// The two paths share one LDS tile per wave.  With kWaves waves per block, the band part
// takes ceil(nband / kWaves) whole blocks (wave w of part block p: band chunk p * kWaves + w)
// and the FP32 part the ceil(nrest / kWaves) blocks after them, each part dealt to the XCDs in
// contiguous stretches (so running chunks spread over every XCD however many tubes have
// finished); waves past a part's end leave at once.  The grid, ceil(nchunks / kWaves) + 1
// blocks, covers both parts whatever the counts; the blocks past them leave at once.
template <int KIND, bool kStats, int kTileP, int kWaves>
__global__ __launch_bounds__(64 * kWaves) __attribute__((amdgpu_waves_per_eu(KIND == 0 ? 5 : 4))) void k_pairs_fused(
    Map mp, const TubeConst* __restrict__ tc, const TubeDyn* __restrict__ td, const int* __restrict__ cur, PBuf B,
    const int* __restrict__ start, const int* __restrict__ wkey, const double* __restrict__ reach,
    double* __restrict__ pr2, double* __restrict__ csav, int* __restrict__ ncount,
    unsigned long long* __restrict__ nhash, unsigned long long* __restrict__ masks, int mask_chunks,
    const unsigned char* __restrict__ pband, const int* __restrict__ perm, int nall, ForceRec* __restrict__ frec) {
    constexpr int kBandTile = band_tile<kTileP>();
    using L32 = TileLds<float, KIND, kStats, kTileP>;
    using L64 = TileLds<double, KIND, kStats, kBandTile>;
    constexpr size_t kTileBytes = ((sizeof(L32) > sizeof(L64) ? sizeof(L32) : sizeof(L64)) + 15) / 16 * 16;
    __shared__ __align__(16) unsigned char raw[kWaves][kTileBytes];
    const int wave = kWaves == 1 ? 0 : uni((int)(threadIdx.x >> 6));
    const int band_all = uni(perm[nall]);
    const int nband = KIND == 0 ? band_all : uni(perm[nall + 1]);
    const int nrest = KIND == 0 ? uni(perm[nall + 2]) : uni(perm[nall + 3]);
    const int bb = (nband + kWaves - 1) / kWaves;
    const int br = (nrest + kWaves - 1) / kWaves;
    const int b = blockIdx.x;
    if (b >= bb + br) return;   // blocks past both parts (finished tubes' chunks, padding)
    if (b < bb) {
        const int e = part_remap(b, 0, bb) * kWaves + wave;
        if (kWaves > 1 && e >= nband) return;
        pairs_chunk<double, KIND, kStats, kBandTile, kWaves>(*reinterpret_cast<L64*>(raw[wave]), uni(perm[e]), mp, tc,
                                                             td, cur, B, start, wkey, reach, pr2, csav, ncount, nhash,
                                                             masks, mask_chunks, pband, frec);
    } else {
        const int e = (part_remap(b, bb, bb + br) - bb) * kWaves + wave;
        if (e >= nrest) return;
        pairs_chunk<float, KIND, kStats, kTileP, kWaves>(*reinterpret_cast<L32*>(raw[wave]), uni(perm[band_all + e]),
                                                         mp, tc, td, cur, B, start, wkey, reach, pr2, csav, ncount,
                                                         nhash, masks, mask_chunks, pband, frec);
    }
}

// kernel.cu:2310-2437 for one component; n_dust 1 or 2, reference operation order.
__device__ __forceinline__ void idic_solve(int nd, double vg, const double* vd, double psi, const double* tb,
                                           const double* ep, double tau, double* vgh, double* vdh) {
    if (nd == 1) {
        double y = vg;
        const double x0 = vg - vd[0];
        y = y + ep[0] * vd[0];
        const double b0 = (tb[0] + tau) / (ep[0] * tau);
        double beta = 1.0;
        beta = beta + 1.0 / b0;
        const double yh = y + tau * psi;
        double xh0 = -1.0 * b0 * beta * (x0 + tau * psi) / (b0 * b0);
        xh0 = xh0 + (x0 + tau * psi) / (b0 * b0);
        xh0 = xh0 * (-1.0 * tb[0]) / (tau * ep[0] * beta);
        double S = 1.0;
        S = S + ep[0];
        double g = yh;
        g = g + ep[0] * xh0;
        double d0 = yh;
        d0 = d0 - (S - ep[0]) * xh0;
        d0 = d0 / S;
        g = g / S;
        *vgh = g;
        vdh[0] = d0;
        return;
    }
    double y = vg;
    const double x0 = vg - vd[0];
    const double x1 = vg - vd[1];
    y = y + ep[0] * vd[0] + ep[1] * vd[1];
    const double b0 = (tb[0] + tau) / (ep[0] * tau);
    const double b1 = (tb[1] + tau) / (ep[1] * tau);
    double beta = 1.0;
    beta = beta + 1.0 / b0 + 1.0 / b1;
    const double yh = y + tau * psi;
    double xh0 = -1.0 * b0 * beta * (x0 + tau * psi) / (b0 * b0);
    xh0 = xh0 + (x0 + tau * psi) / (b0 * b0) + (x1 + tau * psi) / (b0 * b1);
    xh0 = xh0 * (-1.0 * tb[0]) / (tau * ep[0] * beta);
    double xh1 = -1.0 * b1 * beta * (x1 + tau * psi) / (b1 * b1);
    xh1 = xh1 + (x0 + tau * psi) / (b1 * b0) + (x1 + tau * psi) / (b1 * b1);
    xh1 = xh1 * (-1.0 * tb[1]) / (tau * ep[1] * beta);
    double S = 1.0;
    S = S + ep[0] + ep[1];
    double g = yh;
    g = g + ep[0] * xh0;
    double d0 = yh;
    d0 = d0 + ep[1] * xh1;
    d0 = d0 - (S - ep[0]) * xh0;
    d0 = d0 / S;
    g = g + ep[1] * xh1;
    double d1 = yh;
    d1 = d1 + ep[0] * xh0;
    d1 = d1 - (S - ep[1]) * xh1;
    d1 = d1 / S;
    g = g / S;
    *vgh = g;
    vdh[0] = d0;
    vdh[1] = d1;
}

// Visits the particles of species s in averaging cell (ax, ay, az): its f*f rows of f
// neighbour cells are contiguous runs of the sorted order, walked as one run so that the 64
// lanes stay busy across row ends.  fn(q) gets the sorted index.
template <class F>
__device__ __forceinline__ void for_avg_cell(const TubeConst& T, const int* __restrict__ start, int s, int ax,
                                             int ay, int az, int lane, F&& fn) {
    const int f = T.f, nx = T.n[0], ny = T.n[1], nr = f * f;
    auto row_lo = [&](int r) { return T.cell_off[s] + ax * f + nx * ((ay * f + r % f) + ny * (az * f + r / f)); };
    int r = 0, c0 = row_lo(0);
    int lo = start[c0], len = start[c0 + f] - lo, off = 0;
    for (int k = lane;; k += 64) {
        while (k - off >= len) {
            off += len;
            if (++r >= nr) return;
            c0 = row_lo(r);
            lo = start[c0];
            len = start[c0 + f] - lo;
        }
        fn(lo + (k - off));
    }
}

// Sums over one averaging cell: gas count, v (3), A (3); per dust species count, v (3), t_stop.
struct CellSums {
    int ng;
    double sg[6];
    int ndc[kMaxDust];
    double sd[kMaxDust][4];
};
// What the per-particle update of one averaging cell needs.
struct CellDrag {
    double psi[3], vgh[3], vdh[kMaxDust][3], tb[kMaxDust], tmp1, tmpv[3];
};

// Means, fallbacks (kernel.cu:2126-2211), the implicit solve (kernel.cu:2294-2437) and the
// gas velocity terms (IDIC.h:301-309) of one averaging cell.
__device__ __forceinline__ CellDrag cell_drag(const TubeConst& T, int nd, double tau, const CellSums& S) {
    CellDrag R;
    double vg[3], vdm[kMaxDust][3], ep[kMaxDust];
    for (int a = 0; a < 3; ++a) {
        vg[a] = S.ng > 0 ? S.sg[a] / S.ng : S.sg[a];
        R.psi[a] = S.ng > 0 ? S.sg[3 + a] / S.ng : S.sg[3 + a];
    }
    // loops over dust species unrolled so that the arrays stay in registers
#pragma unroll
    for (int k = 0; k < kMaxDust; ++k) {
        vdm[k][0] = vdm[k][1] = vdm[k][2] = 0.0;
        R.tb[k] = 1.0;
        ep[k] = 0.0;
        if (k >= nd) continue;
        for (int a = 0; a < 3; ++a) vdm[k][a] = S.ndc[k] > 0 ? S.sd[k][a] / S.ndc[k] : S.sd[k][a];
        R.tb[k] = S.ndc[k] > 0 ? S.sd[k][3] / S.ndc[k] : T.t_fb;
        const int test = (k == 1 && T.eps2_quirk) ? S.ndc[0] : S.ndc[k];
        ep[k] = (S.ng > 0 && test > 0) ? S.ndc[k] * T.mu[k] / S.ng : T.eps_fb[k];
    }
    for (int a = 0; a < 3; ++a) {
        R.vgh[a] = 0.0;
        for (int k = 0; k < kMaxDust; ++k) R.vdh[k][a] = 0.0;
    }
    if (nd > 0)
        for (int a = 0; a < 3; ++a) {
            double vdc[kMaxDust] = {vdm[0][a], nd > 1 ? vdm[1][a] : 0.0};
            double out[kMaxDust] = {0.0, 0.0};
            idic_solve(nd, vg[a], vdc, R.psi[a], R.tb, ep, tau, &R.vgh[a], out);
#pragma unroll
            for (int k = 0; k < kMaxDust; ++k)
                if (k < nd) R.vdh[k][a] = out[k];
        }
    R.tmp1 = 0.0;
    R.tmpv[0] = R.tmpv[1] = R.tmpv[2] = 0.0;
    if (nd == 1) {
        R.tmp1 = R.tmp1 + ep[0] / R.tb[0];
        for (int a = 0; a < 3; ++a) R.tmpv[a] = R.tmpv[a] + ep[0] * R.vdh[0][a] / R.tb[0];
    } else if (nd == 2) {
        R.tmp1 = R.tmp1 + ep[0] / R.tb[0] + ep[1] / R.tb[1];
        for (int a = 0; a < 3; ++a)
            R.tmpv[a] = R.tmpv[a] + ep[0] * R.vdh[0][a] / R.tb[0] + ep[1] * R.vdh[1][a] / R.tb[1];
    }
    // gas only: the explicit kick of MoveKernel in particle_update
    return R;
}

// Kick (implicit drag, IDIC.h:264-315, or the explicit v += A tau of a gas-only run), drift
// with periodic wrap, energy, smoothing length (IDIC.h:319-346) of sorted particle g of
// species s, and its conservation / CFL contributions.
__device__ __forceinline__ void particle_update(const TubeConst& T, const PBuf& B, long long g, int s, int nd,
                                                double tau, const CellDrag& C, double& px, double& py, double& pz,
                                                double& E, double& mn) {
    // every input is read before the first store: the stores below could alias any field as
    // far as the compiler knows, so loads left after them would each wait for a round trip
    const bool gas_kick = s == 0 && nd == 0;
    double vx = B.vx[g], vy = B.vy[g], vz = B.vz[g];
    const double m = B.m[g];
    const bool active = B.ind[g] == 0;
    const double ax = gas_kick ? B.ax[g] : 0.0, ay = gas_kick ? B.ay[g] : 0.0, az = gas_kick ? B.az[g] : 0.0;
    const double x0 = B.x[g], y0 = B.y[g], z0 = B.z[g];
    const double e0 = s == 0 ? B.e[g] : 0.0, ed = s == 0 ? B.ed[g] : 0.0;
    const double h0 = B.h[g];
    const double rho = T.h_mode == 1 ? B.rho[g] : 1.0;
    double e = 0.0;
    if (active) {
        if (gas_kick) {
            // gas only: v += A tau (MoveKernel, kernel.cu:205-228)
            vx = vx + ax * tau;
            vy = vy + ay * tau;
            vz = vz + az * tau;
        } else if (s == 0) {
            vx = (vx / tau + C.tmpv[0] + C.psi[0]) / (1.0 / tau + C.tmp1);
            vy = (vy / tau + C.tmpv[1] + C.psi[1]) / (1.0 / tau + C.tmp1);
            vz = (vz / tau + C.tmpv[2] + C.psi[2]) / (1.0 / tau + C.tmp1);
        } else {
            const double tk = C.tb[s - 1];
            vx = (vx / tau + C.vgh[0] / tk) / (1.0 / tau + 1.0 / tk);
            vy = (vy / tau + C.vgh[1] / tk) / (1.0 / tau + 1.0 / tk);
            vz = (vz / tau + C.vgh[2] / tk) / (1.0 / tau + 1.0 / tk);
        }
        B.vx[g] = vx;
        B.vy[g] = vy;
        B.vz[g] = vz;
        double x = x0 + vx * tau;
        double y = y0 + vy * tau;
        double z = z0 + vz * tau;
        if (T.per[1]) y = wrap(y, T.o[1], T.L[1], T.hi[1]);
        if (T.per[2]) z = wrap(z, T.o[2], T.L[2], T.hi[2]);
        B.x[g] = x;
        B.y[g] = y;
        B.z[g] = z;
        if (s == 0) {
            e = e0 + ed * tau;
            B.e[g] = e;
        }
        double h = h0;
        if (T.h_mode == 1) {
            h = T.eta * cbrt(m / rho);
            B.h[g] = h;
        }
        const double v = sqrt(vx * vx + vy * vy + vz * vz);
        mn = fmin(mn, h / (T.cs + v));
    } else if (s == 0) {
        e = e0;
    }
    px += m * vx;
    py += m * vy;
    pz += m * vz;
    double kin = 0.5 * m * (vx * vx + vy * vy + vz * vz);
    if (s == 0) kin += m * e;
    E += kin;
}

// t_stop of dust particle k for drag_mode 0 (t_stop_Kernel_IDIC, IDIC.h:78-91): the gas
// particle with the same index, this step's pressure and density; written back.
__device__ __forceinline__ double dust_tstop(const TubeConst& T, const PBuf& B, const int* __restrict__ ginv,
                                             long long wo, int k, int s) {
    double ts = B.ts[wo + k];
    if (T.drag_mode == 0 && B.ind[wo + k] == 0) {
        const long long gq = wo + ginv[T.p_off[0] + B.id[wo + k]];
        const double rg = B.rho[gq];
        ts = T.R[s - 1] / (sqrt(T.gamma * B.p[gq] / rg) * rg);
        B.ts[wo + k] = ts;
    }
    return ts;
}

// Averaging cells of f^3 neighbour cells (f >= 2): one wave per averaging cell, sums over
// its neighbour cells for every species, implicit drag solve, then the per-particle kick,
// drift, energy and h update, and the cell's conservation / CFL partials.
template <int kNd>
__global__ __launch_bounds__(kBS) void k_idic(Map mp, const TubeConst* __restrict__ tc,
                                              const TubeDyn* __restrict__ td, const int* __restrict__ cur, PBuf B,
                                              const int* __restrict__ start, const int* __restrict__ ginv,
                                              double* __restrict__ part) {
    const ItemPos ip = locate(mp, kBS / 64);
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int la = ip.li + wave;
    if (la >= ip.sg.n) return;
    const int t = ip.sg.tube;
    const TubeDyn& D = td[t];
    if (D.done || D.err) return;
    const TubeConst& T = tc[t];
    if (T.f == 1 || T.nsp != kNd + 1) return;   // k_idic_runs, or another instance
    const long long wo = (long long)(1 - *cur) * B.ntot;
    const int ax_ = la % T.na[0];
    const int ay_ = (la / T.na[0]) % T.na[1];
    const int az_ = la / (T.na[0] * T.na[1]);
    constexpr int nd = kNd;
    const double tau = D.dt;

    CellSums S{};
    // species loops unrolled: the per-species arrays are then indexed by constants
#pragma unroll
    for (int s = 0; s <= kNd; ++s) {
        for_avg_cell(T, start, s, ax_, ay_, az_, lane, [&](int k) {
            if (s == 0) {
                ++S.ng;
                S.sg[0] += B.vx[wo + k];
                S.sg[1] += B.vy[wo + k];
                S.sg[2] += B.vz[wo + k];
                S.sg[3] += B.ax[wo + k];
                S.sg[4] += B.ay[wo + k];
                S.sg[5] += B.az[wo + k];
            } else {
                S.ndc[s - 1]++;
                S.sd[s - 1][0] += B.vx[wo + k];
                S.sd[s - 1][1] += B.vy[wo + k];
                S.sd[s - 1][2] += B.vz[wo + k];
                S.sd[s - 1][3] += dust_tstop(T, B, ginv, wo, k, s);
            }
        });
    }
    S.ng = wave_sum_i(S.ng);
    for (int a = 0; a < 6; ++a) S.sg[a] = wave_sum(S.sg[a]);
    for (int k = 0; k < nd; ++k) {
        S.ndc[k] = wave_sum_i(S.ndc[k]);
        for (int a = 0; a < 4; ++a) S.sd[k][a] = wave_sum(S.sd[k][a]);
    }
    const CellDrag C = cell_drag(T, nd, tau, S);
    double px = 0, py = 0, pz = 0, E = 0, mn = 1e300;
#pragma unroll
    for (int s = 0; s <= kNd; ++s) {
        for_avg_cell(T, start, s, ax_, ay_, az_, lane,
                     [&](int k) { particle_update(T, B, wo + k, s, nd, tau, C, px, py, pz, E, mn); });
    }
    px = wave_sum(px);
    py = wave_sum(py);
    pz = wave_sum(pz);
    E = wave_sum(E);
    mn = wave_min(mn);
    if (lane == 0) {
        double* o = part + 5ll * (T.avg_off + la);
        o[0] = px;
        o[1] = py;
        o[2] = pz;
        o[3] = E;
        o[4] = mn;
    }
}

// Averaging cell = one neighbour cell (f == 1): averaging cell la is neighbour cell la, so a
// run of kRunCells consecutive cells is one contiguous range of sorted particles per species.
// One wave per run: lanes take 64 particles at a time; their contributions go to LDS and lane
// j sums cell j's in sorted order, solves cell j and leaves its drag terms in LDS for the
// particles of the cell.  One conservation / CFL partial per block (its runs in order).
// The kernel is latency-bound (a chain of dependent loads, sums and FP64 divisions per wave),
// so registers are capped for 8 waves per SIMD.  The number of dust species is a template
// parameter (one instance per count present), so registers hold only the sums in use.
constexpr int kRunCells = 8;
template <int kNd>
__global__ __launch_bounds__(kBS) __attribute__((amdgpu_waves_per_eu(6))) void k_idic_runs(Map mp, const TubeConst* __restrict__ tc,
                                                   const TubeDyn* __restrict__ td, const int* __restrict__ cur, PBuf B,
                                                   const int* __restrict__ start, const int* __restrict__ ginv,
                                                   double* __restrict__ part) {
    __shared__ double sv[kBS / 64][6][64];
    __shared__ CellDrag cd[kBS / 64][kRunCells];
    __shared__ double bp[kBS / 64][5];
    const ItemPos ip = locate(mp, (kBS / 64) * kRunCells);
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int la0 = ip.li + wave * kRunCells;
    const int t = ip.sg.tube;
    const TubeDyn& D = td[t];
    if (D.done || D.err) return;   // the block's cells are all of tube t
    const TubeConst& T = tc[t];
    if (T.f != 1 || T.nsp != kNd + 1) return;   // k_idic, or another instance
    double px = 0, py = 0, pz = 0, E = 0, mn = 1e300;
    if (la0 < ip.sg.n) {
        const long long wo = (long long)(1 - *cur) * B.ntot;
        const int ncl = min(kRunCells, ip.sg.n - la0);
        constexpr int nd = kNd;
        const double tau = D.dt;
        double (*V)[64] = sv[wave];

        // lane j < ncl: sums of cell la0 + j
        CellSums S{};
#pragma unroll
        for (int s = 0; s <= kNd; ++s) {
            const int c0 = T.cell_off[s] + la0;
            const int p0 = start[c0], p1 = start[c0 + ncl];
            const int jlo = lane < ncl ? start[c0 + lane] : 0, jhi = lane < ncl ? start[c0 + lane + 1] : 0;
            if (lane < ncl) {
                if (s == 0) S.ng = jhi - jlo;
                else S.ndc[s - 1] = jhi - jlo;
            }
            // the next 64 particles' values are read while the current ones are summed
            double nv[6] = {0, 0, 0, 0, 0, 0};
            auto fetch = [&](int base) {
                const int k = base + lane;
                if (k < p1) {
                    const long long g = wo + k;
                    nv[0] = B.vx[g];
                    nv[1] = B.vy[g];
                    nv[2] = B.vz[g];
                    if (s == 0) {
                        nv[3] = B.ax[g];
                        nv[4] = B.ay[g];
                        nv[5] = B.az[g];
                    } else {
                        nv[3] = dust_tstop(T, B, ginv, wo, k, s);
                    }
                }
            };
            if (p0 < p1) fetch(p0);
            for (int base = p0; base < p1; base += 64) {
                if (base + lane < p1) {
                    for (int c = 0; c < (s == 0 ? 6 : 4); ++c) V[c][lane] = nv[c];
                }
                if (base + 64 < p1) fetch(base + 64);
                __builtin_amdgcn_wave_barrier();
                if (lane < ncl) {
                    const int a = max(jlo, base) - base, e = min(jhi, base + 64) - base;
                    for (int u = a; u < e; ++u) {
                        if (s == 0) {
                            for (int c = 0; c < 6; ++c) S.sg[c] += V[c][u];
                        } else {
                            for (int c = 0; c < 4; ++c) S.sd[s - 1][c] += V[c][u];
                        }
                    }
                }
                __builtin_amdgcn_wave_barrier();
            }
        }
        if (lane < ncl) cd[wave][lane] = cell_drag(T, nd, tau, S);
        __builtin_amdgcn_wave_barrier();

#pragma unroll
        for (int s = 0; s <= kNd; ++s) {
            const int c0 = T.cell_off[s] + la0;
            const int p0 = start[c0], p1 = start[c0 + ncl];
            for (int k = p0 + lane; k < p1; k += 64) {
                // the cell of k: the last cell of the run whose start is <= k
                int j = 0;
                for (int jj = 1; jj < ncl; ++jj) j = (start[c0 + jj] <= k) ? jj : j;
                particle_update(T, B, wo + k, s, nd, tau, cd[wave][j], px, py, pz, E, mn);
            }
        }
        px = wave_sum(px);
        py = wave_sum(py);
        pz = wave_sum(pz);
        E = wave_sum(E);
        mn = wave_min(mn);
    }
    if (lane == 0) {
        bp[wave][0] = px;
        bp[wave][1] = py;
        bp[wave][2] = pz;
        bp[wave][3] = E;
        bp[wave][4] = mn;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        double q[4] = {0, 0, 0, 0}, m2 = 1e300;
        for (int w = 0; w < kBS / 64; ++w) {
            for (int c = 0; c < 4; ++c) q[c] += bp[w][c];
            m2 = fmin(m2, bp[w][4]);
        }
        double* o = part + 5ll * (T.avg_off + ip.li / ((kBS / 64) * kRunCells));
        for (int c = 0; c < 4; ++c) o[c] = q[c];
        o[4] = m2;
    }
}

constexpr int kFinBS = 1024;   // threads of the per-tube reduction

// Time, next step size and completion of one tube from its reduced sums (the loop control of
// kernel.cu:1929-1932,2566 plus CFL); sh holds the per-wave sums and minima.
__device__ void finish_tube(const TubeConst& T, TubeDyn& D, int next_cur, bool failed,
                            const double (*sh)[kFinBS / 64]) {
    if (failed) {
        D.done = 1;
        D.done_cur = next_cur;
        return;
    }
    double s[4] = {0, 0, 0, 0}, m2 = 1e300;
    for (int k = 0; k < kFinBS / 64; ++k) {
        for (int c = 0; c < 4; ++c) s[c] += sh[c][k];
        m2 = fmin(m2, sh[4][k]);
    }
    for (int c = 0; c < 4; ++c) D.cons[c] = s[c];
    if (!(isfinite(s[0]) && isfinite(s[1]) && isfinite(s[2]) && isfinite(s[3]))) {
        D.err |= kErrNonFinite;
        D.done = 1;
        D.done_cur = next_cur;
        return;
    }
    D.t = D.t + D.dt;
    D.steps += 1;
    int done = 0;
    if (T.max_steps > 0 && D.steps >= T.max_steps) done = 1;
    if (T.dt_mode == 2) {
        if (D.final_step) {
            done = 1;
        } else {
            const double rem = T.t_end - D.t;
            if (!(rem > 0.0)) {
                done = 1;
            } else {
                const double dtc = T.cfl * m2;
                if (dtc >= rem) {
                    D.dt = rem;
                    D.final_step = 1;
                } else {
                    D.dt = dtc;
                }
            }
        }
    } else {
        if (D.t >= T.t_end) done = 1;
    }
    if (done) {
        D.done = 1;
        D.done_cur = next_cur;
    }
    set_band(T, D);
}

// One block per tube: deterministic reduction of the averaging-cell partials, then
// finish_tube.  The last block to finish flips the current-buffer flag cur[0] for the next
// step (cur[1] counts finished blocks and is reset by that block).
__global__ __launch_bounds__(kFinBS) void k_finalize(const TubeConst* __restrict__ tc, TubeDyn* __restrict__ td,
                                                     int* __restrict__ cur, const double* __restrict__ part) {
    __shared__ double sh[5][kFinBS / 64];
    __shared__ int skip;
    const int t = blockIdx.x;
    const TubeConst& T = tc[t];
    TubeDyn& D = td[t];
    if (threadIdx.x == 0) skip = D.done;
    __syncthreads();
    if (!skip) {
        const bool failed = D.err != 0;
        double a[4] = {0, 0, 0, 0}, mn = 1e300;
        if (!failed) {
            for (int i = threadIdx.x; i < T.npart; i += kFinBS) {
                const double* o = part + 5ll * (T.avg_off + i);
                for (int c = 0; c < 4; ++c) a[c] += o[c];
                mn = fmin(mn, o[4]);
            }
        }
        for (int c = 0; c < 4; ++c) a[c] = wave_sum(a[c]);
        mn = wave_min(mn);
        const int w = threadIdx.x >> 6;
        if ((threadIdx.x & 63) == 0) {
            for (int c = 0; c < 4; ++c) sh[c][w] = a[c];
            sh[4][w] = mn;
        }
        __syncthreads();
        if (threadIdx.x == 0) finish_tube(T, D, 1 - cur[0], failed, sh);
    }
    if (threadIdx.x == 0) {
        // the per-species max of h is rebuilt by the next step's rank pass
        for (int s = 0; s < kMaxSpecies; ++s) D.hmax_bits[s] = 0ull;
        __threadfence();   // this block's reads of cur[0] come before its ticket
        if (atomicAdd(&cur[1], 1) == (int)gridDim.x - 1) {
            cur[1] = 0;
            cur[0] = 1 - cur[0];
        }
    }
}

// The seeded lattice set-up of one (tube, species), one thread per particle, written in
// original order with id = index (what an upload of dsh_tube_ic's arrays would store).
__global__ __launch_bounds__(kBS) void k_tube_ic(TubeIC c, PBuf B, long long off) {
    const int id = blockIdx.x * kBS + threadIdx.x;
    if (id >= c.n) return;
    const ICParticle q = tube_ic_particle(c, id);
    const long long g = off + id;
    B.x[g] = q.x;
    B.y[g] = q.y;
    B.z[g] = q.z;
    B.vx[g] = 0.0;
    B.vy[g] = 0.0;
    B.vz[g] = 0.0;
    B.m[g] = q.m;
    B.h[g] = q.h;
    B.e[g] = q.e;
    B.rho[g] = q.rho;
    B.p[g] = q.p;
    B.ts[g] = q.ts;
    B.ind[g] = q.ind;
    B.id[g] = id;
}

inline dim3 grid1(int n) { return dim3((unsigned)((n + kBS - 1) / kBS)); }

}  // namespace

int scan_blocks(int n) { return (n + kScanTile - 1) / kScanTile; }

// hit masks (one 64-bit word per lane and slot) followed by the x-windows (32 bits each)
size_t mask_words(int gas_chunks) { return (size_t)gas_chunks * kMaskSlots * 64 * 3 / 2; }

hipError_t launch_prepare(const StepArgs& a, double* scratch, hipStream_t s) {
    hipLaunchKernelGGL(k_prepare_part, dim3(a.pmap.nblocks), dim3(kBS), 0, s, a.pmap, a.tc, a.B, scratch);
    hipLaunchKernelGGL(k_prepare, dim3(a.ntubes), dim3(kBS), 0, s, a.pmap, a.tc, a.td, scratch);
    return hipGetLastError();
}

hipError_t launch_key(const StepArgs& a, hipStream_t s) {
    hipLaunchKernelGGL(k_key, dim3(a.pmap.nblocks), dim3(kBS), 0, s, a.pmap, a.tc, a.td, a.cur, a.B, a.key, a.slot,
                       a.count);
    return hipGetLastError();
}

hipError_t launch_scan(const StepArgs& a, hipStream_t s) {
    hipLaunchKernelGGL(k_scan_reduce, dim3(a.scan_nb), dim3(kBS), 0, s, a.count, a.ncells, a.scan_bsum);
    hipLaunchKernelGGL(k_scan_top, dim3(1), dim3(kBS), 0, s, a.scan_bsum, a.scan_nb);
    hipLaunchKernelGGL(k_scan_down, dim3(a.scan_nb), dim3(kBS), 0, s, a.count, a.ncells, a.scan_bsum, a.start);
    return hipGetLastError();
}

hipError_t launch_scatter(const StepArgs& a, hipStream_t s) {
    hipLaunchKernelGGL(k_scatter, dim3(a.pmap.nblocks), dim3(kBS), 0, s, a.pmap, a.td, a.key, a.slot, a.start,
                       a.tsrc, a.tkey);
    return hipGetLastError();
}

hipError_t launch_rank_permute(const StepArgs& a, hipStream_t s) {
    hipLaunchKernelGGL(k_rank_permute, dim3(a.pmap.nblocks), dim3(kBS), 0, s, a.pmap, a.td, a.cur, a.B, a.start,
                       a.tsrc, a.tkey, a.wkey, a.ginv, a.cellh);
    return hipGetLastError();
}

hipError_t launch_radius(const StepArgs& a, hipStream_t s) {
    const dim3 g(a.cmap.nblocks), b(kBS);
    hipLaunchKernelGGL(k_window_max, g, b, 0, s, a.cmap, a.tc, a.td, a.cellh, a.wtmp, 0);
    hipLaunchKernelGGL(k_window_max, g, b, 0, s, a.cmap, a.tc, a.td, a.wtmp, a.wtmp + a.ncells, 1);
    hipLaunchKernelGGL(k_radius, g, b, 0, s, a.cmap, a.tc, a.td, a.cellh, a.wtmp + a.ncells, a.reach);
    // chunk classes (FP64 wall band in mixed precision, finished tubes) and the chunk order of
    // the tiled and fused pair kernels (band chunks first), read by both pair passes
    if (a.pband || a.bperm) {
        const int nall = a.pmap64.nblocks;
        hipLaunchKernelGGL(k_band, dim3(a.pmap.nblocks), b, 0, s, a.pmap, a.td, a.cur, a.B, a.pband, a.pmap64.blk,
                           a.bperm ? a.bperm + nall + 4 : nullptr);
        if (a.bperm && nall > 0) {
            const int nbk = (nall + kPermThreads - 1) / kPermThreads;
            int* cflag = a.bperm + nall + 4;
            int* bcnt = cflag + nall;
            hipLaunchKernelGGL(k_perm_count, dim3(nbk), dim3(kPermThreads), 0, s, cflag, nall, bcnt);
            hipLaunchKernelGGL(k_perm_write, dim3(nbk), dim3(kPermThreads), 0, s, cflag, nall, bcnt,
                               a.gmap64.nblocks, a.bperm);
        }
    }
    return hipGetLastError();
}

template <typename R, int KIND, bool kStats, int kTileP, int kWaves>
static void launch_tiled_t(const StepArgs& a, const Map& m, const unsigned char* pband, hipStream_t s) {
    const int grid = (m.nblocks + kWaves - 1) / kWaves;
    hipLaunchKernelGGL((k_pairs_tiled<R, KIND, kStats, kTileP, kWaves>), dim3(grid), dim3(64 * kWaves), 0, s, m, a.tc,
                       a.td, a.cur, a.B, a.start, a.wkey, a.reach, a.pr2, a.csav, a.ncount, a.nhash, a.masks,
                       a.gmap64.nblocks, pband, a.frec, a.bperm, a.pmap64.nblocks);
}

template <int KIND, bool kStats, int kTileP, int kWaves>
static void launch_fused_t(const StepArgs& a, const Map& m, hipStream_t s) {
    // both parts in whole blocks: ceil(nband / W) + ceil((n - nband) / W) <= ceil(n / W) + 1
    const int grid = kWaves == 1 ? m.nblocks : (m.nblocks + kWaves - 1) / kWaves + 1;
    hipLaunchKernelGGL((k_pairs_fused<KIND, kStats, kTileP, kWaves>), dim3(grid), dim3(64 * kWaves), 0, s, m, a.tc,
                       a.td, a.cur, a.B, a.start, a.wkey, a.reach, a.pr2, a.csav, a.ncount, a.nhash, a.masks,
                       a.gmap64.nblocks, a.pband, a.bperm, a.pmap64.nblocks, a.frec);
}

// Calls f with the block shape (piece size, waves per block) of a as compile-time constants.
template <typename F>
static void with_shape(const StepArgs& a, F&& f) {
    auto waves = [&](auto tile) {
        switch (a.waves) {
            case 2: f(tile, std::integral_constant<int, 2>{}); break;
            case 4: f(tile, std::integral_constant<int, 4>{}); break;
            default: f(tile, std::integral_constant<int, 1>{}); break;
        }
    };
    switch (a.tile) {
        case 128: waves(std::integral_constant<int, 128>{}); break;
        case 192: waves(std::integral_constant<int, 192>{}); break;
        case 224: waves(std::integral_constant<int, 224>{}); break;
        case 256: waves(std::integral_constant<int, 256>{}); break;
        default: waves(std::integral_constant<int, 160>{}); break;
    }
}

template <int KIND, bool kStats>
static void launch_fused(const StepArgs& a, const Map& m, hipStream_t s) {
    with_shape(a, [&](auto tile, auto waves) {
        launch_fused_t<KIND, kStats, decltype(tile)::value, decltype(waves)::value>(a, m, s);
    });
}

template <typename R, int KIND, bool kStats>
static void launch_tiled1(const StepArgs& a, const Map& m, const unsigned char* pband, hipStream_t s) {
    with_shape(a, [&](auto tile, auto waves) {
        launch_tiled_t<R, KIND, kStats, decltype(tile)::value, decltype(waves)::value>(a, m, pband, s);
    });
}

// FP64 mode: the FP64 instance over every chunk.  Mixed without a wall band: the FP32
// instance.  Mixed with a band: the fused kernel, band chunks on its FP64 path first.
template <int KIND, bool kStats>
static hipError_t launch_tiled(const StepArgs& a, const Map& m, int precision, hipStream_t s) {
    if (precision == 0) {
        launch_tiled1<double, KIND, kStats>(a, m, nullptr, s);
    } else if (!a.pband) {
        launch_tiled1<float, KIND, kStats>(a, m, nullptr, s);
    } else {
        launch_fused<KIND, kStats>(a, m, s);
    }
    return hipGetLastError();
}

hipError_t launch_density(const StepArgs& a, int precision, bool stats, hipStream_t s) {
    if (a.variant == kPairTiled) {
        if (a.pmap64.nblocks == 0) return hipSuccess;
        return stats ? launch_tiled<0, true>(a, a.pmap64, precision, s)
                     : launch_tiled<0, false>(a, a.pmap64, precision, s);
    }
    const dim3 g(a.pmap.nblocks), b(kBS);
    if (precision == 0) {
        if (stats)
            hipLaunchKernelGGL((k_density<double, true>), g, b, 0, s, a.pmap, a.tc, a.td, a.cur, a.B, a.start, a.wkey,
                               a.reach, a.pr2, a.csav, a.ncount, a.nhash, nullptr);
        else
            hipLaunchKernelGGL((k_density<double, false>), g, b, 0, s, a.pmap, a.tc, a.td, a.cur, a.B, a.start,
                               a.wkey, a.reach, a.pr2, a.csav, a.ncount, a.nhash, nullptr);
    } else {
        if (stats)
            hipLaunchKernelGGL((k_density<float, true>), g, b, 0, s, a.pmap, a.tc, a.td, a.cur, a.B, a.start, a.wkey,
                               a.reach, a.pr2, a.csav, a.ncount, a.nhash, a.pband);
        else
            hipLaunchKernelGGL((k_density<float, false>), g, b, 0, s, a.pmap, a.tc, a.td, a.cur, a.B, a.start,
                               a.wkey, a.reach, a.pr2, a.csav, a.ncount, a.nhash, a.pband);
    }
    return hipGetLastError();
}

hipError_t launch_force(const StepArgs& a, int precision, hipStream_t s) {
    if (a.variant == kPairTiled) {
        if (a.gmap64.nblocks == 0) return hipSuccess;
        return launch_tiled<1, false>(a, a.gmap64, precision, s);
    }
    if (a.gmap.nblocks == 0) return hipSuccess;
    const dim3 g(a.gmap.nblocks), b(kBS);
    if (precision == 0)
        hipLaunchKernelGGL((k_force<double>), g, b, 0, s, a.gmap, a.tc, a.td, a.cur, a.B, a.start, a.wkey, a.reach,
                           a.pr2, a.csav, nullptr);
    else
        hipLaunchKernelGGL((k_force<float>), g, b, 0, s, a.gmap, a.tc, a.td, a.cur, a.B, a.start, a.wkey, a.reach,
                           a.pr2, a.csav, a.pband);
    return hipGetLastError();
}

hipError_t launch_idic(const StepArgs& a, hipStream_t s) {
    if (a.idic_blocks & 1u)
        hipLaunchKernelGGL(k_idic<0>, dim3(a.amap.nblocks), dim3(kBS), 0, s, a.amap, a.tc, a.td, a.cur, a.B, a.start,
                           a.ginv, a.part);
    if (a.idic_blocks & 2u)
        hipLaunchKernelGGL(k_idic<1>, dim3(a.amap.nblocks), dim3(kBS), 0, s, a.amap, a.tc, a.td, a.cur, a.B, a.start,
                           a.ginv, a.part);
    if (a.idic_blocks & 4u)
        hipLaunchKernelGGL(k_idic<2>, dim3(a.amap.nblocks), dim3(kBS), 0, s, a.amap, a.tc, a.td, a.cur, a.B, a.start,
                           a.ginv, a.part);
    if (a.idic_runs & 1u)
        hipLaunchKernelGGL(k_idic_runs<0>, dim3(a.amap_runs.nblocks), dim3(kBS), 0, s, a.amap_runs, a.tc, a.td, a.cur,
                           a.B, a.start, a.ginv, a.part);
    if (a.idic_runs & 2u)
        hipLaunchKernelGGL(k_idic_runs<1>, dim3(a.amap_runs.nblocks), dim3(kBS), 0, s, a.amap_runs, a.tc, a.td, a.cur,
                           a.B, a.start, a.ginv, a.part);
    if (a.idic_runs & 4u)
        hipLaunchKernelGGL(k_idic_runs<2>, dim3(a.amap_runs.nblocks), dim3(kBS), 0, s, a.amap_runs, a.tc, a.td, a.cur,
                           a.B, a.start, a.ginv, a.part);
    return hipGetLastError();
}

int idic_run_cells() { return kRunCells; }
int idic_block_cells() { return (kBS / 64) * kRunCells; }

hipError_t launch_finalize(const StepArgs& a, hipStream_t s) {
    hipLaunchKernelGGL(k_finalize, dim3(a.ntubes), dim3(kFinBS), 0, s, a.tc, a.td, a.cur, a.part);
    return hipGetLastError();
}

hipError_t launch_tube_ic(const TubeIC& c, const PBuf& B, long long off, hipStream_t s) {
    if (c.n <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_tube_ic, grid1(c.n), dim3(kBS), 0, s, c, B, off);
    return hipGetLastError();
}

}  // namespace dsh
