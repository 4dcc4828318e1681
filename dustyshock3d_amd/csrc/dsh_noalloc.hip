// dsh_noalloc.hip - dsh_step_noalloc: one tube stepped from device arrays through the step
// kernels of dsh_step, with every scratch buffer (sorted state, keys, cell arrays, hit masks,
// partials, launch maps, tube records) carved from one static device array.  It calls no
// device allocation routine, for callers that cannot allocate.
//
// The array is a file-scope __device__ object: each GPU holds its own copy at its own
// address.  Its device address is taken once per device with hipGetSymbolAddress (the host
// symbol is only a stand-in, never passed to a kernel), and every call on that GPU reuses it,
// so each call clears what must start at zero and calls that overlap on one GPU race.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>
#include <cstring>
#include <mutex>

#include "../../include/dsh.h"
#include "dsh_engine.h"
#include "dsh_kernels.h"

namespace dsh {

constexpr size_t kScratchBytes = (size_t)64 << 20;   // 64 MiB per GPU
static __device__ __align__(256) unsigned char scratch[kScratchBytes];

namespace {

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
unsigned char* g_scratch_ptr[kMaxDevices] = {};

// Device address of this GPU's scratch array (taken once per device, device current).
int scratch_base(unsigned char** out) {
    int dev = -1;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kMaxDevices) return DSH_ERR_HIP;
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void* p = nullptr;
        if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess || p == nullptr) return DSH_ERR_HIP;
        g_scratch_ptr[dev] = static_cast<unsigned char*>(p);
    }
    *out = g_scratch_ptr[dev];
    return DSH_OK;
}

// Bump allocation inside the scratch array (256-byte aligned regions).
struct Arena {
    unsigned char* base;
    size_t used = 0;
    bool ok = true;
    template <class T> T* take(size_t count) {
        const size_t bytes = (std::max<size_t>(count, 1) * sizeof(T) + 255) & ~(size_t)255;
        if (used + bytes > kScratchBytes) {
            ok = false;
            return nullptr;
        }
        T* p = reinterpret_cast<T*>(base + used);
        used += bytes;
        return p;
    }
};

__global__ void k_zero(unsigned long long* __restrict__ p, size_t n) {
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        p[i] = 0ull;
}

__global__ void k_iota(int* __restrict__ p, int n) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = i;
}

// Smallest h over the species' input arrays (one block).
struct HIn {
    const double* h[kMaxSpecies];
    int n[kMaxSpecies];
    int nsp;
};
__global__ __launch_bounds__(256) void k_min_h(HIn in, double* __restrict__ out) {
    __shared__ double sh[256];
    double m = 1e300;
    for (int s = 0; s < in.nsp; ++s)
        for (int i = threadIdx.x; i < in.n[s]; i += 256) m = fmin(m, in.h[s][i]);
    sh[threadIdx.x] = m;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) sh[threadIdx.x] = fmin(sh[threadIdx.x], sh[threadIdx.x + w]);
        __syncthreads();
    }
    if (threadIdx.x == 0) *out = sh[0];
}

// Sorted state of one species back to the caller's arrays in original particle order
// (what dsh_download_species fills; nn is not produced here).
__global__ __launch_bounds__(256) void k_unpermute(PBuf B, const int* __restrict__ cur, const TubeDyn* __restrict__ td,
                                                   const int* __restrict__ wkey, int off, int n, int cell_off,
                                                   dsh_soa_view out) {
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= n) return;
    // a tube that finished before the last step requested keeps its state in buffer done_cur
    const long long g = (long long)(td->done ? td->done_cur : *cur) * B.ntot + off + q;
    const int id = B.id[g];
    struct {
        double* o;
        const double* f;
    } fs[] = {{out.x, B.x},   {out.y, B.y},   {out.z, B.z},   {out.vx, B.vx},   {out.vy, B.vy},   {out.vz, B.vz},
              {out.ax, B.ax}, {out.ay, B.ay}, {out.az, B.az}, {out.m, B.m},     {out.rho, B.rho}, {out.p, B.p},
              {out.e, B.e},   {out.h, B.h},   {out.t_stop, B.ts}, {out.edot, B.ed}};
    for (auto& f : fs)
        if (f.o) f.o[id] = f.f[g];
    if (out.ind) out.ind[id] = B.ind[g];
    if (out.cell) out.cell[id] = td->steps > 0 ? wkey[off + q] - cell_off : -1;
}

// Block map of one launch over a single tube's segments (host side, on the stack).
struct HostMap {
    Seg seg[kMaxSpecies];
    int blk[kMaxSpecies + 1];
    int nseg;
};
HostMap host_map(const Seg* segs, int nseg, int per_block) {
    HostMap m{};
    m.nseg = nseg;
    m.blk[0] = 0;
    for (int i = 0; i < nseg; ++i) {
        m.seg[i] = segs[i];
        m.blk[i + 1] = m.blk[i] + (segs[i].n + per_block - 1) / per_block;
    }
    return m;
}

// The launch maps and tube records, copied to the scratch in one transfer.
struct Records {
    HostMap maps[7];   // pmap, gmap, cmap, amap, amap_runs, pmap64, gmap64
    TubeConst tc;
};

int status_of(const TubeDyn& D) {
    if (D.err & kErrGrid) return DSH_ERR_GRID;
    if (D.err & kErrNonFinite) return DSH_ERR_NONFINITE;
    return DSH_OK;
}

}  // namespace

int step_noalloc(const dsh_params* P, const dsh_soa_view* in, const int* counts, int nsteps,
                 const dsh_soa_view* out, double* t_out, double cons_out[4], hipStream_t st) {
    if (!P || !in || !counts || !out || nsteps < 0) return DSH_ERR_ARG;
    if (P->n_dust < 0 || P->n_dust > kMaxDust || (P->drag_mode != 0 && P->drag_mode != 1)) return DSH_ERR_ARG;
    const int nsp = 1 + P->n_dust;
    long long ntot = 0;
    for (int s = 0; s < nsp; ++s) {
        if (counts[s] < 0) return DSH_ERR_ARG;
        if (counts[s] > 0 && (!in[s].x || !in[s].y || !in[s].z || !in[s].m || !in[s].ind || !in[s].h))
            return DSH_ERR_ARG;
        if (out[s].nn) return DSH_ERR_UNSUPPORTED;
        ntot += counts[s];
    }
    if (ntot == 0 || ntot > (1ll << 24)) return ntot == 0 ? DSH_ERR_ARG : DSH_ERR_BUDGET;
    if (P->drag_mode == 0)
        for (int s = 1; s < nsp; ++s)
            if (counts[s] != counts[0]) return DSH_ERR_ARG;
    (void)hipGetLastError();   // an earlier call's error (a caller's query) is not ours
    unsigned char* base = nullptr;
    int rc = scratch_base(&base);
    if (rc) return rc;
    // the per-particle regions alone must fit before any input is read (the cell arrays are
    // checked once the grid is known)
    {
        Arena dry{base};
        const size_t n2 = 2 * (size_t)ntot, ngas = (size_t)counts[0];
        dry.take<Records>(1);
        dry.take<double>(16 * n2);
        dry.take<int>(2 * n2);
        dry.take<int>(6 * (size_t)ntot);
        dry.take<double>(3 * (size_t)ntot);
        dry.take<unsigned long long>(mask_words((int)((ngas + 63) / 64)));
        if (P->precision == DSH_MIXED) dry.take<ForceRec>(ngas);
        dry.take<double>((size_t)kPrepParts * (((size_t)ntot + 255) / 256 + nsp));
        if (!dry.ok) return DSH_ERR_BUDGET;
    }
    Arena A{base};

    // grid from the smallest h of the input (as the engine builds it at its first step)
    double* dmin = A.take<double>(1);
    HIn hin{};
    hin.nsp = nsp;
    for (int s = 0; s < nsp; ++s) {
        hin.h[s] = in[s].h;
        hin.n[s] = counts[s];
    }
    hipLaunchKernelGGL(k_min_h, dim3(1), dim3(256), 0, st, hin, dmin);
    double hmin = 0.0;
    if (hipMemcpyAsync(&hmin, dmin, sizeof(double), hipMemcpyDeviceToHost, st) != hipSuccess ||
        hipStreamSynchronize(st) != hipSuccess)
        return DSH_ERR_HIP;

    Records R{};
    TubeConst& C = R.tc;
    build_grid(*P, hmin, C);
    tube_physics(*P, C);
    if ((long long)C.ncell * nsp > (1ll << 26)) return DSH_ERR_BUDGET;
    Seg pseg[kMaxSpecies], gseg[1], cseg[kMaxSpecies], aseg[1];
    int poff = 0, coff = 0;
    for (int s = 0; s < kMaxSpecies; ++s) {
        if (s >= nsp) {
            C.p_off[s] = C.np[s] = C.cell_off[s] = 0;
            C.pseg[s] = -1;
            continue;
        }
        C.pseg[s] = s;
        C.p_off[s] = poff;
        C.np[s] = counts[s];
        C.cell_off[s] = coff;
        pseg[s] = Seg{poff, counts[s], 0, s};
        cseg[s] = Seg{coff, C.ncell, 0, s};
        poff += counts[s];
        coff += C.ncell;
    }
    gseg[0] = pseg[0];
    C.avg_off = 0;
    C.npart = C.f == 1 ? (C.navg + idic_block_cells() - 1) / idic_block_cells() : C.navg;
    aseg[0] = Seg{0, C.navg, 0, 0};
    R.maps[0] = host_map(pseg, nsp, 256);
    R.maps[1] = host_map(gseg, 1, 256);
    R.maps[2] = host_map(cseg, nsp, 256);
    R.maps[3] = host_map(aseg, 1, 4);
    R.maps[4] = host_map(aseg, 1, idic_block_cells());
    R.maps[5] = host_map(pseg, nsp, 64);
    R.maps[6] = host_map(gseg, 1, 64);

    // scratch layout
    StepArgs a{};
    Records* drec = A.take<Records>(1);
    TubeDyn* dtd = A.take<TubeDyn>(1);
    a.tc = &drec->tc;
    a.td = dtd;
    a.cur = A.take<int>(2);
    a.ntubes = 1;
    PBuf& B = a.B;
    B.ntot = ntot;
    const size_t n2 = 2 * (size_t)ntot;
    double** dfields[] = {&B.x, &B.y, &B.z, &B.vx, &B.vy, &B.vz, &B.m, &B.h, &B.e, &B.rho,
                          &B.p, &B.ts, &B.ax, &B.ay, &B.az, &B.ed};
    // the double-buffered state is one contiguous region (cleared together)
    double* state = A.take<double>(16 * n2);
    for (int f = 0; f < 16; ++f) *dfields[f] = state ? state + (size_t)f * n2 : nullptr;
    B.ind = A.take<int>(n2);
    B.id = A.take<int>(n2);
    int** ifields[] = {&a.key, &a.slot, &a.tsrc, &a.tkey, &a.wkey, &a.ncount};
    for (int** f : ifields) *f = A.take<int>((size_t)ntot);
    a.pr2 = A.take<double>((size_t)ntot);
    a.csav = A.take<double>((size_t)ntot);
    a.nhash = A.take<unsigned long long>((size_t)ntot);
    a.ncells = coff;
    // cell counts and cell maxima of h must start at zero: one region, cleared together
    double* zcells = A.take<double>(2 * (size_t)coff);
    a.cellh = zcells;
    a.count = zcells ? reinterpret_cast<int*>(zcells + coff) : nullptr;
    a.start = A.take<int>((size_t)coff + 1);
    a.reach = A.take<double>((size_t)coff);
    a.wtmp = A.take<double>(2 * (size_t)coff);
    a.scan_nb = scan_blocks(coff);
    a.scan_bsum = A.take<int>((size_t)a.scan_nb);
    a.part = A.take<double>(5 * (size_t)C.navg);
    a.masks = A.take<unsigned long long>(mask_words(R.maps[6].blk[1]));
    a.frec = P->precision == DSH_MIXED ? A.take<ForceRec>((size_t)counts[0]) : nullptr;
    a.ginv = P->drag_mode == 0 ? A.take<int>((size_t)counts[0]) : nullptr;
    a.pband = (P->precision == DSH_MIXED && C.wall_band) ? A.take<unsigned char>((size_t)ntot) : nullptr;
    a.bperm = a.pband ? A.take<int>(bperm_ints(R.maps[5].blk[R.maps[5].nseg])) : nullptr;
    double* prep = A.take<double>((size_t)kPrepParts * R.maps[0].blk[nsp]);
    Map* dmaps[] = {&a.pmap, &a.gmap, &a.cmap, &a.amap, &a.amap_runs, &a.pmap64, &a.gmap64};
    for (int i = 0; i < 7; ++i) {
        dmaps[i]->seg = drec->maps[i].seg;
        dmaps[i]->blk = drec->maps[i].blk;
        dmaps[i]->nseg = R.maps[i].nseg;
        dmaps[i]->nblocks = R.maps[i].blk[R.maps[i].nseg];
    }
    if (!A.ok) return DSH_ERR_BUDGET;   // the tube does not fit the scratch array
    a.idic_blocks = C.f == 1 ? 0u : 1u << (nsp - 1);
    a.idic_runs = C.f == 1 ? 1u << (nsp - 1) : 0u;
    a.variant = kPairTiled;
    a.tile = 192;
    a.waves = 1;

    // set-up: records, cleared state and cells, inputs in buffer 0 in original order
    const int zero2[2] = {0, 0};
    if (hipMemcpyAsync(drec, &R, sizeof(Records), hipMemcpyHostToDevice, st) != hipSuccess ||
        hipMemcpyAsync(a.cur, zero2, sizeof(zero2), hipMemcpyHostToDevice, st) != hipSuccess)
        return DSH_ERR_HIP;
    hipLaunchKernelGGL(k_zero, dim3(1024), dim3(256), 0, st, reinterpret_cast<unsigned long long*>(state), 16 * n2);
    hipLaunchKernelGGL(k_zero, dim3(256), dim3(256), 0, st, reinterpret_cast<unsigned long long*>(zcells),
                       2 * (size_t)coff);
    for (int s = 0; s < nsp; ++s) {
        const size_t n = (size_t)counts[s], off = (size_t)C.p_off[s];
        if (n == 0) continue;
        struct {
            double* dst;
            const double* src;
        } cp[] = {{B.x, in[s].x},   {B.y, in[s].y},   {B.z, in[s].z},   {B.vx, in[s].vx},     {B.vy, in[s].vy},
                  {B.vz, in[s].vz}, {B.m, in[s].m},   {B.h, in[s].h},   {B.e, in[s].e},       {B.rho, in[s].rho},
                  {B.p, in[s].p},   {B.ts, in[s].t_stop}};
        for (auto& c : cp)
            if (c.src && hipMemcpyAsync(c.dst + off, c.src, n * sizeof(double), hipMemcpyDeviceToDevice, st) != hipSuccess)
                return DSH_ERR_HIP;
        if (hipMemcpyAsync(B.ind + off, in[s].ind, n * sizeof(int), hipMemcpyDeviceToDevice, st) != hipSuccess)
            return DSH_ERR_HIP;
        hipLaunchKernelGGL(k_iota, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, B.id + off, (int)n);
    }
    if (hipGetLastError() != hipSuccess || launch_prepare(a, prep, st) != hipSuccess) return DSH_ERR_HIP;

    // the steps of dsh_step: a finished tube's kernels skip it, so extra steps change nothing
    for (int k = 0; k < nsteps; ++k) {
        hipError_t e;
        if ((e = launch_key(a, st)) != hipSuccess || (e = launch_scan(a, st)) != hipSuccess ||
            (e = launch_scatter(a, st)) != hipSuccess || (e = launch_rank_permute(a, st)) != hipSuccess ||
            (e = launch_radius(a, st)) != hipSuccess || (e = launch_density(a, P->precision, false, st)) != hipSuccess ||
            (e = launch_force(a, P->precision, st)) != hipSuccess || (e = launch_idic(a, st)) != hipSuccess ||
            (e = launch_finalize(a, st)) != hipSuccess)
            return DSH_ERR_HIP;
    }
    for (int s = 0; s < nsp; ++s) {
        if (counts[s] == 0) continue;
        hipLaunchKernelGGL(k_unpermute, dim3((unsigned)((counts[s] + 255) / 256)), dim3(256), 0, st, B, a.cur, a.td,
                           a.wkey, C.p_off[s], counts[s], C.cell_off[s], out[s]);
    }
    TubeDyn D{};
    if (hipGetLastError() != hipSuccess ||
        hipMemcpyAsync(&D, dtd, sizeof(TubeDyn), hipMemcpyDeviceToHost, st) != hipSuccess ||
        hipStreamSynchronize(st) != hipSuccess)
        return DSH_ERR_HIP;
    if (t_out) *t_out = D.t;
    if (cons_out)
        for (int c = 0; c < 4; ++c) cons_out[c] = D.cons[c];
    return status_of(D);
}

}  // namespace dsh

extern "C" int dsh_step_noalloc(const dsh_params* params, const dsh_soa_view* in, const int* counts, int nsteps,
                                const dsh_soa_view* out, double* t_out, double cons_out[4], void* stream) {
    return dsh::step_noalloc(params, in, counts, nsteps, out, t_out, cons_out, static_cast<hipStream_t>(stream));
}
