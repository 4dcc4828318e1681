// dsh_engine.cpp - host side of the batch engine: grids, layout, upload/download,
// the per-step launch sequence (optionally captured once in a hipGraph), and timing.
#include "dsh_engine.h"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <numeric>

namespace dsh {

void build_grid(const dsh_params& P, double hmin, TubeConst& T) {
    const double lo[3] = {P.x_min, P.y_min, P.z_min};
    const double up[3] = {P.x_max, P.y_max, P.z_max};
    T.per[0] = 0;
    T.per[1] = P.periodic_y ? 1 : 0;
    T.per[2] = P.periodic_z ? 1 : 0;
    T.f = P.avg_factor < 1 ? 1 : P.avg_factor;
    const double ct = P.cell_size > 0.0 ? P.cell_size : 2.0 * hmin * (1.0 + P.cell_slack);
    const double block = ct * (double)T.f;
    T.cmin = 1e300;
    for (int a = 0; a < 3; ++a) {
        const double L = up[a] - lo[a];
        int n;
        double c;
        if (T.per[a]) {
            n = (int)std::floor(L / block) * T.f;
            if (n < T.f) n = T.f;
            c = L / (double)n;
        } else {
            n = (int)std::ceil(L / block) * T.f;
            if (n < T.f) n = T.f;
            c = ct;
        }
        T.n[a] = n;
        T.c[a] = c;
        T.o[a] = lo[a];
        T.ic[a] = 1.0 / c;
        T.L[a] = L;
        T.hi[a] = lo[a] + L;
        T.cmin = std::min(T.cmin, c);
        T.na[a] = n / T.f;
    }
    T.ncell = T.n[0] * T.n[1] * T.n[2];
    T.navg = T.na[0] * T.na[1] * T.na[2];
}

void tube_physics(const dsh_params& P, TubeConst& C) {
    C.nsp = 1 + P.n_dust;
    C.gamma = P.gamma;
    C.cs = P.cs;
    C.alpha = P.alpha;
    C.beta = P.beta;
    C.eps = P.eps;
    C.eta = P.eta;
    for (int k = 0; k < kMaxDust; ++k) {
        C.K[k] = P.K[k];
        C.mu[k] = P.mass_ratio[k];
        C.eps_fb[k] = P.eps_fallback[k];
        C.R[k] = P.R_dust[k];
    }
    C.drag_mode = P.drag_mode;
    C.t_fb = P.tstop_fallback;
    C.tau = P.tau;
    C.cfl = P.cfl;
    C.t_end = P.t_end;
    C.eos = P.eos_type;
    C.h_mode = P.h_mode;
    C.eps2_quirk = P.eps2_uses_count1;
    C.dt_mode = P.dt_mode;
    C.max_steps = P.max_steps;
    C.precision = P.precision;
    C.wall_band = P.fp64_wall_band ? 1 : 0;
    C.band_speed = P.wall_band_speed > 0.0 ? P.wall_band_speed : 0.0;   // 0: the fixed part of the band alone
    C.band_horizon = P.wall_band_horizon > 0 ? P.wall_band_horizon : 0;  // 0: the whole run
}

Engine::Engine(const std::vector<dsh_params>& params)
    : params_(params), staged_(params.size() * kMaxSpecies), migr_(params.size()) {}

int Engine::adopt(int tube, const Engine& src, int src_tube) {
    if (prepared_) return fail(DSH_ERR_STATE, "adopt after the first step");
    if (tube < 0 || tube >= ntubes()) return fail(DSH_ERR_ARG, "tube index out of range");
    if (!src.prepared() || src_tube < 0 || src_tube >= src.ntubes())
        return fail(DSH_ERR_ARG, "source tube is not on a prepared engine");
    if (src.params_[src_tube].precision != params_[tube].precision ||
        src.params_[src_tube].n_dust != params_[tube].n_dust)
        return fail(DSH_ERR_ARG, "adopted tube must keep its parameters");
    const TubeConst& C = src.tc_h_[src_tube];
    for (int s = 0; s < kMaxSpecies; ++s) {
        HostSpecies& q = staged_[(size_t)tube * kMaxSpecies + s];
        q = HostSpecies{};
        if (s < C.nsp) {
            q.n = C.np[s];
            q.set = true;
        }
    }
    migr_[tube] = Migrant{&src, src_tube};
    return DSH_OK;
}

int Engine::device_fields(int tube, int species, int* cur_out, long long* off_out) const {
    if (!prepared_ || tube < 0 || tube >= ntubes() || species < 0 || species >= tc_h_[tube].nsp) return DSH_ERR_ARG;
    int cur = 0;
    if (hipMemcpy(&cur, a_.cur, sizeof(int), hipMemcpyDeviceToHost) != hipSuccess) return DSH_ERR_HIP;
    TubeDyn d;
    if (hipMemcpy(&d, a_.td + tube, sizeof(TubeDyn), hipMemcpyDeviceToHost) != hipSuccess) return DSH_ERR_HIP;
    *cur_out = d.done ? d.done_cur : cur;   // a finished tube's state stays in done_cur
    *off_out = tc_h_[tube].p_off[species];
    return DSH_OK;
}

Engine::~Engine() {
    for (hipEvent_t e : events_) (void)hipEventDestroy(e);
    if (graph_exec_) (void)hipGraphExecDestroy(graph_exec_);
    if (graph_) (void)hipGraphDestroy(graph_);
    if (graph_exec_n_) (void)hipGraphExecDestroy(graph_exec_n_);
    if (graph_n_) (void)hipGraphDestroy(graph_n_);
    for (void* p : allocs_) (void)hipFree(p);
    if (stream_) (void)hipStreamDestroy(stream_);
}

int Engine::fail(int code, const std::string& msg) {
    err_ = msg;
    return code;
}

int Engine::hip_check(hipError_t e, const char* what) {
    if (e == hipSuccess) return DSH_OK;
    return fail(e == hipErrorOutOfMemory ? DSH_ERR_NOMEM : DSH_ERR_HIP,
                std::string(what) + ": " + hipGetErrorString(e));
}

template <class T> int Engine::alloc(T** p, size_t count) {
    const size_t bytes = std::max<size_t>(count, 1) * sizeof(T);
    void* q = nullptr;
    const hipError_t e = hipMalloc(&q, bytes);
    if (e != hipSuccess) return hip_check(e, "hipMalloc");
    allocs_.push_back(q);
    dev_bytes_ += bytes;
    *p = static_cast<T*>(q);
    return DSH_OK;
}

long long Engine::particles(int tube) const {
    long long n = 0;
    for (int s = 0; s < kMaxSpecies; ++s) n += staged_[(size_t)tube * kMaxSpecies + s].n;
    return n;
}

int Engine::upload(int tube, int species, const dsh_soa_view* v, int pm) {
    if (prepared_) return fail(DSH_ERR_STATE, "upload after the first step");
    if (tube < 0 || tube >= ntubes()) return fail(DSH_ERR_ARG, "tube index out of range");
    const dsh_params& P = params_[tube];
    if (species < 0 || species > P.n_dust) return fail(DSH_ERR_ARG, "species index out of range");
    if (!v || pm < -1) return fail(DSH_ERR_ARG, "null view or negative count");
    if (!v->x || !v->y || !v->z || !v->m || !v->ind) return fail(DSH_ERR_ARG, "view needs x, y, z, m and ind");
    if (P.h_mode == 1 && !v->h && !v->rho) return fail(DSH_ERR_ARG, "h_mode 1 needs h or rho");
    const int n = pm + 1;
    HostSpecies& q = staged_[(size_t)tube * kMaxSpecies + species];
    q.n = n;
    q.set = true;
    auto take = [n](std::vector<double>& dst, const double* src, double dflt) {
        dst.resize(n);
        for (int i = 0; i < n; ++i) dst[i] = src ? src[i] : dflt;
    };
    take(q.x, v->x, 0.0);
    take(q.y, v->y, 0.0);
    take(q.z, v->z, 0.0);
    take(q.vx, v->vx, 0.0);
    take(q.vy, v->vy, 0.0);
    take(q.vz, v->vz, 0.0);
    take(q.m, v->m, 0.0);
    take(q.rho, v->rho, 0.0);
    take(q.p, v->p, 0.0);
    take(q.e, v->e, 0.0);
    take(q.ts, v->t_stop, 0.0);
    q.ind.assign(v->ind, v->ind + n);
    q.h.resize(n);
    for (int i = 0; i < n; ++i) {
        if (v->h)
            q.h[i] = v->h[i];
        else if (P.h_mode == 1)
            q.h[i] = P.eta * std::cbrt(q.m[i] / q.rho[i]);
        else
            q.h[i] = P.h;
    }
    return DSH_OK;
}

int Engine::stage_ic(int tube, int species, const TubeIC& ic) {
    if (prepared_) return fail(DSH_ERR_STATE, "set-up after the first step");
    if (tube < 0 || tube >= ntubes()) return fail(DSH_ERR_ARG, "tube index out of range");
    if (species < 0 || species > params_[tube].n_dust) return fail(DSH_ERR_ARG, "species index out of range");
    if (ic.n < 0 || ic.n0 < 0 || ic.n0 > ic.n) return fail(DSH_ERR_ARG, "bad set-up counts");
    HostSpecies& q = staged_[(size_t)tube * kMaxSpecies + species];
    q = HostSpecies{};
    q.n = ic.n;
    q.set = true;
    q.device_ic = true;
    q.ic = ic;
    return DSH_OK;
}

int Engine::prepare() {
    if (prepared_) return DSH_OK;
    const int T = ntubes();
    if (T == 0) return fail(DSH_ERR_ARG, "no tubes");
    precision_ = params_[0].precision;
    for (int t = 0; t < T; ++t) {
        const dsh_params& P = params_[t];
        if (P.precision != precision_) return fail(DSH_ERR_ARG, "tubes of one batch must share the precision");
        if (P.n_dust < 0 || P.n_dust > kMaxDust) return fail(DSH_ERR_ARG, "n_dust must be 0, 1 or 2");
        if (P.drag_mode != 0 && P.drag_mode != 1) return fail(DSH_ERR_ARG, "drag_mode must be 0 or 1");
        for (int s = 0; s <= P.n_dust; ++s)
            if (!staged_[(size_t)t * kMaxSpecies + s].set) return fail(DSH_ERR_STATE, "species not uploaded");
        if (P.drag_mode == 0)
            for (int s = 1; s <= P.n_dust; ++s)
                if (staged_[(size_t)t * kMaxSpecies + s].n != staged_[(size_t)t * kMaxSpecies].n)
                    return fail(DSH_ERR_ARG, "drag_mode 0 pairs dust index i with gas index i: counts must match");
    }
    tc_h_.assign(T, TubeConst{});
    for (int t = 0; t < T; ++t) {
        const dsh_params& P = params_[t];
        TubeConst& C = tc_h_[t];
        if (migr_[t].src) {
            C = migr_[t].src->tc_h_[migr_[t].tube];   // keeps the grid the tube was started on
        } else {
            double hmin = 1e300;
            for (int s = 0; s <= P.n_dust; ++s) {
                const HostSpecies& q = staged_[(size_t)t * kMaxSpecies + s];
                if (q.device_ic) {
                    if (q.ic.n0 > 0) hmin = std::min(hmin, q.ic.hloc[0]);
                    if (q.ic.n > q.ic.n0) hmin = std::min(hmin, q.ic.hloc[1]);
                } else {
                    for (int i = 0; i < q.n; ++i) hmin = std::min(hmin, q.h[i]);
                }
            }
            build_grid(P, hmin, C);
        }
        tube_physics(P, C);
        if ((long long)C.ncell * C.nsp > 2000000000ll) return fail(DSH_ERR_ARG, "grid too large");
    }
    // segments: species-major, then tube
    pseg_.clear();
    gseg_.clear();
    cseg_.clear();
    aseg_.clear();
    long long poff = 0, coff = 0, aoff = 0;
    for (int s = 0; s < kMaxSpecies; ++s)
        for (int t = 0; t < T; ++t) {
            TubeConst& C = tc_h_[t];
            if (s >= C.nsp) {
                C.p_off[s] = 0;
                C.np[s] = 0;
                C.cell_off[s] = 0;
                C.pseg[s] = -1;
                continue;
            }
            C.pseg[s] = (int)pseg_.size();
            const int n = staged_[(size_t)t * kMaxSpecies + s].n;
            C.p_off[s] = (int)poff;
            C.np[s] = n;
            C.cell_off[s] = (int)coff;
            pseg_.push_back(Seg{(int)poff, n, t, s});
            if (s == 0) gseg_.push_back(Seg{(int)poff, n, t, s});
            cseg_.push_back(Seg{(int)coff, C.ncell, t, s});
            poff += n;
            coff += C.ncell;
            if (poff > 2000000000ll || coff > 2000000000ll) return fail(DSH_ERR_ARG, "batch too large");
        }
    a_.idic_blocks = a_.idic_runs = 0u;
    for (int t = 0; t < T; ++t) {
        TubeConst& C = tc_h_[t];
        C.avg_off = (int)aoff;
        C.npart = C.f == 1 ? (C.navg + idic_block_cells() - 1) / idic_block_cells() : C.navg;
        (C.f == 1 ? a_.idic_runs : a_.idic_blocks) |= 1u << (C.nsp - 1);
        aseg_.push_back(Seg{(int)aoff, C.navg, t, 0});
        aoff += C.navg;
    }
    ntot_ = poff;
    int rc;
    if ((rc = hip_check(hipStreamCreateWithFlags(&stream_, hipStreamNonBlocking), "hipStreamCreate"))) return rc;

    // block maps
    auto make_map = [&](const std::vector<Seg>& segs, int per_block, Map& m) -> int {
        std::vector<int> blk(segs.size() + 1, 0);
        for (size_t i = 0; i < segs.size(); ++i) blk[i + 1] = blk[i] + (segs[i].n + per_block - 1) / per_block;
        Seg* dseg;
        int* dblk;
        int r;
        if ((r = alloc(&dseg, segs.size()))) return r;
        if ((r = alloc(&dblk, blk.size()))) return r;
        if (!segs.empty() &&
            (r = hip_check(hipMemcpy(dseg, segs.data(), segs.size() * sizeof(Seg), hipMemcpyHostToDevice),
                           "hipMemcpy seg")))
            return r;
        if ((r = hip_check(hipMemcpy(dblk, blk.data(), blk.size() * sizeof(int), hipMemcpyHostToDevice),
                           "hipMemcpy blk")))
            return r;
        m.seg = dseg;
        m.blk = dblk;
        m.nseg = (int)segs.size();
        m.nblocks = blk.back();
        return DSH_OK;
    };
    if ((rc = make_map(pseg_, 256, a_.pmap))) return rc;
    if ((rc = make_map(gseg_, 256, a_.gmap))) return rc;
    if ((rc = make_map(cseg_, 256, a_.cmap))) return rc;
    if ((rc = make_map(aseg_, 4, a_.amap))) return rc;
    if ((rc = make_map(aseg_, idic_block_cells(), a_.amap_runs))) return rc;
    if ((rc = make_map(pseg_, 64, a_.pmap64))) return rc;
    if ((rc = make_map(gseg_, 64, a_.gmap64))) return rc;
    a_.variant = pair_variant_;
    a_.tile = tile_;
    a_.waves = waves_;
    if ((rc = alloc(&a_.masks, mask_words(a_.gmap64.nblocks)))) return rc;
    a_.pband = nullptr;
    a_.bperm = nullptr;
    for (int t = 0; t < T; ++t)
        if (tc_h_[t].precision == DSH_MIXED && tc_h_[t].wall_band) {
            if ((rc = alloc(&a_.pband, (size_t)ntot_))) return rc;
            break;
        }
    // the pair kernels' chunk order: band chunks first (fused kernels), finished tubes' chunks
    // left out (a batch of several tubes)
    if (a_.pband || T > 1)
        if ((rc = alloc(&a_.bperm, bperm_ints(a_.pmap64.nblocks)))) return rc;
    a_.frec = nullptr;
    for (int t = 0; t < T; ++t)
        if (tc_h_[t].precision == DSH_MIXED) {
            // gas occupies the front of the sorted space (species-major)
            long long ngas = 0;
            for (const Seg& g : gseg_) ngas += g.n;
            if ((rc = alloc(&a_.frec, (size_t)std::max(ngas, 1ll)))) return rc;
            break;
        }
    a_.ginv = nullptr;
    for (int t = 0; t < T; ++t)
        if (tc_h_[t].drag_mode == 0) {
            long long ngas = 0;
            for (const Seg& g : gseg_) ngas += g.n;
            if ((rc = alloc(&a_.ginv, (size_t)ngas))) return rc;
            break;
        }

    // particle buffers (two copies of each persistent field)
    PBuf& B = a_.B;
    B.ntot = ntot_;
    const size_t n2 = 2 * (size_t)ntot_;
    double** dfields[] = {&B.x, &B.y, &B.z, &B.vx, &B.vy, &B.vz, &B.m, &B.h, &B.e, &B.rho,
                          &B.p, &B.ts, &B.ax, &B.ay, &B.az, &B.ed};
    for (double** f : dfields) {
        if ((rc = alloc(f, n2))) return rc;
        // fields a species does not use (gas t_stop, dust e/p/A) stay zero in both buffers
        if ((rc = hip_check(hipMemset(*f, 0, n2 * sizeof(double)), "hipMemset"))) return rc;
    }
    if ((rc = alloc(&B.ind, n2))) return rc;
    if ((rc = alloc(&B.id, n2))) return rc;
    int** ifields[] = {&a_.key, &a_.slot, &a_.tsrc, &a_.tkey, &a_.wkey, &a_.ncount};
    for (int** f : ifields)
        if ((rc = alloc(f, (size_t)ntot_))) return rc;
    if ((rc = alloc(&a_.pr2, (size_t)ntot_))) return rc;
    if ((rc = alloc(&a_.csav, (size_t)ntot_))) return rc;
    if ((rc = alloc(&a_.nhash, (size_t)ntot_))) return rc;
    a_.ncells = (int)coff;
    if ((rc = alloc(&a_.count, (size_t)coff))) return rc;
    if ((rc = alloc(&a_.start, (size_t)coff + 1))) return rc;
    if ((rc = alloc(&a_.cellh, (size_t)coff))) return rc;
    // cell counts and cell maxima of h start at zero; each step's scan and radius passes clear
    // them again after use
    if ((rc = hip_check(hipMemset(a_.count, 0, (size_t)coff * sizeof(int)), "clear counts"))) return rc;
    if ((rc = hip_check(hipMemset(a_.cellh, 0, (size_t)coff * sizeof(double)), "clear cell h"))) return rc;
    if ((rc = alloc(&a_.reach, (size_t)coff))) return rc;
    if ((rc = alloc(&a_.wtmp, 2 * (size_t)coff))) return rc;
    a_.scan_nb = scan_blocks((int)coff);
    if ((rc = alloc(&a_.scan_bsum, (size_t)a_.scan_nb))) return rc;
    if ((rc = alloc(&a_.part, 5 * (size_t)aoff))) return rc;
    TubeConst* dtc;
    TubeDyn* dtd;
    if ((rc = alloc(&dtc, (size_t)T))) return rc;
    if ((rc = alloc(&dtd, (size_t)T))) return rc;
    if ((rc = alloc(&a_.cur, 2))) return rc;   // [0] current buffer, [1] finalize ticket
    a_.tc = dtc;
    a_.td = dtd;
    a_.ntubes = T;
    if ((rc = hip_check(hipMemcpy(dtc, tc_h_.data(), sizeof(TubeConst) * T, hipMemcpyHostToDevice), "H2D tubes")))
        return rc;
    const int zero[2] = {0, 0};
    if ((rc = hip_check(hipMemcpy(a_.cur, zero, sizeof(zero), hipMemcpyHostToDevice), "H2D cur"))) return rc;

    // adopted tubes: device-to-device copy of the source's current buffer (its sorted
    // order and ids are kept; the first step re-sorts to the same order)
    std::vector<TubeDyn> adopted_dyn;
    for (int t = 0; t < T; ++t) {
        if (!migr_[t].src) continue;
        const Engine& S = *migr_[t].src;
        const int st = migr_[t].tube;
        if ((rc = hip_check(hipStreamSynchronize(S.stream_), "source sync"))) return rc;
        int scur = 0;
        if ((rc = hip_check(hipMemcpy(&scur, S.a_.cur, sizeof(int), hipMemcpyDeviceToHost), "D2H cur"))) return rc;
        TubeDyn sdyn;
        if ((rc = hip_check(hipMemcpy(&sdyn, S.a_.td + st, sizeof(TubeDyn), hipMemcpyDeviceToHost), "D2H dyn")))
            return rc;
        if (sdyn.done) scur = sdyn.done_cur;   // a finished tube's state stays in done_cur
        for (int s = 0; s < tc_h_[t].nsp; ++s) {
            const size_t n = (size_t)tc_h_[t].np[s];
            const size_t so = (size_t)scur * S.ntot_ + (size_t)S.tc_h_[st].p_off[s];
            const size_t off = (size_t)tc_h_[t].p_off[s];
            const PBuf& SB = S.a_.B;
            const double* sd[] = {SB.x, SB.y, SB.z, SB.vx, SB.vy, SB.vz, SB.m, SB.h, SB.e, SB.rho, SB.p, SB.ts};
            double* dd[] = {B.x, B.y, B.z, B.vx, B.vy, B.vz, B.m, B.h, B.e, B.rho, B.p, B.ts};
            for (int f = 0; f < 12; ++f)
                if ((rc = hip_check(hipMemcpy(dd[f] + off, sd[f] + so, n * sizeof(double), hipMemcpyDeviceToDevice),
                                    "D2D field")))
                    return rc;
            if ((rc = hip_check(hipMemcpy(B.ind + off, SB.ind + so, n * sizeof(int), hipMemcpyDeviceToDevice),
                                "D2D ind")))
                return rc;
            if ((rc = hip_check(hipMemcpy(B.id + off, SB.id + so, n * sizeof(int), hipMemcpyDeviceToDevice),
                                "D2D id")))
                return rc;
        }
    }
    // uploaded tubes: buffer 0, original order, id = index
    for (int t = 0; t < T; ++t)
        for (int s = 0; s < tc_h_[t].nsp; ++s) {
            if (migr_[t].src) continue;
            const HostSpecies& q = staged_[(size_t)t * kMaxSpecies + s];
            const size_t off = (size_t)tc_h_[t].p_off[s];
            const size_t n = (size_t)q.n;
            if (n == 0) continue;
            if (q.device_ic) {
                if ((rc = hip_check(launch_tube_ic(q.ic, B, (long long)off, stream_), "k_tube_ic"))) return rc;
                continue;
            }
            struct {
                double* dst;
                const std::vector<double>& src;
            } cp[] = {{B.x, q.x},   {B.y, q.y},   {B.z, q.z}, {B.vx, q.vx}, {B.vy, q.vy}, {B.vz, q.vz},
                      {B.m, q.m},   {B.h, q.h},   {B.e, q.e}, {B.rho, q.rho}, {B.p, q.p}, {B.ts, q.ts}};
            for (auto& c : cp)
                if ((rc = hip_check(hipMemcpy(c.dst + off, c.src.data(), n * sizeof(double), hipMemcpyHostToDevice),
                                    "H2D field")))
                    return rc;
            if ((rc = hip_check(hipMemcpy(B.ind + off, q.ind.data(), n * sizeof(int), hipMemcpyHostToDevice),
                                "H2D ind")))
                return rc;
            std::vector<int> ids(n);
            std::iota(ids.begin(), ids.end(), 0);
            if ((rc = hip_check(hipMemcpy(B.id + off, ids.data(), n * sizeof(int), hipMemcpyHostToDevice),
                                "H2D id")))
                return rc;
            std::vector<double> zeros(n, 0.0);
            for (double* d : {B.ax, B.ay, B.az, B.ed})
                for (int b = 0; b < 2; ++b)
                    if ((rc = hip_check(hipMemcpy(d + b * ntot_ + off, zeros.data(), n * sizeof(double),
                                                  hipMemcpyHostToDevice),
                                        "H2D zero")))
                        return rc;
        }
    double* prep = nullptr;
    if ((rc = alloc(&prep, kPrepParts * (size_t)a_.pmap.nblocks))) return rc;
    if ((rc = hip_check(launch_prepare(a_, prep, stream_), "k_prepare"))) return rc;
    if ((rc = hip_check(hipStreamSynchronize(stream_), "prepare sync"))) return rc;
    // adopted tubes continue with their own clock and step size
    for (int t = 0; t < T; ++t) {
        if (!migr_[t].src) continue;
        TubeDyn sd{};
        if ((rc = hip_check(hipMemcpy(&sd, migr_[t].src->a_.td + migr_[t].tube, sizeof(TubeDyn),
                                      hipMemcpyDeviceToHost),
                            "D2H source dyn")))
            return rc;
        TubeDyn d{};
        d.t = sd.t;
        d.dt = sd.dt;
        d.steps = sd.steps;
        d.done = sd.done;
        d.final_step = sd.final_step;
        d.err = sd.err;
        d.done_cur = 0;
        for (int c = 0; c < 4; ++c) d.cons[c] = sd.cons[c];
        d.xw_lo = sd.xw_lo;
        d.xw_hi = sd.xw_hi;
        d.band_lo = sd.band_lo;
        d.band_hi = sd.band_hi;
        if ((rc = hip_check(hipMemcpy(a_.td + t, &d, sizeof(TubeDyn), hipMemcpyHostToDevice), "H2D dyn"))) return rc;
        migr_[t].src = nullptr;
    }
    prepared_ = true;
    return refresh_dyn();
}

namespace {
// rank_permute also takes each cell's largest h
const char* const kStageNames[] = {"key",     "scan",  "scatter", "rank_permute", "radius",
                                   "density", "force", "idic",    "finalize"};
constexpr int kStages = (int)(sizeof(kStageNames) / sizeof(kStageNames[0]));
}  // namespace

// Enqueue one step.  With timed, an event is recorded before the first and after every
// stage on the step stream; the elapsed times are collected at the next sync().
int Engine::enqueue_step(bool timed) {
    hipEvent_t* ev = nullptr;
    if (timed) {
        const size_t base = events_.size();
        events_.resize(base + kStages + 1);
        for (size_t i = base; i < events_.size(); ++i)
            if (hipEventCreate(&events_[i]) != hipSuccess) return fail(DSH_ERR_HIP, "hipEventCreate");
        ev = events_.data() + base;
        (void)hipEventRecord(ev[0], stream_);
        ++pending_timed_;
    }
    for (int i = 0; i < kStages; ++i) {
        hipError_t e = hipSuccess;
        switch (i) {
            case 0: e = launch_key(a_, stream_); break;
            case 1: e = launch_scan(a_, stream_); break;
            case 2: e = launch_scatter(a_, stream_); break;
            case 3: e = launch_rank_permute(a_, stream_); break;
            case 4: e = launch_radius(a_, stream_); break;
            case 5: e = launch_density(a_, precision_, stats_, stream_); break;
            case 6: e = launch_force(a_, precision_, stream_); break;
            case 7: e = launch_idic(a_, stream_); break;
            case 8: e = launch_finalize(a_, stream_); break;
        }
        if (e != hipSuccess) return hip_check(e, kStageNames[i]);
        if (timed) (void)hipEventRecord(ev[i + 1], stream_);
    }
    return DSH_OK;
}

int Engine::collect_times() {
    if (pending_timed_ == 0) return DSH_OK;
    std::vector<double> tot(kStages, 0.0);
    for (int k = 0; k < pending_timed_; ++k) {
        hipEvent_t* ev = events_.data() + (size_t)k * (kStages + 1);
        for (int i = 0; i < kStages; ++i) {
            float ms = 0.f;
            if (hipEventElapsedTime(&ms, ev[i], ev[i + 1]) != hipSuccess) return fail(DSH_ERR_HIP, "event time");
            tot[i] += ms;
        }
    }
    ktimes_.clear();
    for (int i = 0; i < kStages; ++i) ktimes_.emplace_back(kStageNames[i], tot[i] / pending_timed_);
    timed_steps_ = pending_timed_;
    for (hipEvent_t e : events_) (void)hipEventDestroy(e);
    events_.clear();
    pending_timed_ = 0;
    return DSH_OK;
}

// Captures `steps` consecutive steps of the step stream into one executable graph.
int Engine::build_graph(int steps, hipGraph_t& graph, hipGraphExec_t& exec) {
    if (exec) {
        (void)hipGraphExecDestroy(exec);
        exec = nullptr;
    }
    if (graph) {
        (void)hipGraphDestroy(graph);
        graph = nullptr;
    }
    int rc;
    if ((rc = hip_check(hipStreamBeginCapture(stream_, hipStreamCaptureModeThreadLocal), "begin capture")))
        return rc;
    int r = 0;
    for (int k = 0; k < steps && r == 0; ++k) r = enqueue_step(false);
    hipGraph_t g = nullptr;
    const hipError_t e = hipStreamEndCapture(stream_, &g);
    if (r) {
        if (g) (void)hipGraphDestroy(g);
        return r;
    }
    if ((rc = hip_check(e, "end capture"))) return rc;
    graph = g;
    return hip_check(hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0), "graph instantiate");
}

int Engine::launch_steps(int nsteps) {
    int rc;
    if (!prepared_ && (rc = prepare())) return rc;
    for (int k = 0; k < nsteps; ++k) {
        if (timing_) {
            if ((rc = enqueue_step(true))) return rc;
        } else if (use_graph_) {
            if (graph_dirty_) {
                if ((rc = build_graph(1, graph_, graph_exec_))) return rc;
                if ((rc = build_graph(kGraphSteps, graph_n_, graph_exec_n_))) return rc;
                graph_dirty_ = false;
            }
            // kGraphSteps steps per launch while that many are left
            const bool multi = nsteps - k >= kGraphSteps;
            if ((rc = hip_check(hipGraphLaunch(multi ? graph_exec_n_ : graph_exec_, stream_), "hipGraphLaunch")))
                return rc;
            if (multi) k += kGraphSteps - 1;
        } else {
            if ((rc = enqueue_step(false))) return rc;
        }
    }
    return DSH_OK;
}

int Engine::sync() {
    int rc;
    if ((rc = hip_check(hipStreamSynchronize(stream_), "step sync"))) return rc;
    if ((rc = collect_times())) return rc;
    return refresh_dyn();
}

// Steps are enqueued in chunks; between chunks the host reads the tube states and stops
// once every tube has finished, so a caller may ask for more steps than a run needs.
int Engine::run(int nsteps) {
    constexpr int kChunk = 64;
    int rc;
    for (int left = nsteps; left > 0;) {
        const int k = left < kChunk ? left : kChunk;
        if ((rc = launch_steps(k))) return rc;
        if ((rc = sync())) return rc;
        left -= k;
        bool all_done = true;
        for (const TubeDyn& d : dyn_h_) all_done = all_done && (d.done || d.err);
        if (all_done) break;
    }
    if (nsteps == 0) {
        if ((rc = launch_steps(0))) return rc;
        if ((rc = sync())) return rc;
    }
    return error_status();
}

int Engine::autotune(int steps_per_trial, bool with_variants, int* tile_out, int* variant_out, double* ms_out) {
    if (steps_per_trial < 1) return fail(DSH_ERR_ARG, "steps_per_trial must be positive");
    int rc;
    if (!prepared_ && (rc = prepare())) return rc;
    struct Cand {
        int tile, waves, variant;
    };
    std::vector<Cand> cands;
    for (int w : kBlockWaves)
        for (int t : kTiles) cands.push_back({t, w, kPairTiled});
    if (with_variants) cands.push_back({192, 1, kPairDirect});
    auto finished = [&]() {
        for (const TubeDyn& d : dyn_h_)
            if (!(d.done || d.err)) return false;
        return true;
    };
    hipEvent_t e0 = nullptr, e1 = nullptr;
    if ((rc = hip_check(hipEventCreate(&e0), "hipEventCreate"))) return rc;
    if ((rc = hip_check(hipEventCreate(&e1), "hipEventCreate"))) {
        (void)hipEventDestroy(e0);
        return rc;
    }
    Cand best{tile_, waves_, pair_variant_};
    double best_ms = -1.0;
    for (const Cand& c : cands) {
        set_tile(c.tile);
        set_block_waves(c.waves);
        set_pair_variant(c.variant);
        if ((rc = run(1))) break;   // builds the graph, warms the caches
        if (finished()) break;
        if ((rc = hip_check(hipEventRecord(e0, stream_), "hipEventRecord"))) break;
        if ((rc = launch_steps(steps_per_trial))) break;
        if ((rc = hip_check(hipEventRecord(e1, stream_), "hipEventRecord"))) break;
        if ((rc = sync())) break;
        float ms = 0.f;
        if ((rc = hip_check(hipEventElapsedTime(&ms, e0, e1), "hipEventElapsedTime"))) break;
        const double per = (double)ms / steps_per_trial;
        if (best_ms < 0.0 || per < best_ms) {
            best_ms = per;
            best = c;
        }
        if (finished()) break;
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    set_tile(best.tile);
    set_block_waves(best.waves);
    set_pair_variant(best.variant);
    if (rc) return rc;
    if (tile_out) *tile_out = best.tile;
    if (variant_out) *variant_out = best.variant;
    if (ms_out) *ms_out = best_ms;
    return error_status();
}

int Engine::refresh_dyn() {
    dyn_h_.resize(ntubes());
    return hip_check(hipMemcpy(dyn_h_.data(), a_.td, sizeof(TubeDyn) * ntubes(), hipMemcpyDeviceToHost), "D2H dyn");
}

int Engine::error_status() {
    for (int t = 0; t < ntubes(); ++t) {
        const int e = dyn_h_[t].err;
        if (e & kErrGrid)
            return fail(DSH_ERR_GRID, "tube " + std::to_string(t) +
                                          ": smoothing length grew beyond the neighbour grid's search radius");
        if (e & kErrNonFinite)
            return fail(DSH_ERR_NONFINITE, "tube " + std::to_string(t) + ": non-finite state");
    }
    return DSH_OK;
}

int Engine::download(int tube, int species, const dsh_soa_view* v) {
    if (tube < 0 || tube >= ntubes()) return fail(DSH_ERR_ARG, "tube index out of range");
    if (!v) return fail(DSH_ERR_ARG, "null view");
    const dsh_params& P = params_[tube];
    if (species < 0 || species > P.n_dust) return fail(DSH_ERR_ARG, "species index out of range");
    int rc;
    if (!prepared_ && (rc = prepare())) return rc;
    if ((rc = hip_check(hipStreamSynchronize(stream_), "download sync"))) return rc;
    if ((rc = refresh_dyn())) return rc;
    int cur = 0;
    if ((rc = hip_check(hipMemcpy(&cur, a_.cur, sizeof(int), hipMemcpyDeviceToHost), "D2H cur"))) return rc;
    const TubeConst& C = tc_h_[tube];
    const TubeDyn& D = dyn_h_[tube];
    // a finished tube is not re-sorted: its final state stays in buffer done_cur
    if (D.done) cur = D.done_cur;
    const int derived = cur;
    const size_t n = (size_t)C.np[species];
    const size_t off = (size_t)C.p_off[species];
    std::vector<int> ids(n), idd(n);
    if ((rc = hip_check(hipMemcpy(ids.data(), a_.B.id + (size_t)cur * ntot_ + off, n * sizeof(int),
                                  hipMemcpyDeviceToHost),
                        "D2H id")))
        return rc;
    if ((rc = hip_check(hipMemcpy(idd.data(), a_.B.id + (size_t)derived * ntot_ + off, n * sizeof(int),
                                  hipMemcpyDeviceToHost),
                        "D2H id")))
        return rc;
    std::vector<double> tmp(n);
    auto get = [&](double* out, const double* base, int buf, const std::vector<int>& order) -> int {
        if (!out) return DSH_OK;
        int r = hip_check(hipMemcpy(tmp.data(), base + (size_t)buf * ntot_ + off, n * sizeof(double),
                                    hipMemcpyDeviceToHost),
                          "D2H field");
        if (r) return r;
        for (size_t q = 0; q < n; ++q) out[order[q]] = tmp[q];
        return DSH_OK;
    };
    const PBuf& B = a_.B;
    struct {
        double* out;
        const double* base;
        bool deriv;
    } fs[] = {{v->x, B.x, false},   {v->y, B.y, false},   {v->z, B.z, false},    {v->vx, B.vx, false},
              {v->vy, B.vy, false}, {v->vz, B.vz, false}, {v->m, B.m, false},    {v->rho, B.rho, false},
              {v->p, B.p, false},   {v->e, B.e, false},   {v->h, B.h, false},    {v->t_stop, B.ts, false},
              {v->ax, B.ax, true},  {v->ay, B.ay, true},  {v->az, B.az, true},   {v->edot, B.ed, true}};
    for (auto& f : fs)
        if ((rc = get(f.out, f.base, f.deriv ? derived : cur, f.deriv ? idd : ids))) return rc;
    if (v->ind) {
        std::vector<int> ti(n);
        if ((rc = hip_check(hipMemcpy(ti.data(), B.ind + (size_t)cur * ntot_ + off, n * sizeof(int),
                                      hipMemcpyDeviceToHost),
                            "D2H ind")))
            return rc;
        for (size_t q = 0; q < n; ++q) v->ind[ids[q]] = ti[q];
    }
    if (v->cell || v->nn) {
        std::vector<int> wk(n);
        const bool stepped = D.steps > 0;
        if (stepped &&
            (rc = hip_check(hipMemcpy(wk.data(), a_.wkey + off, n * sizeof(int), hipMemcpyDeviceToHost), "D2H key")))
            return rc;
        for (size_t q = 0; q < n; ++q) {
            const int id = idd[q];
            if (v->cell) v->cell[id] = stepped ? wk[q] - C.cell_off[species] : -1;
            if (v->nn) v->nn[id] = -1;
        }
        if (v->nn && stepped) {
            // the reference lists run through a cell's members in index order (kernel.cu:1937-1995);
            // the device keeps a cell's members in (x, id) order, so each run of equal keys is
            // re-ordered by id here
            std::vector<int> members;
            for (size_t q0 = 0; q0 < n;) {
                size_t q1 = q0 + 1;
                while (q1 < n && wk[q1] == wk[q0]) ++q1;
                members.assign(idd.begin() + q0, idd.begin() + q1);
                std::sort(members.begin(), members.end());
                for (size_t k = 0; k + 1 < members.size(); ++k) v->nn[members[k]] = members[k + 1];
                q0 = q1;
            }
        }
    }
    return DSH_OK;
}

int Engine::sorted_ids(int tube, int species, int* ids) {
    if (tube < 0 || tube >= ntubes() || species < 0 || !ids) return fail(DSH_ERR_ARG, "bad tube/species");
    if (!prepared_ || dyn_h_.empty() || species >= tc_h_[tube].nsp) return fail(DSH_ERR_STATE, "no step yet");
    int rc;
    if ((rc = hip_check(hipStreamSynchronize(stream_), "sorted ids sync"))) return rc;
    int cur = 0;
    if ((rc = hip_check(hipMemcpy(&cur, a_.cur, sizeof(int), hipMemcpyDeviceToHost), "D2H cur"))) return rc;
    if ((rc = refresh_dyn())) return rc;
    // the current buffer holds the last step's particles in that step's sorted order (a
    // finished tube's, in its last step's order, buffer done_cur)
    if (dyn_h_[tube].done) cur = dyn_h_[tube].done_cur;
    return hip_check(hipMemcpy(ids, a_.B.id + (size_t)cur * ntot_ + tc_h_[tube].p_off[species],
                               sizeof(int) * (size_t)tc_h_[tube].np[species], hipMemcpyDeviceToHost),
                     "D2H sorted ids");
}

int Engine::neighbour_stats(int tube, int species, int* count, uint64_t* hash) {
    if (tube < 0 || tube >= ntubes() || species < 0 || species >= tc_h_[tube].nsp)
        return fail(DSH_ERR_ARG, "bad tube/species");
    int rc;
    if ((rc = hip_check(hipStreamSynchronize(stream_), "stats sync"))) return rc;
    if ((rc = refresh_dyn())) return rc;
    int cur = 0;
    if ((rc = hip_check(hipMemcpy(&cur, a_.cur, sizeof(int), hipMemcpyDeviceToHost), "D2H cur"))) return rc;
    const TubeDyn& D = dyn_h_[tube];
    const int derived = D.done ? D.done_cur : cur;   // a finished tube's state stays in done_cur
    const size_t n = (size_t)tc_h_[tube].np[species];
    const size_t off = (size_t)tc_h_[tube].p_off[species];
    std::vector<int> ids(n), c(n);
    std::vector<unsigned long long> h(n);
    if ((rc = hip_check(hipMemcpy(ids.data(), a_.B.id + (size_t)derived * ntot_ + off, n * sizeof(int),
                                  hipMemcpyDeviceToHost),
                        "D2H id")))
        return rc;
    if ((rc = hip_check(hipMemcpy(c.data(), a_.ncount + off, n * sizeof(int), hipMemcpyDeviceToHost), "D2H cnt")))
        return rc;
    if ((rc = hip_check(hipMemcpy(h.data(), a_.nhash + off, n * sizeof(unsigned long long), hipMemcpyDeviceToHost),
                        "D2H hash")))
        return rc;
    for (size_t q = 0; q < n; ++q) {
        if (count) count[ids[q]] = c[q];
        if (hash) hash[ids[q]] = (uint64_t)h[q];
    }
    return DSH_OK;
}

}  // namespace dsh
