// dsh_batch.cpp - C ABI of a batch of independent tubes over the engine.
#include <cstring>
#include <memory>
#include <string>
#include <vector>

#include "../../include/dsh.h"
#include "dsh_engine.h"

struct dsh_batch {
    std::unique_ptr<dsh::Engine> eng;
    std::string err;
    int set(int rc) {
        if (rc != DSH_OK && eng) err = eng->error();
        return rc;
    }
};

namespace dsh {

// Seeded set-up of one tube into an engine slot: each species is generated on the device
// at prepare() (no host arrays, no upload).
int upload_tube_spec(Engine& eng, int tube, const dsh_tube_spec& spec) {
    for (int s = 0; s < 2; ++s) {
        TubeIC c;
        int rc = make_tube_ic(spec, s, c);
        if (rc) return rc;
        if ((rc = eng.stage_ic(tube, s, c))) return rc;
    }
    return DSH_OK;
}

}  // namespace dsh

extern "C" {

int dsh_batch_create(const dsh_params* params, int ntubes, dsh_batch** out) {
    if (!params || ntubes < 1 || !out) return DSH_ERR_ARG;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) return DSH_ERR_HIP;
    std::vector<dsh_params> p(params, params + ntubes);
    dsh_batch* b = new dsh_batch();
    b->eng.reset(new dsh::Engine(p));
    *out = b;
    return DSH_OK;
}

int dsh_batch_from_specs(const dsh_tube_spec* specs, int ntubes, dsh_batch** out) {
    if (!specs || ntubes < 1 || !out) return DSH_ERR_ARG;
    std::vector<dsh_params> p(ntubes);
    for (int i = 0; i < ntubes; ++i) {
        const int rc = dsh_tube_params(&specs[i], &p[i]);
        if (rc) return rc;
    }
    dsh_batch* b = nullptr;
    int rc = dsh_batch_create(p.data(), ntubes, &b);
    if (rc) return rc;
    for (int i = 0; i < ntubes; ++i) {
        rc = dsh::upload_tube_spec(*b->eng, i, specs[i]);
        if (rc) {
            delete b;
            return rc;
        }
    }
    *out = b;
    return DSH_OK;
}

int dsh_batch_upload(dsh_batch* b, int tube, int species, const dsh_soa_view* view, int pm) {
    if (!b) return DSH_ERR_ARG;
    return b->set(b->eng->upload(tube, species, view, pm));
}

int dsh_batch_run(dsh_batch* b, int nsteps, int* ndone) {
    if (!b || nsteps < 0) return DSH_ERR_ARG;
    const int rc = b->set(b->eng->run(nsteps));
    if (ndone && b->eng->prepared()) {
        int d = 0;
        for (int t = 0; t < b->eng->ntubes(); ++t) d += b->eng->dyn(t).done ? 1 : 0;
        *ndone = d;
    }
    return rc;
}

int dsh_batch_status(dsh_batch* b, int tube, double* t, double* dt, int* steps, int* done) {
    if (!b || tube < 0 || tube >= b->eng->ntubes()) return DSH_ERR_ARG;
    int rc;
    if (!b->eng->prepared() && (rc = b->set(b->eng->prepare()))) return rc;
    const dsh::TubeDyn& D = b->eng->dyn(tube);
    if (t) *t = D.t;
    if (dt) *dt = D.dt;
    if (steps) *steps = D.steps;
    if (done) *done = D.done;
    return DSH_OK;
}

int dsh_batch_download(dsh_batch* b, int tube, int species, const dsh_soa_view* view) {
    if (!b) return DSH_ERR_ARG;
    return b->set(b->eng->download(tube, species, view));
}

int dsh_batch_conservation(dsh_batch* b, int tube, double out[4]) {
    if (!b || !out || tube < 0 || tube >= b->eng->ntubes()) return DSH_ERR_ARG;
    int rc;
    if (!b->eng->prepared() && (rc = b->set(b->eng->prepare()))) return rc;
    if ((rc = b->set(b->eng->sync()))) return rc;
    for (int a = 0; a < 4; ++a) out[a] = b->eng->dyn(tube).cons[a];
    return DSH_OK;
}

int dsh_batch_set_pair_variant(dsh_batch* b, int variant) {
    if (!b || (variant != DSH_PAIR_DIRECT && variant != DSH_PAIR_TILED)) return DSH_ERR_ARG;
    b->eng->set_pair_variant(variant);
    return DSH_OK;
}

int dsh_batch_set_tile(dsh_batch* b, int tile) {
    if (!b || !dsh::valid_tile(tile)) return DSH_ERR_ARG;
    b->eng->set_tile(tile);
    return DSH_OK;
}

int dsh_batch_set_block_waves(dsh_batch* b, int waves) {
    if (!b || !dsh::valid_block_waves(waves)) return DSH_ERR_ARG;
    b->eng->set_block_waves(waves);
    return DSH_OK;
}

int dsh_batch_block_shape(dsh_batch* b, int shape[2]) {
    if (!b || !shape) return DSH_ERR_ARG;
    shape[0] = b->eng->tile();
    shape[1] = b->eng->block_waves();
    return DSH_OK;
}

int dsh_batch_autotune(dsh_batch* b, int steps_per_trial, int with_variants, int chosen[2], double* ms_per_step) {
    if (!b || steps_per_trial < 1) return DSH_ERR_ARG;
    int tile = 0, variant = 0;
    const int rc = b->set(b->eng->autotune(steps_per_trial, with_variants != 0, &tile, &variant, ms_per_step));
    if (chosen) {
        chosen[0] = tile;
        chosen[1] = variant;
    }
    return rc;
}

int dsh_batch_set_timing(dsh_batch* b, int on) {
    if (!b) return DSH_ERR_ARG;
    b->eng->set_timing(on != 0);
    return DSH_OK;
}

int dsh_batch_kernel_times(dsh_batch* b, double* ms, char* names, int max_kernels, int name_len) {
    if (!b || max_kernels < 0) return DSH_ERR_ARG;
    const auto& kt = b->eng->kernel_times();
    const int n = (int)kt.size() < max_kernels ? (int)kt.size() : max_kernels;
    for (int i = 0; i < n; ++i) {
        if (ms) ms[i] = kt[i].second;
        if (names && name_len > 0) {
            std::strncpy(names + (size_t)i * name_len, kt[i].first.c_str(), name_len - 1);
            names[(size_t)i * name_len + name_len - 1] = '\0';
        }
    }
    return (int)kt.size();
}

int dsh_batch_device_bytes(dsh_batch* b, double* bytes) {
    if (!b || !bytes) return DSH_ERR_ARG;
    *bytes = (double)b->eng->device_bytes();
    return DSH_OK;
}

const char* dsh_batch_last_error(dsh_batch* b) { return b ? b->err.c_str() : "null batch"; }

void dsh_batch_destroy(dsh_batch* b) { delete b; }

}  // extern "C"
