// dsh_api.cpp - the single-tube C ABI (include/dsh.h) over the batch engine.
#include <cstring>
#include <memory>
#include <string>

#include "../../include/dsh.h"
#include "dsh_engine.h"

struct dsh_ctx {
    std::unique_ptr<dsh::Engine> eng;
    std::string err;
    int set(int rc) {
        if (rc != DSH_OK && eng) err = eng->error();
        return rc;
    }
};

extern "C" {

int dsh_create(const dsh_params* params, dsh_ctx** out) {
    if (!params || !out) return DSH_ERR_ARG;
    if (params->n_dust < 0 || params->n_dust > DSH_MAX_DUST) return DSH_ERR_ARG;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) return DSH_ERR_HIP;
    dsh_ctx* c = new dsh_ctx();
    c->eng.reset(new dsh::Engine(std::vector<dsh_params>{*params}));
    *out = c;
    return DSH_OK;
}

int dsh_upload_species(dsh_ctx* c, int species, const dsh_soa_view* v, int pm) {
    if (!c) return DSH_ERR_ARG;
    return c->set(c->eng->upload(0, species, v, pm));
}

int dsh_step(dsh_ctx* c, int nsteps, double* t_inout) {
    if (!c || nsteps < 0) return DSH_ERR_ARG;
    int rc = c->set(c->eng->run(nsteps));
    if (t_inout && c->eng->prepared()) *t_inout = c->eng->dyn(0).t;
    return rc;
}

int dsh_download_species(dsh_ctx* c, int species, const dsh_soa_view* v) {
    if (!c) return DSH_ERR_ARG;
    return c->set(c->eng->download(0, species, v));
}

int dsh_conservation(dsh_ctx* c, double out[4]) {
    if (!c || !out) return DSH_ERR_ARG;
    int rc;
    if (!c->eng->prepared() && (rc = c->set(c->eng->prepare()))) return rc;
    if ((rc = c->set(c->eng->sync()))) return rc;
    for (int a = 0; a < 4; ++a) out[a] = c->eng->dyn(0).cons[a];
    return DSH_OK;
}

int dsh_status(dsh_ctx* c, double* t, double* dt, int* steps, int* done) {
    if (!c) return DSH_ERR_ARG;
    int rc;
    if (!c->eng->prepared() && (rc = c->set(c->eng->prepare()))) return rc;
    if ((rc = c->set(c->eng->sync()))) return rc;
    const dsh::TubeDyn& D = c->eng->dyn(0);
    if (t) *t = D.t;
    if (dt) *dt = D.dt;
    if (steps) *steps = D.steps;
    if (done) *done = D.done;
    return DSH_OK;
}

int dsh_grid(dsh_ctx* c, double dbl[9], int ints[5]) {
    if (!c) return DSH_ERR_ARG;
    int rc;
    if (!c->eng->prepared() && (rc = c->set(c->eng->prepare()))) return rc;
    const dsh::TubeConst& T = c->eng->tube_const(0);
    for (int a = 0; a < 3; ++a) {
        if (dbl) {
            dbl[a] = T.o[a];
            dbl[3 + a] = T.c[a];
            dbl[6 + a] = T.L[a];
        }
        if (ints) ints[a] = T.n[a];
    }
    if (ints) {
        ints[3] = T.f;
        ints[4] = T.navg;
    }
    return DSH_OK;
}

int dsh_set_pair_variant(dsh_ctx* c, int variant) {
    if (!c || (variant != DSH_PAIR_DIRECT && variant != DSH_PAIR_TILED)) return DSH_ERR_ARG;
    c->eng->set_pair_variant(variant);
    return DSH_OK;
}

int dsh_set_tile(dsh_ctx* c, int tile) {
    if (!c || !dsh::valid_tile(tile)) return DSH_ERR_ARG;
    c->eng->set_tile(tile);
    return DSH_OK;
}

int dsh_set_block_waves(dsh_ctx* c, int waves) {
    if (!c || !dsh::valid_block_waves(waves)) return DSH_ERR_ARG;
    c->eng->set_block_waves(waves);
    return DSH_OK;
}

int dsh_block_shape(dsh_ctx* c, int shape[2]) {
    if (!c || !shape) return DSH_ERR_ARG;
    shape[0] = c->eng->tile();
    shape[1] = c->eng->block_waves();
    return DSH_OK;
}

int dsh_autotune(dsh_ctx* c, int steps_per_trial, int with_variants, int chosen[2], double* ms_per_step) {
    if (!c || steps_per_trial < 1) return DSH_ERR_ARG;
    int tile = 0, variant = 0;
    const int rc = c->set(c->eng->autotune(steps_per_trial, with_variants != 0, &tile, &variant, ms_per_step));
    if (chosen) {
        chosen[0] = tile;
        chosen[1] = variant;
    }
    return rc;
}

int dsh_set_stats(dsh_ctx* c, int on) {
    if (!c) return DSH_ERR_ARG;
    c->eng->set_stats(on != 0);
    return DSH_OK;
}

int dsh_neighbour_stats(dsh_ctx* c, int species, int* count, uint64_t* hash) {
    if (!c) return DSH_ERR_ARG;
    if (!c->eng->prepared()) return DSH_ERR_STATE;
    return c->set(c->eng->neighbour_stats(0, species, count, hash));
}

int dsh_sorted_ids(dsh_ctx* c, int species, int* ids) {
    if (!c) return DSH_ERR_ARG;
    return c->set(c->eng->sorted_ids(0, species, ids));
}

int dsh_set_timing(dsh_ctx* c, int on) {
    if (!c) return DSH_ERR_ARG;
    c->eng->set_timing(on != 0);
    return DSH_OK;
}

int dsh_kernel_times(dsh_ctx* c, double* ms, char* names, int max_kernels, int name_len) {
    if (!c || max_kernels < 0) return DSH_ERR_ARG;
    const auto& kt = c->eng->kernel_times();
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

const char* dsh_last_error(dsh_ctx* c) {
    if (!c) return "null context";
    return c->err.c_str();
}

void dsh_destroy(dsh_ctx* c) { delete c; }

}  // extern "C"
