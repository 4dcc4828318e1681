// Drives the host-side C++ (Init.txt reader, reference and tube set-up, text output, the
// serving scheduler on its host-only fake device, argument checks of the C ABI) and the
// oracle under AddressSanitizer and UndefinedBehaviorSanitizer.  No GPU call is made: the
// scheduler runs on the fake backend and the context entry points are called only with
// arguments they reject before touching a device.  Exits non-zero on the first failed check;
// the sanitizers abort on the first error they find.
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/dsh.h"
#include "../../oracle/oracle.h"

namespace {

int g_fail = 0;

#define CHECK(cond)                                                          \
    do {                                                                     \
        if (!(cond)) {                                                       \
            std::fprintf(stderr, "%s:%d: check failed: %s\n", __FILE__, __LINE__, #cond); \
            g_fail = 1;                                                      \
        }                                                                    \
    } while (0)

struct Arrays {
    std::vector<double> d[17];
    std::vector<int> ind, nn, cell;
    dsh_soa_view v{};
    explicit Arrays(int n) {
        for (auto& a : d) a.assign((size_t)n, 0.0);
        ind.assign((size_t)n, 0);
        nn.assign((size_t)n, 0);
        cell.assign((size_t)n, 0);
        double** f[17] = {&v.x, &v.y, &v.z, &v.vx, &v.vy, &v.vz, &v.ax, &v.ay, &v.az,
                          &v.m, &v.rho, &v.p, &v.e, &v.h, &v.t_stop, &v.edot, nullptr};
        for (int k = 0; k < 16; ++k) *f[k] = d[k].data();
        v.ind = ind.data();
        v.nn = nn.data();
        v.cell = cell.data();
    }
};

dsh_tube_spec small_tube(int index, int precision, int max_steps) {
    dsh_tube_spec s{};
    s.x_min = -0.5;
    s.x_max = 0.5;
    s.y_width = s.z_width = 0.03;
    s.dx_left = 0.01;
    s.rho_left = 1.0;
    s.dust_to_gas = 1.0;
    s.K = 1.0;
    s.jitter = 1e-3;
    s.frozen_h = 2.5;
    s.t_end = 1e30;
    s.cfl = 0.3;
    s.dt_mode = DSH_DT_CFL_ONCE;
    s.max_steps = max_steps;
    s.precision = precision;
    s.tube_index = index;
    s.seed = 12345u + (uint64_t)index;
    return s;
}

void reference_problem(const char* init_path, const char* out_dir) {
    dsh_params p{};
    CHECK(dsh_read_init(init_path, &p) == DSH_OK);
    CHECK(dsh_read_init("/nonexistent/Init.txt", &p) != DSH_OK);
    CHECK(dsh_read_init(init_path, &p) == DSH_OK);
    int n[3] = {0, 0, 0};
    CHECK(dsh_reference_setup(&p, n) == DSH_OK);
    CHECK(n[0] > 0 && n[1] == n[0] && n[2] == n[0]);
    for (int s = 0; s < 3; ++s) {
        Arrays a(n[s]);
        CHECK(dsh_reference_ic(&p, s, &a.v) == DSH_OK);
        CHECK(dsh_reference_nn(&p, a.v.x, a.v.y, a.v.z, n[s] - 1, a.nn.data()) == DSH_OK);
        // each list ends, and no particle is reached twice
        std::vector<char> seen((size_t)n[s], 0);
        int links = 0;
        for (int i = 0; i < n[s]; ++i) {
            const int j = a.nn[(size_t)i];
            CHECK(j >= -1 && j < n[s]);
            if (j >= 0) {
                CHECK(!seen[(size_t)j]);
                seen[(size_t)j] = 1;
                ++links;
            }
        }
        CHECK(links < n[s]);
        if (s < 2) CHECK(dsh_write_data(out_dir, 0, 0.0, p.tau, p.h, s, &a.v, 999) == DSH_OK);
    }
    CHECK(dsh_reference_ic(&p, 5, nullptr) != DSH_OK);
}

void tube_setup_and_oracle() {
    for (int precision : {DSH_FP64, DSH_MIXED}) {
        const dsh_tube_spec spec = small_tube(1, precision, 4);
        dsh_params ph{}, po{};
        CHECK(dsh_tube_params(&spec, &ph) == DSH_OK);
        CHECK(orc_tube_params(&spec, &po) == DSH_OK);
        int nh[3] = {0, 0, 0}, no[3] = {0, 0, 0};
        CHECK(dsh_tube_count(&spec, nh) == DSH_OK);
        CHECK(orc_tube_count(&spec, no) == DSH_OK);
        CHECK(nh[0] == no[0] && nh[1] == no[1]);
        orc_sim* o = nullptr;
        CHECK(orc_create(&po, &o) == DSH_OK);
        if (!o) return;
        for (int s = 0; s < 2; ++s) {
            Arrays a(nh[s]), b(no[s]);
            CHECK(dsh_tube_ic(&spec, s, &a.v) == DSH_OK);
            CHECK(orc_tube_ic(&spec, s, &b.v) == DSH_OK);
            CHECK(std::memcmp(a.v.x, b.v.x, sizeof(double) * (size_t)nh[s]) == 0);
            CHECK(orc_upload(o, s, &a.v, nh[s] - 1) == DSH_OK);   // pm: last index
        }
        double t = 0.0;
        CHECK(orc_step(o, 4, &t) == DSH_OK);
        double cons[4];
        CHECK(orc_conservation(o, cons) == DSH_OK);
        CHECK(std::isfinite(cons[3]));
        for (int s = 0; s < 2; ++s) {
            Arrays a(nh[s]);
            std::vector<int> cnt((size_t)nh[s]), ids((size_t)nh[s]);
            std::vector<uint64_t> hash((size_t)nh[s]);
            CHECK(orc_download(o, s, &a.v) == DSH_OK);
            CHECK(orc_neighbour_stats(o, s, cnt.data(), hash.data()) == DSH_OK);
            CHECK(orc_sorted_ids(o, s, ids.data()) == DSH_OK);
        }
        const int na = orc_num_avg_cells(o);
        CHECK(na > 0);
        std::vector<int> ng((size_t)na), nd((size_t)na);
        std::vector<double> vg(3 * (size_t)na), vd(3 * (size_t)na), psi(3 * (size_t)na), tb((size_t)na), ep((size_t)na);
        CHECK(orc_avg_cells(o, ng.data(), nd.data(), vg.data(), vd.data(), psi.data(), tb.data(), ep.data()) == DSH_OK);
        CHECK(orc_upload(o, 7, nullptr, 0) != DSH_OK);
        orc_destroy(o);
    }
    const double vg[3] = {1.0, 0.0, -1.0}, vd[6] = {0.5, 0.2, 0.1, -0.3, 0.0, 0.4}, psi[3] = {0.1, 0.2, 0.3};
    const double tb[2] = {0.01, 0.02}, ep[2] = {0.5, 0.25};
    double vgh[3], vdh[6];
    orc_idic_cell(2, vg, vd, psi, tb, ep, 1e-3, vgh, vdh);
    CHECK(std::isfinite(vgh[0]) && std::isfinite(vdh[5]));
    CHECK(orc_W(3.0, 1.0) == 0.0);
}

double g_now = 0.0;
double fake_clock(void*) { return g_now += 1.0; }

void serving_on_the_fake_device() {
    const dsh_tube_spec spec = small_tube(0, DSH_MIXED, 25);
    double fp = 0.0;
    CHECK(dsh_server_footprint(&spec, &fp) == DSH_OK && fp > 0.0);
    for (int fail_at : {0, 2}) {
        dsh_server_opts o{};
        o.hbm_budget = 2.5 * fp;
        o.steps_per_epoch = 10;
        o.backend = DSH_BACKEND_FAKE;
        o.clock = fake_clock;
        o.fake_fail_rebuild = fail_at;
        dsh_server* sv = nullptr;
        CHECK(dsh_server_create(&o, &sv) == DSH_OK);
        if (!sv) return;
        std::vector<uint64_t> ids;
        for (int i = 0; i < 6; ++i) {
            dsh_tube_spec s = small_tube(i, DSH_MIXED, 15 + 5 * i);
            uint64_t id = 0;
            CHECK(dsh_server_submit(sv, &s, 0.0, &id) == DSH_OK);
            ids.push_back(id);
        }
        dsh_tube_spec big = small_tube(9, DSH_MIXED, 10);
        big.dx_left = 0.001;
        uint64_t big_id = 0;
        CHECK(dsh_server_submit(sv, &big, 0.0, &big_id) == DSH_OK);
        CHECK(dsh_server_cancel(sv, ids[5]) == DSH_OK);
        int running = 0, queued = 0, pumps = 0;
        do {
            CHECK(dsh_server_pump(sv, &running, &queued) == DSH_OK || fail_at);
        } while ((running || queued) && ++pumps < 1000);
        CHECK(pumps < 1000);
        int done = 0;
        for (uint64_t id : ids) {
            int state = -1;
            double arr, adm, fin, ps;
            CHECK(dsh_server_poll(sv, id, &state, &arr, &adm, &fin, &ps) == DSH_OK);
            if (state == DSH_REQ_DONE) {
                ++done;
                int n[3] = {0, 0, 0};
                CHECK(dsh_tube_count(&spec, n) == DSH_OK);
            }
        }
        CHECK(done >= 4);
        int state = -1;
        double arr, adm, fin, ps;
        CHECK(dsh_server_poll(sv, big_id, &state, &arr, &adm, &fin, &ps) == DSH_OK);
        CHECK(state == DSH_REQ_REJECTED);
        CHECK(dsh_server_poll(sv, 987654321u, &state, &arr, &adm, &fin, &ps) != DSH_OK);
        double bytes = 0.0, rs = 0.0, as = 0.0;
        int rebuilds = 0, epochs = 0;
        CHECK(dsh_server_stats(sv, &bytes, &rebuilds, &epochs) == DSH_OK);
        CHECK(dsh_server_times(sv, &rs, &as) == DSH_OK);
        CHECK(dsh_server_last_error(sv) != nullptr);
        dsh_server_destroy(sv);
    }
}

void argument_checks() {
    dsh_params p{};
    CHECK(dsh_default_params(&p) == DSH_OK);
    CHECK(dsh_default_params(nullptr) != DSH_OK);
    CHECK(dsh_tube_params(nullptr, &p) != DSH_OK);
    size_t sz[4] = {0, 0, 0, 0};
    CHECK(dsh_struct_sizes(sz) == DSH_OK && sz[0] == sizeof(dsh_params));
    CHECK(dsh_upload_species(nullptr, 0, nullptr, 0) != DSH_OK);
    CHECK(dsh_step(nullptr, 1, nullptr) != DSH_OK);
    CHECK(dsh_batch_run(nullptr, 1, nullptr) != DSH_OK);
    dsh_tube_spec bad = small_tube(0, DSH_FP64, 1);
    bad.dx_left = -1.0;
    int n[3];
    CHECK(dsh_tube_count(&bad, n) != DSH_OK);
    CHECK(dsh_step_noalloc(nullptr, nullptr, nullptr, 1, nullptr, nullptr, nullptr, nullptr) != DSH_OK);
}

}  // namespace

int main(int argc, char** argv) {
    if (argc != 3) {
        std::fprintf(stderr, "usage: %s <Init.txt> <output dir>\n", argv[0]);
        return 2;
    }
    argument_checks();
    tube_setup_and_oracle();
    serving_on_the_fake_device();
    reference_problem(argv[1], argv[2]);
    std::printf(g_fail ? "host driver: FAILED\n" : "host driver: ok\n");
    return g_fail;
}
