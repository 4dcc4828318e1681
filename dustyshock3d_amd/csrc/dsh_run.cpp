// dsh_run - the Init.txt-driven run of the reference program on libdsh.
//
//   dsh_run [Init.txt] [output directory]        (defaults: ./Init.txt, ./Data)
//
// Same sequence as main() (kernel.cu:1696-2566): read Init.txt (kernel.cu:1705-1742), place
// the particles of init_Sod_X_Dust (kernel.cu:1391-1575), write output 0, then step until
// Tm >= T_end, printing "Time %5.3lf" every step and writing output number out_num whenever
// Tm >= out_num * T_out (kernel.cu:2528-2564).  The files are the ones Data_out and
// Data_out_dust write, in their format.  With a fixed time step their Nn column is the
// reference's linked list on its own 2h grid, built from the positions at the start of the
// step that precedes an output (with a CFL step, the lists of libdsh's grid); the t = 0
// files print -1 there (the reference prints unset memory).
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/dsh.h"

namespace {

struct Species {
    int n = 0;
    std::vector<double> x, y, z, vx, vy, vz, ax, ay, az, m, rho, p, e, h, ts, ed;
    std::vector<int> ind, nn, nn_ref;
    dsh_soa_view view(bool gas) {
        dsh_soa_view v;
        std::memset(&v, 0, sizeof v);
        v.x = x.data();
        v.y = y.data();
        v.z = z.data();
        v.vx = vx.data();
        v.vy = vy.data();
        v.vz = vz.data();
        v.m = m.data();
        v.rho = rho.data();
        v.h = h.data();
        v.ind = ind.data();
        v.nn = nn.data();
        if (gas) {
            v.ax = ax.data();
            v.ay = ay.data();
            v.az = az.data();
            v.p = p.data();
            v.e = e.data();
            v.edot = ed.data();
        } else {
            v.t_stop = ts.data();
        }
        return v;
    }
    void resize(int count) {
        n = count;
        for (auto* a : {&x, &y, &z, &vx, &vy, &vz, &ax, &ay, &az, &m, &rho, &p, &e, &h, &ts, &ed})
            a->assign(count, 0.0);
        ind.assign(count, 0);
        nn.assign(count, -1);
        nn_ref.assign(count, -1);
    }
};

int report(const char* what, int rc, dsh_ctx* ctx) {
    std::fprintf(stderr, "dsh_run: %s failed (%d)%s%s\n", what, rc, ctx ? ": " : "", ctx ? dsh_last_error(ctx) : "");
    return 1;
}

}  // namespace

int main(int argc, char** argv) {
    const char* init = argc > 1 ? argv[1] : "Init.txt";
    const char* out = argc > 2 ? argv[2] : "Data";
    dsh_params P;
    int rc = dsh_read_init(init, &P);
    if (rc != DSH_OK) return report("reading Init.txt", rc, nullptr);
    int counts[DSH_MAX_SPECIES];
    if ((rc = dsh_reference_setup(&P, counts)) != DSH_OK) return report("particle placement", rc, nullptr);
    const int nsp = 1 + P.n_dust;
    std::vector<Species> sp(nsp);
    for (int s = 0; s < nsp; ++s) {
        sp[s].resize(counts[s]);
        dsh_soa_view v = sp[s].view(s == 0);
        if ((rc = dsh_reference_ic(&P, s, &v)) != DSH_OK) return report("particle placement", rc, nullptr);
    }
    double Tm = 0.0;
    int out_num = 0;
    // nn_mode 0: -1 (t = 0), 1: the reference's lists, 2: the lists of libdsh's own grid
    auto write_all = [&](int nn_mode) -> int {
        for (int s = 0; s < nsp; ++s) {
            dsh_soa_view v = sp[s].view(s == 0);
            v.nn = nn_mode == 1 ? sp[s].nn_ref.data() : nn_mode == 2 ? sp[s].nn.data() : nullptr;
            // host dust accelerations are never refreshed by the reference: they print 0
            if (s > 0) v.ax = v.ay = v.az = nullptr;
            const int r = dsh_write_data(out, out_num, Tm, P.tau, P.h, s, &v, sp[s].n - 1);
            if (r != DSH_OK) return r;
        }
        return DSH_OK;
    };
    if ((rc = write_all(0)) != DSH_OK) return report("writing output 0", rc, nullptr);
    out_num = out_num + 1;

    dsh_ctx* ctx = nullptr;
    if ((rc = dsh_create(&P, &ctx)) != DSH_OK) return report("dsh_create", rc, ctx);
    for (int s = 0; s < nsp; ++s) {
        dsh_soa_view v = sp[s].view(s == 0);
        v.nn = nullptr;
        if ((rc = dsh_upload_species(ctx, s, &v, sp[s].n - 1)) != DSH_OK) return report("upload", rc, ctx);
    }
    const bool fixed = P.dt_mode == DSH_DT_FIXED;
    while (Tm < P.t_end) {
        // a fixed step tells in advance whether this step ends with an output: keep the
        // positions it starts from for the Nn column
        const bool predicted = fixed && Tm + P.tau >= out_num * P.t_out;
        if (predicted)
            for (int s = 0; s < nsp; ++s) {
                dsh_soa_view v;
                std::memset(&v, 0, sizeof v);
                v.x = sp[s].x.data();
                v.y = sp[s].y.data();
                v.z = sp[s].z.data();
                if ((rc = dsh_download_species(ctx, s, &v)) != DSH_OK) return report("download", rc, ctx);
                if ((rc = dsh_reference_nn(&P, v.x, v.y, v.z, sp[s].n - 1, sp[s].nn_ref.data())) != DSH_OK)
                    return report("Nn lists", rc, ctx);
            }
        const double before = Tm;
        if ((rc = dsh_step(ctx, 1, &Tm)) != DSH_OK) return report("dsh_step", rc, ctx);
        if (Tm == before) break;   // the run reached its step limit
        std::printf("Time %5.3lf \n", Tm);
        if (Tm >= out_num * P.t_out) {
            for (int s = 0; s < nsp; ++s) {
                dsh_soa_view v = sp[s].view(s == 0);
                if ((rc = dsh_download_species(ctx, s, &v)) != DSH_OK) return report("download", rc, ctx);
            }
            if ((rc = write_all(predicted ? 1 : 2)) != DSH_OK) return report("writing output", rc, ctx);
            out_num = out_num + 1;
        }
    }
    dsh_destroy(ctx);
    return 0;
}
