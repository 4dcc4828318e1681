// dsh_io.cpp - host-only output side of the C ABI:
//   - the Data/*.dat writers of the reference driver, byte for byte (Data_out and
//     Data_out_dust, kernel.cu:232-826: only the files those functions actually write);
//   - the next-particle pointers Nn of the reference's own neighbour grid (Clh = 2h,
//     kernel.cu:1767-1778, lists kernel.cu:1937-1995), for drivers that print the Nn column.
#include <cmath>
#include <cstdio>
#include <string>
#include <vector>

#include "../../include/dsh.h"

namespace {

double at(const double* a, int i) { return a ? a[i] : 0.0; }
int at(const int* a, int i) { return a ? a[i] : 0; }

std::string name(const char* dir, const char* stem, int num) {
    char buf[64];
    std::snprintf(buf, sizeof buf, "%s%07d.dat", stem, num);
    std::string d = dir ? dir : "Data";
    if (!d.empty() && d.back() != '/') d += '/';
    return d + buf;
}

void headers(FILE* f, double t, double tau, double h) {
    std::fprintf(f, "t=%5.3f \n", t);
    std::fprintf(f, "tau=%10.8lf \t h=%10.8lf \n", tau, h);
}

// One row: the doubles as "%10.8lf \t ", then the ints separated by " \t ", then " \n".
void row(FILE* f, const double* d, int nd, const int* k, int nk) {
    for (int i = 0; i < nd; ++i) std::fprintf(f, "%10.8lf \t ", d[i]);
    for (int i = 0; i < nk; ++i) std::fprintf(f, i + 1 < nk ? "%d \t " : "%d \n", k[i]);
}

}  // namespace

extern "C" {

int dsh_write_data(const char* dir, int num, double t, double tau, double h, int species, const dsh_soa_view* v,
                   int pm) {
    if (!v || !v->x || !v->y || !v->z || num < 0 || num > 9999999 || pm < -1 || species < 0 ||
        species > DSH_MAX_DUST)
        return DSH_ERR_ARG;
    const double eps_out = 1.0 * h;
    char stem_all[16], stem_line[16];
    if (species == 0) {
        std::snprintf(stem_all, sizeof stem_all, "G");
        std::snprintf(stem_line, sizeof stem_line, "L_X");
    } else {
        std::snprintf(stem_all, sizeof stem_all, "D-%d-", species);
        std::snprintf(stem_line, sizeof stem_line, "D-%d-L_X", species);
    }
    // every particle (kernel.cu:242-266, 444-469, 635-660)
    FILE* f = std::fopen(name(dir, stem_all, num).c_str(), "wt");
    if (!f) return DSH_ERR_IO;
    headers(f, t, tau, h);
    if (species == 0)
        std::fprintf(f, "x \t y \t z \t r \t mas \t rho \t p \t Vx \t Vy \t Vz \t V \t e \t Ind \n");
    else
        std::fprintf(f, "x \t y \t z \t r \t mas \t rho \t t_stop \t Vx \t Vy \t Vz \t V \t Ind \t Nn \n");
    for (int i = 0; i <= pm; i++) {
        const double x = v->x[i], y = v->y[i], z = v->z[i];
        const double vx = at(v->vx, i), vy = at(v->vy, i), vz = at(v->vz, i);
        const double r = std::sqrt(x * x + y * y + z * z);
        const double vv = std::sqrt(vx * vx + vy * vy + vz * vz);
        // the Vz column prints Vy (kernel.cu:262,464,655)
        const int k[2] = {at(v->ind, i), v->nn ? v->nn[i] : -1};
        if (species == 0) {
            const double d[12] = {x, y, z, r, at(v->m, i), at(v->rho, i), at(v->p, i), vx, vy, vy, vv, at(v->e, i)};
            row(f, d, 12, k, 2);
        } else {
            const double d[11] = {x, y, z, r, at(v->m, i), at(v->rho, i), at(v->t_stop, i), vx, vy, vy, vv};
            row(f, d, 11, k, 2);
        }
    }
    if (std::fclose(f) != 0) return DSH_ERR_IO;
    // the line y^2 + z^2 <= h^2 (kernel.cu:269-294, 471-496, 662-687)
    f = std::fopen(name(dir, stem_line, num).c_str(), "wt");
    if (!f) return DSH_ERR_IO;
    headers(f, t, tau, h);
    if (species == 0)
        std::fprintf(f, "x \t y \t z \t r \t mas \t rho \t p \t Vx \t Vy \t Vz \t V \t Ax \t Ay \t Az \t e \t Ind \n");
    else
        std::fprintf(f, "x \t y \t z \t r \t mas \t rho \t Vx \t Vy \t Vz \t V \t Ax \t Ay \t Az \t Ind \n");
    for (int i = 0; i <= pm; i++) {
        const double x = v->x[i], y = v->y[i], z = v->z[i];
        if (!(y * y + z * z <= eps_out * eps_out)) continue;
        const double vx = at(v->vx, i), vy = at(v->vy, i), vz = at(v->vz, i);
        const double r = std::sqrt(x * x + y * y + z * z);
        const double vv = std::sqrt(vx * vx + vy * vy + vz * vz);
        const int k[1] = {at(v->ind, i)};
        if (species == 0) {
            const double d[15] = {x,  y,  z,  r,  at(v->m, i),  at(v->rho, i), at(v->p, i), vx,
                                  vy, vy, vv, at(v->ax, i), at(v->ay, i), at(v->az, i), at(v->e, i)};
            row(f, d, 15, k, 1);
        } else {
            // 14 values under 13 value names: t_stop follows Az (kernel.cu:490-491)
            const double d[14] = {x,  y,  z,  r,  at(v->m, i),  at(v->rho, i), vx,
                                  vy, vy, vv, at(v->ax, i), at(v->ay, i), at(v->az, i), at(v->t_stop, i)};
            row(f, d, 14, k, 1);
        }
    }
    if (std::fclose(f) != 0) return DSH_ERR_IO;
    return DSH_OK;
}

int dsh_reference_nn(const dsh_params* p, const double* x, const double* y, const double* z, int pm, int* nn) {
    if (!p || !x || !y || !z || !nn || pm < -1 || !(p->h > 0.0)) return DSH_ERR_ARG;
    const double Clh = 2.0 * p->h;
    const int Cnx = (int)((p->x_max - p->x_min) / Clh);
    const int Cny = (int)((p->y_max - p->y_min) / Clh);
    const int Cnz = (int)((p->z_max - p->z_min) / Clh);
    const long long ncell = (long long)(Cnx + 1) * (Cny + 1) * (Cnz + 1);
    if (Cnx < 0 || Cny < 0 || Cnz < 0 || ncell > 1000000000ll) return DSH_ERR_ARG;
    std::vector<int> first((size_t)ncell + 1, -1), last((size_t)ncell + 1, -1);
    for (int q = 0; q <= pm; q++) nn[q] = -1;
    for (int q = 0; q <= pm; q++) {
        // key with the Cnx stride (kernel.cu:1949); out-of-range keys are not listed
        const double fx = (x[q] - p->x_min) / Clh, fy = (y[q] - p->y_min) / Clh, fz = (z[q] - p->z_min) / Clh;
        if (!(std::fabs(fx) < 2e9 && std::fabs(fy) < 2e9 && std::fabs(fz) < 2e9)) continue;
        const long long i = (long long)(int)fx + (long long)(int)fy * Cnx + (long long)(int)fz * Cnx * Cny;
        if (i < 0 || i > ncell) continue;
        if (first[(size_t)i] == -1) {
            first[(size_t)i] = q;
        } else {
            nn[last[(size_t)i]] = q;
        }
        last[(size_t)i] = q;
    }
    return DSH_OK;
}

}  // extern "C"
