/*
 * oracle.cpp - FP64 OpenMP restatement of DustyShock3D's per-step dusty-gas SPH path.
 *
 * TEST INFRASTRUCTURE, not product code: only tests/, __graft_entry__.smoke() and
 * bench.py's cpu_baseline leg may load it, as the checker or as the CPU baseline.
 *
 * It follows the reference's arithmetic in the reference's operation order:
 *   W, dW                 kernel.cu:11-38
 *   Eq_State              kernel.cu:92-97
 *   cell lists            kernel.cu:1937-1995 (counting sort, stable in index order)
 *   density               IDIC.h:17-45, launch order kernel.cu:2011-2043
 *   force + energy rate   IDIC.h:198-260, launch order kernel.cu:2233-2266
 *   IDIC cell averages    kernel.cu:2060-2211
 *   IDIC implicit solve   kernel.cu:2294-2437
 *   velocity updates      IDIC.h:264-315
 *   drift, energy         IDIC.h:319-346
 * and adds the extensions the benchmark configurations need (periodic y/z with
 * periodic images, per-particle h, K-based stopping time, CFL time step, seeded
 * lattice set-up).  Deliberate deviations from the reference text:
 *   - distance uses sqrt(d2) instead of pow(d2, 0.5) (IDIC.h:37,219);
 *   - the neighbour predicate is d2 < (h_i + h_j)^2 with d2 = fma(dz,dz,fma(dy,dy,dx*dx)),
 *     which equals the reference's dist < 2h for a global h up to that rounding;
 *   - the neighbour grid is the one the GPU build uses (key = kx + nx(ky + ny kz),
 *     truncation toward zero, clamped), not the aliasing Cnx-stride key of IDIC.h:28;
 *   - averaging cells are avg_factor^3 blocks of neighbour cells.
 */
#include "oracle.h"

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>
#ifdef _OPENMP
#include <omp.h>
#endif

namespace {

constexpr double kPi = 3.14159265;  // kernel.cu:13, data.h:6

// kernel.cu:11-22
double W(double r, double h) {
    double k = std::fabs(r) / h;
    double c = 1.0 / kPi;
    double t = 0.0;
    if (k < 1.0) t = 1.0 - 1.5 * k * k + 0.75 * k * k * k;
    if ((k >= 1.0) && (k <= 2.0)) t = 0.25 * (2.0 - k) * (2.0 - k) * (2.0 - k);
    if (k > 2.0) t = 0.0;
    return c / (h * h * h) * t;
}

// kernel.cu:24-38 (signed r)
double dW(double r, double h) {
    double k = r / h;
    double c = 1.0 / kPi;
    double t = 0.0;
    if (k < -2.0) t = 0.0;
    if ((k >= -2.0) && (k <= -1.0)) t = 0.75 * (2.0 + k) * (2.0 + k);
    if ((k > -1.0) && (k < 0)) t = -3.0 * k - 2.25 * k * k;
    if ((k >= 0) && (k <= 1.0)) t = -3.0 * k + 2.25 * k * k;
    if ((k >= 1.0) && (k <= 2.0)) t = -0.75 * (2.0 - k) * (2.0 - k);
    if (k > 2.0) t = 0.0;
    return c / (h * h * h * h) * t;
}

// kernel.cu:92-97
double eq_state(double rho, double e, int type, double gam, double cs) {
    if (type == 0) return rho * e * (gam - 1);
    return cs * cs * rho;
}

uint64_t mix64(uint64_t x) {
    x += 0x9E3779B97F4A7C15ull;
    x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
    x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
    return x ^ (x >> 31);
}

// (id, image) packed injectively (image offsets within +-127), then the bijective mix64
uint64_t pair_tag(int id, int sy, int sz) {
    return mix64(((uint64_t)(uint32_t)id << 16) | ((uint64_t)(uint32_t)(sy + 128) << 8) | (uint64_t)(uint32_t)(sz + 128));
}

struct Grid {
    double o[3], c[3], ic[3], L[3], hi[3];
    int n[3], per[3], f;
    double cmin;
};

int floordiv(int a, int n) { return (a >= 0) ? a / n : -((-a + n - 1) / n); }

// int((x - o) / c): FP64 divide and truncation as kernel.cu:1949 and IDIC.h:28, clamped.
int axis_cell(double x, double o, double c, int n) {
    double t = (x - o) / c;
    if (!(t >= 0.0)) return 0;
    if (t >= (double)n) return n - 1;
    return (int)t;
}

double wrap(double y, double lo, double L, double hi) {
    if (y < lo) y = y + L;
    if (y >= hi) y = y - L;
    return y;
}

struct Sp {
    int n = 0;
    std::vector<double> x, y, z, vx, vy, vz, m, h, rho, p, e, ts, ax, ay, az, ed, csav;
    std::vector<int> ind;
    std::vector<int> key, order, start;
    std::vector<int> ncount;
    std::vector<uint64_t> nhash;
    double hmax = 0.0;
};

}  // namespace

struct orc_sim {
    dsh_params P{};
    int nsp = 2;
    Sp sp[DSH_MAX_SPECIES];
    Grid g{};
    bool prepared = false;
    double t = 0.0, dt = 0.0;
    int steps = 0, done = 0, final_step = 0;
    double cons[4] = {0, 0, 0, 0};
    // averaging cells of the last step
    int nax = 0, nay = 0, naz = 0, navg = 0;
    std::vector<int> ng, nd;
    std::vector<double> vgh, vdh, psi, tbar, eps;
};

namespace {

Grid make_grid(const dsh_params& p, double hmin) {
    Grid g{};
    const double lo[3] = {p.x_min, p.y_min, p.z_min};
    const double up[3] = {p.x_max, p.y_max, p.z_max};
    g.per[0] = 0;
    g.per[1] = p.periodic_y ? 1 : 0;
    g.per[2] = p.periodic_z ? 1 : 0;
    g.f = p.avg_factor < 1 ? 1 : p.avg_factor;
    double ct = p.cell_size > 0.0 ? p.cell_size : 2.0 * hmin * (1.0 + p.cell_slack);
    double block = ct * (double)g.f;
    g.cmin = 1e300;
    for (int a = 0; a < 3; ++a) {
        double L = up[a] - lo[a];
        int n;
        double c;
        if (g.per[a]) {
            n = (int)std::floor(L / block) * g.f;
            if (n < g.f) n = g.f;
            c = L / (double)n;
        } else {
            n = (int)std::ceil(L / block) * g.f;
            if (n < g.f) n = g.f;
            c = ct;
        }
        g.n[a] = n;
        g.c[a] = c;
        g.o[a] = lo[a];
        g.ic[a] = 1.0 / c;
        g.L[a] = L;
        g.hi[a] = lo[a] + L;
        g.cmin = std::min(g.cmin, c);
    }
    return g;
}

void cell_of(const Grid& g, double x, double y, double z, int k[3]) {
    k[0] = axis_cell(x, g.o[0], g.c[0], g.n[0]);
    k[1] = axis_cell(y, g.o[1], g.c[1], g.n[1]);
    k[2] = axis_cell(z, g.o[2], g.c[2], g.n[2]);
}

// kernel.cu:1937-1995: particles appended to their cell's list in increasing index,
// i.e. a counting sort by cell key that is stable in particle index.
void build_cells(Sp& s, const Grid& g) {
    int nc = g.n[0] * g.n[1] * g.n[2];
    s.key.assign(s.n, 0);
#pragma omp parallel for schedule(static)
    for (int i = 0; i < s.n; ++i) {
        int k[3];
        cell_of(g, s.x[i], s.y[i], s.z[i], k);
        s.key[i] = k[0] + g.n[0] * (k[1] + g.n[1] * k[2]);
    }
    s.start.assign(nc + 1, 0);
    for (int i = 0; i < s.n; ++i) s.start[s.key[i] + 1]++;
    for (int c = 0; c < nc; ++c) s.start[c + 1] += s.start[c];
    std::vector<int> tail(s.start.begin(), s.start.end() - 1);
    s.order.assign(s.n, 0);
    for (int i = 0; i < s.n; ++i) s.order[tail[s.key[i]]++] = i;
    s.hmax = 0.0;
    for (int i = 0; i < s.n; ++i) s.hmax = std::max(s.hmax, s.h[i]);
}

// Visit every (j, periodic image) with d2 < (h_i + h_j)^2.  Cell offsets along each axis
// cover h_i + max_j h_j, so the set is exactly the predicate's.  Offsets are visited x
// outermost, then y, then z, like the 27 launches of kernel.cu:2014-2016; inside a cell
// in increasing particle index (the linked-list order).
template <class F>
void for_neighbours(const Sp& s, const Grid& g, int i, F&& f) {
    const double H = (s.h[i] + s.hmax) * (1.0 + 1e-9);
    int R[3];
    for (int a = 0; a < 3; ++a) R[a] = (int)std::ceil(H * g.ic[a]);
    int k[3];
    cell_of(g, s.x[i], s.y[i], s.z[i], k);
    const double xi = s.x[i], yi = s.y[i], zi = s.z[i], hi = s.h[i];
    for (int ox = -R[0]; ox <= R[0]; ++ox) {
        int cx = k[0] + ox;
        if (cx < 0 || cx >= g.n[0]) continue;
        for (int oy = -R[1]; oy <= R[1]; ++oy) {
            int cy = k[1] + oy, sy = 0;
            if (g.per[1]) {
                sy = floordiv(cy, g.n[1]);
                cy -= sy * g.n[1];
            } else if (cy < 0 || cy >= g.n[1]) {
                continue;
            }
            const double shy = (double)sy * g.L[1];
            for (int oz = -R[2]; oz <= R[2]; ++oz) {
                int cz = k[2] + oz, sz = 0;
                if (g.per[2]) {
                    sz = floordiv(cz, g.n[2]);
                    cz -= sz * g.n[2];
                } else if (cz < 0 || cz >= g.n[2]) {
                    continue;
                }
                const double shz = (double)sz * g.L[2];
                const int c = cx + g.n[0] * (cy + g.n[1] * cz);
                f.begin_cell();
                for (int q = s.start[c]; q < s.start[c + 1]; ++q) {
                    const int j = s.order[q];
                    const double dx = s.x[j] - xi;
                    const double dy = (s.y[j] - yi) + shy;
                    const double dz = (s.z[j] - zi) + shz;
                    const double d2 = std::fma(dz, dz, std::fma(dy, dy, dx * dx));
                    const double H2 = hi + s.h[j];
                    if (d2 < H2 * H2) f.pair(j, dx, dy, dz, d2, sy, sz);
                }
                f.end_cell();
            }
        }
    }
}

// IDIC.h:17-45 (one partial sum per cell, added to rho like the 27 launches)
void density(orc_sim& S, int sidx) {
    Sp& s = S.sp[sidx];
    s.ncount.assign(s.n, 0);
    s.nhash.assign(s.n, 0);
#pragma omp parallel for schedule(dynamic, 64)
    for (int i = 0; i < s.n; ++i) {
        if (s.ind[i] != 0) continue;
        struct Acc {
            const Sp& s;
            int i;
            double rho = 0.0, tmp = 0.0;
            int cnt = 0;
            uint64_t hash = 0;
            void begin_cell() { tmp = 0.0; }
            void pair(int j, double, double, double, double d2, int sy, int sz) {
                double d = std::sqrt(d2);
                tmp = tmp + s.m[j] * W(d, 0.5 * (s.h[i] + s.h[j]));
                cnt++;
                hash += pair_tag(j, sy, sz);
            }
            void end_cell() { rho = rho + tmp; }
        } acc{s, i};
        for_neighbours(s, S.g, i, acc);
        s.rho[i] = acc.rho;
        s.ncount[i] = acc.cnt;
        s.nhash[i] = acc.hash;
    }
}

// IDIC.h:198-260; gas-only runs take de/dt of the plain ForceKernel (kernel.cu:182-187),
// which weights each velocity component with dW of that coordinate difference.
void force(orc_sim& S) {
    Sp& s = S.sp[0];
    const double alpha = S.P.alpha, beta = S.P.beta, eps = S.P.eps;
    const bool plain = S.nsp == 1;
#pragma omp parallel for schedule(dynamic, 64)
    for (int i = 0; i < s.n; ++i) {
        if (s.ind[i] != 0) {
            s.ax[i] = s.ay[i] = s.az[i] = s.ed[i] = 0.0;
            continue;
        }
        struct Acc {
            const Sp& s;
            int i;
            double alpha, beta, eps;
            bool plain;
            double ax = 0.0, ay = 0.0, az = 0.0, ed = 0.0;
            void begin_cell() {}
            void end_cell() {}
            void pair(int j, double dx, double dy, double dz, double d2, int, int) {
                if (!(d2 > 0.0)) return;
                const double dist = std::sqrt(d2);
                const double hh = 0.5 * (s.h[i] + s.h[j]);
                double F_nu = 0.0;
                const double nu1x = (s.vx[i] - s.vx[j]) * (-dx);
                const double nu1y = (s.vy[i] - s.vy[j]) * (-dy);
                const double nu1z = (s.vz[i] - s.vz[j]) * (-dz);
                double nu = nu1x + nu1y + nu1z;
                if (nu < 0.0) {
                    const double Cnu = 0.5 * (s.csav[i] + s.csav[j]);
                    nu = hh * nu / (dist * dist + eps * hh * hh);
                    F_nu = (-alpha * Cnu * nu + beta * nu * nu) / (0.5 * (s.rho[i] + s.rho[j]));
                }
                const double dw = dW(dist, hh);
                if (plain)
                    ed = ed + (s.m[i] * s.p[i] / (s.rho[i] * s.rho[i]) + s.m[i] * F_nu / 2.0) *
                                  ((s.vx[i] - s.vx[j]) * dW(-dx, hh) + (s.vy[i] - s.vy[j]) * dW(-dy, hh) +
                                   (s.vz[i] - s.vz[j]) * dW(-dz, hh));
                else
                    ed = ed + (s.m[i] * s.p[i] / (s.rho[i] * s.rho[i]) + s.m[i] * F_nu / 2.0) *
                                  (nu1x + nu1y + nu1z) * dw / dist;
                const double A = s.m[j] *
                                 (s.p[i] / (s.rho[i] * s.rho[i]) + s.p[j] / (s.rho[j] * s.rho[j]) + F_nu) *
                                 dw;
                ax = ax + dx * A / dist;
                ay = ay + dy * A / dist;
                az = az + dz * A / dist;
            }
        } acc{s, i, alpha, beta, eps, plain};
        for_neighbours(s, S.g, i, acc);
        s.ax[i] = acc.ax;
        s.ay[i] = acc.ay;
        s.az[i] = acc.az;
        s.ed[i] = acc.ed;
    }
}

// kernel.cu:2294-2437, one averaging cell, one velocity component.
void idic_solve(int nd, double vg, const double* vd, double psi, const double* tb, const double* ep,
                double tau, double* vgh, double* vdh) {
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

double signal_dt(const orc_sim& S) {
    double best = 1e300;
    for (int s = 0; s < S.nsp; ++s) {
        const Sp& q = S.sp[s];
        for (int i = 0; i < q.n; ++i) {
            if (q.ind[i] != 0) continue;
            const double v = std::sqrt(q.vx[i] * q.vx[i] + q.vy[i] * q.vy[i] + q.vz[i] * q.vz[i]);
            const double r = q.h[i] / (S.P.cs + v);
            if (r < best) best = r;
        }
    }
    return S.P.cfl * best;
}

void conservation(orc_sim& S) {
    double px = 0, py = 0, pz = 0, E = 0;
    for (int s = 0; s < S.nsp; ++s) {
        const Sp& q = S.sp[s];
        for (int i = 0; i < q.n; ++i) {
            px += q.m[i] * q.vx[i];
            py += q.m[i] * q.vy[i];
            pz += q.m[i] * q.vz[i];
            double k = 0.5 * q.m[i] * (q.vx[i] * q.vx[i] + q.vy[i] * q.vy[i] + q.vz[i] * q.vz[i]);
            if (s == 0) k += q.m[i] * q.e[i];
            E += k;
        }
    }
    S.cons[0] = px;
    S.cons[1] = py;
    S.cons[2] = pz;
    S.cons[3] = E;
}

void prepare(orc_sim& S) {
    double hmin = 1e300;
    for (int s = 0; s < S.nsp; ++s)
        for (int i = 0; i < S.sp[s].n; ++i) hmin = std::min(hmin, S.sp[s].h[i]);
    S.g = make_grid(S.P, hmin);
    S.nax = S.g.n[0] / S.g.f;
    S.nay = S.g.n[1] / S.g.f;
    S.naz = S.g.n[2] / S.g.f;
    S.navg = S.nax * S.nay * S.naz;
    if (S.P.dt_mode == DSH_DT_FIXED)
        S.dt = S.P.tau;
    else
        S.dt = signal_dt(S);
    conservation(S);
    S.prepared = true;
}

int avg_cell_of_key(const orc_sim& S, int key) {
    const Grid& g = S.g;
    int kx = key % g.n[0];
    int r = key / g.n[0];
    int ky = r % g.n[1];
    int kz = r / g.n[1];
    return kx / g.f + S.nax * (ky / g.f + S.nay * (kz / g.f));
}

void one_step(orc_sim& S) {
    const dsh_params& P = S.P;
    const double tau = S.dt;
    const int nd = S.nsp - 1;
    for (int s = 0; s < S.nsp; ++s) build_cells(S.sp[s], S.g);
    // density, pressure, stopping time (kernel.cu:2011-2056)
    for (int s = 0; s < S.nsp; ++s) density(S, s);
    {
        Sp& g = S.sp[0];
        for (int i = 0; i < g.n; ++i)
            if (g.ind[i] == 0) g.p[i] = eq_state(g.rho[i], g.e[i], P.eos_type, P.gamma, P.cs);
        for (int i = 0; i < g.n; ++i) g.csav[i] = std::sqrt(P.gamma * g.p[i] / g.rho[i]);
        for (int k = 0; k < nd; ++k) {
            Sp& d = S.sp[1 + k];
            if (P.drag_mode == 0) {
                // t_stop_Kernel_IDIC (IDIC.h:78-91, launches kernel.cu:2049-2053): dust i uses
                // the gas particle with the same index
                for (int i = 0; i < d.n; ++i)
                    if (d.ind[i] == 0)
                        d.ts[i] = P.R_dust[k] / (std::sqrt(P.gamma * g.p[i] / g.rho[i]) * g.rho[i]);
            } else {
                for (int i = 0; i < d.n; ++i)
                    if (d.ind[i] == 0) d.ts[i] = d.rho[i] / P.K[k];
            }
        }
    }
    // averaging-cell sums (kernel.cu:2114-2190), all particles, index order
    const int A = S.navg;
    std::vector<int> ng(A, 0), ndc((size_t)A * nd, 0);
    std::vector<double> vg(3 * (size_t)A, 0.0), vd(3 * (size_t)A * nd, 0.0), tb((size_t)A * nd, 0.0);
    std::vector<double> ps(3 * (size_t)A, 0.0);
    std::vector<int> cg(S.sp[0].n);
    {
        Sp& g = S.sp[0];
        for (int i = 0; i < g.n; ++i) {
            int c = avg_cell_of_key(S, g.key[i]);
            cg[i] = c;
            ng[c] += 1;
            vg[3 * c + 0] += g.vx[i];
            vg[3 * c + 1] += g.vy[i];
            vg[3 * c + 2] += g.vz[i];
        }
        for (int c = 0; c < A; ++c)
            if (ng[c] > 0)
                for (int a = 0; a < 3; ++a) vg[3 * c + a] = vg[3 * c + a] / ng[c];
    }
    std::vector<std::vector<int>> cd(nd);
    for (int k = 0; k < nd; ++k) {
        Sp& d = S.sp[1 + k];
        cd[k].resize(d.n);
        for (int i = 0; i < d.n; ++i) {
            int c = avg_cell_of_key(S, d.key[i]);
            cd[k][i] = c;
            size_t ck = (size_t)c * nd + k;
            ndc[ck] += 1;
            tb[ck] += d.ts[i];
            vd[3 * ck + 0] += d.vx[i];
            vd[3 * ck + 1] += d.vy[i];
            vd[3 * ck + 2] += d.vz[i];
        }
        for (int c = 0; c < A; ++c) {
            size_t ck = (size_t)c * nd + k;
            if (ndc[ck] > 0) {
                for (int a = 0; a < 3; ++a) vd[3 * ck + a] = vd[3 * ck + a] / ndc[ck];
                tb[ck] = tb[ck] / ndc[ck];
            } else {
                tb[ck] = P.tstop_fallback;
            }
        }
    }
    std::vector<double> ep((size_t)A * nd);
    for (int c = 0; c < A; ++c)
        for (int k = 0; k < nd; ++k) {
            int cnt_test = (k == 1 && P.eps2_uses_count1) ? ndc[(size_t)c * nd + 0] : ndc[(size_t)c * nd + k];
            if (ng[c] > 0 && cnt_test > 0)
                ep[(size_t)c * nd + k] = ndc[(size_t)c * nd + k] * P.mass_ratio[k] / ng[c];
            else
                ep[(size_t)c * nd + k] = P.eps_fallback[k];
        }
    // force (kernel.cu:2233-2266)
    force(S);
    std::vector<double> vgh, vdh;
    if (nd == 0) {
        // gas only: the plain explicit kick of MoveKernel (kernel.cu:205-228)
        Sp& g = S.sp[0];
        for (int i = 0; i < g.n; ++i) {
            if (g.ind[i] != 0) continue;
            g.vx[i] = g.vx[i] + g.ax[i] * tau;
            g.vy[i] = g.vy[i] + g.ay[i] * tau;
            g.vz[i] = g.vz[i] + g.az[i] * tau;
        }
    } else {
        // Psi = cell mean of gas acceleration (kernel.cu:2294-2309)
        {
            Sp& g = S.sp[0];
            for (int i = 0; i < g.n; ++i) {
                int c = cg[i];
                ps[3 * c + 0] += g.ax[i];
                ps[3 * c + 1] += g.ay[i];
                ps[3 * c + 2] += g.az[i];
            }
            for (int c = 0; c < A; ++c)
                if (ng[c] > 0)
                    for (int a = 0; a < 3; ++a) ps[3 * c + a] = ps[3 * c + a] / ng[c];
        }
        // implicit solve (kernel.cu:2310-2437)
        vgh.assign(3 * (size_t)A, 0.0);
        vdh.assign(3 * (size_t)A * nd, 0.0);
        for (int c = 0; c < A; ++c) {
            for (int a = 0; a < 3; ++a) {
                double vdc[DSH_MAX_DUST], tbc[DSH_MAX_DUST], epc[DSH_MAX_DUST], out[DSH_MAX_DUST];
                for (int k = 0; k < nd; ++k) {
                    size_t ck = (size_t)c * nd + k;
                    vdc[k] = vd[3 * ck + a];
                    tbc[k] = tb[ck];
                    epc[k] = ep[ck];
                }
                double g;
                idic_solve(nd, vg[3 * c + a], vdc, ps[3 * c + a], tbc, epc, tau, &g, out);
                vgh[3 * c + a] = g;
                for (int k = 0; k < nd; ++k) vdh[3 * ((size_t)c * nd + k) + a] = out[k];
            }
        }
        // dust velocities (IDIC.h:264-284)
        for (int k = 0; k < nd; ++k) {
            Sp& d = S.sp[1 + k];
            for (int i = 0; i < d.n; ++i) {
                if (d.ind[i] != 0) continue;
                int c = cd[k][i];
                double t = tb[(size_t)c * nd + k];
                double* v[3] = {&d.vx[i], &d.vy[i], &d.vz[i]};
                for (int a = 0; a < 3; ++a) *v[a] = (*v[a] / tau + vgh[3 * c + a] / t) / (1.0 / tau + 1.0 / t);
            }
        }
        // gas velocities (IDIC.h:287-315)
        {
            Sp& g = S.sp[0];
            for (int i = 0; i < g.n; ++i) {
                if (g.ind[i] != 0) continue;
                int c = cg[i];
                double tmp1 = 0.0;
                double tmp[3] = {0.0, 0.0, 0.0};
                if (nd == 1) {
                    tmp1 = tmp1 + ep[c] / tb[c];
                    for (int a = 0; a < 3; ++a) tmp[a] = tmp[a] + ep[c] * vdh[3 * (size_t)c + a] / tb[c];
                } else {
                    size_t c0 = (size_t)c * 2, c1 = c0 + 1;
                    tmp1 = tmp1 + ep[c0] / tb[c0] + ep[c1] / tb[c1];
                    for (int a = 0; a < 3; ++a)
                        tmp[a] = tmp[a] + ep[c0] * vdh[3 * c0 + a] / tb[c0] + ep[c1] * vdh[3 * c1 + a] / tb[c1];
                }
                double* v[3] = {&g.vx[i], &g.vy[i], &g.vz[i]};
                for (int a = 0; a < 3; ++a) *v[a] = (*v[a] / tau + tmp[a] + ps[3 * c + a]) / (1.0 / tau + tmp1);
            }
        }
    }
    // drift (IDIC.h:319-333), periodic wrap, energy (IDIC.h:335-346), smoothing length
    for (int s = 0; s < S.nsp; ++s) {
        Sp& q = S.sp[s];
        for (int i = 0; i < q.n; ++i) {
            if (q.ind[i] != 0) continue;
            q.x[i] = q.x[i] + q.vx[i] * tau;
            q.y[i] = q.y[i] + q.vy[i] * tau;
            q.z[i] = q.z[i] + q.vz[i] * tau;
            if (S.g.per[1]) q.y[i] = wrap(q.y[i], S.g.o[1], S.g.L[1], S.g.hi[1]);
            if (S.g.per[2]) q.z[i] = wrap(q.z[i], S.g.o[2], S.g.L[2], S.g.hi[2]);
            if (s == 0) q.e[i] = q.e[i] + q.ed[i] * tau;
            if (P.h_mode == 1) q.h[i] = P.eta * std::cbrt(q.m[i] / q.rho[i]);
        }
    }
    // keep the last step's cell state for inspection
    S.ng = ng;
    S.nd = ndc;
    S.vgh = vgh;
    S.vdh = vdh;
    S.psi = ps;
    S.tbar = tb;
    S.eps = ep;
    // time, conservation, next step
    S.t = S.t + tau;
    S.steps += 1;
    conservation(S);
    if (P.max_steps > 0 && S.steps >= P.max_steps) S.done = 1;
    if (P.dt_mode == DSH_DT_CFL) {
        if (S.final_step) {
            S.done = 1;
        } else {
            double rem = P.t_end - S.t;
            if (!(rem > 0.0)) {
                S.done = 1;
            } else {
                double dtc = signal_dt(S);
                if (dtc >= rem) {
                    S.dt = rem;
                    S.final_step = 1;
                } else {
                    S.dt = dtc;
                }
            }
        }
    } else {
        if (S.t >= P.t_end) S.done = 1;
    }
}

// ---- shock-tube set-up, restated from the specification of dsh_tube_spec ----
uint64_t rng_u64(uint64_t seed, int tube, int species, int id, int axis) {
    uint64_t z = mix64(seed ^ ((uint64_t)(uint32_t)tube * 0xD1B54A32D192ED03ull));
    z = mix64(z ^ ((uint64_t)(uint32_t)(species * 4 + axis) * 0x8CB92BA72F3D8DD7ull));
    z = mix64(z ^ ((uint64_t)(uint32_t)id * 0xA0761D6478BD642Full));
    return z;
}

double rng_sym(uint64_t seed, int tube, int species, int id, int axis) {
    double u = (double)(rng_u64(seed, tube, species, id, axis) >> 11) * (1.0 / 9007199254740992.0);
    return 2.0 * u - 1.0;
}

struct TubeGeom {
    int nlx, nrx, nyl, nyr, nzl, nzr;
    double dxl, dxr, dyl, dyr, dzl, dzr;
    double m, rhol, rhor, hl, hr;
};

TubeGeom tube_geom(const dsh_tube_spec& t) {
    TubeGeom G{};
    G.dxl = t.dx_left;
    G.dxr = 2.0 * t.dx_left;
    G.nlx = (int)std::floor((0.0 - t.x_min) / G.dxl + 0.5);
    G.nrx = (int)std::floor((t.x_max - 0.0) / G.dxr + 0.5);
    G.nyr = std::max(1, (int)std::floor(t.y_width / G.dxr + 0.5));
    G.nzr = std::max(1, (int)std::floor(t.z_width / G.dxr + 0.5));
    G.nyl = 2 * G.nyr;
    G.nzl = 2 * G.nzr;
    G.dyl = t.y_width / G.nyl;
    G.dyr = t.y_width / G.nyr;
    G.dzl = t.z_width / G.nzl;
    G.dzr = t.z_width / G.nzr;
    G.m = t.rho_left * (G.dxl * G.dyl * G.dzl);
    G.rhol = t.rho_left;
    G.rhor = G.m / (G.dxr * G.dyr * G.dzr);
    return G;
}

}  // namespace

extern "C" {

double orc_W(double r, double h) { return W(r, h); }
double orc_dW(double r, double h) { return dW(r, h); }

void orc_idic_cell(int n_dust, const double* vg, const double* vd, const double* psi, const double* tbar,
                   const double* eps, double tau, double* vg_hat, double* vd_hat) {
    for (int a = 0; a < 3; ++a) {
        double vdc[DSH_MAX_DUST], out[DSH_MAX_DUST];
        for (int k = 0; k < n_dust; ++k) vdc[k] = vd[3 * k + a];
        idic_solve(n_dust, vg[a], vdc, psi[a], tbar, eps, tau, &vg_hat[a], out);
        for (int k = 0; k < n_dust; ++k) vd_hat[3 * k + a] = out[k];
    }
}

void orc_set_threads(int n) {
#ifdef _OPENMP
    if (n > 0) omp_set_num_threads(n);
#else
    (void)n;
#endif
}

int orc_get_threads(void) {
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

int orc_create(const dsh_params* p, orc_sim** out) {
    if (!p || !out) return DSH_ERR_ARG;
    if (p->n_dust < 0 || p->n_dust > DSH_MAX_DUST) return DSH_ERR_ARG;
    if (p->drag_mode != 0 && p->drag_mode != 1) return DSH_ERR_ARG;
    orc_sim* s = new orc_sim();
    s->P = *p;
    s->nsp = 1 + p->n_dust;
    *out = s;
    return DSH_OK;
}

void orc_destroy(orc_sim* s) { delete s; }

int orc_upload(orc_sim* S, int species, const dsh_soa_view* v, int pm) {
    if (!S || !v || species < 0 || species >= S->nsp || pm < -1) return DSH_ERR_ARG;
    if (!v->x || !v->y || !v->z || !v->m || !v->ind) return DSH_ERR_ARG;
    Sp& q = S->sp[species];
    const int n = pm + 1;
    q.n = n;
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
        else if (S->P.h_mode == 1 && v->rho)
            q.h[i] = S->P.eta * std::cbrt(q.m[i] / q.rho[i]);
        else
            q.h[i] = S->P.h;
    }
    q.ax.assign(n, 0.0);
    q.ay.assign(n, 0.0);
    q.az.assign(n, 0.0);
    q.ed.assign(n, 0.0);
    q.csav.assign(n, 0.0);
    q.ncount.assign(n, 0);
    q.nhash.assign(n, 0);
    S->prepared = false;
    return DSH_OK;
}

int orc_step(orc_sim* S, int nsteps, double* t_out) {
    if (!S) return DSH_ERR_ARG;
    if (S->P.drag_mode == 0)
        for (int k = 1; k < S->nsp; ++k)
            if (S->sp[k].n != S->sp[0].n) return DSH_ERR_ARG;   // dust i pairs with gas i
    if (!S->prepared) prepare(*S);
    for (int k = 0; k < nsteps && !S->done; ++k) one_step(*S);
    if (t_out) *t_out = S->t;
    return DSH_OK;
}

int orc_download(orc_sim* S, int species, const dsh_soa_view* v) {
    if (!S || !v || species < 0 || species >= S->nsp) return DSH_ERR_ARG;
    const Sp& q = S->sp[species];
    auto put = [&q](double* dst, const std::vector<double>& src) {
        if (dst && !src.empty()) std::memcpy(dst, src.data(), sizeof(double) * q.n);
    };
    put(v->x, q.x);
    put(v->y, q.y);
    put(v->z, q.z);
    put(v->vx, q.vx);
    put(v->vy, q.vy);
    put(v->vz, q.vz);
    put(v->ax, q.ax);
    put(v->ay, q.ay);
    put(v->az, q.az);
    put(v->m, q.m);
    put(v->rho, q.rho);
    put(v->p, q.p);
    put(v->e, q.e);
    put(v->h, q.h);
    put(v->t_stop, q.ts);
    put(v->edot, q.ed);
    if (v->ind) std::memcpy(v->ind, q.ind.data(), sizeof(int) * q.n);
    if (v->cell && !q.key.empty()) std::memcpy(v->cell, q.key.data(), sizeof(int) * q.n);
    if (v->nn && !q.order.empty()) {
        for (int i = 0; i < q.n; ++i) v->nn[i] = -1;
        for (int k = 0; k + 1 < q.n; ++k) {
            int a = q.order[k], b = q.order[k + 1];
            if (q.key[a] == q.key[b]) v->nn[a] = b;
        }
    }
    return DSH_OK;
}

int orc_conservation(orc_sim* S, double out[4]) {
    if (!S || !out) return DSH_ERR_ARG;
    if (!S->prepared) prepare(*S);
    for (int a = 0; a < 4; ++a) out[a] = S->cons[a];
    return DSH_OK;
}

int orc_status(orc_sim* S, double* t, double* dt, int* steps, int* done) {
    if (!S) return DSH_ERR_ARG;
    if (!S->prepared) prepare(*S);
    if (t) *t = S->t;
    if (dt) *dt = S->dt;
    if (steps) *steps = S->steps;
    if (done) *done = S->done;
    return DSH_OK;
}

int orc_neighbour_stats(orc_sim* S, int species, int* count, uint64_t* hash) {
    if (!S || species < 0 || species >= S->nsp) return DSH_ERR_ARG;
    const Sp& q = S->sp[species];
    if (count) std::memcpy(count, q.ncount.data(), sizeof(int) * q.n);
    if (hash) std::memcpy(hash, q.nhash.data(), sizeof(uint64_t) * q.n);
    return DSH_OK;
}

int orc_sorted_ids(orc_sim* S, int species, int* ids) {
    if (!S || !ids || species < 0 || species >= S->nsp) return DSH_ERR_ARG;
    const Sp& q = S->sp[species];
    if ((int)q.order.size() != q.n) return DSH_ERR_STATE;
    std::memcpy(ids, q.order.data(), sizeof(int) * q.n);
    return DSH_OK;
}

int orc_grid(orc_sim* S, double* dbl, int* ints) {
    if (!S) return DSH_ERR_ARG;
    if (!S->prepared) prepare(*S);
    for (int a = 0; a < 3; ++a) {
        if (dbl) {
            dbl[a] = S->g.o[a];
            dbl[3 + a] = S->g.c[a];
            dbl[6 + a] = S->g.L[a];
        }
        if (ints) ints[a] = S->g.n[a];
    }
    if (ints) ints[3] = S->g.f;
    return DSH_OK;
}

int orc_num_avg_cells(orc_sim* S) {
    if (!S) return DSH_ERR_ARG;
    if (!S->prepared) prepare(*S);
    return S->navg;
}

int orc_avg_cells(orc_sim* S, int* ng, int* nd, double* vg_hat, double* vd_hat, double* psi, double* tbar,
                  double* eps) {
    if (!S || S->ng.empty()) return DSH_ERR_STATE;
    const size_t A = (size_t)S->navg, nd_ = (size_t)(S->nsp - 1);
    if (ng) std::memcpy(ng, S->ng.data(), sizeof(int) * A);
    if (nd) std::memcpy(nd, S->nd.data(), sizeof(int) * A * nd_);
    if (vg_hat) std::memcpy(vg_hat, S->vgh.data(), sizeof(double) * 3 * A);
    if (vd_hat) std::memcpy(vd_hat, S->vdh.data(), sizeof(double) * 3 * A * nd_);
    if (psi) std::memcpy(psi, S->psi.data(), sizeof(double) * 3 * A);
    if (tbar) std::memcpy(tbar, S->tbar.data(), sizeof(double) * A * nd_);
    if (eps) std::memcpy(eps, S->eps.data(), sizeof(double) * A * nd_);
    return DSH_OK;
}

int orc_tube_params(const dsh_tube_spec* t, dsh_params* p) {
    if (!t || !p) return DSH_ERR_ARG;
    std::memset(p, 0, sizeof(*p));
    p->x_min = t->x_min;
    p->x_max = t->x_max;
    p->y_min = 0.0;
    p->y_max = t->y_width;
    p->z_min = 0.0;
    p->z_max = t->z_width;
    p->eos_type = 1;
    p->max_particles = 0;
    p->particles_per_length = 1.0 / t->dx_left;
    p->h = 1.2 * t->dx_left;
    p->avg_cell_coeff = 0.0;
    p->tau = 0.0;
    p->t_end = t->t_end;
    p->t_out = t->t_end;
    p->alpha = 1.0;
    p->beta = 2.0;
    p->eps = 0.01;
    p->periodic_y = 1;
    p->periodic_z = 1;
    p->gamma = 1.4;
    p->cs = 1.0;
    p->h_mode = 1;
    p->eta = 1.2;
    p->n_dust = 1;
    p->drag_mode = 1;
    p->K[0] = t->K;
    p->K[1] = t->K;
    p->R_dust[0] = 0.01;
    p->R_dust[1] = 1.0;
    p->mass_ratio[0] = t->dust_to_gas;
    p->mass_ratio[1] = t->dust_to_gas;
    p->eps_fallback[0] = 0.5;
    p->eps_fallback[1] = 0.33;
    p->tstop_fallback = 1.0;
    p->eps2_uses_count1 = 0;
    p->dt_mode = t->dt_mode;
    p->cfl = t->cfl;
    p->max_steps = t->max_steps;
    p->cell_size = 0.0;
    p->cell_slack = 0.05;
    p->avg_factor = 1;
    p->precision = t->precision;
    p->fp64_wall_band = 1;
    p->wall_band_speed = 0.5;
    p->wall_band_horizon = 50;
    return DSH_OK;
}

int orc_tube_count(const dsh_tube_spec* t, int* n) {
    if (!t || !n) return DSH_ERR_ARG;
    TubeGeom G = tube_geom(*t);
    int c = G.nlx * G.nyl * G.nzl + G.nrx * G.nyr * G.nzr;
    n[0] = c;
    n[1] = c;
    n[2] = 0;
    return DSH_OK;
}

int orc_tube_ic(const dsh_tube_spec* t, int species, const dsh_soa_view* v) {
    if (!t || !v || species < 0 || species > 1) return DSH_ERR_ARG;
    TubeGeom G = tube_geom(*t);
    const double ratio = species == 0 ? 1.0 : t->dust_to_gas;
    const double cs = 1.0, gam = 1.4;
    int id = 0;
    for (int side = 0; side < 2; ++side) {
        const int nx = side == 0 ? G.nlx : G.nrx;
        const int ny = side == 0 ? G.nyl : G.nyr;
        const int nz = side == 0 ? G.nzl : G.nzr;
        const double dx = side == 0 ? G.dxl : G.dxr;
        const double dy = side == 0 ? G.dyl : G.dyr;
        const double dz = side == 0 ? G.dzl : G.dzr;
        const double rho = side == 0 ? G.rhol : G.rhor;
        const double hloc = 1.2 * std::cbrt(dx * dy * dz);
        for (int i = 0; i < nx; ++i)
            for (int j = 0; j < ny; ++j)
                for (int k = 0; k < nz; ++k) {
                    double x = side == 0 ? 0.0 - (double)i * dx : dx + (double)i * dx;
                    double y = 0.0 + (double)j * dy;
                    double z = 0.0 + (double)k * dz;
                    x = x + t->jitter * dx * rng_sym(t->seed, t->tube_index, species, id, 0);
                    y = y + t->jitter * dy * rng_sym(t->seed, t->tube_index, species, id, 1);
                    z = z + t->jitter * dz * rng_sym(t->seed, t->tube_index, species, id, 2);
                    y = wrap(y, 0.0, t->y_width, t->y_width);
                    z = wrap(z, 0.0, t->z_width, t->z_width);
                    const double r = rho * ratio;
                    if (v->x) v->x[id] = x;
                    if (v->y) v->y[id] = y;
                    if (v->z) v->z[id] = z;
                    if (v->vx) v->vx[id] = 0.0;
                    if (v->vy) v->vy[id] = 0.0;
                    if (v->vz) v->vz[id] = 0.0;
                    if (v->ax) v->ax[id] = 0.0;
                    if (v->ay) v->ay[id] = 0.0;
                    if (v->az) v->az[id] = 0.0;
                    if (v->m) v->m[id] = G.m * ratio;
                    if (v->rho) v->rho[id] = r;
                    if (v->p) v->p[id] = species == 0 ? cs * cs * r : 0.0;
                    if (v->e) v->e[id] = species == 0 ? cs * cs / (gam - 1.0) : 0.0;
                    if (v->h) v->h[id] = hloc;
                    if (v->t_stop) v->t_stop[id] = species == 0 ? 0.0 : r / t->K;
                    if (v->edot) v->edot[id] = 0.0;
                    const bool frozen = (x < t->x_min + t->frozen_h * hloc) || (x > t->x_max - t->frozen_h * hloc);
                    if (v->ind) v->ind[id] = frozen ? 1 : 0;
                    if (v->nn) v->nn[id] = -1;
                    if (v->cell) v->cell[id] = -1;
                    ++id;
                }
    }
    return DSH_OK;
}

}  // extern "C"
