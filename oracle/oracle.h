/*
 * oracle.h - FP64 OpenMP restatement of DustyShock3D's per-step path (the parity checker).
 *
 * TEST INFRASTRUCTURE.  Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline
 * leg may load liboracle.so, and only as the checker / CPU baseline; the product path
 * (libdsh.so) never links or calls it.
 */
#ifndef DSH_ORACLE_H_
#define DSH_ORACLE_H_

#include "../include/dsh.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct orc_sim orc_sim;

int orc_create(const dsh_params* p, orc_sim** out);
int orc_upload(orc_sim* s, int species, const dsh_soa_view* v, int pm);
int orc_step(orc_sim* s, int nsteps, double* t_out);
int orc_download(orc_sim* s, int species, const dsh_soa_view* v);
int orc_conservation(orc_sim* s, double out[4]);
int orc_status(orc_sim* s, double* t, double* dt, int* steps, int* done);
/* neighbour fingerprint of the last density pass (original order; frozen particles 0) */
int orc_neighbour_stats(orc_sim* s, int species, int* count, uint64_t* hash);
/* sorted order of the last step: ids[k] = particle at sorted position k (key, then id) */
int orc_sorted_ids(orc_sim* s, int species, int* ids);
/* grid: dbl[0..2] origin, [3..5] cell size, [6..8] periodic length; ints[0..2] cells, [3] avg factor */
int orc_grid(orc_sim* s, double* dbl, int* ints);
/* averaging-cell state of the last step, per cell: n_g, n_d[k] and the solved velocities */
int orc_num_avg_cells(orc_sim* s);
int orc_avg_cells(orc_sim* s, int* ng, int* nd, double* vg_hat, double* vd_hat, double* psi,
                  double* tbar, double* eps);
void orc_destroy(orc_sim* s);
void orc_set_threads(int n);
int orc_get_threads(void);

/* unit-level functions */
double orc_W(double r, double h);
double orc_dW(double r, double h);
/* implicit drag solve of one averaging cell; arrays of 3 components per species
   (vd, vd_hat: [n_dust][3]; tbar, eps: [n_dust]) */
void orc_idic_cell(int n_dust, const double* vg, const double* vd, const double* psi,
                   const double* tbar, const double* eps, double tau,
                   double* vg_hat, double* vd_hat);

/* independent restatement of the shock-tube set-up */
int orc_tube_params(const dsh_tube_spec* spec, dsh_params* out);
int orc_tube_count(const dsh_tube_spec* spec, int* n_per_species);
int orc_tube_ic(const dsh_tube_spec* spec, int species, const dsh_soa_view* out);

#ifdef __cplusplus
}
#endif
#endif
