/*
 * dsh.h - C ABI of the MI355X-native dusty-gas SPH step (gas + dust, IDIC drag).
 *
 * The reference (DustyShock3D) has no library API: its operator boundary is the
 * set of kernel signatures plus the launch protocol inside main() (kernel.cu:1929-2566,
 * IDIC.h:2-346), over the global structure-of-arrays state of data.h.  Every entry
 * point below replaces one part of that protocol; the comment on each names the
 * reference lines it stands in for.
 *
 * Conventions
 *   - plain C types only; every function returns an int status (DSH_OK = 0,
 *     negative on error) and never aborts; dsh_last_error() explains the last failure.
 *   - particle arrays are in the data.h layout: one double (or int) array per field,
 *     indexed by the original particle index 0..Pm (Pm = count - 1, kernel.cu:255).
 *     Uploads borrow the caller's arrays for the duration of the call only.
 *   - species 0 is gas, species 1..n_dust are dust species (data.h:19-20,49-55).
 *   - Ind / ind: 0 = active particle, 1 = frozen wall particle (IDIC.h:8).
 *   - a context owns all of its device memory; one context per host thread.
 */
#ifndef DSH_H_
#define DSH_H_

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define DSH_MAX_DUST 2
#define DSH_MAX_SPECIES (1 + DSH_MAX_DUST)

#define DSH_OK 0
#define DSH_ERR_ARG (-1)
#define DSH_ERR_HIP (-2)
#define DSH_ERR_NOMEM (-3)
#define DSH_ERR_STATE (-4)
#define DSH_ERR_GRID (-5)        /* smoothing lengths outgrew the neighbour grid */
#define DSH_ERR_NONFINITE (-6)   /* a tube produced NaN/Inf */
#define DSH_ERR_IO (-7)
#define DSH_ERR_BUDGET (-8)      /* request larger than the HBM budget */
#define DSH_ERR_NOTFOUND (-9)
#define DSH_ERR_UNSUPPORTED (-10)

/* dt_mode */
#define DSH_DT_FIXED 0       /* tau from Init.txt every step (kernel.cu:1931) */
#define DSH_DT_CFL_ONCE 1    /* CFL step computed from the initial state, then fixed */
#define DSH_DT_CFL 2         /* CFL step recomputed every step, last step lands on t_end */

/* precision */
#define DSH_FP64 0           /* all arithmetic in FP64 */
#define DSH_MIXED 1          /* pair terms in FP32 accumulated in FP64; neighbour test, wall band,
                                drag, positions, conservation sums in FP64 */

/*
 * Run parameters.  The first block is Init.txt lines 2-18 in file order
 * (parser kernel.cu:1705-1722); the second block is the extension lines that
 * follow them in an extended Init.txt (old files simply omit them).
 */
typedef struct dsh_params {
    double x_min, x_max, y_min, y_max, z_min, z_max;
    int    eos_type;              /* 0: p = (gamma-1) rho e, 1: p = cs^2 rho (kernel.cu:92-97) */
    int    max_particles;         /* Maximum_particle (capacity hint) */
    double particles_per_length;  /* Particle_on_length */
    double h;                     /* global smoothing length (h_mode 0) */
    double avg_cell_coeff;        /* Coeff_h_dust_cell: averaging-cell width / h */
    double tau;                   /* fixed time step (dt_mode 0) */
    double t_end;
    double t_out;
    double alpha, beta, eps;      /* Monaghan viscosity (IDIC.h:225-235) */

    int    periodic_y, periodic_z;
    double gamma;                 /* Gam_g (data.h:7); also used by the viscosity sound speed */
    double cs;                    /* Speed_of_sound_gas (kernel.cu:1744) */
    int    h_mode;                /* 0: global h, 1: h_i = eta (m_i / rho_i)^(1/3) */
    double eta;
    int    n_dust;                /* dust species: 1 or 2; 0 runs gas alone with the plain explicit
                                     scheme of kernel.cu:48-228 */
    int    drag_mode;             /* 0: the reference law t_stop = R_dust / (sqrt(gamma p_g / rho_g) rho_g)
                                     with the gas particle of the same index (IDIC.h:78-91); 1: t_stop = rho_dust / K */
    double K[DSH_MAX_DUST];       /* drag_mode 1 */
    double R_dust[DSH_MAX_DUST];  /* drag_mode 0 (R of t_stop_Kernel_IDIC, IDIC.h:78-91) */
    double mass_ratio[DSH_MAX_DUST];   /* m_dust / m_gas: eps = n_d mu / n_g (kernel.cu:2192-2211) */
    double eps_fallback[DSH_MAX_DUST]; /* 0.5, 0.33 in the reference (kernel.cu:2200,2209) */
    double tstop_fallback;             /* 1.0 in the reference (kernel.cu:2160,2188) */
    int    eps2_uses_count1;           /* reproduce the test of kernel.cu:2203 */
    int    dt_mode;
    double cfl;
    int    max_steps;             /* 0: no step limit */
    double cell_size;             /* 0: automatic, 2 h_min (1 + cell_slack) */
    double cell_slack;
    int    avg_factor;            /* averaging cell = avg_factor^3 neighbour cells (1 for the shock tubes) */
    int    precision;             /* DSH_FP64 or DSH_MIXED */
    /* DSH_MIXED only.  A tube's momentum changes only through the force of its frozen x-end
       walls: the FP32 pair terms of two active particles are antisymmetric bit for bit and
       accumulate in FP64, so they cancel (exactly for equal masses).  With fp64_wall_band = 1
       the particles within f * wall_band_speed * cs * (time left) + (3 + f) neighbour cells of
       the frozen layers (f = avg_factor) run their density and force in FP64 (per 64-particle
       chunk; a pair is FP64
       when both particles are in the band or one is frozen), so the wall force, and with it
       the total momentum, follows an FP64 run: to 1e-13 per step.  Over a run it does until
       FP32 rounding from outside the band reaches the wall layers (a cfg0-sized 50-step tube:
       <= 1e-9 for 25 steps at speed 0.5, 4e-11 after 50 steps at speed 2) and until a particle
       is assigned to another averaging cell than in the FP64 run; from then on, cells that take
       the reference's fallback drag (kernel.cu:2160-2211) move the total momentum by
       particle-sized amounts.  The band's time term needs a finite run (t_end < 1e29 or
       max_steps > 0); without one, 3 + f cells.  It looks ahead at most wall_band_horizon
       steps: further ahead the first membership difference decides the momentum difference
       anyway (a cfg1 run: the same 2e-5 after 100 steps for speeds 0.5, 1 and 2). */
    int    fp64_wall_band;        /* 1 (default) or 0 */
    double wall_band_speed;       /* band edge speed in units of cs (default 0.5) */
    int    wall_band_horizon;     /* steps the band's time term covers at most (default 50; 0: the whole run) */
} dsh_params;

/* data.h structure-of-arrays view; NULL fields are skipped. */
typedef struct dsh_soa_view {
    double *x, *y, *z;
    double *vx, *vy, *vz;
    double *ax, *ay, *az;         /* gas acceleration of the last step (download only) */
    double *m, *rho, *p, *e, *h;
    double *t_stop;               /* dust stopping time */
    double *edot;                 /* gas de/dt of the last step (download only) */
    int    *ind;
    int    *nn;                   /* next particle in the same cell, -1 ends (download only) */
    int    *cell;                 /* neighbour-cell key of the last step (download only) */
} dsh_soa_view;

/* One isothermal dusty shock tube: the problem the benchmark configurations use. */
typedef struct dsh_tube_spec {
    double x_min, x_max;          /* -0.5, 0.5 */
    double y_width, z_width;      /* periodic cross-section [0, w) */
    double dx_left;               /* lattice spacing for x <= 0; x > 0 uses 2 dx_left */
    double rho_left;              /* gas density left; right is rho_left / 8 */
    double dust_to_gas;
    double K;                     /* drag coefficient, t_stop = rho_dust / K */
    double jitter;                /* uniform position noise, fraction of local spacing */
    double frozen_h;              /* frozen x-end layers, in units of h */
    double t_end;
    double cfl;
    int    dt_mode;
    int    max_steps;
    int    precision;
    int    tube_index;            /* RNG stream of this tube */
    uint64_t seed;
} dsh_tube_spec;

/* ---- problem set-up (host only; replaces init_Sod_X_Dust, kernel.cu:1391-1693) ---- */
int dsh_default_params(dsh_params* out);
int dsh_tube_params(const dsh_tube_spec* spec, dsh_params* out);
int dsh_tube_count(const dsh_tube_spec* spec, int* n_per_species);
int dsh_tube_ic(const dsh_tube_spec* spec, int species, const dsh_soa_view* out);
/* Init.txt reader (kernel.cu:1705-1742); extension lines after line 18 are optional.
   Without them the parameters are the reference run's: two dust species with the
   reference stopping-time law (drag_mode 0), global h, fixed tau. */
int dsh_read_init(const char* path, dsh_params* out);
/* The reference's own problem, init_Sod_X_Dust (kernel.cu:1391-1575): applies the parameter
   overrides of that function to *p (two dust species, mass ratios 0.75 / 0.5, R_dust 0.01 / 1,
   averaging cell 0.5 h, kernel.cu:1495-1502) and returns the particle count of each species. */
int dsh_reference_setup(dsh_params* p, int* n_per_species);
/* Particle placement of init_Sod_X_Dust for one species (0 gas, 1..2 dust), index order. */
int dsh_reference_ic(const dsh_params* p, int species, const dsh_soa_view* out);

/* ---- output (host only; replaces Data_out / Data_out_dust, kernel.cu:232-826) ---- */
/* Writes <dir>/G%07d.dat and L_X%07d.dat (species 0) or D-k-%07d.dat and D-k-L_X%07d.dat
   (dust species k) in the reference's text format for particles 0..pm of the view; t is
   the time, tau and h the Init.txt values printed in the headers.  NULL fields print 0
   (nn prints -1). */
int dsh_write_data(const char* dir, int num, double t, double tau, double h, int species,
                   const dsh_soa_view* view, int pm);
/* Next-particle pointers of the reference's neighbour lists on its 2h grid
   (kernel.cu:1767-1778, 1937-1995) for positions x, y, z of particles 0..pm. */
int dsh_reference_nn(const dsh_params* p, const double* x, const double* y, const double* z, int pm, int* nn);
/* sizeof of dsh_params, dsh_soa_view, dsh_tube_spec, dsh_server_opts (binding checks) */
int dsh_struct_sizes(size_t out[4]);

/* ---- single-tube context (replaces kernel.cu:1793-1926 set-up and 1929-2566 loop) ---- */
typedef struct dsh_ctx dsh_ctx;
int dsh_create(const dsh_params* params, dsh_ctx** out);
/* Upload species `species` with pm+1 particles (H2D of kernel.cu:1878-1926). */
int dsh_upload_species(dsh_ctx* ctx, int species, const dsh_soa_view* view, int pm);
/* Advance up to nsteps steps (fewer if the run reaches t_end / max_steps).
   t_inout receives the simulated time afterwards (Tm, kernel.cu:1931). */
int dsh_step(dsh_ctx* ctx, int nsteps, double* t_inout);
/* Fill the view in original particle order (D2H of kernel.cu:2531-2558). */
int dsh_download_species(dsh_ctx* ctx, int species, const dsh_soa_view* view);
/* out[0..2] = total momentum, out[3] = total energy (FP64, after the last step). */
int dsh_conservation(dsh_ctx* ctx, double out[4]);
int dsh_status(dsh_ctx* ctx, double* t, double* dt, int* steps, int* done);
/* Neighbour grid built at the first step (the role of Clx..Cnz, kernel.cu:1767-1778):
   dbl[0..2] origin, [3..5] cell size, [6..8] axis length; ints[0..2] cells per axis,
   ints[3] averaging factor, ints[4] averaging cells in total. */
int dsh_grid(dsh_ctx* ctx, double dbl[9], int ints[5]);
/* Neighbour-set fingerprint of the last density pass, original particle order:
   count = neighbours incl. self and periodic images, hash = order-free sum of ids. */
int dsh_set_stats(dsh_ctx* ctx, int on);
/* Pair-kernel variant: DSH_PAIR_TILED (default: one wave per 64-target chunk, LDS-staged
   candidate rows) or DSH_PAIR_DIRECT (one thread per target).  Same neighbour sets. */
#define DSH_PAIR_DIRECT 0
#define DSH_PAIR_TILED 1
int dsh_set_pair_variant(dsh_ctx* ctx, int variant);
/* Candidates staged per LDS piece by the tiled pair kernels: 128, 160, 192 (default), 224 or 256.
   Pieces split each candidate row in order, so every size gives the same results bit for bit. */
int dsh_set_tile(dsh_ctx* ctx, int tile);
/* Waves per block of the tiled pair kernels: 1 (default), 2 or 4 (no reference counterpart;
   the reference launches fixed blocks, kernel.cu:1867-1868).  Each wave keeps its own
   64-target chunk and LDS piece, so every block shape gives the same results bit for bit. */
int dsh_set_block_waves(dsh_ctx* ctx, int waves);
/* The block shape in use: shape[0] = piece size, shape[1] = waves per block. */
int dsh_block_shape(dsh_ctx* ctx, int shape[2]);
/* Autotuner (no reference counterpart): runs steps_per_trial steps with each block shape
   (piece size x waves per block; with with_variants also the direct pair kernels, whose
   sums round in another order) on the live state, keeps the fastest and returns
   chosen[0] = tile, chosen[1] = variant and its time per step (dsh_block_shape gives the
   chosen waves).  The trial steps are ordinary steps of the run. */
int dsh_autotune(dsh_ctx* ctx, int steps_per_trial, int with_variants, int chosen[2], double* ms_per_step);
int dsh_neighbour_stats(dsh_ctx* ctx, int species, int* count, uint64_t* hash);
/* Particle indices in the device's sorted order of the last step: by neighbour cell, inside a
   cell by x at the start of the step, then by index (the reference's lists, kernel.cu:1937-1995,
   are the same cells in index order; dsh_download_species rebuilds them as nn). */
int dsh_sorted_ids(dsh_ctx* ctx, int species, int* ids);
/* Per-kernel GPU time of the last timed step, ms (HIP events on the step stream). */
int dsh_set_timing(dsh_ctx* ctx, int on);
int dsh_kernel_times(dsh_ctx* ctx, double* ms, char* names, int max_kernels, int name_len);
const char* dsh_last_error(dsh_ctx* ctx);
void dsh_destroy(dsh_ctx* ctx);

/* No counterpart in the reference: the variant of dsh_step (with dsh_create,
   dsh_upload_species and dsh_download_species around it) for callers that cannot allocate
   device memory, such as code under HIP graph capture or tight serving loops.  One tube of
   counts[s] particles per species s is read from the caller's DEVICE arrays in[s] (data.h
   layout, original index order; x, y, z, m, h, ind required, other fields 0 when NULL),
   stepped up to nsteps steps from t = 0 with the kernels of dsh_step (tiled pair kernels,
   piece size 192), and written to the caller's DEVICE arrays out[s] in original order (nn
   must be NULL; cell as dsh_download_species).  *t_out and cons_out (optional) receive the
   time and the conservation sums.  The step's scratch (sorted state, keys, cell arrays, hit
   masks, partials) comes from one static 64 MiB array per GPU, not from an allocation: a tube
   that does not fit (roughly 2 x 40,000 particles on a few thousand cells) is rejected with
   DSH_ERR_BUDGET before anything runs.  Runs on `stream` (a hipStream_t, NULL for the null
   stream) and waits for it at the end.  Calls that overlap on one GPU share its array and race. */
int dsh_step_noalloc(const dsh_params* params, const dsh_soa_view* in, const int* counts, int nsteps,
                     const dsh_soa_view* out, double* t_out, double cons_out[4], void* stream);

/* ---- batch of independent tubes (no reference counterpart: the reference runs one
   problem per process; the batch runs many in one launch sequence per step) ---- */
typedef struct dsh_batch dsh_batch;
/* ntubes tubes with their own parameters; every tube must be uploaded before stepping */
int dsh_batch_create(const dsh_params* params, int ntubes, dsh_batch** out);
/* convenience: parameters and seeded set-up of every tube from its spec, generated on the device */
int dsh_batch_from_specs(const dsh_tube_spec* specs, int ntubes, dsh_batch** out);
int dsh_batch_upload(dsh_batch* b, int tube, int species, const dsh_soa_view* view, int pm);
/* advance every unfinished tube by up to nsteps steps; *ndone = finished tubes */
int dsh_batch_run(dsh_batch* b, int nsteps, int* ndone);
int dsh_batch_status(dsh_batch* b, int tube, double* t, double* dt, int* steps, int* done);
int dsh_batch_download(dsh_batch* b, int tube, int species, const dsh_soa_view* view);
int dsh_batch_conservation(dsh_batch* b, int tube, double out[4]);
int dsh_batch_set_pair_variant(dsh_batch* b, int variant);
int dsh_batch_set_tile(dsh_batch* b, int tile);
int dsh_batch_set_block_waves(dsh_batch* b, int waves);
int dsh_batch_block_shape(dsh_batch* b, int shape[2]);
int dsh_batch_autotune(dsh_batch* b, int steps_per_trial, int with_variants, int chosen[2], double* ms_per_step);
int dsh_batch_set_timing(dsh_batch* b, int on);
int dsh_batch_kernel_times(dsh_batch* b, double* ms, char* names, int max_kernels, int name_len);
/* bytes of device memory the batch holds */
int dsh_batch_device_bytes(dsh_batch* b, double* bytes);
const char* dsh_batch_last_error(dsh_batch* b);
void dsh_batch_destroy(dsh_batch* b);

/* ---- serving: continuous batching of tube requests (no reference counterpart) ----
   Requests queue in arrival order.  Each pump() admits queued requests while their HBM
   footprint fits the budget (and the rebuild transient fits the capacity), rebuilds the running batch (running tubes move device to
   device, keeping their state and clock), advances it steps_per_epoch steps and retires
   finished tubes, copying their results to host memory. */
typedef struct dsh_server dsh_server;
typedef double (*dsh_clock_fn)(void* user);
#define DSH_BACKEND_GPU 0
#define DSH_BACKEND_FAKE 1     /* host-only simulated device, for scheduler tests */
#define DSH_REQ_QUEUED 0
#define DSH_REQ_RUNNING 1
#define DSH_REQ_DONE 2
#define DSH_REQ_CANCELLED 3
#define DSH_REQ_FAILED 4
#define DSH_REQ_REJECTED 5     /* larger than the budget */
typedef struct dsh_server_opts {
    double hbm_budget;         /* bytes the running batch may hold */
    double hbm_capacity;       /* bytes the old and the new batch may hold together during a
                                  rebuild; 0: 90% of device memory (GPU) or 2 x budget (fake) */
    int steps_per_epoch;
    int backend;               /* DSH_BACKEND_GPU or DSH_BACKEND_FAKE */
    int pair_variant;
    dsh_clock_fn clock;        /* NULL: monotonic wall clock in seconds */
    void* clock_user;
    int fake_fail_rebuild;     /* fake backend: the n-th rebuild (1-based) fails; 0 never */
    /* Heavy requests: estimated work (particles x CFL steps to t_end, or max_steps) above
       heavy_work.  At most max_heavy of them run at once (0 counts as 1), started in arrival
       order; the queue waits behind a heavy request that cannot start, as behind one that does
       not fit the budget.  Light requests keep batching.  Two heavy requests that share the GPU
       both finish late; served one after the other, only the second waits.  heavy_work = 0:
       no distinction. */
    double heavy_work;
    int max_heavy;
} dsh_server_opts;
int dsh_server_create(const dsh_server_opts* opts, dsh_server** out);
int dsh_server_submit(dsh_server* sv, const dsh_tube_spec* spec, double arrival, uint64_t* id);
int dsh_server_pump(dsh_server* sv, int* running, int* queued);
int dsh_server_poll(dsh_server* sv, uint64_t id, int* state, double* arrival, double* admitted,
                    double* finished, double* particle_steps);
int dsh_server_result(dsh_server* sv, uint64_t id, int species, const dsh_soa_view* view);
int dsh_server_cancel(dsh_server* sv, uint64_t id);
int dsh_server_footprint(const dsh_tube_spec* spec, double* bytes);
/* Estimated work of a request: particles x estimated steps (max_steps, or t_end over the CFL
   step of the set-up at rest), the figure heavy_work is compared with.  A lower bound on the
   steps (the CFL step shrinks as the flow speeds up) that orders requests by cost. */
int dsh_server_work(const dsh_tube_spec* spec, double* work);
int dsh_server_stats(dsh_server* sv, double* bytes_in_use, int* rebuilds, int* epochs);
/* Host wall-clock seconds spent in batch rebuilds (allocation, adoption, set-up) and in
   advancing epochs (step launches and the wait for them), since creation. */
int dsh_server_times(dsh_server* sv, double* rebuild_s, double* advance_s);
const char* dsh_server_last_error(dsh_server* sv);
void dsh_server_destroy(dsh_server* sv);

#ifdef __cplusplus
}
#endif
#endif /* DSH_H_ */
