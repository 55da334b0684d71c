/*
 * gade.h -- MI355X-native Crank-Nicolson propagator for the GaDE hot path.
 *
 * C ABI of libgade.so: plain pointers and sizes only.  Complex numbers are
 * interleaved (re, im) doubles, i.e. the layout of C99 `double complex`,
 * which is the reference's csr_data_t (tools/types.h:28-35).  Vectors are
 * block-major: psi = [psi_0; ...; psi_{nblk-1}], psi_r of length blkdim.
 *
 * Device pointers ("_dev") live in HBM of the current HIP device; the
 * library runs every kernel on the stream given to gade_model_create
 * (0 = the null stream) and returns only after the work is complete unless
 * a function says otherwise.
 *
 * Every function returns 0 on success and a negative code on error;
 * gade_last_error() returns a message for the last failure on this thread.
 *
 * Entry points with a counterpart in the reference cite it; those that the
 * reference lacks say so.
 */
#ifndef GADE_H
#define GADE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define GADE_OK 0
#define GADE_EINVAL -1
#define GADE_EHIP -2
#define GADE_ENOMEM -3
#define GADE_EIO -4
#define GADE_ENOTCONV -5
#define GADE_EPIVOT -6

typedef struct gade_model gade_model_t;

/* Host-side description of the coupling files that dumpCouplings and
 * tools/main.py write (one shared radial pattern for g, h0 and s).
 * All pointers are host pointers, complex arrays interleaved (re, im). */
typedef struct {
    int nblk;            /* Hall block rows kept (H.csr dim) */
    const int *hall_Ap;  /* [nblk+1] */
    const int *hall_Ai;  /* [hall_nnz] */
    const int *hall_perm;/* [nblk] compressed -> original angular index, NULL = identity */
    int lmax;            /* number of l values (Nl) */
    int nang;            /* dimension of H0l/H1l (original angular space) */
    const int *const *H_Ap;     /* [2*lmax]: H0l0, H1l0, H0l1, ... */
    const int *const *H_Ai;
    const double *const *H_Ax;
    int blkdim;          /* radial dimension */
    const int *rad_Ap;   /* [blkdim+1] shared radial pattern */
    const int *rad_Ai;   /* [nnz_b] */
    const double *const *g;  /* [24*lmax] values, index alpha*4*lmax + l*4 + n (tools/dirac_solver.c:303-313) */
    const double *const *h0; /* [4] */
    const double *const *s;  /* [3] */
} gade_inputs_t;

const char *gade_last_error(void);

/* Device helpers so that callers without another GPU runtime can stage data. */
int gade_malloc(void **ptr_dev, size_t bytes);
int gade_free(void *ptr_dev);
int gade_memcpy_h2d(void *dst_dev, const void *src, size_t bytes);
int gade_memcpy_d2h(void *dst, const void *src_dev, size_t bytes);
int gade_memset(void *ptr_dev, int value, size_t bytes);
int gade_memcpy_d2d(void *dst_dev, const void *src_dev, size_t bytes);
int gade_device_sync(void);

/* Builds the device model: uploads couplings once, precomputes the block
 * gathers (tools/gpu_solver.cu:137-236), the kappa-grouped row tiles and the
 * per-kappa diagonal operators S(k) +- i h dt/2 H0(k) (tools/solver.c:409-458).
 * stream: a hipStream_t or NULL. */
int gade_model_create(const gade_inputs_t *in, double h, double dt, void *stream, gade_model_t **out);
/* A batch of B instances (pulse scans) sharing the couplings, S and H0: vectors are [B][nblk*blkdim]
 * (instance-major); each instance has its own f(t), BiCGSTAB scalars and convergence state.  B = 1 is
 * gade_model_create.  Batched models need the banded radial pattern. */
int gade_model_create_batch(const gade_inputs_t *in, int batch, double h, double dt, void *stream, gade_model_t **out);
int gade_model_load_batch(const char *dir, int lmax, int batch, double h, double dt, void *stream, gade_model_t **out);
int gade_model_batch(const gade_model_t *m);
/* Same, reading H.csr, H0l*, H1l*, g*, h0*, s* from a directory
 * (tools/dirac_solver.c:223-342). */
int gade_model_load(const char *dir, int lmax, double h, double dt, void *stream, gade_model_t **out);
int gade_model_destroy(gade_model_t *m);
/* nblk, blkdim, hall_nnz, nnz_b, nnz of the equivalent assembled Hfull, nkappa (distinct) */
int gade_model_info(const gade_model_t *m, int *nblk, int *blkdim, int *hall_nnz, int *nnz_b,
                    long long *nnz_full, int *nkappa);

/* Per-step G-sums (tools/solver.c:236-287; tools/gpu_solver.cu:50-82) for
 * time coefficients ft[6] (complex, interleaved: 12 doubles). */
int gade_set_time(gade_model_t *m, const double *ft);
/* f(t) per instance of a batched model: ft [batch][12] */
int gade_set_time_batch(gade_model_t *m, const double *ft);

/* Fused operator, no assembly (tools/solver.c:327-395 evaluated matrix-free):
 *   y = (S + cH * Hfull(t)) x + beta * y     with cH in {+1, -1, 0}.
 * Hfull includes the stationary i h dt/2 H0 blocks.  beta = 0 never reads y. */
int gade_apply(gade_model_t *m, const double *x_dev, double *y_dev, double cH, double beta);

/* Preconditioner P = S + i h dt/2 H0 per block (tools/solver.c:460-485):
 * exact LU of each distinct-kappa block on the GPU (FP64), then x = P^{-1} rhs
 * with chunked level-scheduled triangular solves.  gade_precond_setup is
 * called lazily by gade_psolve / gade_bicgstab. */
int gade_precond_setup(gade_model_t *m);
int gade_psolve(gade_model_t *m, const double *rhs_dev, double *x_dev);

/* Right-preconditioned BiCGSTAB of tools/gpu_bicgstab.c:133-307 on A = S + Hfull(t),
 * entirely on the device.  *iters: in max iterations, out iterations
 * (break-before-increment convention of the reference); *tol_error: in
 * relative tolerance, out ||r||/||b||.  napply (optional) returns the number
 * of A applications.  The workspace persists across calls, as the
 * reference's does. */
int gade_bicgstab(gade_model_t *m, const double *rhs_dev, double *x_dev, int *iters, double *tol_error,
                  int *napply);
int gade_reset_workspace(gade_model_t *m);

/* One CN step (tools/dirac_solver.c:503-577): given f(t_{n+1}) in ft,
 * rhs = (S - Hfull) psi, psi <- BiCGSTAB(S + Hfull, rhs, x0 = psi). */
int gade_cn_step(gade_model_t *m, double *psi_dev, const double *ft, double tol, int maxit, int *iters,
                 double *tol_error);

/* nsteps CN steps of the sin^2 pulse (E0, omega, cycles) starting at time t0
 * (time accumulated as time += dt, tools/dirac_solver.c:509).  Per-step
 * iterations, tolerances and microseconds are written when non-NULL. */
int gade_cn_run(gade_model_t *m, double *psi_dev, double E0, double omega, double cycles, double t0,
                int nsteps, double tol, int maxit, int *iters, double *tol_out, double *usec);

/* Batched pulse scan: nsteps CN steps of every instance of a batched model from t0 (one dt for all).
 * params [batch][4] = (E0, omega, cycles, cep) of each instance's pulse; nsteps_b [batch] steps of each
 * instance (NULL: all run nsteps; an instance idles after its own count, later entries untouched).
 * psi_dev [batch][N1] holds x0 and the results; iters / tol_out [batch][nsteps] (NULL allowed).
 * All instances advance in one set of launches per step; each converges on its own. */
int gade_cn_run_batch(gade_model_t *m, double *psi_dev, const double *params, const int *nsteps_b, double t0,
                      int nsteps, double tol, int maxit, int *iters, double *tol_out);

/* Autotuner of the device-driven stepping (no reference counterpart): times nsteps probe steps from
 * psi_dev (left unchanged) with the pulses params [batch][4] (as gade_cn_run_batch, t0 the start time),
 * best of reps runs per candidate.  Stage 1 (banded radial kernel): the tile shapes (dia_rows,
 * dia_outputs) in {(8,2), (4,2), (4,4), (8,1), (16,1), (4,1)} x output halves per workgroup dia_halves in
 * {1, 2} x the angular multipole group ang_lb in {4, 2, 8}, and the fused angular + radial form
 * (ang_fused 1) for the shapes with up to 2 outputs per lane, with the model's stepping (with the FP64 MFMA
 * radial kernel in FP64 only ang_lb varies); stage 2, with the fastest shapes:
 * window in {2,4,8,16,32}, iter_batch in {0,1,2,4}, graph in {0,1} (the graph only with the banded
 * kernel and when the model's "graph" parameter is 1).  The fastest stage-2 candidate is kept.  Every candidate
 * must give the same trajectory, iterations and tolerances bit for bit, else GADE_EINVAL with the
 * message naming it.  report [GADE_TUNE_MAX][GADE_TUNE_COLS] (NULL allowed) receives (window,
 * iter_batch, graph, dia_rows, dia_outputs, ang_lb, ms, dia_halves, ang_fused) per candidate and *ncand their count; the
 * chosen index
 * goes to *chosen.  Every probe starts from the
 * model's current solver state (BiCGSTAB workspace, scalars, guess histories), which is restored on
 * return: a run after tuning equals the same run without it. */
#define GADE_TUNE_MAX 96
#define GADE_TUNE_COLS 9

/* ---- learned initial guess (no reference counterpart) ----------------------------------------------
 * guess_order GADE_GUESS_LEARNED: x0 = x2 + s sum_k w_k D_k/|D_k| + s U c per step, with D_k the backward
 * differences of psi_n..psi_{n-5}, x2 = psi_n + D_1 + D_2 (order-2 extrapolation, also the fallback while
 * the history is short), s = |D_3|, U the POD basis and (w, c) the outputs of an MLP (GADE_LG_IN inputs,
 * 3 hidden layers of GADE_LG_HIDDEN with GELU, GADE_LG_OUT outputs: 6 weights then GADE_LG_POD complex
 * coefficients), BF16 weights with FP32 accumulation on MFMA.  Runs in the device-driven stepping with the
 * banded radial kernel.  Inputs (raw, standardized by gade_lg_set_norm): f_alpha(t_{n+1}) (6), log10 of
 * the last 6 final residuals relative to their bounds tol |rhs|, log10 |D_k|/|psi_n| (k = 1..5), U^H D_3 / s
 * and U^H D_4 / s
 * (re, im interleaved), zero padding.  Training sample (GADE_LG_SAMPLE floats): inputs, then
 * G = Re(Dh^H Dh) (6x6), a = Re(Dh^H eh) (6), P = U^H Dh (npod x 6 complex), q = U^H eh (npod complex),
 * beta = |eh|^2, valid, with Dh_k = D_k/|D_k| and eh = (psi_{n+1} - x2)/s; its loss is
 * |eh - Dh w - U c|^2 / beta, the guess error relative to the extrapolation's. */
#define GADE_GUESS_LEARNED 4
#define GADE_LG_IN 288
#define GADE_LG_HIDDEN 256
#define GADE_LG_OUT 144
#define GADE_LG_POD 64
#define GADE_LG_SAMPLE 1232
#define GADE_LG_PARAMS 242576

/* After a step run with set_param("lg_features", 1) and guess_order 2: write each instance's training
 * sample (samples [batch][GADE_LG_SAMPLE], device, may be NULL; needs the POD basis) and its normalized
 * extrapolation defect as snapshot column snap_col + b of snaps [snap_cols][N1] (device complex, may be
 * NULL).  psi_dev holds psi_{n+1}.  Instances that did not step, or whose history is shorter than 6,
 * give valid = 0 and no snapshot. */
int gade_lg_record(gade_model_t *m, const double *psi_dev, float *samples_dev, double *snaps_dev, int snap_col,
                   int snap_cols);
/* POD basis: the top npod (<= GADE_LG_POD) left singular vectors of the nsnap snapshots (device complex
 * [nsnap][N1]), by Householder QR and an SVD of R on the device; singular values to sv (host, nsnap). */
int gade_lg_pod_build(gade_model_t *m, const double *snaps_dev, int nsnap, int npod, double *sv);
int gade_lg_basis(gade_model_t *m, double *U_host);   /* [npod][N1] complex */
/* Installs a basis (host complex [npod][N1], orthonormal columns, e.g. from gade_lg_basis of another model of
 * the same couplings): with gade_lg_set_params / gade_lg_set_norm it moves a trained guess between models. */
int gade_lg_set_basis(gade_model_t *m, const double *U_host, int npod);
int gade_lg_set_norm(gade_model_t *m, const float *mean, const float *stdv);  /* [GADE_LG_IN] each */
int gade_lg_init(gade_model_t *m, unsigned seed);   /* He-normal hidden layers, zero output layer */
int gade_lg_set_params(gade_model_t *m, const float *params);  /* host FP32 [GADE_LG_PARAMS] */
int gade_lg_get_params(gade_model_t *m, float *params);
/* W1 [256][288] b1 W2 [256][256] b2 W3 b3 W4 [144][256] b4, row-major [out][in] */
int gade_lg_forward(gade_model_t *m, const float *x_dev, int M, float *y_dev);
int gade_lg_grad(gade_model_t *m, const float *samples_dev, const int *idx_dev, int M, float *loss,
                 float *grad_dev);
int gade_lg_train(gade_model_t *m, const float *samples_dev, const int *idx_dev, int M, int nsteps, float lr,
                  float *loss_hist);
/* Range gate of the learned guess: a step whose standardized inputs reach beyond set_param("lg_zmax", z)
 * standard deviations of the training inputs (default 8; 0 turns the gate off) takes the order-2
 * extrapolation instead of the network.  counts (host int [B][2]): per instance, the steps since the last
 * workspace reset that took the network and those the gate sent to the extrapolation.  No counterpart in the
 * reference (the learned guess is an addition). */
int gade_lg_uses(gade_model_t *m, int *counts);
int gade_autotune(gade_model_t *m, const double *psi_dev, const double *params, double t0, int nsteps,
                  double tol, int maxit, int reps, double *report, int *ncand, int *chosen);

/* Checkpoint / resume of a CN run (the reference's -s flag only shifts the start time,
 * tools/dirac_solver.c:203-209, and restarts from psi0.vec).  gade_checkpoint_save writes psi_dev, the step
 * number and the accumulated time (exact bits) with every piece of solver state that carries from one step to
 * the next (the BiCGSTAB workspace, scalars, control words and the extrapolation history); after
 * gade_checkpoint_load into a model of the same couplings, batch, precision and guess order, gade_cn_run from
 * t0 = *time continues the run bit for bit as if it had not stopped.  The learned guess's state is not saved
 * (GADE_EINVAL while it is active).  No reference counterpart for the state. */
int gade_checkpoint_save(gade_model_t *m, const double *psi_dev, int step, double time, const char *file);
int gade_checkpoint_load(gade_model_t *m, double *psi_dev, int *step, double *time, const char *file);

/* Pulse coefficients f_alpha(t), alpha = 0..5 (tools/solver_common.c:46-65),
 * as 12 doubles (complex with zero imaginary part). */
void gade_pulse(double E0, double omega, double cycles, double t, double *ft);
/* the same with carrier phase -pi/2 + cep (the reference fixes cep = 0, tools/solver_common.c:55) */
void gade_pulse_cep(double E0, double omega, double cycles, double cep, double t, double *ft);
int gade_ik(int i);       /* tools/solver_common.c:30-37 */
double gade_imu(int i);   /* tools/solver_common.c:39-44 */

/* S-norm <x|S|x> and per-block populations |x_r|^2_S (FP64 on the device).
 * pop_host (optional) receives nblk doubles.  No reference counterpart on the
 * solver side: the reference computes these offline (src/hamiltonian.h:1268-1307). */
int gade_snorm(gade_model_t *m, const double *x_dev, double *norm, double *pop_host);

/* Initial guess by polynomial extrapolation (no reference counterpart; the
 * reference starts each solve from psi_n, tools/dirac_solver.c:548): pushes
 * psi_n (in psi_dev) into the model's history and overwrites psi_dev with
 * sum_{k=0..p} (-1)^k binom(p+1,k+1) psi_{n-k}, p = min(order, history length),
 * order <= 3.  With set_param("guess_order", p) every CN step does this before
 * r = b - A x0; the converged solution at the same tolerance is unchanged in
 * meaning, only the iteration count moves.  gade_reset_workspace clears the history. */
int gade_extrapolate(gade_model_t *m, double *psi_dev, int order);

/* Populations by eigen-projection (src/hamiltonian.h:872-943, 1268-1307; driven
 * by src/process_output.cpp:332-355).  For every distinct kappa the generalized
 * eigenpairs H0(k) v = E S(k) v with v^H S v = 1 are computed once per model on
 * the device (gade_eig_setup, done on first use).  gade_populations projects
 * S(k_r) psi_r of every block r on them: population_j = |v_j^H S psi_r|^2,
 * summed into bound (-c^2 <= E < 0) and ionised (E >= 0); norm = psi^H S psi.
 * per_block (optional) receives [nblk][4]: negative continuum (E < -c^2),
 * bound, ionised, psi_r^H S psi_r. */
int gade_eig_setup(gade_model_t *m);
/* The dense solver behind gade_eig_setup (host, no reference counterpart; the reference uses Eigen's
 * GeneralizedSelfAdjointEigenSolver, src/hamiltonian.h:872-943): real symmetric A, symmetric positive
 * definite B, both n x n row-major; E [n] ascending eigenvalues, V [n][n] row-major with column j the
 * eigenvector of E[j], V^T B V = I, each column's largest-magnitude entry positive.  GADE_EINVAL when B is
 * not positive definite, GADE_ENOTCONV when the iteration does not converge. */
int gade_eig_dense(int n, const double *A, const double *B, double *E, double *V);
/* distinct kappa values [nkappa] and energies [nkappa][blkdim] (ascending per kappa) */
int gade_eig_values(gade_model_t *m, int *nkappa, int *kappa, double *energies);
int gade_populations(gade_model_t *m, const double *x_dev, double *norm, double *bound, double *ionised,
                     double *per_block);
/* Eigen-state amplitudes of a state, as src/hamiltonian.h:1246-1331 (eigProj) computes them for
 * savePsievs: psiev [nblk][blkdim] complex (host), psiev[r][j] = psi_r^H S(k_r) v_j over the eigenpairs
 * of block r's kappa (E ascending, gade_eig_values). */
int gade_psiev(gade_model_t *m, const double *x_dev, double *psiev);
/* The eigenpair files of src/dumpEigvals.cpp:239-252 for every distinct kappa of the model, in dir:
 * evl<kappa>.mat (int64 rows = blkdim, int64 cols = 1, double energies ascending) and evc<kappa>.mat
 * (int64 blkdim, int64 blkdim, complex128 column-major: column j the S-normalized eigenvector of energy j). */
int gade_eig_write(gade_model_t *m, const char *dir);

/* Assembled reference formulation (kept as the plain-FP64 HBM comparator):
 * builds Hfull Ap/Ai/Ai_sub_map on the device (tools/csr.c:653-743), assembles
 * Hfull.Ax each step (tools/gpu_solver.cu:84-135) and multiplies it with a
 * hand-written CSR SpMV (replacing hipsparseSpMV, tools/gpu_sparse.c:267-313). */
int gade_assembled_setup(gade_model_t *m);
int gade_assembled_timedep(gade_model_t *m);  /* uses the ft of the last gade_set_time */
/* Same, writing Hfull.Ax into a caller's device CSR with the expanded row
 * pointers Ap_dev (the drop-in gpu_compute_timedep_matrices uses it). */
int gade_assembled_timedep_into(gade_model_t *m, const int *Ap_dev, double *Ax_dev);
/* y = alpha * Hfull x + beta * y, then y += S x when with_S
 * (the rhs of tools/dirac_solver.c:535-536 is alpha = -1, beta = 0, with_S = 1) */
int gade_assembled_spmv(gade_model_t *m, const double *x_dev, double *y_dev, double alpha, double beta, int with_S);
/* device views of the assembled structure (int32) for parity checks */
int gade_assembled_get(gade_model_t *m, int *Ap, int *Ai, int *Ai_sub_map, double *Ax);

/* Tunable launch parameters (no reference counterpart; used by the autotuner):
 *   "tile_r"  row blocks per radial tile of the ELL kernel (1, 2, 4, 8);
 *   "radial"  radial-phase kernel: 0 banded when the pattern allows (default),
 *             1 general ELL kernel, 2 banded (DIA) kernel; reads back 1 or 2;
 *   "device_loop"  1 (default): gade_cn_run decides convergence on the device and
 *             synchronizes once per window of steps; 0: one host round trip per
 *             BiCGSTAB iteration (the reference's control flow);
 *   "window"  steps enqueued per host synchronization (default 8);
 *   "iter_batch"  BiCGSTAB iterations enqueued per step, 0 = adaptive (default);
 *             a step needing more is finished by the host loop, same result;
 *   "profile_period"  time every n-th launch of a profiled family (default 1);
 *   "profile_floor_ns"  timed launches shorter than this (those a device-side
 *             gate ended at once) are left out of the family totals (default 0);
 *   "guess_order"  initial guess of each CN step: 0 psi_n (default, the
 *             reference), 1..3 polynomial extrapolation (gade_extrapolate),
 *             GADE_GUESS_LEARNED the learned guess;
 *   "dia_rows", "dia_outputs"  banded tile shape: rows per tile and outputs per
 *             lane, (8,1) by default; (4,1), (4,2), (4,4), (8,2), (16,1) also;
 *   "dia_halves"  1 (default): one wave per output half of a tile; 2: both halves
 *             in one two-wave workgroup sharing each staged window;
 *   "ang_lb"  multipoles per angular block (2 by default; 4, 8);
 *   "ang_list"  1 (default): angular phase over the coupling lists; 0: Hall rows;
 *   "radial_mfma"  1: FP64 banded products on the matrix cores (0 by default);
 *   "psolve_chunk"  preconditioner level height, 16 (default) or 32 rows;
 *   "precision"  0 (default): FP64 solve; 1: FP32 inner solves with FP64 defect
 *             correction (gade_cn_run / gade_cn_run_batch);
 *   "graph"  1 (default): device-driven windows replay a captured step graph;
 *   "ang_fused"  1: the banded radial kernel forms its windows of the angular
 *             phase from x and the coupling lists itself (no angular kernel, no V in
 *             HBM; tile shapes with up to 2 outputs per lane); 0 (default): the
 *             angular kernel writes V, the radial kernel reads it;
 *   "lg_zmax"  range gate of the learned guess's inputs (gade_lg_uses).
 * Results do not depend on the choice. */
int gade_set_param(gade_model_t *m, const char *name, int value);
int gade_get_param(const gade_model_t *m, const char *name, int *value);

/* Per-family kernel timing with HIP events on the model's stream.  'mask'
 * selects the families (GADE_PROF_*; GADE_PROF_ALL for every one, 0 stops and
 * clears).  Events are recorded without synchronizing; gade_profile_get
 * waits for the recorded launches and returns their accumulated time. */
enum {
    GADE_PROF_GSUMS = 1, GADE_PROF_ANGULAR = 2, GADE_PROF_RADIAL = 4, GADE_PROF_PSOLVE = 8, GADE_PROF_BLAS1 = 16,
    GADE_PROF_ASSEMBLE = 32, GADE_PROF_CSR_SPMV = 64, GADE_PROF_RADIAL_SUM = 128,
    GADE_PROF_ALL = 0x7fffffff
};
int gade_profile(gade_model_t *m, int mask);
int gade_profile_get(gade_model_t *m, const char *name, double *total_ms, long long *count);

#ifdef __cplusplus
}
#endif
#endif
