/*
 * gade_oracle.h -- CPU oracle for the GaDE Crank-Nicolson hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This is a plain-C restatement of the reference's
 * CPU path (tools/dirac_solver.c, tools/solver.c, tools/csr.c,
 * tools/bicgstab.c, tools/solver_common.c, tools/sorted_list.c).  Only tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it, and
 * only as the checker; the product path never links or calls it.
 *
 * Pinning status per function is recorded next to each declaration and in
 * the test that pins it (tests/test_oracle_*.py).
 */
#ifndef GADE_ORACLE_H
#define GADE_ORACLE_H

#include <complex.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef double complex ocplx;

typedef struct {
    int nrows, nnz;
    int *Ap, *Ai;
    ocplx *Ax;
    int *perm;          /* optional trailer of H.csr: compressed -> original index */
    int npart;
    int *part;
} or_csr_t;

/* tools/solver_common.c:30-44.  Pinned: bit-exact vs oracle/_ref (reference source). */
int or_ik(int i);
double or_imu(int i);
/* tools/solver_common.c:46-65.  Pinned: closed forms + bit-exact vs oracle/_ref. */
void or_pulse(double E0, double omega, double cycles, double t, ocplx out[6]);
/* tools/sorted_list.c:32-57.  Pinned: brute force + bit-exact vs oracle/_ref. */
int or_locate(const int *list, int n, int value);

/* tools/csr.c:262-318.  Returns 0 on success, nonzero on a malformed file. */
int or_csr_read(const char *fname, or_csr_t *out);
void or_csr_free(or_csr_t *sp);
/* tools/csr.c:945-952 */
ocplx or_csr_get_value(const or_csr_t *sp, int row, int col);
/* tools/csr.c:974-1003: result[i] += (A x)[i] for rows [row_beg,row_end) */
void or_csr_spmv(int row_beg, int row_end, const int *Ap, const int *Ai, const ocplx *Ax,
                 const ocplx *x, ocplx *result);

/* tools/csr.c:653-743: unblocked Ap[nblk*blkdim+1], Ai[nnz], Ai_sub_map[nnz].
   Pinned: brute-force dense expansion + permutation invariants (tests). */
long long or_unblock_nnz(int nblk, const int *bAp, const int *subAp, int blkdim);
void or_unblock(int nblk, const int *bAp, const int *bAi, int blkdim, const int *subAp, const int *subAi,
                int *Ap, int *Ai, int *Ai_sub_map);

typedef struct or_model or_model_t;

/* Loads H.csr, H0l/H1l, g, h0, s from dir (tools/dirac_solver.c:223-370) and
   builds Hfull/S structure.  Returns NULL on error (message on stderr). */
or_model_t *or_model_load(const char *dir, int lmax, double h, double dt, int nthreads);
/* threads of the per-step assembly (nthreads > 0), and the all-cores baseline: the SpMV rows and the
   preconditioner blocks also run on them (the reference runs them serially); results are the serial ones
   bit for bit */
void or_model_set_threads(or_model_t *m, int nthreads, int all_cores);
void or_model_free(or_model_t *m);
void or_model_info(const or_model_t *m, int *nblk, int *blkdim, int *hall_nnz, long long *nnz_full,
                   int *nnz_b);
const int *or_model_hfull_Ap(const or_model_t *m);
const int *or_model_hfull_Ai(const or_model_t *m);
const int *or_model_hfull_submap(const or_model_t *m);
const ocplx *or_model_hfull_Ax(const or_model_t *m);
const ocplx *or_model_S_Ax(const or_model_t *m);
const int *or_model_S_Ap(const or_model_t *m);
const int *or_model_S_Ai(const or_model_t *m);
/* per Hall non-zero gathers of tools/gpu_solver.cu:137-188 */
void or_model_row_col(const or_model_t *m, int *Ai_orig, int *Aj_orig, int *Ki, int *Kj, ocplx *H0, ocplx *H1);

/* tools/solver.c:196-407: assemble Hfull.Ax for coefficients ft[6]. */
void or_model_timedep(or_model_t *m, const ocplx ft[6]);
/* out += Hfull x + S x   (tools/dirac_solver.c:115-126) */
void or_model_spmv_HS(const or_model_t *m, const ocplx *x, ocplx *out);
void or_model_spmv_H(const or_model_t *m, const ocplx *x, ocplx *out);
void or_model_spmv_S(const or_model_t *m, const ocplx *x, ocplx *out);
/* tools/solver.c:538-759: out += Hfull(ft) x computed without assembly. */
void or_model_matfree_H(or_model_t *m, const ocplx ft[6], const ocplx *x, ocplx *out);

/* Block-diagonal preconditioner P = S + i h dt/2 H0 (tools/solver.c:409-485):
   exact LU per distinct kappa, banded after interleaving the two radial halves. */
int or_model_precond_setup(or_model_t *m);
void or_model_psolve(const or_model_t *m, const ocplx *rhs, ocplx *x);

/* tools/bicgstab.c:125-295 with the restart of tools/gpu_bicgstab.c:222-241.
   The workspace persists across calls as in the reference (v, p are not reset).
   napply (optional) returns the number of A applications. */
void or_bicgstab(or_model_t *m, const ocplx *rhs, ocplx *x, int *iters, double *tol_error, int *napply);
void or_reset_workspace(or_model_t *m);

/* tools/dirac_solver.c:496-577 (CPU branch) for exactly nsteps steps from time t0.
   x holds psi on entry and on exit.  Per-step iters/tol/usec are optional. */
int or_cn_run(or_model_t *m, double E0, double omega, double cycles, double t0, int nsteps,
              double tol, int maxit, ocplx *x, int *iters_out, double *tol_out, long *usec_out);


/* psi0.vec (tools/dirac_solver.c:68-108; tools/main.py:151-155) */
int or_vec_read(const char *fname, int n, ocplx *out, int *perm);

#ifdef __cplusplus
}
#endif
#endif
