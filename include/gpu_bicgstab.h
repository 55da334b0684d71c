/*
 * gpu_bicgstab.h -- right-preconditioned BiCGSTAB on the device
 * (drop-in for tools/gpu_bicgstab.h:27-61).
 *
 * The caller supplies the operator and preconditioner as callbacks, exactly as
 * in the reference; dots, norms and vector updates are libgade's fused
 * kernels with device-resident scalars (one host read per iteration).
 */
#ifndef GADE_GPU_BICGSTAB_H
#define GADE_GPU_BICGSTAB_H

#include "gpu_sparse.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct {
    gpu_dense_vec_t r, r0;
    gpu_dense_vec_t v, p;
    gpu_dense_vec_t y, z;
    gpu_dense_vec_t s, t;
} gpu_solver_workspace_t;

/* out = alpha * A x + beta * out */
typedef void (*gpu_spmv_fun)(const void *mat, gpu_dense_vec_t *x, gpu_dense_vec_t *out, csr_data_t alpha, csr_data_t beta);
/* x = M^{-1} rhs */
typedef void (*gpu_precond_fun)(const void *precond, const gpu_dense_vec_t *rhs, gpu_dense_vec_t *x);

void gpu_blas_init(void); /* tools/gpu_bicgstab.c:30 */

/* tools/gpu_bicgstab.c:133-307.  *iters: in max iterations, out iterations;
 * *tol_error: in tolerance, out ||r||/||b||.  The workspace is allocated on
 * first use and reused across calls. */
void gpu_bicgstab(gpu_spmv_fun spmv, const void *mat, const gpu_dense_vec_t *rhs, gpu_dense_vec_t *x, int nrow, int ncol,
                  int local_col_beg, gpu_precond_fun psolve, const void *precond, gpu_solver_workspace_t *wsp, int *iters,
                  double *tol_error);

/* No counterpart in the reference.  For callers that cannot allocate device memory (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc): the same solve as gpu_bicgstab called with a
 * fresh (zeroed) workspace, with the workspace, scalars and reduction partials in a fixed device array of the
 * library instead of allocated memory (32 MiB per GPU, held until the process ends).  It calls no hipMalloc,
 * hipFree or allocating routine; the callbacks are the caller's.  Limit: ncol <= 262144 (2^18) rows; larger
 * sizes return GADE_EINVAL (-1) with the reason in gade_last_error() and nothing is launched.  Every call on one
 * GPU uses that GPU's one array and starts by clearing it: calls that overlap on one GPU (other threads or
 * streams) race.  Returns 0 on success. */
int gpu_bicgstab_noalloc(gpu_spmv_fun spmv, const void *mat, const gpu_dense_vec_t *rhs, gpu_dense_vec_t *x, int nrow,
                         int ncol, int local_col_beg, gpu_precond_fun psolve, const void *precond, int *iters,
                         double *tol_error);

#ifdef __cplusplus
}
#endif
#endif
