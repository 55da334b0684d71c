/*
 * bicgstab.h -- right-preconditioned BiCGSTAB on host vectors (drop-in for tools/bicgstab.h:24-38).
 *
 * Same workspace type, callback types and argument meanings as the reference: spmv(mat, x, out)
 * accumulates A x into out (the caller's out is zeroed where the algorithm needs it), psolve writes
 * x = M^{-1} rhs; *iters is the maximum on input and the iteration count on output (break before the
 * increment), *tol_error the relative tolerance on input and ||r||/||b|| on output.  The workspace
 * vectors are allocated on first use and reused by later calls.
 */
#ifndef GADE_BICGSTAB_H
#define GADE_BICGSTAB_H

#include "types.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct {
    void *r, *r0;
    void *v, *p;
    void *y, *z;
    void *s, *t;
} solver_workspace_t;

typedef void (*spmv_fun)(const void *mat, cdouble_t *x, cdouble_t *out);
typedef void (*precond_fun)(const void *precond, const cdouble_t *rhs, cdouble_t *x);

void bicgstab(spmv_fun spmv, const void *mat, const cdouble_t *rhs, cdouble_t *x, int nrow, int ncol, int local_col_beg,
              precond_fun psolve, const void *precond, solver_workspace_t *wsp, int *iters, double *tol_error);

#ifdef __cplusplus
}
#endif
#endif
