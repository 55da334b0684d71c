/*
 * bicgstab.c -- BiCGSTAB on host vectors (drop-in for tools/bicgstab.c:125-295).
 *
 * The recurrences, stopping rule and restart rule are the Eigen-style ones of the reference (they are the
 * algorithm, and iteration counts and tol_error are part of its output): r = b - A x, r0 = r; stop when
 * |r|^2 < tol^2 |b|^2, tested after the full update; restart with r0 = r = b - A x when
 * |rho| < 1e-24 |r0|^2; iterations counted break-before-increment.  The BLAS-1 steps are plain loops of
 * this file.  On a restart the residual is recomputed into a zeroed vector, as tools/gpu_bicgstab.c:229-238
 * does (tools/bicgstab.c:217 accumulates A x into the stale r).
 */
#include <complex.h>
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "bicgstab.h"

static cdouble_t *ws_vec(void **slot, int n)
{
    if (!*slot) {
        *slot = calloc(n > 0 ? (size_t)n : 1, sizeof(cdouble_t));
        if (!*slot) {
            fprintf(stderr, "bicgstab: out of memory\n");
            exit(1);
        }
    }
    return (cdouble_t *)*slot;
}

static cdouble_t dotc(int n, const cdouble_t *a, const cdouble_t *b)
{
    cdouble_t s = 0;
    for (int i = 0; i < n; i++) s += conj(a[i]) * b[i];
    return s;
}

static double sqnorm(int n, const cdouble_t *a) { return creal(dotc(n, a, a)); }

void bicgstab(spmv_fun spmv, const void *mat, const cdouble_t *rhs, cdouble_t *x, int nrow, int ncol, int local_col_beg,
              precond_fun psolve, const void *precond, solver_workspace_t *wsp, int *iters, double *tol_error)
{
    const double tol = *tol_error;
    const int maxit = *iters;
    cdouble_t *v = ws_vec(&wsp->v, nrow), *p = ws_vec(&wsp->p, nrow);
    cdouble_t *y = ws_vec(&wsp->y, ncol), *z = ws_vec(&wsp->z, ncol);
    cdouble_t *s = ws_vec(&wsp->s, nrow), *t = ws_vec(&wsp->t, nrow);
    cdouble_t *r = ws_vec(&wsp->r, nrow), *r0 = ws_vec(&wsp->r0, nrow);
    cdouble_t *xl = x + local_col_beg, *yl = y + local_col_beg, *zl = z + local_col_beg;
    cdouble_t rho = 1, alpha = 1, w = 1;

    memset(r, 0, sizeof(cdouble_t) * (size_t)nrow);
    spmv(mat, x, r);
    for (int i = 0; i < nrow; i++) {
        r[i] = rhs[i] - r[i];
        r0[i] = r[i];
    }
    double r0_sq = sqnorm(nrow, r0);
    const double rhs_sq = sqnorm(nrow, rhs);
    if (rhs_sq == 0) {
        memset(xl, 0, sizeof(cdouble_t) * (size_t)nrow);
        *iters = 0;
        *tol_error = 0;
        return;
    }
    const double tol2 = tol * tol * rhs_sq, eps2 = 1e-12 * 1e-12;
    if (r0_sq < tol2) {
        *iters = 0;
        *tol_error = sqrt(r0_sq / rhs_sq);
        return;
    }
    int it = 0;
    while (it < maxit) {
        const cdouble_t rho_old = rho;
        rho = dotc(nrow, r0, r);
        if (cabs(rho) < eps2 * r0_sq) {
            memset(r, 0, sizeof(cdouble_t) * (size_t)nrow);
            spmv(mat, x, r);
            for (int i = 0; i < nrow; i++) {
                r[i] = rhs[i] - r[i];
                r0[i] = r[i];
            }
            rho = r0_sq = sqnorm(nrow, r);
        }
        const cdouble_t beta = (rho / rho_old) * (alpha / w);
        for (int i = 0; i < nrow; i++) p[i] = r[i] + beta * (p[i] - w * v[i]);
        psolve(precond, p, yl);
        memset(v, 0, sizeof(cdouble_t) * (size_t)nrow);
        spmv(mat, y, v);
        alpha = rho / dotc(nrow, r0, v);
        for (int i = 0; i < nrow; i++) s[i] = r[i] - alpha * v[i];
        psolve(precond, s, zl);
        memset(t, 0, sizeof(cdouble_t) * (size_t)nrow);
        spmv(mat, z, t);
        const double tt = sqnorm(nrow, t);
        w = tt > 0 ? dotc(nrow, t, s) / tt : 0;
        for (int i = 0; i < nrow; i++) {
            zl[i] = alpha * yl[i] + w * zl[i];
            xl[i] += zl[i];
            r[i] = s[i] - w * t[i];
        }
        if (sqnorm(nrow, r) < tol2) break;
        ++it;
    }
    *tol_error = sqrt(sqnorm(nrow, r) / rhs_sq);
    *iters = it;
}
