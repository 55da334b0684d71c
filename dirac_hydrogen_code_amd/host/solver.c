/*
 * solver.c -- setup of the CN operators (drop-in for tools/solver.c:196-485).
 *
 * compute_stationary_matrices: S_rr = s0 + k s1 + k^2 s2 and
 * Hst_rr = i h dt/2 (h00 + k h01 + k^2 h02 + k^3 h03) for each diagonal block
 * r, with k = kappa of the block's original angular index (tools/solver.c:409-458).
 * compute_preconditioner: exact LU of every distinct block of P = S + Hst on
 * the device (the reference hands P to SuperLU's ILU).
 * compute_timedep_matrices: Hfull assembled on the device through the
 * gpu_solver.h entry points, values copied back into Hfull.Ax.
 */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "gade.h"
#include "gpu_solver.h"
#include "solver.h"

#define SOLVER_FAIL(...)              \
    do {                              \
        fprintf(stderr, __VA_ARGS__); \
        exit(1);                      \
    } while (0)

/* values of the diagonal block (row, row) of an unblocked block-diagonal matrix,
 * in the order of the submatrix pattern */
static void put_block(sparse_csr_t *A, csr_index_t row, const sparse_csr_t *sub, const csr_data_t *vals)
{
    const csr_index_t b = sub->nrows;
    for (csr_index_t a = 0; a < b; a++) {
        const csr_index_t dst = row * b + a;
        const csr_index_t len = sub->Ap[a + 1] - sub->Ap[a];
        if (A->Ap[dst + 1] - A->Ap[dst] != len || (len && A->Ai[A->Ap[dst]] != row * b + sub->Ai[sub->Ap[a]]))
            SOLVER_FAIL("cant set matrix value: block (%d,%d) row %d not present in CSR.", row, row, a);
        memcpy(A->Ax + A->Ap[dst], vals + sub->Ap[a], sizeof(csr_data_t) * (size_t)len);
    }
}

void compute_stationary_matrices(double h, double dt, sparse_csr_t *submatrix, sparse_csr_t *S_blk, sparse_csr_t *S,
                                 sparse_csr_t *Hst_blk, sparse_csr_t *Hst, const sparse_csr_t *h0, const sparse_csr_t *s0)
{
    (void)Hst_blk;
    const csr_index_t nz = csr_nnz(submatrix);
    const csr_data_t ihdt = I * h * dt / 2;
    csr_data_t *sub = malloc(sizeof(csr_data_t) * (size_t)(nz ? nz : 1));
    if (!sub) SOLVER_FAIL("compute_stationary_matrices: out of memory\n");
    for (csr_index_t row = 0; row < csr_nrowblocks(S_blk); row++) {
        const int k = ik(S_blk->perm ? S_blk->perm[row] : row);
        for (csr_index_t i = 0; i < nz; i++) sub[i] = s0[0].Ax[i] + s0[1].Ax[i] * k + s0[2].Ax[i] * k * k;
        put_block(S, row, submatrix, sub);
        for (csr_index_t i = 0; i < nz; i++)
            sub[i] = ihdt * (h0[0].Ax[i] + h0[1].Ax[i] * k + h0[2].Ax[i] * k * k + h0[3].Ax[i] * k * k * k);
        put_block(Hst, row, submatrix, sub);
    }
    free(sub);
}

/* smallest block size b (dividing n) such that every entry lies in its diagonal block */
static csr_index_t infer_blkdim(const sparse_csr_t *A)
{
    const csr_index_t n = A->nrows;
    csr_index_t reach = 1;
    for (csr_index_t i = 0; i < n; i++)
        for (csr_index_t k = A->Ap[i]; k < A->Ap[i + 1]; k++) {
            csr_index_t d = A->Ai[k] > i ? A->Ai[k] - i : i - A->Ai[k];
            if (d + 1 > reach) reach = d + 1;
        }
    for (csr_index_t b = reach; b <= n; b++) {
        if (n % b) continue;
        int ok = 1;
        for (csr_index_t i = 0; i < n && ok; i++)
            for (csr_index_t k = A->Ap[i]; k < A->Ap[i + 1]; k++)
                if (A->Ai[k] / b != i / b) { ok = 0; break; }
        if (ok) return b;
    }
    return n;
}

gade_precond_t compute_preconditioner(const sparse_csr_t *S, const sparse_csr_t *Hst)
{
    gade_precond_t out = {0};
    if (S->nnz != Hst->nnz) SOLVER_FAIL("compute_preconditioner: S and Hst patterns differ\n");
    csr_data_t *P = malloc(sizeof(csr_data_t) * (size_t)(Hst->nnz ? Hst->nnz : 1));
    if (!P) SOLVER_FAIL("compute_preconditioner: out of memory\n");
    for (csr_index_t k = 0; k < Hst->nnz; k++) P[k] = S->Ax[k] + Hst->Ax[k];
    const csr_index_t b = infer_blkdim(Hst);
    out.blkdim = b;
    out.nblk = Hst->nrows / b;
    if (gade_blocklu_create(out.nblk, b, Hst->Ap, Hst->Ai, (const double *)P, &out.lu))
        SOLVER_FAIL("compute_preconditioner: %s\n", gade_last_error());
    free(P);
    return out;
}

void precond_lu_solve(const gade_precond_t *lu, const gpu_dense_vec_t *rhs, gpu_dense_vec_t *x)
{
    double *out = (double *)(x->desc_local ? x->x + x->local_offset : x->x);
    if (gade_blocklu_solve(lu->lu, (const double *)rhs->x, out)) SOLVER_FAIL("precond_lu_solve: %s\n", gade_last_error());
}

void compute_timedep_matrices(double h, double dt, sparse_csr_t *submatrix, csr_data_t *ft, int lmax,
                              sparse_csr_t *Hfull_blk, sparse_csr_t *Hfull, const sparse_csr_t *h0, const sparse_csr_t *H,
                              const sparse_csr_t *g, const sparse_csr_t *gt)
{
    static int ready = 0;
    static gpu_sparse_csr_t dev;
    (void)submatrix;
    if (!ready) {
        gpu_init_model_matrices(24 * lmax, g, gt, h0);
        gpu_compute_row_col(lmax, H, Hfull_blk, Hfull);
        memset(&dev, 0, sizeof dev);
        ready = 1;
    }
    gpu_compute_timedep_matrices(h, dt, ft, lmax, Hfull_blk, Hfull, &dev);
    if (gade_memcpy_d2h(Hfull->Ax, dev.Ax, sizeof(csr_data_t) * (size_t)csr_nnz(Hfull)))
        SOLVER_FAIL("compute_timedep_matrices: %s\n", gade_last_error());
}
