/*
 * solver.h -- setup of the CN operators (drop-in for tools/solver.h:32-58), in libgade_host.so.
 *
 * compute_stationary_matrices fills S and Hst = i h dt/2 H0 (host, once).
 * compute_preconditioner factors P = S + Hst exactly, block by block, on the
 * device and returns a handle of this library's own type gade_precond_t (the
 * reference returns SuperLU's ILU factors as slu_LU_t, a name this header
 * leaves to SuperLU); precond_lu_solve applies it.  compute_timedep_matrices
 * assembles Hfull on the device and copies the values back into Hfull.Ax.
 * matfree_spmv_S / matfree_spmv_H_init / matfree_spmv_H are the CPU
 * matrix-free products of tools/solver.c:486-759.
 */
#ifndef GADE_SOLVER_H
#define GADE_SOLVER_H

#include "csr.h"
#include "gpu_sparse.h"
#include "solver_common.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct gade_blocklu gade_blocklu_t;
typedef struct {
    gade_blocklu_t *lu;  /* exact block LU of P on the device */
    int blkdim, nblk;
} gade_precond_t;  /* the reference's compute_preconditioner returns slu_LU_t (tools/superlu.h:41) */

void compute_timedep_matrices(double h, double dt, sparse_csr_t *submatrix, csr_data_t *ft, int lmax,
                              sparse_csr_t *Hfull_blk, sparse_csr_t *Hfull, const sparse_csr_t *h0, const sparse_csr_t *H,
                              const sparse_csr_t *g, const sparse_csr_t *gt);              /* tools/solver.c:196-407 */
void compute_stationary_matrices(double h, double dt, sparse_csr_t *submatrix, sparse_csr_t *S_blk, sparse_csr_t *S,
                                 sparse_csr_t *Hst_blk, sparse_csr_t *Hst, const sparse_csr_t *h0,
                                 const sparse_csr_t *s0);                                   /* tools/solver.c:409-458 */
gade_precond_t compute_preconditioner(const sparse_csr_t *S, const sparse_csr_t *Hst);            /* tools/solver.c:460-485 */
/* x = P^{-1} rhs on device vectors: the gpu_precond_fun role of tools/dirac_solver.c:158-165 */
void precond_lu_solve(const gade_precond_t *lu, const gpu_dense_vec_t *rhs, gpu_dense_vec_t *x);

/* result += S x over rows [row_l, row_u) (whole blocks), computed block by block from s0..s2 (tools/solver.c:491-541) */
void matfree_spmv_S(csr_index_t row_l, csr_index_t row_u, const sparse_csr_t *S_blk, const csr_data_t *x, csr_data_t *result,
                    double h, double dt, const sparse_csr_t *s0, sparse_csr_t *submatrix);
/* G-sums of f(t) for matfree_spmv_H, kept by the library (tools/solver.c:543-630) */
void matfree_spmv_H_init(const sparse_csr_t *Hfull_blk, double h, double dt, sparse_csr_t *submatrix, csr_data_t *ft,
                         int lmax, const sparse_csr_t *g, const sparse_csr_t *gt);
/* result += Hfull(t) x over rows [row_l, row_u) without assembly (tools/solver.c:632-759) */
void matfree_spmv_H(csr_index_t row_l, csr_index_t row_u, const sparse_csr_t *Hfull_blk, const csr_data_t *x,
                    csr_data_t *result, double h, double dt, sparse_csr_t *submatrix, csr_data_t *ft, int lmax,
                    const sparse_csr_t *h0, const sparse_csr_t *H, const sparse_csr_t *g, const sparse_csr_t *gt);

/* Generic block-diagonal banded LU (no reference counterpart; used by
 * compute_preconditioner): P given as unblocked CSR with nblk diagonal blocks
 * of dimension blkdim; identical blocks are factored once. */
int gade_blocklu_create(int nblk, int blkdim, const int *Ap, const int *Ai, const double *Ax, gade_blocklu_t **out);
int gade_blocklu_solve(const gade_blocklu_t *lu, const double *b_dev, double *x_dev);
int gade_blocklu_destroy(gade_blocklu_t *lu);

#ifdef __cplusplus
}
#endif
#endif
