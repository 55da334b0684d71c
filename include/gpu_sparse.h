/*
 * gpu_sparse.h -- device CSR matrices and vectors (drop-in for tools/gpu_sparse.h:35-77).
 *
 * Same type and function names as the reference.  The hipSPARSE descriptor
 * fields are kept as opaque pointers owned by libgade: SpMV and the
 * triangular solves are libgade's own kernels.  Failures print a message and
 * exit(1), as the reference's CHECK_GPU* macros do (tools/hip_translate.h:101-126).
 * One rank: local_offset is 0 and the local part is the whole vector.
 */
#ifndef GADE_GPU_SPARSE_H
#define GADE_GPU_SPARSE_H

#include "csr.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef csr_data_t gpu_complex_t;

typedef struct {
    csr_index_t nrows, ncols, nnz;
    csr_index_t *Ap;         /* device */
    csr_index_t *Ai;         /* device */
    csr_data_t *Ax;          /* device */
    csr_index_t *Ai_sub_map; /* device, when the host matrix has one */
    void *desc;              /* libgade triangular-solve schedule (gpu_lu_analyze) */
    void *spsvDescr;         /* unused */
    void *cuBuffer;          /* unused */
} gpu_sparse_csr_t;

typedef struct {
    void *desc;
    void *desc_local;
    csr_index_t dim, local_dim;
    csr_index_t local_offset;
    gpu_complex_t *x;        /* device */
} gpu_dense_vec_t;

typedef struct {
    gpu_sparse_csr_t *L, *U;
    gpu_dense_vec_t *temp;
} gpu_lu_t;

void gpu_sparse_init(void);                                                       /* tools/gpu_sparse.c:33 */
void gpu_sparse_fini(void);                                                       /* tools/gpu_sparse.c:52 */
void gpu_put_csr(gpu_sparse_csr_t *Agpu, const sparse_csr_t *Ahost);             /* tools/gpu_sparse.c:60-109 */
void gpu_free_csr(gpu_sparse_csr_t *Agpu);                                        /* tools/gpu_sparse.c:111 */
void gpu_put_vec(gpu_dense_vec_t *xgpu, const csr_data_t *xhost, csr_index_t dim);/* tools/gpu_sparse.c:120-136 */
void gpu_vec_local_part(gpu_dense_vec_t *xgpu, csr_index_t dim, csr_index_t local_offset); /* tools/gpu_sparse.c:138 */
void gpu_get_vec(csr_data_t *xhost, const gpu_dense_vec_t *xgpu);                 /* tools/gpu_sparse.c:145-152 */
void gpu_free_vec(gpu_dense_vec_t *xgpu);                                         /* tools/gpu_sparse.c:154 */
/* y = alpha A x + beta y  (tools/gpu_sparse.c:267-288) */
void gpu_spmv(gpu_sparse_csr_t *Hfull, gpu_dense_vec_t *x, gpu_dense_vec_t *y, csr_data_t alpha, csr_data_t beta);
/* same on the local part of x (tools/gpu_sparse.c:292-313) */
void gpu_spmv_local(gpu_sparse_csr_t *Hfull, gpu_dense_vec_t *x, gpu_dense_vec_t *y, csr_data_t alpha, csr_data_t beta);
/* test helpers of tools/gpu_sparse.c:315-388 */
void gpu_spmv_block(sparse_csr_t H_blk, csr_data_t *x, csr_data_t *y, sparse_csr_t *g);
void gpu_spmb_block_test(sparse_csr_t H_blk, csr_data_t *x, csr_data_t *yfull, sparse_csr_t *g);
void gpu_spmv_test(sparse_csr_t Hfull, csr_data_t *x, csr_data_t *yfull);
/* level schedules for L (lower) and U (upper), then y = U^{-1} L^{-1} x (tools/gpu_sparse.c:162-263) */
void gpu_lu_analyze(gpu_sparse_csr_t *L, gpu_sparse_csr_t *U, gpu_dense_vec_t *x, gpu_dense_vec_t *y, gpu_dense_vec_t *temp);
void gpu_lu_solve(const gpu_sparse_csr_t *L, const gpu_sparse_csr_t *U, const gpu_dense_vec_t *x, gpu_dense_vec_t *y,
                  gpu_dense_vec_t *temp);

#ifdef __cplusplus
}
#endif
#endif
