/*
 * csr.h -- host CSR container and file I/O of the GaDE hot path.
 *
 * Drop-in for tools/csr.h:32-90: same structure fields and function names.
 * The MPI halo fields are kept (as plain pointers) so that the reference
 * driver compiles against this header; on one GPU they stay NULL/zero.
 * Errors print a message and exit(1), as the reference's ERROR macro does
 * (tools/utils.h:42-46).
 */
#ifndef GADE_CSR_H
#define GADE_CSR_H

#include <stdio.h>
#include <stdlib.h>
#include "types.h"

#ifdef __cplusplus
/* csr_data_t is std::complex<double> in C++: same layout and return convention as double complex */
#pragma clang diagnostic ignored "-Wreturn-type-c-linkage"
extern "C" {
#endif

typedef struct {
    csr_index_t nrows, ncols, nnz;
    csr_index_t blk_dim, blk_nnz;
    csr_index_t npart, row_beg, row_end, local_offset;
    csr_index_t *perm, *row_cpu_dist;
    csr_index_t **comm_pattern, *comm_pattern_size, *n_comm_entries;
    csr_data_t **recv_ptr, **send_ptr, *send_vec;
    void *send_type;
    csr_index_t *Ap;
    csr_index_t *Ai;
    csr_data_t *Ax;
    csr_index_t *Ai_sub_map;
    int is_link;
} sparse_csr_t;

void csr_allocate(sparse_csr_t *out, csr_index_t nrows, csr_index_t ncols, csr_index_t nnz); /* tools/csr.c:51 */
void csr_free(sparse_csr_t *sp);                                                           /* tools/csr.c:79 */
void csr_copy(sparse_csr_t *out, const sparse_csr_t *in);                                  /* tools/csr.c:108 */
void csr_zero(sparse_csr_t *sp);                                                           /* tools/csr.c:257 */
void csr_diag(sparse_csr_t *sp, csr_index_t dim);                                          /* tools/csr.c:151 */

csr_index_t csr_nnz(const sparse_csr_t *sp_blk);          /* tools/csr.c:252 */
csr_index_t csr_nrows(const sparse_csr_t *sp_blk);        /* tools/csr.c:232 */
csr_index_t csr_ncols(const sparse_csr_t *sp_blk);        /* tools/csr.c:227 */
csr_index_t csr_ncolblocks(const sparse_csr_t *sp_blk);   /* tools/csr.c:237 */
csr_index_t csr_nrowblocks(const sparse_csr_t *sp_blk);   /* tools/csr.c:242 */
csr_index_t csr_local_rowoffset(const sparse_csr_t *sp_blk); /* tools/csr.c:247 */

void csr_read(const char *fname, sparse_csr_t *sp);       /* tools/csr.c:262-318 */
void csr_write(const char *fname, const sparse_csr_t *sp);/* tools/csr.c:320-342 */
/* Same as csr_read but returns nonzero instead of exiting (no reference counterpart). */
int csr_try_read(const char *fname, sparse_csr_t *sp);

void csr_block_params(sparse_csr_t *sp, csr_index_t blk_dim, csr_index_t blk_nnz);            /* tools/csr.c:163 */
void csr_unblock_matrix(sparse_csr_t *out, const sparse_csr_t *in, const sparse_csr_t *sp_blk); /* tools/csr.c:653-743 */
void csr_unblock_comm_info(sparse_csr_t *out, const sparse_csr_t *in, int rank, int nranks);  /* tools/csr.c:780; no-op on one rank */

csr_data_t csr_get_value(const sparse_csr_t *sp, csr_index_t row, csr_index_t col);          /* tools/csr.c:945-952 */
void csr_set_value(const sparse_csr_t *sp, csr_index_t row, csr_index_t col, csr_data_t val); /* tools/csr.c:954-961 */
void csr_conj_transpose(sparse_csr_t *out, const sparse_csr_t *in);                           /* tools/csr.c:963-972 */
void csr_spmv(csr_index_t row_l, csr_index_t row_u, const sparse_csr_t *sp, const csr_data_t *x, csr_data_t *result); /* tools/csr.c:974-1003 */
void csr_coo2csr(sparse_csr_t *out, const csr_index_t *rowidx, const csr_index_t *colidx, const csr_data_t *val,
                 csr_index_t matrix_dim, csr_index_t nnz);                                    /* tools/csr.c:1027-1122 */
/* one rank: the local partition is the whole matrix (tools/csr.c:609-651); exit(1) for more ranks */
void csr_get_partition(sparse_csr_t *out, const sparse_csr_t *sp, int rank, int nranks);
void csr_print_comm_info(sparse_csr_t *sp, int rank, int nranks, int blkdim);                   /* tools/csr.c:427-439 */
void csr_block_link(sparse_csr_t *sp_blk, sparse_csr_t *sp, csr_index_t row, csr_index_t col);  /* tools/csr.c:215-225 */
/* one rank: the halo exchange of tools/csr.c:858-943 has nothing to do */
void csr_init_communication(sparse_csr_t *sp, csr_data_t *px, int rank, int nranks);
void csr_comm(const sparse_csr_t *sp, int rank, int nranks);

#ifdef __cplusplus
}
#endif
#endif
