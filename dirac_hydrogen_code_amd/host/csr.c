/*
 * csr.c -- host CSR container and file I/O (drop-in for tools/csr.c entry points).
 *
 * Differences in method, not in results: rows are located by binary search
 * (std semantics of a sorted unique list, same answers as
 * tools/sorted_list.c:32-57), and csr_unblock_matrix computes the expanded
 * row pointers and the block-major map in closed form instead of by the
 * reference's two cursor passes.
 */
#include "csr.h"

#include <string.h>

#define CSR_FAIL(...)                 \
    do {                              \
        fprintf(stderr, __VA_ARGS__); \
        exit(1);                      \
    } while (0)

static void *xmalloc(size_t bytes)
{
    void *p = malloc(bytes ? bytes : 1);
    if (!p) CSR_FAIL("csr: out of memory (%zu bytes)\n", bytes);
    return p;
}

static void clear_comm(sparse_csr_t *sp)
{
    sp->row_cpu_dist = NULL;
    sp->comm_pattern = NULL;
    sp->comm_pattern_size = NULL;
    sp->n_comm_entries = NULL;
    sp->recv_ptr = NULL;
    sp->send_ptr = NULL;
    sp->send_vec = NULL;
    sp->send_type = NULL;
    sp->npart = 1;
    sp->row_beg = 0;
    sp->row_end = 0;
    sp->local_offset = 0;
}

void csr_allocate(sparse_csr_t *out, csr_index_t nrows, csr_index_t ncols, csr_index_t nnz)
{
    out->nnz = nnz;
    out->nrows = nrows;
    out->ncols = ncols;
    out->blk_nnz = 1;
    out->blk_dim = 1;
    out->is_link = 0;
    out->perm = NULL;
    out->Ap = calloc((size_t)nrows + 1, sizeof(csr_index_t));
    out->Ai = calloc(nnz ? (size_t)nnz : 1, sizeof(csr_index_t));
    out->Ax = calloc(nnz ? (size_t)nnz : 1, sizeof(csr_data_t));
    if (!out->Ap || !out->Ai || !out->Ax) CSR_FAIL("csr_allocate: out of memory\n");
    out->Ai_sub_map = NULL;
    clear_comm(out);
}

void csr_free(sparse_csr_t *sp)
{
    free(sp->perm);
    free(sp->Ap);
    free(sp->Ai);
    free(sp->Ai_sub_map);
    if (!sp->is_link) free(sp->Ax);
    free(sp->row_cpu_dist);
    sp->perm = NULL;
    sp->Ap = NULL;
    sp->Ai = NULL;
    sp->Ai_sub_map = NULL;
    sp->Ax = NULL;
    sp->nnz = sp->nrows = sp->ncols = 0;
    sp->blk_nnz = sp->blk_dim = 1;
    sp->is_link = 0;
    clear_comm(sp);
}

/* tools/csr.c:108-149: same nnz -> values only; otherwise a deep copy of the structure */
void csr_copy(sparse_csr_t *out, const sparse_csr_t *in)
{
    if (out->nnz == in->nnz && out->Ax) {
        memcpy(out->Ax, in->Ax, sizeof(csr_data_t) * (size_t)out->nnz * out->blk_nnz);
        return;
    }
    out->nnz = in->nnz;
    out->nrows = in->nrows;
    out->ncols = in->ncols;
    out->blk_nnz = in->blk_nnz;
    out->blk_dim = in->blk_dim;
    out->is_link = 0;
    out->Ap = xmalloc(sizeof(csr_index_t) * ((size_t)out->nrows + 1));
    memcpy(out->Ap, in->Ap, sizeof(csr_index_t) * ((size_t)out->nrows + 1));
    out->Ai = xmalloc(sizeof(csr_index_t) * (size_t)out->nnz);
    memcpy(out->Ai, in->Ai, sizeof(csr_index_t) * (size_t)out->nnz);
    out->Ax = xmalloc(sizeof(csr_data_t) * (size_t)out->nnz * out->blk_nnz);
    if (in->Ax) memcpy(out->Ax, in->Ax, sizeof(csr_data_t) * (size_t)out->nnz * out->blk_nnz);
    out->perm = NULL;
    if (in->perm) {
        out->perm = xmalloc(sizeof(csr_index_t) * (size_t)out->ncols);
        memcpy(out->perm, in->perm, sizeof(csr_index_t) * (size_t)out->ncols);
    }
    out->Ai_sub_map = NULL;
    clear_comm(out);
    out->row_beg = in->row_beg;
    out->row_end = in->row_end;
    out->local_offset = in->local_offset;
}

void csr_zero(sparse_csr_t *sp)
{
    memset(sp->Ax, 0, sizeof(csr_data_t) * (size_t)sp->nnz * sp->blk_nnz);
}

void csr_diag(sparse_csr_t *sp, csr_index_t dim)
{
    csr_allocate(sp, dim, dim, dim);
    for (csr_index_t i = 0; i < dim; i++) {
        sp->Ai[i] = i;
        sp->Ap[i] = i;
    }
    sp->Ap[dim] = dim;
    sp->row_end = dim;
}

csr_index_t csr_nnz(const sparse_csr_t *sp) { return sp->nnz * sp->blk_nnz; }
csr_index_t csr_nrows(const sparse_csr_t *sp) { return sp->nrows * sp->blk_dim; }
csr_index_t csr_ncols(const sparse_csr_t *sp) { return sp->ncols * sp->blk_dim; }
csr_index_t csr_ncolblocks(const sparse_csr_t *sp) { return sp->ncols; }
csr_index_t csr_nrowblocks(const sparse_csr_t *sp) { return sp->nrows; }
csr_index_t csr_local_rowoffset(const sparse_csr_t *sp) { return sp->local_offset; }

int csr_try_read(const char *fname, sparse_csr_t *sp)
{
    memset(sp, 0, sizeof(*sp));
    sp->blk_dim = sp->blk_nnz = 1;
    FILE *fd = fopen(fname, "rb");
    if (!fd) {
        fprintf(stderr, "cant open %s\n", fname);
        return 1;
    }
    int hdr[2];
    if (fread(hdr, sizeof(int), 2, fd) != 2 || hdr[0] < 0 || hdr[1] < 0) goto bad;
    sp->nrows = sp->ncols = hdr[0];
    sp->nnz = hdr[1];
    sp->npart = 1;
    sp->row_end = sp->nrows;
    sp->Ap = xmalloc(sizeof(csr_index_t) * ((size_t)sp->nrows + 1));
    sp->Ai = xmalloc(sizeof(csr_index_t) * (size_t)sp->nnz);
    sp->Ax = xmalloc(sizeof(csr_data_t) * (size_t)sp->nnz);
    if (fread(sp->Ap, sizeof(csr_index_t), (size_t)sp->nrows + 1, fd) != (size_t)sp->nrows + 1) goto bad;
    if (sp->Ap[sp->nrows] != sp->nnz || sp->Ap[0] != 0) goto bad;
    for (csr_index_t r = 0; r < sp->nrows; r++)
        if (sp->Ap[r + 1] < sp->Ap[r]) goto bad;
    if (fread(sp->Ai, sizeof(csr_index_t), (size_t)sp->nnz, fd) != (size_t)sp->nnz) goto bad;
    for (csr_index_t k = 0; k < sp->nnz; k++)
        if (sp->Ai[k] < 0 || sp->Ai[k] >= sp->ncols) goto bad;
    if (fread(sp->Ax, sizeof(csr_data_t), (size_t)sp->nnz, fd) != (size_t)sp->nnz) goto bad;
    sp->perm = xmalloc(sizeof(csr_index_t) * ((size_t)sp->ncols + 1));
    if (fread(sp->perm, sizeof(csr_index_t), (size_t)sp->ncols, fd) != (size_t)sp->ncols) {
        free(sp->perm);
        sp->perm = NULL;
    } else {
        if (fread(&sp->npart, sizeof(csr_index_t), 1, fd) != 1 || sp->npart < 1) goto bad;
        sp->row_cpu_dist = xmalloc(sizeof(csr_index_t) * ((size_t)sp->npart + 1));
        if (fread(sp->row_cpu_dist, sizeof(csr_index_t), (size_t)sp->npart + 1, fd) != (size_t)sp->npart + 1) goto bad;
    }
    fclose(fd);
    return 0;
bad:
    fprintf(stderr, "wrong file format in %s\n", fname);
    fclose(fd);
    csr_free(sp);
    return 2;
}

void csr_read(const char *fname, sparse_csr_t *sp)
{
    if (csr_try_read(fname, sp)) exit(1);
}

void csr_write(const char *fname, const sparse_csr_t *sp)
{
    FILE *fd = fopen(fname, "wb");
    if (!fd) CSR_FAIL("cant open %s\n", fname);
    int ok = fwrite(&sp->nrows, sizeof(csr_index_t), 1, fd) == 1 && fwrite(&sp->nnz, sizeof(csr_index_t), 1, fd) == 1 &&
             fwrite(sp->Ap, sizeof(csr_index_t), (size_t)csr_nrows(sp) + 1, fd) == (size_t)csr_nrows(sp) + 1 &&
             fwrite(sp->Ai, sizeof(csr_index_t), (size_t)sp->nnz, fd) == (size_t)sp->nnz &&
             fwrite(sp->Ax, sizeof(csr_data_t), (size_t)sp->nnz, fd) == (size_t)sp->nnz;
    fclose(fd);
    if (!ok) CSR_FAIL("cant write file %s\n", fname);
}

void csr_block_params(sparse_csr_t *sp, csr_index_t blk_dim, csr_index_t blk_nnz)
{
    sp->blk_dim = blk_dim;
    sp->blk_nnz = blk_nnz;
    if (!sp->is_link) free(sp->Ax);
    sp->Ax = NULL;
    sp->is_link = 0;
}

/* Expanded row (R, a) holds rowlen(R) * sublen(a) entries ordered by (block column, sub column):
 * Ap[R*b + a] = Ap_blk[R] * nnz_sub + rowlen(R) * Ap_sub[a]. */
void csr_unblock_matrix(sparse_csr_t *out, const sparse_csr_t *in, const sparse_csr_t *sp_blk)
{
    const csr_index_t nb = in->nrows, b = in->blk_dim, snz = sp_blk->Ap[b];
    long long total = (long long)in->Ap[nb] * snz;
    if (total > 2147483647LL) CSR_FAIL("csr_unblock_matrix: %lld non-zeros overflow csr_index_t\n", total);
    csr_allocate(out, csr_nrows(in), csr_ncols(in), (csr_index_t)total);
    out->row_beg = 0;
    out->row_end = out->nrows;
    for (csr_index_t R = 0; R < nb; R++) {
        const csr_index_t len = in->Ap[R + 1] - in->Ap[R];
        for (csr_index_t a = 0; a < b; a++) {
            const csr_index_t sl = sp_blk->Ap[a + 1] - sp_blk->Ap[a];
            csr_index_t pos = in->Ap[R] * snz + len * sp_blk->Ap[a];
            out->Ap[R * b + a] = pos;
            for (csr_index_t p = in->Ap[R]; p < in->Ap[R + 1]; p++)
                for (csr_index_t t = 0; t < sl; t++) out->Ai[pos++] = in->Ai[p] * b + sp_blk->Ai[sp_blk->Ap[a] + t];
        }
    }
    out->Ap[(size_t)nb * b] = (csr_index_t)total;
    out->Ai_sub_map = xmalloc(sizeof(csr_index_t) * (size_t)total);
    for (csr_index_t R = 0; R < nb; R++) {
        for (csr_index_t p = in->Ap[R]; p < in->Ap[R + 1]; p++) {
            const csr_index_t pos = p - in->Ap[R];
            for (csr_index_t a = 0; a < b; a++) {
                const csr_index_t sl = sp_blk->Ap[a + 1] - sp_blk->Ap[a];
                const csr_index_t dst = out->Ap[R * b + a] + pos * sl;
                for (csr_index_t t = 0; t < sl; t++)
                    out->Ai_sub_map[(size_t)p * snz + sp_blk->Ap[a] + t] = dst + t;
            }
        }
    }
}

void csr_unblock_comm_info(sparse_csr_t *out, const sparse_csr_t *in, int rank, int nranks)
{
    if (nranks != 1) CSR_FAIL("csr_unblock_comm_info: this build runs on one rank (got %d)\n", nranks);
    (void)rank;
    out->row_beg = in->row_beg * in->blk_dim;
    out->row_end = (in->row_end ? in->row_end : in->nrows) * in->blk_dim;
    out->local_offset = 0;
    out->npart = 1;
}

static csr_index_t find_slot(const sparse_csr_t *sp, csr_index_t row, csr_index_t col)
{
    csr_index_t lo = sp->Ap[row], hi = sp->Ap[row + 1];
    while (lo < hi) {
        csr_index_t mid = lo + (hi - lo) / 2;
        if (sp->Ai[mid] < col) lo = mid + 1;
        else hi = mid;
    }
    return lo;
}

csr_data_t csr_get_value(const sparse_csr_t *sp, csr_index_t row, csr_index_t col)
{
    csr_index_t cp = find_slot(sp, row, col);
    if (cp == sp->Ap[row + 1] || sp->Ai[cp] != col) return 0;
    return sp->Ax[cp];
}

void csr_set_value(const sparse_csr_t *sp, csr_index_t row, csr_index_t col, csr_data_t val)
{
    csr_index_t cp = find_slot(sp, row, col);
    if (cp == sp->Ap[row + 1] || sp->Ai[cp] != col)
        CSR_FAIL("cant set matrix value: (%d,%d) not present in CSR.", row, col);
    sp->Ax[cp] = val;
}

void csr_conj_transpose(sparse_csr_t *out, const sparse_csr_t *in)
{
    for (csr_index_t row = 0; row < in->nrows; row++)
        for (csr_index_t cp = in->Ap[row]; cp < in->Ap[row + 1]; cp++)
            csr_set_value(out, in->Ai[cp], row, conj(in->Ax[cp]));
}

void csr_spmv(csr_index_t row_l, csr_index_t row_u, const sparse_csr_t *sp, const csr_data_t *x, csr_data_t *result)
{
    for (csr_index_t i = row_l; i < row_u; i++) {
        csr_data_t acc = 0;
        for (csr_index_t j = sp->Ap[i]; j < sp->Ap[i + 1]; j++) acc += x[sp->Ai[j]] * sp->Ax[j];
        result[i] += acc;
    }
}

typedef struct {
    csr_index_t row, col, idx;
} coo_key_t;

static int coo_cmp(const void *a, const void *b)
{
    const coo_key_t *x = a, *y = b;
    if (x->row != y->row) return x->row < y->row ? -1 : 1;
    if (x->col != y->col) return x->col < y->col ? -1 : 1;
    return x->idx < y->idx ? -1 : (x->idx > y->idx);
}

/* duplicates are summed in input order, as tools/csr.c:1079-1089 does */
void csr_coo2csr(sparse_csr_t *sp, const csr_index_t *rowidx, const csr_index_t *colidx, const csr_data_t *val,
                 csr_index_t matrix_dim, csr_index_t nnz)
{
    coo_key_t *k = xmalloc(sizeof(coo_key_t) * (size_t)nnz);
    for (csr_index_t i = 0; i < nnz; i++) {
        if (rowidx[i] < 0 || rowidx[i] >= matrix_dim || colidx[i] < 0 || colidx[i] >= matrix_dim)
            CSR_FAIL("csr_coo2csr: entry %d out of range\n", i);
        k[i].row = rowidx[i];
        k[i].col = colidx[i];
        k[i].idx = i;
    }
    qsort(k, (size_t)nnz, sizeof(coo_key_t), coo_cmp);
    csr_index_t uniq = 0;
    for (csr_index_t i = 0; i < nnz; i++)
        if (i == 0 || k[i].row != k[i - 1].row || k[i].col != k[i - 1].col) uniq++;
    csr_allocate(sp, matrix_dim, matrix_dim, uniq);
    csr_index_t pos = -1;
    for (csr_index_t i = 0; i < nnz; i++) {
        if (i == 0 || k[i].row != k[i - 1].row || k[i].col != k[i - 1].col) {
            pos++;
            sp->Ai[pos] = k[i].col;
            sp->Ap[k[i].row + 1]++;
        }
        sp->Ax[pos] += val[k[i].idx];
    }
    for (csr_index_t r = 0; r < matrix_dim; r++) sp->Ap[r + 1] += sp->Ap[r];
    free(k);
}

void csr_init_communication(sparse_csr_t *sp, csr_data_t *px, int rank, int nranks)
{
    (void)sp; (void)px; (void)rank;
    if (nranks != 1) CSR_FAIL("csr_init_communication: this build runs on one rank (got %d)\n", nranks);
}

void csr_comm(const sparse_csr_t *sp, int rank, int nranks)
{
    (void)sp; (void)rank;
    if (nranks != 1) CSR_FAIL("csr_comm: this build runs on one rank (got %d)\n", nranks);
}

/* tools/csr.c:609-651 on one rank: the partition is the whole matrix, so the local part is a copy with
   row_beg 0 and row_end nrows and no halo (every column is local, none is removed).  More ranks: exit(1). */
void csr_get_partition(sparse_csr_t *out, const sparse_csr_t *sp, int rank, int nranks)
{
    if (nranks != 1 || rank != 0) CSR_FAIL("csr_get_partition: this build runs on one rank (got %d)\n", nranks);
    csr_allocate(out, sp->nrows, sp->ncols, sp->nnz);
    memcpy(out->Ap, sp->Ap, sizeof(csr_index_t) * ((size_t)sp->nrows + 1));
    memcpy(out->Ai, sp->Ai, sizeof(csr_index_t) * (size_t)sp->nnz);
    memcpy(out->Ax, sp->Ax, sizeof(csr_data_t) * (size_t)sp->nnz);
    if (sp->perm) {
        out->perm = xmalloc(sizeof(csr_index_t) * (size_t)sp->nrows);
        memcpy(out->perm, sp->perm, sizeof(csr_index_t) * (size_t)sp->nrows);
    }
    out->row_cpu_dist = xmalloc(2 * sizeof(csr_index_t));
    out->row_cpu_dist[0] = 0;
    out->row_cpu_dist[1] = sp->nrows;
    out->npart = 1;
    out->row_beg = 0;
    out->row_end = sp->nrows;
}

/* tools/csr.c:427-439 (its non-MPI message) */
void csr_print_comm_info(sparse_csr_t *sp, int rank, int nranks, int blkdim)
{
    (void)sp; (void)rank; (void)nranks; (void)blkdim;
    printf("no communication with 1 rank");
}

/* tools/csr.c:215-225: sp_blk becomes a view of block (row, col) of the blocked matrix sp */
void csr_block_link(sparse_csr_t *sp_blk, sparse_csr_t *sp, csr_index_t row, csr_index_t col)
{
    const csr_index_t cp = find_slot(sp, row, col);
    if (cp == sp->Ap[row + 1] || sp->Ai[cp] != col)
        CSR_FAIL("cant insert block: (%d,%d) not present in CSR.", row, col);
    sp_blk->is_link = 1;
    sp_blk->Ax = sp->Ax + (size_t)cp * sp->blk_nnz;
    sp_blk->nrows = sp->blk_dim;
    sp_blk->ncols = sp->blk_dim;
    sp_blk->nnz = sp->blk_nnz;
}
