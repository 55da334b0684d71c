/*
 * matfree.c -- matrix-free CPU products of the CN operator (drop-in for tools/solver.h:46-58, the
 * functions tools/matrix_tool.c uses to check the assembled operator against).
 *
 *   matfree_spmv_S:      result += S x            over block rows [row_l/b, row_u/b)
 *   matfree_spmv_H_init: G-sums of f(t) for matfree_spmv_H (kept in this file, one model per process,
 *                        as the reference keeps them in file-static arrays)
 *   matfree_spmv_H:      result += Hfull(t) x     (stationary i h dt/2 H0 blocks included)
 *
 * The mathematics is the operator of tools/solver.c:327-395; the order of evaluation is this project's:
 * per row block r and multipole l the angular contractions come first,
 *   Z0 = sum_c H0l[r,c] x_c,  Z2 = sum_c H0l[r,c] k_c x_c,  W0 = sum_c H1l[r,c] x_c,  W1 = sum_c H1l[r,c] k_c x_c,
 * then one radial product per contraction with the kappa-folded G-sums,
 *   result_r += (Gs0 + k_r Gs1) Z0 + (Gs2 + k_r Gs3) Z2 + (Gts0 + k_r Gts2) W0 + (Gts1 + k_r Gts3) W1,
 * which is the order of the device kernels (4 radial products per (r, l) instead of 8 per (c, l)).
 * Row blocks are independent; they run on OpenMP threads when the library is built with OpenMP.
 * Rows are indexed absolutely (result[i] for row i), as csr_spmv does (tools/csr.c:974-1003).
 */
#include <stdlib.h>
#include <string.h>

#include "solver.h"

#define MF_FAIL(...)                  \
    do {                              \
        fprintf(stderr, __VA_ARGS__); \
        exit(1);                      \
    } while (0)

static struct {
    int lmax, nnzb, nblk;
    csr_data_t *gs;  /* [l*4 + n][nnzb]: SoL * i h dt/2 * sum_a f_a g^n_{a,l} */
    csr_data_t *gts; /* the same over gt */
    int *kap;        /* [nblk] kappa of each (compressed) block */
} mf;

/* tools/solver.c:543-630 */
void matfree_spmv_H_init(const sparse_csr_t *Hfull_blk, double h, double dt, sparse_csr_t *submatrix, csr_data_t *ft,
                         int lmax, const sparse_csr_t *g, const sparse_csr_t *gt)
{
    const int nz = csr_nnz(submatrix), nb = csr_nrowblocks(Hfull_blk);
    if (mf.lmax != lmax || mf.nnzb != nz) {
        free(mf.gs);
        free(mf.gts);
        mf.gs = malloc(sizeof(csr_data_t) * (size_t)lmax * 4 * (size_t)nz);
        mf.gts = malloc(sizeof(csr_data_t) * (size_t)lmax * 4 * (size_t)nz);
        if (!mf.gs || !mf.gts) MF_FAIL("matfree_spmv_H_init: out of memory\n");
        mf.lmax = lmax;
        mf.nnzb = nz;
    }
    if (mf.nblk != nb || !mf.kap) {
        free(mf.kap);
        mf.kap = malloc(sizeof(int) * (size_t)(nb ? nb : 1));
        if (!mf.kap) MF_FAIL("matfree_spmv_H_init: out of memory\n");
        mf.nblk = nb;
    }
    for (int c = 0; c < nb; c++) mf.kap[c] = ik(Hfull_blk->perm ? Hfull_blk->perm[c] : c);
    const csr_data_t ihdt = I * h * dt / 2;
    for (int l = 0; l < lmax; l++)
        for (int n = 0; n < 4; n++) {
            csr_data_t *ds = mf.gs + (size_t)(l * 4 + n) * nz, *dt_ = mf.gts + (size_t)(l * 4 + n) * nz;
            memset(ds, 0, sizeof(csr_data_t) * (size_t)nz);
            memset(dt_, 0, sizeof(csr_data_t) * (size_t)nz);
            for (int a = 0; a < 6; a++) {
                if (a % 2 == l % 2) continue; /* only alpha of the other parity couple to l */
                const csr_data_t c = SoL * ihdt * ft[a];
                const csr_data_t *pg = g[a * 4 * lmax + l * 4 + n].Ax, *pt = gt[a * 4 * lmax + l * 4 + n].Ax;
                for (int i = 0; i < nz; i++) {
                    ds[i] += c * pg[i];
                    dt_[i] += c * pt[i];
                }
            }
        }
}

/* y += (A1 + k A2) v over the shared radial pattern */
static void radial_pair(const sparse_csr_t *sub, const csr_data_t *A1, const csr_data_t *A2, double k,
                        const csr_data_t *v, csr_data_t *y)
{
    for (csr_index_t i = 0; i < sub->nrows; i++) {
        csr_data_t acc = 0;
        for (csr_index_t p = sub->Ap[i]; p < sub->Ap[i + 1]; p++) acc += (A1[p] + k * A2[p]) * v[sub->Ai[p]];
        y[i] += acc;
    }
}

/* tools/solver.c:632-759 */
void matfree_spmv_H(csr_index_t row_l, csr_index_t row_u, const sparse_csr_t *Hfull_blk, const csr_data_t *x,
                    csr_data_t *result, double h, double dt, sparse_csr_t *submatrix, csr_data_t *ft, int lmax,
                    const sparse_csr_t *h0, const sparse_csr_t *H, const sparse_csr_t *g, const sparse_csr_t *gt)
{
    (void)ft; (void)g; (void)gt;
    if (!mf.gs || mf.lmax != lmax || mf.nnzb != csr_nnz(submatrix) || mf.nblk != csr_nrowblocks(Hfull_blk))
        MF_FAIL("matfree_spmv_H: call matfree_spmv_H_init for this model first\n");
    const int b = csr_nrows(Hfull_blk) / csr_nrowblocks(Hfull_blk);
    const int nz = mf.nnzb;
    const csr_data_t ihdt = I * h * dt / 2;
    const int rb = row_l / b, re = row_u / b;
#pragma omp parallel
    {
        csr_data_t *Z = malloc(sizeof(csr_data_t) * 4 * (size_t)b);
        csr_data_t *hst = malloc(sizeof(csr_data_t) * (size_t)nz);
        if (!Z || !hst) MF_FAIL("matfree_spmv_H: out of memory\n");
#pragma omp for schedule(dynamic, 1)
        for (int r = rb; r < re; r++) {
            const int orow = Hfull_blk->perm ? Hfull_blk->perm[csr_local_rowoffset(Hfull_blk) + r] : r;
            const double kr = mf.kap[r];
            csr_data_t *y = result + (size_t)r * b;
            for (int l = 0; l < lmax; l++) {
                csr_data_t *Z0 = Z, *Z2 = Z + b, *W0 = Z + 2 * b, *W1 = Z + 3 * b;
                memset(Z, 0, sizeof(csr_data_t) * 4 * (size_t)b);
                int any0 = 0, any1 = 0;
                for (csr_index_t p = Hfull_blk->Ap[r]; p < Hfull_blk->Ap[r + 1]; p++) {
                    const int c = Hfull_blk->Ai[p];
                    const int ocol = Hfull_blk->perm ? Hfull_blk->perm[c] : c;
                    const csr_data_t a0 = csr_get_value(H + 2 * l, orow, ocol);
                    const csr_data_t a1 = csr_get_value(H + 2 * l + 1, orow, ocol);
                    const double kc = mf.kap[c];
                    const csr_data_t *xc = x + (size_t)c * b;
                    if (a0 != 0) {
                        any0 = 1;
                        for (int i = 0; i < b; i++) {
                            Z0[i] += a0 * xc[i];
                            Z2[i] += a0 * kc * xc[i];
                        }
                    }
                    if (a1 != 0) {
                        any1 = 1;
                        for (int i = 0; i < b; i++) {
                            W0[i] += a1 * xc[i];
                            W1[i] += a1 * kc * xc[i];
                        }
                    }
                }
                const csr_data_t *G = mf.gs + (size_t)l * 4 * nz, *Gt = mf.gts + (size_t)l * 4 * nz;
                if (any0) {
                    radial_pair(submatrix, G, G + nz, kr, Z0, y);             /* (Gs0 + k_r Gs1) Z0 */
                    radial_pair(submatrix, G + 2 * nz, G + 3 * nz, kr, Z2, y); /* (Gs2 + k_r Gs3) Z2 */
                }
                if (any1) {
                    radial_pair(submatrix, Gt, Gt + 2 * nz, kr, W0, y);        /* (Gts0 + k_r Gts2) W0 */
                    radial_pair(submatrix, Gt + nz, Gt + 3 * nz, kr, W1, y);   /* (Gts1 + k_r Gts3) W1 */
                }
            }
            /* stationary part on the diagonal block: i h dt/2 (h00 + k h01 + k^2 h02 + k^3 h03) */
            for (int i = 0; i < nz; i++)
                hst[i] = ihdt * (h0[0].Ax[i] + h0[1].Ax[i] * kr + h0[2].Ax[i] * kr * kr + h0[3].Ax[i] * kr * kr * kr);
            const csr_data_t *xr = x + (size_t)r * b;
            for (csr_index_t i = 0; i < submatrix->nrows; i++) {
                csr_data_t acc = 0;
                for (csr_index_t p = submatrix->Ap[i]; p < submatrix->Ap[i + 1]; p++) acc += hst[p] * xr[submatrix->Ai[p]];
                y[i] += acc;
            }
        }
        free(Z);
        free(hst);
    }
}

/* tools/solver.c:491-541 */
void matfree_spmv_S(csr_index_t row_l, csr_index_t row_u, const sparse_csr_t *S_blk, const csr_data_t *x,
                    csr_data_t *result, double h, double dt, const sparse_csr_t *s0, sparse_csr_t *submatrix)
{
    (void)h; (void)dt;
    const int b = csr_nrows(S_blk) / csr_nrowblocks(S_blk);
    const int rb = row_l / b, re = row_u / b;
#pragma omp parallel for schedule(static)
    for (int r = rb; r < re; r++) {
        const double k = ik(S_blk->perm ? S_blk->perm[r] : r);
        const csr_data_t *xr = x + (size_t)r * b;
        csr_data_t *y = result + (size_t)r * b;
        for (csr_index_t i = 0; i < submatrix->nrows; i++) {
            csr_data_t acc = 0;
            for (csr_index_t p = submatrix->Ap[i]; p < submatrix->Ap[i + 1]; p++)
                acc += (s0[0].Ax[p] + s0[1].Ax[p] * k + s0[2].Ax[p] * k * k) * xr[submatrix->Ai[p]];
            y[i] += acc;
        }
    }
}
