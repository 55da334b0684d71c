/* CPU drop-in check of libgade_host.so, driven the way tools/matrix_tool.c drives the reference's
   matrix-free products: reads a coupling directory with csr_read, blocks Hall with csr_block_params,
   forms gt with csr_conj_transpose, then writes to <out>:
     x (seeded), matfree_spmv_H(x) after matfree_spmv_H_init(f(t)), matfree_spmv_S(x),
     and the BiCGSTAB solution of (S + H) z = b for b = S x - H x, the CPU bicgstab with the matrix-free
     operator and a dense block-diagonal LU of S + i dt/2 H0 as preconditioner, with its iterations and
     tol_error.
   usage: matfree_check <dir> <lmax> <dt> <E0> <omega> <cycles> <t> <tol> <out> */
#include <complex.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "bicgstab.h"
#include "csr.h"
#include "solver.h"
#include "solver_common.h"

typedef struct {
    sparse_csr_t *Hblk, *Sblk, *sub, *h0, *H, *g, *gt, *s0;
    csr_data_t *ft;
    int lmax;
    double dt;
} op_t;

static void HS_matfree(const void *mat, cdouble_t *x, cdouble_t *out)
{
    const op_t *o = mat;
    const int n = csr_nrows(o->Hblk);
    matfree_spmv_H(0, n, o->Hblk, x, out, 1.0, o->dt, o->sub, o->ft, o->lmax, o->h0, o->H, o->g, o->gt);
    matfree_spmv_S(0, n, o->Sblk, x, out, 1.0, o->dt, o->s0, o->sub);
}

/* block-diagonal preconditioner P_r = S(k_r) + i dt/2 H0(k_r): dense LU (partial pivoting) per block */
typedef struct {
    int nblk, b;
    csr_data_t *lu;
    int *piv;
} bdiag_t;

static void bdiag_build(bdiag_t *P, const sparse_csr_t *sub, const sparse_csr_t *s0, const sparse_csr_t *h0,
                        const int *perm, int nblk, double dt)
{
    const int b = sub->nrows;
    P->nblk = nblk;
    P->b = b;
    P->lu = calloc((size_t)nblk * b * b, sizeof(csr_data_t));
    P->piv = malloc(sizeof(int) * (size_t)nblk * b);
    for (int r = 0; r < nblk; r++) {
        const double k = ik(perm ? perm[r] : r);
        csr_data_t *A = P->lu + (size_t)r * b * b;
        int *pv = P->piv + (size_t)r * b;
        for (int i = 0; i < b; i++)
            for (int q = sub->Ap[i]; q < sub->Ap[i + 1]; q++)
                A[i * b + sub->Ai[q]] = s0[0].Ax[q] + s0[1].Ax[q] * k + s0[2].Ax[q] * k * k +
                                        I * dt / 2 * (h0[0].Ax[q] + h0[1].Ax[q] * k + h0[2].Ax[q] * k * k + h0[3].Ax[q] * k * k * k);
        for (int c = 0; c < b; c++) {
            int m = c;
            for (int i = c + 1; i < b; i++)
                if (cabs(A[i * b + c]) > cabs(A[m * b + c])) m = i;
            pv[c] = m;
            if (m != c)
                for (int j = 0; j < b; j++) {
                    csr_data_t t = A[c * b + j];
                    A[c * b + j] = A[m * b + j];
                    A[m * b + j] = t;
                }
            for (int i = c + 1; i < b; i++) {
                A[i * b + c] /= A[c * b + c];
                for (int j = c + 1; j < b; j++) A[i * b + j] -= A[i * b + c] * A[c * b + j];
            }
        }
    }
}

static void bdiag_solve(const void *p, const cdouble_t *rhs, cdouble_t *x)
{
    const bdiag_t *P = p;
    const int b = P->b;
    for (int r = 0; r < P->nblk; r++) {
        const csr_data_t *A = P->lu + (size_t)r * b * b;
        const int *pv = P->piv + (size_t)r * b;
        csr_data_t *y = x + (size_t)r * b;
        memcpy(y, rhs + (size_t)r * b, sizeof(csr_data_t) * b);
        for (int c = 0; c < b; c++) {
            csr_data_t t = y[c];
            y[c] = y[pv[c]];
            y[pv[c]] = t;
            for (int i = c + 1; i < b; i++) y[i] -= A[i * b + c] * y[c];
        }
        for (int i = b - 1; i >= 0; i--) {
            for (int j = i + 1; j < b; j++) y[i] -= A[i * b + j] * y[j];
            y[i] /= A[i * b + i];
        }
    }
}

int main(int argc, char **argv)
{
    if (argc != 10) return 2;
    const char *dir = argv[1];
    const int lmax = atoi(argv[2]);
    const double dt = atof(argv[3]), E0 = atof(argv[4]), omega = atof(argv[5]), cycles = atof(argv[6]);
    const double t = atof(argv[7]), tol = atof(argv[8]);
    char fn[4096];
    sparse_csr_t Hall, Sdiag, *H = calloc(2 * lmax, sizeof(sparse_csr_t)), *g = calloc(24 * lmax, sizeof(sparse_csr_t));
    sparse_csr_t *gt = calloc(24 * lmax, sizeof(sparse_csr_t)), h0[4], s0[3];
    snprintf(fn, sizeof fn, "%s/H.csr", dir);
    csr_read(fn, &Hall);
    for (int l = 0; l < lmax; l++) {
        snprintf(fn, sizeof fn, "%s/H0l%d.csr", dir, l);
        csr_read(fn, H + 2 * l);
        snprintf(fn, sizeof fn, "%s/H1l%d.csr", dir, l);
        csr_read(fn, H + 2 * l + 1);
    }
    for (int a = 0; a < 6; a++)
        for (int l = 0; l < lmax; l++)
            for (int k = 0; k < 4; k++) {
                snprintf(fn, sizeof fn, "%s/g%da%dl%d.csr", dir, k, a, l);
                csr_read(fn, g + a * 4 * lmax + l * 4 + k);
                csr_copy(gt + a * 4 * lmax + l * 4 + k, g + a * 4 * lmax + l * 4 + k); /* as tools/dirac_solver.c:321-322 */
                csr_conj_transpose(gt + a * 4 * lmax + l * 4 + k, g + a * 4 * lmax + l * 4 + k);
            }
    for (int k = 0; k < 4; k++) {
        snprintf(fn, sizeof fn, "%s/h0%d.csr", dir, k);
        csr_read(fn, h0 + k);
    }
    for (int k = 0; k < 3; k++) {
        snprintf(fn, sizeof fn, "%s/s%d.csr", dir, k);
        csr_read(fn, s0 + k);
    }
    /* S is block diagonal: the identity pattern on the Hall blocks (tools/dirac_solver.c:243-248) */
    csr_diag(&Sdiag, Hall.nrows);
    Sdiag.perm = malloc(sizeof(int) * Hall.nrows);
    memcpy(Sdiag.perm, Hall.perm, sizeof(int) * Hall.nrows);
    const int b = g[0].nrows;
    csr_block_params(&Hall, b, csr_nnz(g));
    csr_block_params(&Sdiag, b, csr_nnz(g));
    const int n = csr_nrows(&Hall);
    beyondDipolePulse_t pulse;
    csr_data_t ft[6];
    beyondDipolePulse_init(&pulse, E0, omega, cycles);
    beoyndDipolePulse_axialPart(&pulse, t, ft);
    csr_data_t *x = malloc(sizeof(csr_data_t) * n), *yH = calloc(n, sizeof(csr_data_t)), *yS = calloc(n, sizeof(csr_data_t));
    unsigned long long st = 0x9E3779B97F4A7C15ull;
    for (int i = 0; i < n; i++) {
        double re, im;
        st = st * 6364136223846793005ull + 1442695040888963407ull;
        re = (double)(st >> 11) / 9007199254740992.0 - 0.5;
        st = st * 6364136223846793005ull + 1442695040888963407ull;
        im = (double)(st >> 11) / 9007199254740992.0 - 0.5;
        x[i] = re + I * im;
    }
    matfree_spmv_H_init(&Hall, 1.0, dt, g, ft, lmax, g, gt);
    matfree_spmv_H(0, n, &Hall, x, yH, 1.0, dt, g, ft, lmax, h0, H, g, gt);
    matfree_spmv_S(0, n, &Sdiag, x, yS, 1.0, dt, s0, g);
    /* rhs of the CN step from x: (S - H) x */
    csr_data_t *rhs = malloc(sizeof(csr_data_t) * n), *z = malloc(sizeof(csr_data_t) * n);
    for (int i = 0; i < n; i++) {
        rhs[i] = yS[i] - yH[i];
        z[i] = x[i];
    }
    op_t op = {&Hall, &Sdiag, g, h0, H, g, gt, s0, ft, lmax, dt};
    solver_workspace_t wsp;
    memset(&wsp, 0, sizeof wsp);
    bdiag_t P;
    bdiag_build(&P, g, s0, h0, Hall.perm, Hall.nrows, dt);
    int iters = 2000;
    double terr = tol;
    bicgstab(HS_matfree, &op, rhs, z, n, n, 0, bdiag_solve, &P, &wsp, &iters, &terr);
    FILE *fo = fopen(argv[9], "wb");
    if (!fo) return 3;
    fwrite(&n, sizeof(int), 1, fo);
    fwrite(&iters, sizeof(int), 1, fo);
    fwrite(&terr, sizeof(double), 1, fo);
    fwrite(x, sizeof(csr_data_t), n, fo);
    fwrite(yH, sizeof(csr_data_t), n, fo);
    fwrite(yS, sizeof(csr_data_t), n, fo);
    fwrite(z, sizeof(csr_data_t), n, fo);
    fclose(fo);
    /* everything this program owns goes back, so a leak-checking build reports only the library's own */
    void *wv[] = {wsp.r, wsp.r0, wsp.v, wsp.p, wsp.y, wsp.z, wsp.s, wsp.t};
    for (int i = 0; i < 8; i++) free(wv[i]);
    free(P.lu);
    free(P.piv);
    free(x);
    free(yH);
    free(yS);
    free(rhs);
    free(z);
    csr_free(&Hall);
    csr_free(&Sdiag);
    for (int i = 0; i < 2 * lmax; i++) csr_free(H + i);
    for (int i = 0; i < 24 * lmax; i++) {
        csr_free(g + i);
        csr_free(gt + i);
    }
    for (int k = 0; k < 4; k++) csr_free(h0 + k);
    for (int k = 0; k < 3; k++) csr_free(s0 + k);
    free(H);
    free(g);
    free(gt);
    return 0;
}
