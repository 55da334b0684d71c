/*
 * gade_load.c -- reads the coupling files of one run directory
 * (tools/dirac_solver.c:223-342) and builds the device model.
 */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "csr.h"
#include "gade.h"

static int same_pattern(const sparse_csr_t *a, const sparse_csr_t *b)
{
    return a->nrows == b->nrows && a->nnz == b->nnz &&
           !memcmp(a->Ap, b->Ap, sizeof(csr_index_t) * ((size_t)a->nrows + 1)) &&
           !memcmp(a->Ai, b->Ai, sizeof(csr_index_t) * (size_t)a->nnz);
}

int gade_model_load(const char *dir, int lmax, double h, double dt, void *stream, gade_model_t **out)
{
    return gade_model_load_batch(dir, lmax, 1, h, dt, stream, out);
}

int gade_model_load_batch(const char *dir, int lmax, int batch, double h, double dt, void *stream, gade_model_t **out)
{
    char fname[4096];
    int rc = GADE_EIO, ng = 24 * lmax;
    if (!dir || lmax <= 0 || !out) return GADE_EINVAL;
    sparse_csr_t Hall;
    sparse_csr_t *H = calloc(2 * (size_t)lmax, sizeof(sparse_csr_t));
    sparse_csr_t *g = calloc((size_t)ng, sizeof(sparse_csr_t));
    sparse_csr_t h0[4], s0[3];
    memset(&Hall, 0, sizeof Hall);
    memset(h0, 0, sizeof h0);
    memset(s0, 0, sizeof s0);
    const int **HAp = calloc(2 * (size_t)lmax, sizeof(int *)), **HAi = calloc(2 * (size_t)lmax, sizeof(int *));
    const double **HAx = calloc(2 * (size_t)lmax, sizeof(double *)), **gx = calloc((size_t)ng, sizeof(double *));
    const double *hx[4], *sx[3];

    snprintf(fname, sizeof fname, "%s/H.csr", dir);
    if (csr_try_read(fname, &Hall)) goto done;
    for (int l = 0; l < lmax; l++) {
        snprintf(fname, sizeof fname, "%s/H0l%d.csr", dir, l);
        if (csr_try_read(fname, H + 2 * l)) goto done;
        snprintf(fname, sizeof fname, "%s/H1l%d.csr", dir, l);
        if (csr_try_read(fname, H + 2 * l + 1)) goto done;
    }
    for (int a = 0; a < 6; a++)
        for (int l = 0; l < lmax; l++)
            for (int n = 0; n < 4; n++) {
                snprintf(fname, sizeof fname, "%s/g%da%dl%d.csr", dir, n, a, l);
                if (csr_try_read(fname, g + a * 4 * lmax + l * 4 + n)) goto done;
            }
    for (int n = 0; n < 4; n++) {
        snprintf(fname, sizeof fname, "%s/h0%d.csr", dir, n);
        if (csr_try_read(fname, h0 + n)) goto done;
    }
    for (int n = 0; n < 3; n++) {
        snprintf(fname, sizeof fname, "%s/s%d.csr", dir, n);
        if (csr_try_read(fname, s0 + n)) goto done;
    }
    /* the assembly indexes every radial matrix by the same slot (tools/solver.c:329-335) */
    for (int i = 1; i < ng; i++)
        if (!same_pattern(g, g + i)) { fprintf(stderr, "gade: g matrices do not share one pattern\n"); rc = GADE_EINVAL; goto done; }
    for (int n = 0; n < 4; n++)
        if (!same_pattern(g, h0 + n)) { fprintf(stderr, "gade: h0%d pattern differs from g0a0l0\n", n); rc = GADE_EINVAL; goto done; }
    for (int n = 0; n < 3; n++)
        if (!same_pattern(g, s0 + n)) { fprintf(stderr, "gade: s%d pattern differs from g0a0l0\n", n); rc = GADE_EINVAL; goto done; }
    for (int i = 0; i < 2 * lmax; i++) {
        if (H[i].nrows != H[0].nrows) { fprintf(stderr, "gade: H0l/H1l dimensions differ\n"); rc = GADE_EINVAL; goto done; }
        HAp[i] = H[i].Ap;
        HAi[i] = H[i].Ai;
        HAx[i] = (const double *)H[i].Ax;
    }
    for (int i = 0; i < ng; i++) gx[i] = (const double *)g[i].Ax;
    for (int i = 0; i < 4; i++) hx[i] = (const double *)h0[i].Ax;
    for (int i = 0; i < 3; i++) sx[i] = (const double *)s0[i].Ax;
    {
        gade_inputs_t in;
        memset(&in, 0, sizeof in);
        in.nblk = Hall.nrows;
        in.hall_Ap = Hall.Ap;
        in.hall_Ai = Hall.Ai;
        in.hall_perm = Hall.perm;
        in.lmax = lmax;
        in.nang = H[0].nrows;
        in.H_Ap = HAp;
        in.H_Ai = HAi;
        in.H_Ax = HAx;
        in.blkdim = g[0].nrows;
        in.rad_Ap = g[0].Ap;
        in.rad_Ai = g[0].Ai;
        in.g = gx;
        in.h0 = hx;
        in.s = sx;
        rc = gade_model_create_batch(&in, batch, h, dt, stream, out);
    }
done:
    csr_free(&Hall);
    for (int i = 0; i < 2 * lmax; i++) csr_free(H + i);
    for (int i = 0; i < ng; i++) csr_free(g + i);
    for (int i = 0; i < 4; i++) csr_free(h0 + i);
    for (int i = 0; i < 3; i++) csr_free(s0 + i);
    free(H); free(g); free(HAp); free(HAi); free(HAx); free(gx);
    return rc;
}
