/*
 * dirac_solver.c -- Crank-Nicolson driver over libgade (tools/dirac_solver.c, GPU branch).
 *
 * Flags of the reference (tools/dirac_solver.c:194-214):
 *   -t maxtime  -l lmax  -i intensity(E0)  -o omega  -c cycles  -n Ntime  -s start_step
 * Additions: -e tol (default 1e-12; the reference hard-codes 1e-16), -m maxit
 * (default 500), -d dir (input directory, default "."), -w k (write x<k>.out
 * every k steps, default 2000 as the reference), -f (fused operator, no
 * per-step assembly; default is the reference's assembled formulation).
 *
 * The loop runs exactly Ntime - (start_step - 1) steps with time accumulated as
 * time += dt (the reference's `while (time <= maxtime)` runs Ntime or Ntime+1
 * steps depending on rounding).  Outputs: x<step>.out (int32 nrows, int32
 * ncols=1, complex128[], un-permuted), runstats.txt ("iters, tol, usec" per step).
 *
 * Checkpoint (-k file, with -f): written at every output with psi, the step, the accumulated time and the
 * solver state (gade_checkpoint_save); with -s k > 1 the run resumes from it (its step must be k - 1) and
 * continues bit for bit as the uninterrupted run (the reference's -s only shifts the start time and restarts
 * from psi0.vec, tools/dirac_solver.c:203-209).  Output windows end at the multiples of -w and at the last
 * step; the time is accumulated step by step across windows.
 *
 * Post-processing (-x file, then exit): the eigen-projection of a saved state, as src/process_output.cpp:
 * 320-355 and src/dumpEigvals.cpp:239-252 produce it: evl<kappa>.mat / evc<kappa>.mat for every kappa and
 * psiev_<file> with, per angular state th (kept states in original order), "evl[th] = npy.array((...))" and
 * "psiev[th] = npy.array((...))", psiev_j = psi_th^H S v_j (src/hamiltonian.h:1333-1364).
 */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <sys/time.h>
#include <unistd.h>

#include "csr.h"
#include "gade.h"
#include "gpu_bicgstab.h"
#include "gpu_solver.h"
#include "gpu_sparse.h"
#include "solver.h"
#include "solver_common.h"

#define FAIL(...)                     \
    do {                              \
        fprintf(stderr, __VA_ARGS__); \
        exit(1);                      \
    } while (0)

static double now_us(void)
{
    struct timeval tv;
    gettimeofday(&tv, NULL);
    return tv.tv_sec * 1e6 + tv.tv_usec;
}

static void vec_read(const char *fname, int n, csr_data_t *out, int *perm)
{
    FILE *fd = fopen(fname, "rb");
    if (!fd) FAIL("cant open %s\n", fname);
    int nn;
    if (fread(&nn, sizeof(int), 1, fd) != 1 || nn != n) FAIL("wrong file format in %s\n", fname);
    if (fread(out, sizeof(csr_data_t), (size_t)n, fd) != (size_t)n) FAIL("wrong file format in %s\n", fname);
    if (perm && fread(perm, sizeof(int), (size_t)n, fd) != (size_t)n) FAIL("wrong file format in %s\n", fname);
    fclose(fd);
}

/* src/hamiltonian.h:1338: Eigen IOFormat(FullPrecision, DontAlignCols, ", ", "\n", "(", "),", " = npy.array((\n",
   "\n))\n"); complex entries print as (re,im) */
static void write_psiev(const char *fname, gade_model_t *m, const double *xdev, const int *hall_perm, int nblk, int n)
{
    double *psiev = malloc(sizeof(double) * 2 * (size_t)nblk * n);
    int nk = 0, kap[256];
    gade_model_info(m, NULL, NULL, NULL, NULL, NULL, &nk);
    double *E = malloc(sizeof(double) * (size_t)nk * n);
    if (nk > 256 || gade_psiev(m, xdev, psiev) || gade_eig_values(m, &nk, kap, E))
        FAIL("eigen-projection: %s\n", gade_last_error());
    /* blocks in the order of their original angular index */
    int *order = malloc(sizeof(int) * (size_t)nblk);
    for (int r = 0; r < nblk; r++) order[r] = r;
    for (int a = 1; a < nblk; a++)
        for (int b = a; b > 0 && hall_perm && hall_perm[order[b - 1]] > hall_perm[order[b]]; b--) {
            int t = order[b]; order[b] = order[b - 1]; order[b - 1] = t;
        }
    FILE *fd = fopen(fname, "w");
    if (!fd) FAIL("cant open %s\n", fname);
    for (int th = 0; th < nblk; th++) {
        const int r = order[th];
        const int k = ik(hall_perm ? hall_perm[r] : r);
        int q = 0;
        while (q < nk && kap[q] != k) q++;
        if (q == nk) FAIL("eigen-projection: kappa %d of block %d has no eigenpairs\n", k, r);
        fprintf(fd, "evl[%d] = npy.array((\n", th);
        for (int j = 0; j < n; j++) fprintf(fd, "%s(%.17g),", j ? "\n" : "", E[(size_t)q * n + j]);
        fprintf(fd, "\n))\n\n");
        fprintf(fd, "psiev[%d] = npy.array((\n(", th);
        for (int j = 0; j < n; j++) {
            const double *c = psiev + 2 * ((size_t)r * n + j);
            fprintf(fd, "%s(%.17g,%.17g)", j ? ", " : "", c[0], c[1]);
        }
        fprintf(fd, "),\n))\n\n");
    }
    fclose(fd);
    free(order);
    free(E);
    free(psiev);
    printf("saved %s\n", fname);
}

static void write_out(const char *dir, int step, const csr_data_t *x, const int *perm, int n)
{
    char fname[4096];
    csr_data_t *xo = calloc((size_t)n, sizeof(csr_data_t));
    for (int i = 0; i < n; i++) xo[perm[i]] = x[i];
    snprintf(fname, sizeof fname, "%s/x%d.out", dir, step);
    FILE *fd = fopen(fname, "wb");
    if (!fd) FAIL("cant open %s\n", fname);
    int hdr[2] = {n, 1};
    fwrite(hdr, sizeof(int), 2, fd);
    fwrite(xo, sizeof(csr_data_t), (size_t)n, fd);
    fclose(fd);
    free(xo);
    printf("saving output file %s\n", fname);
}

/* runstats.txt, rewritten at every output as the reference does (tools/dirac_solver.c:622-628): one line
   "iters, tol, usec" per step from step 1; steps before a resumed run's first step are zeros */
static void write_stats(const char *dir, int first_step, int done, const int *iters, const double *tols,
                        const long *usec)
{
    char fname[4096];
    snprintf(fname, sizeof fname, "%s/runstats.txt", dir);
    FILE *fs = fopen(fname, "w");
    if (!fs) FAIL("cant open %s\n", fname);
    for (int s = 1; s < first_step; s++) fprintf(fs, "%d, %e, %ld\n", 0, 0.0, 0L);
    for (int s = 0; s < done; s++) fprintf(fs, "%d, %e, %ld\n", iters[s], tols[s], usec[s]);
    fclose(fs);
}

typedef struct {
    gpu_sparse_csr_t *gpuH, *gpuS;
} gpu_HS_matrices;

/* out = alpha (H + S) x + beta out   (tools/dirac_solver.c:141-156) */
static void gpu_HS_spmv_fun(const void *mat, gpu_dense_vec_t *x, gpu_dense_vec_t *out, csr_data_t alpha, csr_data_t beta)
{
    const gpu_HS_matrices *hs = mat;
    gpu_spmv(hs->gpuH, x, out, alpha, beta);
    gpu_spmv_local(hs->gpuS, x, out, alpha, 1);
}

static void gpu_LU_precond_fun(const void *precond, const gpu_dense_vec_t *rhs, gpu_dense_vec_t *x)
{
    precond_lu_solve((const gade_precond_t *)precond, rhs, x);
}

int main(int argc, char *argv[])
{
    int opt, lmax = 1, maxit = 500, fused = 0, write_every = 2000, iter = 1;
    double maxtime = 0, intensity = 0, omega = 50, cycles = 1, ntime = 0, tol = 1e-12;
    const char *dir = ".", *xfile = NULL, *ckfile = NULL;
    while ((opt = getopt(argc, argv, "t:l:i:o:c:n:s:e:m:d:w:fx:k:")) != -1) {
        switch (opt) {
        case 't': maxtime = atof(optarg); break;
        case 'l': lmax = atoi(optarg); break;
        case 'i': intensity = atof(optarg); break;
        case 'o': omega = atof(optarg); break;
        case 'c': cycles = atof(optarg); break;
        case 'n': ntime = atof(optarg); break;
        case 's': iter = atoi(optarg); break;
        case 'e': tol = atof(optarg); break;
        case 'm': maxit = atoi(optarg); break;
        case 'd': dir = optarg; break;
        case 'w': write_every = atoi(optarg); break;
        case 'f': fused = 1; break;
        case 'x': xfile = optarg; break;
        case 'k': ckfile = optarg; break;
        default: FAIL("Usage: %s -t time -l lmax -i intensity -o omega -c cycles -n Ntime [-s step -e tol -m maxit -d dir -w every -f -x file -k file]\n", argv[0]);
        }
    }
    if (ntime <= 0 || maxtime <= 0 || iter < 1) FAIL("need -t maxtime > 0, -n Ntime > 0 and -s >= 1\n");
    if (ckfile && !fused) FAIL("-k (checkpoint / resume) runs with the fused operator: add -f\n");
    const double dt = maxtime / ntime, h = 1;
    const int nsteps = (int)ntime - (iter - 1);
    printf("Parameters:\n - time       %f\n - lmax       %d\n - intensity  %lf\n - omega      %lf\n - cycles     %lf\n"
           " - iterations %lf\n - dt         %.15lf\n - mode       %s\n",
           maxtime, lmax, intensity, omega, cycles, ntime, dt, fused ? "fused" : "assembled");
    char fname[4096];
    sparse_csr_t Hall;
    snprintf(fname, sizeof fname, "%s/H.csr", dir);
    csr_read(fname, &Hall);
    sparse_csr_t g0;
    snprintf(fname, sizeof fname, "%s/g0a0l0.csr", dir);
    csr_read(fname, &g0);
    const int blkdim = g0.nrows, n = Hall.nrows * blkdim;
    csr_data_t *psi = malloc(sizeof(csr_data_t) * (size_t)n);
    int *perm = malloc(sizeof(int) * (size_t)n);
    snprintf(fname, sizeof fname, "%s/psi0.vec", dir);
    vec_read(fname, n, psi, perm);
    if (xfile) {
        /* post-processing of a saved state: read it (original order), permute into the solver's block order */
        gade_model_t *m = NULL;
        if (gade_model_load(dir, lmax, h, dt, NULL, &m)) FAIL("gade_model_load: %s\n", gade_last_error());
        FILE *fx = fopen(xfile, "rb");
        if (!fx) FAIL("cant open %s\n", xfile);
        int hdr[2];
        csr_data_t *xo = malloc(sizeof(csr_data_t) * (size_t)n);
        if (fread(hdr, sizeof(int), 2, fx) != 2 || hdr[0] != n || hdr[1] != 1 ||
            fread(xo, sizeof(csr_data_t), (size_t)n, fx) != (size_t)n)
            FAIL("wrong file format in %s\n", xfile);
        fclose(fx);
        for (int i = 0; i < n; i++) psi[i] = xo[perm[i]];
        double *dx = NULL;
        if (gade_malloc((void **)&dx, sizeof(csr_data_t) * (size_t)n) || gade_memcpy_h2d(dx, psi, sizeof(csr_data_t) * (size_t)n))
            FAIL("%s\n", gade_last_error());
        if (gade_eig_write(m, dir)) FAIL("gade_eig_write: %s\n", gade_last_error());
        const char *base = strrchr(xfile, '/');
        base = base ? base + 1 : xfile;
        snprintf(fname, sizeof fname, "%s/psiev_%s", dir, base);
        write_psiev(fname, m, dx, Hall.perm, Hall.nrows, blkdim);
        gade_free(dx);
        gade_model_destroy(m);
        free(xo);
        return 0;
    }
    int *itersarr = calloc((size_t)nsteps, sizeof(int));
    double *tolsarr = calloc((size_t)nsteps, sizeof(double));
    long *timearr = calloc((size_t)nsteps, sizeof(long));
    double time = dt * (iter - 1);
    gpu_sparse_init();
    double t_all = now_us();

    if (fused) {
        gade_model_t *m = NULL;
        if (gade_model_load(dir, lmax, h, dt, NULL, &m)) FAIL("gade_model_load: %s\n", gade_last_error());
        double *dpsi = NULL;
        if (gade_malloc((void **)&dpsi, sizeof(csr_data_t) * (size_t)n) ||
            gade_memcpy_h2d(dpsi, psi, sizeof(csr_data_t) * (size_t)n))
            FAIL("%s\n", gade_last_error());
        if (ckfile && iter > 1) {
            /* resume: psi, step, accumulated time and the solver state of the checkpoint */
            int done = 0;
            if (gade_checkpoint_load(m, dpsi, &done, &time, ckfile)) FAIL("resume: %s\n", gade_last_error());
            if (done != iter - 1) FAIL("resume: %s holds step %d, -s %d needs step %d\n", ckfile, done, iter, iter - 1);
            printf("resumed from %s at step %d, time %.17g\n", ckfile, done, time);
        }
        double *usec = calloc((size_t)nsteps, sizeof(double));
        const int last = iter + nsteps - 1;
        for (int s0 = 0; s0 < nsteps;) {
            /* windows end at the output steps (multiples of write_every) and at the last step */
            const int first = iter + s0;
            int end = (first + write_every - 1) / write_every * write_every;
            if (end > last) end = last;
            const int cnt = end - first + 1;
            if (gade_cn_run(m, dpsi, intensity, omega, cycles, time, cnt, tol, maxit, itersarr + s0, tolsarr + s0,
                            usec + s0))
                FAIL("gade_cn_run: %s\n", gade_last_error());
            for (int k = 0; k < cnt; k++) time = time + dt; /* as gade_cn_run accumulated it (tools/dirac_solver.c:509) */
            for (int k = 0; k < cnt; k++) timearr[s0 + k] = (long)usec[s0 + k];
            s0 += cnt;
            if (gade_memcpy_d2h(psi, dpsi, sizeof(csr_data_t) * (size_t)n)) FAIL("%s\n", gade_last_error());
            write_out(dir, end, psi, perm, n);
            write_stats(dir, iter, s0, itersarr, tolsarr, timearr);
            if (ckfile && gade_checkpoint_save(m, dpsi, end, time, ckfile))
                FAIL("checkpoint: %s\n", gade_last_error());
        }
        free(usec);
        gade_free(dpsi);
        gade_model_destroy(m);
    } else {
        /* tools/dirac_solver.c:285-485: matrices, unblocked structures, stationary part, preconditioner */
        sparse_csr_t *H = calloc(2 * (size_t)lmax, sizeof(sparse_csr_t));
        for (int l = 0; l < lmax; l++) {
            snprintf(fname, sizeof fname, "%s/H0l%d.csr", dir, l);
            csr_read(fname, H + 2 * l);
            snprintf(fname, sizeof fname, "%s/H1l%d.csr", dir, l);
            csr_read(fname, H + 2 * l + 1);
        }
        sparse_csr_t *g = calloc(24 * (size_t)lmax, sizeof(sparse_csr_t));
        for (int a = 0; a < 6; a++)
            for (int l = 0; l < lmax; l++)
                for (int gi = 0; gi < 4; gi++) {
                    snprintf(fname, sizeof fname, "%s/g%da%dl%d.csr", dir, gi, a, l);
                    csr_read(fname, g + a * 4 * lmax + l * 4 + gi);
                }
        sparse_csr_t h0[4], s0[3];
        for (int k = 0; k < 4; k++) {
            snprintf(fname, sizeof fname, "%s/h0%d.csr", dir, k);
            csr_read(fname, h0 + k);
        }
        for (int k = 0; k < 3; k++) {
            snprintf(fname, sizeof fname, "%s/s%d.csr", dir, k);
            csr_read(fname, s0 + k);
        }
        sparse_csr_t Sdiag, Hfull_blk = Hall, S_blk, Hst_blk, Hfull, S, Hst, submatrix;
        memset(&submatrix, 0, sizeof submatrix);
        memset(&Hst_blk, 0, sizeof Hst_blk);
        csr_diag(&Sdiag, Hall.nrows);
        if (Hall.perm) {
            Sdiag.perm = malloc(sizeof(int) * (size_t)Hall.nrows);
            memcpy(Sdiag.perm, Hall.perm, sizeof(int) * (size_t)Hall.nrows);
        }
        S_blk = Sdiag;
        csr_copy(&Hst_blk, &S_blk);
        csr_block_params(&Hfull_blk, blkdim, csr_nnz(&g[0]));
        csr_block_params(&S_blk, blkdim, csr_nnz(&s0[0]));
        csr_block_params(&Hst_blk, blkdim, csr_nnz(&s0[0]));
        csr_copy(&submatrix, g);
        csr_unblock_matrix(&Hfull, &Hfull_blk, &submatrix);
        csr_unblock_matrix(&S, &S_blk, &submatrix);
        csr_unblock_matrix(&Hst, &Hst_blk, &submatrix);
        gpu_init_model_matrices(24 * lmax, g, NULL, h0);
        gpu_compute_row_col(lmax, H, &Hfull_blk, &Hfull);
        compute_stationary_matrices(h, dt, &submatrix, &S_blk, &S, &Hst_blk, &Hst, h0, s0);
        gade_precond_t lu = compute_preconditioner(&S, &Hst);
        gpu_blas_init();
        gpu_sparse_csr_t gpuS = {0}, gpuHfull = {0};
        gpu_dense_vec_t xgpu = {0}, rhsgpu = {0};
        gpu_put_vec(&xgpu, psi, n);
        gpu_vec_local_part(&xgpu, n, 0);
        gpu_put_vec(&rhsgpu, NULL, n);
        gpu_put_csr(&gpuS, &S);
        gpu_HS_matrices hs = {&gpuHfull, &gpuS};
        gpu_solver_workspace_t wsp;
        memset(&wsp, 0, sizeof wsp);
        for (int s = 0; s < nsteps; s++) {
            csr_data_t ft[6];
            beyondDipolePulse_t bdpp;
            int iters = maxit;
            double tol_error = tol;
            time = time + dt;
            beyondDipolePulse_init(&bdpp, intensity, omega, cycles);
            beoyndDipolePulse_axialPart(&bdpp, time, ft);
            double t0 = now_us();
            gpu_compute_timedep_matrices(h, dt, ft, lmax, &Hfull_blk, &Hfull, &gpuHfull);
            gpu_spmv(&gpuHfull, &xgpu, &rhsgpu, -1, 0);   /* rhs = (S - H) psi */
            gpu_spmv_local(&gpuS, &xgpu, &rhsgpu, 1, 1);
            gpu_bicgstab(gpu_HS_spmv_fun, &hs, &rhsgpu, &xgpu, n, n, 0, gpu_LU_precond_fun, &lu, &wsp, &iters, &tol_error);
            itersarr[s] = iters;
            tolsarr[s] = tol_error;
            timearr[s] = (long)(now_us() - t0);
            if ((iter + s) % write_every == 0 || s == nsteps - 1) {
                gpu_get_vec(psi, &xgpu);
                write_out(dir, iter + s, psi, perm, n);
            }
            if ((iter + s) % 10 == 0) {
                printf("iteration %d simulation time %lf\n", iter + s, time);
                fflush(stdout);
            }
        }
    }
    printf("Total runtime: %.3f s\n", (now_us() - t_all) * 1e-6);
    write_stats(dir, iter, nsteps, itersarr, tolsarr, timearr);
    return 0;
}
