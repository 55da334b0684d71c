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
    const char *dir = ".";
    while ((opt = getopt(argc, argv, "t:l:i:o:c:n:s:e:m:d:w:f")) != -1) {
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
        default: FAIL("Usage: %s -t time -l lmax -i intensity -o omega -c cycles -n Ntime [-s step -e tol -m maxit -d dir -w every -f]\n", argv[0]);
        }
    }
    if (ntime <= 0 || maxtime <= 0 || iter < 1) FAIL("need -t maxtime > 0, -n Ntime > 0 and -s >= 1\n");
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
        double *usec = calloc((size_t)nsteps, sizeof(double));
        for (int s0 = 0; s0 < nsteps; s0 += write_every) {
            int cnt = nsteps - s0 < write_every ? nsteps - s0 : write_every;
            if (gade_cn_run(m, dpsi, intensity, omega, cycles, time, cnt, tol, maxit, itersarr + s0, tolsarr + s0,
                            usec + s0))
                FAIL("gade_cn_run: %s\n", gade_last_error());
            time += cnt * dt;
            if (gade_memcpy_d2h(psi, dpsi, sizeof(csr_data_t) * (size_t)n)) FAIL("%s\n", gade_last_error());
            write_out(dir, iter + s0 + cnt - 1, psi, perm, n);
        }
        for (int s = 0; s < nsteps; s++) timearr[s] = (long)usec[s];
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
    snprintf(fname, sizeof fname, "%s/runstats.txt", dir);
    FILE *fs = fopen(fname, "w");
    if (!fs) FAIL("cant open %s\n", fname);
    for (int s = 0; s < nsteps; s++) fprintf(fs, "%d, %e, %ld\n", itersarr[s], tolsarr[s], timearr[s]);
    fclose(fs);
    return 0;
}
