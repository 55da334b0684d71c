/* Sanitizer run of the C oracle (test infrastructure): loads a coupling directory and exercises every
   oracle routine the tests use -- pulse, structure, time-dependent assembly, assembled and matrix-free
   products, the block preconditioner, BiCGSTAB and a few CN steps -- so that an AddressSanitizer /
   UndefinedBehaviorSanitizer build reports any out-of-bounds access, leak or undefined operation.
   usage: oracle_run <dir> <lmax> <dt> <nthreads>; prints a checksum of the results. */
#include <complex.h>
#include <math.h>
#include <stdio.h>
#include <stdlib.h>

#include "gade_oracle.h"

int main(int argc, char **argv)
{
    if (argc != 5) {
        fprintf(stderr, "usage: %s dir lmax dt nthreads\n", argv[0]);
        return 2;
    }
    const int lmax = atoi(argv[2]), nth = atoi(argv[4]);
    const double dt = atof(argv[3]);
    or_model_t *m = or_model_load(argv[1], lmax, 1.0, dt, nth);
    if (!m) return 3;
    int nblk, blkdim, hall_nnz, nnz_b;
    long long nnz_full;
    or_model_info(m, &nblk, &blkdim, &hall_nnz, &nnz_full, &nnz_b);
    const int n = nblk * blkdim;
    ocplx *x = malloc(sizeof(ocplx) * n), *y = calloc(n, sizeof(ocplx)), *z = calloc(n, sizeof(ocplx));
    srand(7);
    for (int i = 0; i < n; i++) x[i] = (rand() / (double)RAND_MAX - 0.5) + I * (rand() / (double)RAND_MAX - 0.5);
    ocplx ft[6];
    or_pulse(10.0, 50.0, 1.0, 0.37 * 2 * M_PI / 50.0, ft);
    or_model_timedep(m, ft);
    or_model_spmv_HS(m, x, y);
    or_model_matfree_H(m, ft, x, z); /* accumulates into z (zero) */
    double s = 0, d = 0;
    for (int i = 0; i < n; i++) {
        s += cabs(y[i]);
        y[i] = 0; /* the products accumulate (out += A x), as the reference's csr_spmv */
    }
    or_model_spmv_H(m, x, y);
    for (int i = 0; i < n; i++) d = fmax(d, cabs(y[i] - z[i]));
    if (or_model_precond_setup(m)) return 4;
    or_model_psolve(m, x, y);
    for (int i = 0; i < n; i++) s += cabs(y[i]);
    or_model_set_threads(m, nth, 1);  /* the all-cores paths as well */
    int iters = 500, napply = 0;
    double tol = 1e-12;
    for (int i = 0; i < n; i++) z[i] = x[i];
    or_bicgstab(m, x, z, &iters, &tol, &napply);
    or_reset_workspace(m);
    int *it = malloc(sizeof(int) * 4);
    double *te = malloc(sizeof(double) * 4);
    int rc = or_cn_run(m, 10.0, 50.0, 1.0, 0.0, 4, 1e-12, 500, x, it, te, NULL);
    int ir = or_ik(7);
    double mu = or_imu(7);
    printf("checksum %.6e matfree-vs-assembled %.3e bicgstab %d %.2e cn %d %d %.1f\n", s, d, iters, tol, rc, ir, mu);
    free(it);
    free(te);
    free(x);
    free(y);
    free(z);
    or_model_free(m);
    return d <= 1e-10 * s ? 0 : 5;
}
