/*
 * gpu_solver.h -- per-step assembly of the time-dependent Hamiltonian on the
 * device (drop-in for tools/gpu_solver.h:29-35).
 *
 * gpu_init_model_matrices keeps the radial g matrices (gt is derived on the
 * device as the conjugate transpose) and h0; gpu_compute_row_col keeps the
 * block structure and the angular couplings H0l/H1l;
 * gpu_compute_timedep_matrices writes Hfull.Ax of the device matrix gpu_Hfull
 * for f(t) (G-sums and assembly in libgade kernels).  One model per process,
 * as in the reference.
 */
#ifndef GADE_GPU_SOLVER_H
#define GADE_GPU_SOLVER_H

#include "gpu_sparse.h"

#ifdef __cplusplus
extern "C" {
#endif

void gpu_init_model_matrices(int nmatrices, const sparse_csr_t *cpu_g, const sparse_csr_t *cpu_gt,
                             const sparse_csr_t *cpu_h0);                                   /* tools/gpu_solver.cu:288-327 */
void gpu_compute_row_col(int lmax, const sparse_csr_t *H, sparse_csr_t *Hfull_blk, sparse_csr_t *Hfull); /* :137-236 */
void gpu_compute_timedep_matrices(double h, double dt, csr_data_t *ft, int lmax, const sparse_csr_t *Hfull_blk,
                                  sparse_csr_t *Hfull, gpu_sparse_csr_t *gpu_Hfull);       /* :238-286 */

#ifdef __cplusplus
}
#endif
#endif
