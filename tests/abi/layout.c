/* Prints the size and field offsets of the structs that cross the drop-in boundary.  Compiled twice by
   tests/test_abi.py: once against this repository's include/, once against the reference's own headers
   (with -DUSE_HIP and no MPI, the reference's GPU build); the two outputs must be identical. */
#include <stddef.h>
#include <stdio.h>

#include "csr.h"
#include "gpu_sparse.h"
#include "gpu_bicgstab.h"
#include "bicgstab.h"

#define SZ(T) printf(#T " size %zu\n", sizeof(T))
#define OFF(T, f) printf(#T "." #f " %zu\n", offsetof(T, f))

int main(void)
{
    SZ(csr_data_t);
    SZ(csr_index_t);
    SZ(sparse_csr_t);
    OFF(sparse_csr_t, nrows); OFF(sparse_csr_t, ncols); OFF(sparse_csr_t, nnz);
    OFF(sparse_csr_t, blk_dim); OFF(sparse_csr_t, blk_nnz); OFF(sparse_csr_t, npart);
    OFF(sparse_csr_t, row_beg); OFF(sparse_csr_t, row_end); OFF(sparse_csr_t, local_offset);
    OFF(sparse_csr_t, perm); OFF(sparse_csr_t, row_cpu_dist); OFF(sparse_csr_t, comm_pattern);
    OFF(sparse_csr_t, comm_pattern_size); OFF(sparse_csr_t, n_comm_entries); OFF(sparse_csr_t, recv_ptr);
    OFF(sparse_csr_t, send_ptr); OFF(sparse_csr_t, send_vec); OFF(sparse_csr_t, Ap); OFF(sparse_csr_t, Ai);
    OFF(sparse_csr_t, Ax); OFF(sparse_csr_t, Ai_sub_map); OFF(sparse_csr_t, is_link);
    SZ(gpu_sparse_csr_t);
    OFF(gpu_sparse_csr_t, nrows); OFF(gpu_sparse_csr_t, ncols); OFF(gpu_sparse_csr_t, nnz);
    OFF(gpu_sparse_csr_t, Ap); OFF(gpu_sparse_csr_t, Ai); OFF(gpu_sparse_csr_t, Ax);
    OFF(gpu_sparse_csr_t, Ai_sub_map); OFF(gpu_sparse_csr_t, desc); OFF(gpu_sparse_csr_t, spsvDescr);
    OFF(gpu_sparse_csr_t, cuBuffer);
    SZ(gpu_dense_vec_t);
    OFF(gpu_dense_vec_t, desc); OFF(gpu_dense_vec_t, desc_local); OFF(gpu_dense_vec_t, dim);
    OFF(gpu_dense_vec_t, local_dim); OFF(gpu_dense_vec_t, local_offset); OFF(gpu_dense_vec_t, x);
    SZ(gpu_lu_t);
    OFF(gpu_lu_t, L); OFF(gpu_lu_t, U); OFF(gpu_lu_t, temp);
    SZ(gpu_solver_workspace_t);
    OFF(gpu_solver_workspace_t, r); OFF(gpu_solver_workspace_t, r0); OFF(gpu_solver_workspace_t, v);
    OFF(gpu_solver_workspace_t, p); OFF(gpu_solver_workspace_t, y); OFF(gpu_solver_workspace_t, z);
    OFF(gpu_solver_workspace_t, s); OFF(gpu_solver_workspace_t, t);
    SZ(solver_workspace_t);
    OFF(solver_workspace_t, r); OFF(solver_workspace_t, r0); OFF(solver_workspace_t, v);
    OFF(solver_workspace_t, p); OFF(solver_workspace_t, y); OFF(solver_workspace_t, z);
    OFF(solver_workspace_t, s); OFF(solver_workspace_t, t);
    return 0;
}
