// Norm and populations by eigen-projection (src/hamiltonian.h:872-943, 1268-1307;
// src/process_output.cpp:332-355): for every distinct kappa the generalized
// eigenpairs H0(k) v = E S(k) v (S-orthonormal), and for every block r
// population_j = |v_j^H S(k_r) psi_r|^2, summed into the classes
// E < -c^2 (negative continuum), -c^2 <= E < 0 (bound) and E >= 0 (ionised).
// The eigenpairs are computed once per model on the device (rocSOLVER
// zhegvd, batched over kappa); the projection is a hand-written kernel.
#include <hip/hip_runtime.h>
#include <rocblas/rocblas.h>
#include <rocsolver/rocsolver.h>

#include <cstring>
#include <string>
#include <vector>

#include "gade_device.h"
#include "gade_internal.h"

using gade::set_error;

namespace {

// dense column-major H0(k_q), S(k_q) [q][col][row] from the radial CSR values
__global__ void k_dense_ops(const double2 *__restrict__ s, const double2 *__restrict__ h0,
                            const int *__restrict__ ell_src, const int *__restrict__ ell_col,
                            const int *__restrict__ kq, int nq, int n, int ellw, int nnzb, double2 *__restrict__ A,
                            double2 *__restrict__ B) {
    const long long total = (long long)nq * n;
    for (long long idx = blockIdx.x * (long long)blockDim.x + threadIdx.x; idx < total;
         idx += (long long)gridDim.x * blockDim.x) {
        const int q = (int)(idx / n), a = (int)(idx % n);
        const double k = kq[q];
        const size_t base = (size_t)q * n * n;
        for (int sl = 0; sl < ellw; sl++) {
            const int src = ell_src[(size_t)sl * n + a];
            if (src < 0) continue;
            const int c = ell_col[(size_t)sl * n + a];
            const double2 S = cadd(cadd(s[src], cscale(s[nnzb + src], k)), cscale(s[2 * nnzb + src], k * k));
            const double2 H = cadd(cadd(cadd(h0[src], cscale(h0[nnzb + src], k)), cscale(h0[2 * nnzb + src], k * k)),
                                   cscale(h0[3 * nnzb + src], k * k * k));
            A[base + (size_t)c * n + a] = H;
            B[base + (size_t)c * n + a] = S;
        }
    }
}

// VT[q][i][j] = conj(v_j[i]) from the column-major eigenvectors A[q][j][i]
__global__ void k_conj_transpose(const double2 *__restrict__ A, int n, int nq, double2 *__restrict__ VT) {
    const long long total = (long long)nq * n * n;
    for (long long idx = blockIdx.x * (long long)blockDim.x + threadIdx.x; idx < total;
         idx += (long long)gridDim.x * blockDim.x) {
        const long long q = idx / ((long long)n * n);
        const int rem = (int)(idx % ((long long)n * n));
        const int i = rem / n, j = rem % n;
        double2 v = A[q * n * n + (size_t)j * n + i];
        v.y = -v.y;
        VT[idx] = v;
    }
}

// per block r: w = S(k_r) psi_r (staged in LDS), c_j = sum_i conj(v_j[i]) w_i,
// out[r] = (sum over classes of |c_j|^2, then psi_r^H S psi_r)
__global__ __launch_bounds__(256) void k_eigproj(const double2 *__restrict__ x, const double2 *__restrict__ D,
                                                 const int *__restrict__ ell_col, const int *__restrict__ blk_q,
                                                 const double2 *__restrict__ VT, const double *__restrict__ E,
                                                 int n, int ellw, int nq, double csq, double *__restrict__ out) {
    extern __shared__ double2 w[];
    __shared__ double sh[4 * 16];
    const int r = blockIdx.x;
    const int q = blk_q[r];
    const double2 *Sq = D + ((size_t)2 * nq + q) * ellw * n;  // S(k) in ELL layout
    const double2 *xr = x + (size_t)r * n;
    double acc[4] = {0, 0, 0, 0};
    for (int a = threadIdx.x; a < n; a += blockDim.x) {
        double2 sx = make_double2(0, 0);
        for (int sl = 0; sl < ellw; sl++) sx = cfma(Sq[(size_t)sl * n + a], xr[ell_col[(size_t)sl * n + a]], sx);
        w[a] = sx;
        const double2 xa = xr[a];
        acc[3] += xa.x * sx.x + xa.y * sx.y;
    }
    __syncthreads();
    const double2 *V = VT + (size_t)q * n * n;
    const double *Eq = E + (size_t)q * n;
    for (int j = threadIdx.x; j < n; j += blockDim.x) {
        double2 c0 = make_double2(0, 0), c1 = make_double2(0, 0);
        int i = 0;
        for (; i + 1 < n; i += 2) {
            c0 = cfma(V[(size_t)i * n + j], w[i], c0);
            c1 = cfma(V[(size_t)(i + 1) * n + j], w[i + 1], c1);
        }
        if (i < n) c0 = cfma(V[(size_t)i * n + j], w[i], c0);
        const double2 c = cadd(c0, c1);
        const double p = c.x * c.x + c.y * c.y;
        const double e = Eq[j];
        if (e < -csq) acc[0] += p;
        else if (e < 0) acc[1] += p;
        else acc[2] += p;
    }
    block_sum<4>(acc, sh);
    if (threadIdx.x == 0)
        for (int k = 0; k < 4; k++) out[(size_t)r * 4 + k] = acc[k];
}

}  // namespace

extern "C" {

int gade_eig_setup(gade_model_t *m) {
    if (!m) return GADE_EINVAL;
    if (m->d_evec) return GADE_OK;
    const int n = m->n, nq = m->nq;
    const size_t nn = (size_t)n * n;
    double2 *A = nullptr, *B = nullptr;
    double *Ework = nullptr;
    int *d_kq = nullptr, *info = nullptr;
    int rc = GADE_OK;
    bool ok = false;
    auto cleanup = [&]() {
        for (void *p : {(void *)A, (void *)B, (void *)Ework, (void *)d_kq, (void *)info})
            if (p) (void)hipFree(p);
        if (!ok) {
            for (void *p : {(void *)m->d_eval, (void *)m->d_evec})
                if (p) (void)hipFree(p);
            m->d_eval = nullptr;
            m->d_evec = nullptr;
        }
    };
    if ((rc = gade::hip_alloc((void **)&A, sizeof(double2) * nq * nn)) ||
        (rc = gade::hip_alloc((void **)&B, sizeof(double2) * nq * nn)) ||
        (rc = gade::hip_alloc((void **)&Ework, sizeof(double) * nq * n)) ||
        (rc = gade::hip_alloc((void **)&d_kq, sizeof(int) * nq)) ||
        (rc = gade::hip_alloc((void **)&info, sizeof(int) * nq)) ||
        (rc = gade::hip_alloc((void **)&m->d_eval, sizeof(double) * nq * n)) ||
        (rc = gade::hip_alloc((void **)&m->d_evec, sizeof(double2) * nq * nn))) {
        cleanup();
        return rc;
    }
    hipError_t e = hipMemcpy(d_kq, m->kap_q.data(), sizeof(int) * nq, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemsetAsync(A, 0, sizeof(double2) * nq * nn, m->stream);
    if (e == hipSuccess) e = hipMemsetAsync(B, 0, sizeof(double2) * nq * nn, m->stream);
    if (e == hipSuccess) {
        hipLaunchKernelGGL(k_dense_ops, dim3((nq * n + 255) / 256), dim3(256), 0, m->stream, m->d_s, m->d_h0,
                           m->d_ell_src, m->d_ell_col, d_kq, nq, n, m->ellw, m->nnzb, A, B);
        e = hipGetLastError();
    }
    if (e != hipSuccess) {
        cleanup();
        set_error(std::string("gade_eig_setup: ") + hipGetErrorString(e));
        return GADE_EHIP;
    }
    rocblas_handle h = nullptr;
    if (rocblas_create_handle(&h) != rocblas_status_success) {
        cleanup();
        set_error("gade_eig_setup: rocblas_create_handle failed");
        return GADE_EHIP;
    }
    rocblas_set_stream(h, m->stream);
    rocblas_status st = rocsolver_zhegvd_strided_batched(
        h, rocblas_eform_ax, rocblas_evect_original, rocblas_fill_upper, n, (rocblas_double_complex *)A, n,
        (rocblas_stride)nn, (rocblas_double_complex *)B, n, (rocblas_stride)nn, m->d_eval, (rocblas_stride)n, Ework,
        (rocblas_stride)n, info, nq);
    std::vector<int> info_h(nq, 0);
    e = hipStreamSynchronize(m->stream);
    if (e == hipSuccess) e = hipMemcpy(info_h.data(), info, sizeof(int) * nq, hipMemcpyDeviceToHost);
    rocblas_destroy_handle(h);
    if (st != rocblas_status_success || e != hipSuccess) {
        cleanup();
        set_error("gade_eig_setup: rocsolver_zhegvd failed");
        return GADE_EHIP;
    }
    for (int q = 0; q < nq; q++)
        if (info_h[q] != 0) {
            cleanup();
            set_error("gade_eig_setup: generalized eigenproblem of kappa " + std::to_string(m->kap_q[q]) +
                      " failed (info " + std::to_string(info_h[q]) + "); S must be Hermitian positive definite");
            return GADE_EINVAL;
        }
    hipLaunchKernelGGL(k_conj_transpose, dim3(4096), dim3(256), 0, m->stream, A, n, nq, m->d_evec);
    e = hipGetLastError();
    if (e == hipSuccess) e = hipStreamSynchronize(m->stream);
    ok = e == hipSuccess;
    cleanup();
    if (e != hipSuccess) {
        set_error(std::string("gade_eig_setup: ") + hipGetErrorString(e));
        return GADE_EHIP;
    }
    return GADE_OK;
}

int gade_eig_values(gade_model_t *m, int *nkappa, int *kappa, double *energies) {
    if (!m) return GADE_EINVAL;
    GADE_TRY(gade_eig_setup(m));
    if (nkappa) *nkappa = m->nq;
    if (kappa) memcpy(kappa, m->kap_q.data(), sizeof(int) * m->nq);
    if (energies)
        GADE_HIP(hipMemcpy(energies, m->d_eval, sizeof(double) * m->nq * m->n, hipMemcpyDeviceToHost));
    return GADE_OK;
}

int gade_populations(gade_model_t *m, const double *x, double *norm, double *bound, double *ionised,
                     double *per_block) {
    if (!m || !x) return GADE_EINVAL;
    GADE_TRY(gade_eig_setup(m));
    double *d_out = nullptr;
    GADE_TRY(gade::hip_alloc((void **)&d_out, sizeof(double) * 4 * m->nblk));
    const double csq = gade::kSoL * gade::kSoL;
    hipLaunchKernelGGL(k_eigproj, dim3(m->nblk), dim3(256), sizeof(double2) * m->n, m->stream, (const double2 *)x,
                       m->d_D, m->d_ell_col, m->d_blk_q, m->d_evec, m->d_eval, m->n, m->ellw, m->nq, csq, d_out);
    std::vector<double> out((size_t)4 * m->nblk);
    hipError_t e = hipGetLastError();
    if (e == hipSuccess)
        e = hipMemcpyAsync(out.data(), d_out, sizeof(double) * out.size(), hipMemcpyDeviceToHost, m->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(m->stream);
    (void)hipFree(d_out);
    if (e != hipSuccess) {
        set_error(std::string("gade_populations: ") + hipGetErrorString(e));
        return GADE_EHIP;
    }
    double tb = 0, ti = 0, tn = 0;
    for (int r = 0; r < m->nblk; r++) {
        tb += out[(size_t)r * 4 + 1];
        ti += out[(size_t)r * 4 + 2];
        tn += out[(size_t)r * 4 + 3];
    }
    if (norm) *norm = tn;
    if (bound) *bound = tb;
    if (ionised) *ionised = ti;
    if (per_block) memcpy(per_block, out.data(), sizeof(double) * out.size());
    return GADE_OK;
}

}  // extern "C"
