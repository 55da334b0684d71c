// Norm and populations by eigen-projection (src/hamiltonian.h:872-943, 1268-1307;
// src/process_output.cpp:332-355): for every distinct kappa the generalized
// eigenpairs H0(k) v = E S(k) v (S-orthonormal), and for every block r
// population_j = |v_j^H S(k_r) psi_r|^2, summed into the classes
// E < -c^2 (negative continuum), -c^2 <= E < 0 (bound) and E >= 0 (ionised).
// The eigenpairs are computed once per model on the host by this library's
// dense solver (gade_eig_dense: Cholesky, Householder tridiagonalization,
// implicit QL), one thread per distinct kappa, on the real parts of H0(k) and
// S(k) as the reference does (src/hamiltonian.h:906-943); the projections run
// in hand-written kernels.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <string>
#include <thread>
#include <vector>

#include "gade_device.h"
#include "gade_internal.h"

using gade::set_error;

namespace {

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

// per block r: c_j = v_j^H S(k_r) psi_r for every eigenpair j of its kappa (the amplitudes behind the
// populations); out [nblk][n] complex
__global__ __launch_bounds__(256) void k_eigcoef(const double2 *__restrict__ x, const double2 *__restrict__ D,
                                                 const int *__restrict__ ell_col, const int *__restrict__ blk_q,
                                                 const double2 *__restrict__ VT, int n, int ellw, int nq,
                                                 double2 *__restrict__ out) {
    extern __shared__ double2 w[];
    const int r = blockIdx.x;
    const int q = blk_q[r];
    const double2 *Sq = D + ((size_t)2 * nq + q) * ellw * n;
    const double2 *xr = x + (size_t)r * n;
    for (int a = threadIdx.x; a < n; a += blockDim.x) {
        double2 sx = make_double2(0, 0);
        for (int sl = 0; sl < ellw; sl++) sx = cfma(Sq[(size_t)sl * n + a], xr[ell_col[(size_t)sl * n + a]], sx);
        w[a] = sx;
    }
    __syncthreads();
    const double2 *V = VT + (size_t)q * n * n;
    for (int j = threadIdx.x; j < n; j += blockDim.x) {
        double2 c0 = make_double2(0, 0), c1 = make_double2(0, 0);
        int i = 0;
        for (; i + 1 < n; i += 2) {
            c0 = cfma(V[(size_t)i * n + j], w[i], c0);
            c1 = cfma(V[(size_t)(i + 1) * n + j], w[i + 1], c1);
        }
        if (i < n) c0 = cfma(V[(size_t)i * n + j], w[i], c0);
        out[(size_t)r * n + j] = cadd(c0, c1);
    }
}

}  // namespace

extern "C" {

int gade_eig_setup(gade_model_t *m) {
    if (!m) return GADE_EINVAL;
    if (m->d_evec) return GADE_OK;
    const int n = m->n, nq = m->nq, nnzb = m->nnzb;
    const size_t nn = (size_t)n * n;
    // radial values of s0..s2 and h00..h03 (real parts, as the reference's .real())
    std::vector<double2> hs((size_t)3 * nnzb), hh((size_t)4 * nnzb);
    GADE_HIP(hipMemcpy(hs.data(), m->d_s, sizeof(double2) * hs.size(), hipMemcpyDeviceToHost));
    GADE_HIP(hipMemcpy(hh.data(), m->d_h0, sizeof(double2) * hh.size(), hipMemcpyDeviceToHost));
    std::vector<double> E((size_t)nq * n), V((size_t)nq * nn);
    std::vector<int> rcs(nq, GADE_OK);
    std::atomic<int> next{0};
    auto worker = [&]() {
        std::vector<double> A(nn), B(nn);
        for (int q = next++; q < nq; q = next++) {
            const double k = m->kap_q[q];
            std::fill(A.begin(), A.end(), 0.0);
            std::fill(B.begin(), B.end(), 0.0);
            for (int i = 0; i < n; i++)
                for (int p = m->rad_Ap[i]; p < m->rad_Ap[i + 1]; p++) {
                    const int c = m->rad_Ai[p];
                    B[(size_t)i * n + c] = hs[p].x + k * hs[nnzb + p].x + k * k * hs[2 * nnzb + p].x;
                    A[(size_t)i * n + c] =
                        hh[p].x + k * hh[nnzb + p].x + k * k * hh[2 * nnzb + p].x + k * k * k * hh[3 * nnzb + p].x;
                }
            rcs[q] = gade_eig_dense(n, A.data(), B.data(), E.data() + (size_t)q * n, V.data() + (size_t)q * nn);
        }
    };
    const int nthr = std::max(1, std::min<int>(nq, std::min(16u, std::max(1u, std::thread::hardware_concurrency()))));
    std::vector<std::thread> pool;
    for (int t = 0; t < nthr; t++) pool.emplace_back(worker);
    for (auto &t : pool) t.join();
    for (int q = 0; q < nq; q++)
        if (rcs[q] != GADE_OK) {
            set_error("gade_eig_setup: generalized eigenproblem of kappa " + std::to_string(m->kap_q[q]) +
                      (rcs[q] == GADE_EINVAL ? " failed: S is not positive definite" : " did not converge"));
            return rcs[q];
        }
    // VT[q][i][j] = conj(v_j[i]); the eigenvectors are real
    std::vector<double2> vt((size_t)nq * nn);
    for (size_t idx = 0; idx < vt.size(); idx++) vt[idx] = make_double2(V[idx], 0.0);
    double *d_eval = nullptr;
    double2 *d_evec = nullptr;
    GADE_TRY(gade::hip_alloc((void **)&d_eval, sizeof(double) * nq * n));
    int rc = gade::hip_alloc((void **)&d_evec, sizeof(double2) * vt.size());
    hipError_t e = hipSuccess;
    if (rc == GADE_OK) e = hipMemcpy(d_eval, E.data(), sizeof(double) * E.size(), hipMemcpyHostToDevice);
    if (rc == GADE_OK && e == hipSuccess) e = hipMemcpy(d_evec, vt.data(), sizeof(double2) * vt.size(), hipMemcpyHostToDevice);
    if (rc != GADE_OK || e != hipSuccess) {
        (void)hipFree(d_eval);
        if (d_evec) (void)hipFree(d_evec);
        if (rc != GADE_OK) return rc;
        set_error(std::string("gade_eig_setup: ") + hipGetErrorString(e));
        return GADE_EHIP;
    }
    m->d_eval = d_eval;
    m->d_evec = d_evec;
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

int gade_psiev(gade_model_t *m, const double *x, double *psiev) {
    if (!m || !x || !psiev) return GADE_EINVAL;
    GADE_TRY(gade_eig_setup(m));
    const size_t cnt = (size_t)m->nblk * m->n;
    double2 *d_out = nullptr;
    GADE_TRY(gade::hip_alloc((void **)&d_out, sizeof(double2) * cnt));
    hipLaunchKernelGGL(k_eigcoef, dim3(m->nblk), dim3(256), sizeof(double2) * m->n, m->stream, (const double2 *)x,
                       m->d_D, m->d_ell_col, m->d_blk_q, m->d_evec, m->n, m->ellw, m->nq, d_out);
    hipError_t e = hipGetLastError();
    if (e == hipSuccess) e = hipMemcpyAsync(psiev, d_out, sizeof(double2) * cnt, hipMemcpyDeviceToHost, m->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(m->stream);
    (void)hipFree(d_out);
    if (e != hipSuccess) {
        set_error(std::string("gade_psiev: ") + hipGetErrorString(e));
        return GADE_EHIP;
    }
    // psiev_j = psi^H S v_j = conj(v_j^H S psi)  (src/hamiltonian.h:1318: psiseg.adjoint() * skev)
    for (size_t i = 0; i < cnt; i++) psiev[2 * i + 1] = -psiev[2 * i + 1];
    return GADE_OK;
}

// dense_write of src/dumpEigvals.cpp:87-108: int64 rows, int64 cols, then the column-major data
static int dense_write(const std::string &fname, int64_t rows, int64_t cols, const void *data, size_t elt) {
    FILE *fd = fopen(fname.c_str(), "wb");
    if (!fd) {
        set_error("cant open " + fname);
        return GADE_EIO;
    }
    bool ok = fwrite(&rows, sizeof rows, 1, fd) == 1 && fwrite(&cols, sizeof cols, 1, fd) == 1 &&
              fwrite(data, elt, (size_t)(rows * cols), fd) == (size_t)(rows * cols);
    ok = (fclose(fd) == 0) && ok;
    if (!ok) {
        set_error("write failed: " + fname);
        return GADE_EIO;
    }
    return GADE_OK;
}

int gade_eig_write(gade_model_t *m, const char *dir) {
    if (!m || !dir) return GADE_EINVAL;
    GADE_TRY(gade_eig_setup(m));
    const int n = m->n, nq = m->nq;
    const size_t nn = (size_t)n * n;
    std::vector<double> E((size_t)nq * n);
    std::vector<double2> VT((size_t)nq * nn), evc(nn);
    GADE_HIP(hipMemcpy(E.data(), m->d_eval, sizeof(double) * E.size(), hipMemcpyDeviceToHost));
    GADE_HIP(hipMemcpy(VT.data(), m->d_evec, sizeof(double2) * VT.size(), hipMemcpyDeviceToHost));
    for (int q = 0; q < nq; q++) {
        const int kappa = m->kap_q[q];
        GADE_TRY(dense_write(std::string(dir) + "/evl" + std::to_string(kappa) + ".mat", n, 1, E.data() + (size_t)q * n,
                             sizeof(double)));
        // column j = eigenvector j: evc[j * n + i] = v_j[i] = conj(VT[q][i][j])
        const double2 *V = VT.data() + (size_t)q * nn;
        for (int i = 0; i < n; i++)
            for (int j = 0; j < n; j++) evc[(size_t)j * n + i] = make_double2(V[(size_t)i * n + j].x, -V[(size_t)i * n + j].y);
        GADE_TRY(dense_write(std::string(dir) + "/evc" + std::to_string(kappa) + ".mat", n, n, evc.data(),
                             sizeof(double2)));
    }
    return GADE_OK;
}

}  // extern "C"
