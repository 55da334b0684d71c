// Reference-compatible entry points (gpu_sparse.h, gpu_bicgstab.h,
// gpu_solver.h, the block LU of solver.h) on libgade's own kernels.
// Errors print and exit(1), as the reference's CHECK_GPU macros do.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <complex>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <mutex>
#include <vector>

#include "gade_internal.h"
#include "gade_device.h"
#include "gpu_bicgstab.h"
#include "gpu_solver.h"
#include "gpu_sparse.h"
#include "solver.h"

#define DIE(...)                      \
    do {                              \
        fprintf(stderr, __VA_ARGS__); \
        fprintf(stderr, "\n");        \
        exit(1);                      \
    } while (0)

#define CHECK_HIP(expr)                                                                                 \
    do {                                                                                                \
        hipError_t e_ = (expr);                                                                         \
        if (e_ != hipSuccess) DIE("GPU failed at %s:%d: %s (%s)", __FILE__, __LINE__, #expr,            \
                                  hipGetErrorString(e_));                                               \
    } while (0)

#define CHECK_GADE(expr)                                                                   \
    do {                                                                                   \
        int rc_ = (expr);                                                                  \
        if (rc_ != GADE_OK) DIE("libgade failed at %s:%d: %s", __FILE__, __LINE__, gade_last_error()); \
    } while (0)

static_assert(sizeof(csr_data_t) == sizeof(double2), "csr_data_t must be two doubles");

namespace {

inline double2 d2(csr_data_t z) { return make_double2(z.real(), z.imag()); }

// level schedule of a triangular CSR matrix (lower or upper)
struct TriSched {
    int n = 0, nlev = 0;
    bool lower = true;
    std::vector<int> level_ptr;  // host
    int *d_rows = nullptr;       // rows ordered by level
    int *d_diag = nullptr;       // position of the diagonal entry of each row
};

// one wave per row of the level: x[i] = (b[i] - sum_{j != i} A_ij x_j) / A_ii
__global__ __launch_bounds__(256) void k_trsv_level(const int *__restrict__ Ap, const int *__restrict__ Ai,
                                                    const double2 *__restrict__ Ax, const int *__restrict__ diag,
                                                    const int *__restrict__ rows, int nrows,
                                                    const double2 *__restrict__ b, double2 *__restrict__ x) {
    const int w = (blockIdx.x * blockDim.x + threadIdx.x) >> 6, lane = threadIdx.x & 63;
    if (w >= nrows) return;
    const int i = rows[w];
    double re = 0, im = 0;
    for (int k = Ap[i] + lane; k < Ap[i + 1]; k += 64) {
        const int j = Ai[k];
        if (j == i) continue;
        const double2 a = Ax[k], v = x[j];
        re += a.x * v.x - a.y * v.y;
        im += a.x * v.y + a.y * v.x;
    }
    re = wave_sum(re);
    im = wave_sum(im);
    if (lane == 0) {
        const double2 d = Ax[diag[i]];
        x[i] = cdiv(make_double2(b[i].x - re, b[i].y - im), d);
    }
}

TriSched *analyze(const gpu_sparse_csr_t *A, bool lower) {
    TriSched *s = new TriSched();
    const int n = A->nrows;
    s->n = n;
    s->lower = lower;
    std::vector<int> Ap(n + 1), Ai(A->nnz);
    CHECK_HIP(hipMemcpy(Ap.data(), A->Ap, sizeof(int) * (n + 1), hipMemcpyDeviceToHost));
    CHECK_HIP(hipMemcpy(Ai.data(), A->Ai, sizeof(int) * A->nnz, hipMemcpyDeviceToHost));
    std::vector<int> lev(n, 0), diag(n, -1);
    for (int t = 0; t < n; t++) {
        const int i = lower ? t : n - 1 - t;
        int L = 0;
        for (int k = Ap[i]; k < Ap[i + 1]; k++) {
            const int j = Ai[k];
            if (j == i) {
                diag[i] = k;
            } else if (lower ? j < i : j > i) {
                L = std::max(L, lev[j] + 1);
            } else {
                DIE("gpu_lu_analyze: %s factor has entry (%d,%d) on the wrong side", lower ? "lower" : "upper", i, j);
            }
        }
        if (diag[i] < 0) DIE("gpu_lu_analyze: row %d has no diagonal entry", i);
        lev[i] = L;
    }
    s->nlev = n ? *std::max_element(lev.begin(), lev.end()) + 1 : 0;
    s->level_ptr.assign(s->nlev + 1, 0);
    for (int i = 0; i < n; i++) s->level_ptr[lev[i] + 1]++;
    for (int l = 0; l < s->nlev; l++) s->level_ptr[l + 1] += s->level_ptr[l];
    std::vector<int> rows(n), fill(s->level_ptr.begin(), s->level_ptr.end() - 1);
    for (int i = 0; i < n; i++) rows[fill[lev[i]]++] = i;
    CHECK_HIP(hipMalloc(&s->d_rows, sizeof(int) * std::max(n, 1)));
    CHECK_HIP(hipMalloc(&s->d_diag, sizeof(int) * std::max(n, 1)));
    CHECK_HIP(hipMemcpy(s->d_rows, rows.data(), sizeof(int) * n, hipMemcpyHostToDevice));
    CHECK_HIP(hipMemcpy(s->d_diag, diag.data(), sizeof(int) * n, hipMemcpyHostToDevice));
    return s;
}

void trsv(const gpu_sparse_csr_t *A, const double2 *b, double2 *x) {
    const TriSched *s = (const TriSched *)A->desc;
    if (!s) DIE("gpu_lu_solve: call gpu_lu_analyze first");
    for (int l = 0; l < s->nlev; l++) {
        const int r0 = s->level_ptr[l], nr = s->level_ptr[l + 1] - r0;
        hipLaunchKernelGGL(k_trsv_level, dim3((nr + 3) / 4), dim3(256), 0, 0, A->Ap, A->Ai, (const double2 *)A->Ax,
                           s->d_diag, s->d_rows + r0, nr, b, x);
    }
    CHECK_HIP(hipGetLastError());
}

// BiCGSTAB engine state (tools/gpu_bicgstab.c keeps a BLAS handle; here scalars and partials)
struct Engine {
    double *d_sc = nullptr, *d_part = nullptr, *h_sc = nullptr;
    int nparts = 0;
} g_engine;

void engine_init(long long N) {
    const int np = (int)std::min<long long>(1024, (N + kRedThreads - 1) / kRedThreads);
    if (g_engine.d_sc && np <= g_engine.nparts) return;
    if (g_engine.d_part) CHECK_HIP(hipFree(g_engine.d_part));
    if (!g_engine.d_sc) {
        CHECK_HIP(hipMalloc(&g_engine.d_sc, sizeof(double) * 64));
        CHECK_HIP(hipMemset(g_engine.d_sc, 0, sizeof(double) * 64));
        CHECK_HIP(hipHostMalloc((void **)&g_engine.h_sc, sizeof(double) * 64, hipHostMallocDefault));
    }
    CHECK_HIP(hipMalloc(&g_engine.d_part, sizeof(double) * 4 * np));
    g_engine.nparts = np;
}

template <int OP>
void eop(const Engine &e, long long N, const double2 *b, gpu_solver_workspace_t *w, double2 *x, int off) {
    const int np = (int)std::min<long long>(e.nparts, (N + kRedThreads - 1) / kRedThreads);
    hipLaunchKernelGGL(k_bicg<OP>, dim3(np), dim3(kRedThreads), 0, 0, N, b, (double2 *)w->r.x, (double2 *)w->r0.x,
                       (double2 *)w->p.x, (const double2 *)w->v.x, (double2 *)w->s.x, (const double2 *)w->t.x,
                       (const double2 *)w->y.x + off, (const double2 *)w->z.x + off, x, e.d_sc, e.d_part);
    if (OP != OP_P && OP != OP_S)
        hipLaunchKernelGGL(k_bicg_fin<OP>, dim3(1), dim3(kRedThreads), 0, 0, np, e.d_part, e.d_sc);
    CHECK_HIP(hipGetLastError());
}

void read_sc(const Engine &e) {
    CHECK_HIP(hipMemcpy(e.h_sc, e.d_sc, sizeof(double) * SC_COUNT, hipMemcpyDeviceToHost));
}

// The BiCGSTAB of tools/gpu_bicgstab.c:133-307 on the workspace w with the scalars and partials of e.
void bicgstab_run(gpu_spmv_fun spmv, const void *mat, const gpu_dense_vec_t *rhs, gpu_dense_vec_t *x, int nrow,
                  int local_col_beg, gpu_precond_fun psolve, const void *precond, gpu_solver_workspace_t *wsp,
                  const Engine &e, int *iters, double *tol_error) {
    const double tol = *tol_error;
    const int maxit = *iters;
    const long long N = nrow;
    const csr_data_t one(1, 0), zero(0, 0);
    double2 *xl = (double2 *)(x->x + local_col_beg);
    const double2 *b = (const double2 *)rhs->x;
    // r = rhs - A x, r0 = r
    spmv(mat, x, &wsp->r, one, zero);
    eop<OP_INIT>(e, N, b, wsp, xl, local_col_beg);
    read_sc(e);
    double r0_sq = e.h_sc[SC_R0SQ];
    const double rhs_sq = e.h_sc[SC_RHSSQ];
    if (rhs_sq == 0) {
        CHECK_HIP(hipMemset(x->x, 0, sizeof(csr_data_t) * x->dim));
        *iters = 0;
        *tol_error = 0;
        return;
    }
    const double tol2 = tol * tol * rhs_sq, eps2 = 1e-12 * 1e-12;
    if (r0_sq < tol2) {
        *iters = 0;
        *tol_error = std::sqrt(r0_sq / rhs_sq);
        return;
    }
    int i = 0;
    double rnorm2 = r0_sq, rho_abs = r0_sq;
    while (i < maxit) {
        if (rho_abs < eps2 * r0_sq) {
            spmv(mat, x, &wsp->r, one, zero);
            eop<OP_RESTART>(e, N, b, wsp, xl, local_col_beg);
            read_sc(e);
            r0_sq = e.h_sc[SC_R0SQ];
        }
        eop<OP_P>(e, N, b, wsp, xl, local_col_beg);
        psolve(precond, &wsp->p, &wsp->y);
        spmv(mat, &wsp->y, &wsp->v, one, zero);
        eop<OP_ALPHA>(e, N, b, wsp, xl, local_col_beg);
        eop<OP_S>(e, N, b, wsp, xl, local_col_beg);
        psolve(precond, &wsp->s, &wsp->z);
        spmv(mat, &wsp->z, &wsp->t, one, zero);
        eop<OP_W>(e, N, b, wsp, xl, local_col_beg);
        eop<OP_XR>(e, N, b, wsp, xl, local_col_beg);
        read_sc(e);
        rnorm2 = e.h_sc[SC_RNORM2];
        rho_abs = std::hypot(e.h_sc[SC_RHONEXT], e.h_sc[SC_RHONEXT + 1]);
        if (rnorm2 < tol2) break;
        ++i;
    }
    *tol_error = std::sqrt(rnorm2 / rhs_sq);
    *iters = i;
}

void alloc_vec(gpu_dense_vec_t *v, int n, int nrow, int off, bool local) {
    if (v->x) return;
    gpu_put_vec(v, nullptr, n);
    if (local) gpu_vec_local_part(v, nrow, off);
}

// Scratch of gpu_bicgstab_noalloc: the eight BiCGSTAB vectors (kNoAllocRows complex entries each), then the
// scalars and the reduction partials.  A file-scope device array, so the entry point never allocates; every
// GPU the library runs on holds its own copy until the process ends.
constexpr long long kNoAllocRows = 1 << 18;              // 262,144 rows: 32 MiB of vectors
constexpr int kNoAllocParts = 1024;                       // >= the k_bicg grid for kNoAllocRows rows
constexpr long long kNoAllocTail = 32 + 2 * kNoAllocParts;  // SC scalars (64 doubles), partials (4 per block)
static_assert((kNoAllocRows + kRedThreads - 1) / kRedThreads <= kNoAllocParts, "partials of the largest solve");
static __device__ __align__(256) double2 scratch[8 * kNoAllocRows + kNoAllocTail];

constexpr int kMaxDevices = 64;
double2 *g_scratch_ptr[kMaxDevices] = {};  // device address of 'scratch' on each device (host-side name: stand-in)
std::mutex g_scratch_mu;

// the device address of this device's scratch array, looked up once per device with that device current
int scratch_of_current_device(double2 **out) {
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kMaxDevices) {
        gade::set_error("gpu_bicgstab_noalloc: no current HIP device (or device index beyond 63)");
        return GADE_EHIP;
    }
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void *p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p) {
            gade::set_error(std::string("gpu_bicgstab_noalloc: hipGetSymbolAddress(scratch) failed: ") +
                            hipGetErrorString(e));
            return GADE_EHIP;
        }
        g_scratch_ptr[dev] = (double2 *)p;
    }
    *out = g_scratch_ptr[dev];
    return GADE_OK;
}

__global__ __launch_bounds__(256) void k_zero2(double2 *__restrict__ p, long long n) {
    for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x)
        p[i] = make_double2(0, 0);
}

// one model per process, as the reference keeps its matrices in file-static globals
struct DropinModel {
    int nmat = 0, nnzb = 0, n = 0, lmax = 0, nblk = 0, nang = 0;
    std::vector<int> rAp, rAi, hAp, hAi, perm;
    std::vector<std::vector<int>> HAp, HAi;
    std::vector<std::vector<double>> HAx, g, h0;
    bool have_g = false, have_rc = false;
    gade_model_t *m = nullptr;
} g_dm;

}  // namespace

extern "C" {

// ---------------------------------------------------------------- gpu_sparse.h
void gpu_sparse_init(void) {
    int device = 0;
    hipDeviceProp_t prop;
    CHECK_HIP(hipGetDevice(&device));
    CHECK_HIP(hipGetDeviceProperties(&prop, device));
    printf("rank 0: device %s pciBusID %d\n", prop.name, prop.pciBusID);
    fflush(stdout);
}

void gpu_sparse_fini(void) {}

void gpu_put_csr(gpu_sparse_csr_t *Agpu, const sparse_csr_t *Ahost) {
    const long long nrows = csr_nrows(Ahost), nnz = csr_nnz(Ahost);
    if (!Agpu->Ax) {
        CHECK_HIP(hipMalloc(&Agpu->Ap, sizeof(int) * (nrows + 1)));
        CHECK_HIP(hipMalloc(&Agpu->Ai, sizeof(int) * std::max(nnz, 1LL)));
        CHECK_HIP(hipMalloc(&Agpu->Ax, sizeof(csr_data_t) * std::max(nnz, 1LL)));
        Agpu->nrows = (int)nrows;
        Agpu->ncols = csr_ncols(Ahost);
        Agpu->nnz = (int)nnz;
        Agpu->desc = nullptr;
        Agpu->cuBuffer = nullptr;
        Agpu->spsvDescr = nullptr;
        CHECK_HIP(hipMemcpy(Agpu->Ap, Ahost->Ap, sizeof(int) * (nrows + 1), hipMemcpyHostToDevice));
        CHECK_HIP(hipMemcpy(Agpu->Ai, Ahost->Ai, sizeof(int) * nnz, hipMemcpyHostToDevice));
        Agpu->Ai_sub_map = nullptr;
        if (Ahost->Ai_sub_map) {
            CHECK_HIP(hipMalloc(&Agpu->Ai_sub_map, sizeof(int) * std::max(nnz, 1LL)));
            CHECK_HIP(hipMemcpy(Agpu->Ai_sub_map, Ahost->Ai_sub_map, sizeof(int) * nnz, hipMemcpyHostToDevice));
        }
    }
    if (Ahost->Ax) CHECK_HIP(hipMemcpy(Agpu->Ax, Ahost->Ax, sizeof(csr_data_t) * nnz, hipMemcpyHostToDevice));
}

void gpu_free_csr(gpu_sparse_csr_t *Agpu) {
    TriSched *s = (TriSched *)Agpu->desc;
    if (s) {
        CHECK_HIP(hipFree(s->d_rows));
        CHECK_HIP(hipFree(s->d_diag));
        delete s;
    }
    void *ptrs[] = {Agpu->Ap, Agpu->Ai, Agpu->Ax, Agpu->Ai_sub_map};
    for (void *p : ptrs)
        if (p) CHECK_HIP(hipFree(p));
    memset(Agpu, 0, sizeof(*Agpu));
}

void gpu_put_vec(gpu_dense_vec_t *xgpu, const csr_data_t *xhost, csr_index_t dim) {
    CHECK_HIP(hipMalloc(&xgpu->x, sizeof(csr_data_t) * std::max(dim, 1)));
    if (xhost)
        CHECK_HIP(hipMemcpy(xgpu->x, xhost, sizeof(csr_data_t) * dim, hipMemcpyHostToDevice));
    else
        CHECK_HIP(hipMemset(xgpu->x, 0, sizeof(csr_data_t) * dim));
    xgpu->dim = dim;
    xgpu->desc = nullptr;
    xgpu->desc_local = nullptr;
    xgpu->local_offset = 0;
}

void gpu_vec_local_part(gpu_dense_vec_t *xgpu, csr_index_t dim, csr_index_t local_offset) {
    xgpu->desc_local = xgpu;  // marks that the local part is defined
    xgpu->local_offset = local_offset;
    xgpu->local_dim = dim;
}

void gpu_get_vec(csr_data_t *xhost, const gpu_dense_vec_t *xgpu) {
    if (xgpu->local_dim)
        CHECK_HIP(hipMemcpy(xhost, xgpu->x + xgpu->local_offset, sizeof(csr_data_t) * xgpu->local_dim,
                            hipMemcpyDeviceToHost));
    else
        CHECK_HIP(hipMemcpy(xhost, xgpu->x, sizeof(csr_data_t) * xgpu->dim, hipMemcpyDeviceToHost));
}

void gpu_free_vec(gpu_dense_vec_t *xgpu) {
    if (xgpu->x) CHECK_HIP(hipFree(xgpu->x));
    xgpu->x = nullptr;
}

static void spmv_impl(gpu_sparse_csr_t *A, const csr_data_t *x, gpu_dense_vec_t *y, csr_data_t alpha, csr_data_t beta) {
    const long long N = A->nrows;
    hipLaunchKernelGGL(k_csr_spmv, dim3((unsigned)((N * 64 + 255) / 256)), dim3(256), 0, 0, A->Ap, A->Ai,
                       (const double2 *)A->Ax, (const double2 *)x, N, d2(alpha), d2(beta), (double2 *)y->x);
    CHECK_HIP(hipGetLastError());
}

void gpu_spmv(gpu_sparse_csr_t *Hfull, gpu_dense_vec_t *x, gpu_dense_vec_t *y, csr_data_t alpha, csr_data_t beta) {
    spmv_impl(Hfull, x->x, y, alpha, beta);
}

void gpu_spmv_local(gpu_sparse_csr_t *Hfull, gpu_dense_vec_t *x, gpu_dense_vec_t *y, csr_data_t alpha,
                    csr_data_t beta) {
    spmv_impl(Hfull, x->x + x->local_offset, y, alpha, beta);
}

// tools/gpu_sparse.c:315-340 walks the blocks and multiplies nothing
void gpu_spmv_block(sparse_csr_t H_blk, csr_data_t *x, csr_data_t *y, sparse_csr_t *g) {
    (void)H_blk; (void)x; (void)y; (void)g;
}

// tools/gpu_sparse.c:342-358: compares yfull with the (empty) blocked product; prints mismatches
void gpu_spmb_block_test(sparse_csr_t H_blk, csr_data_t *x, csr_data_t *yfull, sparse_csr_t *g) {
    (void)H_blk; (void)x; (void)g;
    long long bad = 0;
    for (long long i = 0; i < csr_nrows(&H_blk); i++)
        if (std::abs(yfull[i]) > 0) bad++;
    printf("gpu_spmb_block_test: %lld entries differ from the blocked product\n", bad);
}

// tools/gpu_sparse.c:361-388: yfull += Hfull x on the GPU
void gpu_spmv_test(sparse_csr_t Hfull, csr_data_t *x, csr_data_t *yfull) {
    gpu_sparse_csr_t A = {};
    gpu_put_csr(&A, &Hfull);
    gpu_dense_vec_t xg = {}, yg = {};
    gpu_put_vec(&xg, x, csr_ncols(&Hfull));
    gpu_put_vec(&yg, yfull, csr_nrows(&Hfull));
    gpu_spmv(&A, &xg, &yg, csr_data_t(1, 0), csr_data_t(1, 0));
    gpu_get_vec(yfull, &yg);
    gpu_free_csr(&A);
    gpu_free_vec(&xg);
    gpu_free_vec(&yg);
}

void gpu_lu_analyze(gpu_sparse_csr_t *L, gpu_sparse_csr_t *U, gpu_dense_vec_t *x, gpu_dense_vec_t *y,
                    gpu_dense_vec_t *temp) {
    (void)x; (void)y; (void)temp;
    if (!L->desc) L->desc = analyze(L, true);
    if (!U->desc) U->desc = analyze(U, false);
}

void gpu_lu_solve(const gpu_sparse_csr_t *L, const gpu_sparse_csr_t *U, const gpu_dense_vec_t *x, gpu_dense_vec_t *y,
                  gpu_dense_vec_t *temp) {
    trsv(L, (const double2 *)x->x, (double2 *)temp->x);
    double2 *out = (double2 *)(y->desc_local ? y->x + y->local_offset : y->x);
    trsv(U, (const double2 *)temp->x, out);
}

// ---------------------------------------------------------------- gpu_bicgstab.h
void gpu_blas_init(void) { engine_init(1); }

void gpu_bicgstab(gpu_spmv_fun spmv, const void *mat, const gpu_dense_vec_t *rhs, gpu_dense_vec_t *x, int nrow, int ncol,
                  int local_col_beg, gpu_precond_fun psolve, const void *precond, gpu_solver_workspace_t *wsp,
                  int *iters, double *tol_error) {
    engine_init(nrow);
    alloc_vec(&wsp->v, nrow, nrow, 0, false);
    alloc_vec(&wsp->p, nrow, nrow, 0, false);
    alloc_vec(&wsp->y, ncol, nrow, local_col_beg, true);
    alloc_vec(&wsp->z, ncol, nrow, local_col_beg, true);
    alloc_vec(&wsp->s, nrow, nrow, 0, false);
    alloc_vec(&wsp->t, nrow, nrow, 0, false);
    alloc_vec(&wsp->r, nrow, nrow, 0, false);
    alloc_vec(&wsp->r0, nrow, nrow, 0, false);
    bicgstab_run(spmv, mat, rhs, x, nrow, local_col_beg, psolve, precond, wsp, g_engine, iters, tol_error);
}

int gpu_bicgstab_noalloc(gpu_spmv_fun spmv, const void *mat, const gpu_dense_vec_t *rhs, gpu_dense_vec_t *x, int nrow,
                         int ncol, int local_col_beg, gpu_precond_fun psolve, const void *precond, int *iters,
                         double *tol_error) {
    if (!spmv || !psolve || !rhs || !x || !iters || !tol_error || nrow < 0 || ncol < nrow || local_col_beg < 0 ||
        local_col_beg + nrow > ncol) {
        gade::set_error("gpu_bicgstab_noalloc: invalid arguments");
        return GADE_EINVAL;
    }
    if (ncol > kNoAllocRows) {
        gade::set_error("gpu_bicgstab_noalloc: " + std::to_string(ncol) + " columns exceed the scratch maximum of " +
                        std::to_string(kNoAllocRows) + " (use gpu_bicgstab)");
        return GADE_EINVAL;
    }
    double2 *base = nullptr;
    GADE_TRY(scratch_of_current_device(&base));
    // the eight vectors at fixed slots, then the scalars and partials; a fresh workspace is all zeros
    gpu_solver_workspace_t w;
    memset(&w, 0, sizeof w);
    gpu_dense_vec_t *slots[8] = {&w.r, &w.r0, &w.v, &w.p, &w.y, &w.z, &w.s, &w.t};
    for (int k = 0; k < 8; k++) {
        const bool local = slots[k] == &w.y || slots[k] == &w.z;
        slots[k]->x = (gpu_complex_t *)(base + (size_t)k * kNoAllocRows);
        slots[k]->dim = local ? ncol : nrow;
        slots[k]->local_dim = local ? nrow : 0;  // as gpu_put_vec leaves a zeroed workspace's vectors
        slots[k]->local_offset = local ? local_col_beg : 0;
        slots[k]->desc_local = local ? (void *)slots[k] : nullptr;  // marks the local view, as gpu_vec_local_part
    }
    double h_sc[SC_STRIDE];
    Engine e;
    e.d_sc = (double *)(base + 8 * kNoAllocRows);
    e.d_part = e.d_sc + 64;
    e.h_sc = h_sc;
    e.nparts = kNoAllocParts;
    for (int k = 0; k < 8; k++) {
        const long long n = slots[k]->dim;
        if (n) hipLaunchKernelGGL(k_zero2, dim3((unsigned)std::min<long long>(1024, (n + 255) / 256)), dim3(256), 0, 0,
                                  (double2 *)slots[k]->x, n);
    }
    hipLaunchKernelGGL(k_zero2, dim3(8), dim3(256), 0, 0, (double2 *)e.d_sc, kNoAllocTail);
    GADE_HIP(hipGetLastError());
    bicgstab_run(spmv, mat, rhs, x, nrow, local_col_beg, psolve, precond, &w, e, iters, tol_error);
    return GADE_OK;
}

// ---------------------------------------------------------------- gpu_solver.h
void gpu_init_model_matrices(int nmatrices, const sparse_csr_t *cpu_g, const sparse_csr_t *cpu_gt,
                             const sparse_csr_t *cpu_h0) {
    (void)cpu_gt;  // the conjugate transposes are formed on the device from g
    if (nmatrices <= 0 || nmatrices % 24) DIE("gpu_init_model_matrices: expected 24*lmax matrices, got %d", nmatrices);
    g_dm.nmat = nmatrices;
    g_dm.n = cpu_g[0].nrows;
    g_dm.nnzb = cpu_g[0].nnz;
    g_dm.rAp.assign(cpu_g[0].Ap, cpu_g[0].Ap + g_dm.n + 1);
    g_dm.rAi.assign(cpu_g[0].Ai, cpu_g[0].Ai + g_dm.nnzb);
    g_dm.g.resize(nmatrices);
    for (int i = 0; i < nmatrices; i++) {
        if (cpu_g[i].nnz != g_dm.nnzb) DIE("gpu_init_model_matrices: g matrices do not share one pattern");
        const double *v = (const double *)cpu_g[i].Ax;
        g_dm.g[i].assign(v, v + 2 * (size_t)g_dm.nnzb);
    }
    g_dm.h0.resize(4);
    for (int i = 0; i < 4; i++) {
        const double *v = (const double *)cpu_h0[i].Ax;
        g_dm.h0[i].assign(v, v + 2 * (size_t)g_dm.nnzb);
    }
    g_dm.have_g = true;
}

void gpu_compute_row_col(int lmax, const sparse_csr_t *H, sparse_csr_t *Hfull_blk, sparse_csr_t *Hfull) {
    (void)Hfull;
    g_dm.lmax = lmax;
    g_dm.nblk = Hfull_blk->nrows;
    g_dm.hAp.assign(Hfull_blk->Ap, Hfull_blk->Ap + g_dm.nblk + 1);
    g_dm.hAi.assign(Hfull_blk->Ai, Hfull_blk->Ai + Hfull_blk->nnz);
    g_dm.perm.clear();
    if (Hfull_blk->perm) g_dm.perm.assign(Hfull_blk->perm, Hfull_blk->perm + g_dm.nblk);
    g_dm.nang = H[0].nrows;
    g_dm.HAp.resize(2 * lmax);
    g_dm.HAi.resize(2 * lmax);
    g_dm.HAx.resize(2 * lmax);
    for (int i = 0; i < 2 * lmax; i++) {
        g_dm.HAp[i].assign(H[i].Ap, H[i].Ap + H[i].nrows + 1);
        g_dm.HAi[i].assign(H[i].Ai, H[i].Ai + H[i].nnz);
        const double *v = (const double *)H[i].Ax;
        g_dm.HAx[i].assign(v, v + 2 * (size_t)H[i].nnz);
    }
    g_dm.have_rc = true;
}

void gpu_compute_timedep_matrices(double h, double dt, csr_data_t *ft, int lmax, const sparse_csr_t *Hfull_blk,
                                  sparse_csr_t *Hfull, gpu_sparse_csr_t *gpu_Hfull) {
    (void)Hfull_blk;
    if (!g_dm.have_g || !g_dm.have_rc) DIE("gpu_compute_timedep_matrices: call gpu_init_model_matrices and gpu_compute_row_col first");
    if (lmax != g_dm.lmax) DIE("gpu_compute_timedep_matrices: lmax %d differs from gpu_compute_row_col (%d)", lmax, g_dm.lmax);
    if (!g_dm.m) {
        std::vector<const int *> HAp(2 * lmax), HAi(2 * lmax);
        std::vector<const double *> HAx(2 * lmax), gp(g_dm.nmat), hp(4), sp(3);
        for (int i = 0; i < 2 * lmax; i++) {
            HAp[i] = g_dm.HAp[i].data();
            HAi[i] = g_dm.HAi[i].data();
            HAx[i] = g_dm.HAx[i].data();
        }
        for (int i = 0; i < g_dm.nmat; i++) gp[i] = g_dm.g[i].data();
        for (int i = 0; i < 4; i++) hp[i] = g_dm.h0[i].data();
        std::vector<double> szero(2 * (size_t)g_dm.nnzb, 0.0);  // S is not part of the assembled Hfull
        for (int i = 0; i < 3; i++) sp[i] = szero.data();
        gade_inputs_t in = {};
        in.nblk = g_dm.nblk;
        in.hall_Ap = g_dm.hAp.data();
        in.hall_Ai = g_dm.hAi.data();
        in.hall_perm = g_dm.perm.empty() ? nullptr : g_dm.perm.data();
        in.lmax = lmax;
        in.nang = g_dm.nang;
        in.H_Ap = HAp.data();
        in.H_Ai = HAi.data();
        in.H_Ax = HAx.data();
        in.blkdim = g_dm.n;
        in.rad_Ap = g_dm.rAp.data();
        in.rad_Ai = g_dm.rAi.data();
        in.g = gp.data();
        in.h0 = hp.data();
        in.s = sp.data();
        CHECK_GADE(gade_model_create(&in, h, dt, nullptr, &g_dm.m));
        long long nf = 0;
        gade_model_info(g_dm.m, nullptr, nullptr, nullptr, nullptr, &nf, nullptr);
        if (nf != csr_nnz(Hfull)) DIE("gpu_compute_timedep_matrices: Hfull has %d non-zeros, the blocks give %lld", csr_nnz(Hfull), nf);
    }
    if (!gpu_Hfull->Ax) gpu_put_csr(gpu_Hfull, Hfull);
    CHECK_GADE(gade_set_time(g_dm.m, (const double *)ft));
    CHECK_GADE(gade_assembled_timedep_into(g_dm.m, gpu_Hfull->Ap, (double *)gpu_Hfull->Ax));
    CHECK_HIP(hipDeviceSynchronize());
}

// ---------------------------------------------------------------- block LU of solver.h
int gade_blocklu_create(int nblk, int blkdim, const int *Ap, const int *Ai, const double *Ax, gade_blocklu_t **out) {
    if (nblk <= 0 || blkdim <= 0 || !Ap || !Ai || !Ax || !out) return GADE_EINVAL;
    const int n = blkdim;
    // pattern of block 0, which every block must repeat
    std::vector<int> rAp(n + 1, 0), rAi;
    for (int a = 0; a < n; a++) {
        for (int k = Ap[a]; k < Ap[a + 1]; k++) rAi.push_back(Ai[k]);
        rAp[a + 1] = (int)rAi.size();
    }
    const int nnzb = rAp[n];
    for (int r = 0; r < nblk; r++)
        for (int a = 0; a < n; a++) {
            const int row = r * n + a;
            if (Ap[row + 1] - Ap[row] != rAp[a + 1] - rAp[a]) {
                gade::set_error("block LU: diagonal blocks do not share one pattern");
                return GADE_EINVAL;
            }
            for (int k = 0; k < rAp[a + 1] - rAp[a]; k++)
                if (Ai[Ap[row] + k] != r * n + rAi[rAp[a] + k]) {
                    gade::set_error("block LU: matrix is not block diagonal with the given block size");
                    return GADE_EINVAL;
                }
        }
    // identical blocks share one factorisation
    std::vector<int> blk_q(nblk), rep;
    std::vector<double2> Pq;
    for (int r = 0; r < nblk; r++) {
        const double *vals = Ax + 2 * (size_t)Ap[r * n];
        int found = -1;
        for (size_t q = 0; q < rep.size(); q++)
            if (!memcmp(vals, Ax + 2 * (size_t)Ap[rep[q] * n], sizeof(double) * 2 * nnzb)) {
                found = (int)q;
                break;
            }
        if (found < 0) {
            found = (int)rep.size();
            rep.push_back(r);
            for (int k = 0; k < nnzb; k++) Pq.push_back(make_double2(vals[2 * k], vals[2 * k + 1]));
        }
        blk_q[r] = found;
    }
    gade_blocklu *lu = nullptr;
    int rc = gade::blocklu_build(nblk, n, rAp.data(), rAi.data(), (int)rep.size(), blk_q.data(), Pq, nullptr, &lu);
    if (rc) return rc;
    *out = lu;
    return GADE_OK;
}

int gade_blocklu_solve(const gade_blocklu_t *lu, const double *b, double *x) {
    GADE_TRY(gade::blocklu_solve_async(lu, (const double2 *)b, (double2 *)x, nullptr));
    GADE_HIP(hipDeviceSynchronize());
    return GADE_OK;
}

int gade_blocklu_destroy(gade_blocklu_t *lu) {
    gade::blocklu_free(lu);
    return GADE_OK;
}

}  // extern "C"
