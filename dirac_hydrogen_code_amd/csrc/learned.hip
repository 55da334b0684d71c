// Learned initial guess for the CN solve (no reference counterpart).
//
// A small MLP (3 hidden layers of 256, GELU, BF16 weights, FP32 accumulation on
// MFMA) predicts the initial guess of every step from f_alpha(t_{n+1}), the last
// six log residual norms, the sizes of the backward differences of the solution
// history and its last two higher differences in a POD basis:
//
//   x0 = x2 + s * sum_{k=0..5} w_k D_k / |D_k| + s * U c
//
// with D_k = nabla^k psi_n (backward differences of psi_n..psi_{n-5}),
// x2 = psi_n + D_1 + D_2 the order-2 polynomial extrapolation, s = |D_3| (the
// size of its leading error term), U the npod POD vectors of extrapolation
// defects (top left singular vectors of solver snapshots) and (w, c) the
// network's outputs (6 real mixing weights, npod complex POD coefficients).  A
// zero output is the extrapolation itself, which is also the fallback while the
// history is short or an output is not finite.  The converged solve at the
// unchanged tolerance judges every guess: only the iteration count can change.
//
// Training minimises the relative squared error |psi_{n+1} - x0|^2 / |e2|^2
// (e2 = psi_{n+1} - x2), which is a quadratic form in the outputs; each training
// sample therefore holds the small projected quantities of that form (G, a, P, q,
// beta) next to its inputs, and no full vector is touched while training.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>
#include <cstring>
#include <random>
#include <string>
#include <vector>

#include "gade_device.h"
#include "gade_internal.h"

using gade::hip_alloc;
using gade::set_error;

namespace gade {
int svd_left_jacobi(int k, const double *Rcm, double *S, double *Ucm);  // host/eigsolve.cpp
}

namespace {

constexpr int kIn = GADE_LG_IN, kHid = GADE_LG_HIDDEN, kOut = GADE_LG_OUT, kPod = GADE_LG_POD;
constexpr int kSample = GADE_LG_SAMPLE;
constexpr int kNV = 7;  // vectors of the difference stage: D_0..D_5, e2 (record)
constexpr int kHhChunks = 64;  // row chunks of a reflector's dot products
constexpr int kNG = kNV * (kNV + 1) / 2;  // Gram entries Re(V_i^H V_j), i <= j
// sample layout (floats)
constexpr int S_FEAT = 0, S_G = kIn, S_A = S_G + 36, S_P = S_A + 6, S_Q = S_P + kPod * 6 * 2, S_BETA = S_Q + kPod * 2,
              S_VALID = S_BETA + 1;
static_assert(S_VALID < kSample, "sample layout");
// feature layout
constexpr int F_F = 0, F_LRES = 6, F_DN = 12, F_Z3 = 17, F_Z4 = F_Z3 + 2 * kPod, F_END = F_Z4 + 2 * kPod;
static_assert(F_END <= kIn, "feature layout");
// parameter layout (floats): W1 b1 W2 b2 W3 b3 W4 b4, W_l row-major [out][in]
constexpr size_t P_W1 = 0, P_B1 = P_W1 + (size_t)kHid * kIn, P_W2 = P_B1 + kHid, P_B2 = P_W2 + (size_t)kHid * kHid,
                 P_W3 = P_B2 + kHid, P_B3 = P_W3 + (size_t)kHid * kHid, P_W4 = P_B3 + kHid,
                 P_B4 = P_W4 + (size_t)kOut * kHid, P_TOTAL = P_B4 + kOut;
constexpr int kLayers = 4;
const size_t kWoff[kLayers] = {P_W1, P_W2, P_W3, P_W4}, kBoff[kLayers] = {P_B1, P_B2, P_B3, P_B4};
const int kLin[kLayers] = {kIn, kHid, kHid, kHid}, kLout[kLayers] = {kHid, kHid, kHid, kOut};

__host__ __device__ inline int pad16(int x) { return (x + 15) & ~15; }
inline unsigned grid_for_lg(long long n) {
    long long g = (n + 255) / 256;
    return (unsigned)(g < 1 ? 1 : (g > 2048 ? 2048 : g));
}

// ---------------------------------------------------------------------------
// bf16 (round to nearest even) and GELU
// ---------------------------------------------------------------------------
__device__ __forceinline__ uint16_t f2bf(float x) {
    uint32_t u = __float_as_uint(x);
    if ((u & 0x7f800000u) == 0x7f800000u) return (uint16_t)((u >> 16) | ((u & 0xffffu) ? 0x40u : 0u));
    u += 0x7fffu + ((u >> 16) & 1u);
    return (uint16_t)(u >> 16);
}
__device__ __forceinline__ float gelu(float z) { return 0.5f * z * (1.0f + erff(z * 0.70710678118654752f)); }
__device__ __forceinline__ float gelu_d(float z) {
    return 0.5f * (1.0f + erff(z * 0.70710678118654752f)) + z * 0.39894228040143268f * expf(-0.5f * z * z);
}

// ---------------------------------------------------------------------------
// BF16 MFMA GEMM: C(m,n) = sum_k A(m,k) B(k,n), FP32 accumulation, with
// A(m,k) = a[m*sam + k*sak], B(k,n) = b[k*sbk + n*sbn] (bf16 bits).  One block
// per 16x16 tile of C (v_mfma_f32_16x16x16_bf16: lane l holds A[l%16][4(l/16)..+3],
// B[4(l/16)..+3][l%16] and C[4(l/16)+i][l%16]).  The block's KS waves take the
// 16-wide k steps in turn (wave w: steps w, w + KS, ...) and their accumulators
// are added in wave order; KS = 1 for the training sizes, KS = 4 for the few
// rows of inference (short per-wave chains).  M, N, K multiples of 16.
// ---------------------------------------------------------------------------
typedef short bf4 __attribute__((ext_vector_type(4)));
typedef float fl4 __attribute__((ext_vector_type(4)));
enum { EPI_STORE = 0, EPI_BIAS = 1, EPI_BIAS_GELU = 2 };

template <int EPI, int KS = 1>
__global__ __launch_bounds__(64 * KS) void k_gemm_bf16(int M, int N, int K, const uint16_t *__restrict__ a, long sam,
                                                       long sak, const uint16_t *__restrict__ b, long sbk, long sbn,
                                                       float *__restrict__ c, int ldc, const float *__restrict__ bias,
                                                       uint16_t *__restrict__ act, int ldact) {
    const int tiles_n = N / 16;
    const int tm = blockIdx.x / tiles_n, tn = blockIdx.x % tiles_n;
    if (tm * 16 >= M) return;
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6, r = lane & 15, g = lane >> 4;
    const uint16_t *ap = a + (long)(tm * 16 + r) * sam;
    const uint16_t *bp = b + (long)(tn * 16 + r) * sbn;
    fl4 acc = {0.f, 0.f, 0.f, 0.f};
    if (sak == 1 && sbk == 1) {
        // both operands contiguous along k (the forward layers): each lane's four k are one 8-byte load
        for (int k0 = 16 * w; k0 < K; k0 += 16 * KS) {
            bf4 av, bv;
            __builtin_memcpy(&av, reinterpret_cast<const uint2 *>(ap + k0 + 4 * g), sizeof(bf4));
            __builtin_memcpy(&bv, reinterpret_cast<const uint2 *>(bp + k0 + 4 * g), sizeof(bf4));
            acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(av, bv, acc, 0, 0, 0);
        }
    } else {
        for (int k0 = 16 * w; k0 < K; k0 += 16 * KS) {
            bf4 av, bv;
#pragma unroll
            for (int j = 0; j < 4; j++) {
                const long kk = k0 + 4 * g + j;
                av[j] = (short)ap[kk * sak];
                bv[j] = (short)bp[kk * sbk];
            }
            acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(av, bv, acc, 0, 0, 0);
        }
    }
    if constexpr (KS > 1) {
        __shared__ fl4 red[KS][64];
        red[w][lane] = acc;
        __syncthreads();
        if (w != 0) return;
#pragma unroll
        for (int q = 1; q < KS; q++) acc += red[q][lane];
    }
#pragma unroll
    for (int i = 0; i < 4; i++) {
        const int m = tm * 16 + 4 * g + i, n = tn * 16 + r;
        float v = acc[i];
        if (EPI != EPI_STORE) v += bias[n];
        c[(long)m * ldc + n] = v;
        if (EPI == EPI_BIAS_GELU) act[(long)m * ldact + n] = f2bf(gelu(v));
    }
}

// ---------------------------------------------------------------------------
// difference stage: V_0 = psi_n, V_k = nabla^k psi_n (k = 1..5), V_6 = e2 (record:
// psi_{n+1} - (psi_n + V_1 + V_2)), and the partial Gram sums Re(V_i^H V_j).
// Inference (rec = 0, before psi_n enters the ring): psi_n = psi, psi_{n-k} in
// ring slot (have - k) mod S.  Record (rec = 1, after): psi_n in slot (have - 1),
// psi_{n-k} in slot (have - 1 - k), psi_{n+1} = psi.
// ---------------------------------------------------------------------------
// Inference (REC false) uses only the norms |V_k|^2: the off-diagonal Gram partials are written as zeros.
template <bool REC>
__global__ __launch_bounds__(256) void k_lg_diff(long long N1, long long Ntot, const double2 *__restrict__ psi,
                                                 const double2 *__restrict__ hist, const int *__restrict__ ctl0,
                                                 double2 *__restrict__ V, double *__restrict__ part0,
                                                 const int *gate, int gstride) {
    constexpr int rec = REC ? 1 : 0;
    const int b = blockIdx.y;
    GADE_GATED_B(gate, gstride, b);
    __shared__ double sh[kNG * 4];
    const int have = ctl0[(size_t)b * CT_STRIDE + CT_HIST];
    const double2 *h[6];
#pragma unroll
    for (int k = 0; k < 6; k++) {
        const int back = rec ? k + 1 : k;
        if (!rec && k == 0) {
            h[k] = psi;
        } else {
            const int sl = (((have - back) % kHistSlots) + kHistSlots) % kHistSlots;
            h[k] = hist + (size_t)sl * Ntot;
        }
    }
    double acc[kNG];
#pragma unroll
    for (int t = 0; t < kNG; t++) acc[t] = 0;
    const int nv = rec ? kNV : kNV - 1;
    for (long long j = blockIdx.x * (long long)blockDim.x + threadIdx.x; j < N1; j += (long long)gridDim.x * blockDim.x) {
        const long long i = (long long)b * N1 + j;
        double2 x[6];
#pragma unroll
        for (int k = 0; k < 6; k++) x[k] = h[k][i];
        // backward differences nabla^k psi_n = sum_j (-1)^j binom(k, j) psi_{n-j}, by repeated differencing
        double2 v[kNV];
        v[0] = x[0];
#pragma unroll
        for (int k = 1; k < 6; k++) {
#pragma unroll
            for (int q = 5; q >= k; q--) x[q] = csub(x[q - 1], x[q]);
            v[k] = x[k];
        }
        v[6] = rec ? csub(psi[i], cadd(v[0], cadd(v[1], v[2]))) : make_double2(0, 0);
        // D_0 = psi_n is kept only when recording (its projection is a sample input); inference reads psi
#pragma unroll
        for (int k = 0; k < kNV; k++)
            if (k < nv && (rec || k > 0)) V[(size_t)k * Ntot + i] = v[k];
        int t = 0;
#pragma unroll
        for (int p = 0; p < kNV; p++)
#pragma unroll
            for (int q = p; q < kNV; q++, t++)
                if (REC || q == p) acc[t] += v[p].x * v[q].x + v[p].y * v[q].y;
    }
    // block reduction of the kNG sums (the used ones)
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    {
        int t = 0;
#pragma unroll
        for (int p = 0; p < kNV; p++)
#pragma unroll
            for (int q = p; q < kNV; q++, t++) {
                const double s = (REC || q == p) ? wave_sum(acc[t]) : 0.0;
                if (lane == 0) sh[wid * kNG + t] = s;
            }
    }
    __syncthreads();
    if (threadIdx.x < kNG) {
        double s = 0;
        for (int w = 0; w < (int)(blockDim.x >> 6); w++) s += sh[w * kNG + threadIdx.x];
        part0[((size_t)b * gridDim.x + blockIdx.x) * kNG + threadIdx.x] = s;
    }
}

// gram[b][e] = sum of the block partials: eight interleaved groups of partials, then the groups in order
// (lanes e = t % 32, group t / 32; 256 threads)
__global__ __launch_bounds__(256) void k_lg_gram_fin(int nparts, const double *__restrict__ part0,
                                                     double *__restrict__ gram, const int *gate, int gstride) {
    const int b = blockIdx.x;
    GADE_GATED_B(gate, gstride, b);
    __shared__ double sh[8][32];
    const int e = threadIdx.x & 31, g = threadIdx.x >> 5;
    double s = 0;
    if (e < kNG)
#pragma unroll 4
        for (int p = g; p < nparts; p += 8) s += part0[((size_t)b * nparts + p) * kNG + e];
    sh[g][e] = s;
    __syncthreads();
    if (g == 0 && e < kNG) {
        double t = sh[0][e];
#pragma unroll
        for (int q = 1; q < 8; q++) t += sh[q][e];
        gram[(size_t)b * kNG + e] = t;
    }
}

__device__ __forceinline__ int gidx(int p, int q) {  // index of Re(V_p^H V_q), p <= q
    return p * kNV - p * (p - 1) / 2 + (q - p);
}

// partial[b][c][v][j] = sum over chunk c of conj(UT[k][j]) V_{v0+v}[b][k]: one block per (chunk, b), lanes
// (j = t % 64, s = t / 64), UT the k-major FP32 copy of the basis (coalesced over j; the products and sums in
// FP64), V shared by the 64 lanes
// Grid (B, nchunk): the instances of one chunk are neighbouring blocks, so the chunk of the basis is read from
// memory once and from cache for the other instances.
constexpr int kProjChunk = 256;  // shortest chunk
inline int proj_chunks(long long N1, int B) {  // chunks per instance: ~4k blocks in all, 64..1024 per instance
    const long long want = std::max(64, std::min(1024, 4096 / std::max(1, B)));
    return (int)std::max<long long>(1, std::min<long long>(want, (N1 + kProjChunk - 1) / kProjChunk));
}
template <int NV>
__global__ __launch_bounds__(256) void k_lg_proj_part(long long N1, long long Ntot, int npod,
                                                      const float2 *__restrict__ UT, const double2 *__restrict__ V,
                                                      int v0, double2 *__restrict__ part, const int *gate,
                                                      int gstride, long long clen) {
    const int b = blockIdx.x, c = blockIdx.y, nch = gridDim.y;
    GADE_GATED_B(gate, gstride, b);
    __shared__ double2 sh[4][NV][64];
    const int j = threadIdx.x & 63, sl = threadIdx.x >> 6;
    double2 acc[NV];
#pragma unroll
    for (int v = 0; v < NV; v++) acc[v] = make_double2(0, 0);
    // wave sl takes a contiguous quarter of the chunk; its V values are loaded 64 at a time (one per lane) and
    // broadcast with lane reads, so only the basis loads are in the inner loop
    const long long cb = (long long)c * clen, ce = cb + clen < N1 ? cb + clen : N1;
    const long long q = (clen + 3) / 4;
    const long long k0 = cb + sl * q, k1 = k0 + q < ce ? k0 + q : ce;
    const int jj = j < npod ? j : npod - 1;
    for (long long kb = k0; kb < k1; kb += 64) {
        const int cnt = (int)(k1 - kb < 64 ? k1 - kb : 64);
        double2 vv[NV];
#pragma unroll
        for (int v = 0; v < NV; v++)
            vv[v] = j < cnt ? V[(size_t)(v0 + v) * Ntot + (size_t)b * N1 + kb + j] : make_double2(0, 0);
        // sixteen basis loads in flight before they are used (clamped to the run, zero-weighted beyond it)
        for (int i0 = 0; i0 < cnt; i0 += 16) {
            float2 ub[16];
#pragma unroll
            for (int u = 0; u < 16; u++) ub[u] = UT[(size_t)(kb + min(i0 + u, cnt - 1)) * npod + jj];
#pragma unroll
            for (int u = 0; u < 16; u++) {
                const int i = i0 + u;
                const double2 uu = i < cnt ? to_f64(ub[u]) : make_double2(0, 0);
#pragma unroll
                for (int v = 0; v < NV; v++) {
                    const double2 x = make_double2(__shfl(vv[v].x, i & 63, 64), __shfl(vv[v].y, i & 63, 64));
                    acc[v] = cfmaconj(uu, x, acc[v]);
                }
            }
        }
    }
#pragma unroll
    for (int v = 0; v < NV; v++) sh[sl][v][j] = acc[v];
    __syncthreads();
    if (sl == 0 && j < npod)
#pragma unroll
        for (int v = 0; v < NV; v++)
            part[(((size_t)b * nch + c) * kNV + v) * 64 + j] =
                cadd(cadd(sh[0][v][j], sh[1][v][j]), cadd(sh[2][v][j], sh[3][v][j]));
}

// out[b][v][j] = sum over chunks of the partials in a fixed order: sixteen interleaved chunk groups, then the
// groups in order.  Grid (B, nv, 4): a block takes 16 of the 64 basis indices, lanes (j, group).
__global__ __launch_bounds__(256) void k_lg_proj_fin(int nch, int npod, int nv, const double2 *__restrict__ part,
                                                     double2 *__restrict__ out, const int *gate, int gstride) {
    const int b = blockIdx.x, v = blockIdx.y;
    GADE_GATED_B(gate, gstride, b);
    __shared__ double2 sh[16][16];
    const int jl = threadIdx.x & 15, g = threadIdx.x >> 4, j = blockIdx.z * 16 + jl;
    double2 s = make_double2(0, 0);
#pragma unroll 4
    for (int c = g; c < nch; c += 16) s = cadd(s, part[(((size_t)b * nch + c) * kNV + v) * 64 + j]);
    sh[g][jl] = s;
    __syncthreads();
    if (g == 0 && j < npod) {
        double2 t = sh[0][jl];
#pragma unroll
        for (int q = 1; q < 16; q++) t = cadd(t, sh[q][jl]);
        out[((size_t)b * kNV + v) * npod + j] = t;
    }
}

// squared norm of snapshot column c (one block per column)
__global__ __launch_bounds__(256) void k_lg_colnorm(long long n, const double2 *__restrict__ X, double *__restrict__ out) {
    __shared__ double sh[4 * 16];
    const double2 *x = X + (size_t)blockIdx.x * n;
    double acc[1] = {0};
    for (long long i = threadIdx.x; i < n; i += blockDim.x) acc[0] += x[i].x * x[i].x + x[i].y * x[i].y;
    block_sum<1>(acc, sh);
    if (threadIdx.x == 0) out[blockIdx.x] = acc[0];
}

// ---------------------------------------------------------------------------
// Householder QR of the snapshot matrix for the POD (n x k, column-major), in the LAPACK convention
// (zgeqr2 / zlarfg): H_j = I - tau_j v_j v_j^H with v_j = (0 .. 0, 1, A[j+1:n][j]) stored below the diagonal
// and A[j][j] = beta real; the trailing columns take H_j^H = I - conj(tau_j) v_j v_j^H.  Q C = H_0 H_1 ..
// H_{k-1} C applies the reflectors last to first.  Every sum runs in a fixed order (chunk partials added
// in chunk order), so the factorisation is deterministic.
// ---------------------------------------------------------------------------

// reflector of column j: one workgroup of 1024 threads
__global__ __launch_bounds__(1024) void k_hh_reflector(double2 *__restrict__ A, long long n, int j,
                                                       double2 *__restrict__ tau) {
    __shared__ double sh[4 * 16];
    __shared__ double2 scale_sh;
    double2 *x = A + (size_t)j * n;
    double acc[1] = {0};
    for (long long r = j + 1 + threadIdx.x; r < n; r += blockDim.x) acc[0] += x[r].x * x[r].x + x[r].y * x[r].y;
    block_sum<1>(acc, sh);
    if (threadIdx.x == 0) {
        const double2 al = x[j];
        double2 t = make_double2(0, 0), sc = make_double2(0, 0);
        if (acc[0] > 0 || al.y != 0) {
            const double beta = -copysign(sqrt(al.x * al.x + al.y * al.y + acc[0]), al.x);
            t = make_double2((beta - al.x) / beta, -al.y / beta);
            sc = cdiv(make_double2(1, 0), make_double2(al.x - beta, al.y));
            x[j] = make_double2(beta, 0);
        }
        tau[j] = t;
        scale_sh = sc;
    }
    __syncthreads();
    const double2 sc = scale_sh;
    if (sc.x != 0 || sc.y != 0)
        for (long long r = j + 1 + threadIdx.x; r < n; r += blockDim.x) x[r] = cmul(x[r], sc);
}

// part[c][ch] = sum over the rows of chunk ch (rows >= j) of conj(v_r) M[r][c0 + c]
__global__ __launch_bounds__(256) void k_hh_dots(const double2 *__restrict__ v, long long n, int j,
                                                 const double2 *__restrict__ M, int c0, long long rows_per, int nch,
                                                 double2 *__restrict__ part) {
    __shared__ double sh[4 * 16];
    const int c = blockIdx.x, ch = blockIdx.y;
    const double2 *mc = M + (size_t)(c0 + c) * n;
    const long long r0 = j + (long long)ch * rows_per, r1 = min(n, r0 + rows_per);
    double acc[2] = {0, 0};
    for (long long r = r0 + threadIdx.x; r < r1; r += blockDim.x) {
        const double2 vr = r == j ? make_double2(1, 0) : v[r];
        const double2 mr = mc[r];
        acc[0] += vr.x * mr.x + vr.y * mr.y;
        acc[1] += vr.x * mr.y - vr.y * mr.x;
    }
    block_sum<2>(acc, sh);
    if (threadIdx.x == 0) part[(size_t)c * nch + ch] = make_double2(acc[0], acc[1]);
}

// M[r][c0 + c] -= v_r w_c for rows r >= j, w_c = t sum_ch part[c][ch], t = conj(tau_j) (R) or tau_j (Q C)
__global__ __launch_bounds__(256) void k_hh_update(const double2 *__restrict__ v, long long n, int j,
                                                   double2 *__restrict__ M, int c0, const double2 *__restrict__ part,
                                                   int nch, const double2 *__restrict__ tau, int conj_tau,
                                                   long long rows_per) {
    __shared__ double2 w_sh;
    const int c = blockIdx.x;
    if (threadIdx.x == 0) {
        double2 sum = make_double2(0, 0);
        for (int ch = 0; ch < nch; ch++) sum = cadd(sum, part[(size_t)c * nch + ch]);
        double2 t = tau[j];
        if (conj_tau) t.y = -t.y;
        w_sh = cmul(t, sum);
    }
    __syncthreads();
    const double2 w = w_sh;
    double2 *mc = M + (size_t)(c0 + c) * n;
    const long long r0 = j + (long long)blockIdx.y * rows_per, r1 = min(n, r0 + rows_per);
    for (long long r = r0 + threadIdx.x; r < r1; r += blockDim.x) {
        const double2 vr = r == j ? make_double2(1, 0) : v[r];
        mc[r] = csub(mc[r], cmul(vr, w));
    }
}

// the per-step copies of the basis, rounded to FP32: U32[j][k] = UT32[k][j] = U[j][k].  Each step reads
// the basis twice (projections, guess); FP32 halves those reads.  The correction s U c it forms is a small
// part of the guess and the converged solve at the unchanged tolerance judges every guess, so the rounding
// (6e-8 relative in U) changes no result beyond the solve's own tolerance.
__global__ void k_lg_basis32(long long N1, int npod, const double2 *__restrict__ U, float2 *__restrict__ U32,
                             float2 *__restrict__ UT32) {
    for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < N1 * npod; i += (long long)gridDim.x * blockDim.x) {
        const long long k = i / npod;
        const int j = (int)(i % npod);
        const float2 u = to_f32(U[(size_t)j * N1 + k]);
        UT32[i] = u;
        U32[(size_t)j * N1 + k] = u;
    }
}

// The inputs of one step, one block of 256 threads per instance:
// 1. the residual ring [b][8] (newest first) takes the previous solve's final residual relative to its
//    requested bound tol |rhs| (0 where the solve ended just at the bound, below 0 the further it went under
//    it): the same scale at any tolerance, so a network trained on solves at one tolerance reads the same
//    inputs at another;
// 2. the Gram sums of the difference stage, in k_lg_gram_fin's order (eight interleaved groups of partials,
//    then the groups);
// 3. the raw inputs and their standardized bf16 row.  f_alpha(t_{n+1}) comes from the window's coefficient
//    table (c_alpha = hd * i * f_alpha, f real).  Range gate: oor[b] = 1 when a standardized input lies beyond
//    *zmax standard deviations of the training inputs (or is not finite), *zmax > 0; such a step takes the
//    extrapolation instead of the network.
__global__ __launch_bounds__(256) void k_lg_inputs(const double *__restrict__ sc0, double *__restrict__ lres,
                                                   int nparts, const double *__restrict__ gpart,
                                                   double *__restrict__ gram, const Coef6 *__restrict__ cf_dev,
                                                   const int *__restrict__ gctl, int B, double hd,
                                                   const double2 *__restrict__ z, int npod,
                                                   const float *__restrict__ mean, const float *__restrict__ stdv,
                                                   float *__restrict__ feat, uint16_t *__restrict__ xin,
                                                   const float *__restrict__ zmax, int *__restrict__ oor,
                                                   const int *gate, int gstride) {
    const int b = blockIdx.x;
    GADE_GATED_B(gate, gstride, b);
    __shared__ double gsh[8][32];
    __shared__ double g[kNG];
    __shared__ double lr[6];
    __shared__ float zw[4];
    if (threadIdx.x == 0) {
        const double *sc = sc0 + (size_t)b * SC_STRIDE;
        double *r = lres + (size_t)b * 8;
        for (int k = 7; k > 0; k--) r[k] = r[k - 1];
        const double t = sc[SC_TOL2] > 0 ? sqrt(sc[SC_RNORM2] / sc[SC_TOL2]) : 0.0;
        r[0] = log10(fmax(t, 1e-30));
        for (int k = 0; k < 6; k++) lr[k] = r[k];
    }
    {
        const int e = threadIdx.x & 31, gi = threadIdx.x >> 5;
        double acc = 0;
        if (e < kNG)
#pragma unroll 4
            for (int p = gi; p < nparts; p += 8) acc += gpart[((size_t)b * nparts + p) * kNG + e];
        gsh[gi][e] = acc;
        __syncthreads();
        if (gi == 0 && e < kNG) {
            double t = gsh[0][e];
#pragma unroll
            for (int q = 1; q < 8; q++) t += gsh[q][e];
            g[e] = t;
            gram[(size_t)b * kNG + e] = t;
        }
        __syncthreads();
    }
    float zabs = 0.f;
    const Coef6 *cf = cf_dev + (size_t)gctl[GC_WSLOT] * B + b;
    const double n0 = sqrt(fmax(g[gidx(0, 0)], 1e-300));
    const double s = sqrt(g[gidx(3, 3)]);
    const double inv_s = s > 0 ? 1.0 / s : 0.0;
    float *fr = feat + (size_t)b * kIn;
    for (int t = threadIdx.x; t < kIn; t += blockDim.x) {
        double v = 0;
        if (t < F_LRES) v = cf->c[t - F_F].y / hd;
        else if (t < F_DN) v = lr[t - F_LRES];
        else if (t < F_Z3) v = 0.5 * log10(fmax(g[gidx(t - F_DN + 1, t - F_DN + 1)], 1e-300)) - log10(n0);
        else if (t < F_Z4) {
            const int q = (t - F_Z3) >> 1;
            const double2 zz = q < npod ? z[((size_t)b * kNV + 0) * npod + q] : make_double2(0, 0);
            v = ((t - F_Z3) & 1 ? zz.y : zz.x) * inv_s;
        } else if (t < F_END) {
            const int q = (t - F_Z4) >> 1;
            const double2 zz = q < npod ? z[((size_t)b * kNV + 1) * npod + q] : make_double2(0, 0);
            v = ((t - F_Z4) & 1 ? zz.y : zz.x) * inv_s;
        }
        fr[t] = (float)v;
        const float zt = t < F_END ? ((float)v - mean[t]) / stdv[t] : 0.f;
        zabs = (zt == zt) ? fmaxf(zabs, fabsf(zt)) : INFINITY;
        xin[(size_t)b * kIn + t] = t < F_END ? f2bf(zt) : (uint16_t)0;
    }
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) zabs = fmaxf(zabs, __shfl_xor(zabs, off));
    if ((threadIdx.x & 63) == 0) zw[threadIdx.x >> 6] = zabs;
    __syncthreads();
    if (threadIdx.x == 0) {
        const float zm = fmaxf(fmaxf(zw[0], zw[1]), fmaxf(zw[2], zw[3]));
        oor[b] = (*zmax > 0.f && !(zm <= *zmax)) ? 1 : 0;
    }
}

// the guess of instance b into psi (psi_n enters the ring); fallback: order-2 extrapolation (or lower)
// (the network's term needs a full history, a non-zero s, finite outputs and inputs inside the range gate);
// uses[b] counts the steps that took the network (0) and those the range gate sent to the extrapolation (1)
__global__ __launch_bounds__(256) void k_lg_guess(long long N1, long long Ntot, double2 *__restrict__ psi,
                                                  double2 *__restrict__ hist, const double2 *__restrict__ V,
                                                  const double *__restrict__ gram, const float *__restrict__ Y,
                                                  const float2 *__restrict__ U, int npod,
                                                  const int *__restrict__ ctl0, const int *__restrict__ oor,
                                                  int *__restrict__ uses, const int *gate, int gstride) {
    const int b = blockIdx.y;
    GADE_GATED_B(gate, gstride, b);
    __shared__ double2 cs[kPod];
    __shared__ double ws[6];
    __shared__ int ok_s;
    const int have = ctl0[(size_t)b * CT_STRIDE + CT_HIST];
    const double *g = gram + (size_t)b * kNG;
    const double s = sqrt(g[gidx(3, 3)]);
    const float *y = Y + (size_t)b * kOut;
    if (threadIdx.x == 0) {
        bool ok = have >= 5 && s > 0 && isfinite(s);
        for (int k = 0; k < 6 + 2 * npod; k++) ok = ok && isfinite(y[k]);
        if (ok && blockIdx.x == 0) uses[2 * b + (oor[b] ? 1 : 0)] += 1;
        ok_s = ok && oor[b] == 0;
    }
    __syncthreads();
    const bool ok = ok_s != 0;
    if (ok) {
        for (int k = threadIdx.x; k < 6; k += blockDim.x) {
            const double nk = sqrt(g[gidx(k, k)]);
            ws[k] = nk > 0 ? s * y[k] / nk : 0.0;
        }
        for (int q = threadIdx.x; q < npod; q += blockDim.x) cs[q] = make_double2(s * y[6 + 2 * q], s * y[7 + 2 * q]);
    }
    __syncthreads();
    const int p = have < 2 ? have : 2;  // extrapolation order of the fallback
    const int slot0 = have % kHistSlots;
    for (long long j = blockIdx.x * (long long)blockDim.x + threadIdx.x; j < N1; j += (long long)gridDim.x * blockDim.x) {
        const long long i = (long long)b * N1 + j;
        const double2 xn = psi[i];
        double2 x;
        if (ok) {
            // x2 + sum_k w_k V_k + U c
            x = cadd(xn, cadd(V[(size_t)1 * Ntot + i], V[(size_t)2 * Ntot + i]));
            x = cadd(x, cscale(xn, ws[0]));
#pragma unroll
            for (int k = 1; k < 6; k++) x = cadd(x, cscale(V[(size_t)k * Ntot + i], ws[k]));
#pragma unroll 8
            for (int q = 0; q < npod; q++) x = cfma(cs[q], to_f64(U[(size_t)q * N1 + j]), x);
        } else {
            // Lagrange extrapolation of order p through psi_n..psi_{n-p}
            const double c2[3][3] = {{1, 0, 0}, {2, -1, 0}, {3, -3, 1}};
            x = cscale(xn, c2[p][0]);
            for (int k = 1; k <= p; k++) {
                const int sl = ((slot0 - k) % kHistSlots + kHistSlots) % kHistSlots;
                x = cadd(x, cscale(hist[(size_t)sl * Ntot + i], c2[p][k]));
            }
        }
        hist[(size_t)slot0 * Ntot + i] = xn;
        psi[i] = x;
    }
}

// training sample of instance b after its step (record mode of k_lg_diff / k_lg_proj): the projected
// quadratic form of the normalized error, next to the inputs saved at the step's start.  Valid: a full
// history, an active instance and a converged solve (a solve that ran out of iterations is no target)
__global__ void k_lg_record(const double *__restrict__ gram, const double2 *__restrict__ proj, int npod,
                            const float *__restrict__ feat, const int *__restrict__ ctl0, float *__restrict__ out) {
    const int b = blockIdx.x;
    const double *g = gram + (size_t)b * kNG;
    const int *ctl = ctl0 + (size_t)b * CT_STRIDE;
    float *o = out + (size_t)b * kSample;
    double nrm[kNV];
    for (int k = 0; k < kNV; k++) nrm[k] = sqrt(fmax(g[gidx(k, k)], 0.0));
    const double s = nrm[3];
    const bool valid = ctl[CT_HIST] >= 6 && ctl[CT_PSKIP] == 0 && s > 0 && isfinite(s) &&
                       ctl[CT_STATUS] == 1;
    for (int t = threadIdx.x; t < kSample; t += blockDim.x) {
        double v = 0;
        if (!valid) {
            v = 0;
        } else if (t < S_G) {
            v = feat[(size_t)b * kIn + t];
        } else if (t < S_A) {
            const int k = (t - S_G) / 6, l = (t - S_G) % 6;
            v = g[k <= l ? gidx(k, l) : gidx(l, k)] / (nrm[k] * nrm[l]);
        } else if (t < S_P) {
            const int k = t - S_A;
            v = g[gidx(k, 6)] / (nrm[k] * s);
        } else if (t < S_Q) {
            const int e = t - S_P, q = e / 12, k = (e % 12) >> 1;
            const double2 pv = proj[((size_t)b * kNV + k) * npod + q];
            v = ((e & 1) ? pv.y : pv.x) / nrm[k];
        } else if (t < S_BETA) {
            const int e = t - S_Q, q = e >> 1;
            const double2 qv = proj[((size_t)b * kNV + 6) * npod + q];
            v = ((e & 1) ? qv.y : qv.x) / s;
        } else if (t == S_BETA) {
            v = g[gidx(6, 6)] / (s * s);
        }
        if (t == S_VALID) v = valid ? 1.0 : 0.0;
        o[t] = (float)v;
    }
}

// normalized extrapolation defect e2 / |e2| of instance b into snapshot column col0 + b (valid instances)
__global__ void k_lg_snap(long long N1, long long Ntot, const double2 *__restrict__ V,
                          const double *__restrict__ gram, const int *__restrict__ ctl0, double2 *__restrict__ snaps,
                          int col0, int ncols) {
    const int b = blockIdx.y;
    const int *ctl = ctl0 + (size_t)b * CT_STRIDE;
    if (col0 + b >= ncols || ctl[CT_HIST] < 6 || ctl[CT_PSKIP] != 0) return;
    const double ne = sqrt(gram[(size_t)b * kNG + gidx(6, 6)]);
    if (!(ne > 0)) return;
    for (long long j = blockIdx.x * (long long)blockDim.x + threadIdx.x; j < N1; j += (long long)gridDim.x * blockDim.x)
        snaps[(size_t)(col0 + b) * N1 + j] = cscale(V[(size_t)6 * Ntot + (size_t)b * N1 + j], 1.0 / ne);
}

// ---------------------------------------------------------------------------
// training kernels
// ---------------------------------------------------------------------------
// X[m] = standardized bf16 inputs of sample idx[m]
__global__ void k_lg_gather(const float *__restrict__ samples, const int *__restrict__ idx, const float *mean,
                            const float *stdv, uint16_t *__restrict__ X) {
    const int m = blockIdx.x;
    const float *s = samples + (size_t)idx[m] * kSample;
    for (int t = threadIdx.x; t < kIn; t += blockDim.x)
        X[(size_t)m * kIn + t] = t < F_END ? f2bf((s[S_FEAT + t] - mean[t]) / stdv[t]) : (uint16_t)0;
}

// X[m] = standardized bf16 inputs of raw row m (stride kIn)
__global__ void k_lg_gather_rows(const float *__restrict__ x, const float *mean, const float *stdv,
                                 uint16_t *__restrict__ X) {
    const int m = blockIdx.x;
    for (int t = threadIdx.x; t < kIn; t += blockDim.x)
        X[(size_t)m * kIn + t] = t < F_END ? f2bf((x[(size_t)m * kIn + t] - mean[t]) / stdv[t]) : (uint16_t)0;
}

// per sample: loss L = |e_hat - D w - U c|^2 / beta (the quadratic form) and dL/dY / M
__global__ __launch_bounds__(64) void k_lg_loss(const float *__restrict__ Y, const float *__restrict__ samples,
                                                const int *__restrict__ idx, int M, float *__restrict__ dY,
                                                float *__restrict__ loss) {
    const int m = blockIdx.x, t = threadIdx.x;
    const float *y = Y + (size_t)m * kOut;
    const float *s = samples + (size_t)idx[m] * kSample;
    __shared__ float w[6], pw[2 * kPod], part[64];
    if (t < 6) w[t] = y[t];
    __syncthreads();
    const float beta = s[S_BETA];
    const float ib = beta > 0 ? 1.0f / beta : 0.0f;
    float contrib = 0;
    // POD part, one coefficient per lane
    {
        const int q = t;
        float pr = 0, pi = 0;
        for (int k = 0; k < 6; k++) {
            pr += s[S_P + (q * 6 + k) * 2] * w[k];
            pi += s[S_P + (q * 6 + k) * 2 + 1] * w[k];
        }
        pw[2 * q] = pr;
        pw[2 * q + 1] = pi;
        const float cr = y[6 + 2 * q], ci = y[7 + 2 * q];
        const float qr = s[S_Q + 2 * q], qi = s[S_Q + 2 * q + 1];
        contrib += -2.f * (cr * qr + ci * qi) + cr * cr + ci * ci + 2.f * (cr * pr + ci * pi);
        dY[(size_t)m * kOut + 6 + 2 * q] = 2.f * ib * (cr - qr + pr) / M;
        dY[(size_t)m * kOut + 7 + 2 * q] = 2.f * ib * (ci - qi + pi) / M;
    }
    __syncthreads();
    if (t < 6) {
        float gw = 0;
        for (int l = 0; l < 6; l++) gw += s[S_G + t * 6 + l] * w[l];
        float pc = 0;
        for (int q = 0; q < kPod; q++) pc += s[S_P + (q * 6 + t) * 2] * y[6 + 2 * q] + s[S_P + (q * 6 + t) * 2 + 1] * y[7 + 2 * q];
        contrib += -2.f * w[t] * s[S_A + t] + w[t] * gw;
        dY[(size_t)m * kOut + t] = 2.f * ib * (-s[S_A + t] + gw + pc) / M;
    }
    if (t < kOut - (6 + 2 * kPod)) dY[(size_t)m * kOut + 6 + 2 * kPod + t] = 0.f;  // padding outputs
    part[t] = contrib;
    __syncthreads();
    if (t == 0) {
        float tot = 0;
        for (int k = 0; k < 64; k++) tot += part[k];
        loss[m] = beta > 0 ? 1.0f + tot * ib : 0.f;
    }
}

// mean of loss[0..M) into out[0]
__global__ void k_lg_mean(const float *__restrict__ loss, int M, float *__restrict__ out) {
    __shared__ float sh[256];
    float s = 0;
    for (int m = threadIdx.x; m < M; m += blockDim.x) s += loss[m];
    sh[threadIdx.x] = s;
    __syncthreads();
    if (threadIdx.x == 0) {
        float t = 0;
        for (int k = 0; k < (int)blockDim.x; k++) t += sh[k];
        out[0] = t / M;
    }
}

// bf16 copy of an FP32 matrix
__global__ void k_lg_tobf(long n, const float *__restrict__ x, uint16_t *__restrict__ y) {
    for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) y[i] = f2bf(x[i]);
}

// dZ = dA * gelu'(Z), and its bf16 copy
__global__ void k_lg_gelu_back(long n, const float *__restrict__ dA, const float *__restrict__ Z, float *__restrict__ dZ,
                               uint16_t *__restrict__ G) {
    for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
        const float v = dA[i] * gelu_d(Z[i]);
        dZ[i] = v;
        G[i] = f2bf(v);
    }
}

// column sums db[n] = sum_m dZ[m][n]: 64 columns per block, 4 row groups
__global__ __launch_bounds__(256) void k_lg_colsum(int M, int N, const float *__restrict__ dZ, float *__restrict__ db) {
    __shared__ float sh[4][64];
    const int c = threadIdx.x & 63, g = threadIdx.x >> 6;
    const int n = blockIdx.x * 64 + c;
    float s = 0;
    if (n < N)
        for (int m = g; m < M; m += 4) s += dZ[(size_t)m * N + n];
    sh[g][c] = s;
    __syncthreads();
    if (g == 0 && n < N) db[n] = (sh[0][c] + sh[1][c]) + (sh[2][c] + sh[3][c]);
}

// Adam step on the FP32 parameters, refreshing the bf16 weights
__global__ void k_lg_adam(long n, float *__restrict__ p, const float *__restrict__ g, float *__restrict__ m1,
                          float *__restrict__ m2, float lr, float b1, float b2, float eps, float c1, float c2,
                          uint16_t *__restrict__ pbf) {
    for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
        const float gi = g[i];
        const float a = b1 * m1[i] + (1.f - b1) * gi;
        const float v = b2 * m2[i] + (1.f - b2) * gi * gi;
        m1[i] = a;
        m2[i] = v;
        const float pi = p[i] - lr * (a / c1) / (sqrtf(v / c2) + eps);
        p[i] = pi;
        pbf[i] = f2bf(pi);
    }
}

}  // namespace

// ---------------------------------------------------------------------------
// model-side state
// ---------------------------------------------------------------------------
struct gade_lg {
    int npod = 0;
    double2 *U = nullptr;         // [npod][N1] POD basis (orthonormal columns)
    float2 *U32 = nullptr;        // [npod][N1] FP32 copy (guess assembly)
    float2 *UT32 = nullptr;       // [N1][npod] FP32 copy, k-major (projections)
    double2 *ppart = nullptr;     // [B][nchunk][7][64] projection partials
    int nchunk = 0;
    double2 *V = nullptr;         // [7][N] difference stage vectors
    double *part = nullptr;       // [B][nparts][kNG]
    int nparts = 0;
    double *gram = nullptr;       // [B][kNG]
    double2 *proj = nullptr;      // [B][7][npod]
    double *lres = nullptr;       // [B][8]
    float *feat = nullptr;        // [B][kIn] raw inputs of the current step
    float *mean = nullptr, *stdv = nullptr;  // [kIn]
    float *zmax = nullptr;        // range gate of the inputs (device scalar, standard deviations; 0: off)
    float zmax_set = -1.f;        // the value last written there
    int *oor = nullptr;           // [B] this step's inputs lie outside the range gate
    int *uses = nullptr;          // [B][2] steps that took the network / that the range gate sent to the fallback
    int Mi = 0;                   // inference rows (B padded to 16)
    uint16_t *xin = nullptr;      // [Mi][kIn]
    float *zi[3] = {}, *yi = nullptr;  // inference pre-activations [Mi][kHid], outputs [Mi][kOut]
    uint16_t *ai[3] = {};         // inference activations
    // network
    bool net = false;
    float *p = nullptr, *g = nullptr, *m1 = nullptr, *m2 = nullptr;  // [P_TOTAL]
    uint16_t *pbf = nullptr;      // bf16 copy (weights)
    long adam_t = 0;
    // training buffers (batch Mt)
    int Mt = 0;
    uint16_t *xt = nullptr, *at[3] = {}, *gt[4] = {};
    float *zt[3] = {}, *yt = nullptr, *dy = nullptr, *da = nullptr, *dz = nullptr, *loss = nullptr;
};

namespace gade {

void lg_free(gade_lg *lg) {
    if (!lg) return;
    std::vector<void *> ptrs = {lg->U, lg->U32, lg->UT32, lg->ppart, lg->V, lg->part, lg->gram, lg->proj, lg->lres, lg->feat, lg->mean, lg->stdv,
                                lg->zmax, lg->oor, lg->uses,
                                lg->xin, lg->yi, lg->p, lg->g, lg->m1, lg->m2, lg->pbf, lg->xt, lg->yt, lg->dy,
                                lg->da, lg->dz, lg->loss};
    for (int k = 0; k < 3; k++) {
        ptrs.push_back(lg->zi[k]);
        ptrs.push_back(lg->ai[k]);
        ptrs.push_back(lg->at[k]);
        ptrs.push_back(lg->zt[k]);
    }
    for (int k = 0; k < 4; k++) ptrs.push_back(lg->gt[k]);
    for (void *q : ptrs)
        if (q) (void)hipFree(q);
    delete lg;
}

}  // namespace gade

namespace {

// the FP32 per-step copies of the model's basis (after the FP64 basis is built or installed)
int lg_basis32(gade_model *m) {
    gade_lg *lg = m->lg;
    const size_t n = (size_t)m->N1;
    if (!lg->U32) GADE_TRY(hip_alloc((void **)&lg->U32, sizeof(float2) * n * kPod));
    if (!lg->UT32) GADE_TRY(hip_alloc((void **)&lg->UT32, sizeof(float2) * n * kPod));
    hipLaunchKernelGGL(k_lg_basis32, dim3(grid_for_lg((long long)n * lg->npod)), dim3(256), 0, m->stream,
                       (long long)n, lg->npod, (const double2 *)lg->U, lg->U32, lg->UT32);
    GADE_HIP(hipGetLastError());
    GADE_HIP(hipStreamSynchronize(m->stream));
    return GADE_OK;
}

int lg_ensure(gade_model *m) {
    if (m->lg) return GADE_OK;
    gade_lg *lg = new gade_lg();
    m->lg = lg;
    const size_t N = (size_t)m->N, B = (size_t)m->B;
    lg->nparts = (int)std::min<long long>(512, std::max<long long>(1, (m->N1 + 511) / 512));
    lg->Mi = pad16(m->B);
    GADE_TRY(hip_alloc((void **)&lg->V, sizeof(double2) * kNV * N));
    GADE_TRY(hip_alloc((void **)&lg->part, sizeof(double) * B * lg->nparts * kNG));
    GADE_TRY(hip_alloc((void **)&lg->gram, sizeof(double) * B * kNG));
    GADE_TRY(hip_alloc((void **)&lg->proj, sizeof(double2) * B * kNV * kPod));
    lg->nchunk = proj_chunks(m->N1, m->B);
    GADE_TRY(hip_alloc((void **)&lg->ppart, sizeof(double2) * B * lg->nchunk * kNV * 64));
    GADE_TRY(hip_alloc((void **)&lg->lres, sizeof(double) * B * 8));
    GADE_TRY(hip_alloc((void **)&lg->feat, sizeof(float) * B * kIn));
    GADE_TRY(hip_alloc((void **)&lg->mean, sizeof(float) * kIn));
    GADE_TRY(hip_alloc((void **)&lg->stdv, sizeof(float) * kIn));
    GADE_TRY(hip_alloc((void **)&lg->xin, sizeof(uint16_t) * lg->Mi * kIn));
    GADE_TRY(hip_alloc((void **)&lg->yi, sizeof(float) * lg->Mi * kOut));
    for (int k = 0; k < 3; k++) {
        GADE_TRY(hip_alloc((void **)&lg->zi[k], sizeof(float) * lg->Mi * kHid));
        GADE_TRY(hip_alloc((void **)&lg->ai[k], sizeof(uint16_t) * lg->Mi * kHid));
    }
    GADE_HIP(hipMemset(lg->xin, 0, sizeof(uint16_t) * lg->Mi * kIn));
    GADE_HIP(hipMemset(lg->lres, 0, sizeof(double) * B * 8));
    GADE_TRY(hip_alloc((void **)&lg->zmax, sizeof(float)));
    GADE_TRY(hip_alloc((void **)&lg->oor, sizeof(int) * B));
    GADE_TRY(hip_alloc((void **)&lg->uses, sizeof(int) * B * 2));
    GADE_HIP(hipMemset(lg->oor, 0, sizeof(int) * B));
    GADE_HIP(hipMemset(lg->zmax, 0, sizeof(float)));
    lg->zmax_set = 0.f;
    GADE_HIP(hipMemset(lg->uses, 0, sizeof(int) * B * 2));
    GADE_HIP(hipMemset(lg->proj, 0, sizeof(double2) * B * kNV * kPod));
    std::vector<float> one(kIn, 1.0f);
    GADE_HIP(hipMemset(lg->mean, 0, sizeof(float) * kIn));
    GADE_HIP(hipMemcpy(lg->stdv, one.data(), sizeof(float) * kIn, hipMemcpyHostToDevice));
    if (!m->d_hist) GADE_TRY(hip_alloc((void **)&m->d_hist, sizeof(double2) * kHistSlots * m->N));
    return GADE_OK;
}

int lg_net_alloc(gade_lg *lg) {
    if (lg->net) return GADE_OK;
    for (float **q : {&lg->p, &lg->g, &lg->m1, &lg->m2}) GADE_TRY(hip_alloc((void **)q, sizeof(float) * P_TOTAL));
    GADE_TRY(hip_alloc((void **)&lg->pbf, sizeof(uint16_t) * P_TOTAL));
    GADE_HIP(hipMemset(lg->g, 0, sizeof(float) * P_TOTAL));
    lg->net = true;
    return GADE_OK;
}

int lg_train_alloc(gade_lg *lg, int Mt) {
    if (lg->Mt >= Mt) return GADE_OK;
    for (void *q : {(void *)lg->xt, (void *)lg->yt, (void *)lg->dy, (void *)lg->da, (void *)lg->dz, (void *)lg->loss})
        if (q) (void)hipFree(q);
    for (int k = 0; k < 3; k++) {
        if (lg->at[k]) (void)hipFree(lg->at[k]);
        if (lg->zt[k]) (void)hipFree(lg->zt[k]);
    }
    for (int k = 0; k < 4; k++)
        if (lg->gt[k]) (void)hipFree(lg->gt[k]);
    lg->Mt = Mt;
    GADE_TRY(hip_alloc((void **)&lg->xt, sizeof(uint16_t) * Mt * kIn));
    GADE_TRY(hip_alloc((void **)&lg->yt, sizeof(float) * Mt * kOut));
    GADE_TRY(hip_alloc((void **)&lg->dy, sizeof(float) * Mt * kOut));
    GADE_TRY(hip_alloc((void **)&lg->da, sizeof(float) * Mt * kHid));
    GADE_TRY(hip_alloc((void **)&lg->dz, sizeof(float) * Mt * kHid));
    GADE_TRY(hip_alloc((void **)&lg->loss, sizeof(float) * Mt));
    for (int k = 0; k < 3; k++) {
        GADE_TRY(hip_alloc((void **)&lg->at[k], sizeof(uint16_t) * Mt * kHid));
        GADE_TRY(hip_alloc((void **)&lg->zt[k], sizeof(float) * Mt * kHid));
    }
    for (int k = 0; k < 4; k++) GADE_TRY(hip_alloc((void **)&lg->gt[k], sizeof(uint16_t) * Mt * kHid));
    return GADE_OK;
}

// out[b][v][j] = U_j^H V_{v0+v} for v < nv
// projections of V_{v0..v0+nv-1} on the basis (nv: 2 at inference, kNV when recording)
void lg_project(gade_model *m, int v0, int nv, const int *gate, int gstride) {
    gade_lg *lg = m->lg;
    const long long clen = (m->N1 + lg->nchunk - 1) / lg->nchunk;
    if (nv == 2)
        hipLaunchKernelGGL(k_lg_proj_part<2>, dim3(m->B, lg->nchunk), dim3(256), 0, m->stream, m->N1, m->N, lg->npod,
                           (const float2 *)lg->UT32, (const double2 *)lg->V, v0, lg->ppart, gate, gstride, clen);
    else
        hipLaunchKernelGGL(k_lg_proj_part<kNV>, dim3(m->B, lg->nchunk), dim3(256), 0, m->stream, m->N1, m->N,
                           lg->npod, (const float2 *)lg->UT32, (const double2 *)lg->V, v0, lg->ppart, gate, gstride,
                           clen);
    hipLaunchKernelGGL(k_lg_proj_fin, dim3(m->B, nv, 4), dim3(256), 0, m->stream, lg->nchunk, lg->npod, nv,
                       (const double2 *)lg->ppart, lg->proj, gate, gstride);
}

// forward of M rows: X (bf16 [M][kIn]) -> Z_l, A_l, Y; up to kSplitKRows rows (inference) each tile's k steps are
// split over four waves
constexpr int kSplitKRows = 64;
void mlp_forward(const gade_lg *lg, hipStream_t st, int M, const uint16_t *X, float *const Z[3], uint16_t *const A[3],
                 float *Y) {
    const uint16_t *in = X;
    for (int l = 0; l < kLayers; l++) {
        const int K = kLin[l], N = kLout[l];
        const dim3 grid((M / 16) * (N / 16));
        const uint16_t *W = lg->pbf + kWoff[l];
        const float *bias = lg->p + kBoff[l];
        if (M <= kSplitKRows) {
            if (l < 3)
                hipLaunchKernelGGL((k_gemm_bf16<EPI_BIAS_GELU, 4>), grid, dim3(256), 0, st, M, N, K, in, (long)K, 1L,
                                   W, 1L, (long)K, Z[l], N, bias, A[l], N);
            else
                hipLaunchKernelGGL((k_gemm_bf16<EPI_BIAS, 4>), grid, dim3(256), 0, st, M, N, K, in, (long)K, 1L, W,
                                   1L, (long)K, Y, N, bias, (uint16_t *)nullptr, 0);
        } else if (l < 3) {
            hipLaunchKernelGGL(k_gemm_bf16<EPI_BIAS_GELU>, grid, dim3(64), 0, st, M, N, K, in, (long)K, 1L, W, 1L,
                               (long)K, Z[l], N, bias, A[l], N);
        } else {
            hipLaunchKernelGGL(k_gemm_bf16<EPI_BIAS>, grid, dim3(64), 0, st, M, N, K, in, (long)K, 1L, W, 1L, (long)K,
                               Y, N, bias, (uint16_t *)nullptr, 0);
        }
        if (l < 3) in = A[l];
    }
}

// backward of M rows from dY (FP32 [M][kOut]): gradients of every parameter into lg->g
void mlp_backward(gade_lg *lg, hipStream_t st, int M, const uint16_t *X) {
    hipLaunchKernelGGL(k_lg_tobf, dim3(64), dim3(256), 0, st, (long)M * kOut, (const float *)lg->dy, lg->gt[3]);
    const float *dZ = lg->dy;
    for (int l = kLayers - 1; l >= 0; l--) {
        const int K = kLin[l], N = kLout[l];
        const uint16_t *Gl = lg->gt[l];
        const uint16_t *Ain = l > 0 ? lg->at[l - 1] : X;
        // dW[n][k] = sum_m dZ[m][n] A[m][k]
        hipLaunchKernelGGL(k_gemm_bf16<EPI_STORE>, dim3((N / 16) * (K / 16)), dim3(64), 0, st, N, K, M, Gl, 1L,
                           (long)N, Ain, (long)K, 1L, lg->g + kWoff[l], K, (const float *)nullptr,
                           (uint16_t *)nullptr, 0);
        hipLaunchKernelGGL(k_lg_colsum, dim3((N + 63) / 64), dim3(256), 0, st, M, N, dZ, lg->g + kBoff[l]);
        if (l > 0) {
            // dA[m][k] = sum_n dZ[m][n] W[n][k], then dZ_{l-1} = dA * gelu'(Z_{l-1})
            hipLaunchKernelGGL(k_gemm_bf16<EPI_STORE>, dim3((M / 16) * (K / 16)), dim3(64), 0, st, M, K, N, Gl,
                               (long)N, 1L, lg->pbf + kWoff[l], (long)K, 1L, lg->da, K, (const float *)nullptr,
                               (uint16_t *)nullptr, 0);
            hipLaunchKernelGGL(k_lg_gelu_back, dim3(64), dim3(256), 0, st, (long)M * K, (const float *)lg->da,
                               (const float *)lg->zt[l - 1], lg->dz, lg->gt[l - 1]);
            dZ = lg->dz;
        }
    }
}

}  // namespace

// ---------------------------------------------------------------------------
// hooks used by the CN step
// ---------------------------------------------------------------------------
namespace gade {

// buffers of the learned guess (no allocation may happen while a step graph is captured)
// before a run's steps are enqueued (outside any capture): the range gate's bound into device memory, where
// the steps read it; stream-ordered, after the launches of earlier runs that read the old bound
int lg_prepare(gade_model *m) {
    GADE_TRY(lg_ensure(m));
    gade_lg *lg = m->lg;
    const float z = (float)m->lg_zmax;
    if (z != lg->zmax_set) {
        lg->zmax_set = z;
        GADE_HIP(hipMemcpyAsync(lg->zmax, &lg->zmax_set, sizeof(float), hipMemcpyHostToDevice, m->stream));
        GADE_HIP(hipStreamSynchronize(m->stream));
    }
    return GADE_OK;
}

// device state of the learned guess that carries from step to step (the residual history of its inputs)
void lg_state(gade_model *m, void **ptr, size_t *bytes) {
    *ptr = m->lg ? (void *)m->lg->lres : nullptr;
    *bytes = m->lg ? sizeof(double) * m->B * 8 : 0;
}

int lg_reset(gade_model *m) {
    if (!m->lg) return GADE_OK;
    GADE_HIP(hipMemsetAsync(m->lg->lres, 0, sizeof(double) * m->B * 8, m->stream));
    GADE_HIP(hipMemsetAsync(m->lg->uses, 0, sizeof(int) * m->B * 2, m->stream));
    return GADE_OK;
}

// Step prologue of the learned guess (guess_order == GADE_GUESS_LEARNED) or of the training-data mode
// (lg_features with an extrapolated guess): inputs of this step, then the guess into psi.
int lg_step_prologue(gade_model *m, double2 *psi) {
    GADE_TRY(lg_ensure(m));
    gade_lg *lg = m->lg;
    const bool learned = m->guess_order == GADE_GUESS_LEARNED;
    if (!m->d_cfwin || m->radial_kind != 2) {
        set_error("learned guess: runs in the device-driven stepping with the banded radial kernel");
        return GADE_EINVAL;
    }
    if (learned && (!lg->net || lg->npod == 0)) {
        set_error("learned guess: no trained network (gade_lg_train) or POD basis (gade_lg_pod_build)");
        return GADE_EINVAL;
    }
    const int B = m->B;
    const long long N1 = m->N1, N = m->N;
    const dim3 gd(lg->nparts, B);
    hipLaunchKernelGGL(k_lg_diff<false>, gd, dim3(256), 0, m->stream, N1, N, (const double2 *)psi,
                       (const double2 *)m->d_hist, (const int *)m->d_ctl, lg->V, lg->part, m->gate, m->gstride);
    if (lg->npod > 0) lg_project(m, 3, 2, m->gate, m->gstride);
    const double hd = kSoL * m->h * m->dt / 2;
    hipLaunchKernelGGL(k_lg_inputs, dim3(B), dim3(256), 0, m->stream, (const double *)m->d_sc, lg->lres, lg->nparts,
                       (const double *)lg->part, lg->gram, (const Coef6 *)m->d_cfwin, (const int *)m->d_gctl, B, hd,
                       (const double2 *)lg->proj, lg->npod, (const float *)lg->mean, (const float *)lg->stdv,
                       lg->feat, lg->xin, (const float *)lg->zmax, lg->oor, m->gate, m->gstride);
    GADE_HIP(hipGetLastError());
    if (learned) {
        mlp_forward(lg, m->stream, lg->Mi, lg->xin, lg->zi, lg->ai, lg->yi);
        hipLaunchKernelGGL(k_lg_guess, dim3(grid_for_lg(N1), B), dim3(256), 0, m->stream, N1, N, psi, m->d_hist,
                           (const double2 *)lg->V, (const double *)lg->gram, (const float *)lg->yi,
                           (const float2 *)lg->U32, lg->npod, (const int *)m->d_ctl, (const int *)lg->oor, lg->uses,
                           m->gate, m->gstride);
    } else {
        hipLaunchKernelGGL(k_extrap, dim3(grid_for_lg(N1), B), dim3(256), 0, m->stream, N1, N, psi, m->d_hist,
                           m->guess_order, (const int *)m->d_ctl, m->gate, m->gstride);
    }
    hipLaunchKernelGGL(k_hist_count, dim3(B), dim3(64), 0, m->stream, m->d_ctl, m->gate, m->gstride);
    GADE_HIP(hipGetLastError());
    return GADE_OK;
}

}  // namespace gade

// ---------------------------------------------------------------------------
// C ABI
// ---------------------------------------------------------------------------
extern "C" {

int gade_lg_record(gade_model_t *m, const double *psi, float *samples, double *snaps, int snap_col, int snap_cols) {
    if (!m || !psi) return GADE_EINVAL;
    GADE_TRY(lg_ensure(m));
    gade_lg *lg = m->lg;
    const int B = m->B;
    const long long N1 = m->N1, N = m->N;
    hipLaunchKernelGGL(k_lg_diff<true>, dim3(lg->nparts, B), dim3(256), 0, m->stream, N1, N, (const double2 *)psi,
                       (const double2 *)m->d_hist, (const int *)m->d_ctl, lg->V, lg->part, (const int *)nullptr, 0);
    hipLaunchKernelGGL(k_lg_gram_fin, dim3(B), dim3(256), 0, m->stream, lg->nparts, (const double *)lg->part,
                       lg->gram, (const int *)nullptr, 0);
    if (snaps)
        hipLaunchKernelGGL(k_lg_snap, dim3(grid_for_lg(N1), B), dim3(256), 0, m->stream, N1, N,
                           (const double2 *)lg->V, (const double *)lg->gram, (const int *)m->d_ctl,
                           (double2 *)snaps, snap_col, snap_cols);
    if (samples) {
        if (lg->npod == 0) {
            set_error("gade_lg_record: samples need the POD basis (gade_lg_pod_build)");
            return GADE_EINVAL;
        }
        lg_project(m, 0, kNV, (const int *)nullptr, 0);
        hipLaunchKernelGGL(k_lg_record, dim3(B), dim3(256), 0, m->stream, (const double *)lg->gram,
                           (const double2 *)lg->proj, lg->npod, (const float *)lg->feat, (const int *)m->d_ctl,
                           samples);
    }
    GADE_HIP(hipGetLastError());
    GADE_HIP(hipStreamSynchronize(m->stream));
    return GADE_OK;
}

// M[:, c0 .. c0+nc) <- H_j^H M (conj_tau) or H_j M with the reflector of column j of A (n rows)
static int hh_apply(gade_model *m, const double2 *A, long long n, int j, double2 *M, int c0, int nc,
                    double2 *part, const double2 *tau, int conj_tau) {
    if (nc <= 0) return GADE_OK;
    const long long rows = n - j;
    const int nch = std::max(1, std::min(kHhChunks, (2048 + nc - 1) / nc));
    const long long per = (rows + nch - 1) / nch;
    const double2 *v = A + (size_t)j * n;
    hipLaunchKernelGGL(k_hh_dots, dim3(nc, nch), dim3(256), 0, m->stream, v, n, j, (const double2 *)M, c0, per, nch,
                       part);
    const int nrb = std::max(1, std::min(256, (4096 + nc - 1) / nc));
    const long long per_b = (rows + nrb - 1) / nrb;
    hipLaunchKernelGGL(k_hh_update, dim3(nc, nrb), dim3(256), 0, m->stream, v, n, j, M, c0, (const double2 *)part,
                       nch, tau, conj_tau, per_b);
    GADE_HIP(hipGetLastError());
    return GADE_OK;
}

int gade_lg_pod_build(gade_model_t *m, const double *snaps, int nsnap, int npod, double *sv) {
    if (!m || !snaps || npod <= 0 || npod > kPod || nsnap < npod) return GADE_EINVAL;
    GADE_TRY(lg_ensure(m));
    gade_lg *lg = m->lg;
    const int n = (int)m->N1;
    if (nsnap > n) return GADE_EINVAL;
    gade::drop_step_graph(m);  // the graph reads the basis being replaced
    double2 *A = nullptr, *tau = nullptr, *part = nullptr, *C = nullptr;
    auto cleanup = [&]() {
        for (void *q : {(void *)A, (void *)tau, (void *)part, (void *)C})
            if (q) (void)hipFree(q);
    };
    int rc = GADE_OK;
    // snapshot columns that are zero (unused slots) or not finite are left out: they carry nothing and an
    // exact zero column only adds a zero singular value
    std::vector<int> keep;
    {
        double *d_nrm = nullptr;
        GADE_TRY(hip_alloc((void **)&d_nrm, sizeof(double) * nsnap));
        hipLaunchKernelGGL(k_lg_colnorm, dim3(nsnap), dim3(256), 0, m->stream, (long long)n, (const double2 *)snaps,
                           d_nrm);
        std::vector<double> nrm(nsnap);
        hipError_t e = hipGetLastError();
        if (e == hipSuccess) e = hipMemcpyAsync(nrm.data(), d_nrm, sizeof(double) * nsnap, hipMemcpyDeviceToHost, m->stream);
        if (e == hipSuccess) e = hipStreamSynchronize(m->stream);
        (void)hipFree(d_nrm);
        if (e != hipSuccess) {
            set_error(std::string("gade_lg_pod_build: ") + hipGetErrorString(e));
            return GADE_EHIP;
        }
        for (int c = 0; c < nsnap; c++)
            if (nrm[c] > 0 && std::isfinite(nrm[c])) keep.push_back(c);
    }
    if ((int)keep.size() < npod) {
        set_error("gade_lg_pod_build: " + std::to_string(keep.size()) + " non-zero snapshots, fewer than the basis");
        return GADE_EINVAL;
    }
    const int nused = (int)keep.size();
    const size_t ns = (size_t)nused;
    const int ncol = std::max(nused, npod);
    if ((rc = hip_alloc((void **)&A, sizeof(double2) * n * ns)) || (rc = hip_alloc((void **)&tau, sizeof(double2) * ns)) ||
        (rc = hip_alloc((void **)&part, sizeof(double2) * (size_t)ncol * kHhChunks)) ||
        (rc = hip_alloc((void **)&C, sizeof(double2) * n * npod))) {
        cleanup();
        return rc;
    }
    // snapshots [col][n] are the column-major n x nused matrix X: X = Q R (Householder, own kernels),
    // R = Ur S Vr^H (one-sided Jacobi on the host), U = Q Ur[:, :npod]
    bool ok = true;
    for (size_t c = 0; c < ns && ok; c++)
        ok = hipMemcpyAsync(A + c * n, (const double2 *)snaps + (size_t)keep[c] * n, sizeof(double2) * n,
                            hipMemcpyDeviceToDevice, m->stream) == hipSuccess;
    for (int j = 0; j < nused && ok && rc == GADE_OK; j++) {
        hipLaunchKernelGGL(k_hh_reflector, dim3(1), dim3(1024), 0, m->stream, A, (long long)n, j, tau);
        rc = hh_apply(m, A, n, j, A, j + 1, nused - j - 1, part, tau, 1);
    }
    // R: upper triangle of the first nused rows (a strided copy), below the diagonal the reflectors
    std::vector<double2> Rh(ns * ns, make_double2(0, 0));
    if (ok && rc == GADE_OK)
        ok = hipMemcpy2DAsync(Rh.data(), sizeof(double2) * ns, A, sizeof(double2) * n, sizeof(double2) * ns, ns,
                              hipMemcpyDeviceToHost, m->stream) == hipSuccess &&
             hipStreamSynchronize(m->stream) == hipSuccess;
    std::vector<double> Sh(ns);
    std::vector<double2> Uh(ns * ns);
    int svd_rc = GADE_OK;
    if (ok && rc == GADE_OK) {
        for (size_t c = 0; c < ns; c++)
            for (size_t r = c + 1; r < ns; r++) Rh[c * ns + r] = make_double2(0, 0);
        svd_rc = gade::svd_left_jacobi(nused, (const double *)Rh.data(), Sh.data(), (double *)Uh.data());
    }
    // C = [Ur[:, :npod]; 0] (n x npod), then U = Q C
    if (ok && rc == GADE_OK && svd_rc == GADE_OK) {
        std::vector<double2> Ch((size_t)n * npod, make_double2(0, 0));
        for (int c = 0; c < npod; c++)
            for (size_t r = 0; r < ns; r++) Ch[(size_t)c * n + r] = Uh[(size_t)c * ns + r];
        ok = hipMemcpy(C, Ch.data(), sizeof(double2) * n * npod, hipMemcpyHostToDevice) == hipSuccess;
        for (int j = nused - 1; j >= 0 && ok && rc == GADE_OK; j--) rc = hh_apply(m, A, n, j, C, 0, npod, part, tau, 0);
        if (ok && rc == GADE_OK) ok = hipStreamSynchronize(m->stream) == hipSuccess;
    }
    if (!ok || rc != GADE_OK || svd_rc != GADE_OK) {
        cleanup();
        set_error(svd_rc != GADE_OK ? "gade_lg_pod_build: the SVD of R did not converge"
                                    : "gade_lg_pod_build: QR of the snapshots failed");
        return rc != GADE_OK ? rc : svd_rc != GADE_OK ? svd_rc : GADE_EHIP;
    }
    if (sv) {  // singular values of the used snapshots, then zeros for the columns left out
        memcpy(sv, Sh.data(), sizeof(double) * ns);
        for (int c = nused; c < nsnap; c++) sv[c] = 0.0;
    }
    if (lg->U) (void)hipFree(lg->U);
    lg->U = C;
    C = nullptr;
    lg->npod = npod;
    cleanup();
    return lg_basis32(m);
}

int gade_lg_set_basis(gade_model_t *m, const double *U, int npod) {
    if (!m || !U || npod <= 0 || npod > kPod) return GADE_EINVAL;
    GADE_TRY(lg_ensure(m));
    gade_lg *lg = m->lg;
    gade::drop_step_graph(m);  // the graph reads the basis being replaced
    const size_t n = (size_t)m->N1;
    if (lg->U) (void)hipFree(lg->U);
    lg->U = nullptr;
    GADE_TRY(hip_alloc((void **)&lg->U, sizeof(double2) * n * npod));
    GADE_HIP(hipMemcpy(lg->U, U, sizeof(double2) * n * npod, hipMemcpyHostToDevice));
    lg->npod = npod;
    return lg_basis32(m);
}

int gade_lg_basis(gade_model_t *m, double *U) {
    if (!m || !U || !m->lg || !m->lg->U) return GADE_EINVAL;
    GADE_HIP(hipMemcpy(U, m->lg->U, sizeof(double2) * m->N1 * m->lg->npod, hipMemcpyDeviceToHost));
    return GADE_OK;
}

int gade_lg_set_norm(gade_model_t *m, const float *mean, const float *stdv) {
    if (!m || !mean || !stdv) return GADE_EINVAL;
    GADE_TRY(lg_ensure(m));
    for (int t = 0; t < kIn; t++)
        if (!(stdv[t] > 0)) {
            set_error("gade_lg_set_norm: every standard deviation must be positive");
            return GADE_EINVAL;
        }
    GADE_HIP(hipMemcpy(m->lg->mean, mean, sizeof(float) * kIn, hipMemcpyHostToDevice));
    GADE_HIP(hipMemcpy(m->lg->stdv, stdv, sizeof(float) * kIn, hipMemcpyHostToDevice));
    return GADE_OK;
}

int gade_lg_init(gade_model_t *m, unsigned seed) {
    if (!m) return GADE_EINVAL;
    GADE_TRY(lg_ensure(m));
    GADE_TRY(lg_net_alloc(m->lg));
    // He-normal hidden layers, zero output layer and biases: the untrained network is the extrapolation
    std::vector<float> p(P_TOTAL, 0.0f);
    std::mt19937 rng(seed);
    for (int l = 0; l < 3; l++) {
        std::normal_distribution<float> nd(0.0f, std::sqrt(2.0f / kLin[l]));
        for (size_t i = 0; i < (size_t)kLout[l] * kLin[l]; i++) p[kWoff[l] + i] = nd(rng);
    }
    return gade_lg_set_params(m, p.data());
}

int gade_lg_set_params(gade_model_t *m, const float *params) {
    if (!m || !params) return GADE_EINVAL;
    GADE_TRY(lg_ensure(m));
    gade_lg *lg = m->lg;
    GADE_TRY(lg_net_alloc(lg));
    GADE_HIP(hipMemcpy(lg->p, params, sizeof(float) * P_TOTAL, hipMemcpyHostToDevice));
    GADE_HIP(hipMemset(lg->m1, 0, sizeof(float) * P_TOTAL));
    GADE_HIP(hipMemset(lg->m2, 0, sizeof(float) * P_TOTAL));
    lg->adam_t = 0;
    hipLaunchKernelGGL(k_lg_tobf, dim3(256), dim3(256), 0, m->stream, (long)P_TOTAL, (const float *)lg->p, lg->pbf);
    GADE_HIP(hipGetLastError());
    GADE_HIP(hipStreamSynchronize(m->stream));
    return GADE_OK;
}

int gade_lg_get_params(gade_model_t *m, float *params) {
    if (!m || !params || !m->lg || !m->lg->net) return GADE_EINVAL;
    GADE_HIP(hipMemcpy(params, m->lg->p, sizeof(float) * P_TOTAL, hipMemcpyDeviceToHost));
    return GADE_OK;
}

// forward of M rows of raw inputs (device FP32 [M][GADE_LG_IN], M a multiple of 16) into Y [M][GADE_LG_OUT]
int gade_lg_uses(gade_model_t *m, int *counts) {
    if (!m || !counts) return GADE_EINVAL;
    if (!m->lg) {
        for (int k = 0; k < 2 * m->B; k++) counts[k] = 0;
        return GADE_OK;
    }
    GADE_HIP(hipStreamSynchronize(m->stream));
    GADE_HIP(hipMemcpy(counts, m->lg->uses, sizeof(int) * 2 * m->B, hipMemcpyDeviceToHost));
    return GADE_OK;
}

int gade_lg_forward(gade_model_t *m, const float *x, int M, float *y) {
    if (!m || !x || !y || M <= 0 || M % 16) return GADE_EINVAL;
    if (!m->lg || !m->lg->net) return GADE_EINVAL;
    gade_lg *lg = m->lg;
    GADE_TRY(lg_train_alloc(lg, M));
    hipLaunchKernelGGL(k_lg_gather_rows, dim3(M), dim3(64), 0, m->stream, x, (const float *)lg->mean,
                       (const float *)lg->stdv, lg->xt);
    mlp_forward(lg, m->stream, M, lg->xt, lg->zt, lg->at, lg->yt);
    hipError_t e = hipMemcpyAsync(y, lg->yt, sizeof(float) * M * kOut, hipMemcpyDeviceToDevice, m->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(m->stream);
    if (e != hipSuccess) {
        set_error(std::string("gade_lg_forward: ") + hipGetErrorString(e));
        return GADE_EHIP;
    }
    return GADE_OK;
}

// loss and gradient (no update) of the batch idx [M] of samples: loss_out = mean loss, grad (device FP32
// [GADE_LG_PARAMS], may be null) = its gradient
int gade_lg_grad(gade_model_t *m, const float *samples, const int *idx, int M, float *loss_out, float *grad) {
    if (!m || !samples || !idx || M <= 0 || M % 16) return GADE_EINVAL;
    if (!m->lg || !m->lg->net) return GADE_EINVAL;
    gade_lg *lg = m->lg;
    GADE_TRY(lg_train_alloc(lg, M));
    float *d_mean = nullptr;
    GADE_TRY(hip_alloc((void **)&d_mean, sizeof(float)));
    hipLaunchKernelGGL(k_lg_gather, dim3(M), dim3(64), 0, m->stream, samples, idx, (const float *)lg->mean,
                       (const float *)lg->stdv, lg->xt);
    mlp_forward(lg, m->stream, M, lg->xt, lg->zt, lg->at, lg->yt);
    hipLaunchKernelGGL(k_lg_loss, dim3(M), dim3(64), 0, m->stream, (const float *)lg->yt, samples, idx, M, lg->dy,
                       lg->loss);
    hipLaunchKernelGGL(k_lg_mean, dim3(1), dim3(256), 0, m->stream, (const float *)lg->loss, M, d_mean);
    mlp_backward(lg, m->stream, M, lg->xt);
    hipError_t e = hipGetLastError();
    if (e == hipSuccess && grad)
        e = hipMemcpyAsync(grad, lg->g, sizeof(float) * P_TOTAL, hipMemcpyDeviceToDevice, m->stream);
    if (e == hipSuccess && loss_out)
        e = hipMemcpyAsync(loss_out, d_mean, sizeof(float), hipMemcpyDeviceToHost, m->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(m->stream);
    (void)hipFree(d_mean);
    if (e != hipSuccess) {
        set_error(std::string("gade_lg_grad: ") + hipGetErrorString(e));
        return GADE_EHIP;
    }
    return GADE_OK;
}

// nsteps Adam steps; step k uses the samples idx[k*M .. (k+1)*M) (device ints); loss_hist [nsteps] (host,
// may be null) receives each step's mean loss
int gade_lg_train(gade_model_t *m, const float *samples, const int *idx, int M, int nsteps, float lr,
                  float *loss_hist) {
    if (!m || !samples || !idx || M <= 0 || M % 16 || nsteps <= 0 || !(lr > 0)) return GADE_EINVAL;
    if (!m->lg || !m->lg->net) {
        set_error("gade_lg_train: initialise the network first (gade_lg_init)");
        return GADE_EINVAL;
    }
    gade_lg *lg = m->lg;
    GADE_TRY(lg_train_alloc(lg, M));
    gade::drop_step_graph(m);
    float *d_hist = nullptr;
    GADE_TRY(hip_alloc((void **)&d_hist, sizeof(float) * nsteps));
    const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
    for (int k = 0; k < nsteps; k++) {
        const int *ik = idx + (size_t)k * M;
        hipLaunchKernelGGL(k_lg_gather, dim3(M), dim3(64), 0, m->stream, samples, ik, (const float *)lg->mean,
                           (const float *)lg->stdv, lg->xt);
        mlp_forward(lg, m->stream, M, lg->xt, lg->zt, lg->at, lg->yt);
        hipLaunchKernelGGL(k_lg_loss, dim3(M), dim3(64), 0, m->stream, (const float *)lg->yt, samples, ik, M, lg->dy,
                           lg->loss);
        hipLaunchKernelGGL(k_lg_mean, dim3(1), dim3(256), 0, m->stream, (const float *)lg->loss, M, d_hist + k);
        mlp_backward(lg, m->stream, M, lg->xt);
        lg->adam_t++;
        const float c1 = 1.f - std::pow(b1, (float)lg->adam_t), c2 = 1.f - std::pow(b2, (float)lg->adam_t);
        hipLaunchKernelGGL(k_lg_adam, dim3(256), dim3(256), 0, m->stream, (long)P_TOTAL, lg->p, (const float *)lg->g,
                           lg->m1, lg->m2, lr, b1, b2, eps, c1, c2, lg->pbf);
    }
    hipError_t e = hipGetLastError();
    if (e == hipSuccess && loss_hist)
        e = hipMemcpyAsync(loss_hist, d_hist, sizeof(float) * nsteps, hipMemcpyDeviceToHost, m->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(m->stream);
    (void)hipFree(d_hist);
    if (e != hipSuccess) {
        set_error(std::string("gade_lg_train: ") + hipGetErrorString(e));
        return GADE_EHIP;
    }
    return GADE_OK;
}

}  // extern "C"
