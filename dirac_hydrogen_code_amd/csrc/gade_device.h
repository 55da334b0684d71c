// Device helpers and kernels shared by the libgade translation units
// (complex arithmetic, BiCGSTAB vector kernels, CSR SpMV).
#pragma once

#include <hip/hip_runtime.h>

// ---------------------------------------------------------------------------
// complex helpers (double2 = (re, im))
// ---------------------------------------------------------------------------
__device__ __forceinline__ double2 cadd(double2 a, double2 b) { return make_double2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ double2 csub(double2 a, double2 b) { return make_double2(a.x - b.x, a.y - b.y); }
__device__ __forceinline__ double2 cmul(double2 a, double2 b) {
    return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__device__ __forceinline__ double2 cscale(double2 a, double s) { return make_double2(a.x * s, a.y * s); }
// acc + a*b
__device__ __forceinline__ double2 cfma(double2 a, double2 b, double2 acc) {
    acc.x = fma(a.x, b.x, acc.x);
    acc.x = fma(-a.y, b.y, acc.x);
    acc.y = fma(a.x, b.y, acc.y);
    acc.y = fma(a.y, b.x, acc.y);
    return acc;
}
// acc + conj(a)*b
__device__ __forceinline__ double2 cfmaconj(double2 a, double2 b, double2 acc) {
    acc.x = fma(a.x, b.x, acc.x);
    acc.x = fma(a.y, b.y, acc.x);
    acc.y = fma(a.x, b.y, acc.y);
    acc.y = fma(-a.y, b.x, acc.y);
    return acc;
}
__device__ __forceinline__ double2 cdiv(double2 a, double2 b) {
    double d = b.x * b.x + b.y * b.y;
    return make_double2((a.x * b.x + a.y * b.y) / d, (a.y * b.x - a.x * b.y) / d);
}
__device__ __forceinline__ bool cnz(double2 a) { return a.x != 0.0 || a.y != 0.0; }
// the same for single precision (the inner solve of the mixed-precision form)
__device__ __forceinline__ float2 cadd(float2 a, float2 b) { return make_float2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ float2 csub(float2 a, float2 b) { return make_float2(a.x - b.x, a.y - b.y); }
__device__ __forceinline__ float2 cmul(float2 a, float2 b) {
    return make_float2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__device__ __forceinline__ float2 cscale(float2 a, float s) { return make_float2(a.x * s, a.y * s); }
__device__ __forceinline__ float2 cfma(float2 a, float2 b, float2 acc) {
    acc.x = fmaf(a.x, b.x, acc.x);
    acc.x = fmaf(-a.y, b.y, acc.x);
    acc.y = fmaf(a.x, b.y, acc.y);
    acc.y = fmaf(a.y, b.x, acc.y);
    return acc;
}
__device__ __forceinline__ float2 cfmaconj(float2 a, float2 b, float2 acc) {
    acc.x = fmaf(a.x, b.x, acc.x);
    acc.x = fmaf(a.y, b.y, acc.x);
    acc.y = fmaf(a.x, b.y, acc.y);
    acc.y = fmaf(-a.y, b.x, acc.y);
    return acc;
}
__device__ __forceinline__ bool cnz(float2 a) { return a.x != 0.0f || a.y != 0.0f; }

// complex type traits: real type and constructor of double2 / float2
template <class T> struct cx;
template <> struct cx<double2> {
    using R = double;
    static __device__ __forceinline__ double2 mk(double re, double im) { return make_double2(re, im); }
};
template <> struct cx<float2> {
    using R = float;
    static __device__ __forceinline__ float2 mk(float re, float im) { return make_float2(re, im); }
};
template <class T> __device__ __forceinline__ T czero() { return cx<T>::mk(0, 0); }
__device__ __forceinline__ float2 to_f32(double2 a) { return make_float2((float)a.x, (float)a.y); }
__device__ __forceinline__ double2 to_f64(float2 a) { return make_double2((double)a.x, (double)a.y); }
__device__ __forceinline__ double2 to_f64(double2 a) { return a; }
template <class T> __device__ __forceinline__ T from_f64(double2 a);
template <> __device__ __forceinline__ double2 from_f64<double2>(double2 a) { return a; }
template <> __device__ __forceinline__ float2 from_f64<float2>(double2 a) { return to_f32(a); }

typedef double dvec2 __attribute__((ext_vector_type(2)));
// streaming (read-once) load of a complex value
__device__ __forceinline__ double2 ldnt(const double2 *p) {
    dvec2 v = __builtin_nontemporal_load((const dvec2 *)p);
    return make_double2(v.x, v.y);
}

// ---------------------------------------------------------------------------
// gated launches: a kernel given a non-null 'gate' returns at once when the
// device word it points at is non-zero.  With several instances (a batch of
// pulses sharing the couplings) each instance has its own words, 'gstride'
// ints apart, and a workgroup checks the word of the instance it serves.
// The device-driven CN stepping uses this to enqueue a step's iterations
// ahead of knowing how many each instance needs.
// ---------------------------------------------------------------------------
__device__ __forceinline__ bool gated(const int *gate, int gstride, int b) {
    return gate != nullptr && gate[(size_t)b * gstride] != 0;
}
#define GADE_GATED_B(gate, gstride, b)             \
    do {                                           \
        if (gated((gate), (gstride), (b))) return; \
    } while (0)
#define GADE_GATED(gate) GADE_GATED_B(gate, 0, 0)

// pulse coefficients c_alpha = SoL * (i h dt / 2) * f_alpha of one instance (alpha = 0..5)
struct Coef6 {
    double2 c[6];
};

// per-instance control words of the device-driven solve (int, CT_STRIDE apart)
enum {
    CT_STATUS = 0,  // 0 iterating, 1 converged, 2 needs the host (restart, zero rhs, non-finite), 3 maxit reached
    CT_ITER = 1,    // completed iterations of the current solve
    CT_SKIP = 2,    // iteration gate: set once the solve is decided, the instance is idle or the run stalled
    CT_PSKIP = 3,   // step-prologue gate: the instance is idle this step or the run stalled
    CT_HSEL = 4,    // gate of work the host enqueues for selected instances (0 = selected)
    CT_MAXIT = 5,
    CT_HIST = 6,    // previous solutions held in the guess history
    CT_ISKIP = 7,   // inner-solve gate of the mixed-precision form (set once the inner solve stops)
    CT_STRIDE = 8
};
// run-wide words (int, in the global control array)
enum { GC_STALLED = 0, GC_STALL_STEP = 1, GC_STEP = 2, GC_WSLOT = 3, GC_COUNT = 4 };

constexpr int kMaxGuessOrder = 3;  // polynomial extrapolation order of the initial guess
constexpr int kHistSlots = 6;  // previous solutions kept (the learned guess mixes psi_n..psi_{n-5})

// Initial guess by polynomial extrapolation of the step history:
// x0 = sum_{k=0..p} c_k psi_{n-k}, c_k = (-1)^k binom(p+1, k+1) (Lagrange at
// t_{n+1} through equally spaced t_n..t_{n-p}), with p = min(order, history).
// psi holds psi_n on entry and x0 on exit; psi_n enters the history ring.
// blockIdx.y = instance; hist is [slot][total vector].
[[maybe_unused]] static __global__ __launch_bounds__(256) void k_extrap(long long N1, long long Ntot,
                                                                       double2 *__restrict__ psi,
                                                                       double2 *__restrict__ hist, int order,
                                                                       const int *__restrict__ ctl, const int *gate,
                                                                       int gstride) {
    const int b = blockIdx.y;
    GADE_GATED_B(gate, gstride, b);
    const int have = ctl[(size_t)b * CT_STRIDE + CT_HIST];
    const int p = order < have ? order : have;
    // ring slot of psi_{n-k} before the push: (have - k) mod kHistSlots; psi_n goes to slot have mod kHistSlots
    const int slot0 = have % kHistSlots;
    double c[kMaxGuessOrder + 1];
    {
        // binom(p+1, k+1) with alternating signs
        double bnm = p + 1;
        for (int k = 0; k <= kMaxGuessOrder; k++) {
            c[k] = (k <= p) ? ((k % 2) ? -bnm : bnm) : 0.0;
            bnm = bnm * (p - k) / (k + 2);
        }
    }
    for (long long j = blockIdx.x * (long long)blockDim.x + threadIdx.x; j < N1; j += (long long)gridDim.x * blockDim.x) {
        const long long i = (long long)b * N1 + j;
        const double2 xn = psi[i];
        double2 x = cscale(xn, c[0]);
#pragma unroll
        for (int k = 1; k <= kMaxGuessOrder; k++) {
            if (k > p) break;
            const int sl = ((slot0 - k) % kHistSlots + kHistSlots) % kHistSlots;
            x = cadd(x, cscale(hist[(size_t)sl * Ntot + i], c[k]));
        }
        hist[(size_t)slot0 * Ntot + i] = xn;
        psi[i] = x;
    }
}

// history count update after k_extrap (one thread per instance, ordered after it on the stream)
[[maybe_unused]] static __global__ void k_hist_count(int *ctl, const int *gate, int gstride) {
    const int b = blockIdx.x;
    GADE_GATED_B(gate, gstride, b);
    if (threadIdx.x == 0) ctl[(size_t)b * CT_STRIDE + CT_HIST] += 1;
}

// ---------------------------------------------------------------------------
// BiCGSTAB vector kernels with deterministic two-kernel reductions, one
// instance per blockIdx.y (k_bicg) or blockIdx.x (k_bicg_fin)
// ---------------------------------------------------------------------------
// scalar slots in d_sc (doubles, SC_STRIDE per instance)
enum {
    SC_RHO = 0,        // 2
    SC_ALPHA = 2,      // 2
    SC_W = 4,          // 2
    SC_R0SQ = 6,
    SC_RHSSQ = 7,
    SC_RNORM2 = 8,
    SC_RHONEXT = 9,    // 2
    SC_TOL = 11,       // requested relative tolerance (device-driven solve)
    SC_TOL2 = 12,      // tol^2 |rhs|^2
    SC_COUNT = 16,
    SC_STRIDE = 16
};
constexpr double kRestartEps2 = 1e-24;  // restart when |rho| < 1e-24 |r0|^2 (tools/gpu_bicgstab.c:195)
enum RedOp { OP_INIT = 0, OP_RESTART, OP_P, OP_ALPHA, OP_S, OP_W, OP_XR };

constexpr int kRedThreads = 256;

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}

// Block-reduce NV values; thread 0 returns the sums in out[].
template <int NV>
__device__ __forceinline__ void block_sum(double (&v)[NV], double *sh) {
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
#pragma unroll
    for (int i = 0; i < NV; i++) v[i] = wave_sum(v[i]);
    if (lane == 0)
#pragma unroll
        for (int i = 0; i < NV; i++) sh[wid * NV + i] = v[i];
    __syncthreads();
    if (threadIdx.x == 0) {
        const int nw = blockDim.x >> 6;
#pragma unroll
        for (int i = 0; i < NV; i++) {
            double s = 0;
            for (int w2 = 0; w2 < nw; w2++) s += sh[w2 * NV + i];
            v[i] = s;
        }
    }
    __syncthreads();
}

// One BiCGSTAB vector step over the N1 entries of instance blockIdx.y (vectors
// are instance-major, N1 apart); partial sums to part[instance][block][4].
// FOLD (OP_S, OP_XR): the scalar this step needs (alpha, w) is finished here from the previous reduction's
// partials fold_part[instance][fold_nparts][4] instead of by its own k_bicg_fin: every block sums them in
// k_bicg_fin's order (the same value bit for bit) and block 0 records it (and rho for OP_S) in the scalars.
template <int OP, class T = double2, bool FOLD = false>
__global__ __launch_bounds__(kRedThreads) void k_bicg(long long N1, const T *__restrict__ b, T *__restrict__ r,
                                                       T *__restrict__ r0, T *__restrict__ p,
                                                       const T *__restrict__ v, T *__restrict__ s,
                                                       const T *__restrict__ t, const T *__restrict__ yv,
                                                       const T *__restrict__ zv, T *__restrict__ x,
                                                       double *__restrict__ sc0, double *__restrict__ part0,
                                                       const int *gate = nullptr, int gstride = 0,
                                                       const double *__restrict__ fold_part = nullptr,
                                                       int fold_nparts = 0) {
    static_assert(!FOLD || OP == OP_S || OP == OP_XR, "folded finish: alpha into s, w into x/r");
    const int inst = blockIdx.y;
    GADE_GATED_B(gate, gstride, inst);
    __shared__ double sh[4 * 16];
    double *sc = sc0 + (size_t)inst * SC_STRIDE;
    double *part = part0 + ((size_t)inst * gridDim.x) * 4;
    const long long o = (long long)inst * N1;
    double acc[4] = {0, 0, 0, 0};
    double2 beta = make_double2(0, 0), w = make_double2(0, 0), alpha = make_double2(0, 0);
    if (OP == OP_P) {
        // beta = (rho_new / rho_old) * (alpha / w)
        double2 rho_new = make_double2(sc[SC_RHONEXT], sc[SC_RHONEXT + 1]);
        double2 rho_old = make_double2(sc[SC_RHO], sc[SC_RHO + 1]);
        alpha = make_double2(sc[SC_ALPHA], sc[SC_ALPHA + 1]);
        w = make_double2(sc[SC_W], sc[SC_W + 1]);
        beta = cmul(cdiv(rho_new, rho_old), cdiv(alpha, w));
    }
    if constexpr (FOLD) {
        __shared__ double2 fs;
        double tot[4] = {0, 0, 0, 0};
        const double *fp = fold_part + (size_t)inst * fold_nparts * 4;
        for (int bidx = threadIdx.x; bidx < fold_nparts; bidx += blockDim.x)
            for (int i = 0; i < 4; i++) tot[i] += fp[(size_t)bidx * 4 + i];
        block_sum<4>(tot, sh);
        if (threadIdx.x == 0) {
            if (OP == OP_S) {  // k_bicg_fin<OP_ALPHA>
                const double2 rho = make_double2(sc[SC_RHONEXT], sc[SC_RHONEXT + 1]);
                const double2 al = cdiv(rho, make_double2(tot[0], tot[1]));
                fs = al;
                if (blockIdx.x == 0) {
                    sc[SC_ALPHA] = al.x;
                    sc[SC_ALPHA + 1] = al.y;
                    sc[SC_RHO] = rho.x;
                    sc[SC_RHO + 1] = rho.y;
                }
            } else {  // k_bicg_fin<OP_W>
                double2 ww = make_double2(0, 0);
                if (tot[0] > 0) ww = make_double2(tot[1] / tot[0], tot[2] / tot[0]);
                fs = ww;
                if (blockIdx.x == 0) {
                    sc[SC_W] = ww.x;
                    sc[SC_W + 1] = ww.y;
                }
            }
        }
        __syncthreads();
        if (OP == OP_S) alpha = fs;
        if (OP == OP_XR) {
            alpha = make_double2(sc[SC_ALPHA], sc[SC_ALPHA + 1]);
            w = fs;
        }
    } else {
        if (OP == OP_S || OP == OP_XR) alpha = make_double2(sc[SC_ALPHA], sc[SC_ALPHA + 1]);
        if (OP == OP_XR) w = make_double2(sc[SC_W], sc[SC_W + 1]);
    }
    // the vector updates run in T, the dot products accumulate in double
    const T betaT = from_f64<T>(beta), wT = from_f64<T>(w), alphaT = from_f64<T>(alpha);
    for (long long j = blockIdx.x * (long long)blockDim.x + threadIdx.x; j < N1; j += (long long)gridDim.x * blockDim.x) {
        const long long i = o + j;
        if (OP == OP_INIT || OP == OP_RESTART) {
            const T riT = csub(b[i], r[i]);
            r[i] = riT;
            r0[i] = riT;
            const double2 ri = to_f64(riT);
            acc[0] = fma(ri.x, ri.x, fma(ri.y, ri.y, acc[0]));
            if (OP == OP_INIT) {
                const double2 bi = to_f64(b[i]);
                acc[1] = fma(bi.x, bi.x, fma(bi.y, bi.y, acc[1]));
            }
        } else if (OP == OP_P) {
            const T pi = csub(p[i], cmul(wT, v[i]));
            p[i] = cadd(r[i], cmul(betaT, pi));
        } else if (OP == OP_ALPHA) {
            double2 d = cfmaconj(to_f64(r0[i]), to_f64(v[i]), make_double2(acc[0], acc[1]));
            acc[0] = d.x;
            acc[1] = d.y;
        } else if (OP == OP_S) {
            s[i] = csub(r[i], cmul(alphaT, v[i]));
        } else if (OP == OP_W) {
            const double2 ti = to_f64(t[i]);
            acc[0] = fma(ti.x, ti.x, fma(ti.y, ti.y, acc[0]));
            double2 d = cfmaconj(ti, to_f64(s[i]), make_double2(acc[1], acc[2]));
            acc[1] = d.x;
            acc[2] = d.y;
        } else if (OP == OP_XR) {
            T xi = x[i];
            xi = cadd(xi, cadd(cmul(alphaT, yv[i]), cmul(wT, zv[i])));
            x[i] = xi;
            const T riT = csub(s[i], cmul(wT, t[i]));
            r[i] = riT;
            const double2 ri = to_f64(riT);
            acc[0] = fma(ri.x, ri.x, fma(ri.y, ri.y, acc[0]));
            double2 d = cfmaconj(to_f64(r0[i]), ri, make_double2(acc[1], acc[2]));
            acc[1] = d.x;
            acc[2] = d.y;
        }
    }
    if (OP == OP_P || OP == OP_S) return;
    block_sum<4>(acc, sh);
    if (threadIdx.x == 0)
        for (int i = 0; i < 4; i++) part[(size_t)blockIdx.x * 4 + i] = acc[i];
}

// The last step of an apply fused with the BiCGSTAB reduction that follows it: out = sum over the lgrid
// multipole partials of the radial phase (in multipole order, as k_radial_sum with beta 0), then for
// OP_ALPHA the partial r0 . out, for OP_W the partials |out|^2 and out . s (out = t).  The grid and every
// thread's order of entries are k_bicg's, so out and the partial sums equal the unfused pair's bit for bit.
template <int OP>
__global__ __launch_bounds__(kRedThreads) void k_sum_bicg(long long N1, long long N, int lgrid,
                                                           const double2 *__restrict__ ypart,
                                                           double2 *__restrict__ out, const double2 *__restrict__ r0,
                                                           const double2 *__restrict__ s, double *__restrict__ part0,
                                                           const int *gate, int gstride) {
    static_assert(OP == OP_ALPHA || OP == OP_W, "fused sums: alpha or w");
    const int inst = blockIdx.y;
    GADE_GATED_B(gate, gstride, inst);
    __shared__ double sh[4 * 16];
    double *part = part0 + ((size_t)inst * gridDim.x) * 4;
    const long long o = (long long)inst * N1;
    double acc[4] = {0, 0, 0, 0};
    for (long long j = blockIdx.x * (long long)blockDim.x + threadIdx.x; j < N1; j += (long long)gridDim.x * blockDim.x) {
        const long long i = o + j;
        double2 v = ypart[i];
        for (int l = 1; l < lgrid; l++) v = cadd(v, ypart[(size_t)l * N + i]);
        out[i] = v;
        if (OP == OP_ALPHA) {
            double2 d = cfmaconj(r0[i], v, make_double2(acc[0], acc[1]));
            acc[0] = d.x;
            acc[1] = d.y;
        } else {
            acc[0] = fma(v.x, v.x, fma(v.y, v.y, acc[0]));
            double2 d = cfmaconj(v, s[i], make_double2(acc[1], acc[2]));
            acc[1] = d.x;
            acc[2] = d.y;
        }
    }
    block_sum<4>(acc, sh);
    if (threadIdx.x == 0)
        for (int i = 0; i < 4; i++) part[(size_t)blockIdx.x * 4 + i] = acc[i];
}

// The paired apply's two multipole sums fused with OP_INIT: b = sum of ypart (the rhs (S - H) x),
// ax = sum of ypart2 (A x0 = (S + H) x), r = r0 = b - ax, partials |r|^2 and |b|^2.  Each sum runs in
// multipole order and the grid and every thread's order of entries are k_bicg's: the same results as two
// k_radial_sum launches then k_bicg<OP_INIT>, bit for bit.
[[maybe_unused]] static __global__ __launch_bounds__(kRedThreads) void k_sum2_init(long long N1, long long N, int lgrid,
                                                            const double2 *__restrict__ ypart,
                                                            const double2 *__restrict__ ypart2,
                                                            double2 *__restrict__ b, double2 *__restrict__ r,
                                                            double2 *__restrict__ r0, double *__restrict__ part0,
                                                            const int *gate, int gstride) {
    const int inst = blockIdx.y;
    GADE_GATED_B(gate, gstride, inst);
    __shared__ double sh[4 * 16];
    double *part = part0 + ((size_t)inst * gridDim.x) * 4;
    const long long o = (long long)inst * N1;
    double acc[4] = {0, 0, 0, 0};
    for (long long j = blockIdx.x * (long long)blockDim.x + threadIdx.x; j < N1; j += (long long)gridDim.x * blockDim.x) {
        const long long i = o + j;
        double2 bs = ypart[i], ax = ypart2[i];
        for (int l = 1; l < lgrid; l++) {
            bs = cadd(bs, ypart[(size_t)l * N + i]);
            ax = cadd(ax, ypart2[(size_t)l * N + i]);
        }
        b[i] = bs;
        const double2 ri = csub(bs, ax);
        r[i] = ri;
        r0[i] = ri;
        acc[0] = fma(ri.x, ri.x, fma(ri.y, ri.y, acc[0]));
        acc[1] = fma(bs.x, bs.x, fma(bs.y, bs.y, acc[1]));
    }
    block_sum<4>(acc, sh);
    if (threadIdx.x == 0)
        for (int i = 0; i < 4; i++) part[(size_t)blockIdx.x * 4 + i] = acc[i];
}

// Sums the partials of instance blockIdx.x in a fixed order and updates its
// scalars; with ctl, also takes the host loop's decisions for it.
// INNER (the FP32 inner solve of the mixed-precision form): the only decision is when the inner
// solve stops (CT_ISKIP), and CT_ITER counts every inner iteration of the step.
template <int OP, bool INNER = false>
__global__ __launch_bounds__(kRedThreads) void k_bicg_fin(int nparts, const double *__restrict__ part0, double *sc0,
                                                           const int *gate = nullptr, int gstride = 0,
                                                           int *ctl0 = nullptr) {
    const int inst = blockIdx.x;
    GADE_GATED_B(gate, gstride, inst);
    __shared__ double sh[4 * 16];
    const double *part = part0 + (size_t)inst * nparts * 4;
    double *sc = sc0 + (size_t)inst * SC_STRIDE;
    double tot[4] = {0, 0, 0, 0};
    for (int bidx = threadIdx.x; bidx < nparts; bidx += blockDim.x)
        for (int i = 0; i < 4; i++) tot[i] += part[(size_t)bidx * 4 + i];
    block_sum<4>(tot, sh);
    if (threadIdx.x != 0) return;
    if (OP == OP_INIT) {
        sc[SC_R0SQ] = tot[0];
        sc[SC_RHSSQ] = tot[1];
        sc[SC_RNORM2] = tot[0];
        sc[SC_RHONEXT] = tot[0];
        sc[SC_RHONEXT + 1] = 0;
        sc[SC_RHO] = 1; sc[SC_RHO + 1] = 0;
        sc[SC_ALPHA] = 1; sc[SC_ALPHA + 1] = 0;
        sc[SC_W] = 1; sc[SC_W + 1] = 0;
    } else if (OP == OP_RESTART) {
        sc[SC_R0SQ] = tot[0];
        sc[SC_RNORM2] = tot[0];
        sc[SC_RHONEXT] = tot[0];
        sc[SC_RHONEXT + 1] = 0;
    } else if (OP == OP_ALPHA) {
        double2 rho = make_double2(sc[SC_RHONEXT], sc[SC_RHONEXT + 1]);
        double2 al = cdiv(rho, make_double2(tot[0], tot[1]));
        sc[SC_ALPHA] = al.x;
        sc[SC_ALPHA + 1] = al.y;
        sc[SC_RHO] = rho.x;
        sc[SC_RHO + 1] = rho.y;
    } else if (OP == OP_W) {
        double2 ww = make_double2(0, 0);
        if (tot[0] > 0) ww = make_double2(tot[1] / tot[0], tot[2] / tot[0]);
        sc[SC_W] = ww.x;
        sc[SC_W + 1] = ww.y;
    } else if (OP == OP_XR) {
        sc[SC_RNORM2] = tot[0];
        sc[SC_RHONEXT] = tot[1];
        sc[SC_RHONEXT + 1] = tot[2];
    }
    if (ctl0 == nullptr) return;
    int *ctl = ctl0 + (size_t)inst * CT_STRIDE;
    if constexpr (INNER) {
        if (OP == OP_XR) {
            const int it = ctl[CT_ITER] + 1;
            ctl[CT_ITER] = it;
            if (!isfinite(tot[0]) || tot[0] < sc[SC_TOL2] || it >= ctl[CT_MAXIT] ||
                hypot(tot[1], tot[2]) < kRestartEps2 * sc[SC_R0SQ])
                ctl[CT_ISKIP] = 1;
        }
        return;
    }
    // device-driven solve: the same decisions as the host loop (break-before-increment)
    int st = -1;
    if (OP == OP_INIT) {
        const double tol = sc[SC_TOL];
        const double tol2 = tol * tol * tot[1];
        sc[SC_TOL2] = tol2;
        st = (tot[1] == 0) ? 2 : (tot[0] < tol2 ? 1 : 0);
        ctl[CT_ITER] = 0;
    } else if (OP == OP_RESTART) {
        st = 0;
    } else if (OP == OP_XR) {
        st = 0;
        if (!isfinite(tot[0])) {
            st = 2;
        } else if (tot[0] < sc[SC_TOL2]) {
            st = 1;
        } else {
            const int it = ctl[CT_ITER] + 1;
            ctl[CT_ITER] = it;
            if (it >= ctl[CT_MAXIT]) st = 3;
            // the top of the next iteration restarts when |rho| < eps^2 |r0|^2: left to the host
            else if (hypot(tot[1], tot[2]) < kRestartEps2 * sc[SC_R0SQ]) st = 2;
        }
    }
    if (st >= 0) {
        ctl[CT_STATUS] = st;
        ctl[CT_SKIP] = st != 0;
    }
}

// ---------------------------------------------------------------------------
// mixed precision: FP64 defect correction around an FP32 inner BiCGSTAB.
// Per sweep: d = A^{-1} r approximately in FP32 (from d = 0), x += d in FP64,
// r = b - A x in FP64, and the FP64 residual decides convergence.
// ---------------------------------------------------------------------------
// inner start of instance blockIdx.y: ri = r0i = fp32(r), d = p = v = 0; partial |ri|^2
[[maybe_unused]] static __global__ __launch_bounds__(kRedThreads) void k_mp_start(
    long long N1, const double2 *__restrict__ r, float2 *__restrict__ ri, float2 *__restrict__ r0i,
    float2 *__restrict__ d, float2 *__restrict__ p, float2 *__restrict__ v, double *__restrict__ part0,
    const int *gate, int gstride) {
    const int inst = blockIdx.y;
    GADE_GATED_B(gate, gstride, inst);
    __shared__ double sh[4 * 16];
    double *part = part0 + ((size_t)inst * gridDim.x) * 4;
    double acc[4] = {0, 0, 0, 0};
    for (long long j = blockIdx.x * (long long)blockDim.x + threadIdx.x; j < N1; j += (long long)gridDim.x * blockDim.x) {
        const long long i = (long long)inst * N1 + j;
        const float2 rv = to_f32(r[i]);
        ri[i] = rv;
        r0i[i] = rv;
        d[i] = make_float2(0.f, 0.f);
        p[i] = make_float2(0.f, 0.f);  // a fresh Krylov start for every sweep
        v[i] = make_float2(0.f, 0.f);
        const double2 vd = to_f64(rv);
        acc[0] = fma(vd.x, vd.x, fma(vd.y, vd.y, acc[0]));
    }
    block_sum<4>(acc, sh);
    if (threadIdx.x == 0)
        for (int k = 0; k < 4; k++) part[(size_t)blockIdx.x * 4 + k] = acc[k];
}

// inner scalars of instance blockIdx.x (as OP_INIT, with the inner tolerance), and its inner gate:
// closed when the outer solve is decided or the rounded residual is zero
[[maybe_unused]] static __global__ __launch_bounds__(kRedThreads) void k_mp_start_fin(
    int nparts, const double *__restrict__ part0, double *sci0, int *ctl0, double tol_in) {
    const int inst = blockIdx.x;
    __shared__ double sh[4 * 16];
    const double *part = part0 + (size_t)inst * nparts * 4;
    double tot[4] = {0, 0, 0, 0};
    for (int bidx = threadIdx.x; bidx < nparts; bidx += blockDim.x) tot[0] += part[(size_t)bidx * 4];
    block_sum<4>(tot, sh);
    if (threadIdx.x != 0) return;
    int *ctl = ctl0 + (size_t)inst * CT_STRIDE;
    if (ctl[CT_SKIP] != 0) {
        ctl[CT_ISKIP] = 1;
        return;
    }
    double *sc = sci0 + (size_t)inst * SC_STRIDE;
    sc[SC_R0SQ] = tot[0];
    sc[SC_RHSSQ] = tot[0];
    sc[SC_RNORM2] = tot[0];
    sc[SC_RHONEXT] = tot[0];
    sc[SC_RHONEXT + 1] = 0;
    sc[SC_RHO] = 1; sc[SC_RHO + 1] = 0;
    sc[SC_ALPHA] = 1; sc[SC_ALPHA + 1] = 0;
    sc[SC_W] = 1; sc[SC_W + 1] = 0;
    sc[SC_TOL2] = tol_in * tol_in * tot[0];
    ctl[CT_ISKIP] = tot[0] > 0 ? 0 : 1;
}

// x += d (FP64 update with the FP32 correction), instance blockIdx.y
[[maybe_unused]] static __global__ __launch_bounds__(256) void k_mp_update(long long N1, double2 *__restrict__ x,
                                                                          const float2 *__restrict__ d,
                                                                          const int *gate, int gstride) {
    const int inst = blockIdx.y;
    GADE_GATED_B(gate, gstride, inst);
    for (long long j = blockIdx.x * (long long)blockDim.x + threadIdx.x; j < N1; j += (long long)gridDim.x * blockDim.x) {
        const long long i = (long long)inst * N1 + j;
        x[i] = cadd(x[i], to_f64(d[i]));
    }
}

// r = b - A x (FP64) and partial |r|^2, instance blockIdx.y
[[maybe_unused]] static __global__ __launch_bounds__(kRedThreads) void k_mp_resid(
    long long N1, const double2 *__restrict__ b, const double2 *__restrict__ ax, double2 *__restrict__ r,
    double *__restrict__ part0, const int *gate, int gstride) {
    const int inst = blockIdx.y;
    GADE_GATED_B(gate, gstride, inst);
    __shared__ double sh[4 * 16];
    double *part = part0 + ((size_t)inst * gridDim.x) * 4;
    double acc[4] = {0, 0, 0, 0};
    for (long long j = blockIdx.x * (long long)blockDim.x + threadIdx.x; j < N1; j += (long long)gridDim.x * blockDim.x) {
        const long long i = (long long)inst * N1 + j;
        const double2 v = csub(b[i], ax[i]);
        r[i] = v;
        acc[0] = fma(v.x, v.x, fma(v.y, v.y, acc[0]));
    }
    block_sum<4>(acc, sh);
    if (threadIdx.x == 0)
        for (int k = 0; k < 4; k++) part[(size_t)blockIdx.x * 4 + k] = acc[k];
}

// outer decision of instance blockIdx.x after a sweep: converged (FP64 tolerance), non-finite (host),
// iteration budget spent, or another sweep
[[maybe_unused]] static __global__ __launch_bounds__(kRedThreads) void k_mp_resid_fin(
    int nparts, const double *__restrict__ part0, double *sc0, int *ctl0, const int *gate, int gstride) {
    const int inst = blockIdx.x;
    GADE_GATED_B(gate, gstride, inst);
    __shared__ double sh[4 * 16];
    const double *part = part0 + (size_t)inst * nparts * 4;
    double tot[4] = {0, 0, 0, 0};
    for (int bidx = threadIdx.x; bidx < nparts; bidx += blockDim.x) tot[0] += part[(size_t)bidx * 4];
    block_sum<4>(tot, sh);
    if (threadIdx.x != 0) return;
    double *sc = sc0 + (size_t)inst * SC_STRIDE;
    int *ctl = ctl0 + (size_t)inst * CT_STRIDE;
    sc[SC_RNORM2] = tot[0];
    const int st = !isfinite(tot[0]) ? 2 : (tot[0] < sc[SC_TOL2] ? 1 : (ctl[CT_ITER] >= ctl[CT_MAXIT] ? 3 : 0));
    ctl[CT_STATUS] = st;
    ctl[CT_SKIP] = st != 0;
}

// y = fp32(x) over n entries, blockIdx.y = instance of 'stride' entries (gated per instance)
[[maybe_unused]] static __global__ __launch_bounds__(256) void k_to_f32(long long n, size_t stride,
                                                                       const double2 *__restrict__ x,
                                                                       float2 *__restrict__ y, const int *gate,
                                                                       int gstride) {
    GADE_GATED_B(gate, gstride, blockIdx.y);
    const size_t o = (size_t)blockIdx.y * stride;
    for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x)
        y[o + i] = to_f32(x[o + i]);
}

// Start of a device-driven CN step for instance blockIdx.x: idle when the step
// is past its pulse or the run has stalled.
// (step < 0: the step index is the run-wide GC_STEP word, as in a replayed step graph)
[[maybe_unused]] static __global__ void k_step_begin(int *ctl0, const int *__restrict__ gctl, int step,
                                                     const int *__restrict__ nsteps) {
    if (threadIdx.x != 0) return;
    if (step < 0) step = gctl[GC_STEP];
    const int b = blockIdx.x;
    int *ctl = ctl0 + (size_t)b * CT_STRIDE;
    const int idle = gctl[GC_STALLED] != 0 || (nsteps != nullptr && step >= nsteps[b]);
    ctl[CT_PSKIP] = idle;
    if (idle) ctl[CT_SKIP] = 1;
}

// End of a device-driven CN step: record iterations and tolerance of each
// finished solve, or mark the run stalled at this step (any instance unfinished)
// so the host continues it.
[[maybe_unused]] static __global__ void k_step_end(int *ctl0, int *gctl, const double *__restrict__ sc0, int step,
                                                   int nsteps_cap, int *iters, double *tols) {
    if (threadIdx.x != 0) return;
    if (step < 0) step = gctl[GC_STEP];
    const int b = blockIdx.x;
    const int *ctl = ctl0 + (size_t)b * CT_STRIDE;
    const double *sc = sc0 + (size_t)b * SC_STRIDE;
    if (ctl[CT_PSKIP] != 0 || gctl[GC_STALLED] != 0) return;
    const int st = ctl[CT_STATUS];
    if (st == 1 || st == 3) {
        iters[(size_t)b * nsteps_cap + step] = ctl[CT_ITER];
        tols[(size_t)b * nsteps_cap + step] = sqrt(sc[SC_RNORM2] / sc[SC_RHSSQ]);
    } else {
        gctl[GC_STALLED] = 1;  // every unfinished instance writes the same values
        gctl[GC_STALL_STEP] = step;
    }
}

// End of the extra iterations the host enqueues for a stalled step: record every decided solve.
[[maybe_unused]] static __global__ void k_step_end_resume(int *ctl0, int *gctl, const double *__restrict__ sc0,
                                                          int step, int nsteps_cap, int *iters, double *tols) {
    if (threadIdx.x != 0) return;
    const int b = blockIdx.x;
    const int *ctl = ctl0 + (size_t)b * CT_STRIDE;
    const double *sc = sc0 + (size_t)b * SC_STRIDE;
    const int st = ctl[CT_STATUS];
    if (st == 1 || st == 3) {
        iters[(size_t)b * nsteps_cap + step] = ctl[CT_ITER];
        tols[(size_t)b * nsteps_cap + step] = sc[SC_RHSSQ] > 0 ? sqrt(sc[SC_RNORM2] / sc[SC_RHSSQ]) : 0.0;
    }
    (void)gctl;
}

// Next step of a window (after k_stall_spread): advances the run-wide step words.
[[maybe_unused]] static __global__ void k_step_advance(int *gctl) {
    if (threadIdx.x == 0) {
        gctl[GC_STEP] += 1;
        gctl[GC_WSLOT] += 1;
    }
}
[[maybe_unused]] static __global__ void k_window_begin(int *gctl, int step) {
    if (threadIdx.x == 0) {
        gctl[GC_STEP] = step;
        gctl[GC_WSLOT] = 0;
    }
}

// After k_step_end: once the run has stalled, every instance idles.
[[maybe_unused]] static __global__ void k_stall_spread(int *ctl0, const int *__restrict__ gctl) {
    if (threadIdx.x != 0 || gctl[GC_STALLED] == 0) return;
    int *ctl = ctl0 + (size_t)blockIdx.x * CT_STRIDE;
    ctl[CT_SKIP] = 1;
    ctl[CT_PSKIP] = 1;
}

// y = alpha A x + beta y for a CSR matrix: one wave per row, streaming loads.
[[maybe_unused]] static __global__ __launch_bounds__(256) void k_csr_spmv(const int *__restrict__ Ap, const int *__restrict__ Ai,
                                                  const double2 *__restrict__ Ax, const double2 *__restrict__ x,
                                                  long long nrows, double2 alpha, double2 beta,
                                                  double2 *__restrict__ y) {
    const long long row = (blockIdx.x * (long long)blockDim.x + threadIdx.x) >> 6;
    const int lane = threadIdx.x & 63;
    if (row >= nrows) return;
    const int b = Ap[row], e = Ap[row + 1];
    double2 acc0 = make_double2(0, 0), acc1 = make_double2(0, 0);
    int i = b + lane;
    for (; i + 64 < e; i += 128) {
        double2 a0 = ldnt(Ax + i);
        int c0 = __builtin_nontemporal_load(Ai + i);
        double2 a1 = ldnt(Ax + i + 64);
        int c1 = __builtin_nontemporal_load(Ai + i + 64);
        acc0 = cfma(a0, x[c0], acc0);
        acc1 = cfma(a1, x[c1], acc1);
    }
    if (i < e) {
        double2 a0 = ldnt(Ax + i);
        int c0 = __builtin_nontemporal_load(Ai + i);
        acc0 = cfma(a0, x[c0], acc0);
    }
    double re = wave_sum(acc0.x + acc1.x), im = wave_sum(acc0.y + acc1.y);
    if (lane == 0) {
        double2 res = cmul(alpha, make_double2(re, im));
        if (beta.x != 0.0 || beta.y != 0.0) res = cadd(res, cmul(beta, y[row]));
        y[row] = res;
    }
}


// ---------------------------------------------------------------------------
// preconditioner: banded LU per distinct kappa (interleaved radial order)
// ---------------------------------------------------------------------------
// AB[q][i][w], w = 2*band+1, A(i,j) at AB[i][j-i+band].  No pivoting: P has a
// positive-definite Hermitian part (S), so the factorisation exists.
[[maybe_unused]] static __global__ __launch_bounds__(256) void k_band_lu(double2 *AB, int n, int band, int *flag) {
    const int q = blockIdx.x;
    const int w = 2 * band + 1;
    double2 *A = AB + (size_t)q * n * w;
    for (int k = 0; k < n; k++) {
        const double2 piv = A[(size_t)k * w + band];
        if (threadIdx.x == 0 && !cnz(piv)) atomicOr(flag, 1);
        const int iend = min(n - 1, k + band);
        for (int t = threadIdx.x; t < iend - k; t += blockDim.x) {
            const int i = k + 1 + t;
            double2 *lik = A + (size_t)i * w + (k - i + band);
            *lik = cdiv(*lik, piv);
        }
        __syncthreads();
        const int m = iend - k;
        for (int idx = threadIdx.x; idx < m * m; idx += blockDim.x) {
            const int i = k + 1 + idx / m, j = k + 1 + idx % m;
            const double2 lik = A[(size_t)i * w + (k - i + band)];
            const double2 ukj = A[(size_t)k * w + (j - k + band)];
            double2 *aij = A + (size_t)i * w + (j - i + band);
            *aij = csub(*aij, cmul(lik, ukj));
        }
        __syncthreads();
    }
}

// Per (kappa, chunk): inverses of the CHxCH diagonal chunks of L (unit) and U,
// stored [col][row], and the band coupling to the neighbouring chunk.
template <int CH>
__global__ __launch_bounds__(CH) void k_chunk_inverse(const double2 *AB, int n, int band, int nch, double2 *Linv,
                                                      double2 *Uinv, double2 *Loff, double2 *Uoff) {
    const int q = blockIdx.x, ch = blockIdx.y, j = threadIdx.x;
    const int w = 2 * band + 1;
    const double2 *A = AB + (size_t)q * n * w;
    const int base = ch * CH;
    double2 *Li = Linv + ((size_t)q * nch + ch) * CH * CH + (size_t)j * CH;
    double2 *Ui = Uinv + ((size_t)q * nch + ch) * CH * CH + (size_t)j * CH;
    auto at = [&](int gi, int gj) -> double2 {  // band entry (gi, gj), zero outside
        if (gi < 0 || gj < 0 || gi >= n || gj >= n || abs(gi - gj) > band) return make_double2(0, 0);
        return A[(size_t)gi * w + (gj - gi + band)];
    };
    // column j of L_cc^{-1}: forward substitution (rows beyond n act as identity)
    for (int i = 0; i < CH; i++) {
        double2 v = make_double2(i == j ? 1.0 : 0.0, 0);
        if (i > j && base + i < n) {
            for (int m2 = max(j, i - band); m2 < i; m2++) v = csub(v, cmul(at(base + i, base + m2), Li[m2]));
        }
        Li[i] = (i < j) ? make_double2(0, 0) : v;
    }
    // column j of U_cc^{-1}: back substitution
    for (int i = CH - 1; i >= 0; i--) {
        double2 v = make_double2(0, 0);
        if (i <= j) {
            v = make_double2(i == j ? 1.0 : 0.0, 0);
            if (base + i < n) {
                for (int m2 = i + 1; m2 <= min(j, i + band); m2++) v = csub(v, cmul(at(base + i, base + m2), Ui[m2]));
                v = cdiv(v, at(base + i, base + i));
            }
        }
        Ui[i] = v;
    }
    // band couplings [t][row], CH rows of t with zeros for t >= band: lane j is a row of the chunk
    const int gi = base + j;
    double2 *Lo = Loff + ((size_t)q * nch + ch) * CH * CH;
    double2 *Uo = Uoff + ((size_t)q * nch + ch) * CH * CH;
    for (int t = 0; t < CH; t++) {
        const bool in = gi < n && t < band;
        Lo[(size_t)t * CH + j] = in ? at(gi, base - band + t) : make_double2(0, 0);
        Uo[(size_t)t * CH + j] = in ? at(gi, base + CH + t) : make_double2(0, 0);
    }
}

// radial index of interleaved position gi (the two halves of the radial basis alternate)
static __device__ __forceinline__ int natural_index(int gi, int n, int ilv) {
    return ilv ? ((gi & 1) ? (n >> 1) + (gi >> 1) : (gi >> 1)) : gi;
}

// sum of a complex value over the four lanes of a quad (DPP quad permutes), in every lane
__device__ __forceinline__ double dpp_quad_swap(double v, int ctrl_is_xor2) {
    const int lo = __double2loint(v), hi = __double2hiint(v);
    int l2, h2;
    if (ctrl_is_xor2) {
        l2 = __builtin_amdgcn_update_dpp(0, lo, 0x4E, 0xF, 0xF, false);  // quad_perm [2,3,0,1]
        h2 = __builtin_amdgcn_update_dpp(0, hi, 0x4E, 0xF, 0xF, false);
    } else {
        l2 = __builtin_amdgcn_update_dpp(0, lo, 0xB1, 0xF, 0xF, false);  // quad_perm [1,0,3,2]
        h2 = __builtin_amdgcn_update_dpp(0, hi, 0xB1, 0xF, 0xF, false);
    }
    return __hiloint2double(h2, l2);
}
__device__ __forceinline__ float dpp_quad_swap(float v, int ctrl_is_xor2) {
    const int i = __float_as_int(v);
    const int o = ctrl_is_xor2 ? __builtin_amdgcn_update_dpp(0, i, 0x4E, 0xF, 0xF, false)
                               : __builtin_amdgcn_update_dpp(0, i, 0xB1, 0xF, 0xF, false);
    return __int_as_float(o);
}
template <class T>
__device__ __forceinline__ T quad_sum(T v) {
    v.x += dpp_quad_swap(v.x, 0);
    v.y += dpp_quad_swap(v.y, 0);
    v.x += dpp_quad_swap(v.x, 1);
    v.y += dpp_quad_swap(v.y, 1);
    return v;
}
// sum over the L (2 or 4) neighbouring lanes that share a row
template <int L, class T>
__device__ __forceinline__ T lanes_sum(T v) {
    if constexpr (L == 4) return quad_sum(v);
    v.x += dpp_quad_swap(v.x, 0);
    v.y += dpp_quad_swap(v.y, 0);
    return v;
}

// x = P^{-1} b with CH-row levels (CH = 16 or 32) and one block per workgroup of LPR*CH lanes: lane = LPR*row
// + k, and each of the LPR lanes of a row takes 1/LPR of the band coupling (terms (16/LPR)k..) and of the level
// inverse (terms (CH/LPR)k ..), summed over the lanes of the row with DPP.  LPR 4: 16-row levels in one wave
// or 32-row levels in two; LPR 2: 32-row levels in one wave (half the levels of 16-row ones, twice the terms
// per lane), the form launched for 32-row levels (16-row ones use the twisted form, k_psolve_tw).  b is
// staged in LDS; the factors of level ch+PF are loaded as soon as the register slot of level ch is free, so
// PF levels of loads are in flight.  Requires band <= 16 (coupling rows padded to CH, zero beyond band).
template <int PF, class T = double2, int CH = 16, int LPR = 4>
__global__ __launch_bounds__(LPR * CH) void k_psolve_q(const T *__restrict__ b, T *__restrict__ xo,
                                                     const int *__restrict__ order, const int *__restrict__ blk_q,
                                                     int ilv, const T *__restrict__ Linv, const T *__restrict__ Uinv,
                                                     const T *__restrict__ Loff, const T *__restrict__ Uoff, int n,
                                                     int band, int nch, const int *gate = nullptr, int gstride = 0,
                                                     int nblk1 = 1) {
    constexpr int NT = LPR * CH, QI = CH / LPR, CT = 16 / LPR;  // lanes; inverse and coupling terms per lane
    extern __shared__ __align__(16) char smem_raw[];
    T *smem = reinterpret_cast<T *>(smem_raw);
    const int lane = threadIdx.x, i = lane / LPR, k = lane % LPR;
    const int r = order[blockIdx.x];
    if (r < 0) return;  // padding slot (whole workgroup)
    GADE_GATED_B(gate, gstride, r / nblk1);
    const int q = blk_q[r];
    T *yv = smem;              // [nch*CH]: b, then y, then x (interleaved order)
    T *acc = smem + nch * CH;  // [CH]
    const T *bb = b + (size_t)r * n;
    const int ntot = nch * CH;
    for (int c0 = 0; c0 < ntot; c0 += NT * 8) {
        T t[8];
#pragma unroll
        for (int u = 0; u < 8; u++) {
            const int c = c0 + NT * u + lane;
            t[u] = (c < n) ? bb[natural_index(c, n, ilv)] : czero<T>();
        }
#pragma unroll
        for (int u = 0; u < 8; u++) {
            const int c = c0 + NT * u + lane;
            if (c < ntot) yv[c] = t[u];
        }
    }
    __syncthreads();
    const size_t cs = (size_t)CH * CH;
    T fc[PF][CT], fv[PF][QI];
    // factors of level ch into slot s: coupling rows CT*k.. and inverse columns QI*k.. of row i
    auto load = [&](const T *Off, const T *Inv, int ch, T (&c)[CT], T (&v)[QI]) {
        const T *po = Off + ((size_t)q * nch + ch) * cs + (size_t)(CT * k) * CH + i;
        const T *pi = Inv + ((size_t)q * nch + ch) * cs + (size_t)(QI * k) * CH + i;
#pragma unroll
        for (int u = 0; u < CT; u++) c[u] = po[(size_t)u * CH];
#pragma unroll
        for (int u = 0; u < QI; u++) v[u] = pi[(size_t)u * CH];
    };
    // one level: val = b_i - sum_t C[t][i] y[cbase + t];  y_i = sum_j Inv[j][i] val_j
    auto level = [&](int ch, int cbase, bool couple, const T (&c)[CT], const T (&v)[QI]) {
        const int gi = ch * CH + i;
        T s = czero<T>();
        if (couple) {
#pragma unroll
            for (int u = 0; u < CT; u++) s = cfma(c[u], yv[cbase + CT * k + u], s);
            s = lanes_sum<LPR>(s);
        }
        const T val = csub(yv[gi], s);
        if (k == 0) acc[i] = val;
        __syncthreads();
        T y = czero<T>();
#pragma unroll
        for (int u = 0; u < QI; u++) y = cfma(v[u], acc[QI * k + u], y);
        y = lanes_sum<LPR>(y);
        if (k == 0) yv[gi] = y;
        __syncthreads();
    };
    // forward sweep: L y = b
#pragma unroll
    for (int u = 0; u < PF; u++)
        if (u < nch) load(Loff, Linv, u, fc[u], fv[u]);
    for (int ch0 = 0; ch0 < nch; ch0 += PF) {
#pragma unroll
        for (int u = 0; u < PF; u++) {
            const int ch = ch0 + u;
            if (ch >= nch) break;
            level(ch, ch * CH - band, ch > 0, fc[u], fv[u]);
            if (ch + PF < nch) load(Loff, Linv, ch + PF, fc[u], fv[u]);
        }
    }
    // backward sweep: U x = y, level ch couples to level ch+1
#pragma unroll
    for (int u = 0; u < PF; u++)
        if (nch - 1 - u >= 0) load(Uoff, Uinv, nch - 1 - u, fc[u], fv[u]);
    for (int ch0 = nch - 1; ch0 >= 0; ch0 -= PF) {
#pragma unroll
        for (int u = 0; u < PF; u++) {
            const int ch = ch0 - u;
            if (ch < 0) break;
            level(ch, (ch + 1) * CH, ch < nch - 1, fc[u], fv[u]);
            if (ch - PF >= 0) load(Uoff, Uinv, ch - PF, fc[u], fv[u]);
        }
    }
    for (int c = lane; c < n; c += NT) xo[(size_t)r * n + natural_index(c, n, ilv)] = yv[c];
}

// LDS exchange inside one wave: a wave's LDS operations complete in issue order, so ordering the compiler's
// memory operations is all that is needed between a lane's write and another lane's read
__device__ __forceinline__ void wave_lds_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// x = P^{-1} b from the twisted factorization over 16-row levels (twisted_factors): one block per workgroup of
// two waves.  Wave 0 runs the top levels (g_c = b_c - X_c g_{c-1}, c = 1..m-1), wave 1 the bottom ones
// (h_c = b_c - X_c h_{c+1}, c = nch-2..m+1) at the same time; after a barrier wave 0 solves the meeting level
// x_m = Y_m (b_m - X_m g_{m-1} - W_m h_{m+1}); after another, both run their back sweeps outwards
// (x_c = Y_c g_c - Z_c x_{c+1} above m, x_c = Y_c h_c - Z_c x_{c-1} below).  So each wave's dependency chain is
// about half the levels of one sweep pair, with no workgroup barrier inside a sweep.  Per level, lane 4i + k of
// a wave takes terms 4k..4k+3 of row i, summed over the quad with DPP; b, g/h and x share one LDS array in
// place (a level reads its neighbour's values and then writes its own rows).  The factors of the level PF
// steps ahead load as soon as a register slot is free; the meeting level's factors load at the start.
template <int PF, class T = double2>
__global__ __launch_bounds__(128) void k_psolve_tw(const T *__restrict__ b, T *__restrict__ xo,
                                                   const int *__restrict__ order, const int *__restrict__ blk_q,
                                                   int ilv, const T *__restrict__ F, int n, int nch, int m,
                                                   const int *gate = nullptr, int gstride = 0, int nblk1 = 1) {
    constexpr int CH = 16, MS = CH * CH;
    extern __shared__ __align__(16) char smem_raw[];
    T *v = reinterpret_cast<T *>(smem_raw);  // [nch*16], interleaved order
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63, i = lane >> 2, k = lane & 3;
    const int r = order[blockIdx.x];
    if (r < 0) return;  // padding slot (whole workgroup)
    GADE_GATED_B(gate, gstride, r / nblk1);
    const int q = blk_q[r];
    const int ntot = nch * CH;
    // terms 4k..4k+3 of row i of matrix j of level c
    const T *Fq = F + (size_t)q * nch * 4 * MS + (size_t)(4 * k) * CH + i;
    auto ld4 = [&](int c, int j, T (&f)[4]) {
        const T *p = Fq + ((size_t)c * 4 + j) * MS;
#pragma unroll
        for (int u = 0; u < 4; u++) f[u] = p[(size_t)u * CH];
    };
    // this wave's levels: nf forward steps, then nb back steps
    const int nf = wave == 0 ? max(m - 1, 0) : max(nch - 2 - m, 0);
    const int nb = wave == 0 ? m : nch - 1 - m;
    const int dir = wave == 0 ? -1 : 1;  // neighbour of a forward level (the back sweep uses -dir)
    auto fwd_level = [&](int s) { return wave == 0 ? s + 1 : nch - 2 - s; };
    auto back_level = [&](int s) { return wave == 0 ? m - 1 - s : m + 1 + s; };
    T fm[3][4];  // meeting level: X_m, W_m, Y_m (wave 0)
    if (wave == 0) {
        ld4(m, 0, fm[0]);
        ld4(m, 3, fm[1]);
        ld4(m, 1, fm[2]);
    }
    T ff[PF][4];
#pragma unroll
    for (int u = 0; u < PF; u++)
        if (u < nf) ld4(fwd_level(u), 0, ff[u]);
    // b: eight loads per thread in flight before the LDS writes
    const T *bb = b + (size_t)r * n;
    for (int c0 = 0; c0 < ntot; c0 += 128 * 8) {
        T t[8];
#pragma unroll
        for (int u = 0; u < 8; u++) {
            const int c = c0 + 128 * u + (int)threadIdx.x;
            t[u] = c < n ? bb[natural_index(c, n, ilv)] : czero<T>();
        }
#pragma unroll
        for (int u = 0; u < 8; u++) {
            const int c = c0 + 128 * u + (int)threadIdx.x;
            if (c < ntot) v[c] = t[u];
        }
    }
    __syncthreads();
    // forward: v_c -= X_c v_{c+dir}
    for (int s0 = 0; s0 < nf; s0 += PF) {
#pragma unroll
        for (int u = 0; u < PF; u++) {
            const int s = s0 + u;
            if (s >= nf) break;
            const int c = fwd_level(s), nbr = (c + dir) * CH;
            T t = czero<T>();
#pragma unroll
            for (int e = 0; e < 4; e++) t = cfma(ff[u][e], v[nbr + 4 * k + e], t);
            t = quad_sum(t);
            if (k == 0) v[c * CH + i] = csub(v[c * CH + i], t);
            wave_lds_sync();
            if (s + PF < nf) ld4(fwd_level(s + PF), 0, ff[u]);
        }
    }
    T fb[PF][2][4];  // back steps: Y_c, Z_c
#pragma unroll
    for (int u = 0; u < PF; u++)
        if (u < nb) {
            ld4(back_level(u), 1, fb[u][0]);
            ld4(back_level(u), 2, fb[u][1]);
        }
    __syncthreads();
    if (wave == 0) {
        // meeting level: v_m = Y_m (v_m - X_m v_{m-1} - W_m v_{m+1})
        T t = czero<T>();
        if (m >= 1) {
#pragma unroll
            for (int e = 0; e < 4; e++) t = cfma(fm[0][e], v[(m - 1) * CH + 4 * k + e], t);
        }
        if (m + 1 < nch) {
#pragma unroll
            for (int e = 0; e < 4; e++) t = cfma(fm[1][e], v[(m + 1) * CH + 4 * k + e], t);
        }
        t = quad_sum(t);
        if (k == 0) v[m * CH + i] = csub(v[m * CH + i], t);
        wave_lds_sync();
        T y = czero<T>();
#pragma unroll
        for (int e = 0; e < 4; e++) y = cfma(fm[2][e], v[m * CH + 4 * k + e], y);
        y = quad_sum(y);
        if (k == 0) v[m * CH + i] = y;
    }
    __syncthreads();
    // back: v_c = Y_c v_c - Z_c v_{c-dir}
    for (int s0 = 0; s0 < nb; s0 += PF) {
#pragma unroll
        for (int u = 0; u < PF; u++) {
            const int s = s0 + u;
            if (s >= nb) break;
            const int c = back_level(s), nbr = (c - dir) * CH;
            T t = czero<T>();
#pragma unroll
            for (int e = 0; e < 4; e++) t = cfma(fb[u][0][e], v[c * CH + 4 * k + e], t);
#pragma unroll
            for (int e = 0; e < 4; e++) t = cfma(cx<T>::mk(-fb[u][1][e].x, -fb[u][1][e].y), v[nbr + 4 * k + e], t);
            t = quad_sum(t);
            if (k == 0) v[c * CH + i] = t;
            wave_lds_sync();
            if (s + PF < nb) {
                ld4(back_level(s + PF), 1, fb[u][0]);
                ld4(back_level(s + PF), 2, fb[u][1]);
            }
        }
    }
    __syncthreads();
    for (int c = threadIdx.x; c < n; c += 128) xo[(size_t)r * n + natural_index(c, n, ilv)] = v[c];
}

// x = P^{-1} b: one wave serves 64/CH blocks (preferably of one kappa), lane =
// (block slot, row of the chunk).  Levels are chunks of CH rows: each chunk
// takes the band coupling to the neighbouring chunk, then its diagonal-block
// inverse.  b is staged in LDS before the sweeps, and the factors of the next
// level are loaded one phase ahead (coupling factors during the inverse phase,
// inverse factors during the coupling phase) so their latency overlaps work.
template <int CH>
__global__ __launch_bounds__(64) void k_psolve(const double2 *__restrict__ b, double2 *__restrict__ xo,
                                               const int *__restrict__ wblocks, const int *__restrict__ blk_q,
                                               int ilv, const double2 *__restrict__ Linv,
                                               const double2 *__restrict__ Uinv, const double2 *__restrict__ Loff,
                                               const double2 *__restrict__ Uoff, int n, int band, int nch,
                                               const int *__restrict__ gate = nullptr) {
    GADE_GATED(gate);
    constexpr int G = 64 / CH;
    constexpr int NB = 16;  // b loads in flight while staging
    extern __shared__ double2 smem[];
    const int lane = threadIdx.x, g = lane / CH, i = lane % CH;
    const int r = wblocks[blockIdx.x * G + g];
    const int q = r >= 0 ? blk_q[r] : 0;
    double2 *yv = smem + (size_t)g * (nch * CH + CH);
    double2 *acc = yv + nch * CH;
    const size_t cs = (size_t)CH * CH;
    const double2 *bb = b + (size_t)(r >= 0 ? r : 0) * n;
    for (int c0 = 0; c0 < nch; c0 += NB) {
        double2 t[NB];
#pragma unroll
        for (int u = 0; u < NB; u++) {
            const int gi = (c0 + u) * CH + i;
            t[u] = (r >= 0 && c0 + u < nch && gi < n) ? bb[natural_index(gi, n, ilv)] : make_double2(0, 0);
        }
#pragma unroll
        for (int u = 0; u < NB; u++)
            if (c0 + u < nch) yv[(c0 + u) * CH + i] = t[u];
    }
    __syncthreads();
    if constexpr (CH <= 16) {
        double2 fc[CH], fi[CH];
        auto load_cpl = [&](const double2 *Off, int ch) {
            const double2 *p = Off + ((size_t)q * nch + ch) * cs + i;
#pragma unroll
            for (int t = 0; t < CH; t++) fc[t] = p[(size_t)t * CH];
        };
        auto load_inv = [&](const double2 *Inv, int ch) {
            const double2 *p = Inv + ((size_t)q * nch + ch) * cs + i;
#pragma unroll
            for (int j = 0; j < CH; j++) fi[j] = p[(size_t)j * CH];
        };
        // v - sum_t fc[t] * y[base + t] over all CH rows of the coupling (rows t >= band are zero and
        // meet finite staged values), two interleaved partial sums
        auto coupling = [&](double2 v, int base) {
            double2 s0 = make_double2(0, 0), s1 = make_double2(0, 0);
#pragma unroll
            for (int t = 0; t < CH; t += 2) {
                s0 = cfma(fc[t], yv[base + t], s0);
                s1 = cfma(fc[t + 1], yv[base + t + 1], s1);
            }
            return csub(v, cadd(s0, s1));
        };
        // sum_j fi[j] * acc[j], four interleaved partial sums
        auto inverse = [&]() {
            double2 s[4];
#pragma unroll
            for (int k = 0; k < 4; k++) s[k] = make_double2(0, 0);
#pragma unroll
            for (int j = 0; j < CH; j++) s[j % 4] = cfma(fi[j], acc[j], s[j % 4]);
            return cadd(cadd(s[0], s[1]), cadd(s[2], s[3]));
        };
        // forward sweep: L y = b.  Level ch uses fc = coupling of ch and fi = inverse of ch;
        // each is reloaded for ch+1 as soon as its phase of level ch is done.
        load_inv(Linv, 0);
        for (int ch = 0; ch < nch; ch++) {
            const int gi = ch * CH + i;
            double2 v = yv[gi];
            if (ch > 0) v = coupling(v, ch * CH - band);
            acc[i] = v;
            __syncthreads();
            if (ch + 1 < nch) load_cpl(Loff, ch + 1);
            const double2 yi = inverse();
            if (ch + 1 < nch) load_inv(Linv, ch + 1);
            yv[gi] = yi;
            __syncthreads();
        }
        // backward sweep (in place): U x = y; chunk ch couples to chunk ch+1
        load_inv(Uinv, nch - 1);
        for (int ch = nch - 1; ch >= 0; ch--) {
            const int gi = ch * CH + i;
            double2 v = yv[gi];
            if (ch < nch - 1) v = coupling(v, (ch + 1) * CH);
            acc[i] = v;
            __syncthreads();
            if (ch >= 1) load_cpl(Uoff, ch - 1);
            const double2 xi = inverse();
            if (ch >= 1) load_inv(Uinv, ch - 1);
            yv[gi] = xi;
            __syncthreads();
        }
    } else {
        // wide chunks: factors streamed from memory 16 at a time (no register prefetch)
        constexpr int B = 16;
        for (int ch = 0; ch < nch; ch++) {
            const int gi = ch * CH + i;
            double2 v = yv[gi];
            if (ch > 0) {
                const double2 *Lo = Loff + ((size_t)q * nch + ch) * cs + i;
                for (int t0 = 0; t0 < band; t0 += B) {
                    double2 f[B];
#pragma unroll
                    for (int u = 0; u < B; u++) f[u] = (t0 + u < band) ? Lo[(size_t)(t0 + u) * CH] : make_double2(0, 0);
#pragma unroll
                    for (int u = 0; u < B; u++)
                        if (t0 + u < band) v = csub(v, cmul(f[u], yv[ch * CH - band + t0 + u]));
                }
            }
            acc[i] = v;
            __syncthreads();
            const double2 *Li = Linv + ((size_t)q * nch + ch) * cs + i;
            double2 yi = make_double2(0, 0);
            for (int j0 = 0; j0 < CH; j0 += B) {
                double2 f[B];
#pragma unroll
                for (int u = 0; u < B; u++) f[u] = Li[(size_t)(j0 + u) * CH];
#pragma unroll
                for (int u = 0; u < B; u++) yi = cfma(f[u], acc[j0 + u], yi);
            }
            yv[gi] = yi;
            __syncthreads();
        }
        for (int ch = nch - 1; ch >= 0; ch--) {
            const int gi = ch * CH + i;
            double2 v = yv[gi];
            if (ch < nch - 1) {
                const double2 *Uo = Uoff + ((size_t)q * nch + ch) * cs + i;
                for (int t0 = 0; t0 < band; t0 += B) {
                    double2 f[B];
#pragma unroll
                    for (int u = 0; u < B; u++) f[u] = (t0 + u < band) ? Uo[(size_t)(t0 + u) * CH] : make_double2(0, 0);
#pragma unroll
                    for (int u = 0; u < B; u++)
                        if (t0 + u < band) v = csub(v, cmul(f[u], yv[(ch + 1) * CH + t0 + u]));
                }
            }
            acc[i] = v;
            __syncthreads();
            const double2 *Ui = Uinv + ((size_t)q * nch + ch) * cs + i;
            double2 xi = make_double2(0, 0);
            for (int j0 = 0; j0 < CH; j0 += B) {
                double2 f[B];
#pragma unroll
                for (int u = 0; u < B; u++) f[u] = Ui[(size_t)(j0 + u) * CH];
#pragma unroll
                for (int u = 0; u < B; u++) xi = cfma(f[u], acc[j0 + u], xi);
            }
            yv[gi] = xi;
            __syncthreads();
        }
    }
    if (r < 0) return;
    for (int gi = i; gi < n; gi += CH) xo[(size_t)r * n + natural_index(gi, n, ilv)] = yv[gi];
}

