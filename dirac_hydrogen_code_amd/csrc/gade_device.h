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
typedef double dvec2 __attribute__((ext_vector_type(2)));
// streaming (read-once) load of a complex value
__device__ __forceinline__ double2 ldnt(const double2 *p) {
    dvec2 v = __builtin_nontemporal_load((const dvec2 *)p);
    return make_double2(v.x, v.y);
}

// ---------------------------------------------------------------------------
// BiCGSTAB vector kernels with deterministic last-block reductions
// ---------------------------------------------------------------------------
// scalar slots in d_sc
enum {
    SC_RHO = 0,        // 2
    SC_ALPHA = 2,      // 2
    SC_W = 4,          // 2
    SC_R0SQ = 6,
    SC_RHSSQ = 7,
    SC_RNORM2 = 8,
    SC_RHONEXT = 9,    // 2
    SC_COUNT = 12
};
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

template <int OP>
__global__ __launch_bounds__(kRedThreads) void k_bicg(long long N, const double2 *__restrict__ b, double2 *__restrict__ r,
                                                       double2 *__restrict__ r0, double2 *__restrict__ p,
                                                       const double2 *__restrict__ v, double2 *__restrict__ s,
                                                       const double2 *__restrict__ t, const double2 *__restrict__ yv,
                                                       const double2 *__restrict__ zv, double2 *__restrict__ x,
                                                       const double *__restrict__ sc, double *__restrict__ part) {
    __shared__ double sh[4 * 16];
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
    if (OP == OP_S || OP == OP_XR) alpha = make_double2(sc[SC_ALPHA], sc[SC_ALPHA + 1]);
    if (OP == OP_XR) w = make_double2(sc[SC_W], sc[SC_W + 1]);
    for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < N; i += (long long)gridDim.x * blockDim.x) {
        if (OP == OP_INIT || OP == OP_RESTART) {
            double2 ri = csub(b[i], r[i]);
            r[i] = ri;
            r0[i] = ri;
            acc[0] = fma(ri.x, ri.x, fma(ri.y, ri.y, acc[0]));
            if (OP == OP_INIT) {
                double2 bi = b[i];
                acc[1] = fma(bi.x, bi.x, fma(bi.y, bi.y, acc[1]));
            }
        } else if (OP == OP_P) {
            double2 pi = csub(p[i], cmul(w, v[i]));
            p[i] = cadd(r[i], cmul(beta, pi));
        } else if (OP == OP_ALPHA) {
            double2 d = cfmaconj(r0[i], v[i], make_double2(acc[0], acc[1]));
            acc[0] = d.x;
            acc[1] = d.y;
        } else if (OP == OP_S) {
            s[i] = csub(r[i], cmul(alpha, v[i]));
        } else if (OP == OP_W) {
            double2 ti = t[i];
            acc[0] = fma(ti.x, ti.x, fma(ti.y, ti.y, acc[0]));
            double2 d = cfmaconj(ti, s[i], make_double2(acc[1], acc[2]));
            acc[1] = d.x;
            acc[2] = d.y;
        } else if (OP == OP_XR) {
            double2 xi = x[i];
            xi = cadd(xi, cadd(cmul(alpha, yv[i]), cmul(w, zv[i])));
            x[i] = xi;
            double2 ri = csub(s[i], cmul(w, t[i]));
            r[i] = ri;
            acc[0] = fma(ri.x, ri.x, fma(ri.y, ri.y, acc[0]));
            double2 d = cfmaconj(r0[i], ri, make_double2(acc[1], acc[2]));
            acc[1] = d.x;
            acc[2] = d.y;
        }
    }
    if (OP == OP_P || OP == OP_S) return;
    block_sum<4>(acc, sh);
    if (threadIdx.x == 0)
        for (int i = 0; i < 4; i++) part[(size_t)blockIdx.x * 4 + i] = acc[i];
}

// Sums the partials in a fixed order and updates the scalars (one block).
template <int OP>
__global__ __launch_bounds__(kRedThreads) void k_bicg_fin(int nparts, const double *__restrict__ part, double *sc) {
    __shared__ double sh[4 * 16];
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
}

// y = alpha A x + beta y for a CSR matrix: one wave per row, streaming loads.
static __global__ __launch_bounds__(256) void k_csr_spmv(const int *__restrict__ Ap, const int *__restrict__ Ai,
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

