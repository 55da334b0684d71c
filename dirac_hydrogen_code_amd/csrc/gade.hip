// libgade: MI355X-native Crank-Nicolson hot path (fused operator, block-LU
// preconditioner, device-resident BiCGSTAB).  See include/gade.h for the ABI.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <complex>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <chrono>
#include <cstring>
#include <string>
#include <vector>

#include "gade_internal.h"
#include "gade_device.h"

using gade::kSoL;

namespace gade {
static thread_local std::string g_err;
void set_error(const std::string &msg) { g_err = msg; }
}  // namespace gade

extern "C" const char *gade_last_error(void) { return gade::g_err.c_str(); }

// ---------------------------------------------------------------------------
// profiling helper
// ---------------------------------------------------------------------------
namespace {
struct Prof {
    gade_model *m;
    const char *name;
    bool on;
    hipEvent_t a{}, b{};
    static hipEvent_t take(gade_model *m) {
        hipEvent_t e{};
        if (!m->event_pool.empty()) {
            e = m->event_pool.back();
            m->event_pool.pop_back();
        } else if (hipEventCreate(&e) != hipSuccess) {
            e = nullptr;
        }
        return e;
    }
    Prof(gade_model *mm, const char *nm, int bit) : m(mm), name(nm), on((mm->profile & bit) != 0) {
        if (on && m->prof_period > 1) on = (m->prof_seen[nm]++ % m->prof_period) == 0;
        if (on) {
            a = take(m);
            b = take(m);
            on = a && b;
            if (on) (void)hipEventRecord(a, m->stream);
        }
    }
    ~Prof() {
        if (on) {
            (void)hipEventRecord(b, m->stream);
            m->pending.push_back({name, a, b});
        }
    }
};

void resolve_pending(gade_model *m) {
    for (auto &p : m->pending) {
        (void)hipEventSynchronize(p.b);
        float ms = 0;
        (void)hipEventElapsedTime(&ms, p.a, p.b);
        auto &t = m->timers[p.name];
        if ((double)ms * 1e6 >= (double)m->prof_floor_ns) {  // launches ended by a device gate are dropped
            t.total_ms += ms;
            t.count += 1;
        }
        m->event_pool.push_back(p.a);
        m->event_pool.push_back(p.b);
    }
    m->pending.clear();
}
}  // namespace

__global__ __launch_bounds__(256) void k_copy(const double2 *__restrict__ src, double2 *__restrict__ dst, long long n) {
    for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x)
        dst[i] = src[i];
}

// ---------------------------------------------------------------------------
// kernels: per-step G-sums and stationary diagonal operators
// ---------------------------------------------------------------------------
// the pulse coefficients of up to kCoefGroup instances, passed by value to one launch
constexpr int kCoefGroup = 32;
struct CoefGroup {
    Coef6 c[kCoefGroup];
};

// GE[l][m][s][a]: m<4: Gs_{l,m} = sum_alpha c_alpha g^m_{alpha,l};
// m>=4: Gts_{l,m-4} with gt = conj-transpose of g (tools/solver.c:236-287).
__global__ void k_gsums(const double2 *__restrict__ g, const int *__restrict__ ell_src, const int *__restrict__ tmap,
                        Coef6 cf, int lmax, int ellw, int n, int nnzb, double2 *__restrict__ GE,
                  const int *__restrict__ gate) {
    GADE_GATED(gate);
    long long total = (long long)lmax * 8 * ellw * n;
    for (long long idx = blockIdx.x * (long long)blockDim.x + threadIdx.x; idx < total;
         idx += (long long)gridDim.x * blockDim.x) {
        int sa = (int)(idx % ((long long)ellw * n));
        int lm = (int)(idx / ((long long)ellw * n));
        int m = lm & 7, l = lm >> 3;
        int src = ell_src[sa];
        double2 acc = make_double2(0, 0);
        if (src >= 0) {
            int nn = m & 3;
            bool tr = m >= 4;
            int k = tr ? tmap[src] : src;
            for (int a = (l % 2 == 0) ? 1 : 0; a < 6; a += 2) {
                double2 gv = g[(size_t)(a * 4 * lmax + l * 4 + nn) * nnzb + k];
                if (tr) gv.y = -gv.y;
                acc = cadd(acc, cmul(cf.c[a], gv));
            }
        }
        GE[idx] = acc;
    }
}

// D[sel][q][s][a]: sel 0: S(k)+ihdt H0(k), 1: S(k)-ihdt H0(k), 2: S(k)  (tools/solver.c:409-458)
__global__ void k_build_D(const double2 *__restrict__ s, const double2 *__restrict__ h0, const int *__restrict__ ell_src,
                          const int *__restrict__ kq, int nq, int ellw, int n, int nnzb, double hdt2,
                          double2 *__restrict__ D) {
    long long total = (long long)nq * ellw * n;
    for (long long idx = blockIdx.x * (long long)blockDim.x + threadIdx.x; idx < total;
         idx += (long long)gridDim.x * blockDim.x) {
        int sa = (int)(idx % ((long long)ellw * n));
        int q = (int)(idx / ((long long)ellw * n));
        int src = ell_src[sa];
        double2 S = make_double2(0, 0), H = make_double2(0, 0);
        if (src >= 0) {
            double k = kq[q];
            S = cadd(cadd(s[src], cscale(s[nnzb + src], k)), cscale(s[2 * nnzb + src], k * k));
            H = cadd(cadd(cadd(h0[src], cscale(h0[nnzb + src], k)), cscale(h0[2 * nnzb + src], k * k)),
                     cscale(h0[3 * nnzb + src], k * k * k));
        }
        // i*hdt2*H
        double2 iH = make_double2(-hdt2 * H.y, hdt2 * H.x);
        D[((long long)0 * nq * ellw * n) + idx] = cadd(S, iH);
        D[((long long)1 * nq * ellw * n) + idx] = csub(S, iH);
        D[((long long)2 * nq * ellw * n) + idx] = S;
    }
}

// ---------------------------------------------------------------------------
// fused operator, phase A: angular contraction per row block
//   Z0_l(r) = sum_c H0l[r,c] x_c      Z2_l(r) = sum_c H0l[r,c] k_c x_c
//   W0_l(r) = sum_c H1l[r,c] x_c      W1_l(r) = sum_c H1l[r,c] k_c x_c
// V layout [l][v][r][a].  blockIdx.z selects a group of LB multipoles so the
// accumulators stay small; the row's coefficients are staged in LDS and the
// x loads of four block columns are in flight together.
// ---------------------------------------------------------------------------
// XCD-aware workgroup order: workgroup b of a 1-D grid of G runs on XCD b % 8 (round-robin dispatch over the 8
// XCDs, each with its own L2); the logical index returned gives XCD j the contiguous range [j G/8, (j+1) G/8)
// (remainder spread over the first XCDs), so an XCD works through a few neighbouring slices of the work, and
// their data, at a time.
__device__ __forceinline__ int xcd_order(int b, int G) {
    const int j = b & 7, k = b >> 3, q = G >> 3, r = G & 7;
    return j * q + min(j, r) + k;
}

constexpr int kAngThreads = 256;
constexpr int kAngLB = 2;  // default multipoles per block

template <int LB, class T = double2>
__global__ __launch_bounds__(kAngThreads) void k_angular(const T *__restrict__ x, const int *__restrict__ hAp,
                                                         const int *__restrict__ hAi,
                                                         const T *__restrict__ coef,
                                                         const int *__restrict__ kap, int nblk, int n, int lmax,
                                                         int maxrow, T *__restrict__ V,
                                                         const int *gate, int gstride, int nblk1) {
    const int r = blockIdx.x;
    GADE_GATED_B(gate, gstride, r / nblk1);
    using Rt = typename cx<T>::R;
    extern __shared__ __align__(16) char ang_raw[];
    T *shc = reinterpret_cast<T *>(ang_raw);  // [maxrow][LB][2] coefficients, then [maxrow] (col, kappa) packed
    const int a = blockIdx.y * blockDim.x + threadIdx.x;
    const int l0 = blockIdx.z * LB;
    const int nlb = min(LB, lmax - l0);
    const int pb = hAp[r], pe = hAp[r + 1];
    const int len = pe - pb;
    int2 *shk = (int2 *)(shc + (size_t)maxrow * LB * 2);
    for (int e = threadIdx.x; e < len * LB * 2; e += blockDim.x) {
        const int p = e / (LB * 2), rem = e % (LB * 2), l = rem / 2, w = rem % 2;
        shc[e] = (l < nlb) ? coef[((size_t)(pb + p) * lmax + l0 + l) * 2 + w] : czero<T>();
    }
    for (int e = threadIdx.x; e < len; e += blockDim.x) {
        const int c = hAi[pb + e];
        shk[e] = make_int2(c, kap[c]);
    }
    __syncthreads();
    T z0[LB], z2[LB], w0[LB], w1[LB];
#pragma unroll
    for (int l = 0; l < LB; l++) z0[l] = z2[l] = w0[l] = w1[l] = czero<T>();
    const int ac = a < n ? a : n - 1;
    constexpr int U = 4;
    for (int p0 = 0; p0 < len; p0 += U) {
        T xv[U];
#pragma unroll
        for (int u = 0; u < U; u++) {
            const int p = min(p0 + u, len - 1);
            xv[u] = x[(size_t)shk[p].x * n + ac];
        }
#pragma unroll
        for (int u = 0; u < U; u++) {
            const int p = p0 + u;
            if (p >= len) break;
            const T kx = cscale(xv[u], (Rt)shk[p].y);
            const T *cp = shc + (size_t)p * LB * 2;
#pragma unroll
            for (int l = 0; l < LB; l++) {
                const T h0 = cp[2 * l], h1 = cp[2 * l + 1];
                if (cnz(h0)) {
                    z0[l] = cfma(h0, xv[u], z0[l]);
                    z2[l] = cfma(h0, kx, z2[l]);
                }
                if (cnz(h1)) {
                    w0[l] = cfma(h1, xv[u], w0[l]);
                    w1[l] = cfma(h1, kx, w1[l]);
                }
            }
        }
    }
    if (a >= n) return;
    const size_t plane = (size_t)nblk * n;
#pragma unroll
    for (int l = 0; l < LB; l++) {
        if (l >= nlb) break;
        size_t base = (size_t)((l0 + l) * 4) * plane + (size_t)r * n + a;
        V[base] = z0[l];
        V[base + plane] = z2[l];
        V[base + 2 * plane] = w0[l];
        V[base + 3 * plane] = w1[l];
    }
}

// Phase A from the coupling lists: per (row, multipole) only the columns with a non-zero H0l or H1l
// coefficient, in increasing column order, each with both coefficients (h0 for Z0, Z2; h1 for W0, W1; either
// may be 0), padded to a multiple of kAngListU entries with zero coefficients (on the list's last column).  The
// sums start at +0 and round to nearest, so they never hold -0, and a zero-coefficient term (+0 or -0 added)
// leaves them unchanged: every V entry is the same sum of the same non-zero terms in the same order as in
// k_angular, bit for bit, without testing coefficients.  Each x value is loaded once for the four sums.  The
// list entries are read at addresses uniform over the block.  A block's lists (its multipole group) are one
// contiguous stream; the next group's x loads are issued before the current group is summed.  Lists are per
// instance-local row; x is offset to the instance.
constexpr int kAngListU = 4;
template <class T>
__device__ __forceinline__ void ang_list_group(const T (&xv)[kAngListU], const int2 *__restrict__ ck,
                                               const T *__restrict__ cf, T &z0, T &z2, T &w0, T &w1) {
    using Rt = typename cx<T>::R;
#pragma unroll
    for (int u = 0; u < kAngListU; u++) {
        const T h0 = cf[2 * u], h1 = cf[2 * u + 1];
        const T kx = cscale(xv[u], (Rt)ck[u].y);
        z0 = cfma(h0, xv[u], z0);
        z2 = cfma(h0, kx, z2);
        w0 = cfma(h1, xv[u], w0);
        w1 = cfma(h1, kx, w1);
    }
}

template <int LB, class T = double2>
__global__ __launch_bounds__(kAngThreads) void k_angular_list(const T *__restrict__ x, const int *__restrict__ aoff,
                                                              const int2 *__restrict__ ack,
                                                              const T *__restrict__ acf, int n, int lmax,
                                                              int nblk, T *__restrict__ V, const int *gate,
                                                              int gstride, int nblk1) {
    // rows are the fastest block index: the blocks in flight together write different rows of the V planes
    // (spread over the memory channels; one row's blocks together would hit the same offset of 16 planes)
    const int r = blockIdx.x;
    const int inst = r / nblk1;
    GADE_GATED_B(gate, gstride, inst);
    const int rl = r - inst * nblk1;
    const int a = blockIdx.y * blockDim.x + threadIdx.x;
    const int l0 = blockIdx.z * LB;
    const int nlb = min(LB, lmax - l0);
    const int *off = aoff + (size_t)rl * lmax + l0;
    const T *xi = x + (size_t)inst * nblk1 * n;
    const int ac = a < n ? a : n - 1;
    T z0[LB], z2[LB], w0[LB], w1[LB];
#pragma unroll
    for (int l = 0; l < LB; l++) z0[l] = z2[l] = w0[l] = w1[l] = czero<T>();
    const int E0 = off[0], E1 = off[nlb];
    if (E1 > E0) {
        T cur[kAngListU];
#pragma unroll
        for (int u = 0; u < kAngListU; u++) cur[u] = xi[(size_t)ack[E0 + u].x * n + ac];
        int g = E0;
#pragma unroll
        for (int li = 0; li < LB; li++) {
            if (li >= nlb) break;
            const int e = off[li + 1];
            // two groups per pass so the two register sets alternate (no copies that wait on the loads)
            for (; g < e; g += 2 * kAngListU) {
                T nxt[kAngListU];
#pragma unroll
                for (int u = 0; u < kAngListU; u++)
                    nxt[u] = xi[(size_t)ack[min(g + kAngListU + u, E1 - 1)].x * n + ac];
                ang_list_group(cur, ack + g, acf + 2 * g, z0[li], z2[li], w0[li], w1[li]);
                if (g + kAngListU >= e) {  // odd number of groups: the loaded group opens the next list
#pragma unroll
                    for (int u = 0; u < kAngListU; u++) cur[u] = nxt[u];
                    g += kAngListU;
                    break;
                }
#pragma unroll
                for (int u = 0; u < kAngListU; u++)
                    cur[u] = xi[(size_t)ack[min(g + 2 * kAngListU + u, E1 - 1)].x * n + ac];
                ang_list_group(nxt, ack + g + kAngListU, acf + 2 * (g + kAngListU), z0[li], z2[li], w0[li],
                               w1[li]);
            }
        }
    }
    if (a >= n) return;
    const size_t plane = (size_t)nblk * n;
#pragma unroll
    for (int l = 0; l < LB; l++) {
        if (l >= nlb) break;
        size_t base = (size_t)((l0 + l) * 4) * plane + (size_t)r * n + a;
        V[base] = z0[l];
        V[base + plane] = z2[l];
        V[base + 2 * plane] = w0[l];
        V[base + 3 * plane] = w1[l];
    }
}

// ---------------------------------------------------------------------------
// fused operator, phase B: radial SpMM on a kappa-grouped tile of row blocks
//   y_r = D_sel(k) x_r + cH * sum_l [ (Gs0+k Gs1) Z0 + (Gs2+k Gs3) Z2 + (Gts0+k Gts2) W0 + (Gts1+k Gts3) W1 ]
//         + beta y_r
// One wave per (tile, 64-row radial chunk).  The chunk's column window of each
// V vector is staged in LDS once and read by every slot; the slot loads are
// issued eight at a time.  The chunk index is the fastest-varying part of
// blockIdx so that with 8 chunks each XCD keeps one chunk's slice of the
// G-sums resident in its L2.
// ---------------------------------------------------------------------------
constexpr int kMaxSlots = 32;   // ELL width supported by the fused radial kernel
constexpr int kMaxLocW = 4;     // LDS window <= 64 * kMaxLocW columns per chunk

template <int R>
__global__ __launch_bounds__(64) void k_radial(const double2 *__restrict__ x, const double2 *__restrict__ V,
                                               const double2 *__restrict__ GE, const double2 *__restrict__ D,
                                               const int *__restrict__ ell_lcol, const int *__restrict__ locj,
                                               const int *__restrict__ nloc, const int *__restrict__ tiles,
                                               const int *__restrict__ tile_q, const double *__restrict__ tile_k,
                                               int nblk, int n, int ellw, int lgrid, int nq, int nchunk, int maxloc,
                                               int sel, double cH, double beta, double2 *__restrict__ y,
                                               double2 *__restrict__ ypart,
                  const int *__restrict__ gate) {
    GADE_GATED(gate);
    extern __shared__ double2 sm[];  // [4][R][maxloc]
    const int chunk = blockIdx.x % nchunk;
    const int rest = blockIdx.x / nchunk;
    const int l = rest % lgrid, tile = rest / lgrid;
    const int lane = threadIdx.x;
    const int a = chunk * 64 + lane;
    const bool valid = a < n;
    const int ac = valid ? a : n - 1;
    int rows[R];
    bool live[R];
#pragma unroll
    for (int i = 0; i < R; i++) {
        const int rr = tiles[tile * R + i];
        live[i] = rr >= 0;
        rows[i] = rr >= 0 ? rr : tiles[tile * R];  // padded rows re-read row 0 and are never stored
    }
    const double k = tile_k[tile];
    const int q = tile_q[tile];
    const int nl = nloc[chunk];
    int ljr[kMaxLocW];
#pragma unroll
    for (int t = 0; t < kMaxLocW; t++) {
        const int idx = lane + 64 * t;
        ljr[t] = idx < nl ? locj[(size_t)chunk * maxloc + idx] : -1;
    }
    const int *lc = ell_lcol + ac;
    const size_t plane = (size_t)nblk * n;
    const size_t slab = (size_t)ellw * n;
    double2 acc[R];
#pragma unroll
    for (int i = 0; i < R; i++) acc[i] = make_double2(0, 0);

    if (cH != 0.0) {
        // stage the chunk windows of Z0, Z2, W0, W1 of this multipole for every row of the tile
#pragma unroll
        for (int v = 0; v < 4; v++) {
            double2 tmp[R][kMaxLocW];
#pragma unroll
            for (int i = 0; i < R; i++)
#pragma unroll
                for (int t = 0; t < kMaxLocW; t++)
                    if (ljr[t] >= 0) tmp[i][t] = V[(size_t)(l * 4 + v) * plane + (size_t)rows[i] * n + ljr[t]];
#pragma unroll
            for (int i = 0; i < R; i++)
#pragma unroll
                for (int t = 0; t < kMaxLocW; t++)
                    if (ljr[t] >= 0) sm[(v * R + i) * maxloc + lane + 64 * t] = tmp[i][t];
        }
        __syncthreads();
        // flattened (v, slot batch) loop; the next batch's loads are in flight while this one computes
        constexpr int SB = 4;
        const int nbv = (ellw + SB - 1) / SB;
        const int nb = 4 * nbv;
        int lcN[SB];
        double2 g1N[SB], g2N[SB];
        auto fetch = [&](int j) {
            const int v = j / nbv, s0 = (j % nbv) * SB;
            // (Gs0,Gs1) (Gs2,Gs3) (Gts0,Gts2) (Gts1,Gts3) pair with Z0, Z2, W0, W1
            const int m1 = (v < 2) ? 2 * v : 4 + (v - 2);
            const int m2 = (v < 2) ? 2 * v + 1 : 6 + (v - 2);
            const double2 *G1 = GE + (size_t)(l * 8 + m1) * slab + ac;
            const double2 *G2 = GE + (size_t)(l * 8 + m2) * slab + ac;
#pragma unroll
            for (int u = 0; u < SB; u++) {
                const int s2 = s0 + u;
                const bool ok = s2 < ellw;
                lcN[u] = ok ? lc[(size_t)s2 * n] : 0;
                g1N[u] = ok ? G1[(size_t)s2 * n] : make_double2(0, 0);
                g2N[u] = ok ? G2[(size_t)s2 * n] : make_double2(0, 0);
            }
        };
        fetch(0);
#pragma unroll 1
        for (int j = 0; j < nb; j++) {
            int lcC[SB];
            double2 g1C[SB], g2C[SB];
#pragma unroll
            for (int u = 0; u < SB; u++) {
                lcC[u] = lcN[u];
                g1C[u] = g1N[u];
                g2C[u] = g2N[u];
            }
            if (j + 1 < nb) fetch(j + 1);
            const double2 *smv = sm + (size_t)(j / nbv) * R * maxloc;
#pragma unroll
            for (int u = 0; u < SB; u++) {
                const double2 M = make_double2(fma(k, g2C[u].x, g1C[u].x), fma(k, g2C[u].y, g1C[u].y));
#pragma unroll
                for (int i = 0; i < R; i++) acc[i] = cfma(M, smv[i * maxloc + lcC[u]], acc[i]);
            }
        }
    }
    double2 dacc[R];
#pragma unroll
    for (int i = 0; i < R; i++) dacc[i] = make_double2(0, 0);
    if (l == 0) {
        // stationary diagonal operator on x_r (carried by the first multipole slice)
        __syncthreads();
        {
            double2 tmp[R][kMaxLocW];
#pragma unroll
            for (int i = 0; i < R; i++)
#pragma unroll
                for (int t = 0; t < kMaxLocW; t++)
                    if (ljr[t] >= 0) tmp[i][t] = x[(size_t)rows[i] * n + ljr[t]];
#pragma unroll
            for (int i = 0; i < R; i++)
#pragma unroll
                for (int t = 0; t < kMaxLocW; t++)
                    if (ljr[t] >= 0) sm[i * maxloc + lane + 64 * t] = tmp[i][t];
        }
        __syncthreads();
        const double2 *Dq = D + ((size_t)sel * nq + q) * slab + ac;
#pragma unroll 2
        for (int s2 = 0; s2 < ellw; s2++) {
            const int lcol = lc[(size_t)s2 * n];
            const double2 d = Dq[(size_t)s2 * n];
#pragma unroll
            for (int i = 0; i < R; i++) dacc[i] = cfma(d, sm[i * maxloc + lcol], dacc[i]);
        }
    }
    if (!valid) return;
#pragma unroll
    for (int i = 0; i < R; i++) {
        if (!live[i]) continue;
        const size_t o = (size_t)rows[i] * n + a;
        double2 res = cadd(dacc[i], cscale(acc[i], cH));
        if (ypart) {
            ypart[(size_t)l * plane + o] = res;  // summed over l by k_radial_sum
        } else {
            if (beta != 0.0) res = cadd(res, cscale(y[o], beta));
            y[o] = res;
        }
    }
}

// y = sum_l ypart[l] + beta y  (fixed summation order); blockIdx.y = instance of N1 entries
template <class T = double2>
__global__ __launch_bounds__(256) void k_radial_sum(const T *__restrict__ ypart, int lgrid, long long N,
                                                    double beta, T *__restrict__ y, const int *gate,
                                                    int gstride, long long N1) {
    GADE_GATED_B(gate, gstride, blockIdx.y);
    using Rt = typename cx<T>::R;
    for (long long j = blockIdx.x * (long long)blockDim.x + threadIdx.x; j < N1; j += (long long)gridDim.x * blockDim.x) {
        const long long i = (long long)blockIdx.y * N1 + j;
        T s = ypart[i];
        for (int l = 1; l < lgrid; l++) s = cadd(s, ypart[(size_t)l * N + i]);
        if (beta != 0.0) s = cadd(s, cscale(y[i], (Rt)beta));
        y[i] = s;
    }
}

// ---------------------------------------------------------------------------
// banded (DIA) form of the fused radial phase
//
// The radial pattern of the DKB basis is a 2x2 block of banded matrices: with
// nh = n/2, row h*nh+i couples to columns h'*nh+i+d-BW (d = 0..2BW) of both
// halves h'.  Coefficients are stored per quadrant qd = 2h+h' and diagonal d
// as contiguous runs over i (padded to nhp, zeros outside the band), so a
// radial product is a sum over d of coefficient[d][i] * window[i+d]: no column
// indices, and the window of a wave is one contiguous run staged in LDS.
//
// One wave owns R row blocks of one kappa (kDiaR by default), one output half h,
// 64*P consecutive outputs (P per lane, so each staged window value is read
// once per lane and reused across the diagonals) and one multipole l; the grid
// is ordered so that each XCD works through a few (segment, multipole) slices
// at a time.  The per-multipole results go to ypart and are summed in a fixed
// order.
// ---------------------------------------------------------------------------
constexpr int kDiaR = 8;        // default rows per tile
constexpr int kDiaSegMax = 256; // longest output run of a wave (64 lanes x 4 outputs); nhp is a multiple of it
constexpr int kDiaBW = 7;

// LDS slot of window position p: swaps neighbours on alternate 8-slot rows so
// that lanes reading positions 2*lane+t hit distinct banks
// (P = 2); for P = 4 (lanes reading positions 4*lane+t) the low two bits are XORed instead.  Both are
// conflict-free for the sliding reads and for the stride-1 staging writes (64 banks, 16 lanes per pass).
// With one output per lane (P = 1) the lanes read consecutive positions: no swizzle.
template <int P>
__device__ __forceinline__ int dia_swz(int p) {
    if constexpr (P == 1) return p;
    if constexpr (P == 4) return p ^ (((p >> 2) ^ (p >> 4)) & 3);
    return p ^ (((p >> 3) ^ (p >> 4)) & 1);
}

// GD[l][v][qd][d][ip] = (G_m1, G_m2) of the kappa-folded pair of v, from the
// DIA-ordered couplings gs[l][j][nn][qd][d][ip] (g of alpha = 2j + (l even))
// and gts (conjugated transposes).  One thread per (l, v, qd, d, ip): the pair of sums that multiplies V
// component v (six loads each).
// Coefficients: by value (cg, instances b0 + blockIdx.y), or from the device window table
// cf_dev[slot][B] at slot gctl[GC_WSLOT] (the form a replayed step graph uses).
__global__ __launch_bounds__(256) void k_gsums_dia(const double2 *__restrict__ gs, const double2 *__restrict__ gts,
                                                   CoefGroup cg, int b0, size_t gd_stride, int lmax, int nd, int nhp,
                                                   double2 *__restrict__ GD0, const int *gate, int gstride,
                                                   const Coef6 *__restrict__ cf_dev, const int *__restrict__ gctl,
                                                   int B) {
    const int b = b0 + blockIdx.y;
    GADE_GATED_B(gate, gstride, b);
    const Coef6 cf = cf_dev ? cf_dev[(size_t)gctl[GC_WSLOT] * B + b] : cg.c[blockIdx.y];
    // coefficients of the alpha of each multipole parity, held with static indices (no private-memory copy)
    double2 c_even[3], c_odd[3];
#pragma unroll
    for (int j = 0; j < 3; j++) {
        c_even[j] = cf.c[2 * j + 1];
        c_odd[j] = cf.c[2 * j];
    }
    double2 *__restrict__ GD = GD0 + (size_t)b * gd_stride;
    const int S = 4 * nd * nhp;  // entries per (l, j, nn) plane
    const long long total = (long long)lmax * 4 * S;
    for (long long idx = blockIdx.x * (long long)blockDim.x + threadIdx.x; idx < total;
         idx += (long long)gridDim.x * blockDim.x) {
        const int lv = (int)(idx / S);
        const int e = (int)(idx - (long long)lv * S);  // (qd, d, ip)
        const int l = lv >> 2, v = lv & 3;
        // (Gs0,Gs1) (Gs2,Gs3) (Gts0,Gts2) (Gts1,Gts3) pair with Z0, Z2, W0, W1
        const double2 *src = v < 2 ? gs : gts;
        const int n1 = v == 0 ? 0 : v == 1 ? 2 : v == 2 ? 0 : 1, n2 = v == 0 ? 1 : v == 1 ? 3 : v == 2 ? 2 : 3;
        double2 G1 = make_double2(0, 0), G2 = make_double2(0, 0);
#pragma unroll
        for (int j = 0; j < 3; j++) {
            const double2 c = (l % 2 == 0) ? c_even[j] : c_odd[j];
            const size_t o = (size_t)(l * 3 + j) * 4 * S + e;
            G1 = cfma(c, src[o + (size_t)n1 * S], G1);
            G2 = cfma(c, src[o + (size_t)n2 * S], G2);
        }
        const size_t o = ((size_t)lv * S + e) * 2;
        GD[o] = G1;
        GD[o + 1] = G2;
    }
}

// gs/gts gathered into DIA order once at setup; dia_src[(qd*nd+d)*nhp+ip] is
// the radial CSR index of that entry or -1.
__global__ void k_build_gdia(const double2 *__restrict__ g, const int *__restrict__ dia_src,
                             const int *__restrict__ tmap, int lmax, int nd, int nhp, int nnzb,
                             double2 *__restrict__ gs, double2 *__restrict__ gts) {
    const int S = 4 * nd * nhp;
    const long long total = (long long)lmax * 12 * S;
    for (long long idx = blockIdx.x * (long long)blockDim.x + threadIdx.x; idx < total;
         idx += (long long)gridDim.x * blockDim.x) {
        const int e = (int)(idx % S);
        const int ljn = (int)(idx / S);  // (l*3 + j)*4 + nn
        const int nn = ljn % 4, j = (ljn / 4) % 3, l = ljn / 12;
        const int alpha = 2 * j + ((l % 2 == 0) ? 1 : 0);
        const int k = dia_src[e];
        double2 a = make_double2(0, 0), b = make_double2(0, 0);
        if (k >= 0) {
            const size_t plane = (size_t)(alpha * 4 * lmax + l * 4 + nn) * nnzb;
            a = g[plane + k];
            b = g[plane + tmap[k]];
            b.y = -b.y;
        }
        gs[idx] = a;
        gts[idx] = b;
    }
}

// DD[sel][q][qd][d][ip] from the ELL-layout diagonal operators
__global__ void k_build_DD(const double2 *__restrict__ D, const int *__restrict__ dia_src,
                           const int *__restrict__ ell_pos, int nq, int ellw, int n, int nd, int nhp,
                           double2 *__restrict__ DD) {
    const int S = 4 * nd * nhp;
    const long long total = (long long)3 * nq * S;
    for (long long idx = blockIdx.x * (long long)blockDim.x + threadIdx.x; idx < total;
         idx += (long long)gridDim.x * blockDim.x) {
        const int e = (int)(idx % S);
        const long long sq = idx / S;  // sel*nq + q
        const int k = dia_src[e];
        DD[idx] = k >= 0 ? D[(size_t)sq * ellw * n + ell_pos[k]] : make_double2(0, 0);
    }
}

// DUAL: a second output y2/ypart2 = (D_sel2 + G-part) x from the same G sweeps (beta must be 0); with
// cH = -1, sel = 1, sel2 = 0 this is the CN pair (S - H) x, (S + H) x, each bit-identical to its own launch.
// HW = 2: one workgroup of two waves serves both output halves h of a tile (wave h), staging each window once
// for both; HW = 1: one wave per (tile, h).
// The G-part sums the source halves hp and parts v (Z0, Z2, W0, W1) in the order (hp, v).
// FUSED: no V input.  The windows of V_{l,v} (all four parts of a half hp) are formed from x by the coupling
// lists of phase A (aoff, ack, acf: the lists of k_angular_list), with the same terms in the same order as
// k_angular_list forms V, so the results are bit-identical to the two-kernel form and V never goes to HBM.
template <int BW, bool DUAL, class T, int R, int P, int HW = 1, bool FUSED = false>
__global__ __launch_bounds__(64 * HW) void k_radial_dia(
    const T *__restrict__ x, const T *__restrict__ V, const T *__restrict__ GD, const T *__restrict__ DD,
    const int *__restrict__ tiles, const int *__restrict__ tile_q, const double *__restrict__ tile_k, int nblk, int n,
    int nh, int nhp, int nseg, int lgrid, int nq, int sel, double cH, double beta, T *__restrict__ y,
    T *__restrict__ ypart, const int *gate, int gstride, const int *__restrict__ tile_b, size_t gd_stride,
    T *__restrict__ y2, T *__restrict__ ypart2, int sel2, const int *__restrict__ aoff = nullptr,
    const int2 *__restrict__ ack = nullptr, const T *__restrict__ acf = nullptr, int lmax = 0, int nblk1 = 0) {
    constexpr int SEG = 64 * P, ND = 2 * BW + 1, W = SEG + 2 * BW;
    constexpr int NT = 64 * HW;  // threads of the workgroup
    using Rt = typename cx<T>::R;
    __shared__ T sm[(FUSED ? 4 : 1) * R * W];  // FUSED: the windows of the four parts
    const int lane = threadIdx.x & 63;
    // (segment, multipole, tile [, half]) with tiles fastest, in XCD order: an XCD works through a few
    // (segment, multipole) slices at a time, whose G-sums and x windows stay in its L2
    const int nt2 = (int)gridDim.x / (nseg * lgrid);
    int rest = xcd_order((int)blockIdx.x, (int)gridDim.x);
    const int t2 = rest % nt2;
    rest /= nt2;
    const int l = rest % lgrid;
    const int seg = rest / lgrid;
    const int h = HW == 2 ? __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6)) : (t2 & 1);
    const int tile = HW == 2 ? t2 : t2 >> 1;
    const int inst = tile_b[tile];
    GADE_GATED_B(gate, gstride, inst);
    GD += (size_t)inst * gd_stride;  // this instance's G-sums
    const Rt kap = (Rt)tile_k[tile];
    const int q = tile_q[tile];
    const int i0 = seg * SEG;
    int rows[R];
#pragma unroll
    for (int i = 0; i < R; i++) rows[i] = tiles[tile * R + i];
    const size_t plane = (size_t)nblk * n;
    const int S = 4 * ND * nhp;

    T acc[R][P];
#pragma unroll
    for (int i = 0; i < R; i++)
#pragma unroll
        for (int j = 0; j < P; j++) acc[i][j] = czero<T>();

    // window of half hp of 'src' for every row of the tile: position p <-> half index i0 - BW + p.
    // All loads of the window are issued before the first LDS write.
    auto stage = [&](const T *src, int hp) {
        constexpr int NC = (W + NT - 1) / NT;
        T t[R][NC];
#pragma unroll
        for (int i = 0; i < R; i++) {
            const T *rp = src + (size_t)(rows[i] >= 0 ? rows[i] : 0) * n + hp * nh;
#pragma unroll
            for (int c = 0; c < NC; c++) {
                const int p = c * NT + (int)threadIdx.x;
                const int b = i0 - BW + p;
                t[i][c] = czero<T>();
                if (rows[i] >= 0 && p < W && b >= 0 && b < nh) t[i][c] = rp[b];
            }
        }
        __syncthreads();
#pragma unroll
        for (int i = 0; i < R; i++)
#pragma unroll
            for (int c = 0; c < NC; c++) {
                const int p = c * NT + (int)threadIdx.x;
                if (p < W) sm[i * W + dia_swz<P>(p)] = t[i][c];
            }
        __syncthreads();
    };
    // FUSED: the windows of V_{l,v} (v = 0..3) for half hp of every row into sm[v][i][p], each entry the sums
    // Z0 = sum_e h0_e x_{c_e}, Z2 = sum_e h0_e (k_{c_e} x_{c_e}), W0, W1 likewise with h1, over the row's list
    // (l) in list order from +0, as k_angular_list forms them (positions outside the half are 0).  Wave h of a
    // two-wave workgroup takes rows h, h + 2, ...; each lane positions lane + 64c.  The list entries are
    // wave-uniform (scalar loads); each x value is loaded once for the four sums.
    auto gather = [&](int hp) {
        constexpr int NCG = (W + 63) / 64;
        constexpr int U = kAngListU;
        const int w0 = HW == 2 ? h : 0;
        const T *xh = x + (size_t)inst * nblk1 * n + hp * nh;
        int bc[NCG];
#pragma unroll
        for (int c = 0; c < NCG; c++) bc[c] = min(max(i0 - BW + c * 64 + lane, 0), nh - 1);
#pragma unroll
        for (int k = 0; k < R / HW; k++) {
            const int i = w0 + HW * k;
            T z[4][NCG];
#pragma unroll
            for (int v = 0; v < 4; v++)
#pragma unroll
                for (int c = 0; c < NCG; c++) z[v][c] = czero<T>();
            if (rows[i] >= 0) {
                const int *off = aoff + (size_t)(rows[i] - inst * nblk1) * lmax + l;
                const int e1 = off[1];
#pragma unroll 2
                for (int e = off[0]; e < e1; e += U) {
                    T xv[U][NCG];
#pragma unroll
                    for (int u = 0; u < U; u++) {
                        const T *xc = xh + (size_t)ack[e + u].x * n;
#pragma unroll
                        for (int c = 0; c < NCG; c++) xv[u][c] = xc[bc[c]];
                    }
#pragma unroll
                    for (int u = 0; u < U; u++) {
                        const T h0 = acf[2 * (e + u)], h1 = acf[2 * (e + u) + 1];
                        const Rt kc = (Rt)ack[e + u].y;
#pragma unroll
                        for (int c = 0; c < NCG; c++) {
                            const T kx = cscale(xv[u][c], kc);
                            z[0][c] = cfma(h0, xv[u][c], z[0][c]);
                            z[1][c] = cfma(h0, kx, z[1][c]);
                            z[2][c] = cfma(h1, xv[u][c], z[2][c]);
                            z[3][c] = cfma(h1, kx, z[3][c]);
                        }
                    }
                }
            }
            if (k == 0) __syncthreads();  // the previous sweeps' reads of sm are done
#pragma unroll
            for (int c = 0; c < NCG; c++) {
                const int p = c * 64 + lane;
                const int b = i0 - BW + p;
                const bool ok = rows[i] >= 0 && b >= 0 && b < nh;
                if (p < W) {
#pragma unroll
                    for (int v = 0; v < 4; v++) sm[(v * R + i) * W + dia_swz<P>(p)] = ok ? z[v][c] : czero<T>();
                }
            }
        }
        __syncthreads();
    };
    // acc[i][j] += sum_d M_d[i0 + P*lane + j] * win[i][P*lane + j + d] (win: R windows of W); with two
    // coefficient streams M = C1 + kap*C2 (pairs), else M = C1.  Coefficients are fetched PF diagonals ahead.
    auto sweep = [&](const T *win, const T *C, bool pair) {
        constexpr int PF = P >= 4 ? 2 : P == 2 ? 3 : (R <= 8 ? 8 : 4);  // coefficient diagonals in flight
        const int stride = pair ? 2 : 1;
        const T *cp = C + (size_t)(i0 + P * lane) * stride;
        T w[R][P];
#pragma unroll
        for (int i = 0; i < R; i++)
#pragma unroll
            for (int j = 0; j < P; j++) w[i][j] = win[i * W + dia_swz<P>(P * lane + j)];
        T c1[PF][P], c2[PF][P];
#pragma unroll
        for (int d = 0; d < PF - 1; d++)
#pragma unroll
            for (int j = 0; j < P; j++) {
                const T *np = cp + (size_t)d * nhp * stride;
                c1[d][j] = np[j * stride];
                c2[d][j] = pair ? np[j * stride + 1] : czero<T>();
            }
#pragma unroll
        for (int d = 0; d < ND; d++) {
            if (d + PF - 1 < ND) {
                const T *np = cp + (size_t)(d + PF - 1) * nhp * stride;
#pragma unroll
                for (int j = 0; j < P; j++) {
                    c1[(d + PF - 1) % PF][j] = np[j * stride];
                    c2[(d + PF - 1) % PF][j] = pair ? np[j * stride + 1] : czero<T>();
                }
            }
            T M[P];
#pragma unroll
            for (int j = 0; j < P; j++) {
                const T a = c1[d % PF][j], g = c2[d % PF][j];
                M[j] = pair ? cx<T>::mk(fma(kap, g.x, a.x), fma(kap, g.y, a.y)) : a;
            }
#pragma unroll
            for (int i = 0; i < R; i++) {
#pragma unroll
                for (int j = 0; j < P; j++) acc[i][j] = cfma(M[j], w[i][j], acc[i][j]);
                if (d + 1 < ND) {
#pragma unroll
                    for (int j = 0; j + 1 < P; j++) w[i][j] = w[i][j + 1];
                    w[i][P - 1] = win[i * W + dia_swz<P>(P * lane + P + d)];
                }
            }
        }
    };

    // acc -> output (per-multipole partial, or y itself when there is a single multipole grid)
#define DIA_EMIT(YO, YPO)                                                     \
    _Pragma("unroll") for (int i = 0; i < R; i++) {                           \
        if (rows[i] < 0) continue;                                            \
        _Pragma("unroll") for (int j = 0; j < P; j++) {                       \
            const int ip = i0 + P * lane + j;                                 \
            if (ip >= nh) continue;                                           \
            const size_t o = (size_t)rows[i] * n + h * nh + ip;               \
            if (YPO) {                                                        \
                (YPO)[(size_t)l * plane + o] = acc[i][j];                     \
            } else {                                                          \
                T r_ = acc[i][j];                                             \
                if (beta != 0.0) r_ = cadd(r_, cscale((YO)[o], (Rt)beta));    \
                (YO)[o] = r_;                                                 \
            }                                                                 \
        }                                                                     \
    }

    if (cH != 0.0) {
        // k = (hp, v): one copy of the sweep in the code (its own loops are unrolled)
#pragma unroll 1
        for (int k = 0; k < 8; k++) {
            const int hp = k >> 2, v = k & 3;
            if constexpr (FUSED) {
                if (v == 0) gather(hp);
            } else {
                stage(V + (size_t)(l * 4 + v) * plane, hp);
            }
            sweep(FUSED ? sm + v * R * W : sm, GD + ((size_t)(l * 4 + v) * S + (size_t)(2 * h + hp) * ND * nhp) * 2,
                  true);
        }
    }
    // the second output's unscaled G-part: final for l > 0, parked in its own output slot for l = 0
    if constexpr (DUAL) {
        DIA_EMIT(y2, ypart2)
    }
    if (cH != 0.0 && cH != 1.0) {
#pragma unroll
        for (int i = 0; i < R; i++)
#pragma unroll
            for (int j = 0; j < P; j++) acc[i][j] = cscale(acc[i][j], (Rt)cH);
    }
    if (l == 0) {
        for (int hp = 0; hp < 2; hp++) {
            stage(x, hp);
            sweep(sm, DD + ((size_t)(sel * nq + q) * S + (size_t)(2 * h + hp) * ND * nhp), false);
        }
    }
    DIA_EMIT(y, ypart)
    if constexpr (DUAL) {
        if (l == 0) {
            // reload this lane's parked values (written by this thread above), then add D_sel2 x
#pragma unroll
            for (int i = 0; i < R; i++)
#pragma unroll
                for (int j = 0; j < P; j++) {
                    const int ip = i0 + P * lane + j;
                    acc[i][j] = czero<T>();
                    if (rows[i] >= 0 && ip < nh) {
                        const size_t o = (size_t)rows[i] * n + h * nh + ip;
                        acc[i][j] = ypart2 ? ypart2[o] : y2[o];
                    }
                }
            for (int hp = 0; hp < 2; hp++) {
                stage(x, hp);
                sweep(sm, DD + ((size_t)(sel2 * nq + q) * S + (size_t)(2 * h + hp) * ND * nhp), false);
            }
            DIA_EMIT(y2, ypart2)
        }
    }
#undef DIA_EMIT
}

// ---------------------------------------------------------------------------
// fused operator, phase B on the FP64 matrix cores (v_mfma_f64_4x4x4_4b_f64)
//
// One wave owns 16 rows of the banded radial product in two halves of 8 rows, each half of one kappa, one
// output half h, one multipole l and a run of NIT "i-steps" of 4 consecutive outputs.  The MFMA's four
// blocks b are four groups of 4 rows; within a block, M = 4 outputs (ii), N = 4 rows (j), K = 4 window
// positions (kk).  Lanes: A lane = kk*16 + b*4 + ii, B lane = kk*16 + b*4 + j, D lane = ii*16 + b*4 + j
// (measured on gfx950).  For i-step mi and K-step s (s = 0..4), block b computes
//   D[ii][j] += sum_kk M[4 mi + ii][d = 4 s + kk - ii] * V[row_j][4 mi + 4 s + kk - 7]
// over the band d = 0..14 (M = 0 outside it): 20 window positions for the 18 an i-step needs.  Every lane
// loads and folds its own A entry (M = C1 + kappa C2 with its half's kappa; blocks 0-1 and 2-3 hold equal
// A values), so the instruction runs without A broadcast (cbsz 0: with cbsz 1 the chained products come
// out wrong on gfx950).  The B fragment of (mi, s) is the fragment of
// (mi + 1, s - 1): each wave loads every window fragment once and keeps five in a ring.  Complex products
// are four real MFMAs: Re += Mr Vr + Mi (-Vi), Im += Mr Vi + Mi Vr.
// The coefficients come in fragment order: GF[l][v][qd][mi][s][e = kk*4 + ii][C1re, C1im, C2re, C2im]
// (zero outside the band), so the A loads of a wave are 512 consecutive bytes; the stationary operators
// D_sel(kappa) likewise as DF[sel][q][qd][mi][s][e] (complex).
// ---------------------------------------------------------------------------
constexpr int kMfmaNIT = 8;     // i-steps (4 outputs each) per wave
constexpr int kMfmaFrag = 16;   // entries per coefficient fragment (kk, ii)

// GF[b][l][v][qd][mi][s][e] from the DIA-ordered couplings: the (G, G') pair of k_gsums_dia at
// (l, v, qd, d = 4 s + kk - ii, i = 4 mi + ii), zero outside the band
__global__ __launch_bounds__(256) void k_gsums_frag(const double2 *__restrict__ gs, const double2 *__restrict__ gts,
                                                    CoefGroup cg, int b0, size_t gf_stride, int lmax, int nd, int nhp,
                                                    int nmi, double *__restrict__ GF0, const int *gate, int gstride,
                                                    const Coef6 *__restrict__ cf_dev, const int *__restrict__ gctl,
                                                    int B) {
    const int b = b0 + blockIdx.y;
    GADE_GATED_B(gate, gstride, b);
    const Coef6 cf = cf_dev ? cf_dev[(size_t)gctl[GC_WSLOT] * B + b] : cg.c[blockIdx.y];
    double2 c_even[3], c_odd[3];
#pragma unroll
    for (int j = 0; j < 3; j++) {
        c_even[j] = cf.c[2 * j + 1];
        c_odd[j] = cf.c[2 * j];
    }
    double *__restrict__ GF = GF0 + (size_t)b * gf_stride;
    const int S = 4 * nd * nhp;                 // entries per (l, j, nn) plane of gs / gts
    const int F = 4 * nmi * 5 * kMfmaFrag;      // fragment entries per (l, v)
    const long long total = (long long)lmax * 4 * F;
    for (long long idx = blockIdx.x * (long long)blockDim.x + threadIdx.x; idx < total;
         idx += (long long)gridDim.x * blockDim.x) {
        const int lv = (int)(idx / F);
        int rem = (int)(idx - (long long)lv * F);
        const int e = rem % kMfmaFrag;
        rem /= kMfmaFrag;
        const int sk = rem % 5;
        rem /= 5;
        const int mi = rem % nmi;
        const int qd = rem / nmi;
        const int ii = e & 3, kk = e >> 2;
        const int d = 4 * sk + kk - ii, ip = 4 * mi + ii;
        double2 G1 = make_double2(0, 0), G2 = make_double2(0, 0);
        if (d >= 0 && d < nd) {
            const int l = lv >> 2, v = lv & 3;
            const double2 *src = v < 2 ? gs : gts;
            const int n1 = v == 0 ? 0 : v == 1 ? 2 : v == 2 ? 0 : 1, n2 = v == 0 ? 1 : v == 1 ? 3 : v == 2 ? 2 : 3;
            const int ent = (qd * nd + d) * nhp + ip;
#pragma unroll
            for (int j = 0; j < 3; j++) {
                const double2 c = (l % 2 == 0) ? c_even[j] : c_odd[j];
                const size_t o = (size_t)(l * 3 + j) * 4 * S + ent;
                G1 = cfma(c, src[o + (size_t)n1 * S], G1);
                G2 = cfma(c, src[o + (size_t)n2 * S], G2);
            }
        }
        double *dst = GF + (size_t)idx * 4;
        dst[0] = G1.x;
        dst[1] = G1.y;
        dst[2] = G2.x;
        dst[3] = G2.y;
    }
}

// DF[sel][q][qd][mi][s][e] from DD[sel][q][qd][d][i]
__global__ void k_build_DF(const double2 *__restrict__ DD, int nq, int nd, int nhp, int nmi, double2 *__restrict__ DF) {
    const long long total = (long long)3 * nq * 4 * nmi * 5 * kMfmaFrag;
    for (long long idx = blockIdx.x * (long long)blockDim.x + threadIdx.x; idx < total;
         idx += (long long)gridDim.x * blockDim.x) {
        long long rem = idx;
        const int e = (int)(rem % kMfmaFrag);
        rem /= kMfmaFrag;
        const int sk = (int)(rem % 5);
        rem /= 5;
        const int mi = (int)(rem % nmi);
        const long long sq = rem / nmi;  // (sel * nq + q) * 4 + qd
        const int ii = e & 3, kk = e >> 2;
        const int d = 4 * sk + kk - ii, ip = 4 * mi + ii;
        DF[idx] = (d >= 0 && d < nd) ? DD[((size_t)sq * nd + d) * nhp + ip] : make_double2(0, 0);
    }
}

__device__ __forceinline__ double mfma4(double a, double b, double c) {
    return __builtin_amdgcn_mfma_f64_4x4x4f64(a, b, c, 0, 0, 0);  // per-block A, no broadcast
}

template <bool DUAL, int NIT>
__global__ __launch_bounds__(64) void k_radial_mfma(const double2 *__restrict__ x, const double2 *__restrict__ V,
                                                    const double *__restrict__ GF, const double2 *__restrict__ DF,
                                                    const int *__restrict__ grows, const int *__restrict__ gq,
                                                    const double *__restrict__ gk, const int *__restrict__ gb,
                                                    int n, int nh, int nmi, int nseg, int lgrid, int nq, int sel,
                                                    double cH, double beta, double2 *__restrict__ y,
                                                    double2 *__restrict__ ypart, size_t plane, const int *gate,
                                                    int gstride, size_t gf_stride, double2 *__restrict__ y2,
                                                    double2 *__restrict__ ypart2, int sel2) {
    const int lane = threadIdx.x;
    int rest = blockIdx.x;
    const int l = rest % lgrid;
    rest /= lgrid;
    const int seg = rest % nseg;
    rest /= nseg;
    const int h = rest & 1;
    const int g = rest >> 1;
    const int inst = gb[g];
    GADE_GATED_B(gate, gstride, inst);
    const int jj = lane & 3, blk = (lane >> 2) & 3, kk = lane >> 4;
    const int half = blk >> 1;
    const int row = grows[g * 16 + blk * 4 + jj];      // this lane's B / D row, -1 pads
    const double kap = gk[g * 2 + half];
    const int q = gq[g * 2 + half];
    const int e = kk * 4 + jj;                         // this lane's A entry (jj plays ii)
    const int mi0 = seg * NIT;
    const size_t frag = (size_t)5 * kMfmaFrag;         // entries per i-step
    GF += (size_t)inst * gf_stride;

    double accR[NIT], accI[NIT];
#pragma unroll
    for (int it = 0; it < NIT; it++) accR[it] = accI[it] = 0.0;

    // acc += M * window over the i-steps of this wave; A entries from coef (fragment order, 'pair': C1 + kap C2
    // from 4 doubles, else one complex), window values src[0..nh) of this lane's row (nullptr: zero)
    auto sweep = [&](const double *coef, bool pair, const double2 *src) {
        double bR[5], bI[5], bN[5];
        auto load_b = [&](int m, int slot) {
            const int pos = 4 * m + kk - 7;
            double2 v = make_double2(0.0, 0.0);
            if (src && pos >= 0 && pos < nh) v = src[pos];
            bR[slot] = v.x;
            bI[slot] = v.y;
            bN[slot] = -v.y;
        };
#pragma unroll
        for (int k = 0; k < 4; k++) load_b(mi0 + k, k);
#pragma unroll
        for (int it = 0; it < NIT; it++) {
            load_b(mi0 + it + 4, (it + 4) % 5);
            const int mi = mi0 + it;
#pragma unroll
            for (int sk = 0; sk < 5; sk++) {
                const size_t ent = (size_t)mi * frag + (size_t)sk * kMfmaFrag + e;
                double mr, mv;
                if (pair) {
                    const double4 c = *(const double4 *)(coef + ent * 4);
                    mr = fma(kap, c.z, c.x);
                    mv = fma(kap, c.w, c.y);
                } else {
                    const double2 c = *(const double2 *)(coef + ent * 2);
                    mr = c.x;
                    mv = c.y;
                }
                const int sl = (it + sk) % 5;
                accR[it] = mfma4(mr, bR[sl], accR[it]);
                accR[it] = mfma4(mv, bN[sl], accR[it]);
                accI[it] = mfma4(mr, bI[sl], accI[it]);
                accI[it] = mfma4(mv, bR[sl], accI[it]);
            }
        }
    };

    // D lanes: output row 'row', position 4 (mi0 + it) + (lane >> 4) of half h
    const int ii_out = lane >> 4;
    auto emit = [&](double2 *YO, double2 *YPO) {
        if (row < 0) return;
#pragma unroll
        for (int it = 0; it < NIT; it++) {
            const int ip = 4 * (mi0 + it) + ii_out;
            if (ip >= nh) continue;
            const size_t o = (size_t)row * n + h * nh + ip;
            double2 r = make_double2(accR[it], accI[it]);
            if (YPO) {
                YPO[(size_t)l * plane + o] = r;
            } else {
                if (beta != 0.0) r = cadd(r, cscale(YO[o], beta));
                YO[o] = r;
            }
        }
    };

    if (cH != 0.0) {
        for (int v = 0; v < 4; v++)
            for (int hp = 0; hp < 2; hp++) {
                const double *coef = GF + ((size_t)((l * 4 + v) * 4 + 2 * h + hp) * nmi) * frag * 4;
                const double2 *src = row >= 0 ? V + (size_t)(l * 4 + v) * plane + (size_t)row * n + hp * nh : nullptr;
                sweep(coef, true, src);
            }
    }
    if constexpr (DUAL) emit(y2, ypart2);
    if (cH != 0.0 && cH != 1.0) {
#pragma unroll
        for (int it = 0; it < NIT; it++) {
            accR[it] *= cH;
            accI[it] *= cH;
        }
    }
    const size_t qplane = (size_t)4 * nmi * frag;  // DF entries per (sel, q)
    if (l == 0) {
        for (int hp = 0; hp < 2; hp++) {
            const double2 *coef = DF + ((size_t)sel * nq + q) * qplane + (size_t)(2 * h + hp) * nmi * frag;
            sweep((const double *)coef, false, row >= 0 ? x + (size_t)row * n + hp * nh : nullptr);
        }
    }
    emit(y, ypart);
    if constexpr (DUAL) {
        if (l == 0) {
            // reload the parked second output (written by this lane above), then add D_sel2 x
#pragma unroll
            for (int it = 0; it < NIT; it++) {
                const int ip = 4 * (mi0 + it) + ii_out;
                accR[it] = accI[it] = 0.0;
                if (row >= 0 && ip < nh) {
                    const size_t o = (size_t)row * n + h * nh + ip;
                    const double2 p = ypart2 ? ypart2[o] : y2[o];
                    accR[it] = p.x;
                    accI[it] = p.y;
                }
            }
            for (int hp = 0; hp < 2; hp++) {
                const double2 *coef = DF + ((size_t)sel2 * nq + q) * qplane + (size_t)(2 * h + hp) * nmi * frag;
                sweep((const double *)coef, false, row >= 0 ? x + (size_t)row * n + hp * nh : nullptr);
            }
            emit(y2, ypart2);
        }
    }
}

// per-block S-norm contributions pop[r] = x_r^H S(k_r) x_r
__global__ __launch_bounds__(256) void k_snorm(const double2 *__restrict__ x, const double2 *__restrict__ D,
                                               const int *__restrict__ ell_col, const int *__restrict__ blk_q,
                                               int n, int ellw, int nq, double *__restrict__ pop) {
    __shared__ double sh[4 * 16];
    const int r = blockIdx.x;
    const double2 *Dq = D + ((size_t)2 * nq + blk_q[r]) * ellw * n;
    const double2 *xr = x + (size_t)r * n;
    double acc[1] = {0};
    for (int a = threadIdx.x; a < n; a += blockDim.x) {
        double2 sx = make_double2(0, 0);
        for (int s = 0; s < ellw; s++) sx = cfma(Dq[(size_t)s * n + a], xr[ell_col[(size_t)s * n + a]], sx);
        double2 xa = xr[a];
        acc[0] += xa.x * sx.x + xa.y * sx.y;  // Re(conj(x) Sx); S Hermitian
    }
    block_sum<1>(acc, sh);
    if (threadIdx.x == 0) pop[r] = acc[0];
}

// ---------------------------------------------------------------------------
// assembled reference formulation
// ---------------------------------------------------------------------------
// Hfull row (R, a) has hall_rowlen(R) * sub_rowlen(a) entries ordered by (block col, sub col):
// Ap = hAp[R]*nnzb + (hAp[R+1]-hAp[R]) * subAp[a]  (closed form of tools/csr.c:653-700)
__global__ void k_full_Ap(const int *hAp, const int *subAp, int nblk, int n, int nnzb, int *fAp) {
    long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x;
    long long N = (long long)nblk * n;
    if (i > N) return;
    if (i == N) { fAp[N] = hAp[nblk] * nnzb; return; }
    int R = (int)(i / n), a = (int)(i % n);
    fAp[i] = hAp[R] * nnzb + (hAp[R + 1] - hAp[R]) * subAp[a];
}

__global__ void k_full_Ai(const int *hAp, const int *hAi, const int *subAp, const int *subAi, const int *fAp,
                          int nblk, int n, int *fAi) {
    // one block per Hfull row
    long long row = blockIdx.x + (long long)blockIdx.y * gridDim.x;
    if (row >= (long long)nblk * n) return;
    int R = (int)(row / n), a = (int)(row % n);
    int len = subAp[a + 1] - subAp[a];
    int nb = hAp[R + 1] - hAp[R];
    int base = fAp[row];
    for (int e = threadIdx.x; e < nb * len; e += blockDim.x) {
        int pos = e / len, t = e % len;
        fAi[base + e] = hAi[hAp[R] + pos] * n + subAi[subAp[a] + t];
    }
}

// Ai_sub_map[p*nnzb + k] = slot of (Hall nz p, sub entry k) in the expanded CSR (tools/csr.c:702-739)
__global__ void k_full_submap(const int *hAp, const int *subAp, const int *fAp, int nblk, int n, int nnzb,
                              int *rowof, int *submap) {
    // one block per Hall nz p; rowof[p] = block row of p
    int p = blockIdx.x;
    int R = rowof[p];
    int pos = p - hAp[R];
    for (int k = threadIdx.x; k < nnzb; k += blockDim.x) {
        // row a of sub entry k: binary search in subAp
        int lo = 0, hi = n;
        while (hi - lo > 1) {
            int mid = (lo + hi) >> 1;
            if (subAp[mid] <= k) lo = mid; else hi = mid;
        }
        int a = lo, len = subAp[a + 1] - subAp[a];
        submap[(size_t)p * nnzb + k] = fAp[(size_t)R * n + a] + pos * len + (k - subAp[a]);
    }
}

// Hfull.Ax in row order: entry e of row (R, a) -> Hall nz p = hAp[R] + e/len, sub entry k
// (tools/gpu_solver.cu:84-135 evaluated per destination so the stores are coalesced)
template <int LM>
__global__ __launch_bounds__(256) void k_assemble(const int *__restrict__ hAp, const int *__restrict__ hAi,
                                                  const int *__restrict__ kap, const double2 *__restrict__ coef,
                                                  const int *__restrict__ diag, const int *__restrict__ subAp,
                                                  const int *__restrict__ ell_pos, const double2 *__restrict__ GE,
                                                  const double2 *__restrict__ D, const int *__restrict__ blk_q,
                                                  const int *__restrict__ fAp, int nblk, int n, int ellw, int nq,
                                                  double2 *__restrict__ fAx) {
    long long row = blockIdx.x + (long long)blockIdx.y * gridDim.x;
    if (row >= (long long)nblk * n) return;
    const int R = (int)(row / n), a = (int)(row % n);
    const int len = subAp[a + 1] - subAp[a];
    const int nb = hAp[R + 1] - hAp[R];
    const double ki = kap[R];
    const size_t slab = (size_t)ellw * n;
    const double2 *Dq = D + ((size_t)0 * nq + blk_q[R]) * slab;  // S + ihdt H0: take the H0 part below
    const double2 *Sq = D + ((size_t)2 * nq + blk_q[R]) * slab;
    const int base = fAp[row];
    for (int e = threadIdx.x; e < nb * len; e += blockDim.x) {
        const int pos = e / len, t = e % len;
        const int p = hAp[R] + pos;
        const double kj = kap[hAi[p]];
        const int k = subAp[a] + t;
        const size_t ep = (size_t)ell_pos[k];
        double2 val = make_double2(0, 0);
        if (diag[p]) val = csub(Dq[ep], Sq[ep]);  // ihdt*(h00 + k h01 + k^2 h02 + k^3 h03)
        const double2 *cp = coef + (size_t)p * LM * 2;
#pragma unroll
        for (int l = 0; l < LM; l++) {
            const double2 *G = GE + (size_t)(l * 8) * slab + ep;
            double2 h0 = cp[2 * l], h1 = cp[2 * l + 1];
            if (cnz(h0)) {
                double2 g = cadd(cadd(G[0], cscale(G[slab], ki)), cadd(cscale(G[2 * slab], kj), cscale(G[3 * slab], ki * kj)));
                val = cfma(h0, g, val);
            }
            if (cnz(h1)) {
                double2 g = cadd(cadd(G[4 * slab], cscale(G[5 * slab], kj)), cadd(cscale(G[6 * slab], ki), cscale(G[7 * slab], ki * kj)));
                val = cfma(h1, g, val);
            }
        }
        fAx[base + e] = val;
    }
}

// ===========================================================================
// host side
// ===========================================================================
namespace {

}  // namespace

int gade::hip_alloc(void **p, size_t bytes) {
    if (bytes == 0) bytes = 16;
    hipError_t e = hipMalloc(p, bytes);
    if (e != hipSuccess) {
        gade::set_error(std::string("hipMalloc(") + std::to_string(bytes) + "): " + hipGetErrorString(e));
        return GADE_ENOMEM;
    }
    return GADE_OK;
}

namespace {
using gade::hip_alloc;

template <typename T>
int upload(T **dst, const T *src, size_t count) {
    GADE_TRY(hip_alloc((void **)dst, sizeof(T) * count));
    if (count) GADE_HIP(hipMemcpy(*dst, src, sizeof(T) * count, hipMemcpyHostToDevice));
    return GADE_OK;
}

int locate(const int *list, int n, int value) {  // first index with list[i] >= value
    return (int)(std::lower_bound(list, list + n, value) - list);
}

std::complex<double> csr_value(const int *Ap, const int *Ai, const double *Ax, int row, int col) {
    int b = Ap[row], n = Ap[row + 1] - b;
    int cp = b + locate(Ai + b, n, col);
    if (cp == Ap[row + 1] || Ai[cp] != col) return {0, 0};
    return {Ax[2 * cp], Ax[2 * cp + 1]};
}

unsigned grid_for(long long total, int threads, unsigned cap = 65535u * 4) {
    long long g = (total + threads - 1) / threads;
    if (g < 1) g = 1;
    if (g > cap) g = cap;
    return (unsigned)g;
}

void free_all(gade_model *m) {
    if (m->stream) (void)hipStreamSynchronize(m->stream);  // no launch may still read what is freed below
    gade::blocklu_free(m->pre);
    m->pre = nullptr;
    void *ptrs[] = {m->d_ell_lcol, m->d_locj, m->d_nloc, m->d_ell_col, m->d_ell_src, m->d_tmap, m->d_ell_pos, m->d_g, m->d_s, m->d_h0, m->d_GE, m->d_D,
                    m->d_hAp, m->d_hAi, m->d_kap, m->d_blk_q, m->d_coef, m->d_diag, m->d_tiles, m->d_tile_q, m->d_tile_k,
                    m->d_V, m->d_ypart, m->d_flag, m->d_dia_src, m->d_gs, m->d_gts, m->d_GD, m->d_DD, m->d_dtiles, m->d_dtile_q, m->d_dtile_k,
                    m->r, m->r0, m->v, m->p, m->y, m->z, m->s, m->t, m->tmp, m->d_sc, m->d_part, m->d_part2,
                    m->d_fAp, m->d_fAi, m->d_fAx, m->d_SAp, m->d_SAi, m->d_SAx};
    for (void *p : ptrs)
        if (p) hipFree(p);
    if (m->h_sc) hipHostFree(m->h_sc);
    if (m->h_ctl) hipHostFree(m->h_ctl);
    if (m->h_gctl) hipHostFree(m->h_gctl);
    if (m->h_cfwin) hipHostFree(m->h_cfwin);
    gade::lg_free(m->lg);
    m->lg = nullptr;
    if (m->d_cfwin) hipFree(m->d_cfwin);
    if (m->graph_exec) (void)hipGraphExecDestroy(m->graph_exec);
    if (m->graph) (void)hipGraphDestroy(m->graph);
    if (m->cap_stream) (void)hipStreamDestroy(m->cap_stream);
    for (void *p : {(void *)m->d_ctl, (void *)m->d_iters, (void *)m->d_tols, (void *)m->d_evec, (void *)m->d_eval,
                    (void *)m->d_hist, (void *)m->d_gctl, (void *)m->d_nsteps, (void *)m->d_dtile_b,
                    (void *)m->d_ypart2, (void *)m->d_V32, (void *)m->d_ypart32, (void *)m->d_GD32,
                    (void *)m->d_DD32, (void *)m->d_coef32, (void *)m->lu32[0], (void *)m->lu32[1],
                    (void *)m->lu32[2], (void *)m->lu32[3], (void *)m->tw32, (void *)m->mp_r, (void *)m->mp_r0, (void *)m->mp_p,
                    (void *)m->mp_v, (void *)m->mp_y, (void *)m->mp_z, (void *)m->mp_s, (void *)m->mp_t,
                    (void *)m->mp_d, (void *)m->d_mp_ax, (void *)m->d_sc32, (void *)m->d_aoff, (void *)m->d_ack,
                    (void *)m->d_acf, (void *)m->d_acf32, (void *)m->d_GF, (void *)m->d_DF, (void *)m->d_mg_rows,
                    (void *)m->d_mg_q, (void *)m->d_mg_b, (void *)m->d_mg_k})
        if (p) hipFree(p);
    resolve_pending(m);
    for (hipEvent_t e : m->event_pool) (void)hipEventDestroy(e);
    m->event_pool.clear();
}

// the angular kernel stages a row's coefficients for its multipole group in LDS
bool ang_lds_ok(const gade_model *m, int lb) {
    return (size_t)m->maxrow * (lb * 2 * sizeof(double2) + sizeof(int2)) <= 60 * 1024;
}

// the angular kernel with the model's multipole group size (2, 4 or 8 multipoles per block; each V entry is
// summed over the row's couplings in the same order for every group size)
template <class T>
int launch_angular_t(gade_model *m, const T *x, const T *coef, const T *acf, T *V) {
    const int LB = m->ang_lb;
    dim3 grid(m->nblk, (m->n + kAngThreads - 1) / kAngThreads, (m->lmax + LB - 1) / LB);
    if (m->ang_list) {
#define GADE_ANGL(LB_)                                                                                        \
    hipLaunchKernelGGL((k_angular_list<LB_, T>), grid, dim3(kAngThreads), 0, m->stream, x, m->d_aoff, m->d_ack, \
                       acf, m->n, m->lmax, m->nblk, V, m->gate, m->gstride, m->nblk1)
        switch (LB) {
            case 2: GADE_ANGL(2); break;
            case 4: GADE_ANGL(4); break;
            case 8: GADE_ANGL(8); break;
            default:
                gade::set_error("angular multipole group must be 2, 4 or 8");
                return GADE_EINVAL;
        }
#undef GADE_ANGL
        GADE_HIP(hipGetLastError());
        return GADE_OK;
    }
    size_t shm = sizeof(T) * (size_t)m->maxrow * LB * 2 + sizeof(int2) * (size_t)m->maxrow;
#define GADE_ANG(LB_)                                                                                            \
    hipLaunchKernelGGL((k_angular<LB_, T>), grid, dim3(kAngThreads), shm, m->stream, x, m->d_hAp, m->d_hAi, coef, \
                       m->d_kap, m->nblk, m->n, m->lmax, m->maxrow, V, m->gate, m->gstride, m->nblk1)
    switch (LB) {
        case 2: GADE_ANG(2); break;
        case 4: GADE_ANG(4); break;
        case 8: GADE_ANG(8); break;
        default:
            gade::set_error("angular multipole group must be 2, 4 or 8");
            return GADE_EINVAL;
    }
#undef GADE_ANG
    GADE_HIP(hipGetLastError());
    return GADE_OK;
}

int launch_angular(gade_model *m, const double2 *x) {
    return launch_angular_t<double2>(m, x, m->d_coef, m->d_acf, m->d_V);
}

template <class T>
bool ang_in_radial(const gade_model *m);

// phase A of an apply into V, unless the banded radial kernel forms its windows from x itself
int angular_phase(gade_model *m, const double2 *x) {
    if (ang_in_radial<double2>(m)) return GADE_OK;
    Prof pr(m, "angular", GADE_PROF_ANGULAR);
    return launch_angular(m, x);
}

// tile shapes (rows per tile, outputs per lane) with an instantiated banded kernel: rows x outputs <= 16
// complex accumulators per lane
bool dia_shape_ok(int r, int p) {
    return (r == 4 && (p == 1 || p == 2 || p == 4)) || (r == 8 && (p == 1 || p == 2)) || (r == 16 && p == 1);
}

// phase A inside the banded radial kernel (its windows formed from x and the coupling lists, no V): the
// FP64 form when the MFMA kernel is off, the FP32 form of the mixed solve always (it has no MFMA kernel)
template <class T>
bool ang_in_radial(const gade_model *m) {
    return m->ang_fused && m->radial_kind == 2 && (sizeof(T) == sizeof(float2) || !m->radial_mfma);
}

// tile shapes with a fused form: one row's window sums per lane fit the registers for up to 2 outputs per lane
bool fused_shape_ok(int r, int p) { return dia_shape_ok(r, p) && p <= 2; }

// the banded radial kernel with the model's tile shape (rows per tile, outputs per lane); every shape sums
// each output in the same order, so all give the same results bit for bit
template <bool DUAL, class T>
int launch_dia(gade_model *m, int lgrid, const T *x, const T *V, const T *GD, const T *DD, int sel, double cH,
               double beta, T *y, T *yp, T *y2, T *yp2, int sel2) {
    const int P = m->dia_p;
    const int nseg = (m->nh + 64 * P - 1) / (64 * P);
    const bool fused = ang_in_radial<T>(m);
    // the paired (DUAL) two-kernel form runs one wave per output half whatever the model's choice: with both
    // halves in one workgroup it measured 30-40 % slower at every shape (its two accumulator sets per wave)
    const int hw = DUAL && !fused ? 1 : m->dia_hw;
    const dim3 grid((unsigned)m->ndtiles * (2 / hw) * nseg * lgrid);
    const T *acf = sizeof(T) == sizeof(float2) ? (const T *)m->d_acf32 : (const T *)m->d_acf;
    if (fused && !fused_shape_ok(m->dia_r, P)) {
        gade::set_error("the fused angular + radial form needs at most 2 outputs per lane");
        return GADE_EINVAL;
    }
#define GADE_DIA1(R_, P_, HW_, F_)                                                                               \
    hipLaunchKernelGGL((k_radial_dia<kDiaBW, DUAL, T, R_, P_, HW_, F_>), grid, dim3(64 * HW_), 0, m->stream, x, V, \
                       GD, DD, m->d_dtiles, m->d_dtile_q, m->d_dtile_k, m->nblk, m->n, m->nh, m->nhp, nseg, lgrid, \
                       m->nq, sel, cH, beta, y, yp, m->gate, m->gstride, m->d_dtile_b, m->gd_stride, y2, yp2, sel2, \
                       m->d_aoff, (const int2 *)m->d_ack, acf, m->lmax, m->nblk1)
#define GADE_DIA(R_, P_)                 \
    if (hw == 2 && fused) {              \
        GADE_DIA1(R_, P_, 2, true);      \
    } else if (fused) {                  \
        GADE_DIA1(R_, P_, 1, true);      \
    } else if (hw == 2) {                \
        GADE_DIA1(R_, P_, 2, false);     \
    } else {                             \
        GADE_DIA1(R_, P_, 1, false);     \
    }
    switch (m->dia_r * 10 + P) {
        case 82: GADE_DIA(8, 2); break;
        case 42: GADE_DIA(4, 2); break;
        case 44:
            if (hw == 2)
                GADE_DIA1(4, 4, 2, false);
            else
                GADE_DIA1(4, 4, 1, false);
            break;
        case 81: GADE_DIA(8, 1); break;
        case 41: GADE_DIA(4, 1); break;
        case 161: GADE_DIA(16, 1); break;
        default:
            gade::set_error("unsupported banded tile shape (rows, outputs per lane)");
            return GADE_EINVAL;
    }
#undef GADE_DIA
#undef GADE_DIA1
    GADE_HIP(hipGetLastError());
    return GADE_OK;
}

// FP64 banded radial phase: the MFMA kernel when enabled, else k_radial_dia with the model's tile shape
template <bool DUAL>
int launch_dia64(gade_model *m, int lgrid, const double2 *x, int sel, double cH, double beta, double2 *y,
                 double2 *yp, double2 *y2, double2 *yp2, int sel2) {
    if (!m->radial_mfma)
        return launch_dia<DUAL, double2>(m, lgrid, x, m->d_V, m->d_GD, m->d_DD, sel, cH, beta, y, yp, y2, yp2, sel2);
    const int nseg = m->nmi / kMfmaNIT;
    const dim3 grid((unsigned)m->nmg * 2 * nseg * lgrid);
    hipLaunchKernelGGL((k_radial_mfma<DUAL, kMfmaNIT>), grid, dim3(64), 0, m->stream, x, (const double2 *)m->d_V,
                       (const double *)m->d_GF, (const double2 *)m->d_DF, m->d_mg_rows, m->d_mg_q, m->d_mg_k,
                       m->d_mg_b, m->n, m->nh, m->nmi, nseg, lgrid, m->nq, sel, cH, beta, y, yp,
                       (size_t)m->nblk * m->n, m->gate, m->gstride, m->gf_stride, y2, yp2, sel2);
    GADE_HIP(hipGetLastError());
    return GADE_OK;
}

// y2 (banded kernel only): also y2 = (D_sel2 + G-part) x from the same pass, beta 0
int launch_radial(gade_model *m, const double2 *x, double2 *y, int sel, double cH, double beta,
                  double2 *y2 = nullptr, int sel2 = 0) {
    const bool inter = cH != 0.0;
    if (m->radial_kind == 2) {
        const int lgrid = inter ? m->lmax : 1;
        double2 *yp = inter && lgrid > 1 ? m->d_ypart : nullptr;
        double2 *yp2 = y2 && yp ? m->d_ypart2 : nullptr;  // allocated with the model when lmax > 1
        if (y2) {
            Prof pr(m, "radial_pair", GADE_PROF_RADIAL);
            GADE_TRY((launch_dia64<true>(m, lgrid, x, sel, cH, 0.0, y, yp, y2, yp2, sel2)));
        } else {
            Prof pr(m, "radial", GADE_PROF_RADIAL);
            GADE_TRY((launch_dia64<false>(m, lgrid, x, sel, cH, beta, y, yp, nullptr, nullptr, 0)));
        }
        if (yp) {
            Prof pr(m, "radial_sum", GADE_PROF_RADIAL_SUM);
            hipLaunchKernelGGL(k_radial_sum<double2>, dim3(grid_for(m->N1, 256, 2048), m->B), dim3(256), 0, m->stream,
                               m->d_ypart, lgrid, m->N, y2 ? 0.0 : beta, y, m->gate, m->gstride, m->N1);
            if (y2)
                hipLaunchKernelGGL(k_radial_sum<double2>, dim3(grid_for(m->N1, 256, 2048), m->B), dim3(256), 0, m->stream,
                                   m->d_ypart2, lgrid, m->N, 0.0, y2, m->gate, m->gstride, m->N1);
            GADE_HIP(hipGetLastError());
        }
        return GADE_OK;
    }
    const int lgrid = inter ? m->lmax : 1;
    dim3 grid(m->ntiles * lgrid * m->nchunk);
    size_t shm = sizeof(double2) * 4 * (size_t)m->tile_r * m->maxloc;
    double2 *part = inter ? m->d_ypart : nullptr;
#define RAD(R) \
    case R: hipLaunchKernelGGL(k_radial<R>, grid, dim3(64), shm, m->stream, x, m->d_V, m->d_GE, m->d_D, m->d_ell_lcol, m->d_locj, m->d_nloc, m->d_tiles, m->d_tile_q, m->d_tile_k, m->nblk, m->n, m->ellw, lgrid, m->nq, m->nchunk, m->maxloc, sel, cH, beta, y, part, m->gate); break;
    {
        Prof pr(m, "radial", GADE_PROF_RADIAL);
        switch (m->tile_r) {
            RAD(1) RAD(2) RAD(4) RAD(8)
            default:
                gade::set_error("unsupported tile width");
                return GADE_EINVAL;
        }
    }
#undef RAD
    GADE_HIP(hipGetLastError());
    if (inter) {
        Prof pr(m, "radial_sum", GADE_PROF_RADIAL_SUM);
        hipLaunchKernelGGL(k_radial_sum<double2>, dim3(grid_for(m->N1, 256, 2048), m->B), dim3(256), 0, m->stream, m->d_ypart,
                           lgrid, m->N, beta, y, m->gate, m->gstride, m->N1);
        GADE_HIP(hipGetLastError());
    }
    return GADE_OK;
}

// group blocks by distinct kappa index and split each group into tiles of R rows (-1 pads)
// per instance, group its blocks by distinct kappa index and split each group into tiles of R rows (-1 pads);
// a tile never mixes instances
int kappa_tiles(const gade_model *m, int R, int **d_tiles, int **d_tq, double **d_tk, int *ntiles,
                int **d_tb = nullptr) {
    std::vector<int> tiles, tq, tb;
    std::vector<double> tk;
    for (int b = 0; b < m->B; b++)
        for (int q = 0; q < m->nq; q++) {
            std::vector<int> rows;
            for (int r = b * m->nblk1; r < (b + 1) * m->nblk1; r++)
                if (m->blk_q[r] == q) rows.push_back(r);
            for (size_t i = 0; i < rows.size(); i += R) {
                for (int j = 0; j < R; j++) tiles.push_back(i + j < rows.size() ? rows[i + j] : -1);
                tq.push_back(q);
                tk.push_back((double)m->kap_q[q]);
                tb.push_back(b);
            }
        }
    *ntiles = (int)tq.size();
    for (void *p : {(void *)*d_tiles, (void *)*d_tq, (void *)*d_tk, d_tb ? (void *)*d_tb : nullptr})
        if (p) (void)hipFree(p);
    *d_tiles = nullptr;
    *d_tq = nullptr;
    *d_tk = nullptr;
    GADE_TRY(upload(d_tiles, tiles.data(), tiles.size()));
    GADE_TRY(upload(d_tq, tq.data(), tq.size()));
    GADE_TRY(upload(d_tk, tk.data(), tk.size()));
    if (d_tb) {
        *d_tb = nullptr;
        GADE_TRY(upload(d_tb, tb.data(), tb.size()));
    }
    return GADE_OK;
}

int build_tiles(gade_model *m) { return kappa_tiles(m, m->tile_r, &m->d_tiles, &m->d_tile_q, &m->d_tile_k, &m->ntiles); }

// Banded form of the radial pattern: n even, every entry (h*nh+i, h'*nh+j)
// with |i-j| <= kDiaBW.  Builds the DIA maps and the static DIA operators.
int setup_dia(gade_model *m) {
    const int n = m->n;
    m->dia_ok = false;
    if (n % 2) return GADE_OK;
    const int nh = n / 2;
    for (int a = 0; a < n; a++)
        for (int k = m->rad_Ap[a]; k < m->rad_Ap[a + 1]; k++)
            if (std::abs(m->rad_Ai[k] % nh - a % nh) > kDiaBW) return GADE_OK;
    m->nh = nh;
    m->nd = 2 * kDiaBW + 1;
    m->nhp = (nh + kDiaSegMax - 1) / kDiaSegMax * kDiaSegMax;
    const int S = 4 * m->nd * m->nhp;
    std::vector<int> dia_src(S, -1);
    for (int qd = 0; qd < 4; qd++)
        for (int d = 0; d < m->nd; d++)
            for (int ip = 0; ip < nh; ip++) {
                const int a = (qd >> 1) * nh + ip, jc = ip + d - kDiaBW;
                if (jc < 0 || jc >= nh) continue;
                const int col = (qd & 1) * nh + jc;
                const int b = m->rad_Ap[a], len = m->rad_Ap[a + 1] - b;
                const int p = b + locate(m->rad_Ai.data() + b, len, col);
                if (p < m->rad_Ap[a + 1] && m->rad_Ai[p] == col) dia_src[((size_t)qd * m->nd + d) * m->nhp + ip] = p;
            }
    const int lmax = m->lmax;
    GADE_TRY(upload(&m->d_dia_src, dia_src.data(), dia_src.size()));
    GADE_TRY(hip_alloc((void **)&m->d_gs, sizeof(double2) * (size_t)lmax * 12 * S));
    GADE_TRY(hip_alloc((void **)&m->d_gts, sizeof(double2) * (size_t)lmax * 12 * S));
    m->gd_stride = (size_t)lmax * 4 * S * 2;
    GADE_TRY(hip_alloc((void **)&m->d_GD, sizeof(double2) * m->gd_stride * m->B));
    GADE_HIP(hipMemset(m->d_GD, 0, sizeof(double2) * m->gd_stride * m->B));  // f = 0 until the first set_time
    GADE_TRY(hip_alloc((void **)&m->d_DD, sizeof(double2) * (size_t)3 * m->nq * S));
    const long long tg = (long long)lmax * 12 * S, td = (long long)3 * m->nq * S;
    hipLaunchKernelGGL(k_build_gdia, dim3(grid_for(tg, 256)), dim3(256), 0, m->stream, m->d_g, m->d_dia_src, m->d_tmap,
                       lmax, m->nd, m->nhp, m->nnzb, m->d_gs, m->d_gts);
    hipLaunchKernelGGL(k_build_DD, dim3(grid_for(td, 256)), dim3(256), 0, m->stream, m->d_D, m->d_dia_src,
                       m->d_ell_pos, m->nq, m->ellw, n, m->nd, m->nhp, m->d_DD);
    GADE_HIP(hipGetLastError());
    GADE_HIP(hipStreamSynchronize(m->stream));
    GADE_TRY(kappa_tiles(m, kDiaR, &m->d_dtiles, &m->d_dtile_q, &m->d_dtile_k, &m->ndtiles, &m->d_dtile_b));
    m->dia_ok = true;
    return GADE_OK;
}

// MFMA form of the banded radial phase: per instance, the rows of each kappa in halves of 8 (padded), two
// halves per wave (one instance per wave); the coefficient fragments GF and the stationary fragments DF.
int setup_mfma(gade_model *m) {
    m->mfma_ok = false;
    if (!m->dia_ok || m->nhp % 4 || (m->nhp / 4) % kMfmaNIT) return GADE_OK;
    m->nmi = m->nhp / 4;
    std::vector<int> rows, qs, bs;
    std::vector<double> ks;
    for (int b = 0; b < m->B; b++) {
        std::vector<std::pair<int, std::vector<int>>> halves;  // (q, up to 8 rows)
        for (int q = 0; q < m->nq; q++) {
            std::vector<int> rq;
            for (int r = b * m->nblk1; r < (b + 1) * m->nblk1; r++)
                if (m->blk_q[r] == q) rq.push_back(r);
            for (size_t i = 0; i < rq.size(); i += 8)
                halves.emplace_back(q, std::vector<int>(rq.begin() + i, rq.begin() + std::min(rq.size(), i + 8)));
        }
        for (size_t w = 0; w < halves.size(); w += 2) {
            for (int hf = 0; hf < 2; hf++) {
                const bool have = w + hf < halves.size();
                const auto &hv = halves[have ? w + hf : w];
                for (int i = 0; i < 8; i++) rows.push_back(have && i < (int)hv.second.size() ? hv.second[i] : -1);
                qs.push_back(hv.first);
                ks.push_back((double)m->kap_q[hv.first]);
            }
            bs.push_back(b);
        }
    }
    m->nmg = (int)bs.size();
    GADE_TRY(upload(&m->d_mg_rows, rows.data(), rows.size()));
    GADE_TRY(upload(&m->d_mg_q, qs.data(), qs.size()));
    GADE_TRY(upload(&m->d_mg_k, ks.data(), ks.size()));
    GADE_TRY(upload(&m->d_mg_b, bs.data(), bs.size()));
    m->gf_stride = (size_t)m->lmax * 4 * 4 * m->nmi * 5 * kMfmaFrag * 4;
    GADE_TRY(hip_alloc((void **)&m->d_GF, sizeof(double) * m->gf_stride * m->B));
    GADE_HIP(hipMemset(m->d_GF, 0, sizeof(double) * m->gf_stride * m->B));  // f = 0 until the first set_time
    const long long tdf = (long long)3 * m->nq * 4 * m->nmi * 5 * kMfmaFrag;
    GADE_TRY(hip_alloc((void **)&m->d_DF, sizeof(double2) * tdf));
    hipLaunchKernelGGL(k_build_DF, dim3(grid_for(tdf, 256)), dim3(256), 0, m->stream, m->d_DD, m->nq, m->nd, m->nhp,
                       m->nmi, m->d_DF);
    GADE_HIP(hipGetLastError());
    GADE_HIP(hipStreamSynchronize(m->stream));
    m->mfma_ok = true;
    return GADE_OK;
}

// G-sums of the current f(t) in the layout of the active radial kernel
// pulse coefficients c_alpha = SoL * (i h dt/2) * f_alpha of instance b
Coef6 coef_of(const gade_model *m, int b) {
    Coef6 cf;
    const double hd = kSoL * m->h * m->dt / 2;
    const double *ft = m->ft.data() + (size_t)12 * b;
    for (int a = 0; a < 6; a++) cf.c[a] = make_double2(-hd * ft[2 * a + 1], hd * ft[2 * a]);
    return cf;
}

// the DIA-layout G-sums feed k_radial_dia (FP64 when the MFMA kernel is off) and the FP32 copies of the
// mixed-precision form
bool dia_gsums_needed(const gade_model *m) { return !m->radial_mfma || m->precision == 1; }

// G-sums of the current f(t) of every instance in the layout of the active radial kernel
int launch_gsums(gade_model *m, bool ell) {
    Prof pr(m, "gsums", GADE_PROF_GSUMS);
    if (ell) {
        long long total = (long long)m->lmax * 8 * m->ellw * m->n;
        hipLaunchKernelGGL(k_gsums, dim3(grid_for(total, 256)), dim3(256), 0, m->stream, m->d_g, m->d_ell_src,
                           m->d_tmap, coef_of(m, 0), m->lmax, m->ellw, m->n, m->nnzb, m->d_GE, m->gate);
    } else {
        long long total = (long long)m->lmax * 16 * m->nd * m->nhp;  // (l, v, qd, d, ip)
        const long long totf = (long long)m->lmax * 16 * m->nmi * 5 * kMfmaFrag;
        for (int b0 = 0; b0 < m->B; b0 += kCoefGroup) {
            const int nb = std::min(kCoefGroup, m->B - b0);
            CoefGroup cg{};
            for (int k = 0; k < nb; k++) cg.c[k] = coef_of(m, b0 + k);
            if (dia_gsums_needed(m))
                hipLaunchKernelGGL(k_gsums_dia, dim3(grid_for(total, 256, 4096), nb), dim3(256), 0, m->stream,
                                   m->d_gs, m->d_gts, cg, b0, m->gd_stride, m->lmax, m->nd, m->nhp, m->d_GD, m->gate,
                                   m->gstride, (const Coef6 *)nullptr, (const int *)nullptr, m->B);
            if (m->radial_mfma)
                hipLaunchKernelGGL(k_gsums_frag, dim3(grid_for(totf, 256, 4096), nb), dim3(256), 0, m->stream,
                                   m->d_gs, m->d_gts, cg, b0, m->gf_stride, m->lmax, m->nd, m->nhp, m->nmi, m->d_GF,
                                   m->gate, m->gstride, (const Coef6 *)nullptr, (const int *)nullptr, m->B);
        }
    }
    GADE_HIP(hipGetLastError());
    return GADE_OK;
}

}  // namespace

// ---------------------------------------------------------------------------
// C ABI
// ---------------------------------------------------------------------------
extern "C" {

int gade_malloc(void **ptr, size_t bytes) { return hip_alloc(ptr, bytes); }
int gade_free(void *ptr) {
    GADE_HIP(hipFree(ptr));
    return GADE_OK;
}
int gade_memcpy_h2d(void *dst, const void *src, size_t bytes) {
    GADE_HIP(hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice));
    return GADE_OK;
}
int gade_memcpy_d2h(void *dst, const void *src, size_t bytes) {
    GADE_HIP(hipMemcpy(dst, src, bytes, hipMemcpyDeviceToHost));
    return GADE_OK;
}
int gade_memcpy_d2d(void *dst, const void *src, size_t bytes) {
    if (!dst || !src) return GADE_EINVAL;
    if (bytes % 16 == 0 && ((uintptr_t)dst | (uintptr_t)src) % 16 == 0) {
        // plain streaming copy: reads and writes exactly 'bytes' (used to calibrate traffic counters)
        const long long n = (long long)(bytes / 16);
        hipLaunchKernelGGL(k_copy, dim3(grid_for(n, 256, 8192)), dim3(256), 0, nullptr, (const double2 *)src,
                           (double2 *)dst, n);
        GADE_HIP(hipGetLastError());
        GADE_HIP(hipDeviceSynchronize());
    } else {
        GADE_HIP(hipMemcpy(dst, src, bytes, hipMemcpyDeviceToDevice));
    }
    return GADE_OK;
}
int gade_memset(void *ptr, int value, size_t bytes) {
    GADE_HIP(hipMemset(ptr, value, bytes));
    return GADE_OK;
}
int gade_device_sync(void) {
    GADE_HIP(hipDeviceSynchronize());
    return GADE_OK;
}

int gade_model_create(const gade_inputs_t *in, double h, double dt, void *stream, gade_model_t **out) {
    return gade_model_create_batch(in, 1, h, dt, stream, out);
}

int gade_model_create_batch(const gade_inputs_t *in, int B, double h, double dt, void *stream, gade_model_t **out) {
    if (!in || !out || in->nblk <= 0 || in->blkdim <= 0 || in->lmax <= 0 || B < 1) {
        gade::set_error("gade_model_create: invalid arguments");
        return GADE_EINVAL;
    }
    if (in->lmax > gade::kMaxL) {
        gade::set_error("gade_model_create: lmax > 16 is not supported");
        return GADE_EINVAL;
    }
    gade_model *m = new gade_model();
    *out = nullptr;
    auto fail = [&](int rc) {
        free_all(m);
        delete m;
        return rc;
    };
    m->nblk = in->nblk;  // one instance until the structure is replicated below
    m->nblk1 = in->nblk;
    m->B = B;
    m->n = in->blkdim;
    m->lmax = in->lmax;
    m->h = h;
    m->dt = dt;
    m->stream = (hipStream_t)stream;
    m->N = (long long)m->nblk * m->n;
    const int n = m->n, lmax = m->lmax;
    m->rad_Ap.assign(in->rad_Ap, in->rad_Ap + n + 1);
    m->nnzb = m->rad_Ap[n];
    m->rad_Ai.assign(in->rad_Ai, in->rad_Ai + m->nnzb);
    m->hall_Ap.assign(in->hall_Ap, in->hall_Ap + m->nblk + 1);
    m->hnnz = m->hall_Ap[m->nblk];
    m->hall_Ai.assign(in->hall_Ai, in->hall_Ai + m->hnnz);
    // structural checks
    for (int a = 0; a < n; a++) {
        if (m->rad_Ap[a + 1] < m->rad_Ap[a]) { gade::set_error("radial Ap not monotone"); return fail(GADE_EINVAL); }
        for (int k = m->rad_Ap[a]; k < m->rad_Ap[a + 1]; k++) {
            if (m->rad_Ai[k] < 0 || m->rad_Ai[k] >= n) { gade::set_error("radial Ai out of range"); return fail(GADE_EINVAL); }
            if (k > m->rad_Ap[a] && m->rad_Ai[k] <= m->rad_Ai[k - 1]) { gade::set_error("radial Ai not sorted"); return fail(GADE_EINVAL); }
        }
    }
    for (int r = 0; r < m->nblk; r++)
        for (int p = m->hall_Ap[r]; p < m->hall_Ap[r + 1]; p++)
            if (m->hall_Ai[p] < 0 || m->hall_Ai[p] >= m->nblk) { gade::set_error("Hall Ai out of range"); return fail(GADE_EINVAL); }
    // ELL of the radial pattern
    int ellw = 0;
    for (int a = 0; a < n; a++) ellw = std::max(ellw, m->rad_Ap[a + 1] - m->rad_Ap[a]);
    m->ellw = ellw;
    std::vector<int> ell_col((size_t)ellw * n), ell_src((size_t)ellw * n), ell_pos(m->nnzb), tmap(m->nnzb);
    for (int a = 0; a < n; a++)
        for (int s = 0; s < ellw; s++) {
            int k = m->rad_Ap[a] + s;
            bool ok = k < m->rad_Ap[a + 1];
            ell_col[(size_t)s * n + a] = ok ? m->rad_Ai[k] : a;
            ell_src[(size_t)s * n + a] = ok ? k : -1;
            if (ok) ell_pos[k] = s * n + a;
        }
    // per 64-row chunk: distinct columns (staged in LDS by the radial kernel)
    m->nchunk = (n + 63) / 64;
    std::vector<std::vector<int>> cols(m->nchunk);
    for (int c = 0; c < m->nchunk; c++) {
        for (int a = c * 64; a < std::min(n, c * 64 + 64); a++)
            for (int s2 = 0; s2 < ellw; s2++) cols[c].push_back(ell_col[(size_t)s2 * n + a]);
        std::sort(cols[c].begin(), cols[c].end());
        cols[c].erase(std::unique(cols[c].begin(), cols[c].end()), cols[c].end());
        m->maxloc = std::max(m->maxloc, (int)cols[c].size());
    }
    std::vector<int> locj((size_t)m->nchunk * m->maxloc, 0), nloc(m->nchunk), ell_lcol((size_t)ellw * n);
    for (int c = 0; c < m->nchunk; c++) {
        nloc[c] = (int)cols[c].size();
        std::copy(cols[c].begin(), cols[c].end(), locj.begin() + (size_t)c * m->maxloc);
    }
    for (int a = 0; a < n; a++)
        for (int s2 = 0; s2 < ellw; s2++) {
            const auto &cc = cols[a / 64];
            ell_lcol[(size_t)s2 * n + a] = (int)(std::lower_bound(cc.begin(), cc.end(), ell_col[(size_t)s2 * n + a]) - cc.begin());
        }
    if (m->maxloc > 64 * kMaxLocW) {
        gade::set_error("radial pattern too wide for the LDS window of the fused operator");
        return fail(GADE_EINVAL);
    }
    // transpose map: entry k=(a,j) -> position of (j,a)  (tools/csr.c:963-972 needs structural symmetry)
    for (int a = 0; a < n; a++)
        for (int k = m->rad_Ap[a]; k < m->rad_Ap[a + 1]; k++) {
            int j = m->rad_Ai[k];
            int b = m->rad_Ap[j], len = m->rad_Ap[j + 1] - b;
            int cp = b + locate(m->rad_Ai.data() + b, len, a);
            if (cp == m->rad_Ap[j + 1] || m->rad_Ai[cp] != a) {
                gade::set_error("radial pattern is not structurally symmetric");
                return fail(GADE_EINVAL);
            }
            tmap[k] = cp;
        }
    if (ellw > kMaxSlots) {
        gade::set_error("radial rows with more than 32 non-zeros are not supported by the fused operator");
        return fail(GADE_EINVAL);
    }
    for (int r = 0; r < m->nblk; r++) m->maxrow = std::max(m->maxrow, m->hall_Ap[r + 1] - m->hall_Ap[r]);
    if (!ang_lds_ok(m, kAngLB)) {
        gade::set_error("Hall rows too long for the angular kernel's LDS staging");
        return fail(GADE_EINVAL);
    }
    // kappa per block and distinct kappa values
    m->kap_blk.resize(m->nblk);
    m->blk_q.resize(m->nblk);
    for (int r = 0; r < m->nblk; r++) {
        int orig = in->hall_perm ? in->hall_perm[r] : r;
        if (orig < 0 || orig >= in->nang) { gade::set_error("Hall perm out of range"); return fail(GADE_EINVAL); }
        m->kap_blk[r] = gade_ik(orig);
        auto it = std::find(m->kap_q.begin(), m->kap_q.end(), m->kap_blk[r]);
        if (it == m->kap_q.end()) {
            m->blk_q[r] = (int)m->kap_q.size();
            m->kap_q.push_back(m->kap_blk[r]);
        } else {
            m->blk_q[r] = (int)(it - m->kap_q.begin());
        }
    }
    m->nq = (int)m->kap_q.size();
    // angular coefficients per Hall nz (tools/gpu_solver.cu:156-188)
    std::vector<double> coef((size_t)m->hnnz * lmax * 4);
    std::vector<int> diag(m->hnnz);
    for (int r = 0; r < m->nblk; r++)
        for (int p = m->hall_Ap[r]; p < m->hall_Ap[r + 1]; p++) {
            int c = m->hall_Ai[p];
            int orow = in->hall_perm ? in->hall_perm[r] : r;
            int ocol = in->hall_perm ? in->hall_perm[c] : c;
            diag[p] = (orow == ocol);
            for (int l = 0; l < lmax; l++) {
                auto v0 = csr_value(in->H_Ap[2 * l], in->H_Ai[2 * l], in->H_Ax[2 * l], orow, ocol);
                auto v1 = csr_value(in->H_Ap[2 * l + 1], in->H_Ai[2 * l + 1], in->H_Ax[2 * l + 1], orow, ocol);
                double *cp = &coef[((size_t)p * lmax + l) * 4];
                cp[0] = v0.real(); cp[1] = v0.imag(); cp[2] = v1.real(); cp[3] = v1.imag();
            }
        }
    // coupling lists of phase A (one instance's rows): per row and multipole, the columns whose H0l or H1l
    // coefficient is non-zero, in the Hall row's (increasing column) order, each with both coefficients (h0,
    // h1; one of them may be 0)
    std::vector<int2> ack;
    std::vector<double> acf;
    m->h_aoff.assign((size_t)m->nblk * lmax + 1, 0);
    for (int r = 0; r < m->nblk; r++)
        for (int l = 0; l < lmax; l++) {
            m->h_aoff[(size_t)r * lmax + l] = (int)ack.size();
            const size_t first = ack.size();
            for (int p = m->hall_Ap[r]; p < m->hall_Ap[r + 1]; p++) {
                const double *cp = &coef[((size_t)p * lmax + l) * 4];
                if (cp[0] != 0.0 || cp[1] != 0.0 || cp[2] != 0.0 || cp[3] != 0.0) {
                    const int c = m->hall_Ai[p];
                    ack.push_back(make_int2(c, m->kap_blk[c]));
                    acf.insert(acf.end(), cp, cp + 4);
                }
            }
            while ((ack.size() - first) % kAngListU) {  // zero terms on the last column
                ack.push_back(ack.back());
                acf.insert(acf.end(), 4, 0.0);
            }
        }
    m->h_aoff.back() = (int)ack.size();
    if (ack.empty()) {  // no coupling: one unread entry keeps the uploads non-empty
        ack.push_back(make_int2(0, 0));
        acf.assign(4, 0.0);
    }
    // B instances: the Hall structure, kappa maps and angular coefficients repeat with row offsets
    m->N1 = (long long)m->nblk1 * n;
    if (B > 1) {
        const int nb1 = m->nblk1, hn1 = m->hnnz;
        std::vector<int> Ap((size_t)B * nb1 + 1), Ai((size_t)B * hn1), kb((size_t)B * nb1), bq((size_t)B * nb1);
        std::vector<double> cf((size_t)B * coef.size());
        std::vector<int> dg((size_t)B * hn1);
        for (int b = 0; b < B; b++) {
            for (int r = 0; r < nb1; r++) {
                Ap[(size_t)b * nb1 + r] = b * hn1 + m->hall_Ap[r];
                kb[(size_t)b * nb1 + r] = m->kap_blk[r];
                bq[(size_t)b * nb1 + r] = m->blk_q[r];
            }
            for (int p = 0; p < hn1; p++) {
                Ai[(size_t)b * hn1 + p] = b * nb1 + m->hall_Ai[p];
                dg[(size_t)b * hn1 + p] = diag[p];
            }
            std::copy(coef.begin(), coef.end(), cf.begin() + (size_t)b * coef.size());
        }
        Ap[(size_t)B * nb1] = B * hn1;
        m->hall_Ap.swap(Ap);
        m->hall_Ai.swap(Ai);
        m->kap_blk.swap(kb);
        m->blk_q.swap(bq);
        coef.swap(cf);
        diag.swap(dg);
        m->nblk = B * nb1;
        m->hnnz = B * hn1;
    }
    m->N = (long long)m->nblk * n;
    // uploads
    int rc;
    std::vector<double> gbuf((size_t)24 * lmax * m->nnzb * 2);
    for (int i = 0; i < 24 * lmax; i++) memcpy(&gbuf[(size_t)i * m->nnzb * 2], in->g[i], sizeof(double) * 2 * m->nnzb);
    std::vector<double> sbuf((size_t)3 * m->nnzb * 2), hbuf((size_t)4 * m->nnzb * 2);
    for (int i = 0; i < 3; i++) memcpy(&sbuf[(size_t)i * m->nnzb * 2], in->s[i], sizeof(double) * 2 * m->nnzb);
    for (int i = 0; i < 4; i++) memcpy(&hbuf[(size_t)i * m->nnzb * 2], in->h0[i], sizeof(double) * 2 * m->nnzb);
    if ((rc = upload(&m->d_ell_col, ell_col.data(), ell_col.size())) ||
        (rc = upload(&m->d_ell_lcol, ell_lcol.data(), ell_lcol.size())) ||
        (rc = upload(&m->d_locj, locj.data(), locj.size())) ||
        (rc = upload(&m->d_nloc, nloc.data(), nloc.size())) ||
        (rc = upload(&m->d_ell_src, ell_src.data(), ell_src.size())) ||
        (rc = upload(&m->d_ell_pos, ell_pos.data(), ell_pos.size())) ||
        (rc = upload(&m->d_tmap, tmap.data(), tmap.size())) ||
        (rc = upload(&m->d_g, (const double2 *)gbuf.data(), gbuf.size() / 2)) ||
        (rc = upload(&m->d_s, (const double2 *)sbuf.data(), sbuf.size() / 2)) ||
        (rc = upload(&m->d_h0, (const double2 *)hbuf.data(), hbuf.size() / 2)) ||
        (rc = upload(&m->d_hAp, m->hall_Ap.data(), m->hall_Ap.size())) ||
        (rc = upload(&m->d_hAi, m->hall_Ai.data(), m->hall_Ai.size())) ||
        (rc = upload(&m->d_kap, m->kap_blk.data(), m->kap_blk.size())) ||
        (rc = upload(&m->d_blk_q, m->blk_q.data(), m->blk_q.size())) ||
        (rc = upload(&m->d_coef, (const double2 *)coef.data(), coef.size() / 2)) ||
        (rc = upload(&m->d_aoff, m->h_aoff.data(), m->h_aoff.size())) ||
        (rc = upload(&m->d_ack, ack.data(), ack.size())) ||
        (rc = upload(&m->d_acf, (const double2 *)acf.data(), acf.size() / 2)) ||
        (rc = upload(&m->d_diag, diag.data(), diag.size())))
        return fail(rc);
    if ((rc = hip_alloc((void **)&m->d_GE, sizeof(double2) * (size_t)lmax * 8 * ellw * n)) ||
        (rc = hip_alloc((void **)&m->d_D, sizeof(double2) * (size_t)3 * m->nq * ellw * n)) ||
        (rc = hip_alloc((void **)&m->d_V, sizeof(double2) * (size_t)lmax * 4 * m->N)) ||
        (rc = hip_alloc((void **)&m->d_ypart, sizeof(double2) * (size_t)lmax * m->N)) ||
        (lmax > 1 && (rc = hip_alloc((void **)&m->d_ypart2, sizeof(double2) * (size_t)lmax * m->N))))
        return fail(rc);
    if (hipMemset(m->d_GE, 0, sizeof(double2) * (size_t)lmax * 8 * ellw * n) != hipSuccess)  // f = 0 until set_time
        return fail(GADE_EHIP);
    {
        int *d_kq = nullptr;
        if ((rc = upload(&d_kq, m->kap_q.data(), m->kap_q.size()))) return fail(rc);
        long long total = (long long)m->nq * ellw * n;
        hipLaunchKernelGGL(k_build_D, dim3(grid_for(total, 256)), dim3(256), 0, m->stream, m->d_s, m->d_h0,
                           m->d_ell_src, d_kq, m->nq, ellw, n, m->nnzb, h * dt / 2, m->d_D);
        hipError_t e = hipStreamSynchronize(m->stream);
        hipFree(d_kq);
        if (e != hipSuccess) {
            gade::set_error(std::string("k_build_D: ") + hipGetErrorString(e));
            return fail(GADE_EHIP);
        }
    }
    if ((rc = build_tiles(m))) return fail(rc);
    if ((rc = setup_dia(m))) return fail(rc);
    if ((rc = setup_mfma(m))) return fail(rc);
    m->radial_kind = m->dia_ok ? 2 : 1;
    if (B > 1 && !m->dia_ok) {
        gade::set_error("batched models need a banded radial pattern (2x2 bands of half-width <= 7)");
        return fail(GADE_EINVAL);
    }
    // BiCGSTAB workspace and scalars
    double2 **ws[] = {&m->r, &m->r0, &m->v, &m->p, &m->y, &m->z, &m->s, &m->t, &m->tmp};
    for (auto w : ws) {
        if ((rc = hip_alloc((void **)w, sizeof(double2) * m->N))) return fail(rc);
        if (hipMemset(*w, 0, sizeof(double2) * m->N) != hipSuccess) return fail(GADE_EHIP);
    }
    m->red_blocks = (int)std::min<long long>(1024, (m->N1 + kRedThreads - 1) / kRedThreads);
    const size_t nsc = (size_t)B * SC_STRIDE, nct = (size_t)B * CT_STRIDE;
    if ((rc = hip_alloc((void **)&m->d_sc, sizeof(double) * nsc)) ||
        (rc = hip_alloc((void **)&m->d_part, sizeof(double) * 4 * (size_t)B * m->red_blocks)) ||
        (rc = hip_alloc((void **)&m->d_part2, sizeof(double) * 4 * (size_t)B * m->red_blocks)) ||
        (rc = hip_alloc((void **)&m->d_flag, sizeof(int) * 4)) ||
        (rc = hip_alloc((void **)&m->d_ctl, sizeof(int) * nct)) ||
        (rc = hip_alloc((void **)&m->d_gctl, sizeof(int) * GC_COUNT)) ||
        (rc = hip_alloc((void **)&m->d_nsteps, sizeof(int) * B)))
        return fail(rc);
    if (hipMemset(m->d_sc, 0, sizeof(double) * nsc) != hipSuccess ||
        hipMemset(m->d_ctl, 0, sizeof(int) * nct) != hipSuccess || hipMemset(m->d_gctl, 0, sizeof(int) * GC_COUNT) != hipSuccess)
        return fail(GADE_EHIP);
    if (hipHostMalloc((void **)&m->h_ctl, sizeof(int) * nct, hipHostMallocDefault) != hipSuccess ||
        hipHostMalloc((void **)&m->h_gctl, sizeof(int) * GC_COUNT, hipHostMallocDefault) != hipSuccess ||
        hipHostMalloc((void **)&m->h_sc, sizeof(double) * nsc, hipHostMallocDefault) != hipSuccess) {
        gade::set_error("hipHostMalloc failed");
        return fail(GADE_ENOMEM);
    }
    m->ft.assign((size_t)12 * B, 0.0);
    *out = m;
    return GADE_OK;
}

int gade_model_destroy(gade_model_t *m) {
    if (!m) return GADE_OK;
    hipDeviceSynchronize();
    free_all(m);
    delete m;
    return GADE_OK;
}

int gade_model_info(const gade_model_t *m, int *nblk, int *blkdim, int *hall_nnz, int *nnz_b, long long *nnz_full,
                    int *nkappa) {
    if (!m) return GADE_EINVAL;
    if (nblk) *nblk = m->nblk1;  // blocks of one instance
    if (blkdim) *blkdim = m->n;
    if (hall_nnz) *hall_nnz = m->hnnz / m->B;
    if (nnz_b) *nnz_b = m->nnzb;
    if (nnz_full) *nnz_full = (long long)(m->hnnz / m->B) * m->nnzb;
    if (nkappa) *nkappa = m->nq;
    return GADE_OK;
}

int gade_set_time(gade_model_t *m, const double *ft) {
    if (!m || !ft) return GADE_EINVAL;
    for (int b = 0; b < m->B; b++) std::copy(ft, ft + 12, m->ft.begin() + (size_t)12 * b);
    return launch_gsums(m, m->radial_kind != 2);
}
int gade_set_time_batch(gade_model_t *m, const double *ft) {
    if (!m || !ft) return GADE_EINVAL;
    std::copy(ft, ft + (size_t)12 * m->B, m->ft.begin());
    return launch_gsums(m, m->radial_kind != 2);
}

static int apply_async(gade_model *m, const double2 *x, double2 *y, double cH, double beta) {
    int sel;
    if (cH == 1.0) sel = 0;
    else if (cH == -1.0) sel = 1;
    else if (cH == 0.0) sel = 2;
    else {
        gade::set_error("gade_apply: cH must be +1, -1 or 0");
        return GADE_EINVAL;
    }
    if (sel != 2) GADE_TRY(angular_phase(m, x));
    return launch_radial(m, x, y, sel, cH, beta);
}

// the CN pair from one angular pass: ym = (S - H) x (the rhs), yp = (S + H) x (A x), each equal bit for
// bit to its own apply_async; two applies when the radial phase is not banded
static int apply_pair_async(gade_model *m, const double2 *x, double2 *ym, double2 *yp) {
    if (m->radial_kind != 2) {
        GADE_TRY(apply_async(m, x, ym, -1.0, 0.0));
        return apply_async(m, x, yp, 1.0, 0.0);
    }
    GADE_TRY(angular_phase(m, x));
    return launch_radial(m, x, ym, 1, -1.0, 0.0, yp, 0);
}

int gade_apply(gade_model_t *m, const double *x, double *y, double cH, double beta) {
    if (!m || !x || !y) return GADE_EINVAL;
    GADE_TRY(apply_async(m, (const double2 *)x, (double2 *)y, cH, beta));
    GADE_HIP(hipStreamSynchronize(m->stream));
    return GADE_OK;
}

int gade_precond_setup(gade_model_t *m) {
    if (!m) return GADE_EINVAL;
    if (m->pre) return GADE_OK;
    const int n = m->n, nq = m->nq;
    // P(k) = S(k) + ihdt H0(k) per distinct kappa, from the device D operators (sel 0), in CSR order
    std::vector<double2> Dh((size_t)nq * m->ellw * n);
    GADE_HIP(hipMemcpy(Dh.data(), m->d_D, sizeof(double2) * Dh.size(), hipMemcpyDeviceToHost));
    std::vector<double2> Pq((size_t)nq * m->nnzb);
    for (int q = 0; q < nq; q++)
        for (int a = 0; a < n; a++)
            for (int k = m->rad_Ap[a]; k < m->rad_Ap[a + 1]; k++)
                Pq[(size_t)q * m->nnzb + k] = Dh[((size_t)q * m->ellw + (k - m->rad_Ap[a])) * n + a];
    return gade::blocklu_build(m->nblk, n, m->rad_Ap.data(), m->rad_Ai.data(), nq, m->blk_q.data(), Pq, m->stream,
                               &m->pre, m->psolve_ch);
}

}  // extern "C"

// ---------------------------------------------------------------------------
// preconditioner apply, BiCGSTAB, CN stepping
// ---------------------------------------------------------------------------
namespace {

int psolve_async(gade_model *m, const double2 *b, double2 *x) {
    Prof pr(m, "psolve", GADE_PROF_PSOLVE);
    return gade::blocklu_solve_async(m->pre, b, x, m->stream, m->gate, m->gstride, m->nblk1);
}

template <int OP>
int bicg_op(gade_model *m, const double2 *b, double2 *x = nullptr) {
    Prof pr(m, "blas1", GADE_PROF_BLAS1);
    int *ctl = m->ctl_on ? m->d_ctl : nullptr;
    hipLaunchKernelGGL(k_bicg<OP>, dim3(m->red_blocks, m->B), dim3(kRedThreads), 0, m->stream, m->N1, b, m->r, m->r0,
                       m->p, m->v, m->s, m->t, m->y, m->z, x, m->d_sc, m->d_part, m->gate, m->gstride);
    if (OP != OP_P && OP != OP_S)
        hipLaunchKernelGGL(k_bicg_fin<OP>, dim3(m->B), dim3(kRedThreads), 0, m->stream, m->red_blocks, m->d_part,
                           m->d_sc, m->gate, m->gstride, ctl);
    GADE_HIP(hipGetLastError());
    return GADE_OK;
}

// out = A x followed by the reduction OP (OP_ALPHA: out = v; OP_W: out = t), the radial phase's partial sum
// and the reduction in one kernel (partials into 'part'); the same results as apply_async then bicg_op<OP>,
// bit for bit.  Needs the banded radial phase with multipole partials (fused_sums()).  fin: finish the
// reduction in its own k_bicg_fin (else the next vector step folds it in).
bool fused_sums(const gade_model *m) { return m->radial_kind == 2 && m->lmax >= 2; }

template <int OP>
int apply_dot(gade_model *m, const double2 *x, double2 *out, double *part, bool fin) {
    GADE_TRY(angular_phase(m, x));
    {
        Prof pr(m, "radial", GADE_PROF_RADIAL);
        GADE_TRY((launch_dia64<false>(m, m->lmax, x, 0, 1.0, 0.0, out, m->d_ypart, nullptr, nullptr, 0)));
    }
    Prof pr(m, "radial_sum", GADE_PROF_RADIAL_SUM);
    hipLaunchKernelGGL(k_sum_bicg<OP>, dim3(m->red_blocks, m->B), dim3(kRedThreads), 0, m->stream, m->N1, m->N,
                       m->lmax, (const double2 *)m->d_ypart, out, (const double2 *)m->r0, (const double2 *)m->s,
                       part, m->gate, m->gstride);
    if (fin) {
        int *ctl = m->ctl_on ? m->d_ctl : nullptr;
        hipLaunchKernelGGL(k_bicg_fin<OP>, dim3(m->B), dim3(kRedThreads), 0, m->stream, m->red_blocks, part,
                           m->d_sc, m->gate, m->gstride, ctl);
    }
    GADE_HIP(hipGetLastError());
    return GADE_OK;
}

// OP_S / OP_XR with the previous reduction (alpha / w) finished inside (its partials in fold_part), then for
// OP_XR its own reduction and finish
template <int OP>
int bicg_op_fold(gade_model *m, const double *fold_part, double2 *x = nullptr) {
    Prof pr(m, "blas1", GADE_PROF_BLAS1);
    hipLaunchKernelGGL((k_bicg<OP, double2, true>), dim3(m->red_blocks, m->B), dim3(kRedThreads), 0, m->stream,
                       m->N1, (const double2 *)m->tmp, m->r, m->r0, m->p, m->v, m->s, m->t, m->y, m->z, x, m->d_sc,
                       m->d_part, m->gate, m->gstride, fold_part, m->red_blocks);
    if (OP == OP_XR) {
        int *ctl = m->ctl_on ? m->d_ctl : nullptr;
        hipLaunchKernelGGL(k_bicg_fin<OP>, dim3(m->B), dim3(kRedThreads), 0, m->stream, m->red_blocks, m->d_part,
                           m->d_sc, m->gate, m->gstride, ctl);
    }
    GADE_HIP(hipGetLastError());
    return GADE_OK;
}

int read_scalars(gade_model *m) {
    GADE_HIP(hipMemcpyAsync(m->h_sc, m->d_sc, sizeof(double) * SC_COUNT, hipMemcpyDeviceToHost, m->stream));
    GADE_HIP(hipStreamSynchronize(m->stream));
    return GADE_OK;
}

// A x into out:  out = (S + Hfull) x
int applyA(gade_model *m, const double2 *x, double2 *out, int *napply) {
    if (napply) (*napply)++;
    return apply_async(m, x, out, 1.0, 0.0);
}

// Host-driven BiCGSTAB iterations from iteration i, with r, r0, p, v, x and the
// scalars as they stand on the device (tools/gpu_bicgstab.c:133-307).
int bicg_host_loop(gade_model *m, const double2 *rhs, double2 *x, int i, double tol, int maxit, int *iters,
                   double *tol_error, int *napply) {
    GADE_TRY(read_scalars(m));
    double r0_sq = m->h_sc[SC_R0SQ];
    const double rhs_sq = m->h_sc[SC_RHSSQ];
    const double tol2 = tol * tol * rhs_sq;
    double rnorm2 = m->h_sc[SC_RNORM2];
    double rho_next_abs = std::hypot(m->h_sc[SC_RHONEXT], m->h_sc[SC_RHONEXT + 1]);
    if (!std::isfinite(rnorm2)) {
        gade::set_error("BiCGSTAB: non-finite residual");
        return GADE_ENOTCONV;
    }
    while (i < maxit) {
        if (rho_next_abs < kRestartEps2 * r0_sq) {
            // restart: r = rhs - A x, r0 = r, rho = |r|^2 (tools/gpu_bicgstab.c:222-241)
            GADE_TRY(applyA(m, x, m->r, napply));
            GADE_TRY(bicg_op<OP_RESTART>(m, rhs));
            GADE_TRY(read_scalars(m));
            r0_sq = m->h_sc[SC_R0SQ];
        }
        GADE_TRY(bicg_op<OP_P>(m, rhs));                 // p = r + beta (p - w v)
        GADE_TRY(psolve_async(m, m->p, m->y));           // y = M^-1 p
        GADE_TRY(applyA(m, m->y, m->v, napply));         // v = A y
        GADE_TRY(bicg_op<OP_ALPHA>(m, rhs));             // alpha = rho / (r0 . v)
        GADE_TRY(bicg_op<OP_S>(m, rhs));                 // s = r - alpha v
        GADE_TRY(psolve_async(m, m->s, m->z));           // z = M^-1 s
        GADE_TRY(applyA(m, m->z, m->t, napply));         // t = A z
        GADE_TRY(bicg_op<OP_W>(m, rhs));                 // w = (t . s) / |t|^2
        GADE_TRY(bicg_op<OP_XR>(m, rhs, x));             // x += alpha y + w z ; r = s - w t
        GADE_TRY(read_scalars(m));
        rnorm2 = m->h_sc[SC_RNORM2];
        rho_next_abs = std::hypot(m->h_sc[SC_RHONEXT], m->h_sc[SC_RHONEXT + 1]);
        if (!std::isfinite(rnorm2)) {
            gade::set_error("BiCGSTAB: non-finite residual");
            return GADE_ENOTCONV;
        }
        if (rnorm2 < tol2) break;
        ++i;
    }
    *tol_error = std::sqrt(rnorm2 / rhs_sq);
    *iters = i;
    return GADE_OK;
}

int bicgstab_impl(gade_model *m, const double2 *rhs, double2 *x, int *iters, double *tol_error, int *napply) {
    if (m->B != 1) {
        gade::set_error("the host-driven BiCGSTAB serves one instance; batched models step with gade_cn_run_batch");
        return GADE_EINVAL;
    }
    const double tol = *tol_error;
    const int maxit = *iters;
    GADE_TRY(gade_precond_setup(m));
    if (napply) *napply = 0;
    // r = rhs - A x ; r0 = r
    GADE_TRY(applyA(m, x, m->r, napply));
    GADE_TRY(bicg_op<OP_INIT>(m, rhs));
    GADE_TRY(read_scalars(m));
    const double r0_sq = m->h_sc[SC_R0SQ];
    const double rhs_sq = m->h_sc[SC_RHSSQ];
    if (rhs_sq == 0) {
        GADE_HIP(hipMemsetAsync(x, 0, sizeof(double2) * m->N, m->stream));
        GADE_HIP(hipStreamSynchronize(m->stream));
        *iters = 0;
        *tol_error = 0;
        return GADE_OK;
    }
    if (r0_sq < tol * tol * rhs_sq) {
        *iters = 0;
        *tol_error = std::sqrt(r0_sq / rhs_sq);
        return GADE_OK;
    }
    return bicg_host_loop(m, rhs, x, 0, tol, maxit, iters, tol_error, napply);
}

// ---------------------------------------------------------------------------
// device-driven CN stepping
//
// A step is enqueued for every instance as f(t) -> G-sums, rhs = (S - H) psi,
// [x0 by extrapolation], r = rhs - A x0 and K BiCGSTAB iterations whose
// launches are gated on each instance's control words: the reductions take
// the host loop's decisions (converged, maxit, restart, non-finite, zero rhs)
// on the device, and once an instance's solve is decided its remaining
// launches return at once.  k_step_end records iterations and tolerance, or
// marks the run stalled when an instance needs more than K iterations or the
// host (restart, non-finite, zero rhs); every later launch is then void until
// the host has finished that step: it restarts or zeroes the instances that
// need it and enqueues more iterations.  The host synchronizes once per window
// of steps.
// ---------------------------------------------------------------------------
__global__ void k_hist_reset(int *ctl0) {
    if (threadIdx.x == 0) ctl0[(size_t)blockIdx.x * CT_STRIDE + CT_HIST] = 0;
}

__global__ void k_ctl_set(int *ctl0, int *gctl, double *sc0, double tol, int maxit) {
    if (threadIdx.x != 0) return;
    int *ctl = ctl0 + (size_t)blockIdx.x * CT_STRIDE;
    for (int k = 0; k < CT_STRIDE; k++)
        if (k != CT_HIST) ctl[k] = 0;
    ctl[CT_MAXIT] = maxit;
    ctl[CT_HSEL] = 1;
    sc0[(size_t)blockIdx.x * SC_STRIDE + SC_TOL] = tol;
    if (blockIdx.x == 0)
        for (int k = 0; k < GC_COUNT; k++) gctl[k] = 0;
}

// x0 by extrapolation of each instance's history (order m->guess_order); no-op for order 0
int launch_guess(gade_model *m, double2 *psi) {
    if (m->guess_order == GADE_GUESS_LEARNED || (m->lg_features && m->guess_order > 0))
        return gade::lg_step_prologue(m, psi);
    if (m->guess_order <= 0) return GADE_OK;
    if (!m->d_hist) GADE_TRY(hip_alloc((void **)&m->d_hist, sizeof(double2) * kHistSlots * m->N));
    hipLaunchKernelGGL(k_extrap, dim3(grid_for(m->N1, 256, 2048), m->B), dim3(256), 0, m->stream, m->N1, m->N, psi,
                       m->d_hist, m->guess_order, m->d_ctl, m->gate, m->gstride);
    hipLaunchKernelGGL(k_hist_count, dim3(m->B), dim3(64), 0, m->stream, m->d_ctl, m->gate, m->gstride);
    GADE_HIP(hipGetLastError());
    return GADE_OK;
}

struct GateScope {
    gade_model *m;
    explicit GateScope(gade_model *mm) : m(mm) {
        m->ctl_on = true;
        m->gstride = CT_STRIDE;
    }
    ~GateScope() {
        m->gate = nullptr;
        m->gstride = 0;
        m->ctl_on = false;
    }
};

}  // namespace
namespace gade {
// the step graph holds device addresses of the model's buffers: drop it before any of them is replaced
void drop_step_graph(gade_model *m) {
    // a replay may still be running on the model's stream: let it finish before the executable goes
    if (m->graph_exec && m->stream) (void)hipStreamSynchronize(m->stream);
    if (m->graph_exec) (void)hipGraphExecDestroy(m->graph_exec);
    if (m->graph) (void)hipGraphDestroy(m->graph);
    m->graph_exec = nullptr;
    m->graph = nullptr;
    memset(m->graph_key, 0, sizeof(m->graph_key));
}
}  // namespace gade
using gade::drop_step_graph;
namespace {

int ensure_step_buffers(gade_model *m, int nsteps) {
    if (nsteps <= m->steps_cap) return GADE_OK;
    drop_step_graph(m);  // it writes the per-step outputs being replaced
    if (m->d_iters) (void)hipFree(m->d_iters);
    if (m->d_tols) (void)hipFree(m->d_tols);
    m->d_iters = nullptr;
    m->d_tols = nullptr;
    m->steps_cap = 0;
    GADE_TRY(hip_alloc((void **)&m->d_iters, sizeof(int) * (size_t)m->B * nsteps));
    GADE_TRY(hip_alloc((void **)&m->d_tols, sizeof(double) * (size_t)m->B * nsteps));
    m->steps_cap = nsteps;
    return GADE_OK;
}

// K gated BiCGSTAB iterations on the current step of every instance
int enqueue_iterations(gade_model *m, double2 *psi, int K) {
    m->gate = m->d_ctl + CT_SKIP;
    const bool fused = fused_sums(m);
    for (int k = 0; k < K; k++) {
        GADE_TRY(bicg_op<OP_P>(m, m->tmp));
        GADE_TRY(psolve_async(m, m->p, m->y));
        if (fused) {
            // the radial sums carry the dots; alpha and w are finished by the vector steps that use them (w's
            // partials in their own buffer: x/r's reduction reuses d_part while other blocks still read them)
            GADE_TRY(apply_dot<OP_ALPHA>(m, m->y, m->v, m->d_part, false));
            GADE_TRY(bicg_op_fold<OP_S>(m, m->d_part));
            GADE_TRY(psolve_async(m, m->s, m->z));
            GADE_TRY(apply_dot<OP_W>(m, m->z, m->t, m->d_part2, false));
            GADE_TRY(bicg_op_fold<OP_XR>(m, m->d_part2, psi));
        } else {
            GADE_TRY(apply_async(m, m->y, m->v, 1.0, 0.0));
            GADE_TRY(bicg_op<OP_ALPHA>(m, m->tmp));
            GADE_TRY(bicg_op<OP_S>(m, m->tmp));
            GADE_TRY(psolve_async(m, m->s, m->z));
            GADE_TRY(apply_async(m, m->z, m->t, 1.0, 0.0));
            GADE_TRY(bicg_op<OP_W>(m, m->tmp));
            GADE_TRY(bicg_op<OP_XR>(m, m->tmp, psi));
        }
    }
    return GADE_OK;
}

// ---------------------------------------------------------------------------
// mixed precision: FP32 copies of the operator and preconditioner, FP32 inner BiCGSTAB, FP64 defect
// correction (residual, update, tolerance test and every accumulator stay in FP64)
// ---------------------------------------------------------------------------
int mp_setup(gade_model *m) {
    if (m->mp_ready && m->lu32[0]) return GADE_OK;
    if (m->radial_kind != 2 || !m->pre || !((m->pre->ch == 16 || m->pre->ch == 32) && m->pre->band <= 16)) {
        gade::set_error("mixed precision needs the banded radial kernel and 16- or 32-row preconditioner levels");
        return GADE_EINVAL;
    }
    const size_t nlu = (size_t)m->nq * m->pre->nch * m->pre->ch * m->pre->ch;
    auto conv = [&](const double2 *src, float2 *dst, size_t n) {
        hipLaunchKernelGGL(k_to_f32, dim3(grid_for((long long)n, 256, 8192), 1), dim3(256), 0, m->stream,
                           (long long)n, (size_t)0, src, dst, (const int *)nullptr, 0);
    };
    const size_t ntw = (size_t)m->nq * m->pre->nch * 4 * 256;
    if (m->mp_ready) {
        // the preconditioner was rebuilt: its FP32 copies only
        for (auto &f : m->lu32) GADE_TRY(hip_alloc((void **)&f, sizeof(float2) * nlu));
        conv(m->pre->d_Linv, m->lu32[0], nlu);
        conv(m->pre->d_Uinv, m->lu32[1], nlu);
        conv(m->pre->d_Loff, m->lu32[2], nlu);
        conv(m->pre->d_Uoff, m->lu32[3], nlu);
        if (m->pre->d_TW) {
            GADE_TRY(hip_alloc((void **)&m->tw32, sizeof(float2) * ntw));
            conv(m->pre->d_TW, m->tw32, ntw);
        }
        GADE_HIP(hipGetLastError());
        GADE_HIP(hipStreamSynchronize(m->stream));
        return GADE_OK;
    }
    const size_t N = (size_t)m->N, L = (size_t)m->lmax;
    const size_t S = (size_t)4 * m->nd * m->nhp;
    const size_t nDD = 3 * (size_t)m->nq * S, nco = (size_t)m->hnnz * L * 2;
    GADE_TRY(hip_alloc((void **)&m->d_V32, sizeof(float2) * L * 4 * N));
    GADE_TRY(hip_alloc((void **)&m->d_ypart32, sizeof(float2) * L * N));
    GADE_TRY(hip_alloc((void **)&m->d_GD32, sizeof(float2) * m->gd_stride * m->B));
    GADE_TRY(hip_alloc((void **)&m->d_DD32, sizeof(float2) * nDD));
    GADE_TRY(hip_alloc((void **)&m->d_coef32, sizeof(float2) * nco));
    const size_t nacf = 2 * (size_t)m->h_aoff.back();  // (h0, h1) per list entry
    GADE_TRY(hip_alloc((void **)&m->d_acf32, sizeof(float2) * std::max<size_t>(nacf, 1)));
    for (auto &f : m->lu32) GADE_TRY(hip_alloc((void **)&f, sizeof(float2) * nlu));
    for (float2 **v : {&m->mp_r, &m->mp_r0, &m->mp_p, &m->mp_v, &m->mp_y, &m->mp_z, &m->mp_s, &m->mp_t, &m->mp_d})
        GADE_TRY(hip_alloc((void **)v, sizeof(float2) * N));
    GADE_TRY(hip_alloc((void **)&m->d_mp_ax, sizeof(double2) * N));
    GADE_TRY(hip_alloc((void **)&m->d_sc32, sizeof(double) * (size_t)m->B * SC_STRIDE));
    conv(m->d_DD, m->d_DD32, nDD);
    conv(m->d_coef, m->d_coef32, nco);
    if (nacf) conv(m->d_acf, m->d_acf32, nacf);
    conv(m->pre->d_Linv, m->lu32[0], nlu);
    conv(m->pre->d_Uinv, m->lu32[1], nlu);
    conv(m->pre->d_Loff, m->lu32[2], nlu);
    conv(m->pre->d_Uoff, m->lu32[3], nlu);
    if (m->pre->d_TW) {
        GADE_TRY(hip_alloc((void **)&m->tw32, sizeof(float2) * ntw));
        conv(m->pre->d_TW, m->tw32, ntw);
    }
    GADE_HIP(hipGetLastError());
    GADE_HIP(hipStreamSynchronize(m->stream));
    m->mp_ready = true;
    return GADE_OK;
}

// y = (S + H) x in FP32 (the FP32 G-sums of the current step)
int apply32_async(gade_model *m, const float2 *x, float2 *y) {
    if (!ang_in_radial<float2>(m)) {
        Prof pr(m, "angular_f32", GADE_PROF_ANGULAR);
        GADE_TRY(launch_angular_t<float2>(m, x, (const float2 *)m->d_coef32, (const float2 *)m->d_acf32, m->d_V32));
    }
    const int lgrid = m->lmax;
    float2 *yp = lgrid > 1 ? m->d_ypart32 : nullptr;
    {
        Prof pr(m, "radial_f32", GADE_PROF_RADIAL);
        GADE_TRY((launch_dia<false, float2>(m, lgrid, x, (const float2 *)m->d_V32, (const float2 *)m->d_GD32,
                                            (const float2 *)m->d_DD32, 0, 1.0, 0.0, y, yp, (float2 *)nullptr,
                                            (float2 *)nullptr, 0)));
    }
    if (yp) {
        Prof pr(m, "radial_sum_f32", GADE_PROF_RADIAL_SUM);
        hipLaunchKernelGGL(k_radial_sum<float2>, dim3(grid_for(m->N1, 256, 2048), m->B), dim3(256), 0, m->stream,
                           yp, lgrid, m->N, 0.0, y, m->gate, m->gstride, m->N1);
        GADE_HIP(hipGetLastError());
    }
    return GADE_OK;
}

int psolve32_async(gade_model *m, const float2 *b, float2 *x) {
    Prof pr(m, "psolve_f32", GADE_PROF_PSOLVE);
    const gade_blocklu *lu = m->pre;
    if (lu->d_TW) {
        hipLaunchKernelGGL((k_psolve_tw<4, float2>), dim3(lu->norder), dim3(128), sizeof(float2) * (size_t)lu->nch * 16,
                           m->stream, b, x, lu->d_order, lu->d_blk_q, lu->ilv, (const float2 *)m->tw32, lu->n,
                           lu->nch, lu->tw_m, m->gate, m->gstride, m->nblk1);
        GADE_HIP(hipGetLastError());
        return GADE_OK;
    }
    // 32-row levels (mixed precision needs 16- or 32-row levels and a band of at most 16)
    const size_t shq = sizeof(float2) * ((size_t)lu->nch * lu->ch + lu->ch);
    hipLaunchKernelGGL((k_psolve_q<2, float2, 32, 2>), dim3(lu->norder), dim3(64), shq, m->stream, b, x,
                       lu->d_order, lu->d_blk_q, lu->ilv, m->lu32[0], m->lu32[1], m->lu32[2], m->lu32[3], lu->n,
                       lu->band, lu->nch, m->gate, m->gstride, m->nblk1);
    GADE_HIP(hipGetLastError());
    return GADE_OK;
}

template <int OP>
int bicg32_op(gade_model *m) {
    Prof pr(m, "blas1_f32", GADE_PROF_BLAS1);
    hipLaunchKernelGGL((k_bicg<OP, float2>), dim3(m->red_blocks, m->B), dim3(kRedThreads), 0, m->stream, m->N1,
                       (const float2 *)nullptr, m->mp_r, m->mp_r0, m->mp_p, m->mp_v, m->mp_s, m->mp_t, m->mp_y,
                       m->mp_z, m->mp_d, m->d_sc32, m->d_part, m->gate, m->gstride);
    if (OP != OP_P && OP != OP_S)
        hipLaunchKernelGGL((k_bicg_fin<OP, true>), dim3(m->B), dim3(kRedThreads), 0, m->stream, m->red_blocks,
                           m->d_part, m->d_sc32, m->gate, m->gstride, m->d_ctl);
    GADE_HIP(hipGetLastError());
    return GADE_OK;
}

// J defect-correction sweeps of every instance still solving (gated per instance)
int enqueue_sweeps(gade_model *m, double2 *psi, int J) {
    const dim3 red(m->red_blocks, m->B);
    for (int j = 0; j < J; j++) {
        m->gate = m->d_ctl + CT_SKIP;
        hipLaunchKernelGGL(k_mp_start, red, dim3(kRedThreads), 0, m->stream, m->N1, (const double2 *)m->r, m->mp_r,
                           m->mp_r0, m->mp_d, m->mp_p, m->mp_v, m->d_part, m->gate, m->gstride);
        hipLaunchKernelGGL(k_mp_start_fin, dim3(m->B), dim3(kRedThreads), 0, m->stream, m->red_blocks, m->d_part,
                           m->d_sc32, m->d_ctl, m->mp_inner_tol);
        GADE_HIP(hipGetLastError());
        m->gate = m->d_ctl + CT_ISKIP;
        for (int k = 0; k < m->mp_inner_iters; k++) {
            GADE_TRY(bicg32_op<OP_P>(m));
            GADE_TRY(psolve32_async(m, m->mp_p, m->mp_y));
            GADE_TRY(apply32_async(m, m->mp_y, m->mp_v));
            GADE_TRY(bicg32_op<OP_ALPHA>(m));
            GADE_TRY(bicg32_op<OP_S>(m));
            GADE_TRY(psolve32_async(m, m->mp_s, m->mp_z));
            GADE_TRY(apply32_async(m, m->mp_z, m->mp_t));
            GADE_TRY(bicg32_op<OP_W>(m));
            GADE_TRY(bicg32_op<OP_XR>(m));
        }
        m->gate = m->d_ctl + CT_SKIP;
        hipLaunchKernelGGL(k_mp_update, dim3(grid_for(m->N1, 256, 2048), m->B), dim3(256), 0, m->stream, m->N1, psi,
                           (const float2 *)m->mp_d, m->gate, m->gstride);
        GADE_TRY(apply_async(m, psi, m->d_mp_ax, 1.0, 0.0));
        hipLaunchKernelGGL(k_mp_resid, red, dim3(kRedThreads), 0, m->stream, m->N1, (const double2 *)m->tmp,
                           (const double2 *)m->d_mp_ax, m->r, m->d_part, m->gate, m->gstride);
        hipLaunchKernelGGL(k_mp_resid_fin, dim3(m->B), dim3(kRedThreads), 0, m->stream, m->red_blocks, m->d_part,
                           m->d_sc, m->d_ctl, m->gate, m->gstride);
        GADE_HIP(hipGetLastError());
    }
    return GADE_OK;
}

// the FP32 G-sums of the step (after the FP64 ones), per instance under the step-prologue gate
int mp_convert_gsums(gade_model *m) {
    hipLaunchKernelGGL(k_to_f32, dim3(grid_for((long long)m->gd_stride, 256, 4096), m->B), dim3(256), 0, m->stream,
                       (long long)m->gd_stride, m->gd_stride, (const double2 *)m->d_GD, m->d_GD32, m->gate,
                       m->gstride);
    GADE_HIP(hipGetLastError());
    return GADE_OK;
}

// the solve part of a device-driven step after OP_INIT: FP64 iterations or mixed sweeps
int enqueue_solve(gade_model *m, double2 *psi, int K) {
    return m->precision == 1 ? enqueue_sweeps(m, psi, K) : enqueue_iterations(m, psi, K);
}

// rhs = (S - H) psi into tmp (tools/dirac_solver.c:531-536), x0 into psi (psi_n unless extrapolating),
// A x0 into r; with x0 = psi_n the rhs and A x0 come from one paired apply
int enqueue_rhs_ax0(gade_model *m, double2 *psi) {
    if (m->guess_order == 0) return apply_pair_async(m, psi, m->tmp, m->r);
    GADE_TRY(apply_async(m, psi, m->tmp, -1.0, 0.0));
    GADE_TRY(launch_guess(m, psi));
    return apply_async(m, psi, m->r, 1.0, 0.0);
}

// the rhs, A x0 and the start of the solve (OP_INIT); with x0 = psi and multipole partials, the paired
// apply's sums and OP_INIT are one kernel (the same results)
int enqueue_rhs_init(gade_model *m, double2 *psi) {
    if (m->guess_order != 0 || !fused_sums(m)) {
        GADE_TRY(enqueue_rhs_ax0(m, psi));
        return bicg_op<OP_INIT>(m, m->tmp);
    }
    GADE_TRY(angular_phase(m, psi));
    {
        Prof pr(m, "radial_pair", GADE_PROF_RADIAL);
        GADE_TRY((launch_dia64<true>(m, m->lmax, psi, 1, -1.0, 0.0, m->tmp, m->d_ypart, m->r, m->d_ypart2, 0)));
    }
    Prof pr(m, "radial_sum", GADE_PROF_RADIAL_SUM);
    hipLaunchKernelGGL(k_sum2_init, dim3(m->red_blocks, m->B), dim3(kRedThreads), 0, m->stream, m->N1, m->N, m->lmax,
                       (const double2 *)m->d_ypart, (const double2 *)m->d_ypart2, m->tmp, m->r, m->r0, m->d_part,
                       m->gate, m->gstride);
    int *ctl = m->ctl_on ? m->d_ctl : nullptr;
    hipLaunchKernelGGL(k_bicg_fin<OP_INIT>, dim3(m->B), dim3(kRedThreads), 0, m->stream, m->red_blocks, m->d_part,
                       m->d_sc, m->gate, m->gstride, ctl);
    GADE_HIP(hipGetLastError());
    return GADE_OK;
}

int enqueue_step_end(gade_model *m, int step) {
    hipLaunchKernelGGL(k_step_end, dim3(m->B), dim3(64), 0, m->stream, m->d_ctl, m->d_gctl, m->d_sc, step,
                       m->steps_cap, m->d_iters, m->d_tols);
    hipLaunchKernelGGL(k_stall_spread, dim3(m->B), dim3(64), 0, m->stream, m->d_ctl, m->d_gctl);
    GADE_HIP(hipGetLastError());
    return GADE_OK;
}

// ft: [B][12] pulse coefficients of this step
int enqueue_step(gade_model *m, double2 *psi, const double *ft, int step, int K, bool per_instance_steps) {
    GateScope gs(m);
    std::copy(ft, ft + (size_t)12 * m->B, m->ft.begin());
    hipLaunchKernelGGL(k_step_begin, dim3(m->B), dim3(64), 0, m->stream, m->d_ctl, m->d_gctl, step,
                       per_instance_steps ? m->d_nsteps : nullptr);
    GADE_HIP(hipGetLastError());
    m->gate = m->d_ctl + CT_PSKIP;
    GADE_TRY(launch_gsums(m, m->radial_kind != 2));
    if (m->precision == 1) GADE_TRY(mp_convert_gsums(m));
    GADE_TRY(enqueue_rhs_init(m, psi));
    GADE_TRY(enqueue_solve(m, psi, K));
    return enqueue_step_end(m, step);
}

// the G-sums of the current window slot from the device coefficient table
int launch_gsums_window(gade_model *m) {
    Prof pr(m, "gsums", GADE_PROF_GSUMS);
    long long total = (long long)m->lmax * 16 * m->nd * m->nhp;  // (l, v, qd, d, ip)
    CoefGroup cg{};
    if (dia_gsums_needed(m))
        hipLaunchKernelGGL(k_gsums_dia, dim3(grid_for(total, 256, 4096), m->B), dim3(256), 0, m->stream, m->d_gs,
                           m->d_gts, cg, 0, m->gd_stride, m->lmax, m->nd, m->nhp, m->d_GD, m->gate, m->gstride,
                           (const Coef6 *)m->d_cfwin, m->d_gctl, m->B);
    if (m->radial_mfma) {
        const long long totf = (long long)m->lmax * 16 * m->nmi * 5 * kMfmaFrag;
        hipLaunchKernelGGL(k_gsums_frag, dim3(grid_for(totf, 256, 4096), m->B), dim3(256), 0, m->stream, m->d_gs,
                           m->d_gts, cg, 0, m->gf_stride, m->lmax, m->nd, m->nhp, m->nmi, m->d_GF, m->gate,
                           m->gstride, (const Coef6 *)m->d_cfwin, m->d_gctl, m->B);
    }
    GADE_HIP(hipGetLastError());
    return GADE_OK;
}

// One CN step whose per-step inputs are on the device (the window's coefficient table and the run-wide
// step words), so the same launches serve every step: the body of the step graph.
int enqueue_step_window(gade_model *m, double2 *psi, int K, bool per_instance_steps) {
    GateScope gs(m);
    hipLaunchKernelGGL(k_step_begin, dim3(m->B), dim3(64), 0, m->stream, m->d_ctl, m->d_gctl, -1,
                       per_instance_steps ? m->d_nsteps : nullptr);
    GADE_HIP(hipGetLastError());
    m->gate = m->d_ctl + CT_PSKIP;
    GADE_TRY(launch_gsums_window(m));
    if (m->precision == 1) GADE_TRY(mp_convert_gsums(m));
    GADE_TRY(enqueue_rhs_init(m, psi));
    GADE_TRY(enqueue_solve(m, psi, K));
    hipLaunchKernelGGL(k_step_end, dim3(m->B), dim3(64), 0, m->stream, m->d_ctl, m->d_gctl, m->d_sc, -1,
                       m->steps_cap, m->d_iters, m->d_tols);
    hipLaunchKernelGGL(k_stall_spread, dim3(m->B), dim3(64), 0, m->stream, m->d_ctl, m->d_gctl);
    hipLaunchKernelGGL(k_step_advance, dim3(1), dim3(64), 0, m->stream, m->d_gctl);
    GADE_HIP(hipGetLastError());
    return GADE_OK;
}

// the step graph for this K (captured again when anything it bakes in has changed)
int step_graph(gade_model *m, double2 *psi, int K, bool per_instance_steps) {
    if (m->guess_order > 0 && !m->d_hist)  // no allocation while capturing
        GADE_TRY(hip_alloc((void **)&m->d_hist, sizeof(double2) * kHistSlots * m->N));
    if (m->guess_order == GADE_GUESS_LEARNED || m->lg_features) GADE_TRY(gade::lg_prepare(m));
    // everything the captured launches bake in that can change between runs (buffers that are replaced
    // on growth also drop the graph themselves)
    intptr_t tol_bits = 0;
    memcpy(&tol_bits, &m->mp_inner_tol, sizeof(double) < sizeof(intptr_t) ? sizeof(double) : sizeof(intptr_t));
    const void *key[12] = {psi,      (const void *)(intptr_t)K, (const void *)(intptr_t)m->guess_order,
                           (const void *)(intptr_t)(per_instance_steps ? 1 : 0), m->d_iters,
                           (const void *)(intptr_t)m->steps_cap, m->d_cfwin, m->d_hist,
                           (const void *)(intptr_t)m->precision, (const void *)(intptr_t)m->mp_inner_iters,
                           (const void *)tol_bits, (const void *)(intptr_t)m->lg_features};
    static_assert(sizeof(key) == sizeof(m->graph_key), "graph key size");
    if (m->graph_exec && !memcmp(key, m->graph_key, sizeof(key))) return GADE_OK;
    drop_step_graph(m);
    if (!m->cap_stream) GADE_HIP(hipStreamCreateWithFlags(&m->cap_stream, hipStreamNonBlocking));
    hipStream_t saved = m->stream;
    m->stream = m->cap_stream;
    GADE_HIP(hipStreamBeginCapture(m->cap_stream, hipStreamCaptureModeThreadLocal));
    int rc = enqueue_step_window(m, psi, K, per_instance_steps);
    hipGraph_t g = nullptr;
    hipError_t e = hipStreamEndCapture(m->cap_stream, &g);
    m->stream = saved;
    GADE_TRY(rc);
    GADE_HIP(e);
    m->graph = g;
    GADE_HIP(hipGraphInstantiate(&m->graph_exec, m->graph, nullptr, nullptr, 0));
    memcpy(m->graph_key, key, sizeof(key));
    return GADE_OK;
}

int ensure_cfwin(gade_model *m, int slots) {
    if (slots <= m->cfwin_cap) return GADE_OK;
    drop_step_graph(m);  // it reads the table being replaced
    if (m->d_cfwin) (void)hipFree(m->d_cfwin);
    if (m->h_cfwin) (void)hipHostFree(m->h_cfwin);
    m->d_cfwin = m->h_cfwin = nullptr;
    m->cfwin_cap = 0;
    const size_t bytes = sizeof(Coef6) * (size_t)slots * m->B;
    GADE_TRY(hip_alloc((void **)&m->d_cfwin, bytes));
    GADE_HIP(hipHostMalloc((void **)&m->h_cfwin, bytes, hipHostMallocDefault));
    m->cfwin_cap = slots;
    return GADE_OK;
}

// The run stalled at 'step': finish it for every instance from the device state.
// it_h/tol_h: [B][steps_cap] host copies of the per-step outputs.
int finish_stalled_step(gade_model *m, double2 *psi, int step, int K, std::vector<int> &it_h,
                        std::vector<double> &tol_h) {
    const int B = m->B;
    for (int round = 0;; round++) {
        GADE_HIP(hipMemcpyAsync(m->h_ctl, m->d_ctl, sizeof(int) * B * CT_STRIDE, hipMemcpyDeviceToHost, m->stream));
        GADE_HIP(hipMemcpyAsync(m->h_sc, m->d_sc, sizeof(double) * B * SC_STRIDE, hipMemcpyDeviceToHost, m->stream));
        GADE_HIP(hipStreamSynchronize(m->stream));
        // instances idle at this step (pulse over) keep PSKIP set by k_step_begin; k_stall_spread set it for
        // the others, so activity is read from the step's status instead
        bool more = false, restart = false;
        for (int b = 0; b < B; b++) {
            int *c = m->h_ctl + (size_t)b * CT_STRIDE;
            const double *sc = m->h_sc + (size_t)b * SC_STRIDE;
            const int st = c[CT_STATUS];
            c[CT_HSEL] = 1;
            c[CT_PSKIP] = 1;
            if (st == 2) {
                if (sc[SC_RHSSQ] == 0) {
                    // zero rhs: x = 0 (tools/gpu_bicgstab.c:183-189)
                    GADE_HIP(hipMemsetAsync(psi + (size_t)b * m->N1, 0, sizeof(double2) * m->N1, m->stream));
                    c[CT_STATUS] = 1;
                    c[CT_ITER] = 0;
                } else if (!std::isfinite(sc[SC_RNORM2])) {
                    gade::set_error("BiCGSTAB: non-finite residual");
                    return GADE_ENOTCONV;
                } else if (m->precision == 1) {
                    c[CT_STATUS] = 0;  // no restart in the defect correction: another sweep
                    more = true;
                } else {
                    // restart: r = rhs - A x, r0 = r, rho = |r|^2 (tools/gpu_bicgstab.c:222-241)
                    c[CT_HSEL] = 0;
                    c[CT_STATUS] = 0;
                    restart = more = true;
                }
            } else if (st == 0) {
                more = true;
            }
            c[CT_SKIP] = c[CT_STATUS] != 0;
        }
        if (round > 4096) {
            gade::set_error("device-driven solve does not finish");
            return GADE_ENOTCONV;
        }
        // the decisions above go back to the device in every round, the last one included, and the
        // run is no longer stalled
        m->h_gctl[GC_STALLED] = 0;
        m->h_gctl[GC_STALL_STEP] = 0;
        GADE_HIP(hipMemcpyAsync(m->d_ctl, m->h_ctl, sizeof(int) * B * CT_STRIDE, hipMemcpyHostToDevice, m->stream));
        GADE_HIP(hipMemcpyAsync(m->d_gctl, m->h_gctl, sizeof(int) * GC_COUNT, hipMemcpyHostToDevice, m->stream));
        if (more) {
            GateScope gs(m);
            if (restart) {
                m->gate = m->d_ctl + CT_HSEL;
                GADE_TRY(apply_async(m, psi, m->r, 1.0, 0.0));
                GADE_TRY(bicg_op<OP_RESTART>(m, m->tmp));
            }
            GADE_TRY(enqueue_solve(m, psi, K));
        }
        // record every decided solve of this step (zero-rhs instances with 0 iterations, tolerance 0)
        hipLaunchKernelGGL(k_step_end_resume, dim3(B), dim3(64), 0, m->stream, m->d_ctl, m->d_gctl, m->d_sc, step,
                           m->steps_cap, m->d_iters, m->d_tols);
        GADE_HIP(hipGetLastError());
        GADE_HIP(hipMemcpyAsync(it_h.data(), m->d_iters, sizeof(int) * B * m->steps_cap, hipMemcpyDeviceToHost,
                                m->stream));
        GADE_HIP(hipMemcpyAsync(tol_h.data(), m->d_tols, sizeof(double) * B * m->steps_cap, hipMemcpyDeviceToHost,
                                m->stream));
        if (!more) {
            GADE_HIP(hipStreamSynchronize(m->stream));
            break;
        }
    }
    // normal stepping resumes with the next step: k_step_begin recomputes every instance's gates
    return GADE_OK;
}

// Steps 0..nsteps-1 for every instance at the given times; params [B][4] = (E0, omega, cycles, cep) of
// each instance's pulse; nsteps_b [B] (null: every instance runs all steps).  psi [B][N1] is x0 and
// result.  iters/tol_out [B][nsteps] (entries past an instance's pulse are left untouched).
int cn_run_device(gade_model *m, double2 *psi, const double *times, int nsteps, const double *params,
                  const int *nsteps_b, double tol, int maxit, int *iters, double *tol_out, double *usec) {
    if (nsteps == 0) return GADE_OK;
    const int B = m->B;
    GADE_TRY(gade_precond_setup(m));
    if (m->precision == 1) GADE_TRY(mp_setup(m));
    GADE_TRY(ensure_step_buffers(m, nsteps));
    if (m->guess_order == GADE_GUESS_LEARNED || m->lg_features) GADE_TRY(gade::lg_prepare(m));
    const size_t cap = (size_t)m->steps_cap;
    std::vector<int> it_h((size_t)B * cap, 0);
    std::vector<double> tol_h((size_t)B * cap, 0.0);
    GADE_HIP(hipMemsetAsync(m->d_iters, 0, sizeof(int) * B * cap, m->stream));
    GADE_HIP(hipMemsetAsync(m->d_tols, 0, sizeof(double) * B * cap, m->stream));
    if (nsteps_b) GADE_HIP(hipMemcpyAsync(m->d_nsteps, nsteps_b, sizeof(int) * B, hipMemcpyHostToDevice, m->stream));
    std::vector<hipEvent_t> ev;
    auto free_events = [&]() {
        for (hipEvent_t e : ev) (void)hipEventDestroy(e);
        ev.clear();
    };
    hipLaunchKernelGGL(k_ctl_set, dim3(B), dim3(64), 0, m->stream, m->d_ctl, m->d_gctl, m->d_sc, tol, maxit);
    GADE_HIP(hipGetLastError());
    std::vector<double> ft((size_t)12 * B);
    int s = 0;
    while (s < nsteps) {
        const int W = std::min(m->window, nsteps - s);
        // iterations (FP64) or sweeps (mixed) enqueued per step; more are handed to the host
        const int K = m->iter_batch > 0 ? m->iter_batch : (m->precision == 1 ? 2 : m->last_k);
        if (usec) {
            free_events();
            ev.resize(W + 1);
            for (auto &e : ev) GADE_HIP(hipEventCreate(&e));
            GADE_HIP(hipEventRecord(ev[0], m->stream));
        }
        // f(t) of every step of the window, [W][B][12]
        ft.resize((size_t)12 * B * W);
        for (int w = 0; w < W; w++)
            for (int b = 0; b < B; b++) {
                const double *pp = params + (size_t)4 * b;
                gade_pulse_cep(pp[0], pp[1], pp[2], pp[3], times[s + w], ft.data() + (size_t)12 * (w * B + b));
            }
        // DIA: the window's pulse coefficients go to a device table that the steps index by the
        // run-wide step words, so every step is the same launches (one graph); ELL: per-step launches
        const bool table = m->radial_kind == 2;
        if (table) {
            GADE_TRY(ensure_cfwin(m, W));
            Coef6 *tab = (Coef6 *)m->h_cfwin;
            for (int w = 0; w < W; w++) {
                std::copy(ft.data() + (size_t)12 * B * w, ft.data() + (size_t)12 * B * (w + 1), m->ft.begin());
                for (int b = 0; b < B; b++) tab[(size_t)w * B + b] = coef_of(m, b);
            }
            GADE_HIP(hipMemcpyAsync(m->d_cfwin, m->h_cfwin, sizeof(Coef6) * (size_t)W * B, hipMemcpyHostToDevice,
                                    m->stream));
            hipLaunchKernelGGL(k_window_begin, dim3(1), dim3(64), 0, m->stream, m->d_gctl, s);
            GADE_HIP(hipGetLastError());
        }
        const bool graph = table && m->use_graph && m->profile == 0;
        if (graph) GADE_TRY(step_graph(m, psi, K, nsteps_b != nullptr));
        for (int w = 0; w < W; w++) {
            int rc = graph   ? (hipGraphLaunch(m->graph_exec, m->stream) == hipSuccess ? GADE_OK : GADE_EHIP)
                     : table ? enqueue_step_window(m, psi, K, nsteps_b != nullptr)
                             : enqueue_step(m, psi, ft.data() + (size_t)12 * B * w, s + w, K, nsteps_b != nullptr);
            if (rc == GADE_OK && usec) rc = hipEventRecord(ev[w + 1], m->stream) == hipSuccess ? GADE_OK : GADE_EHIP;
            if (rc) {
                if (graph && rc == GADE_EHIP) gade::set_error("step graph launch failed");
                free_events();
                return rc;
            }
        }
        GADE_HIP(hipMemcpyAsync(m->h_gctl, m->d_gctl, sizeof(int) * GC_COUNT, hipMemcpyDeviceToHost, m->stream));
        GADE_HIP(hipMemcpyAsync(it_h.data(), m->d_iters, sizeof(int) * B * cap, hipMemcpyDeviceToHost, m->stream));
        GADE_HIP(hipMemcpyAsync(tol_h.data(), m->d_tols, sizeof(double) * B * cap, hipMemcpyDeviceToHost, m->stream));
        GADE_HIP(hipStreamSynchronize(m->stream));
        int done = s + W;
        if (m->h_gctl[GC_STALLED]) {
            done = m->h_gctl[GC_STALL_STEP];
            if (done < s || done >= s + W) {
                gade::set_error("device-driven stepping: inconsistent stall record");
                return GADE_EHIP;
            }
            GADE_TRY(finish_stalled_step(m, psi, done, K, it_h, tol_h));
            GADE_HIP(hipStreamSynchronize(m->stream));
            done += 1;
        }
        int kmax = 0;
        for (int b = 0; b < B; b++)
            for (int w = s; w < done; w++) kmax = std::max(kmax, it_h[(size_t)b * cap + w] + 1);
        m->last_k = std::min(std::max(kmax + 1, 2), 16);
        if (usec) {
            for (int w = 0; w < done - s; w++) {
                float ms = 0;
                (void)hipEventElapsedTime(&ms, ev[w], ev[w + 1]);
                usec[s + w] = ms * 1000.0;
            }
            if (done - s < W) usec[done - 1] = -1;  // finished on the host: no device timing
        }
        s = done;
    }
    free_events();
    GADE_HIP(hipStreamSynchronize(m->stream));
    for (int b = 0; b < B; b++) {
        const int nb = nsteps_b ? std::min(nsteps_b[b], nsteps) : nsteps;
        for (int w = 0; w < nb; w++) {
            if (iters) iters[(size_t)b * nsteps + w] = it_h[(size_t)b * cap + w];
            if (tol_out) tol_out[(size_t)b * nsteps + w] = tol_h[(size_t)b * cap + w];
        }
    }
    return GADE_OK;
}

}  // namespace

extern "C" {

int gade_psolve(gade_model_t *m, const double *rhs, double *x) {
    if (!m || !rhs || !x) return GADE_EINVAL;
    GADE_TRY(gade_precond_setup(m));
    GADE_TRY(psolve_async(m, (const double2 *)rhs, (double2 *)x));
    GADE_HIP(hipStreamSynchronize(m->stream));
    return GADE_OK;
}

int gade_bicgstab(gade_model_t *m, const double *rhs, double *x, int *iters, double *tol_error, int *napply) {
    if (!m || !rhs || !x || !iters || !tol_error) return GADE_EINVAL;
    return bicgstab_impl(m, (const double2 *)rhs, (double2 *)x, iters, tol_error, napply);
}

int gade_reset_workspace(gade_model_t *m) {
    if (!m) return GADE_EINVAL;
    double2 *ws[] = {m->r, m->r0, m->v, m->p, m->y, m->z, m->s, m->t};
    for (auto w : ws) GADE_HIP(hipMemsetAsync(w, 0, sizeof(double2) * m->N, m->stream));
    // scalars too (every solve re-initialises its own; the learned guess's inputs read the last residual)
    GADE_HIP(hipMemsetAsync(m->d_sc, 0, sizeof(double) * m->B * SC_STRIDE, m->stream));
    hipLaunchKernelGGL(k_hist_reset, dim3(m->B), dim3(64), 0, m->stream, m->d_ctl);  // the guess histories restart
    GADE_HIP(hipGetLastError());
    if (m->lg) GADE_TRY(gade::lg_reset(m));
    GADE_HIP(hipStreamSynchronize(m->stream));
    return GADE_OK;
}

// ---- checkpoint / resume (tools/dirac_solver.c:203-209 only shifts the start time; this restores the state)
// File: "GADECKP1", int32 [version, B, nblk, blkdim, step, precision, guess_order, has_hist], double time,
// then complex128 psi [B N1], the eight BiCGSTAB vectors r r0 v p y z s t [B N1] each (p and v carry into
// the next solve, tools/gpu_bicgstab.c:36-47), double scalars [B][SC_STRIDE], int32 control words
// [B][CT_STRIDE] and, with has_hist, the guess history [kHistSlots][B N1].
namespace {
constexpr char kCkpMagic[8] = {'G', 'A', 'D', 'E', 'C', 'K', 'P', '1'};

int ckp_io(FILE *fd, bool save, void *host, size_t bytes) {
    const size_t k = save ? fwrite(host, 1, bytes, fd) : fread(host, 1, bytes, fd);
    if (k != bytes) {
        gade::set_error(save ? "checkpoint: write failed" : "checkpoint: file is truncated");
        return GADE_EIO;
    }
    return GADE_OK;
}

// copies one device array through the host in either direction
int ckp_dev(FILE *fd, bool save, void *dev, size_t bytes, std::vector<char> &buf) {
    buf.resize(bytes);
    if (save) GADE_HIP(hipMemcpy(buf.data(), dev, bytes, hipMemcpyDeviceToHost));
    GADE_TRY(ckp_io(fd, save, buf.data(), bytes));
    if (!save) GADE_HIP(hipMemcpy(dev, buf.data(), bytes, hipMemcpyHostToDevice));
    return GADE_OK;
}

int checkpoint(gade_model *m, double2 *psi, int *step, double *time, const char *file, bool save) {
    if (m->guess_order == GADE_GUESS_LEARNED || m->lg_features) {
        gade::set_error("checkpoint: the learned guess's state is not saved; checkpoint runs without it");
        return GADE_EINVAL;
    }
    GADE_HIP(hipStreamSynchronize(m->stream));
    FILE *fd = fopen(file, save ? "wb" : "rb");
    if (!fd) {
        gade::set_error(std::string("checkpoint: cannot open ") + file);
        return GADE_EIO;
    }
    struct Closer {
        FILE *f;
        ~Closer() { if (f) fclose(f); }
    } closer{fd};
    char magic[8];
    int hdr[8] = {1, m->B, m->nblk1, m->n, step ? *step : 0, m->precision, m->guess_order, m->d_hist ? 1 : 0};
    double t = time ? *time : 0;
    if (save) {
        GADE_TRY(ckp_io(fd, true, (void *)kCkpMagic, 8));
        GADE_TRY(ckp_io(fd, true, hdr, sizeof hdr));
        GADE_TRY(ckp_io(fd, true, &t, sizeof t));
    } else {
        GADE_TRY(ckp_io(fd, false, magic, 8));
        int fh[8];
        GADE_TRY(ckp_io(fd, false, fh, sizeof fh));
        GADE_TRY(ckp_io(fd, false, &t, sizeof t));
        if (memcmp(magic, kCkpMagic, 8) || fh[0] != 1) {
            gade::set_error(std::string("checkpoint: ") + file + " is not a checkpoint of this library");
            return GADE_EINVAL;
        }
        if (fh[1] != hdr[1] || fh[2] != hdr[2] || fh[3] != hdr[3] || fh[5] != hdr[5] || fh[6] != hdr[6]) {
            gade::set_error("checkpoint: saved for batch " + std::to_string(fh[1]) + ", " + std::to_string(fh[2]) +
                            " x " + std::to_string(fh[3]) + " blocks, precision " + std::to_string(fh[5]) +
                            ", guess order " + std::to_string(fh[6]) + "; the model differs");
            return GADE_EINVAL;
        }
        if (fh[7] && !m->d_hist) GADE_TRY(hip_alloc((void **)&m->d_hist, sizeof(double2) * kHistSlots * m->N));
        hdr[7] = fh[7];
        if (step) *step = fh[4];
        if (time) *time = t;
    }
    std::vector<char> buf;
    const size_t vb = sizeof(double2) * m->N;
    GADE_TRY(ckp_dev(fd, save, psi, vb, buf));
    double2 *ws[] = {m->r, m->r0, m->v, m->p, m->y, m->z, m->s, m->t};
    for (double2 *w : ws) GADE_TRY(ckp_dev(fd, save, w, vb, buf));
    GADE_TRY(ckp_dev(fd, save, m->d_sc, sizeof(double) * m->B * SC_STRIDE, buf));
    GADE_TRY(ckp_dev(fd, save, m->d_ctl, sizeof(int) * m->B * CT_STRIDE, buf));
    if (hdr[7]) GADE_TRY(ckp_dev(fd, save, m->d_hist, vb * kHistSlots, buf));
    if (save && fflush(fd) != 0) {
        gade::set_error("checkpoint: write failed");
        return GADE_EIO;
    }
    return GADE_OK;
}
}  // namespace

int gade_checkpoint_save(gade_model_t *m, const double *psi, int step, double time, const char *file) {
    if (!m || !psi || !file) return GADE_EINVAL;
    return checkpoint(m, (double2 *)psi, &step, &time, file, true);
}

int gade_checkpoint_load(gade_model_t *m, double *psi, int *step, double *time, const char *file) {
    if (!m || !psi || !file) return GADE_EINVAL;
    return checkpoint(m, (double2 *)psi, step, time, file, false);
}

int gade_extrapolate(gade_model_t *m, double *psi, int order) {
    if (!m || !psi || order < 0 || order > kMaxGuessOrder) return GADE_EINVAL;
    const int saved = m->guess_order;
    m->guess_order = order;
    int rc = order > 0 ? launch_guess(m, (double2 *)psi) : GADE_OK;
    m->guess_order = saved;
    GADE_TRY(rc);
    GADE_HIP(hipStreamSynchronize(m->stream));
    return GADE_OK;
}

int gade_cn_step(gade_model_t *m, double *psi, const double *ft, double tol, int maxit, int *iters, double *tol_error) {
    if (!m || !psi || !ft) return GADE_EINVAL;
    if (m->precision != 0 || m->guess_order == GADE_GUESS_LEARNED || m->lg_features) {
        gade::set_error("gade_cn_step is the FP64 host-driven step; mixed precision and the learned guess run "
                        "through gade_cn_run / gade_cn_run_batch");
        return GADE_EINVAL;
    }
    GADE_TRY(gade_set_time(m, ft));
    // rhs = (S - Hfull) psi  (tools/dirac_solver.c:531-536)
    GADE_TRY(apply_async(m, (const double2 *)psi, m->tmp, -1.0, 0.0));
    GADE_TRY(launch_guess(m, (double2 *)psi));
    int it = maxit;
    double te = tol;
    GADE_TRY(bicgstab_impl(m, m->tmp, (double2 *)psi, &it, &te, nullptr));
    if (iters) *iters = it;
    if (tol_error) *tol_error = te;
    return GADE_OK;
}
int gade_cn_run(gade_model_t *m, double *psi, double E0, double omega, double cycles, double t0, int nsteps,
                double tol, int maxit, int *iters, double *tol_out, double *usec) {
    if (!m || !psi || nsteps < 0) return GADE_EINVAL;
    GADE_TRY(gade_precond_setup(m));
    // step times accumulate as in the reference loop (tools/dirac_solver.c:509)
    std::vector<double> times(nsteps);
    double time = t0;
    for (int step = 0; step < nsteps; step++) {
        time = time + m->dt;
        times[step] = time;
    }
    if (m->device_loop || m->B > 1 || m->precision == 1) {
        std::vector<double> params((size_t)4 * m->B);
        for (int b = 0; b < m->B; b++) {
            params[(size_t)4 * b] = E0;
            params[(size_t)4 * b + 1] = omega;
            params[(size_t)4 * b + 2] = cycles;
            params[(size_t)4 * b + 3] = 0.0;
        }
        return cn_run_device(m, (double2 *)psi, times.data(), nsteps, params.data(), nullptr, tol, maxit, iters,
                             tol_out, usec);
    }
    hipEvent_t ea, eb;
    GADE_HIP(hipEventCreate(&ea));
    GADE_HIP(hipEventCreate(&eb));
    for (int step = 0; step < nsteps; step++) {
        double ft[12];
        gade_pulse(E0, omega, cycles, times[step], ft);
        if (usec) GADE_HIP(hipEventRecord(ea, m->stream));
        int it = 0;
        double te = 0;
        int rc = gade_cn_step(m, psi, ft, tol, maxit, &it, &te);
        if (rc) {
            hipEventDestroy(ea);
            hipEventDestroy(eb);
            return rc;
        }
        if (usec) {
            GADE_HIP(hipEventRecord(eb, m->stream));
            GADE_HIP(hipEventSynchronize(eb));
            float ms = 0;
            hipEventElapsedTime(&ms, ea, eb);
            usec[step] = ms * 1000.0;
        }
        if (iters) iters[step] = it;
        if (tol_out) tol_out[step] = te;
    }
    hipEventDestroy(ea);
    hipEventDestroy(eb);
    return GADE_OK;
}

int gade_model_batch(const gade_model_t *m) { return m ? m->B : 0; }

int gade_cn_run_batch(gade_model_t *m, double *psi, const double *params, const int *nsteps_b, double t0, int nsteps,
                      double tol, int maxit, int *iters, double *tol_out) {
    if (!m || !psi || !params || nsteps < 0) return GADE_EINVAL;
    if (nsteps_b)
        for (int b = 0; b < m->B; b++)
            if (nsteps_b[b] < 0) return GADE_EINVAL;
    std::vector<double> times(nsteps);
    double time = t0;
    for (int step = 0; step < nsteps; step++) {
        time = time + m->dt;  // tools/dirac_solver.c:509
        times[step] = time;
    }
    return cn_run_device(m, (double2 *)psi, times.data(), nsteps, params, nsteps_b, tol, maxit, iters, tol_out,
                         nullptr);
}

int gade_autotune(gade_model_t *m, const double *psi, const double *params, double t0, int nsteps, double tol,
                  int maxit, int reps, double *report, int *ncand, int *chosen) {
    if (!m || !psi || !params || nsteps <= 0 || reps <= 0) return GADE_EINVAL;
    GADE_TRY(gade_precond_setup(m));
    std::vector<double> times(nsteps);
    double time = t0;
    for (int step = 0; step < nsteps; step++) {
        time = time + m->dt;
        times[step] = time;
    }
    struct Cand {
        int window, iter_batch, graph, dia_r, dia_p, ang_lb, dia_hw, fused;
        double ms;
    };
    // stage 1: kernel shapes (banded tile shape x angular multipole group) with the model's stepping;
    // stage 2: stepping with the best shapes
    // (the FP64 MFMA radial kernel has one shape: only the angular group is tuned then; the tile shape still
    // matters to the FP32 kernels of the mixed-precision form)
    // (the fused angular + radial form: the shapes with up to 2 outputs per lane, the angular group unused)
    std::vector<Cand> cands;
    const bool one_tile = m->radial_mfma && m->precision == 0;
    const int g0 = m->use_graph ? 1 : 0;
    if (m->radial_kind == 2)
        for (auto rp : {std::pair<int, int>{8, 2}, {4, 2}, {4, 4}, {8, 1}, {16, 1}, {4, 1}}) {
            if (one_tile && (rp.first != m->dia_r || rp.second != m->dia_p)) continue;
            for (int hw : {1, 2}) {
                if (one_tile && hw != m->dia_hw) continue;
                for (int lb : {4, 2, 8})
                    if (ang_lds_ok(m, lb))
                        cands.push_back({m->window, m->iter_batch, g0, rp.first, rp.second, lb, hw, 0, 0.0});
                if (!one_tile && fused_shape_ok(rp.first, rp.second))
                    cands.push_back({m->window, m->iter_batch, g0, rp.first, rp.second, m->ang_lb, hw, 1, 0.0});
            }
        }
    const size_t nshape = cands.size();
    for (int g : {1, 0}) {
        if (g && (m->radial_kind != 2 || !m->use_graph)) continue;  // graph 0 on the model excludes graphs
        for (int w : {2, 4, 8, 16, 32})
            for (int k : {0, 1, 2, 4}) cands.push_back({w, k, g, -1, -1, -1, -1, -1, 0.0});
    }
    const size_t N = (size_t)m->N, B = (size_t)m->B;
    // the solver state a run starts from (BiCGSTAB workspace, which carries over between solves, scalars,
    // control words and guess histories): saved once, restored before every probe and on return, so that
    // tuning leaves the next run exactly as it would have been
    double2 *ws[] = {m->r, m->r0, m->v, m->p, m->y, m->z, m->s, m->t};
    constexpr int NWS = 8;
    const bool hist = m->d_hist != nullptr;
    const size_t vec_bytes = sizeof(double2) * N;
    void *lg_ptr = nullptr;
    size_t lg_bytes = 0;
    gade::lg_state(m, &lg_ptr, &lg_bytes);
    const size_t state_bytes = vec_bytes * (NWS + (hist ? kHistSlots : 0)) + sizeof(double) * B * SC_STRIDE +
                               sizeof(int) * B * CT_STRIDE + lg_bytes;
    double2 *work = nullptr;
    char *state = nullptr;
    GADE_TRY(hip_alloc((void **)&work, vec_bytes));
    if (int rc = hip_alloc((void **)&state, state_bytes)) {
        (void)hipFree(work);
        return rc;
    }
    auto copy_state = [&](bool save) -> int {
        char *q = state;
        auto mv = [&](void *dev, size_t bytes) {
            hipError_t e = save ? hipMemcpyAsync(q, dev, bytes, hipMemcpyDeviceToDevice, m->stream)
                                : hipMemcpyAsync(dev, q, bytes, hipMemcpyDeviceToDevice, m->stream);
            q += bytes;
            return e;
        };
        hipError_t e = hipSuccess;
        for (int k = 0; k < NWS && e == hipSuccess; k++) e = mv(ws[k], vec_bytes);
        if (hist && e == hipSuccess) e = mv(m->d_hist, vec_bytes * kHistSlots);
        if (e == hipSuccess) e = mv(m->d_sc, sizeof(double) * B * SC_STRIDE);
        if (e == hipSuccess) e = mv(m->d_ctl, sizeof(int) * B * CT_STRIDE);
        if (e == hipSuccess && lg_ptr) e = mv(lg_ptr, lg_bytes);
        if (e != hipSuccess) {
            gade::set_error(std::string("gade_autotune: state copy: ") + hipGetErrorString(e));
            return GADE_EHIP;
        }
        return GADE_OK;
    };
    const int sv_window = m->window, sv_batch = m->iter_batch, sv_profile = m->profile, sv_k = m->last_k;
    const int sv_r = m->dia_r, sv_p = m->dia_p, sv_lb = m->ang_lb, sv_hw = m->dia_hw, sv_fused = m->ang_fused;
    const bool sv_graph = m->use_graph;
    auto set_shape = [&](int r, int pp, int lb, int hw, int fused) -> int {
        if (lb != m->ang_lb || hw != m->dia_hw || fused != m->ang_fused) {
            drop_step_graph(m);
            m->ang_lb = lb;
            m->dia_hw = hw;
            m->ang_fused = fused;
        }
        if (!m->dia_ok || (r == m->dia_r && pp == m->dia_p)) return GADE_OK;
        drop_step_graph(m);
        if (r != m->dia_r)
            GADE_TRY(kappa_tiles(m, r, &m->d_dtiles, &m->d_dtile_q, &m->d_dtile_k, &m->ndtiles, &m->d_dtile_b));
        m->dia_r = r;
        m->dia_p = pp;
        return GADE_OK;
    };
    auto restore = [&]() {
        m->window = sv_window;
        m->iter_batch = sv_batch;
        m->use_graph = sv_graph;
        m->profile = sv_profile;
        m->last_k = sv_k;
        (void)set_shape(sv_r, sv_p, sv_lb, sv_hw, sv_fused);
        const int rc = copy_state(false);
        (void)hipStreamSynchronize(m->stream);
        (void)hipFree(work);
        (void)hipFree(state);
        return rc;
    };
    if (int rc = copy_state(true)) {
        (void)hipFree(work);
        (void)hipFree(state);
        return rc;
    }
    m->profile = 0;
    std::vector<double2> ref(N), cur(N);
    std::vector<int> ref_it(B * nsteps), it(B * nsteps);
    std::vector<double> ref_tol(B * nsteps), tl(B * nsteps);
    size_t best_shape = 0;
    for (size_t c = 0; c < cands.size(); c++) {
        if (c == nshape && nshape > 0) {
            for (size_t k = 1; k < nshape; k++)
                if (cands[k].ms < cands[best_shape].ms) best_shape = k;
            if (int rc = set_shape(cands[best_shape].dia_r, cands[best_shape].dia_p, cands[best_shape].ang_lb,
                                   cands[best_shape].dia_hw, cands[best_shape].fused)) {
                restore();
                return rc;
            }
        }
        if (cands[c].dia_r < 0) {
            cands[c].dia_r = m->dia_r;
            cands[c].dia_p = m->dia_p;
            cands[c].ang_lb = m->ang_lb;
            cands[c].dia_hw = m->dia_hw;
            cands[c].fused = m->ang_fused;
        } else if (int rc = set_shape(cands[c].dia_r, cands[c].dia_p, cands[c].ang_lb, cands[c].dia_hw,
                                      cands[c].fused)) {
            restore();
            return rc;
        }
        m->window = cands[c].window;
        m->iter_batch = cands[c].iter_batch;
        m->use_graph = cands[c].graph != 0;
        double best = 1e300;
        for (int r = 0; r < reps; r++) {
            int rc = GADE_OK;
            if (hipMemcpyAsync(work, psi, sizeof(double2) * N, hipMemcpyDeviceToDevice, m->stream) != hipSuccess)
                rc = GADE_EHIP;
            if (rc == GADE_OK) rc = copy_state(false);
            m->last_k = sv_k;
            const auto t_a = std::chrono::steady_clock::now();
            if (rc == GADE_OK)
                rc = cn_run_device(m, work, times.data(), nsteps, params, nullptr, tol, maxit, it.data(), tl.data(),
                                   nullptr);
            if (rc == GADE_OK && hipStreamSynchronize(m->stream) != hipSuccess) rc = GADE_EHIP;
            const auto t_b = std::chrono::steady_clock::now();
            if (rc != GADE_OK) {
                restore();
                return rc;
            }
            best = std::min(best, std::chrono::duration<double, std::milli>(t_b - t_a).count());
        }
        cands[c].ms = best;
        if (hipMemcpy(cur.data(), work, sizeof(double2) * N, hipMemcpyDeviceToHost) != hipSuccess) {
            restore();
            gade::set_error("gade_autotune: copy of the probe result failed");
            return GADE_EHIP;
        }
        if (c == 0) {
            ref.swap(cur);
            ref_it = it;
            ref_tol = tl;
        } else if (memcmp(ref.data(), cur.data(), sizeof(double2) * N) || ref_it != it ||
                   memcmp(ref_tol.data(), tl.data(), sizeof(double) * tl.size())) {
            restore();
            gade::set_error("gade_autotune: candidate window " + std::to_string(cands[c].window) + " iter_batch " +
                            std::to_string(cands[c].iter_batch) + " graph " + std::to_string(cands[c].graph) +
                            " tile " + std::to_string(cands[c].dia_r) + "x" + std::to_string(cands[c].dia_p) +
                            " halves " + std::to_string(cands[c].dia_hw) + " angular group " +
                            std::to_string(cands[c].ang_lb) + " fused " + std::to_string(cands[c].fused) +
                            " changed the trajectory");
            return GADE_EINVAL;
        }
    }
    size_t pick = nshape < cands.size() ? nshape : 0;  // the fastest stepping (with the best shape)
    for (size_t c = pick + 1; c < cands.size(); c++)
        if (cands[c].ms < cands[pick].ms) pick = c;
    GADE_TRY(restore());
    m->window = cands[pick].window;
    m->iter_batch = cands[pick].iter_batch;
    m->use_graph = cands[pick].graph != 0;
    GADE_TRY(set_shape(cands[pick].dia_r, cands[pick].dia_p, cands[pick].ang_lb, cands[pick].dia_hw,
                       cands[pick].fused));
    if (report)
        for (size_t c = 0; c < cands.size() && c < GADE_TUNE_MAX; c++) {
            double *r = report + (size_t)GADE_TUNE_COLS * c;
            r[0] = cands[c].window;
            r[1] = cands[c].iter_batch;
            r[2] = cands[c].graph;
            r[3] = cands[c].dia_r;
            r[4] = cands[c].dia_p;
            r[5] = cands[c].ang_lb;
            r[6] = cands[c].ms;
            r[7] = cands[c].dia_hw;
            r[8] = cands[c].fused;
        }
    if (ncand) *ncand = (int)std::min<size_t>(cands.size(), GADE_TUNE_MAX);
    if (chosen) *chosen = (int)pick;
    return GADE_OK;
}

int gade_snorm(gade_model_t *m, const double *x, double *norm, double *pop_host) {
    if (!m || !x) return GADE_EINVAL;
    double *d_pop = nullptr;
    GADE_TRY(hip_alloc((void **)&d_pop, sizeof(double) * m->nblk));
    hipLaunchKernelGGL(k_snorm, dim3(m->nblk), dim3(256), 0, m->stream, (const double2 *)x, m->d_D, m->d_ell_col,
                       m->d_blk_q, m->n, m->ellw, m->nq, d_pop);
    std::vector<double> pop(m->nblk);
    hipError_t e = hipMemcpyAsync(pop.data(), d_pop, sizeof(double) * m->nblk, hipMemcpyDeviceToHost, m->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(m->stream);
    hipFree(d_pop);
    if (e != hipSuccess) {
        gade::set_error(std::string("gade_snorm: ") + hipGetErrorString(e));
        return GADE_EHIP;
    }
    double tot = 0;
    for (double v : pop) tot += v;
    if (norm) *norm = tot;
    if (pop_host) memcpy(pop_host, pop.data(), sizeof(double) * m->nblk);
    return GADE_OK;
}

// ---------------------------------------------------------------------------
// assembled reference formulation
// ---------------------------------------------------------------------------
int gade_assembled_setup(gade_model_t *m) {
    if (!m) return GADE_EINVAL;
    if (m->asm_ready) return GADE_OK;
    if (m->B != 1) {
        gade::set_error("the assembled formulation serves one instance");
        return GADE_EINVAL;
    }
    m->nnz_full = (long long)m->hnnz * m->nnzb;
    if (m->nnz_full > 2147483647LL) {
        gade::set_error("assembled Hfull exceeds int32 indexing (csr_index_t, tools/types.h:37)");
        return GADE_EINVAL;
    }
    const long long N = m->N;
    int *d_subAp = nullptr, *d_subAi = nullptr;
    GADE_TRY(upload(&d_subAp, m->rad_Ap.data(), m->rad_Ap.size()));
    GADE_TRY(upload(&d_subAi, m->rad_Ai.data(), m->rad_Ai.size()));
    GADE_TRY(hip_alloc((void **)&m->d_fAp, sizeof(int) * (N + 1)));
    GADE_TRY(hip_alloc((void **)&m->d_fAi, sizeof(int) * m->nnz_full));
    GADE_TRY(hip_alloc((void **)&m->d_fAx, sizeof(double2) * m->nnz_full));
    hipLaunchKernelGGL(k_full_Ap, dim3(grid_for(N + 1, 256, 1u << 30)), dim3(256), 0, m->stream, m->d_hAp, d_subAp,
                       m->nblk, m->n, m->nnzb, m->d_fAp);
    GADE_HIP(hipGetLastError());
    dim3 g2((unsigned)std::min<long long>(N, 65535), (unsigned)((N + 65534) / 65535));
    hipLaunchKernelGGL(k_full_Ai, g2, dim3(256), 0, m->stream, m->d_hAp, m->d_hAi, d_subAp, d_subAi, m->d_fAp, m->nblk,
                       m->n, m->d_fAi);
    GADE_HIP(hipGetLastError());
    // S: block diagonal, same expansion with a diagonal block pattern
    std::vector<int> dAp(m->nblk + 1), dAi(m->nblk);
    for (int i = 0; i < m->nblk; i++) { dAp[i] = i; dAi[i] = i; }
    dAp[m->nblk] = m->nblk;
    int *d_dAp = nullptr, *d_dAi = nullptr;
    GADE_TRY(upload(&d_dAp, dAp.data(), dAp.size()));
    GADE_TRY(upload(&d_dAi, dAi.data(), dAi.size()));
    const long long snnz = (long long)m->nblk * m->nnzb;
    GADE_TRY(hip_alloc((void **)&m->d_SAp, sizeof(int) * (N + 1)));
    GADE_TRY(hip_alloc((void **)&m->d_SAi, sizeof(int) * snnz));
    GADE_TRY(hip_alloc((void **)&m->d_SAx, sizeof(double2) * snnz));
    hipLaunchKernelGGL(k_full_Ap, dim3(grid_for(N + 1, 256, 1u << 30)), dim3(256), 0, m->stream, d_dAp, d_subAp,
                       m->nblk, m->n, m->nnzb, m->d_SAp);
    hipLaunchKernelGGL(k_full_Ai, g2, dim3(256), 0, m->stream, d_dAp, d_dAi, d_subAp, d_subAi, m->d_SAp, m->nblk, m->n,
                       m->d_SAi);
    GADE_HIP(hipGetLastError());
    // S values: block r copies S(k_r) in CSR order
    std::vector<double2> Sh((size_t)snnz);
    std::vector<double2> Dh((size_t)m->nq * m->ellw * m->n);
    GADE_HIP(hipMemcpy(Dh.data(), m->d_D + (size_t)2 * m->nq * m->ellw * m->n, sizeof(double2) * Dh.size(),
                       hipMemcpyDeviceToHost));
    for (int r = 0; r < m->nblk; r++)
        for (int a = 0; a < m->n; a++)
            for (int k = m->rad_Ap[a]; k < m->rad_Ap[a + 1]; k++)
                Sh[(size_t)r * m->nnzb + k] = Dh[((size_t)m->blk_q[r] * m->ellw + (k - m->rad_Ap[a])) * m->n + a];
    GADE_HIP(hipMemcpy(m->d_SAx, Sh.data(), sizeof(double2) * snnz, hipMemcpyHostToDevice));
    GADE_HIP(hipStreamSynchronize(m->stream));
    hipFree(d_subAp); hipFree(d_subAi); hipFree(d_dAp); hipFree(d_dAi);
    m->asm_ready = true;
    return GADE_OK;
}

int gade_assembled_timedep_into(gade_model_t *m, const int *Ap_dev, double *Ax_dev) {
    if (!m || !Ap_dev || !Ax_dev) return GADE_EINVAL;
    if (m->radial_kind == 2) GADE_TRY(launch_gsums(m, true));  // the fused path keeps only the DIA layout
    int *d_subAp = nullptr;
    GADE_TRY(upload(&d_subAp, m->rad_Ap.data(), m->rad_Ap.size()));
    const long long N = m->N;
    dim3 g2((unsigned)std::min<long long>(N, 65535), (unsigned)((N + 65534) / 65535));
    {
        Prof pr(m, "assemble", GADE_PROF_ASSEMBLE);
#define ASM(LM) \
    case LM: hipLaunchKernelGGL(k_assemble<LM>, g2, dim3(256), 0, m->stream, m->d_hAp, m->d_hAi, m->d_kap, m->d_coef, m->d_diag, d_subAp, m->d_ell_pos, m->d_GE, m->d_D, m->d_blk_q, Ap_dev, m->nblk, m->n, m->ellw, m->nq, (double2 *)Ax_dev); break;
        switch (m->lmax) {
            ASM(1) ASM(2) ASM(3) ASM(4) ASM(5) ASM(6) ASM(7) ASM(8) ASM(9) ASM(10) ASM(11) ASM(12) ASM(13) ASM(14) ASM(15) ASM(16)
            default: (void)hipFree(d_subAp); return GADE_EINVAL;
        }
#undef ASM
    }
    hipError_t e = hipGetLastError();
    if (e == hipSuccess) e = hipStreamSynchronize(m->stream);
    (void)hipFree(d_subAp);
    if (e != hipSuccess) {
        gade::set_error(std::string("k_assemble: ") + hipGetErrorString(e));
        return GADE_EHIP;
    }
    return GADE_OK;
}

int gade_assembled_timedep(gade_model_t *m) {
    if (!m) return GADE_EINVAL;
    GADE_TRY(gade_assembled_setup(m));
    return gade_assembled_timedep_into(m, m->d_fAp, (double *)m->d_fAx);
}

int gade_assembled_spmv(gade_model_t *m, const double *x, double *y, double alpha, double beta, int with_S) {
    if (!m || !x || !y || !m->asm_ready) {
        gade::set_error("gade_assembled_spmv: call gade_assembled_timedep first");
        return GADE_EINVAL;
    }
    const long long N = m->N;
    unsigned grid = (unsigned)((N * 64 + 255) / 256);
    {
        Prof pr(m, "csr_spmv", GADE_PROF_CSR_SPMV);
        hipLaunchKernelGGL(k_csr_spmv, dim3(grid), dim3(256), 0, m->stream, m->d_fAp, m->d_fAi, m->d_fAx,
                           (const double2 *)x, N, make_double2(alpha, 0), make_double2(beta, 0), (double2 *)y);
    }
    if (with_S) {
        Prof pr(m, "csr_spmv_S", GADE_PROF_CSR_SPMV);
        hipLaunchKernelGGL(k_csr_spmv, dim3(grid), dim3(256), 0, m->stream, m->d_SAp, m->d_SAi, m->d_SAx,
                           (const double2 *)x, N, make_double2(1.0, 0), make_double2(1.0, 0), (double2 *)y);
    }
    GADE_HIP(hipGetLastError());
    GADE_HIP(hipStreamSynchronize(m->stream));
    return GADE_OK;
}

int gade_assembled_get(gade_model_t *m, int *Ap, int *Ai, int *Ai_sub_map, double *Ax) {
    if (!m || !m->asm_ready) return GADE_EINVAL;
    if (Ap) GADE_HIP(hipMemcpy(Ap, m->d_fAp, sizeof(int) * (m->N + 1), hipMemcpyDeviceToHost));
    if (Ai) GADE_HIP(hipMemcpy(Ai, m->d_fAi, sizeof(int) * m->nnz_full, hipMemcpyDeviceToHost));
    if (Ax) GADE_HIP(hipMemcpy(Ax, m->d_fAx, sizeof(double2) * m->nnz_full, hipMemcpyDeviceToHost));
    if (Ai_sub_map) {
        int *d_rowof = nullptr, *d_map = nullptr, *d_subAp = nullptr;
        std::vector<int> rowof(m->hnnz);
        for (int r = 0; r < m->nblk; r++)
            for (int p = m->hall_Ap[r]; p < m->hall_Ap[r + 1]; p++) rowof[p] = r;
        GADE_TRY(upload(&d_rowof, rowof.data(), rowof.size()));
        GADE_TRY(upload(&d_subAp, m->rad_Ap.data(), m->rad_Ap.size()));
        GADE_TRY(hip_alloc((void **)&d_map, sizeof(int) * m->nnz_full));
        hipLaunchKernelGGL(k_full_submap, dim3(m->hnnz), dim3(256), 0, m->stream, m->d_hAp, d_subAp, m->d_fAp,
                           m->nblk, m->n, m->nnzb, d_rowof, d_map);
        hipError_t e = hipGetLastError();
        if (e == hipSuccess) e = hipMemcpy(Ai_sub_map, d_map, sizeof(int) * m->nnz_full, hipMemcpyDeviceToHost);
        hipFree(d_rowof); hipFree(d_map); hipFree(d_subAp);
        if (e != hipSuccess) {
            gade::set_error(std::string("k_full_submap: ") + hipGetErrorString(e));
            return GADE_EHIP;
        }
    }
    return GADE_OK;
}

int gade_set_param(gade_model_t *m, const char *name, int value) {
    if (!m || !name) return GADE_EINVAL;
    if (!strcmp(name, "tile_r")) {
        if (value != 1 && value != 2 && value != 4 && value != 8) {
            gade::set_error("tile_r must be 1, 2, 4 or 8");
            return GADE_EINVAL;
        }
        GADE_HIP(hipStreamSynchronize(m->stream));
        m->tile_r = value;
        return build_tiles(m);
    }
    if (!strcmp(name, "device_loop")) {
        m->device_loop = value != 0;
        return GADE_OK;
    }
    if (!strcmp(name, "graph")) {
        m->use_graph = value != 0;
        return GADE_OK;
    }
    if (!strcmp(name, "psolve_chunk")) {
        if (value != 16 && value != 32) {
            gade::set_error("psolve_chunk must be 16 or 32");
            return GADE_EINVAL;
        }
        if (value != m->psolve_ch) {
            GADE_HIP(hipStreamSynchronize(m->stream));
            drop_step_graph(m);
            gade::blocklu_free(m->pre);  // rebuilt with the new level height on next use
            m->pre = nullptr;
            for (auto &f : m->lu32) {
                if (f) (void)hipFree(f);
                f = nullptr;
            }
            if (m->tw32) (void)hipFree(m->tw32);
            m->tw32 = nullptr;
            m->psolve_ch = value;
        }
        return GADE_OK;
    }
    if (!strcmp(name, "ang_list")) {
        if (value != 0 && value != 1) {
            gade::set_error("ang_list must be 0 (Hall rows) or 1 (coupling lists)");
            return GADE_EINVAL;
        }
        if (m->ang_list != value) drop_step_graph(m);
        m->ang_list = value;
        return GADE_OK;
    }
    if (!strcmp(name, "ang_lb")) {
        if ((value != 2 && value != 4 && value != 8) || !ang_lds_ok(m, value)) {
            gade::set_error("ang_lb must be 2, 4 or 8 (and the Hall rows short enough for its LDS staging)");
            return GADE_EINVAL;
        }
        drop_step_graph(m);
        m->ang_lb = value;
        return GADE_OK;
    }
    if (!strcmp(name, "ang_fused")) {
        // 1: the banded radial kernel forms each V window from x and the coupling lists (no angular kernel, no
        // V in HBM; bit-identical results), 0: the angular kernel writes V and the radial kernel reads it
        if ((value != 0 && value != 1) || (value == 1 && !fused_shape_ok(m->dia_r, m->dia_p))) {
            gade::set_error("ang_fused must be 0 or 1 (1 needs a tile shape with at most 2 outputs per lane)");
            return GADE_EINVAL;
        }
        if (value != m->ang_fused) drop_step_graph(m);
        m->ang_fused = value;
        return GADE_OK;
    }
    if (!strcmp(name, "dia_halves")) {
        // output halves per banded workgroup: 1 (one wave per half) or 2 (two waves share each window)
        if (value != 1 && value != 2) {
            gade::set_error("dia_halves must be 1 or 2");
            return GADE_EINVAL;
        }
        if (value != m->dia_hw) drop_step_graph(m);
        m->dia_hw = value;
        return GADE_OK;
    }
    if (!strcmp(name, "dia_rows") || !strcmp(name, "dia_outputs")) {
        const bool rows = name[4] == 'r';
        const int r = rows ? value : m->dia_r, pp = rows ? m->dia_p : value;
        if (!dia_shape_ok(r, pp)) {
            gade::set_error("banded tile shape (dia_rows, dia_outputs) must be (4,1), (4,2), (4,4), (8,1), (8,2) "
                            "or (16,1)");
            return GADE_EINVAL;
        }
        if (m->ang_fused && !fused_shape_ok(r, pp)) {
            gade::set_error("the fused angular + radial form (ang_fused 1) has no (4,4) tile shape");
            return GADE_EINVAL;
        }
        if (!m->dia_ok) {
            gade::set_error("the radial pattern is not banded");
            return GADE_EINVAL;
        }
        drop_step_graph(m);
        if (rows && r != m->dia_r)
            GADE_TRY(kappa_tiles(m, r, &m->d_dtiles, &m->d_dtile_q, &m->d_dtile_k, &m->ndtiles, &m->d_dtile_b));
        m->dia_r = r;
        m->dia_p = pp;
        return GADE_OK;
    }
    if (!strcmp(name, "precision")) {
        if (value != 0 && value != 1) {
            gade::set_error("precision must be 0 (FP64) or 1 (mixed)");
            return GADE_EINVAL;
        }
        m->precision = value;
        return GADE_OK;
    }
    if (!strcmp(name, "radial_mfma")) {
        // 1: the FP64 banded radial products on the matrix cores (when the pattern allows), 0: VALU (default:
        // on gfx950 the FP64 matrix rate equals the vector rate, and the 4x4x4 band form carries zero padding)
        if ((value != 0 && value != 1) || (value == 1 && !m->mfma_ok)) {
            gade::set_error(value == 1 ? "the radial pattern does not fit the MFMA kernel" : "radial_mfma must be 0 or 1");
            return GADE_EINVAL;
        }
        if (value != m->radial_mfma) {
            drop_step_graph(m);
            m->radial_mfma = value;
            if (m->radial_kind == 2) GADE_TRY(launch_gsums(m, false));  // the current f(t) in the kernel's layout
        }
        return GADE_OK;
    }
    if (!strcmp(name, "mp_inner_iters")) {
        if (value < 1 || value > 64) {
            gade::set_error("mp_inner_iters must be 1..64");
            return GADE_EINVAL;
        }
        m->mp_inner_iters = value;
        return GADE_OK;
    }
    if (!strcmp(name, "mp_inner_tol_exp")) {
        if (value < 1 || value > 12) {
            gade::set_error("mp_inner_tol_exp must be 1..12 (inner tolerance 10^-value)");
            return GADE_EINVAL;
        }
        m->mp_inner_tol = std::pow(10.0, -value);
        return GADE_OK;
    }
    if (!strcmp(name, "guess_order")) {
        if (value < 0 || (value > kMaxGuessOrder && value != GADE_GUESS_LEARNED)) {
            gade::set_error("guess_order must be 0..3 (extrapolation) or 4 (learned)");
            return GADE_EINVAL;
        }
        m->guess_order = value;
        return GADE_OK;
    }
    if (!strcmp(name, "lg_features")) {
        m->lg_features = value != 0;
        return GADE_OK;
    }
    if (!strcmp(name, "lg_zmax")) {
        if (value < 0 || value > 1000) {
            gade::set_error("lg_zmax must be 0 (no range gate) .. 1000");
            return GADE_EINVAL;
        }
        m->lg_zmax = value;
        return GADE_OK;
    }
    if (!strcmp(name, "window") || !strcmp(name, "iter_batch")) {
        if (value < (name[0] == 'w' ? 1 : 0) || value > 4096) {
            gade::set_error(std::string(name) + " out of range");
            return GADE_EINVAL;
        }
        (name[0] == 'w' ? m->window : m->iter_batch) = value;
        return GADE_OK;
    }
    if (!strcmp(name, "profile_floor_ns")) {
        if (value < 0) {
            gade::set_error("profile_floor_ns must be >= 0");
            return GADE_EINVAL;
        }
        m->prof_floor_ns = value;
        return GADE_OK;
    }
    if (!strcmp(name, "profile_period")) {
        if (value < 1) {
            gade::set_error("profile_period must be >= 1");
            return GADE_EINVAL;
        }
        m->prof_period = value;
        return GADE_OK;
    }
    if (!strcmp(name, "radial")) {
        // 0: banded kernel when the pattern allows it, 1: ELL kernel, 2: banded kernel
        if (value < 0 || value > 2 || (value == 2 && !m->dia_ok)) {
            gade::set_error(value == 2 ? "the radial pattern is not banded: no DIA kernel" : "radial must be 0, 1 or 2");
            return GADE_EINVAL;
        }
        const int kind = value == 0 ? (m->dia_ok ? 2 : 1) : value;
        if (kind == 1 && m->B > 1) {
            gade::set_error("batched models need the banded radial kernel");
            return GADE_EINVAL;
        }
        if (kind != m->radial_kind) {
            m->radial_kind = kind;
            GADE_TRY(launch_gsums(m, kind != 2));  // coefficients of the current f(t) in the new layout
        }
        return GADE_OK;
    }
    gade::set_error(std::string("unknown parameter ") + name);
    return GADE_EINVAL;
}

int gade_get_param(const gade_model_t *m, const char *name, int *value) {
    if (!m || !name || !value) return GADE_EINVAL;
    if (!strcmp(name, "tile_r")) {
        *value = m->tile_r;
        return GADE_OK;
    }
    if (!strcmp(name, "radial")) {
        *value = m->radial_kind;
        return GADE_OK;
    }
    if (!strcmp(name, "device_loop")) {
        *value = m->device_loop;
        return GADE_OK;
    }
    if (!strcmp(name, "guess_order")) {
        *value = m->guess_order;
        return GADE_OK;
    }
    if (!strcmp(name, "graph")) {
        *value = m->use_graph;
        return GADE_OK;
    }
    if (!strcmp(name, "precision")) {
        *value = m->precision;
        return GADE_OK;
    }
    if (!strcmp(name, "radial_mfma")) {
        *value = m->radial_mfma;
        return GADE_OK;
    }
    if (!strcmp(name, "dia_rows")) {
        *value = m->dia_r;
        return GADE_OK;
    }
    if (!strcmp(name, "ang_list")) {
        *value = m->ang_list;
        return GADE_OK;
    }
    if (!strcmp(name, "ang_lb")) {
        *value = m->ang_lb;
        return GADE_OK;
    }
    if (!strcmp(name, "psolve_chunk")) {
        *value = m->pre ? m->pre->ch : m->psolve_ch;
        return GADE_OK;
    }
    if (!strcmp(name, "dia_halves")) {
        *value = m->dia_hw;
        return GADE_OK;
    }
    if (!strcmp(name, "ang_fused")) {
        *value = m->ang_fused;
        return GADE_OK;
    }
    if (!strcmp(name, "dia_outputs")) {
        *value = m->dia_p;
        return GADE_OK;
    }
    if (!strcmp(name, "lg_features")) {
        *value = m->lg_features;
        return GADE_OK;
    }
    if (!strcmp(name, "lg_zmax")) {
        *value = m->lg_zmax;
        return GADE_OK;
    }
    if (!strcmp(name, "mp_inner_iters")) {
        *value = m->mp_inner_iters;
        return GADE_OK;
    }
    if (!strcmp(name, "mp_inner_tol_exp")) {
        *value = (int)std::lround(-std::log10(m->mp_inner_tol));
        return GADE_OK;
    }
    if (!strcmp(name, "window")) {
        *value = m->window;
        return GADE_OK;
    }
    if (!strcmp(name, "iter_batch")) {
        *value = m->iter_batch;
        return GADE_OK;
    }
    if (!strcmp(name, "profile_period")) {
        *value = m->prof_period;
        return GADE_OK;
    }
    gade::set_error(std::string("unknown parameter ") + name);
    return GADE_EINVAL;
}

int gade_profile(gade_model_t *m, int mask) {
    if (!m) return GADE_EINVAL;
    m->profile = mask;
    if (!mask) {
        resolve_pending(m);
        m->timers.clear();
    }
    return GADE_OK;
}

int gade_profile_get(gade_model_t *m, const char *name, double *total_ms, long long *count) {
    if (!m || !name) return GADE_EINVAL;
    resolve_pending(m);
    auto it = m->timers.find(name);
    if (total_ms) *total_ms = it == m->timers.end() ? 0 : it->second.total_ms;
    if (count) *count = it == m->timers.end() ? 0 : it->second.count;
    return GADE_OK;
}

}  // extern "C"

// ---------------------------------------------------------------------------
// block-diagonal banded LU (shared by the model and the drop-in compute_preconditioner)
// ---------------------------------------------------------------------------
namespace gade {

void blocklu_free(gade_blocklu *lu) {
    if (!lu) return;
    void *ptrs[] = {lu->d_AB,      lu->d_Linv, lu->d_Uinv, lu->d_Loff,  lu->d_Uoff, lu->d_blk_q,
                    lu->d_wblocks, lu->d_flag, lu->d_order, lu->d_TW};
    for (void *p : ptrs)
        if (p) (void)hipFree(p);
    delete lu;
}

// Twisted (burn-at-both-ends) block factorization of each distinct block over 16-row levels c = 0..nch-1 of the
// interleaved band (band <= 16, so level c couples only to c-1 and c+1): P = tridiag(A_c, D_c, B_c), rows
// beyond n identity.  Top levels c < m: S_c = D_c - A_c S_{c-1}^-1 B_{c-1}; bottom levels c > m:
// T_c = D_c - B_c T_{c+1}^-1 A_{c+1}; the meeting level m: M = D_m - A_m S_{m-1}^-1 B_{m-1} - B_m T_{m+1}^-1 A_{m+1}.
// Per level four 16 x 16 matrices (column-major: entry (i, t) at t*16 + i):
//   X_c = A_c S_{c-1}^-1 (0 < c <= m), B_c T_{c+1}^-1 (m < c < nch-1);   Y_c = S_c^-1, M^-1, T_c^-1;
//   Z_c = S_c^-1 B_c (c < m), T_c^-1 A_c (c > m);   W_m = B_m T_{m+1}^-1 (zero elsewhere).
// The solve (k_psolve_tw) runs g_c = b_c - X_c g_{c-1} down to m-1 and h_c = b_c - X_c h_{c+1} up to m+1, then
// x_m = Y_m (b_m - X_m g_{m-1} - W_m h_{m+1}), then x_c = Y_c g_c - Z_c x_{c+1} (c < m) and
// x_c = Y_c h_c - Z_c x_{c-1} (c > m).  Inverses by Gauss-Jordan with partial pivoting.
int twisted_factors(const std::vector<double2> &AB, int nq, int n, int band, std::vector<double2> &out,
                    int *mout) {
    using cd = std::complex<double>;
    constexpr int C = 16;
    const int nch = (n + C - 1) / C, w = 2 * band + 1, m = (nch - 1) / 2;
    struct Mat {
        cd a[C][C];
    };
    auto zero = [] {
        Mat r;
        for (auto &row : r.a)
            for (auto &v : row) v = 0.0;
        return r;
    };
    auto mul = [&](const Mat &x, const Mat &y) {
        Mat r = zero();
        for (int i = 0; i < C; i++)
            for (int k = 0; k < C; k++) {
                const cd xv = x.a[i][k];
                if (xv == 0.0) continue;
                for (int j = 0; j < C; j++) r.a[i][j] += xv * y.a[k][j];
            }
        return r;
    };
    auto sub = [](Mat x, const Mat &y) {
        for (int i = 0; i < C; i++)
            for (int j = 0; j < C; j++) x.a[i][j] -= y.a[i][j];
        return x;
    };
    bool singular = false;
    auto inv = [&](const Mat &x) {
        cd g[C][2 * C];
        for (int i = 0; i < C; i++)
            for (int j = 0; j < 2 * C; j++) g[i][j] = j < C ? x.a[i][j] : cd(i == j - C ? 1.0 : 0.0);
        for (int k = 0; k < C; k++) {
            int p = k;
            for (int i = k + 1; i < C; i++)
                if (std::abs(g[i][k]) > std::abs(g[p][k])) p = i;
            if (g[p][k] == 0.0) {
                singular = true;
                return zero();
            }
            if (p != k)
                for (int j = 0; j < 2 * C; j++) std::swap(g[p][j], g[k][j]);
            const cd d = 1.0 / g[k][k];
            for (int j = 0; j < 2 * C; j++) g[k][j] *= d;
            for (int i = 0; i < C; i++)
                if (i != k && g[i][k] != 0.0) {
                    const cd f = g[i][k];
                    for (int j = 0; j < 2 * C; j++) g[i][j] -= f * g[k][j];
                }
        }
        Mat r;
        for (int i = 0; i < C; i++)
            for (int j = 0; j < C; j++) r.a[i][j] = g[i][C + j];
        return r;
    };
    out.assign((size_t)nq * nch * 4 * C * C, make_double2(0, 0));
    std::vector<Mat> D(nch), A(nch), B(nch), Si(nch), Ti(nch);
    for (int q = 0; q < nq; q++) {
        // entry (gi, gj) of block q, identity beyond n
        auto at = [&](int gi, int gj) -> cd {
            if (gi >= n || gj >= n) return cd(gi == gj ? 1.0 : 0.0);
            if (gi < 0 || gj < 0 || std::abs(gi - gj) > band) return 0.0;
            const double2 v = AB[((size_t)q * n + gi) * w + (gj - gi + band)];
            return cd(v.x, v.y);
        };
        for (int c = 0; c < nch; c++)
            for (int i = 0; i < C; i++)
                for (int j = 0; j < C; j++) {
                    D[c].a[i][j] = at(c * C + i, c * C + j);
                    A[c].a[i][j] = c > 0 ? at(c * C + i, (c - 1) * C + j) : 0.0;
                    B[c].a[i][j] = c + 1 < nch ? at(c * C + i, (c + 1) * C + j) : 0.0;
                }
        auto put = [&](int c, int slot, const Mat &x) {
            double2 *o = out.data() + (((size_t)q * nch + c) * 4 + slot) * C * C;
            for (int i = 0; i < C; i++)
                for (int t = 0; t < C; t++) o[t * C + i] = make_double2(x.a[i][t].real(), x.a[i][t].imag());
        };
        for (int c = 0; c < m; c++) {
            if (c > 0) {
                const Mat X = mul(A[c], Si[c - 1]);
                put(c, 0, X);
                Si[c] = inv(sub(D[c], mul(X, B[c - 1])));
            } else {
                Si[c] = inv(D[c]);
            }
        }
        for (int c = nch - 1; c > m; c--) {
            if (c < nch - 1) {
                const Mat X = mul(B[c], Ti[c + 1]);
                put(c, 0, X);
                Ti[c] = inv(sub(D[c], mul(X, A[c + 1])));
            } else {
                Ti[c] = inv(D[c]);
            }
        }
        Mat M = D[m];
        if (m >= 1) {
            const Mat X = mul(A[m], Si[m - 1]);
            put(m, 0, X);
            M = sub(M, mul(X, B[m - 1]));
        }
        if (m + 1 < nch) {
            const Mat Wm = mul(B[m], Ti[m + 1]);
            put(m, 3, Wm);
            M = sub(M, mul(Wm, A[m + 1]));
        }
        put(m, 1, inv(M));
        for (int c = 0; c < nch; c++) {
            if (c < m) {
                put(c, 1, Si[c]);
                put(c, 2, mul(Si[c], B[c]));
            } else if (c > m) {
                put(c, 1, Ti[c]);
                put(c, 2, mul(Ti[c], A[c]));
            }
        }
        if (singular) {
            gade::set_error("preconditioner: singular level block in the twisted factorization");
            return GADE_EPIVOT;
        }
    }
    *mout = m;
    return GADE_OK;
}

int blocklu_build(int nblk, int n, const int *rAp, const int *rAi, int nq, const int *blk_q,
                  const std::vector<double2> &Pq, hipStream_t stream, gade_blocklu **out, int min_ch) {
    const int nnzb = rAp[n];
    // interleave the two radial halves so that banded quadrants form one band
    std::vector<int> ilv(n);
    const int inter = (n % 2 == 0) ? 1 : 0;
    for (int i = 0; i < n; i++) ilv[i] = inter ? (i < n / 2 ? 2 * i : 2 * (i - n / 2) + 1) : i;
    int band = 0;
    for (int a = 0; a < n; a++)
        for (int k = rAp[a]; k < rAp[a + 1]; k++) band = std::max(band, std::abs(ilv[a] - ilv[rAi[k]]));
    if (band > gade::kMaxBand) {
        gade::set_error("preconditioner: radial half-bandwidth " + std::to_string(band) + " exceeds 63");
        return GADE_EINVAL;
    }
    gade_blocklu *lu = new gade_blocklu();
    auto fail = [&](int rc) {
        blocklu_free(lu);
        return rc;
    };
    lu->n = n;
    lu->nblk = nblk;
    lu->nq = nq;
    lu->ilv = inter;
    lu->band = std::max(band, 1);
    lu->ch = lu->band < 16 ? 16 : (lu->band < 32 ? 32 : 64);
    if (min_ch > lu->ch) lu->ch = min_ch >= 64 ? 64 : (min_ch >= 32 ? 32 : 16);
    const int CH = lu->ch;
    lu->nch = (n + CH - 1) / CH;
    const int w = 2 * lu->band + 1;
    std::vector<double2> AB((size_t)nq * n * w, make_double2(0, 0));
    for (int q = 0; q < nq; q++)
        for (int a = 0; a < n; a++)
            for (int k = rAp[a]; k < rAp[a + 1]; k++) {
                const int i = ilv[a], j = ilv[rAi[k]];
                AB[((size_t)q * n + i) * w + (j - i + lu->band)] = Pq[(size_t)q * nnzb + k];
            }
    int rc;
    if ((rc = upload(&lu->d_AB, AB.data(), AB.size()))) return fail(rc);
    if (CH == 16 && lu->band <= 16) {
        std::vector<double2> tw;
        if ((rc = twisted_factors(AB, nq, n, lu->band, tw, &lu->tw_m))) return fail(rc);
        if ((rc = upload(&lu->d_TW, tw.data(), tw.size()))) return fail(rc);
    }
    const size_t cs = (size_t)CH * CH;
    if ((rc = hip_alloc((void **)&lu->d_Linv, sizeof(double2) * nq * lu->nch * cs)) ||
        (rc = hip_alloc((void **)&lu->d_Uinv, sizeof(double2) * nq * lu->nch * cs)) ||
        (rc = hip_alloc((void **)&lu->d_Loff, sizeof(double2) * nq * lu->nch * cs)) ||
        (rc = hip_alloc((void **)&lu->d_Uoff, sizeof(double2) * nq * lu->nch * cs)) ||
        (rc = hip_alloc((void **)&lu->d_flag, sizeof(int))))
        return fail(rc);
    if (hipMemsetAsync(lu->d_flag, 0, sizeof(int), stream) != hipSuccess) return fail(GADE_EHIP);
    hipLaunchKernelGGL(k_band_lu, dim3(nq), dim3(256), 0, stream, lu->d_AB, n, lu->band, lu->d_flag);
#define INV(C) \
    case C: hipLaunchKernelGGL(k_chunk_inverse<C>, dim3(nq, lu->nch), dim3(C), 0, stream, lu->d_AB, n, lu->band, lu->nch, lu->d_Linv, lu->d_Uinv, lu->d_Loff, lu->d_Uoff); break;
    switch (CH) { INV(16) INV(32) INV(64) }
#undef INV
    int flag = 0;
    hipError_t e = hipGetLastError();
    if (e == hipSuccess) e = hipMemcpyAsync(&flag, lu->d_flag, sizeof(int), hipMemcpyDeviceToHost, stream);
    if (e == hipSuccess) e = hipStreamSynchronize(stream);
    if (e != hipSuccess) {
        gade::set_error(std::string("block LU: ") + hipGetErrorString(e));
        return fail(GADE_EHIP);
    }
    if (flag) {
        gade::set_error("preconditioner: zero pivot in block LU");
        return fail(GADE_EPIVOT);
    }
    // waves of 64/CH blocks, blocks of one distinct factor together so their loads coincide
    const int G = 64 / CH;
    std::vector<int> wb;
    for (int q = 0; q < nq; q++) {
        std::vector<int> rows;
        for (int r = 0; r < nblk; r++)
            if (blk_q[r] == q) rows.push_back(r);
        for (size_t i = 0; i < rows.size(); i += G)
            for (int j = 0; j < G; j++) wb.push_back(i + j < rows.size() ? rows[i + j] : -1);
    }
    lu->npw = (int)wb.size() / G;
    // one block per wave: the blocks of one distinct factor go to one XCD (workgroups are dealt to the 8
    // XCDs round-robin), so each XCD's L2 holds the factors of its own kappa groups only
    constexpr int kXcd = 8;
    std::vector<std::vector<int>> bins(kXcd);
    {
        std::vector<std::vector<int>> groups(nq);
        for (int r = 0; r < nblk; r++) groups[blk_q[r]].push_back(r);
        std::vector<int> gi(nq);
        for (int q = 0; q < nq; q++) gi[q] = q;
        std::stable_sort(gi.begin(), gi.end(), [&](int a, int b) { return groups[a].size() > groups[b].size(); });
        for (int q : gi) {
            int best = 0;
            for (int x = 1; x < kXcd; x++)
                if (bins[x].size() < bins[best].size()) best = x;
            bins[best].insert(bins[best].end(), groups[q].begin(), groups[q].end());
        }
    }
    size_t maxlen = 0;
    for (auto &b : bins) maxlen = std::max(maxlen, b.size());
    std::vector<int> order(maxlen * kXcd, -1);
    for (int x = 0; x < kXcd; x++)
        for (size_t k = 0; k < bins[x].size(); k++) order[k * kXcd + x] = bins[x][k];
    lu->norder = (int)order.size();
    if ((rc = upload(&lu->d_wblocks, wb.data(), wb.size())) || (rc = upload(&lu->d_blk_q, blk_q, (size_t)nblk)) ||
        (rc = upload(&lu->d_order, order.data(), order.size())))
        return fail(rc);
    *out = lu;
    return GADE_OK;
}

int blocklu_solve_async(const gade_blocklu *lu, const double2 *b, double2 *x, hipStream_t stream, const int *gate,
                        int gstride, int nblk1) {
    if (!lu) {
        gade::set_error("preconditioner not set up");
        return GADE_EINVAL;
    }
    if (lu->d_TW) {
        // twisted factorization: one block per workgroup of two waves, meeting at level tw_m
        const size_t sh = sizeof(double2) * (size_t)lu->nch * 16;
        hipLaunchKernelGGL((k_psolve_tw<4, double2>), dim3(lu->norder), dim3(128), sh, stream, b, x, lu->d_order,
                           lu->d_blk_q, lu->ilv, lu->d_TW, lu->n, lu->nch, lu->tw_m, gate, gstride, nblk1);
        GADE_HIP(hipGetLastError());
        return GADE_OK;
    }
    const bool quad = (lu->ch == 16 || lu->ch == 32) && lu->band <= 16;
    if (!quad && gstride != 0) {
        gade::set_error("gated (batched) preconditioner solves need 16- or 32-row levels and a band of at most 16");
        return GADE_EINVAL;
    }
    if (quad) {
        // 32-row levels (16-row ones with this band have the twisted factors above): one wave per block, two
        // lanes per row
        const size_t shq = sizeof(double2) * ((size_t)lu->nch * lu->ch + lu->ch);
        hipLaunchKernelGGL((k_psolve_q<2, double2, 32, 2>), dim3(lu->norder), dim3(64), shq, stream, b, x,
                           lu->d_order, lu->d_blk_q, lu->ilv, lu->d_Linv, lu->d_Uinv, lu->d_Loff, lu->d_Uoff, lu->n,
                           lu->band, lu->nch, gate, gstride, nblk1);
        GADE_HIP(hipGetLastError());
        return GADE_OK;
    }
    const int G = 64 / lu->ch;
    size_t shm = sizeof(double2) * (size_t)G * ((size_t)lu->nch * lu->ch + lu->ch);
#define PS(C) \
    case C: hipLaunchKernelGGL(k_psolve<C>, dim3(lu->npw), dim3(64), shm, stream, b, x, lu->d_wblocks, lu->d_blk_q, lu->ilv, lu->d_Linv, lu->d_Uinv, lu->d_Loff, lu->d_Uoff, lu->n, lu->band, lu->nch, gate); break;
    switch (lu->ch) { PS(16) PS(32) PS(64) default: return GADE_EINVAL; }
#undef PS
    GADE_HIP(hipGetLastError());
    return GADE_OK;
}

}  // namespace gade
