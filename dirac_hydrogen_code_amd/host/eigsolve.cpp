// Dense real symmetric-definite generalized eigenproblem A v = E B v (B positive definite), the per-kappa
// radial problem H0(k) v = E S(k) v of the populations and of the evl / evc files (the reference solves it with
// Eigen's GeneralizedSelfAdjointEigenSolver on the real parts, src/hamiltonian.h:872-943).
//
// Method: Cholesky B = L L^T; C = L^{-1} A L^{-T}; Householder reduction of C to tridiagonal form with the
// transformation accumulated; implicit QL iteration with Wilkinson-type shifts on the tridiagonal matrix
// (Bowdler, Martin, Reinsch and Wilkinson, Numer. Math. 11 (1968) 293-306; Golub and Van Loan, Matrix
// Computations, sections 8.3.1-8.3.3); eigenvectors v = L^{-T} y, so that v^T B v = 1.  Eigenvalues ascending;
// each eigenvector's largest-magnitude component (the first, on a tie) is made positive, so the output does not
// depend on the order of the iteration's sign choices.  Setup only: it runs once per model.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <limits>
#include <numeric>
#include <condition_variable>
#include <mutex>
#include <thread>
#include <vector>

#include "gade.h"

namespace {

// Householder tridiagonalization of the symmetric n x n matrix C (row-major, overwritten); Q (row-major)
// receives the orthogonal transformation with Q^T C Q = T; d, e the diagonal and subdiagonal (e[i] = T[i+1][i]).
void tridiagonalize(int n, std::vector<double> &C, std::vector<double> &Q, std::vector<double> &d,
                    std::vector<double> &e) {
    Q.assign((size_t)n * n, 0.0);
    for (int i = 0; i < n; i++) Q[(size_t)i * n + i] = 1.0;
    std::vector<double> v(n), p(n), w(n);
    for (int k = 0; k + 2 < n; k++) {
        // x = C[k+1:n][k]; v = x + sign(x0) |x| e1; beta = 2 / v^T v
        double alpha = 0;
        for (int i = k + 1; i < n; i++) alpha += C[(size_t)i * n + k] * C[(size_t)i * n + k];
        alpha = std::sqrt(alpha);
        if (alpha == 0) continue;
        const double x0 = C[(size_t)(k + 1) * n + k];
        const double s = x0 >= 0 ? 1.0 : -1.0;
        for (int i = k + 1; i < n; i++) v[i] = C[(size_t)i * n + k];
        v[k + 1] += s * alpha;
        double vv = 0;
        for (int i = k + 1; i < n; i++) vv += v[i] * v[i];
        const double beta = 2.0 / vv;
        // p = beta C22 v, w = p - (beta p^T v / 2) v; C22 -= v w^T + w v^T
        for (int i = k + 1; i < n; i++) {
            double acc = 0;
            const double *ci = &C[(size_t)i * n];
            for (int j = k + 1; j < n; j++) acc += ci[j] * v[j];
            p[i] = beta * acc;
        }
        double pv = 0;
        for (int i = k + 1; i < n; i++) pv += p[i] * v[i];
        const double half = 0.5 * beta * pv;
        for (int i = k + 1; i < n; i++) w[i] = p[i] - half * v[i];
        for (int i = k + 1; i < n; i++) {
            double *ci = &C[(size_t)i * n];
            for (int j = k + 1; j < n; j++) ci[j] -= v[i] * w[j] + w[i] * v[j];
        }
        // the reflected column: C[k+1][k] = -s alpha, zeros below
        C[(size_t)(k + 1) * n + k] = C[(size_t)k * n + k + 1] = -s * alpha;
        for (int i = k + 2; i < n; i++) C[(size_t)i * n + k] = C[(size_t)k * n + i] = 0;
        // Q <- Q (I - beta v v^T) on columns k+1..n-1
        for (int r = 0; r < n; r++) {
            double *qr = &Q[(size_t)r * n];
            double acc = 0;
            for (int j = k + 1; j < n; j++) acc += qr[j] * v[j];
            acc *= beta;
            for (int j = k + 1; j < n; j++) qr[j] -= acc * v[j];
        }
    }
    d.resize(n);
    e.assign(n, 0.0);
    for (int i = 0; i < n; i++) d[i] = C[(size_t)i * n + i];
    for (int i = 0; i + 1 < n; i++) e[i] = C[(size_t)(i + 1) * n + i];
}

// Implicit QL iteration on the symmetric tridiagonal (d, e); the plane rotations are applied to the rows of
// Zt (Zt[j][k] = component k of vector j), which then holds the eigenvectors as rows.  Returns false when an
// eigenvalue does not converge.
bool tridiagonal_ql(int n, std::vector<double> &d, std::vector<double> &e, std::vector<double> &Zt) {
    const double eps = std::numeric_limits<double>::epsilon();
    for (int l = 0; l < n; l++) {
        int iter = 0;
        while (true) {
            int m = l;
            for (; m + 1 < n; m++) {
                const double dd = std::fabs(d[m]) + std::fabs(d[m + 1]);
                if (std::fabs(e[m]) <= eps * dd) break;
            }
            if (m == l) break;
            if (++iter > 64) return false;
            // shift: the eigenvalue of the leading 2x2 block closer to d[l]
            double g = (d[l + 1] - d[l]) / (2.0 * e[l]);
            double r = std::hypot(g, 1.0);
            g = d[m] - d[l] + e[l] / (g + (g >= 0 ? r : -r));
            double s = 1.0, c = 1.0, p = 0.0;
            int i = m - 1;
            bool deflated = false;
            for (; i >= l; i--) {
                const double f = s * e[i], b = c * e[i];
                r = std::hypot(f, g);
                e[i + 1] = r;
                if (r == 0.0) {  // underflow: the block splits here
                    d[i + 1] -= p;
                    e[m] = 0.0;
                    deflated = true;
                    break;
                }
                s = f / r;
                c = g / r;
                g = d[i + 1] - p;
                r = (d[i] - g) * s + 2.0 * c * b;
                p = s * r;
                d[i + 1] = g + p;
                g = c * r - b;
                double *zi = &Zt[(size_t)i * n], *zi1 = &Zt[(size_t)(i + 1) * n];
                for (int k = 0; k < n; k++) {
                    const double zf = zi1[k];
                    zi1[k] = s * zi[k] + c * zf;
                    zi[k] = c * zi[k] - s * zf;
                }
            }
            if (deflated) continue;
            d[l] -= p;
            e[l] = g;
            e[m] = 0.0;
        }
    }
    return true;
}

}  // namespace

extern "C" int gade_eig_dense(int n, const double *A, const double *B, double *E, double *V) {
    if (n <= 0 || !A || !B || !E || !V) return GADE_EINVAL;
    const size_t nn = (size_t)n * n;
    // Cholesky B = L L^T (lower, row-major)
    std::vector<double> L(nn, 0.0);
    for (int j = 0; j < n; j++) {
        double djj = B[(size_t)j * n + j];
        for (int k = 0; k < j; k++) djj -= L[(size_t)j * n + k] * L[(size_t)j * n + k];
        if (!(djj > 0)) return GADE_EINVAL;  // not positive definite
        const double ljj = std::sqrt(djj);
        L[(size_t)j * n + j] = ljj;
        for (int i = j + 1; i < n; i++) {
            double a = B[(size_t)i * n + j];
            for (int k = 0; k < j; k++) a -= L[(size_t)i * n + k] * L[(size_t)j * n + k];
            L[(size_t)i * n + j] = a / ljj;
        }
    }
    // W = L^{-1} A (forward substitution, row operations), then C = L^{-1} W^T
    auto lower_solve = [&](const std::vector<double> &R, std::vector<double> &X) {
        X = R;
        for (int i = 0; i < n; i++) {
            double *xi = &X[(size_t)i * n];
            for (int k = 0; k < i; k++) {
                const double lik = L[(size_t)i * n + k];
                if (lik == 0) continue;
                const double *xk = &X[(size_t)k * n];
                for (int c = 0; c < n; c++) xi[c] -= lik * xk[c];
            }
            const double inv = 1.0 / L[(size_t)i * n + i];
            for (int c = 0; c < n; c++) xi[c] *= inv;
        }
    };
    std::vector<double> Am(A, A + nn), W, Wt(nn), C;
    lower_solve(Am, W);
    for (int i = 0; i < n; i++)
        for (int j = 0; j < n; j++) Wt[(size_t)i * n + j] = W[(size_t)j * n + i];
    lower_solve(Wt, C);
    for (int i = 0; i < n; i++)
        for (int j = 0; j < i; j++) {
            const double s = 0.5 * (C[(size_t)i * n + j] + C[(size_t)j * n + i]);
            C[(size_t)i * n + j] = C[(size_t)j * n + i] = s;
        }
    std::vector<double> Q, d, e;
    tridiagonalize(n, C, Q, d, e);
    std::vector<double> Zt(nn);  // eigenvectors of C as rows
    for (int i = 0; i < n; i++)
        for (int j = 0; j < n; j++) Zt[(size_t)j * n + i] = Q[(size_t)i * n + j];
    if (!tridiagonal_ql(n, d, e, Zt)) return GADE_ENOTCONV;
    // v = L^{-T} y for every eigenvector y: back substitution with L^T, rows of Y indexed by component
    std::vector<double> Y(nn);
    for (int i = 0; i < n; i++)
        for (int j = 0; j < n; j++) Y[(size_t)i * n + j] = Zt[(size_t)j * n + i];
    for (int i = n - 1; i >= 0; i--) {
        double *yi = &Y[(size_t)i * n];
        for (int k = i + 1; k < n; k++) {
            const double lki = L[(size_t)k * n + i];
            if (lki == 0) continue;
            const double *yk = &Y[(size_t)k * n];
            for (int c = 0; c < n; c++) yi[c] -= lki * yk[c];
        }
        const double inv = 1.0 / L[(size_t)i * n + i];
        for (int c = 0; c < n; c++) yi[c] *= inv;
    }
    std::vector<int> order(n);
    std::iota(order.begin(), order.end(), 0);
    std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return d[a] < d[b]; });
    for (int j = 0; j < n; j++) {
        const int src = order[j];
        E[j] = d[src];
        int imax = 0;
        for (int i = 1; i < n; i++)
            if (std::fabs(Y[(size_t)i * n + src]) > std::fabs(Y[(size_t)imax * n + src])) imax = i;
        const double sg = Y[(size_t)imax * n + src] < 0 ? -1.0 : 1.0;
        for (int i = 0; i < n; i++) V[(size_t)i * n + j] = sg * Y[(size_t)i * n + src];
    }
    return GADE_OK;
}

// ---------------------------------------------------------------------------------------------------------
// Singular values and left singular vectors of a complex k x k matrix R (column-major), for the POD of the
// learned initial guess: one-sided (Hestenes) Jacobi on the columns of A = R^H, A <- A G_pq, with the same
// rotations accumulated in J (J = I at the start).  At convergence R^H J = W has orthogonal columns, so
// R = J W^H: the singular values are the column norms of W and the left singular vectors the columns of J
// (Demmel and Veselic, SIAM J. Matrix Anal. Appl. 13 (1992) 1204-1245: high relative accuracy, also for
// the small singular values of graded R).  Rotation of the pair (p, q): gamma = a_p^H a_q = |gamma| e,
// zeta = (|a_q|^2 - |a_p|^2) / (2 |gamma|), t = sign(zeta) / (|zeta| + sqrt(1 + zeta^2)), c = 1/sqrt(1 + t^2),
// s = c t; a_p <- c a_p - s conj(e) a_q, a_q <- s a_p + c conj(e) a_q.  Pairs are visited in round-robin
// (tournament) order: the k/2 pairs of a round touch disjoint columns and run on host threads; every thread
// count gives the same result.  Sweeps end when no pair has |gamma| > 4 eps |a_p| |a_q|.
// Output: S [k] descending, U [k][k] column-major (column j for S[j]).  GADE_ENOTCONV after 60 sweeps.
namespace gade {
int svd_left_jacobi(int k, const double *Rcm, double *S, double *Ucm) {
    if (k <= 0 || !Rcm || !S || !Ucm) return GADE_EINVAL;
    const int kp = k + (k & 1);  // an odd order gets a zero column that never rotates
    std::vector<double> A((size_t)2 * kp * k, 0.0), J((size_t)2 * kp * k, 0.0);
    auto a = [&](int col) { return &A[(size_t)2 * k * col]; };
    auto jc = [&](int col) { return &J[(size_t)2 * k * col]; };
    for (int c = 0; c < k; c++)  // A[:, c] = conj(R[c, :])
        for (int r = 0; r < k; r++) {
            const double *z = Rcm + 2 * ((size_t)r * k + c);
            a(c)[2 * r] = z[0];
            a(c)[2 * r + 1] = -z[1];
        }
    for (int c = 0; c < k; c++) jc(c)[2 * c] = 1.0;
    const double tol = 4 * std::numeric_limits<double>::epsilon();
    auto rotate = [&](int p, int q) -> int {
        if (p >= k || q >= k) return 0;
        double al = 0, be = 0, gr = 0, gi = 0;
        const double *x = a(p), *y = a(q);
        for (int r = 0; r < k; r++) {
            const double xr = x[2 * r], xi = x[2 * r + 1], yr = y[2 * r], yi = y[2 * r + 1];
            al += xr * xr + xi * xi;
            be += yr * yr + yi * yi;
            gr += xr * yr + xi * yi;  // conj(x) y
            gi += xr * yi - xi * yr;
        }
        const double g = std::hypot(gr, gi);
        if (g == 0 || al == 0 || be == 0 || g <= tol * std::sqrt(al) * std::sqrt(be)) return 0;
        const double er = gr / g, ei = -gi / g;  // conj(e)
        const double z = (be - al) / (2 * g);
        const double t = (z >= 0 ? 1.0 : -1.0) / (std::fabs(z) + std::sqrt(1 + z * z));
        const double c = 1 / std::sqrt(1 + t * t), s = c * t;
        for (double *M : {&A[0], &J[0]}) {
            double *u = M + (size_t)2 * k * p, *v = M + (size_t)2 * k * q;
            for (int r = 0; r < k; r++) {
                const double ur = u[2 * r], ui = u[2 * r + 1];
                const double wr = er * v[2 * r] - ei * v[2 * r + 1], wi = er * v[2 * r + 1] + ei * v[2 * r];
                u[2 * r] = c * ur - s * wr;
                u[2 * r + 1] = c * ui - s * wi;
                v[2 * r] = s * ur + c * wr;
                v[2 * r + 1] = s * ui + c * wi;
            }
        }
        return 1;
    };
    const int npair = kp / 2;
    const int nth = (int)std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
    // circle method: player 0 fixed, the others rotate; round rd pairs slot i with slot kp - 1 - i
    auto pair_of = [&](int rd, int i, int &p, int &q) {
        auto player = [&](int slot) { return slot == 0 ? 0 : 1 + (slot - 1 + rd) % (kp - 1); };
        p = player(i);
        q = player(kp - 1 - i);
        if (p > q) std::swap(p, q);
    };
    // the threads stay up for the whole iteration and meet at a barrier after every round
    const int T = npair >= 2 * nth ? nth : 1;
    std::mutex mu;
    std::condition_variable cv;
    int arrived = 0, generation = 0;
    auto barrier = [&]() {
        std::unique_lock<std::mutex> lk(mu);
        const int g = generation;
        if (++arrived == T) {
            arrived = 0;
            generation++;
            cv.notify_all();
        } else {
            cv.wait(lk, [&] { return generation != g; });
        }
    };
    std::vector<int> rot(T, 0);
    bool done = false;
    auto run = [&](int tid) {
        for (int sw = 0; sw < 60; sw++) {
            rot[tid] = 0;
            for (int rd = 0; rd < kp - 1; rd++) {
                for (int i = tid; i < npair; i += T) {
                    int p, q;
                    pair_of(rd, i, p, q);
                    rot[tid] += rotate(p, q);
                }
                barrier();
            }
            if (tid == 0) {
                int total = 0;
                for (int v : rot) total += v;
                done = total == 0;
            }
            barrier();  // every thread reads 'done' after thread 0 wrote it
            if (done) return;
            barrier();  // and before thread 0 clears the counts of the next sweep
        }
    };
    std::vector<std::thread> th;
    for (int t = 1; t < T; t++) th.emplace_back(run, t);
    run(0);
    for (auto &x : th) x.join();
    if (!done) return GADE_ENOTCONV;
    std::vector<double> nrm(k);
    for (int c = 0; c < k; c++) {
        double s2 = 0;
        for (int r = 0; r < 2 * k; r++) s2 += a(c)[r] * a(c)[r];
        nrm[c] = std::sqrt(s2);
    }
    std::vector<int> order(k);
    std::iota(order.begin(), order.end(), 0);
    std::stable_sort(order.begin(), order.end(), [&](int x, int y) { return nrm[x] > nrm[y]; });
    for (int j = 0; j < k; j++) {
        S[j] = nrm[order[j]];
        memcpy(Ucm + (size_t)2 * k * j, jc(order[j]), sizeof(double) * 2 * k);
    }
    return GADE_OK;
}
}  // namespace gade
