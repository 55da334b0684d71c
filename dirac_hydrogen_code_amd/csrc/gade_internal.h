// Internal definitions shared by the libgade translation units.
#pragma once

#include <hip/hip_runtime.h>

#include <map>
#include <string>
#include <vector>

#include "gade.h"

namespace gade {

constexpr double kSoL = 137.035999084;  // tools/solver_common.h:37
constexpr int kChunk = 64;              // rows per triangular-solve chunk (one wave)
constexpr int kMaxL = 16;               // largest lmax the angular kernel is instantiated for
constexpr int kMaxBand = 63;            // largest half-bandwidth the chunked solve supports

void set_error(const std::string &msg);

#define GADE_HIP(expr)                                                                      \
    do {                                                                                    \
        hipError_t e_ = (expr);                                                             \
        if (e_ != hipSuccess) {                                                             \
            gade::set_error(std::string(#expr) + ": " + hipGetErrorString(e_));            \
            return GADE_EHIP;                                                               \
        }                                                                                   \
    } while (0)

#define GADE_TRY(expr)                  \
    do {                                \
        int rc_ = (expr);               \
        if (rc_ != GADE_OK) return rc_; \
    } while (0)

struct Timer {
    double total_ms = 0;
    long long count = 0;
};

// one kernel-family launch bracketed by two events, resolved when queried
struct PendingEvent {
    std::string name;
    hipEvent_t a, b;
};

}  // namespace gade

// Block-diagonal preconditioner: exact banded LU of each distinct diagonal
// block (interleaved radial order) with chunk-level-scheduled triangular solves.
struct gade_blocklu {
    int n = 0, nblk = 0, nq = 0, band = 0, ch = 16, nch = 0, npw = 0, ilv = 0;
    double2 *d_AB = nullptr, *d_Linv = nullptr, *d_Uinv = nullptr, *d_Loff = nullptr, *d_Uoff = nullptr;
    int *d_blk_q = nullptr, *d_wblocks = nullptr, *d_flag = nullptr;
    int *d_order = nullptr;  // one block per wave, -1 pads; blocks of one factor on one XCD
    int norder = 0;
    // twisted factorization over 16-row levels (k_psolve_tw): [q][level][X, Y, Z, W][16 x 16, column-major],
    // meeting level tw_m; null when the levels are not 16 rows or the band exceeds 16
    double2 *d_TW = nullptr;
    int tw_m = 0;
};

struct gade_lg;  // learned initial guess (learned.hip)

namespace gade {
// Pq: values of each distinct block in the CSR order of (rAp, rAi), [nq][nnz]
// min_ch: smallest level height (16 or 32 rows); taller levels when the band needs them
int blocklu_build(int nblk, int n, const int *rAp, const int *rAi, int nq, const int *blk_q,
                  const std::vector<double2> &Pq, hipStream_t stream, gade_blocklu **out, int min_ch = 16);
int blocklu_solve_async(const gade_blocklu *lu, const double2 *b, double2 *x, hipStream_t stream,
                        const int *gate = nullptr, int gstride = 0, int nblk1 = 1);
void blocklu_free(gade_blocklu *lu);
int hip_alloc(void **p, size_t bytes);
}  // namespace gade

namespace gade {
void lg_free(gade_lg *lg);
}

struct gade_model {
    // sizes.  A model holds B instances (pulses) that share the couplings: rows are
    // instance-major (row b*nblk1 + r), nblk = B*nblk1 rows, N = B*N1 entries.
    int nblk = 0, n = 0, lmax = 0, nnzb = 0, ellw = 0, hnnz = 0, nq = 0;
    long long N = 0;
    int B = 1, nblk1 = 0;
    long long N1 = 0;
    double h = 1, dt = 0;
    hipStream_t stream = nullptr;

    // host-side structure kept for setup
    std::vector<int> rad_Ap, rad_Ai, hall_Ap, hall_Ai, kap_blk, kap_q;  // kap_q: kappa of each distinct index
    std::vector<int> blk_q;   // distinct-kappa index of each block
    std::vector<double> ft;   // last f(t) of every instance, [B][12] doubles

    // radial pattern (ELL, transposed: [slot][row])
    int *d_ell_col = nullptr, *d_ell_src = nullptr, *d_tmap = nullptr, *d_ell_pos = nullptr;
    // per 64-row radial chunk: distinct columns (LDS window) and ELL columns local to it
    int nchunk = 0, maxloc = 0, maxrow = 0;
    int *d_ell_lcol = nullptr, *d_locj = nullptr, *d_nloc = nullptr;
    double2 *d_g = nullptr;        // [24*lmax][nnzb]
    double2 *d_s = nullptr;        // [3][nnzb]
    double2 *d_h0 = nullptr;       // [4][nnzb]
    double2 *d_GE = nullptr;       // [lmax][8][ellw][n]
    double2 *d_D = nullptr;        // [3][nq][ellw][n]: S+ihdtH0, S-ihdtH0, S
    // Hall and angular coefficients
    int *d_hAp = nullptr, *d_hAi = nullptr, *d_kap = nullptr, *d_blk_q = nullptr;
    double2 *d_coef = nullptr;     // [hnnz][lmax][2]
    // the same couplings as lists (one instance's rows): per (row, multipole, part 0 H0l / 1 H1l) the
    // non-zero coefficients in increasing column order; entries (column, kappa) and values
    std::vector<int> h_aoff;       // [nblk1 * lmax * 2 + 1]
    int *d_aoff = nullptr;
    int2 *d_ack = nullptr;
    double2 *d_acf = nullptr;
    float2 *d_acf32 = nullptr;     // FP32 copy (mixed precision)
    int ang_list = 1;              // phase A from the coupling lists (1) or from the Hall rows (0)
    int *d_diag = nullptr;         // [hnnz] 1 when orig_row == orig_col
    // kappa-grouped row tiles for the radial phase
    int tile_r = 2, ntiles = 0;
    int *d_tiles = nullptr;        // [ntiles][tile_r]
    int *d_tile_q = nullptr;       // [ntiles]
    double *d_tile_k = nullptr;    // [ntiles]
    // banded (DIA) radial phase: quadrant/diagonal-major coefficients, tiles of 8 rows of one kappa
    bool dia_ok = false;           // radial pattern is a 2x2 block of bands of half-width <= 7
    int radial_kind = 0;           // 1: ELL kernel, 2: DIA kernel (active)
    int nh = 0, nhp = 0, nd = 0, ndtiles = 0;
    int *d_dia_src = nullptr;      // [4][nd][nhp] radial CSR index or -1
    double2 *d_gs = nullptr, *d_gts = nullptr;  // [lmax][3][4][4][nd][nhp] DIA-ordered couplings
    double2 *d_GD = nullptr;       // [B][lmax][4][4][nd][nhp][2] kappa-folded G-sum pairs
    double2 *d_DD = nullptr;       // [3][nq][4][nd][nhp] diagonal operators
    int *d_dtiles = nullptr, *d_dtile_q = nullptr;
    double *d_dtile_k = nullptr;
    int *d_dtile_b = nullptr;      // instance of each tile (tiles never mix instances)
    int dia_r = 8, dia_p = 1;      // banded tile shape: rows per tile, outputs per lane
    int dia_hw = 1;                // output halves per banded workgroup (2: both halves share each window)
    // FP64 banded radial phase on the matrix cores: waves of two 8-row halves of one kappa each
    bool mfma_ok = false;          // the banded pattern fits the MFMA kernel (set up)
    int radial_mfma = 0;           // 1: FP64 radial products on MFMA (k_radial_mfma), 0: k_radial_dia (default)
    int nmi = 0, nmg = 0;          // i-steps of 4 outputs per half; MFMA wave groups
    size_t gf_stride = 0;          // GF doubles per instance
    double *d_GF = nullptr;        // [B][lmax][4][4][nmi][5][16][4] G-sum pairs in MFMA fragment order
    double2 *d_DF = nullptr;       // [3][nq][4][nmi][5][16] diagonal operators in fragment order
    int *d_mg_rows = nullptr, *d_mg_q = nullptr, *d_mg_b = nullptr;  // [nmg][16], [nmg][2], [nmg]
    double *d_mg_k = nullptr;      // [nmg][2] kappa of each half
    int ang_lb = 2;                // multipoles per angular block
    int ang_fused = 0;             // 1: the banded radial kernel forms its V windows from x (no V in HBM)
    int psolve_ch = 16;            // preconditioner level height (16 or 32 rows)
    size_t gd_stride = 0;          // GD entries per instance
    double2 *d_V = nullptr;        // [lmax][4][nblk][n]
    double2 *d_ypart = nullptr;    // [lmax][nblk][n] per-multipole partial outputs of the radial phase
    double2 *d_ypart2 = nullptr;   // the same for the second output of a paired apply (lmax > 1)
    // preconditioner
    gade_blocklu *pre = nullptr;
    int *d_flag = nullptr;
    // BiCGSTAB workspace (persists across calls, tools/gpu_bicgstab.c:150-167)
    double2 *r = nullptr, *r0 = nullptr, *v = nullptr, *p = nullptr, *y = nullptr, *z = nullptr,
            *s = nullptr, *t = nullptr, *tmp = nullptr;
    double *d_sc = nullptr;        // device scalars [B][SC_STRIDE]
    double *d_part = nullptr;      // reduction partials [B][red_blocks][4]
    double *d_part2 = nullptr;     // the same, for w while x/r's reduction reuses d_part
    double *h_sc = nullptr;        // pinned host mirror
    int red_blocks = 0;            // reduction blocks per instance
    // assembled reference formulation
    bool asm_ready = false;
    long long nnz_full = 0;
    int *d_fAp = nullptr, *d_fAi = nullptr;
    double2 *d_fAx = nullptr;
    int *d_SAp = nullptr, *d_SAi = nullptr;
    double2 *d_SAx = nullptr;
    // eigen-projection: per distinct kappa, S-orthonormal eigenvectors (conjugate-transposed,
    // [q][i][j] = conj(v_j[i])) and energies [q][j] of H0(k) v = E S(k) v
    double2 *d_evec = nullptr;
    double *d_eval = nullptr;
    // device-driven CN stepping: control words, per-step outputs, launch gate
    int *d_ctl = nullptr;          // [B][CT_STRIDE] per-instance words
    int *h_ctl = nullptr;          // pinned mirror
    int *d_gctl = nullptr;         // [GC_COUNT] run-wide words
    int *h_gctl = nullptr;         // pinned mirror
    int *d_nsteps = nullptr;       // [B] steps of each instance's pulse
    int *d_iters = nullptr;        // [B][steps_cap]
    double *d_tols = nullptr;      // [B][steps_cap]
    int steps_cap = 0;
    const int *gate = nullptr;     // gate passed to every model launch (null: ungated)
    int gstride = 0;               // ints between the gate words of consecutive instances
    bool ctl_on = false;           // reductions update d_ctl
    bool device_loop = true;       // CN steps driven on the device (false: host loop per iteration)
    int window = 8;                // steps enqueued per host synchronization
    int iter_batch = 0;            // iterations enqueued per step (0: adaptive)
    int last_k = 2;                // adaptive iterations per step
    int guess_order = 0;           // initial guess: 0 psi_n (the reference), p: extrapolation of order p
    // per-step pulse coefficients of a window on the device ([slot][B] x 6 complex) and the step graph
    double2 *d_cfwin = nullptr, *h_cfwin = nullptr;
    int cfwin_cap = 0;             // slots
    bool use_graph = true;         // replay one captured hipGraph per CN step of a window
    hipStream_t cap_stream = nullptr;
    hipGraph_t graph = nullptr;
    hipGraphExec_t graph_exec = nullptr;
    const void *graph_key[12] = {};  // what the captured step depends on
    double2 *d_hist = nullptr;     // [kHistSlots][N] history ring of previous solutions (per instance)
    gade_lg *lg = nullptr;         // learned initial guess
    int lg_features = 0;           // compute the learned guess's inputs every step (training data)
    int lg_zmax = 8;               // learned guess: inputs beyond this many standard deviations take the fallback
    // mixed precision: FP32 inner BiCGSTAB with the FP32 operator and preconditioner, FP64 defect correction
    int precision = 0;             // 0: FP64 (the reference), 1: mixed
    int mp_inner_iters = 2;        // inner iterations per sweep
    double mp_inner_tol = 1e-6;    // inner relative tolerance
    bool mp_ready = false;
    float2 *d_V32 = nullptr, *d_ypart32 = nullptr, *d_GD32 = nullptr, *d_DD32 = nullptr, *d_coef32 = nullptr;
    float2 *lu32[4] = {};          // Linv, Uinv, Loff, Uoff
    float2 *tw32 = nullptr;        // FP32 copy of the twisted factors (when the preconditioner has them)
    float2 *mp_r = nullptr, *mp_r0 = nullptr, *mp_p = nullptr, *mp_v = nullptr, *mp_y = nullptr, *mp_z = nullptr,
           *mp_s = nullptr, *mp_t = nullptr, *mp_d = nullptr;
    double2 *d_mp_ax = nullptr;    // A x of the outer residual
    double *d_sc32 = nullptr;      // inner scalars [B][SC_STRIDE]
    // profiling: event pairs are recorded without synchronizing and resolved
    // by gade_profile_get; 'profile' is a mask of GADE_PROF_* families
    int profile = 0;
    std::map<std::string, gade::Timer> timers;
    std::vector<gade::PendingEvent> pending;
    std::vector<hipEvent_t> event_pool;
    int prof_period = 1;                       // time every prof_period-th launch of a family
    int prof_floor_ns = 0;                     // timed launches shorter than this are not counted
    std::map<std::string, long long> prof_seen;
};

namespace gade {
void drop_step_graph(gade_model *m);
int lg_prepare(gade_model *m);
int lg_reset(gade_model *m);  // residual history of the learned guess's inputs
void lg_state(gade_model *m, void **ptr, size_t *bytes);
int lg_step_prologue(gade_model *m, double2 *psi);
}  // namespace gade
