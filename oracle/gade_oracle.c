/*
 * gade_oracle.c -- CPU oracle for the GaDE Crank-Nicolson hot path.
 *
 * TEST INFRASTRUCTURE ONLY (see gade_oracle.h).  A plain-C restatement of the
 * reference's CPU algorithm, written for checking the HIP product path.
 *
 * Deliberate deviations from the reference CPU path, each also used by the
 * product so that both compute the same mathematics:
 *  - the preconditioner is an exact LU of each diagonal block of
 *    P = S + i h dt/2 H0 (tools/solver.c:460-485 hands P to SuperLU's ILU with
 *    drop tolerance 1e-6 and fill factor 2, tools/superlu.c:218-232); this
 *    changes iteration counts only;
 *  - the BiCGSTAB restart recomputes r = b - A x from a zeroed vector, as the
 *    reference's GPU solver does (tools/gpu_bicgstab.c:229-238); the CPU
 *    version accumulates into the stale r (tools/bicgstab.c:217-221);
 *  - tolerance and maximum iterations are parameters (hard-coded 1e-16 / 500
 *    at tools/dirac_solver.c:506-507) and the loop runs exactly nsteps steps.
 */
#include "gade_oracle.h"

#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <sys/time.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define SOL 137.035999084 /* tools/solver_common.h:37 */

/* ------------------------------------------------------------------ */
/* index maps and pulse                                                */
/* ------------------------------------------------------------------ */

/* tools/solver_common.c:30-37 */
int or_ik(int i)
{
    int blk = i / 4;
    int ak = (int)(0.5 * (sqrt(8.0 * blk + 1.0) - 1.0)) + 1;
    int smod = ak * 4 > 4 ? ak * 4 : 4;
    double sfloor = 2 * ak * (ak - 1);
    double q = (i - sfloor) / smod;
    int sgn = (q >= 0.5) - (q < 0.5);
    return ak * sgn;
}

/* tools/solver_common.c:39-44 */
double or_imu(int i)
{
    int ak = abs(or_ik(i));
    int kmod = 2 * ak > 2 ? 2 * ak : 2;
    return i % kmod - ak + 0.5;
}

/* tools/solver_common.c:46-65 (phase -pi/2, T = cycles 2 pi / omega).  The
   reference passes a complex argument to the real sin/cos of <math.h>, i.e.
   its real part, so every f_alpha is real with a +0 imaginary part. */
void or_pulse(double E0, double omega, double cycles, double t, ocplx out[6])
{
    const double pi = 3.14159265358979323846;
    double T = cycles * 2 * pi / omega;
    double phi = -pi / 2;
    out[0] = (-E0 / (4 * omega) * cos(phi + t * (2 * pi / T + omega)));
    out[1] = (-E0 / (4 * omega) * sin(phi + t * (2 * pi / T + omega)));
    out[2] = (-E0 / (4 * omega) * cos(-phi + t * (2 * pi / T - omega)));
    out[3] = (E0 / (4 * omega) * sin(-phi + t * (2 * pi / T - omega)));
    out[4] = (E0 / (2 * omega) * cos(phi + t * omega));
    out[5] = (E0 / (2 * omega) * sin(phi + t * omega));
}

/* tools/sorted_list.c:32-57: bisection down to 32 entries, then linear. */
int or_locate(const int *list, int n, int value)
{
    int lo = 0, hi = n;
    while (hi - lo > 32) {
        int mid = (lo + hi) / 2;
        if (list[mid] > value) hi = mid;
        else lo = mid;
    }
    while (lo < hi && list[lo] < value) lo++;
    return lo;
}

/* ------------------------------------------------------------------ */
/* CSR I/O and primitives                                              */
/* ------------------------------------------------------------------ */

/* tools/csr.c:262-318 */
int or_csr_read(const char *fname, or_csr_t *sp)
{
    memset(sp, 0, sizeof(*sp));
    FILE *fd = fopen(fname, "rb");
    if (!fd) { fprintf(stderr, "oracle: cannot open %s\n", fname); return 1; }
    int hdr[2];
    if (fread(hdr, sizeof(int), 2, fd) != 2 || hdr[0] < 0 || hdr[1] < 0) goto bad;
    sp->nrows = hdr[0];
    sp->nnz = hdr[1];
    sp->Ap = malloc(sizeof(int) * (size_t)(sp->nrows + 1));
    sp->Ai = malloc(sizeof(int) * (size_t)(sp->nnz ? sp->nnz : 1));
    sp->Ax = malloc(sizeof(ocplx) * (size_t)(sp->nnz ? sp->nnz : 1));
    if (fread(sp->Ap, sizeof(int), sp->nrows + 1, fd) != (size_t)(sp->nrows + 1)) goto bad;
    if (sp->Ap[sp->nrows] != sp->nnz) goto bad;
    if (fread(sp->Ai, sizeof(int), sp->nnz, fd) != (size_t)sp->nnz) goto bad;
    if (fread(sp->Ax, sizeof(ocplx), sp->nnz, fd) != (size_t)sp->nnz) goto bad;
    sp->perm = malloc(sizeof(int) * (size_t)(sp->nrows ? sp->nrows : 1));
    if (fread(sp->perm, sizeof(int), sp->nrows, fd) != (size_t)sp->nrows) {
        free(sp->perm);
        sp->perm = NULL;
    } else {
        if (fread(&sp->npart, sizeof(int), 1, fd) != 1 || sp->npart < 1) goto bad;
        sp->part = malloc(sizeof(int) * (size_t)(sp->npart + 1));
        if (fread(sp->part, sizeof(int), sp->npart + 1, fd) != (size_t)(sp->npart + 1)) goto bad;
    }
    fclose(fd);
    return 0;
bad:
    fprintf(stderr, "oracle: wrong file format in %s\n", fname);
    fclose(fd);
    or_csr_free(sp);
    return 2;
}

void or_csr_free(or_csr_t *sp)
{
    free(sp->Ap); free(sp->Ai); free(sp->Ax); free(sp->perm); free(sp->part);
    memset(sp, 0, sizeof(*sp));
}

/* tools/csr.c:945-952 */
ocplx or_csr_get_value(const or_csr_t *sp, int row, int col)
{
    int b = sp->Ap[row], n = sp->Ap[row + 1] - b;
    int cp = b + or_locate(sp->Ai + b, n, col);
    if (cp == sp->Ap[row + 1] || sp->Ai[cp] != col) return 0;
    return sp->Ax[cp];
}

/* tools/csr.c:954-961; returns nonzero when (row,col) is not stored */
static int csr_set_value(or_csr_t *sp, int row, int col, ocplx val)
{
    int b = sp->Ap[row], n = sp->Ap[row + 1] - b;
    int cp = b + or_locate(sp->Ai + b, n, col);
    if (cp == sp->Ap[row + 1] || sp->Ai[cp] != col) return 1;
    sp->Ax[cp] = val;
    return 0;
}

/* tools/csr.c:974-1003 (accumulates into result) */
void or_csr_spmv(int row_beg, int row_end, const int *Ap, const int *Ai, const ocplx *Ax,
                 const ocplx *x, ocplx *result)
{
    for (int i = row_beg; i < row_end; i++) {
        ocplx acc = 0;
        for (int j = Ap[i]; j < Ap[i + 1]; j++) acc += x[Ai[j]] * Ax[j];
        result[i] += acc;
    }
}

/* ------------------------------------------------------------------ */
/* unblocking (tools/csr.c:653-743)                                    */
/* ------------------------------------------------------------------ */
long long or_unblock_nnz(int nblk, const int *bAp, const int *subAp, int blkdim)
{
    return (long long)bAp[nblk] * subAp[blkdim];
}

void or_unblock(int nblk, const int *bAp, const int *bAi, int blkdim, const int *subAp, const int *subAi,
                int *Ap, int *Ai, int *Ai_sub_map)
{
    /* pass 1: expanded rows in (block row, sub row, block col, sub col) order */
    long long pos = 0;
    Ap[0] = 0;
    for (int row = 0; row < nblk; row++) {
        for (int rb = 0; rb < blkdim; rb++) {
            for (int cp = bAp[row]; cp < bAp[row + 1]; cp++) {
                int col = bAi[cp];
                for (int k = subAp[rb]; k < subAp[rb + 1]; k++) Ai[pos++] = col * blkdim + subAi[k];
            }
            Ap[row * blkdim + rb + 1] = (int)pos;
        }
    }
    /* pass 2: block-major map (block nz, sub row, sub nz) -> slot in the expanded CSR */
    int nrows = nblk * blkdim;
    int *cursor = malloc(sizeof(int) * (size_t)(nrows + 1));
    memcpy(cursor, Ap, sizeof(int) * (size_t)(nrows + 1));
    long long loc = 0;
    for (int row = 0; row < nblk; row++)
        for (int cp = bAp[row]; cp < bAp[row + 1]; cp++)
            for (int rb = 0; rb < blkdim; rb++) {
                int dst = row * blkdim + rb;
                for (int k = subAp[rb]; k < subAp[rb + 1]; k++) Ai_sub_map[loc++] = cursor[dst]++;
            }
    free(cursor);
}

/* ------------------------------------------------------------------ */
/* model                                                               */
/* ------------------------------------------------------------------ */
struct or_model {
    int lmax, nthreads;
    int all_cores; /* 1: SpMV rows and preconditioner blocks also run on nthreads (same results) */
    double h, dt;
    ocplx ihdt;
    or_csr_t Hall;
    or_csr_t *H;        /* H[2l] = H0l, H[2l+1] = H1l */
    or_csr_t *g, *gt;   /* index a*4*lmax + l*4 + n */
    or_csr_t h0[4], s0[3];
    int nblk, blkdim, nnzb;
    const int *subAp, *subAi;
    long long nnz_full;
    int *Ap, *Ai, *submap;
    ocplx *Ax;
    int *SAp, *SAi, *Ssub;
    ocplx *SAx, *HstAx;
    int *ikarr;
    ocplx *pgsum, *pgtsum; /* [l*4+n][nnzb] */
    /* preconditioner: banded LU of each distinct kappa block */
    int nkappa_distinct, band, *kidx, *ilv;
    ocplx *lu; /* [nkappa_distinct][blkdim][2*band+1] */
    /* BiCGSTAB workspace (persists across calls) */
    ocplx *r, *r0, *v, *p, *y, *z, *s, *t, *tmp;
    int maxit;
    double tol;
};

static int same_pattern(const or_csr_t *a, const or_csr_t *b)
{
    if (a->nrows != b->nrows || a->nnz != b->nnz) return 0;
    return !memcmp(a->Ap, b->Ap, sizeof(int) * (size_t)(a->nrows + 1)) &&
           !memcmp(a->Ai, b->Ai, sizeof(int) * (size_t)a->nnz);
}

static int read_into(const char *dir, const char *name, or_csr_t *sp)
{
    char fname[4096];
    snprintf(fname, sizeof fname, "%s/%s", dir, name);
    return or_csr_read(fname, sp);
}

/* tools/solver.c:409-458 */
static void compute_stationary(or_model_t *m)
{
    int nb = m->nblk, bd = m->blkdim, nz = m->nnzb;
    ocplx *sub = malloc(sizeof(ocplx) * nz);
    for (int row = 0; row < nb; row++) {
        int ki = m->ikarr[row];
        int *map = m->Ssub + (size_t)row * nz;
        for (int i = 0; i < nz; i++) sub[i] = m->s0[0].Ax[i] + m->s0[1].Ax[i] * ki + m->s0[2].Ax[i] * ki * ki;
        for (int i = 0; i < nz; i++) m->SAx[map[i]] = sub[i];
        for (int i = 0; i < nz; i++)
            sub[i] = m->ihdt * (m->h0[0].Ax[i] + m->h0[1].Ax[i] * ki + m->h0[2].Ax[i] * ki * ki +
                                m->h0[3].Ax[i] * ki * ki * ki);
        for (int i = 0; i < nz; i++) m->HstAx[map[i]] = sub[i];
    }
    (void)bd;
    free(sub);
}

or_model_t *or_model_load(const char *dir, int lmax, double h, double dt, int nthreads)
{
    or_model_t *m = calloc(1, sizeof(*m));
    char name[256];
    m->lmax = lmax;
    m->h = h;
    m->dt = dt;
    m->ihdt = I * h * dt / 2;
    m->nthreads = nthreads > 0 ? nthreads : 1;
    if (read_into(dir, "H.csr", &m->Hall)) goto fail;
    m->H = calloc(2 * lmax, sizeof(or_csr_t));
    for (int l = 0; l < lmax; l++) {
        snprintf(name, sizeof name, "H0l%d.csr", l);
        if (read_into(dir, name, m->H + 2 * l)) goto fail;
        snprintf(name, sizeof name, "H1l%d.csr", l);
        if (read_into(dir, name, m->H + 2 * l + 1)) goto fail;
    }
    int ng = 6 * lmax * 4;
    m->g = calloc(ng, sizeof(or_csr_t));
    m->gt = calloc(ng, sizeof(or_csr_t));
    for (int a = 0; a < 6; a++)
        for (int l = 0; l < lmax; l++)
            for (int n = 0; n < 4; n++) {
                snprintf(name, sizeof name, "g%da%dl%d.csr", n, a, l);
                if (read_into(dir, name, m->g + a * 4 * lmax + l * 4 + n)) goto fail;
            }
    for (int n = 0; n < 4; n++) {
        snprintf(name, sizeof name, "h0%d.csr", n);
        if (read_into(dir, name, m->h0 + n)) goto fail;
    }
    for (int n = 0; n < 3; n++) {
        snprintf(name, sizeof name, "s%d.csr", n);
        if (read_into(dir, name, m->s0 + n)) goto fail;
    }
    /* every radial matrix must share the g0a0l0 pattern (tools/solver.c:329-335 indexes them alike) */
    for (int i = 1; i < ng; i++)
        if (!same_pattern(m->g, m->g + i)) { fprintf(stderr, "oracle: g pattern mismatch\n"); goto fail; }
    for (int n = 0; n < 4; n++)
        if (!same_pattern(m->g, m->h0 + n)) { fprintf(stderr, "oracle: h0 pattern mismatch\n"); goto fail; }
    for (int n = 0; n < 3; n++)
        if (!same_pattern(m->g, m->s0 + n)) { fprintf(stderr, "oracle: s pattern mismatch\n"); goto fail; }
    m->nblk = m->Hall.nrows;
    m->blkdim = m->g[0].nrows;
    m->nnzb = m->g[0].nnz;
    m->subAp = m->g[0].Ap;
    m->subAi = m->g[0].Ai;
    /* conjugate transposes (tools/dirac_solver.c:315-326, tools/csr.c:963-972) */
    for (int i = 0; i < ng; i++) {
        or_csr_t *o = m->gt + i, *in = m->g + i;
        o->nrows = in->nrows;
        o->nnz = in->nnz;
        o->Ap = malloc(sizeof(int) * (in->nrows + 1));
        o->Ai = malloc(sizeof(int) * in->nnz);
        o->Ax = malloc(sizeof(ocplx) * in->nnz);
        memcpy(o->Ap, in->Ap, sizeof(int) * (in->nrows + 1));
        memcpy(o->Ai, in->Ai, sizeof(int) * in->nnz);
        memcpy(o->Ax, in->Ax, sizeof(ocplx) * in->nnz);
        for (int row = 0; row < in->nrows; row++)
            for (int cp = in->Ap[row]; cp < in->Ap[row + 1]; cp++)
                if (csr_set_value(o, in->Ai[cp], row, conj(in->Ax[cp]))) {
                    fprintf(stderr, "oracle: g pattern is not structurally symmetric\n");
                    goto fail;
                }
    }
    /* kappa per block (tools/solver.c:200-210) */
    m->ikarr = malloc(sizeof(int) * m->nblk);
    for (int c = 0; c < m->nblk; c++) m->ikarr[c] = or_ik(m->Hall.perm ? m->Hall.perm[c] : c);
    /* Hfull structure */
    m->nnz_full = or_unblock_nnz(m->nblk, m->Hall.Ap, m->subAp, m->blkdim);
    if (m->nnz_full > 2147483647LL) { fprintf(stderr, "oracle: nnz overflows int32\n"); goto fail; }
    int N = m->nblk * m->blkdim;
    m->Ap = malloc(sizeof(int) * (N + 1));
    m->Ai = malloc(sizeof(int) * m->nnz_full);
    m->submap = malloc(sizeof(int) * m->nnz_full);
    m->Ax = calloc(m->nnz_full, sizeof(ocplx));
    or_unblock(m->nblk, m->Hall.Ap, m->Hall.Ai, m->blkdim, m->subAp, m->subAi, m->Ap, m->Ai, m->submap);
    /* S / Hst: unblocked diagonal (tools/dirac_solver.c:229, 369-370) */
    int *dAp = malloc(sizeof(int) * (m->nblk + 1)), *dAi = malloc(sizeof(int) * m->nblk);
    for (int i = 0; i < m->nblk; i++) { dAp[i] = i; dAi[i] = i; }
    dAp[m->nblk] = m->nblk;
    long long snz = or_unblock_nnz(m->nblk, dAp, m->subAp, m->blkdim);
    m->SAp = malloc(sizeof(int) * (N + 1));
    m->SAi = malloc(sizeof(int) * snz);
    m->Ssub = malloc(sizeof(int) * snz);
    m->SAx = calloc(snz, sizeof(ocplx));
    m->HstAx = calloc(snz, sizeof(ocplx));
    or_unblock(m->nblk, dAp, dAi, m->blkdim, m->subAp, m->subAi, m->SAp, m->SAi, m->Ssub);
    free(dAp); free(dAi);
    compute_stationary(m);
    m->pgsum = calloc((size_t)lmax * 4 * m->nnzb, sizeof(ocplx));
    m->pgtsum = calloc((size_t)lmax * 4 * m->nnzb, sizeof(ocplx));
    m->maxit = 500;
    m->tol = 1e-12;
    return m;
fail:
    or_model_free(m);
    return NULL;
}

void or_model_free(or_model_t *m)
{
    if (!m) return;
    or_csr_free(&m->Hall);
    if (m->H) for (int i = 0; i < 2 * m->lmax; i++) or_csr_free(m->H + i);
    if (m->g) for (int i = 0; i < 24 * m->lmax; i++) or_csr_free(m->g + i);
    if (m->gt) for (int i = 0; i < 24 * m->lmax; i++) or_csr_free(m->gt + i);
    for (int n = 0; n < 4; n++) or_csr_free(m->h0 + n);
    for (int n = 0; n < 3; n++) or_csr_free(m->s0 + n);
    free(m->H); free(m->g); free(m->gt);
    free(m->Ap); free(m->Ai); free(m->submap); free(m->Ax);
    free(m->SAp); free(m->SAi); free(m->Ssub); free(m->SAx); free(m->HstAx);
    free(m->ikarr); free(m->pgsum); free(m->pgtsum);
    free(m->kidx); free(m->ilv); free(m->lu);
    free(m->r); free(m->r0); free(m->v); free(m->p); free(m->y); free(m->z); free(m->s); free(m->t); free(m->tmp);
    free(m);
}

void or_model_info(const or_model_t *m, int *nblk, int *blkdim, int *hall_nnz, long long *nnz_full, int *nnz_b)
{
    if (nblk) *nblk = m->nblk;
    if (blkdim) *blkdim = m->blkdim;
    if (hall_nnz) *hall_nnz = m->Hall.nnz;
    if (nnz_full) *nnz_full = m->nnz_full;
    if (nnz_b) *nnz_b = m->nnzb;
}
const int *or_model_hfull_Ap(const or_model_t *m) { return m->Ap; }
const int *or_model_hfull_Ai(const or_model_t *m) { return m->Ai; }
const int *or_model_hfull_submap(const or_model_t *m) { return m->submap; }
const ocplx *or_model_hfull_Ax(const or_model_t *m) { return m->Ax; }
const ocplx *or_model_S_Ax(const or_model_t *m) { return m->SAx; }
const int *or_model_S_Ap(const or_model_t *m) { return m->SAp; }
const int *or_model_S_Ai(const or_model_t *m) { return m->SAi; }

/* tools/gpu_solver.cu:137-188 */
void or_model_row_col(const or_model_t *m, int *Ai_orig, int *Aj_orig, int *Ki, int *Kj, ocplx *H0, ocplx *H1)
{
    int lmax = m->lmax;
    for (int row = 0; row < m->nblk; row++)
        for (int cp = m->Hall.Ap[row]; cp < m->Hall.Ap[row + 1]; cp++) {
            int col = m->Hall.Ai[cp];
            int orow = m->Hall.perm ? m->Hall.perm[row] : row;
            int ocol = m->Hall.perm ? m->Hall.perm[col] : col;
            Ai_orig[cp] = orow;
            Aj_orig[cp] = ocol;
            Ki[cp] = or_ik(orow);
            Kj[cp] = or_ik(ocol);
            for (int l = 0; l < lmax; l++) {
                H0[cp * lmax + l] = or_csr_get_value(m->H + 2 * l, orow, ocol);
                H1[cp * lmax + l] = or_csr_get_value(m->H + 2 * l + 1, orow, ocol);
            }
        }
}

/* G-sums of tools/solver.c:236-287 */
static void compute_gsums(or_model_t *m, const ocplx ft[6])
{
    int lmax = m->lmax, nz = m->nnzb;
    memset(m->pgsum, 0, sizeof(ocplx) * (size_t)lmax * 4 * nz);
    memset(m->pgtsum, 0, sizeof(ocplx) * (size_t)lmax * 4 * nz);
    for (int l = 0; l < lmax; l++)
        for (int a = 0; a < 6; a++) {
            if (a % 2 == l % 2) continue;
            for (int n = 0; n < 4; n++) {
                const ocplx *pg = m->g[a * 4 * lmax + l * 4 + n].Ax;
                const ocplx *pgt = m->gt[a * 4 * lmax + l * 4 + n].Ax;
                ocplx *gs = m->pgsum + (size_t)(l * 4 + n) * nz;
                ocplx *gts = m->pgtsum + (size_t)(l * 4 + n) * nz;
                for (int i = 0; i < nz; i++) {
                    gs[i] += SOL * m->ihdt * ft[a] * pg[i];
                    gts[i] += SOL * m->ihdt * ft[a] * pgt[i];
                }
            }
        }
}

/* tools/solver.c:196-407 */
void or_model_timedep(or_model_t *m, const ocplx ft[6])
{
    int lmax = m->lmax, nz = m->nnzb;
    compute_gsums(m, ft);
#pragma omp parallel num_threads(m->nthreads)
    {
        ocplx *sub = malloc(sizeof(ocplx) * nz);
        ocplx *H0 = malloc(sizeof(ocplx) * lmax), *H1 = malloc(sizeof(ocplx) * lmax);
#pragma omp for schedule(dynamic, 1)
        for (int row = 0; row < m->nblk; row++) {
            long long loc = (long long)m->Hall.Ap[row] * nz;
            for (int cp = m->Hall.Ap[row]; cp < m->Hall.Ap[row + 1]; cp++) {
                int col = m->Hall.Ai[cp];
                int orow = m->Hall.perm ? m->Hall.perm[row] : row;
                int ocol = m->Hall.perm ? m->Hall.perm[col] : col;
                int ki = m->ikarr[row], kj = m->ikarr[col];
                for (int l = 0; l < lmax; l++) {
                    H0[l] = or_csr_get_value(m->H + 2 * l, orow, ocol);
                    H1[l] = or_csr_get_value(m->H + 2 * l + 1, orow, ocol);
                }
                if (orow == ocol) {
                    for (int i = 0; i < nz; i++)
                        sub[i] = m->ihdt * (m->h0[0].Ax[i] + m->h0[1].Ax[i] * ki + m->h0[2].Ax[i] * ki * ki +
                                            m->h0[3].Ax[i] * ki * ki * ki);
                } else {
                    memset(sub, 0, sizeof(ocplx) * nz);
                }
                for (int l = 0; l < lmax; l++) {
                    const ocplx *g0 = m->pgsum + (size_t)(l * 4 + 0) * nz, *g1 = g0 + nz, *g2 = g1 + nz, *g3 = g2 + nz;
                    const ocplx *t0 = m->pgtsum + (size_t)(l * 4 + 0) * nz, *t1 = t0 + nz, *t2 = t1 + nz, *t3 = t2 + nz;
                    if (H0[l] != 0 && H1[l] != 0) {
                        for (int i = 0; i < nz; i++)
                            sub[i] += H0[l] * (g0[i] + g1[i] * ki + g2[i] * kj + g3[i] * ki * kj) +
                                      H1[l] * (t0[i] + t1[i] * kj + t2[i] * ki + t3[i] * ki * kj);
                    } else {
                        if (H0[l] != 0)
                            for (int i = 0; i < nz; i++)
                                sub[i] += H0[l] * (g0[i] + g1[i] * ki + g2[i] * kj + g3[i] * ki * kj);
                        if (H1[l] != 0)
                            for (int i = 0; i < nz; i++)
                                sub[i] += H1[l] * (t0[i] + t1[i] * kj + t2[i] * ki + t3[i] * ki * kj);
                    }
                }
                for (int i = 0; i < nz; i++) m->Ax[m->submap[loc++]] = sub[i];
            }
        }
        free(sub); free(H0); free(H1);
    }
}

/* Rows are independent and each row sums in the same order, so the threaded form of the all-cores
   baseline gives the serial result bit for bit. */
static void spmv_rows(const or_model_t *m, const int *Ap, const int *Ai, const ocplx *Ax, const ocplx *x,
                      ocplx *out)
{
    int n = m->nblk * m->blkdim;
    if (!m->all_cores || m->nthreads <= 1) {
        or_csr_spmv(0, n, Ap, Ai, Ax, x, out);
        return;
    }
#pragma omp parallel for num_threads(m->nthreads) schedule(static, 256)
    for (int i = 0; i < n; i++) or_csr_spmv(i, i + 1, Ap, Ai, Ax, x, out);
}

void or_model_spmv_H(const or_model_t *m, const ocplx *x, ocplx *out)
{
    spmv_rows(m, m->Ap, m->Ai, m->Ax, x, out);
}

void or_model_spmv_S(const or_model_t *m, const ocplx *x, ocplx *out)
{
    spmv_rows(m, m->SAp, m->SAi, m->SAx, x, out);
}

void or_model_set_threads(or_model_t *m, int nthreads, int all_cores)
{
    if (nthreads > 0) m->nthreads = nthreads;
    m->all_cores = all_cores ? 1 : 0;
}

/* tools/dirac_solver.c:115-126 */
void or_model_spmv_HS(const or_model_t *m, const ocplx *x, ocplx *out)
{
    or_model_spmv_H(m, x, out);
    or_model_spmv_S(m, x, out);
}

/* tools/solver.c:538-759 */
void or_model_matfree_H(or_model_t *m, const ocplx ft[6], const ocplx *x, ocplx *out)
{
    int lmax = m->lmax, nz = m->nnzb, bd = m->blkdim, nb = m->nblk;
    compute_gsums(m, ft);
    size_t vlen = (size_t)nb * bd;
    ocplx *gmv = calloc(vlen * lmax * 4, sizeof(ocplx));
    ocplx *gtmv = calloc(vlen * lmax * 4, sizeof(ocplx));
    for (int i = 0; i < lmax * 4; i++)
        for (int col = 0; col < nb; col++) {
            or_csr_spmv(0, bd, m->subAp, m->subAi, m->pgsum + (size_t)i * nz, x + (size_t)col * bd,
                        gmv + i * vlen + (size_t)col * bd);
            or_csr_spmv(0, bd, m->subAp, m->subAi, m->pgtsum + (size_t)i * nz, x + (size_t)col * bd,
                        gtmv + i * vlen + (size_t)col * bd);
        }
    for (int row = 0; row < nb; row++) {
        ocplx *yp = out + (size_t)row * bd;
        for (int cp = m->Hall.Ap[row]; cp < m->Hall.Ap[row + 1]; cp++) {
            int col = m->Hall.Ai[cp];
            const ocplx *xp = x + (size_t)col * bd;
            int orow = m->Hall.perm ? m->Hall.perm[row] : row;
            int ocol = m->Hall.perm ? m->Hall.perm[col] : col;
            int ki = m->ikarr[row], kj = m->ikarr[col];
            if (orow == ocol)
                for (int sr = 0; sr < bd; sr++)
                    for (int k = m->subAp[sr]; k < m->subAp[sr + 1]; k++) {
                        ocplx mv = m->ihdt * (m->h0[0].Ax[k] + m->h0[1].Ax[k] * ki + m->h0[2].Ax[k] * ki * ki +
                                              m->h0[3].Ax[k] * ki * ki * ki);
                        yp[sr] += xp[m->subAi[k]] * mv;
                    }
            for (int l = 0; l < lmax; l++) {
                ocplx H0 = or_csr_get_value(m->H + 2 * l, orow, ocol);
                ocplx H1 = or_csr_get_value(m->H + 2 * l + 1, orow, ocol);
                const ocplx *a0 = gmv + (size_t)(l * 4) * vlen + (size_t)col * bd;
                const ocplx *b0 = gtmv + (size_t)(l * 4) * vlen + (size_t)col * bd;
                for (int i = 0; i < bd; i++) {
                    if (H0 != 0)
                        yp[i] += H0 * (a0[i] + a0[vlen + i] * ki + a0[2 * vlen + i] * kj + a0[3 * vlen + i] * ki * kj);
                    if (H1 != 0)
                        yp[i] += H1 * (b0[i] + b0[vlen + i] * kj + b0[2 * vlen + i] * ki + b0[3 * vlen + i] * ki * kj);
                }
            }
        }
    }
    free(gmv);
    free(gtmv);
}

/* ------------------------------------------------------------------ */
/* preconditioner: exact banded LU per distinct kappa                  */
/* ------------------------------------------------------------------ */
int or_model_precond_setup(or_model_t *m)
{
    int bd = m->blkdim, nz = m->nnzb, nb = m->nblk;
    /* interleave the two radial halves so the four banded quadrants form one band */
    free(m->ilv);
    m->ilv = malloc(sizeof(int) * bd);
    for (int i = 0; i < bd; i++) m->ilv[i] = (bd % 2 == 0) ? (i < bd / 2 ? 2 * i : 2 * (i - bd / 2) + 1) : i;
    int band = 0;
    for (int r = 0; r < bd; r++)
        for (int k = m->subAp[r]; k < m->subAp[r + 1]; k++) {
            int d = abs(m->ilv[r] - m->ilv[m->subAi[k]]);
            if (d > band) band = d;
        }
    m->band = band;
    /* distinct kappa values */
    free(m->kidx);
    m->kidx = malloc(sizeof(int) * nb);
    int kvals[4096], nk = 0;
    for (int r = 0; r < nb; r++) {
        int j;
        for (j = 0; j < nk; j++) if (kvals[j] == m->ikarr[r]) break;
        if (j == nk) kvals[nk++] = m->ikarr[r];
        m->kidx[r] = j;
    }
    m->nkappa_distinct = nk;
    int w = 2 * band + 1;
    free(m->lu);
    m->lu = calloc((size_t)nk * bd * w, sizeof(ocplx));
    for (int q = 0; q < nk; q++) {
        int r0 = 0;
        while (m->kidx[r0] != q) r0++;
        ocplx *A = m->lu + (size_t)q * bd * w;
        const int *map = m->Ssub + (size_t)r0 * nz;
        for (int r = 0; r < bd; r++)
            for (int k = m->subAp[r]; k < m->subAp[r + 1]; k++) {
                int i = m->ilv[r], j = m->ilv[m->subAi[k]];
                A[(size_t)i * w + (j - i + band)] = m->SAx[map[k]] + m->HstAx[map[k]];
            }
        for (int k = 0; k < bd; k++) {
            ocplx piv = A[(size_t)k * w + band];
            if (piv == 0) { fprintf(stderr, "oracle: zero pivot\n"); return 1; }
            int iend = k + band < bd - 1 ? k + band : bd - 1;
            for (int i = k + 1; i <= iend; i++) {
                ocplx lik = A[(size_t)i * w + (k - i + band)] / piv;
                A[(size_t)i * w + (k - i + band)] = lik;
                for (int j = k + 1; j <= iend; j++)
                    A[(size_t)i * w + (j - i + band)] -= lik * A[(size_t)k * w + (j - k + band)];
            }
        }
    }
    return 0;
}

static void psolve_block(const or_model_t *m, int r, const ocplx *rhs, ocplx *x, ocplx *yv)
{
    int bd = m->blkdim, band = m->band, w = 2 * band + 1;
    const ocplx *A = m->lu + (size_t)m->kidx[r] * bd * w;
    const ocplx *b = rhs + (size_t)r * bd;
    ocplx *xo = x + (size_t)r * bd;
    for (int i = 0; i < bd; i++) yv[m->ilv[i]] = b[i];
    for (int i = 0; i < bd; i++) {
        ocplx acc = yv[i];
        for (int j = (i - band > 0 ? i - band : 0); j < i; j++) acc -= A[(size_t)i * w + (j - i + band)] * yv[j];
        yv[i] = acc;
    }
    for (int i = bd - 1; i >= 0; i--) {
        ocplx acc = yv[i];
        int jend = i + band < bd - 1 ? i + band : bd - 1;
        for (int j = i + 1; j <= jend; j++) acc -= A[(size_t)i * w + (j - i + band)] * yv[j];
        yv[i] = acc / A[(size_t)i * w + band];
    }
    for (int i = 0; i < bd; i++) xo[i] = yv[m->ilv[i]];
}

void or_model_psolve(const or_model_t *m, const ocplx *rhs, ocplx *x)
{
    int bd = m->blkdim;
    if (!m->all_cores || m->nthreads <= 1) {
        ocplx *yv = malloc(sizeof(ocplx) * bd);
        for (int r = 0; r < m->nblk; r++) psolve_block(m, r, rhs, x, yv);
        free(yv);
        return;
    }
#pragma omp parallel num_threads(m->nthreads)
    {
        ocplx *yv = malloc(sizeof(ocplx) * bd);
#pragma omp for schedule(static)
        for (int r = 0; r < m->nblk; r++) psolve_block(m, r, rhs, x, yv);
        free(yv);
    }
}

/* ------------------------------------------------------------------ */
/* BiCGSTAB (tools/bicgstab.c:125-295)                                 */
/* ------------------------------------------------------------------ */
static ocplx zdotc(int n, const ocplx *x, const ocplx *y)
{
    ocplx s = 0;
    for (int i = 0; i < n; i++) s += conj(x[i]) * y[i];
    return s;
}

static double sqnorm(int n, const ocplx *x)
{
    return creal(zdotc(n, x, x));
}

static void apply_A(or_model_t *m, const ocplx *x, ocplx *out, int *napply)
{
    int n = m->nblk * m->blkdim;
    memset(out, 0, sizeof(ocplx) * n);
    or_model_spmv_HS(m, x, out);
    if (napply) (*napply)++;
}

void or_reset_workspace(or_model_t *m)
{
    int n = m->nblk * m->blkdim;
    ocplx **ws[] = {&m->r, &m->r0, &m->v, &m->p, &m->y, &m->z, &m->s, &m->t, &m->tmp};
    for (unsigned i = 0; i < sizeof ws / sizeof ws[0]; i++) {
        free(*ws[i]);
        *ws[i] = calloc(n, sizeof(ocplx));
    }
}

void or_bicgstab(or_model_t *m, const ocplx *rhs, ocplx *x, int *iters, double *tol_error, int *napply)
{
    int n = m->nblk * m->blkdim;
    double tol = *tol_error;
    int maxit = *iters;
    if (!m->r) or_reset_workspace(m);
    ocplx *r = m->r, *r0 = m->r0, *v = m->v, *p = m->p, *y = m->y, *z = m->z, *s = m->s, *t = m->t;
    ocplx rho = 1, alpha = 1, w = 1;
    if (napply) *napply = 0;

    apply_A(m, x, r, napply);
    for (int i = 0; i < n; i++) { r[i] = rhs[i] - r[i]; r0[i] = r[i]; }
    double r0_sq = sqnorm(n, r0);
    double rhs_sq = sqnorm(n, rhs);
    if (rhs_sq == 0) {
        memset(x, 0, sizeof(ocplx) * n);
        *iters = 0;
        *tol_error = 0;
        return;
    }
    double tol2 = tol * tol * rhs_sq;
    double eps2 = 1e-12 * 1e-12;
    if (r0_sq < tol2) {
        *iters = 0;
        *tol_error = sqrt(r0_sq / rhs_sq);
        return;
    }
    int it = 0;
    while (it < maxit) {
        ocplx rho_old = rho;
        rho = zdotc(n, r0, r);
        if (cabs(rho) < eps2 * r0_sq) {
            apply_A(m, x, r, napply);
            for (int i = 0; i < n; i++) { r[i] = rhs[i] - r[i]; r0[i] = r[i]; }
            r0_sq = sqnorm(n, r);
            rho = r0_sq;
        }
        ocplx beta = (rho / rho_old) * (alpha / w);
        for (int i = 0; i < n; i++) p[i] = r[i] + beta * (p[i] - w * v[i]);
        or_model_psolve(m, p, y);
        apply_A(m, y, v, napply);
        alpha = rho / zdotc(n, r0, v);
        for (int i = 0; i < n; i++) s[i] = r[i] - alpha * v[i];
        or_model_psolve(m, s, z);
        apply_A(m, z, t, napply);
        double tt = sqnorm(n, t);
        w = tt > 0 ? zdotc(n, t, s) / tt : 0;
        for (int i = 0; i < n; i++) { z[i] = alpha * y[i] + w * z[i]; x[i] += z[i]; }
        for (int i = 0; i < n; i++) r[i] = s[i] - w * t[i];
        if (sqnorm(n, r) < tol2) break;
        ++it;
    }
    *tol_error = sqrt(sqnorm(n, r) / rhs_sq);
    *iters = it;
}

/* ------------------------------------------------------------------ */
/* CN time loop (tools/dirac_solver.c:496-577, CPU branch)             */
/* ------------------------------------------------------------------ */
int or_cn_run(or_model_t *m, double E0, double omega, double cycles, double t0, int nsteps,
              double tol, int maxit, ocplx *x, int *iters_out, double *tol_out, long *usec_out)
{
    int n = m->nblk * m->blkdim;
    ocplx *rhs = malloc(sizeof(ocplx) * n);
    double time = t0;
    for (int step = 0; step < nsteps; step++) {
        struct timeval a, b;
        gettimeofday(&a, NULL);
        ocplx ft[6];
        time = time + m->dt;
        or_pulse(E0, omega, cycles, time, ft);
        or_model_timedep(m, ft);
        memset(rhs, 0, sizeof(ocplx) * n);
        or_model_spmv_H(m, x, rhs);
        for (int i = 0; i < n; i++) rhs[i] = -1 * rhs[i];
        or_model_spmv_S(m, x, rhs);
        int iters = maxit;
        double terr = tol;
        or_bicgstab(m, rhs, x, &iters, &terr, NULL);
        gettimeofday(&b, NULL);
        if (iters_out) iters_out[step] = iters;
        if (tol_out) tol_out[step] = terr;
        if (usec_out) usec_out[step] = (b.tv_sec - a.tv_sec) * 1000000L + (b.tv_usec - a.tv_usec);
    }
    free(rhs);
    return 0;
}

/* tools/dirac_solver.c:68-108 */
int or_vec_read(const char *fname, int n, ocplx *out, int *perm)
{
    FILE *fd = fopen(fname, "rb");
    if (!fd) return 1;
    int nn;
    if (fread(&nn, sizeof(int), 1, fd) != 1 || nn != n) { fclose(fd); return 2; }
    if (fread(out, sizeof(ocplx), n, fd) != (size_t)n) { fclose(fd); return 2; }
    if (perm && fread(perm, sizeof(int), n, fd) != (size_t)n) { fclose(fd); return 2; }
    fclose(fd);
    return 0;
}


