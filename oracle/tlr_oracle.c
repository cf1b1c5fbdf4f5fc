/*
 * tlr_oracle.c -- CPU restatement of hicma-parsec's TLR Cholesky hot path.
 *
 * TEST INFRASTRUCTURE ONLY (the checker).  Loaded by tests/, smoke() and the
 * cpu_baseline leg of bench.py; never by the product library.
 *
 * Dense BLAS/LAPACK come from MKL (sequential, LP64), as the reference links
 * them (hicma_parsec/CMakeLists.txt:47-55).  The task graph runs as OpenMP
 * tasks with tile dependencies, standing in for PaRSEC's dataflow.
 *
 * Parity notes (pinned by tests/test_oracle.py unless marked):
 *  - points: the build's seeded jittered-grid/Morton spec (golden: numpy).
 *  - kernels: sqexp / exp closed forms of STARS-H ssdata kernels
 *    (hicma_parsec/testing_dpotrf.c:336-358).
 *  - compression: STARS-H starsh_dense_dlrrsdd (randomized SVD, oversample 10,
 *    relative-Frobenius truncation) as called at hicma_parsec/STARSH_gen.jdf:158-171.
 *    STARS-H is an un-vendored dependency; its RNG stream is replaced by a
 *    seeded counter-based normal sketch (parity unpinned for the sketch bits;
 *    pinned by invariants against numpy SVD).
 *  - HCORE kernels (un-vendored): semantics from the call sites in
 *    hicma_parsec/HiCMA_dpotrf_L_2flow.jdf and the workspace/op-count comments;
 *    recompression truncates singular values below acc (absolute), capped at
 *    maxrank (pinned by invariants, parity with HCORE itself unpinned).
 */
#define _GNU_SOURCE
#include "tlr_oracle.h"
#include <math.h>
#include <omp.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

/* ---- Fortran BLAS / LAPACK prototypes (MKL LP64) ---- */
extern void dgemm_(const char *, const char *, const int *, const int *, const int *,
                   const double *, const double *, const int *, const double *,
                   const int *, const double *, double *, const int *);
extern void dsyrk_(const char *, const char *, const int *, const int *, const double *,
                   const double *, const int *, const double *, double *, const int *);
extern void dtrsm_(const char *, const char *, const char *, const char *, const int *,
                   const int *, const double *, const double *, const int *, double *,
                   const int *);
extern void dtrmm_(const char *, const char *, const char *, const char *, const int *,
                   const int *, const double *, const double *, const int *, double *,
                   const int *);
extern void dpotrf_(const char *, const int *, double *, const int *, int *);
extern void dgeqrf_(const int *, const int *, double *, const int *, double *, double *,
                    const int *, int *);
extern void dorgqr_(const int *, const int *, const int *, double *, const int *,
                    const double *, double *, const int *, int *);
extern void dormqr_(const char *, const char *, const int *, const int *, const int *,
                    const double *, const int *, const double *, double *, const int *,
                    double *, const int *, int *);
extern void dgesvd_(const char *, const char *, const int *, const int *, double *,
                    const int *, double *, double *, const int *, double *, const int *,
                    double *, const int *, int *);
extern void dgesdd_(const char *, const int *, const int *, double *, const int *,
                    double *, double *, const int *, double *, const int *, double *,
                    const int *, int *, int *);

/* libmkl_rt: select the sequential threading layer and LP64 interface once
 * at load time (each OpenMP task then runs single-threaded BLAS) */
extern int MKL_Set_Threading_Layer(int);
extern int MKL_Set_Interface_Layer(int);
__attribute__((constructor)) static void orc_mkl_init(void)
{
    MKL_Set_Interface_Layer(0); /* MKL_INTERFACE_LP64 */
    MKL_Set_Threading_Layer(1); /* MKL_THREADING_SEQUENTIAL */
}

static void gemm(char ta, char tb, int m, int n, int k, double al, const double *A,
                 int lda, const double *B, int ldb, double be, double *C, int ldc)
{
    if (m <= 0 || n <= 0) return;
    if (k <= 0) {
        for (int j = 0; j < n; j++)
            for (int i = 0; i < m; i++) C[i + (size_t)j * ldc] *= be;
        return;
    }
    dgemm_(&ta, &tb, &m, &n, &k, &al, A, &lda, B, &ldb, &be, C, &ldc);
}

/* ---- counter-based RNG (splitmix64) ---- */
static uint64_t splitmix64(uint64_t x)
{
    x += 0x9E3779B97F4A7C15ULL;
    uint64_t z = x;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
}
static double uniform01(uint64_t seed, uint64_t ctr)
{
    return (double)(splitmix64(seed * 0x632BE59BD9B4E019ULL + ctr) >> 11) * 0x1.0p-53;
}

static int cmp_code(const void *a, const void *b)
{
    uint64_t x = *(const uint64_t *)a, y = *(const uint64_t *)b;
    return x < y ? -1 : (x > y ? 1 : 0);
}

/* ---- points: jittered grid, Morton order ----
 * Build's spec for STARS-H's UNIFORM placement + z-sort
 * (hicma_parsec/testing_dpotrf.c:309-312, 315-358). */
int orc_points(int N, int ndim, uint64_t seed, double *xyz)
{
    if (N <= 0 || ndim < 1 || ndim > 3) return -1;
    long s = (long)ceil(pow((double)N, 1.0 / ndim) - 1e-9);
    if (s < 1) s = 1;
    for (;;) {
        long p = 1;
        for (int d = 0; d < ndim; d++) p *= s;
        if (p >= N) break;
        s++;
    }
    long total = 1;
    for (int d = 0; d < ndim; d++) total *= s;
    uint64_t *key = malloc(sizeof(uint64_t) * total);
    for (long g = 0; g < total; g++) {
        long rem = g;
        uint64_t code = 0;
        for (int d = 0; d < ndim; d++) {
            uint64_t c = (uint64_t)(rem % s);
            rem /= s;
            for (int b = 0; b < 21; b++) code |= ((c >> b) & 1ULL) << (b * ndim + d);
        }
        key[g] = code;
    }
    /* order of grid points by Morton code: codes are unique, so sort (code,g) */
    uint64_t *pair = malloc(sizeof(uint64_t) * 2 * total);
    for (long g = 0; g < total; g++) { pair[2 * g] = key[g]; pair[2 * g + 1] = (uint64_t)g; }
    qsort(pair, total, 2 * sizeof(uint64_t), cmp_code);
    for (long i = 0; i < N; i++) {
        long g = (long)pair[2 * i + 1];
        long rem = g;
        for (int d = 0; d < ndim; d++) {
            long c = rem % s;
            rem /= s;
            double u = uniform01(seed, (uint64_t)g * ndim + d);
            xyz[(size_t)d * N + i] = ((double)c + 0.5 + 0.8 * u - 0.4) / (double)s;
        }
    }
    free(key); free(pair);
    return 0;
}

void orc_sketch(int ld, int ncol, uint64_t seed, double *omega)
{
    const double twopi = 6.283185307179586476925286766559;
    for (long e = 0; e < (long)ld * ncol; e++) {
        double u1 = uniform01(seed, 2 * (uint64_t)e);
        double u2 = uniform01(seed, 2 * (uint64_t)e + 1);
        omega[e] = sqrt(-2.0 * log(1.0 - u1)) * cos(twopi * u2);
    }
}

static int kind_ndim(int kind) { return kind == ORC_KIND_SQEXP_2D ? 2 : 3; }

/* STARS-H ssdata kernels: sigma2*exp(-r^2/(2 beta^2)) (sqexp) or
 * sigma2*exp(-r/beta) (exp); the nugget is added on the diagonal. */
void orc_kernel_block(int kind, int nrows, int ncols, const double *xyz, int N,
                      int row0, int col0, double sigma2, double beta, double nugget,
                      double *out, int ld)
{
    int nd = kind_ndim(kind);
    for (int j = 0; j < ncols; j++) {
        int gj = col0 + j;
        for (int i = 0; i < nrows; i++) {
            int gi = row0 + i;
            double d2 = 0;
            for (int d = 0; d < nd; d++) {
                double t = xyz[(size_t)d * N + gi] - xyz[(size_t)d * N + gj];
                d2 += t * t;
            }
            double v;
            if (kind == ORC_KIND_EXP_3D) v = sigma2 * exp(-sqrt(d2) / beta);
            else v = sigma2 * exp(-d2 / (2.0 * beta * beta));
            if (gi == gj) v += nugget;
            out[i + (size_t)j * ld] = v;
        }
    }
}

/* ---- descriptor ---- */
struct orc_tlr {
    int N, NB, NT, kind, ndim;
    double sigma2, beta, nugget;
    uint64_t seed;
    double *xyz;
    int band;          /* current band size */
    double **tile;     /* NT*NT, index n*NT+m; band tiles NB*NB, LR compact */
    int *rank;         /* NT*NT, n*NT+m */
    int *cap;          /* LR column capacity */
};

#define TIX(t, m, n) ((size_t)(n) * (t)->NT + (m))
static int tdim(const orc_tlr *t, int m) { return m == t->NT - 1 ? t->N - m * t->NB : t->NB; }

orc_tlr *orc_create(int N, int NB, int kind, double sigma2, double beta, double nugget,
                    uint64_t seed)
{
    if (N <= 0 || NB <= 0) return NULL;
    orc_tlr *t = calloc(1, sizeof(orc_tlr));
    t->N = N; t->NB = NB; t->NT = (N + NB - 1) / NB; t->kind = kind;
    t->ndim = kind_ndim(kind);
    t->sigma2 = sigma2; t->beta = beta; t->nugget = nugget; t->seed = seed;
    t->xyz = malloc(sizeof(double) * (size_t)t->ndim * N);
    orc_points(N, t->ndim, seed, t->xyz);
    t->tile = calloc((size_t)t->NT * t->NT, sizeof(double *));
    t->rank = malloc(sizeof(int) * (size_t)t->NT * t->NT);
    t->cap = calloc((size_t)t->NT * t->NT, sizeof(int));
    for (size_t i = 0; i < (size_t)t->NT * t->NT; i++) t->rank[i] = -1;
    t->band = 1;
    return t;
}

void orc_free(orc_tlr *t)
{
    if (!t) return;
    for (size_t i = 0; i < (size_t)t->NT * t->NT; i++) free(t->tile[i]);
    free(t->tile); free(t->rank); free(t->cap); free(t->xyz); free(t);
}
int orc_NT(const orc_tlr *t) { return t->NT; }
const double *orc_xyz(const orc_tlr *t) { return t->xyz; }
int orc_band_size(const orc_tlr *t) { return t->band; }

static void gen_dense_tile(orc_tlr *t, int m, int n)
{
    size_t ix = TIX(t, m, n);
    free(t->tile[ix]);
    t->tile[ix] = calloc((size_t)t->NB * t->NB, sizeof(double));
    orc_kernel_block(t->kind, tdim(t, m), tdim(t, n), t->xyz, t->N, m * t->NB, n * t->NB,
                     t->sigma2, t->beta, t->nugget, t->tile[ix], t->NB);
}

/* STARS-H dsvfr: smallest rank whose discarded tail satisfies
 * sum_{i>=r} S_i^2 <= tol^2 * sum S_i^2 */
static int dsvfr(int n, const double *S, double tol)
{
    double tot = 0, acc = 0;
    for (int i = 0; i < n; i++) tot += S[i] * S[i];
    double stop = tot * tol * tol;
    for (int i = n - 1; i >= 0; i--) {
        acc += S[i] * S[i];
        if (acc > stop) return i + 1;
    }
    return 0;
}

/* starsh_dense_dlrrsdd restated (hicma_parsec/STARSH_gen.jdf:158-171 call site;
 * workspace sizing :263-271).  D: nrows x ncols, ld nrows. U/V ld NB. */
static int lrrsdd(int nrows, int ncols, const double *D, int NB, const double *omega,
                  int maxrank, int oversample, double tol, double *U, double *V)
{
    int mn = nrows < ncols ? nrows : ncols;
    int mn2 = maxrank + oversample;
    if (mn2 > mn) mn2 = mn;
    double *Q = malloc(sizeof(double) * (size_t)nrows * mn2);
    double *X = malloc(sizeof(double) * (size_t)ncols * mn2);
    double *tau = malloc(sizeof(double) * mn2);
    double *S = malloc(sizeof(double) * mn2);
    double *VT = malloc(sizeof(double) * (size_t)mn2 * mn2);
    int info, lw = -1;
    double wq;
    gemm('N', 'N', nrows, mn2, ncols, 1.0, D, nrows, omega, NB, 0.0, Q, nrows);
    dgeqrf_(&nrows, &mn2, Q, &nrows, tau, &wq, &lw, &info);
    lw = (int)wq + 1;
    double *work = malloc(sizeof(double) * lw);
    dgeqrf_(&nrows, &mn2, Q, &nrows, tau, work, &lw, &info);
    int lw2 = -1;
    dorgqr_(&nrows, &mn2, &mn2, Q, &nrows, tau, &wq, &lw2, &info);
    if ((int)wq + 1 > lw) { lw = (int)wq + 1; work = realloc(work, sizeof(double) * lw); }
    dorgqr_(&nrows, &mn2, &mn2, Q, &nrows, tau, work, &lw, &info);
    gemm('T', 'N', ncols, mn2, nrows, 1.0, D, nrows, Q, nrows, 0.0, X, ncols);
    int *iwork = malloc(sizeof(int) * 8 * mn2);
    int one = 1;
    lw2 = -1;
    dgesdd_("O", &ncols, &mn2, X, &ncols, S, NULL, &one, VT, &mn2, &wq, &lw2, iwork, &info);
    if ((int)wq + 1 > lw) { lw = (int)wq + 1; work = realloc(work, sizeof(double) * lw); }
    dgesdd_("O", &ncols, &mn2, X, &ncols, S, NULL, &one, VT, &mn2, work, &lw, iwork, &info);
    int rank = dsvfr(mn2, S, tol);
    if (rank < mn / 2 && rank <= maxrank) {
        /* U = Q * VT(1:rank,:)^T ; V = X(:,1:rank) * diag(S) */
        gemm('N', 'T', nrows, rank, mn2, 1.0, Q, nrows, VT, mn2, 0.0, U, NB);
        for (int j = 0; j < rank; j++)
            for (int i = 0; i < ncols; i++) V[i + (size_t)j * NB] = X[i + (size_t)j * ncols] * S[j];
    } else {
        rank = -1;
    }
    free(Q); free(X); free(tau); free(S); free(VT); free(work); free(iwork);
    return rank;
}

/* STARSH_gen (hicma_parsec/STARSH_gen.jdf:45-213, 233-312) */
int orc_generate(orc_tlr *t, double tol, int maxrank, int band_size)
{
    if (band_size < 1) return -2;
    int NT = t->NT, NB = t->NB, info = 0;
    t->band = band_size;
    int p = maxrank + 10;
    if (p > NB) p = NB;
    double *omega = malloc(sizeof(double) * (size_t)NB * p);
    orc_sketch(NB, p, 0x5EED0FULL, omega);
#pragma omp parallel for schedule(dynamic) collapse(2)
    for (int n = 0; n < NT; n++)
        for (int m = 0; m < NT; m++) {
            if (m < n) continue;
            if (m - n < band_size) {
                gen_dense_tile(t, m, n);
                continue;
            }
            int rows = tdim(t, m), cols = tdim(t, n);
            double *D = malloc(sizeof(double) * (size_t)rows * cols);
            orc_kernel_block(t->kind, rows, cols, t->xyz, t->N, m * NB, n * NB, t->sigma2,
                             t->beta, t->nugget, D, rows);
            double *U = calloc((size_t)NB * maxrank * 2, sizeof(double));
            double *V = U + (size_t)NB * maxrank;
            int r = lrrsdd(rows, cols, D, NB, omega, maxrank, 10, tol, U, V);
            size_t ix = TIX(t, m, n);
            if (r < 0) {
#pragma omp atomic write
                info = -1;
                t->rank[ix] = -1;
            } else {
                double *buf = calloc((size_t)NB * (r > 0 ? r : 1) * 2, sizeof(double));
                memcpy(buf, U, sizeof(double) * (size_t)NB * r);
                memcpy(buf + (size_t)NB * r, V, sizeof(double) * (size_t)NB * r);
                free(t->tile[ix]);
                t->tile[ix] = buf;
                t->cap[ix] = r;
                t->rank[ix] = r;
            }
            free(U); free(D);
        }
    free(omega);
    return info;
}

/* band_regenerate (hicma_parsec/band_regenerate.jdf:27-64) */
int orc_band_regenerate(orc_tlr *t, int band_size)
{
    t->band = band_size;
#pragma omp parallel for schedule(dynamic) collapse(2)
    for (int m = 0; m < t->NT; m++)
        for (int d = 0; d < band_size; d++) {
            int n = m - d;
            if (n < 0) continue;
            gen_dense_tile(t, m, n);
        }
    return 0;
}

/* rank_gather (hicma_parsec/rank_gather.jdf:23-37) */
void orc_rank_gather(const orc_tlr *t, int *G)
{
    int NT = t->NT;
    memset(G, 0, sizeof(int) * (size_t)NT * NT);
    for (int m = t->band; m < NT; m++)
        for (int n = 0; n <= m - t->band; n++) G[(size_t)n * NT + m] = t->rank[TIX(t, m, n)];
}

/* HICMA_get_stat2 (hicma_parsec/hicma_internal.c:52-72) */
void orc_rank_stat(const int *G, int NT, int band, int *mn, int *mx, double *avg)
{
    int lo = 2147483647, hi = 0;
    long num = 0;
    long long sum = 0;
    for (int i = band; i < NT; i++)
        for (int j = 0; j < i - band + 1; j++) {
            int v = G[(size_t)j * NT + i];
            sum += v; num++;
            if (v < lo) lo = v;
            if (v > hi) hi = v;
        }
    *mn = lo; *mx = hi;
    *avg = (double)sum / (double)num;
}

/* rank_check (hicma_parsec/rank_check.jdf:22-61): counts negative off-band
 * ranks and sets band entries to -1 */
int orc_rank_check(int *G, int NT, int band)
{
    int bad = 0;
    for (int m = band; m < NT; m++)
        for (int n = 0; n <= m - band; n++)
            if (G[(size_t)n * NT + m] < 0 && n < m) bad++;
    for (int m = 0; m < NT; m++)
        for (int n = (m - band + 1 > 0 ? m - band + 1 : 0); n <= m; n++) G[(size_t)n * NT + m] = -1;
    return bad;
}

/* op_counts (hicma_parsec/op_counts.h:10-72); the float casts inside 'q' are
 * part of the formula and kept. */
unsigned long orc_op_counts(char op, unsigned long a, unsigned long b, unsigned long c,
                            unsigned long d)
{
    unsigned long res = 0;
    if (op == 'q') {
        unsigned long m = a, n = b;
        res = 2 * m * n * n - (unsigned long)(2 * n * n * n / 3.0f) + 2 * m * n +
              (unsigned long)(17 * n / 3.0f);
    } else if (op == 'c') {
        unsigned long n = a;
        res = n * n * n / 3 - n * n / 2.0 + n / 6;
    } else if (op == 't') {
        unsigned long m = a, n = b;
        res = (c == 1) ? n * m * m : m * n * n;
    } else if (op == 'm') {
        res = a * b * c * 2;
    } else if (op == 's') {
        res = 22 * a * a * a;
    } else if (op == 'o') {
        unsigned long m = a, n = b, k = c;
        if (d == 1) res = 4 * n * m * k - 2 * n * k * k + 3 * n * k;
        else res = 4 * n * m * k - 2 * m * k * k + 2 * m * k + n * k - k * k / 2 + k / 2;
    } else if (op == 'r') {
        unsigned long m = a, n = b;
        res = (c == 1) ? m * m * n : m * n * n;
    }
    return res;
}
#define OPC orc_op_counts

/* band_size_auto_tuning (hicma_parsec/band_size_auto_tuning.jdf:36-127, 201-318),
 * FLUCTUATION from hicma_parsec/hicma_parsec_internal.h:107 */
int orc_band_autotune(const int *G, int NT, int NB)
{
    const double FLUCT = 0.66666667;
    int opt_min = 0, ind_min = 0;
    for (int b = 2; b < NT; b++) {
        long long tlr = 0, dense = 0;
        for (int m = b - 1; m < NT; m++) {
            int k = m + 1 - b;
            int Ar = G[(size_t)k * NT + m];
            tlr += (long long)NB * NB * Ar;
            dense += (long long)NB * NB * NB;
        }
        for (int k = 0; k <= NT - 3; k++)
            for (int m = k + b; m < NT; m++) {
                int n = m + 1 - b;
                int Ar = G[(size_t)k * NT + m], Br = G[(size_t)k * NT + n], Cr = G[(size_t)n * NT + m];
                int s = Cr + Ar;
                unsigned long qa = OPC('q', NB, s, 0, 0);
                unsigned long qb = OPC('m', Ar, Br, NB, 0);
                qb += OPC('m', Ar, NB, Br, 0);
                qb += OPC('q', NB, s, 0, 0);
                unsigned long sv = OPC('r', s, s, 2, 0);
                sv += OPC('s', s, 0, 0, 0);
                sv += (unsigned long)s * Cr;
                unsigned long nu = OPC('o', NB, Cr, s, 1);
                unsigned long nv = OPC('o', Cr, NB, s, 2);
                tlr += (long long)(qa + qb + sv + nu + nv);
                if (m - k < b) dense += (long long)2 * NB * NB * NB;
                else if (n - k < b && m - n < b) dense += (long long)4 * NB * NB * Ar;
                else if (m - n < b)
                    dense += (long long)2 * NB * NB * (Ar < Br ? Ar : Br) + (long long)4 * NB * Ar * Br;
            }
        if (tlr * FLUCT <= dense && ind_min == 0) {
            ind_min = 1;
            opt_min = b - 1;
        }
        if (tlr <= dense) return opt_min;
    }
    return NT;
}

/* ---- HCORE recompression (case iv / v) ----
 * C' = Cu Cv^T + Pn Qn^T with Pn = Au (mrows x ra), Qn (ncols x ra) the
 * update's right factor (already scaled by alpha).  QR of both stacks,
 * M = Rp Rq^T, SVD(M), truncate sigma < acc (or fixed rk), cap maxrank,
 * U' = Qp [Us S], V' = Qq [Vs].  Returns the new rank; *out receives a
 * compact [U|V] buffer with ld NB. */
static int recompress(int NB, int mrows, int ncols, const double *Cu, const double *Cv,
                      int rc, const double *Au, int ra, const double *Qn, int ldq,
                      double acc, int rk, int maxrank, double **out, int *outcap)
{
    int s = rc + ra, info;
    double *P = malloc(sizeof(double) * (size_t)mrows * s);
    double *Q = malloc(sizeof(double) * (size_t)ncols * s);
    for (int j = 0; j < rc; j++) {
        memcpy(P + (size_t)j * mrows, Cu + (size_t)j * NB, sizeof(double) * mrows);
        memcpy(Q + (size_t)j * ncols, Cv + (size_t)j * NB, sizeof(double) * ncols);
    }
    for (int j = 0; j < ra; j++) {
        memcpy(P + (size_t)(rc + j) * mrows, Au + (size_t)j * NB, sizeof(double) * mrows);
        memcpy(Q + (size_t)(rc + j) * ncols, Qn + (size_t)j * ldq, sizeof(double) * ncols);
    }
    int kp = mrows < s ? mrows : s, kq = ncols < s ? ncols : s;
    double *taup = malloc(sizeof(double) * s), *tauq = malloc(sizeof(double) * s);
    double wq;
    int lw = -1;
    dgeqrf_(&mrows, &s, P, &mrows, taup, &wq, &lw, &info);
    int lwork = (int)wq + 1;
    lw = -1;
    dgeqrf_(&ncols, &s, Q, &ncols, tauq, &wq, &lw, &info);
    if ((int)wq + 1 > lwork) lwork = (int)wq + 1;
    double *work = malloc(sizeof(double) * lwork);
    dgeqrf_(&mrows, &s, P, &mrows, taup, work, &lwork, &info);
    dgeqrf_(&ncols, &s, Q, &ncols, tauq, work, &lwork, &info);
    /* M = Rp Rq^T (kp x kq) */
    double *Rp = calloc((size_t)kp * s, sizeof(double));
    double *Rq = calloc((size_t)kq * s, sizeof(double));
    for (int j = 0; j < s; j++) {
        for (int i = 0; i <= j && i < kp; i++) Rp[i + (size_t)j * kp] = P[i + (size_t)j * mrows];
        for (int i = 0; i <= j && i < kq; i++) Rq[i + (size_t)j * kq] = Q[i + (size_t)j * ncols];
    }
    double *M = malloc(sizeof(double) * (size_t)kp * kq);
    gemm('N', 'T', kp, kq, s, 1.0, Rp, kp, Rq, kq, 0.0, M, kp);
    int kmin = kp < kq ? kp : kq;
    double *S = malloc(sizeof(double) * kmin);
    double *Us = malloc(sizeof(double) * (size_t)kp * kmin);
    double *VsT = malloc(sizeof(double) * (size_t)kmin * kq);
    lw = -1;
    dgesvd_("S", "S", &kp, &kq, M, &kp, S, Us, &kp, VsT, &kmin, &wq, &lw, &info);
    int lw2 = (int)wq + 1;
    double *w2 = malloc(sizeof(double) * lw2);
    dgesvd_("S", "S", &kp, &kq, M, &kp, S, Us, &kp, VsT, &kmin, w2, &lw2, &info);
    int nr;
    if (rk > 0) nr = rk < kmin ? rk : kmin;
    else {
        nr = kmin;
        for (int i = 0; i < kmin; i++)
            if (S[i] < acc) { nr = i; break; }
    }
    if (nr > maxrank) nr = maxrank;
    double *nu = calloc((size_t)mrows * (nr > 0 ? nr : 1), sizeof(double));
    double *nv = calloc((size_t)ncols * (nr > 0 ? nr : 1), sizeof(double));
    for (int j = 0; j < nr; j++) {
        for (int i = 0; i < kp; i++) nu[i + (size_t)j * mrows] = Us[i + (size_t)j * kp] * S[j];
        for (int i = 0; i < kq; i++) nv[i + (size_t)j * ncols] = VsT[j + (size_t)i * kmin];
    }
    if (nr > 0) {
        lw = -1;
        dormqr_("L", "N", &mrows, &nr, &kp, P, &mrows, taup, nu, &mrows, &wq, &lw, &info);
        int lw3 = (int)wq + 1;
        lw = -1;
        dormqr_("L", "N", &ncols, &nr, &kq, Q, &ncols, tauq, nv, &ncols, &wq, &lw, &info);
        if ((int)wq + 1 > lw3) lw3 = (int)wq + 1;
        double *w3 = malloc(sizeof(double) * lw3);
        dormqr_("L", "N", &mrows, &nr, &kp, P, &mrows, taup, nu, &mrows, w3, &lw3, &info);
        dormqr_("L", "N", &ncols, &nr, &kq, Q, &ncols, tauq, nv, &ncols, w3, &lw3, &info);
        free(w3);
    }
    if (nr > *outcap || *out == NULL) {
        free(*out);
        *out = calloc((size_t)NB * (nr > 0 ? nr : 1) * 2, sizeof(double));
        *outcap = nr;
    }
    double *o = *out;
    memset(o, 0, sizeof(double) * (size_t)NB * nr * 2);
    for (int j = 0; j < nr; j++) {
        memcpy(o + (size_t)j * NB, nu + (size_t)j * mrows, sizeof(double) * mrows);
        memcpy(o + (size_t)NB * nr + (size_t)j * NB, nv + (size_t)j * ncols, sizeof(double) * ncols);
    }
    free(P); free(Q); free(taup); free(tauq); free(work); free(Rp); free(Rq); free(M);
    free(S); free(Us); free(VsT); free(w2); free(nu); free(nv);
    return nr;
}

/* ---- the factorization (2-flow semantics; band 1 = 3-flow case set) ---- */
typedef struct {
    orc_tlr *t;
    double acc;
    int rk, maxrank;
    int *info;
    unsigned long *tileops;
    unsigned long *st; /* per-thread x4 */
} fctx;

static void logop(fctx *c, unsigned long cnt, int i, int j, int cls_band, int path)
{
    int th = omp_get_thread_num();
    unsigned long *st = c->st + 4 * (size_t)th;
#pragma omp atomic
    c->tileops[(size_t)i * c->t->NT + j] += cnt;
    st[cls_band ? 0 : 1] += cnt;
    st[path ? 2 : 3] += cnt;
}

/* potrf_dpotrf (HiCMA_dpotrf_L_2flow.jdf:175-245) */
static void t_potrf(fctx *c, int k)
{
    orc_tlr *t = c->t;
    int nn = tdim(t, k), iinfo = 0, ld = t->NB;
    logop(c, OPC('c', nn, 0, 0, 0), k, k, 1, 1);
    dpotrf_("L", &nn, t->tile[TIX(t, k, k)], &ld, &iinfo);
    if (iinfo != 0) {
#pragma omp critical(orc_info)
        if (*c->info == 0) *c->info = k * t->NB + iinfo;
    }
}

/* potrf_dtrsm (HiCMA_dpotrf_L_2flow.jdf:251-401) */
static void t_trsm(fctx *c, int m, int k)
{
    orc_tlr *t = c->t;
    int NB = t->NB, mm = tdim(t, m), kk = tdim(t, k);
    double one = 1.0;
    const double *T = t->tile[TIX(t, k, k)];
    if (m - k < t->band) {
        logop(c, OPC('t', mm, mm, 1, 0), m, k, 1, m - k == 1);
        dtrsm_("R", "L", "T", "N", &mm, &kk, &one, T, &NB, t->tile[TIX(t, m, k)], &NB);
    } else {
        int r = t->rank[TIX(t, m, k)];
        logop(c, OPC('t', mm, r, 1, 0), m, k, 0, m - k == 1);
        if (r == 0) return;
        double *V = t->tile[TIX(t, m, k)] + (size_t)NB * r;
        dtrsm_("L", "L", "N", "N", &kk, &r, &one, T, &NB, V, &NB);
    }
}

/* potrf_dsyrk (HiCMA_dpotrf_L_2flow.jdf:409-565) */
static void t_syrk(fctx *c, int k, int m)
{
    orc_tlr *t = c->t;
    int NB = t->NB, mm = tdim(t, m), kk = tdim(t, k);
    double *T = t->tile[TIX(t, m, m)];
    if (m - k < t->band) {
        double mone = -1.0, one = 1.0;
        logop(c, OPC('m', mm, mm, mm, 0), m, m, 1, m - k == 1);
        dsyrk_("L", "N", &mm, &kk, &mone, t->tile[TIX(t, m, k)], &NB, &one, T, &NB);
    } else {
        int r = t->rank[TIX(t, m, k)];
        unsigned long cnt = OPC('m', r, r, mm, 0) + OPC('m', mm, r, r, 0) + OPC('m', mm, mm, r, 0);
        logop(c, cnt, m, m, 1, m - k == 1);
        if (r == 0) return;
        const double *U = t->tile[TIX(t, m, k)], *V = U + (size_t)NB * r;
        double *Z = malloc(sizeof(double) * (size_t)r * r);
        double *Y = malloc(sizeof(double) * (size_t)mm * r);
        gemm('T', 'N', r, r, kk, 1.0, V, NB, V, NB, 0.0, Z, r);
        gemm('N', 'N', mm, r, r, 1.0, U, NB, Z, r, 0.0, Y, mm);
        gemm('N', 'T', mm, mm, r, -1.0, Y, mm, U, NB, 1.0, T, NB);
        free(Z); free(Y);
    }
}

/* potrf_dgemm (HiCMA_dpotrf_L_2flow.jdf:573-1097), all five cases */
static void t_gemm(fctx *c, int m, int n, int k)
{
    orc_tlr *t = c->t;
    int NB = t->NB, band = t->band, mm = tdim(t, m), nn = tdim(t, n), kk = tdim(t, k);
    size_t iA = TIX(t, m, k), iB = TIX(t, n, k), iC = TIX(t, m, n);
    double *A = t->tile[iA], *B = t->tile[iB], *C = t->tile[iC];
    if (m - k < band) { /* (i) dense */
        logop(c, OPC('m', mm, mm, mm, 0), m, n, 1, 0);
        gemm('N', 'T', mm, nn, kk, -1.0, A, NB, B, NB, 1.0, C, NB);
        return;
    }
    int ra = t->rank[iA];
    const double *Au = A, *Av = A + (size_t)NB * (ra > 0 ? ra : 0);
    if (n - k < band && m - n < band) { /* (ii) A LR, B dense, C dense */
        logop(c, OPC('m', mm, mm, ra, 0) * 2, m, n, 1, 0);
        if (ra == 0) return;
        double *W = malloc(sizeof(double) * (size_t)ra * nn);
        gemm('T', 'T', ra, nn, kk, 1.0, Av, NB, B, NB, 0.0, W, ra);
        gemm('N', 'N', mm, nn, ra, -1.0, Au, NB, W, ra, 1.0, C, NB);
        free(W);
        return;
    }
    if (m - n < band) { /* (iii) A LR, B LR, C dense */
        int rb = t->rank[iB];
        logop(c, OPC('m', mm, mm, ra < rb ? ra : rb, 0) + OPC('m', mm, ra, rb, 0) * 2, m, n, 1, 0);
        if (ra == 0 || rb == 0) return;
        const double *Bu = B, *Bv = B + (size_t)NB * rb;
        double *X = malloc(sizeof(double) * (size_t)ra * rb);
        gemm('T', 'N', ra, rb, kk, 1.0, Av, NB, Bv, NB, 0.0, X, ra);
        if (ra > rb) {
            double *W = malloc(sizeof(double) * (size_t)mm * rb);
            gemm('N', 'N', mm, rb, ra, 1.0, Au, NB, X, ra, 0.0, W, mm);
            gemm('N', 'T', mm, nn, rb, -1.0, W, mm, Bu, NB, 1.0, C, NB);
            free(W);
        } else {
            double *W = malloc(sizeof(double) * (size_t)ra * nn);
            gemm('N', 'T', ra, nn, rb, 1.0, X, ra, Bu, NB, 0.0, W, ra);
            gemm('N', 'N', mm, nn, ra, -1.0, Au, NB, W, ra, 1.0, C, NB);
            free(W);
        }
        free(X);
        return;
    }
    int rc = t->rank[iC];
    const double *Cu = C, *Cv = C + (size_t)NB * rc;
    int s = rc + ra;
    double *Qn = NULL;
    int rb = 0;
    if (n - k < band) { /* (iv) A LR, B dense, C LR: Qn = -(B Av) */
        Qn = malloc(sizeof(double) * (size_t)nn * (ra > 0 ? ra : 1));
        gemm('N', 'N', nn, ra, kk, -1.0, B, NB, Av, NB, 0.0, Qn, nn);
    } else { /* (v) LR x LR -> LR */
        rb = t->rank[iB];
        if (ra == 0 || rb == 0) return;
        const double *Bu = B, *Bv = B + (size_t)NB * rb;
        double *X = malloc(sizeof(double) * (size_t)ra * rb);
        gemm('T', 'N', ra, rb, kk, 1.0, Av, NB, Bv, NB, 0.0, X, ra);
        Qn = malloc(sizeof(double) * (size_t)nn * ra);
        gemm('N', 'T', nn, ra, rb, -1.0, Bu, NB, X, ra, 0.0, Qn, nn);
        free(X);
    }
    double *out = NULL;
    int outcap = 0;
    int nr = recompress(NB, mm, nn, Cu, Cv, rc, Au, ra, Qn, nn, c->acc, c->rk, c->maxrank,
                        &out, &outcap);
    free(Qn);
    free(t->tile[iC]);
    t->tile[iC] = out;
    t->cap[iC] = outcap;
    t->rank[iC] = nr;
    /* operation count, as logged after the HCORE calls */
    /* counted with tile m's rows (tempmmv), as the reference does */
    unsigned long qa = OPC('q', mm, s, 0, 0), qb;
    if (n - k < band) qb = OPC('q', mm, s, 0, 0) + OPC('m', ra, mm, mm, 0);
    else qb = OPC('m', ra, rb, mm, 0) + OPC('m', ra, mm, rb, 0) + OPC('q', mm, s, 0, 0);
    unsigned long sv = OPC('r', s, s, 2, 0) + OPC('s', s, 0, 0, 0) + (unsigned long)s * nr;
    unsigned long nu = OPC('o', mm, nr, s, 1), nv = OPC('o', nr, mm, s, 2);
    logop(c, qa + qb + sv + nu + nv, m, n, 0, 0);
}

int orc_dpotrf(orc_tlr *t, double acc, int rk, int maxrank, int kmax, unsigned long *tileops,
               unsigned long *stats, double *seconds)
{
    int NT = t->NT;
    if (maxrank > t->NB / 2) return -2;
    int info = 0;
    int nth = omp_get_max_threads();
    fctx c = {t, acc, rk, maxrank, &info, tileops, calloc(4 * (size_t)nth, sizeof(unsigned long))};
    unsigned long *dummy = NULL;
    if (!tileops) { dummy = calloc((size_t)NT * NT, sizeof(unsigned long)); c.tileops = dummy; }
    char *dep = calloc((size_t)NT * NT, 1);
    int K = (kmax < 0 || kmax > NT) ? NT : kmax;
    double t0 = omp_get_wtime();
#pragma omp parallel
#pragma omp single
    {
        for (int k = 0; k < K; k++) {
#pragma omp task depend(inout : dep[(size_t)k * NT + k]) firstprivate(k)
            t_potrf(&c, k);
            for (int m = k + 1; m < NT; m++) {
#pragma omp task depend(in : dep[(size_t)k * NT + k]) depend(inout : dep[(size_t)k * NT + m]) firstprivate(m, k)
                t_trsm(&c, m, k);
            }
            for (int m = k + 1; m < NT; m++) {
#pragma omp task depend(in : dep[(size_t)k * NT + m]) depend(inout : dep[(size_t)m * NT + m]) firstprivate(m, k)
                t_syrk(&c, k, m);
            }
            for (int n = k + 1; n < NT; n++)
                for (int m = n + 1; m < NT; m++) {
#pragma omp task depend(in : dep[(size_t)k * NT + m], dep[(size_t)k * NT + n]) depend(inout : dep[(size_t)n * NT + m]) firstprivate(m, n, k)
                    t_gemm(&c, m, n, k);
                }
        }
#pragma omp taskwait
    }
    double t1 = omp_get_wtime();
    if (seconds) *seconds = t1 - t0;
    if (stats) {
        for (int i = 0; i < 4; i++) stats[i] = 0;
        for (int th = 0; th < nth; th++)
            for (int i = 0; i < 4; i++) stats[i] += c.st[4 * (size_t)th + i];
    }
    free(c.st); free(dep); free(dummy);
    return info;
}

/* ---- uncompress and checks (STARSH_check.jdf:90-143; testing_dpotrf.c:111-230) ---- */
void orc_uncompress(const orc_tlr *t, double *D)
{
    int N = t->N, NB = t->NB, NT = t->NT;
    memset(D, 0, sizeof(double) * (size_t)N * N);
#pragma omp parallel for schedule(dynamic) collapse(2)
    for (int n = 0; n < NT; n++)
        for (int m = 0; m < NT; m++) {
            if (m < n) continue;
            int mm = tdim(t, m), nn = tdim(t, n);
            double *dst = D + (size_t)n * NB * N + (size_t)m * NB;
            const double *src = t->tile[TIX(t, m, n)];
            if (m - n < t->band) {
                for (int j = 0; j < nn; j++)
                    for (int i = 0; i < mm; i++) dst[i + (size_t)j * N] = src[i + (size_t)j * NB];
            } else {
                int r = t->rank[TIX(t, m, n)];
                if (r > 0) gemm('N', 'T', mm, nn, r, 1.0, src, NB, src + (size_t)NB * r, NB, 0.0, dst, N);
            }
        }
    for (size_t j = 0; j < (size_t)N; j++)
        for (size_t i = 0; i < j; i++) D[i + j * N] = 0.0;
}

void orc_dense_exact(const orc_tlr *t, double *D)
{
    int N = t->N;
#pragma omp parallel for schedule(dynamic, 16)
    for (int j = 0; j < N; j++) {
        orc_kernel_block(t->kind, N - j, 1, t->xyz, N, j, j, t->sigma2, t->beta, t->nugget,
                         D + (size_t)j * N + j, N);
        for (int i = 0; i < j; i++) D[i + (size_t)j * N] = 0.0;
    }
}

double orc_logdet(const orc_tlr *t)
{
    double s = 0;
    for (int k = 0; k < t->NT; k++) {
        const double *T = t->tile[TIX(t, k, k)];
        int nn = tdim(t, k);
        for (int i = 0; i < nn; i++) s += log(T[i + (size_t)i * t->NB]);
    }
    return 2.0 * s;
}

/* dlansy(F, lower) of a symmetric matrix stored in its lower triangle */
static double lansy_lower(int N, const double *A, int lda)
{
    double s = 0;
#pragma omp parallel for reduction(+ : s)
    for (int j = 0; j < N; j++) {
        double d = A[j + (size_t)j * lda];
        double off = 0;
        for (int i = j + 1; i < N; i++) {
            double v = A[i + (size_t)j * lda];
            off += v * v;
        }
        s += d * d + 2.0 * off;
    }
    return sqrt(s);
}

double orc_check_dpotrf2(int N, const double *L, const double *A0)
{
    double *LL = malloc(sizeof(double) * (size_t)N * N);
    for (size_t j = 0; j < (size_t)N; j++)
        for (size_t i = 0; i < (size_t)N; i++) LL[i + j * N] = i >= j ? L[i + j * N] : 0.0;
    double one = 1.0;
    dtrmm_("R", "L", "T", "N", &N, &N, &one, L, &N, LL, &N);
    for (size_t j = 0; j < (size_t)N; j++)
        for (size_t i = j; i < (size_t)N; i++) LL[i + j * N] -= A0[i + j * N];
    double r = lansy_lower(N, LL, N) / lansy_lower(N, A0, N);
    free(LL);
    return r;
}

double orc_check_diff(int N, const double *L, const double *L0)
{
    double an = lansy_lower(N, L, N);
    double *D = malloc(sizeof(double) * (size_t)N * N);
    for (size_t j = 0; j < (size_t)N; j++)
        for (size_t i = 0; i < (size_t)N; i++) D[i + j * N] = i >= j ? L[i + j * N] - L0[i + j * N] : 0.0;
    double rn = lansy_lower(N, D, N);
    free(D);
    return rn / an;
}

int orc_dense_potrf(int N, double *A)
{
    int info = 0;
    dpotrf_("L", &N, A, &N, &info);
    return info;
}

double orc_dense_logdet(int N, const double *L)
{
    double s = 0;
    for (size_t i = 0; i < (size_t)N; i++) s += log(L[i + i * N]);
    return 2.0 * s;
}

/* ---- triangular solves with the TLR factor ----
 * No counterpart in the reference (it stops at the factor); this is the
 * textbook blocked substitution over the A0 tile layout, the checker of the
 * device solve.  trans = 0: B <- L^{-1} B, trans = 1: B <- L^{-T} B, with
 * L_mk = U V^T for off-band tiles.  B is N x nrhs, column major, ld ldb. */
int orc_tlr_trsm(const orc_tlr *t, int trans, int nrhs, double *B, int ldb)
{
    const int NT = t->NT, NB = t->NB;
    const double one = 1.0;
    if (nrhs <= 0) return 0;
    double *W = malloc(sizeof(double) * (size_t)NB * nrhs);
    for (int s = 0; s < NT; s++) {
        const int k = trans ? NT - 1 - s : s;
        const int kk = tdim(t, k);
        double *Bk = B + (size_t)k * NB;
        dtrsm_("L", "L", trans ? "T" : "N", "N", &kk, &nrhs, &one, t->tile[TIX(t, k, k)], &NB, Bk, &ldb);
        const int lo = trans ? 0 : k + 1, hi = trans ? k : NT;
        for (int o = lo; o < hi; o++) {
            const int m = trans ? k : o, n = trans ? o : k;  /* tile L(m, n) */
            const int mm = tdim(t, m), nn = tdim(t, n);
            const double *T = t->tile[TIX(t, m, n)];
            double *Bo = B + (size_t)o * NB;
            if (m - n < t->band) {
                if (trans) gemm('T', 'N', nn, nrhs, mm, -1.0, T, NB, Bk, ldb, 1.0, Bo, ldb);
                else gemm('N', 'N', mm, nrhs, nn, -1.0, T, NB, Bk, ldb, 1.0, Bo, ldb);
            } else {
                const int r = t->rank[TIX(t, m, n)];
                if (r <= 0) continue;
                const double *U = T, *V = T + (size_t)NB * r;
                if (trans) { /* B_n -= V (U^T B_k) */
                    gemm('T', 'N', r, nrhs, mm, 1.0, U, NB, Bk, ldb, 0.0, W, r);
                    gemm('N', 'N', nn, nrhs, r, -1.0, V, NB, W, r, 1.0, Bo, ldb);
                } else { /* B_m -= U (V^T B_k) */
                    gemm('T', 'N', r, nrhs, nn, 1.0, V, NB, Bk, ldb, 0.0, W, r);
                    gemm('N', 'N', mm, nrhs, r, -1.0, U, NB, W, r, 1.0, Bo, ldb);
                }
            }
        }
    }
    free(W);
    return 0;
}

/* ---- CPU-baseline helpers: thread count and BLAS identification ---- */
extern void mkl_get_version_string(char *buf, int len);
int orc_set_threads(int n)
{
    if (n > 0) omp_set_num_threads(n);
    return omp_get_max_threads();
}
void orc_blas_version(char *buf, int len) { mkl_get_version_string(buf, len); }

/* ---- tile access ---- */
int orc_get_rank(const orc_tlr *t, int m, int n) { return t->rank[TIX(t, m, n)]; }

int orc_get_tile(const orc_tlr *t, int m, int n, double *buf)
{
    const double *src = t->tile[TIX(t, m, n)];
    if (!src) return -1;
    if (m - n < t->band) {
        memcpy(buf, src, sizeof(double) * (size_t)t->NB * t->NB);
        return t->NB;
    }
    int r = t->rank[TIX(t, m, n)];
    if (r > 0) memcpy(buf, src, sizeof(double) * (size_t)t->NB * r * 2);
    return r;
}

int orc_set_band(orc_tlr *t, int band_size)
{
    t->band = band_size;
    return 0;
}

int orc_set_tile(orc_tlr *t, int m, int n, const double *buf, int rank)
{
    size_t ix = TIX(t, m, n);
    free(t->tile[ix]);
    if (m - n < t->band) {
        t->tile[ix] = malloc(sizeof(double) * (size_t)t->NB * t->NB);
        memcpy(t->tile[ix], buf, sizeof(double) * (size_t)t->NB * t->NB);
        return 0;
    }
    size_t cnt = (size_t)t->NB * (rank > 0 ? rank : 1) * 2;
    t->tile[ix] = calloc(cnt, sizeof(double));
    if (rank > 0) memcpy(t->tile[ix], buf, sizeof(double) * (size_t)t->NB * rank * 2);
    t->rank[ix] = rank;
    t->cap[ix] = rank;
    return 0;
}
