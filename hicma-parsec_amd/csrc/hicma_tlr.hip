// hicma_tlr.hip -- host side of the MI355X TLR Cholesky library (C-ABI).
//
// One HIP stream per context; the PaRSEC task graph of HiCMA_dpotrf_L is
// replaced by a panel-step schedule of batched launches whose grids cover
// every task of a step; tile ranks never leave HBM (kernels read them).
#include "tlr_kernels.hpp"
#include "rc_fast.hpp"
#include "rc_wave.hpp"
#include "../../include/hicma_tlr.h"

#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <algorithm>
#include <vector>

#define HIPCHK(x)                                                                                   \
    do {                                                                                            \
        hipError_t e_ = (x);                                                                        \
        if (e_ != hipSuccess) {                                                                     \
            fprintf(stderr, "hicma_tlr: HIP error '%s' at %s:%d\n", hipGetErrorString(e_), __FILE__, \
                    __LINE__);                                                                      \
            return e_ == hipErrorOutOfMemory ? HICMA_ERR_NOMEM : HICMA_ERR_HIP;                     \
        }                                                                                           \
    } while (0)

struct hicma_ctx {
    int device;
    int ncu;
    hipStream_t stream;
    double *ws;
    size_t ws_bytes;
    int *d_info;
    int *d_ovf;
    double *d_scalar;
    size_t lds_max;
    int profile;                     // record events around recompression launches
    std::vector<hipEvent_t> prof_ev; // pairs
    // the recompression variants of one panel step run concurrently on side
    // streams (default priority), forked from and joined back to `stream`
    // (the device exposes 4 hardware queues per process: 4 streams in all)
    hipStream_t side[3];             // [2]: syrk(k) + potrf(k+1), one step ahead
    hipEvent_t ev_fork, ev_join[3];
};

struct hicma_tlr_desc {
    hicma_ctx *ctx;
    TLR t;  // device pointers + sizes
    int band_alloc;
    uint64_t seed;
    double *h_xyz;
    unsigned long long *d_tileops;
    unsigned long long *d_opstat;
    unsigned long last_stats[8];
    int last_ovf;
};

// ---------------------------------------------------------------------------
// point set: jittered grid, Morton order (spec of the build; see DESIGN)
static uint64_t sm64(uint64_t x)
{
    x += 0x9E3779B97F4A7C15ULL;
    uint64_t z = x;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
}
static double u01(uint64_t seed, uint64_t ctr)
{
    return (double)(sm64(seed * 0x632BE59BD9B4E019ULL + ctr) >> 11) * 0x1.0p-53;
}

extern "C" int hicma_points(int N, int ndim, uint64_t seed, double *xyz)
{
    if (N <= 0 || ndim < 1 || ndim > 3 || !xyz) return HICMA_ERR_ARG;
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
    std::vector<std::pair<uint64_t, long>> key((size_t)total);
    for (long g = 0; g < total; g++) {
        long rem = g;
        uint64_t code = 0;
        for (int d = 0; d < ndim; d++) {
            uint64_t c = (uint64_t)(rem % s);
            rem /= s;
            for (int b = 0; b < 21; b++) code |= ((c >> b) & 1ULL) << (b * ndim + d);
        }
        key[(size_t)g] = {code, g};
    }
    std::sort(key.begin(), key.end());
    for (long i = 0; i < N; i++) {
        long g = key[(size_t)i].second, rem = g;
        for (int d = 0; d < ndim; d++) {
            long c = rem % s;
            rem /= s;
            double u = u01(seed, (uint64_t)g * ndim + d);
            xyz[(size_t)d * N + i] = ((double)c + 0.5 + 0.8 * u - 0.4) / (double)s;
        }
    }
    return 0;
}

// standard normal sketch, Box-Muller on the counter-based uniform stream
static void sketch(int ld, int ncol, uint64_t seed, double *om)
{
    const double twopi = 6.283185307179586476925286766559;
    for (long e = 0; e < (long)ld * ncol; e++) {
        double a = u01(seed, 2 * (uint64_t)e), b = u01(seed, 2 * (uint64_t)e + 1);
        om[e] = sqrt(-2.0 * log(1.0 - a)) * cos(twopi * b);
    }
}
#define SKETCH_SEED 0x5EED0FULL

// ---------------------------------------------------------------------------
extern "C" unsigned long hicma_op_counts(char op, unsigned long a, unsigned long b, unsigned long c,
                                         unsigned long d)
{
    unsigned long r = 0;
    if (op == 'q') {
        r = 2 * a * b * b - (unsigned long)((float)(2 * b * b * b) / 3.0f) + 2 * a * b +
            (unsigned long)((float)(17 * b) / 3.0f);
    } else if (op == 'c') {
        r = (unsigned long)((double)(a * a * a / 3) - (double)(a * a) / 2.0 + (double)(a / 6));
    } else if (op == 't') {
        r = (c == 1) ? b * a * a : a * b * b;
    } else if (op == 'm') {
        r = a * b * c * 2;
    } else if (op == 's') {
        r = 22 * a * a * a;
    } else if (op == 'o') {
        if (d == 1) r = 4 * b * a * c - 2 * b * c * c + 3 * b * c;
        else r = 4 * b * a * c - 2 * a * c * c + 2 * a * c + b * c - c * c / 2 + c / 2;
    } else if (op == 'r') {
        r = (c == 1) ? a * a * b : a * b * b;
    }
    return r;
}

// ---------------------------------------------------------------------------
extern "C" int hicma_init(hicma_ctx **ctx, int device)
{
    if (!ctx) return HICMA_ERR_ARG;
    int ndev = 0;
    HIPCHK(hipGetDeviceCount(&ndev));
    if (device < 0 || device >= ndev) return HICMA_ERR_ARG;
    HIPCHK(hipSetDevice(device));
    hicma_ctx *c = new hicma_ctx();
    c->device = device;
    hipDeviceProp_t prop;
    HIPCHK(hipGetDeviceProperties(&prop, device));
    c->ncu = prop.multiProcessorCount;
    c->lds_max = prop.sharedMemPerBlock;
    HIPCHK(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
    for (int i = 0; i < 3; i++) {
        HIPCHK(hipStreamCreateWithFlags(&c->side[i], hipStreamNonBlocking));
        HIPCHK(hipEventCreateWithFlags(&c->ev_join[i], hipEventDisableTiming));
    }
    HIPCHK(hipEventCreateWithFlags(&c->ev_fork, hipEventDisableTiming));
    HIPCHK(hipMalloc(&c->d_info, 4 * sizeof(int)));
    c->d_ovf = c->d_info + 1;
    HIPCHK(hipMalloc(&c->d_scalar, 8 * sizeof(double)));
    *ctx = c;
    return 0;
}

extern "C" int hicma_fini(hicma_ctx *c)
{
    if (!c) return 0;
    hipSetDevice(c->device);
    if (c->ws) hipFree(c->ws);
    hipFree(c->d_info);
    hipFree(c->d_scalar);
    for (auto e : c->prof_ev) (void)hipEventDestroy(e);
    for (int i = 0; i < 3; i++) {
        (void)hipEventDestroy(c->ev_join[i]);
        (void)hipStreamDestroy(c->side[i]);
    }
    (void)hipEventDestroy(c->ev_fork);
    hipStreamDestroy(c->stream);
    delete c;
    return 0;
}

static int ensure_ws(hicma_ctx *c, size_t bytes)
{
    if (c->ws_bytes >= bytes) return 0;
    if (c->ws) HIPCHK(hipFree(c->ws));
    c->ws = nullptr;
    c->ws_bytes = 0;
    HIPCHK(hipMalloc(&c->ws, bytes));
    c->ws_bytes = bytes;
    return 0;
}

static int kind_ndim(int kind) { return kind == HICMA_KIND_SQEXP_2D ? 2 : 3; }

extern "C" int hicma_desc_create(hicma_ctx *ctx, hicma_tlr_desc **out, int N, int NB, int maxrank, int kind,
                                 double sigma2, double beta, double nugget, uint64_t seed)
{
    if (!ctx || !out || N <= 0 || NB <= 0 || maxrank <= 0 || maxrank > NB) return HICMA_ERR_ARG;
    if (kind != HICMA_KIND_SQEXP_2D && kind != HICMA_KIND_SQEXP_3D && kind != HICMA_KIND_EXP_3D)
        return HICMA_ERR_ARG;
    HIPCHK(hipSetDevice(ctx->device));
    hicma_tlr_desc *A = (hicma_tlr_desc *)calloc(1, sizeof(hicma_tlr_desc));
    A->ctx = ctx;
    A->seed = seed;
    TLR &t = A->t;
    t.N = N;
    t.NB = NB;
    t.NT = (N + NB - 1) / NB;
    t.cap = maxrank;
    t.kind = kind;
    t.ndim = kind_ndim(kind);
    t.sigma2 = sigma2;
    t.beta = beta;
    t.nugget = nugget;
    t.band_size = 1;
    A->h_xyz = (double *)malloc(sizeof(double) * (size_t)t.ndim * N);
    hicma_points(N, t.ndim, seed, A->h_xyz);
    double *dx = nullptr;
    int NT = t.NT;
    size_t lr_slots = (size_t)NT * (NT - 1) / 2;
    if (hipMalloc(&dx, sizeof(double) * (size_t)t.ndim * N) != hipSuccess ||
        hipMalloc(&t.rank, sizeof(int) * (size_t)NT * NT) != hipSuccess ||
        (lr_slots && hipMalloc(&t.lr, sizeof(double) * lr_slots * NB * 2 * maxrank) != hipSuccess) ||
        hipMalloc(&A->d_tileops, sizeof(unsigned long long) * (size_t)NT * NT) != hipSuccess ||
        hipMalloc(&A->d_opstat, sizeof(unsigned long long) * 8) != hipSuccess) {
        hipGetLastError();
        if (dx) hipFree(dx);
        if (t.rank) hipFree(t.rank);
        if (t.lr) hipFree(t.lr);
        if (A->d_tileops) hipFree(A->d_tileops);
        free(A->h_xyz);
        free(A);
        return HICMA_ERR_NOMEM;
    }
    HIPCHK(hipMemcpy(dx, A->h_xyz, sizeof(double) * (size_t)t.ndim * N, hipMemcpyHostToDevice));
    t.xyz = dx;
    HIPCHK(hipMemset(t.rank, 0xFF, sizeof(int) * (size_t)NT * NT));
    *out = A;
    return 0;
}

extern "C" int hicma_desc_destroy(hicma_tlr_desc *A)
{
    if (!A) return 0;
    hipSetDevice(A->ctx->device);
    hipFree((void *)A->t.xyz);
    hipFree(A->t.rank);
    if (A->t.lr) hipFree(A->t.lr);
    if (A->t.band) hipFree(A->t.band);
    hipFree(A->d_tileops);
    hipFree(A->d_opstat);
    free(A->h_xyz);
    free(A);
    return 0;
}

extern "C" int hicma_desc_info(const hicma_tlr_desc *A, int *N, int *NB, int *NT, int *band_size, int *maxrank)
{
    if (!A) return HICMA_ERR_ARG;
    if (N) *N = A->t.N;
    if (NB) *NB = A->t.NB;
    if (NT) *NT = A->t.NT;
    if (band_size) *band_size = A->t.band_size;
    if (maxrank) *maxrank = A->t.cap;
    return 0;
}

extern "C" int hicma_get_points(const hicma_tlr_desc *A, double *xyz)
{
    if (!A || !xyz) return HICMA_ERR_ARG;
    memcpy(xyz, A->h_xyz, sizeof(double) * (size_t)A->t.ndim * A->t.N);
    return 0;
}

static int alloc_band(hicma_tlr_desc *A, int band_size)
{
    TLR &t = A->t;
    if (A->band_alloc != band_size) {
        if (t.band) HIPCHK(hipFree(t.band));
        t.band = nullptr;
        A->band_alloc = 0;
        HIPCHK(hipMalloc(&t.band, sizeof(double) * (size_t)band_size * t.NT * t.NB * t.NB));
        A->band_alloc = band_size;
    }
    t.band_size = band_size;
    return 0;
}

static int gen_band_tiles(hicma_tlr_desc *A)
{
    TLR &t = A->t;
    dim3 grid((unsigned)(((size_t)t.NB * t.NB + BS - 1) / BS), (unsigned)(t.band_size * t.NT));
    hipLaunchKernelGGL(k_gen_band, grid, dim3(BS), 0, A->ctx->stream, t, 0);
    HIPCHK(hipGetLastError());
    return 0;
}

static int nslots_for(hicma_ctx *c, size_t slot_bytes, long long ntask)
{
    // persistent grid: about two workgroups per CU, bounded by an HBM budget
    long long want = 2LL * c->ncu;
    size_t budget = (size_t)24 << 30;
    long long cap_mem = (long long)(budget / (slot_bytes ? slot_bytes : 1));
    if (cap_mem < 1) cap_mem = 1;
    if (want > cap_mem) want = cap_mem;
    if (want > ntask) want = ntask;
    if (want < 1) want = 1;
    return (int)want;
}

extern "C" int STARSH_gen(hicma_ctx *ctx, int uplo, hicma_tlr_desc *A, hicma_tlr_desc *Ar, double tol,
                          int maxrank, int band_size, int send_full_tile, int *info)
{
    (void)Ar;
    (void)send_full_tile;
    if (!ctx || !A) return HICMA_ERR_ARG;
    if (uplo != HICMA_LOWER && uplo != HICMA_UPPER) return HICMA_ERR_UPLO;
    if (band_size < 1 || maxrank < 1 || maxrank > A->t.cap) return HICMA_ERR_ARG;
    HIPCHK(hipSetDevice(ctx->device));
    TLR &t = A->t;
    int rc = alloc_band(A, band_size);
    if (rc) return rc;
    HIPCHK(hipMemsetAsync(t.rank, 0xFF, sizeof(int) * (size_t)t.NT * t.NT, ctx->stream));
    HIPCHK(hipMemsetAsync(ctx->d_info, 0, 4 * sizeof(int), ctx->stream));
    if ((rc = gen_band_tiles(A))) return rc;
    int NB = t.NB, p = maxrank + 10;
    if (p > NB) p = NB;
    long long ntask = (long long)t.NT * (t.NT - 1) / 2;
    if (ntask > 0 && t.NT > band_size) {
        std::vector<double> om((size_t)NB * p);
        sketch(NB, p, SKETCH_SEED, om.data());
        double *d_om = nullptr;
        HIPCHK(hipMalloc(&d_om, sizeof(double) * om.size()));
        HIPCHK(hipMemcpyAsync(d_om, om.data(), sizeof(double) * om.size(), hipMemcpyHostToDevice, ctx->stream));
        GenWs g;
        g.p = p;
        g.stride = gen_slot_doubles(NB, p, t.cap);
        int ns = nslots_for(ctx, g.stride * sizeof(double), ntask);
        if ((rc = ensure_ws(ctx, g.stride * sizeof(double) * ns))) return rc;
        size_t lds = sizeof(double) * (BLK_GEMM_SMEM + BLK_RED_SMEM + 2 * QRB * QRB + 3 * (size_t)p + 16);
        HIPCHK(hipFuncSetAttribute((const void *)k_compress, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
        hipLaunchKernelGGL(k_compress, dim3(ns), dim3(BS), lds, ctx->stream, t, (const double *)d_om, ctx->ws, g,
                           tol, maxrank, ctx->d_info);
        HIPCHK(hipGetLastError());
        HIPCHK(hipStreamSynchronize(ctx->stream));
        HIPCHK(hipFree(d_om));
    }
    HIPCHK(hipStreamSynchronize(ctx->stream));
    int hinfo[4];
    HIPCHK(hipMemcpy(hinfo, ctx->d_info, 4 * sizeof(int), hipMemcpyDeviceToHost));
    if (info) *info = hinfo[0];
    return 0;
}

extern "C" int hicma_get_rank_map(const hicma_tlr_desc *A, int *G)
{
    if (!A || !G) return HICMA_ERR_ARG;
    HIPCHK(hipSetDevice(A->ctx->device));
    HIPCHK(hipStreamSynchronize(A->ctx->stream));
    HIPCHK(hipMemcpy(G, A->t.rank, sizeof(int) * (size_t)A->t.NT * A->t.NT, hipMemcpyDeviceToHost));
    return 0;
}

extern "C" int parsec_rank_gather(hicma_ctx *ctx, hicma_tlr_desc *Ar, int *rank_array, int band_size)
{
    (void)ctx;
    if (!Ar || !rank_array) return HICMA_ERR_ARG;
    int NT = Ar->t.NT;
    std::vector<int> G((size_t)NT * NT);
    int rc = hicma_get_rank_map(Ar, G.data());
    if (rc) return rc;
    for (size_t i = 0; i < G.size(); i++) rank_array[i] = 0;
    for (int m = band_size; m < NT; m++)
        for (int n = 0; n <= m - band_size; n++) rank_array[(size_t)n * NT + m] = G[(size_t)n * NT + m];
    return 0;
}

extern "C" void HICMA_get_stat2(const int *G, int lda, int band_size, int *mn, int *mx, double *avg)
{
    int lo = 2147483647, hi = 0;
    long num = 0;
    long long sum = 0;
    for (int i = band_size; i < lda; i++)
        for (int j = 0; j < i - band_size + 1; j++) {
            int v = G[(size_t)j * lda + i];
            sum += v;
            num++;
            if (v < lo) lo = v;
            if (v > hi) hi = v;
        }
    if (mn) *mn = lo;
    if (mx) *mx = hi;
    if (avg) *avg = num ? (double)sum / (double)num : 0.0;
}

extern "C" int parsec_rank_check(hicma_ctx *ctx, hicma_tlr_desc *Ar, int band_size)
{
    if (!ctx || !Ar) return HICMA_ERR_ARG;
    int NT = Ar->t.NT;
    std::vector<int> G((size_t)NT * NT);
    int rc = hicma_get_rank_map(Ar, G.data());
    if (rc) return rc;
    int bad = 0;
    for (int m = band_size; m < NT; m++)
        for (int n = 0; n <= m - band_size; n++)
            if (G[(size_t)n * NT + m] < 0 && n < m) {
                fprintf(stderr, "ERROR rank of (%d, %d) : %d\n", m, n, G[(size_t)n * NT + m]);
                bad++;
            }
    TLR t = Ar->t;
    t.band_size = band_size;
    hipLaunchKernelGGL(k_rank_band, dim3((NT * NT + 255) / 256), dim3(256), 0, ctx->stream, t);
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(ctx->stream));
    return bad;
}

// band_size_auto_tuning.jdf:201-318 (op-count comparison per sub-diagonal)
extern "C" int hicma_band_autotune(const int *G, int NT, int NB)
{
    if (!G || NT <= 0 || NB <= 0) return HICMA_ERR_ARG;
    const double fluct = 0.66666667;
    int opt_min = 0, ind_min = 0;
    for (int b = 2; b < NT; b++) {
        long long tlr = 0, dense = 0;
        for (int m = b - 1; m < NT; m++) {
            int k = m + 1 - b;
            int ar = G[(size_t)k * NT + m];
            tlr += (long long)NB * NB * ar;
            dense += (long long)NB * NB * NB;
        }
        for (int k = 0; k <= NT - 3; k++)
            for (int m = k + b; m < NT; m++) {
                int n = m + 1 - b;
                int ar = G[(size_t)k * NT + m], br = G[(size_t)k * NT + n], cr = G[(size_t)n * NT + m];
                int s = cr + ar;
                unsigned long ops = hicma_op_counts('q', NB, s, 0, 0) + hicma_op_counts('m', ar, br, NB, 0) +
                                    hicma_op_counts('m', ar, NB, br, 0) + hicma_op_counts('q', NB, s, 0, 0) +
                                    hicma_op_counts('r', s, s, 2, 0) + hicma_op_counts('s', s, 0, 0, 0) +
                                    (unsigned long)s * cr + hicma_op_counts('o', NB, cr, s, 1) +
                                    hicma_op_counts('o', cr, NB, s, 2);
                tlr += (long long)ops;
                if (m - k < b) dense += 2LL * NB * NB * NB;
                else if (n - k < b && m - n < b) dense += 4LL * NB * NB * ar;
                else if (m - n < b) dense += 2LL * NB * NB * (ar < br ? ar : br) + 4LL * NB * ar * br;
            }
        if (tlr * fluct <= dense && ind_min == 0) {
            ind_min = 1;
            opt_min = b - 1;
        }
        if (tlr <= dense) return opt_min;
    }
    return NT;
}

extern "C" int parsec_band_size_auto_tuning(hicma_ctx *ctx, hicma_tlr_desc *Ar, hicma_tlr_desc *Fake,
                                            int *G, int NB)
{
    (void)ctx;
    (void)Fake;
    if (!Ar || !G) return HICMA_ERR_ARG;
    return hicma_band_autotune(G, Ar->t.NT, NB);
}

extern "C" int parsec_band_regenerate(hicma_ctx *ctx, int uplo, hicma_tlr_desc *A, int band_size)
{
    if (!ctx || !A) return HICMA_ERR_ARG;
    if (uplo != HICMA_LOWER) return HICMA_ERR_UPLO;
    if (band_size < 1 || band_size > A->t.NT) return HICMA_ERR_ARG;
    HIPCHK(hipSetDevice(ctx->device));
    HIPCHK(hipStreamSynchronize(ctx->stream));
    int rc = alloc_band(A, band_size);
    if (rc) return rc;
    if ((rc = gen_band_tiles(A))) return rc;
    HIPCHK(hipStreamSynchronize(ctx->stream));
    return 0;
}

// ---------------------------------------------------------------------------
// the factorization

// blocked Cholesky of diagonal tile k: per 64-column block, factor + invert
// the diagonal block in LDS, panel = C W^T, lower trailing update.  The block
// inverses (Winv, NB/64 blocks of 64x64) are reused by the low-rank trsm.
static int launch_potrf(hicma_tlr_desc *A, int k, double *Winv, OpLog o, hipStream_t s)
{
    TLR &t = A->t;
    for (int j0 = 0; j0 < t.NB; j0 += PBK) {
        int jb = std::min(PBK, t.NB - j0);
        hipLaunchKernelGGL(k_potrf_diag3, dim3(1), dim3(64), 0, s, t, k, j0, jb, Winv, A->ctx->d_info, o);
        int R = t.NB - j0 - jb;
        if (R > 0) {
            hipLaunchKernelGGL(k_potrf_panel2, dim3((R + PBK - 1) / PBK), dim3(BS), 0, s, t, k, j0, jb,
                               (const double *)Winv);
            int nt = (R + GT - 1) / GT;
            hipLaunchKernelGGL(k_potrf_update, dim3(nt * (nt + 1) / 2), dim3(BS), 0, s, t, k, j0, jb);
        }
    }
    HIPCHK(hipGetLastError());
    return 0;
}

extern "C" int HiCMA_dpotrf_L(hicma_ctx *ctx, int uplo, hicma_tlr_desc *A, hicma_tlr_desc *Ar, hicma_tlr_desc *Rank,
                              double acc, int fixed_rank, int maxrank, int *lookahead, int band_size, int hmb,
                              int compmaxrank, int send_full_tile, int *two_flow, unsigned long *tileopcounters,
                              unsigned long *opcounters, double *critical_path_time, int verbose)
{
    (void)Ar;
    (void)Rank;
    (void)hmb;
    (void)send_full_tile;
    if (!ctx || !A) return HICMA_ERR_ARG;
    // argument checks of HiCMA_dpotrf_L_2flow_New (wrapper.c:218-237), as error codes
    if (uplo != HICMA_LOWER) return HICMA_ERR_UPLO;
    if (compmaxrank <= 0) compmaxrank = maxrank;
    if (maxrank > compmaxrank) return HICMA_ERR_MAXRANK_COMP;
    if (maxrank > A->t.NB / 2) return HICMA_ERR_MAXRANK_NB;
    if (maxrank > A->t.cap || band_size != A->t.band_size) return HICMA_ERR_ARG;
    if (lookahead && *lookahead == -1) *lookahead = band_size;
    // 3-flow when band == 1 and not forced 2-flow, else 2-flow (hicma_parsec.c:517, 560-573)
    bool three_flow = (band_size == 1 && !(two_flow && *two_flow));
    if (!three_flow && two_flow) *two_flow = 1;
    if (verbose) fprintf(stderr, "%s version start\n", three_flow ? "3flow" : "2flow");

    HIPCHK(hipSetDevice(ctx->device));
    hipStream_t st = ctx->stream;
    TLR &t = A->t;
    const int NT = t.NT, NB = t.NB, cap = t.cap, b = band_size;
    HIPCHK(hipMemsetAsync(ctx->d_info, 0, 4 * sizeof(int), st));
    HIPCHK(hipMemsetAsync(A->d_tileops, 0, sizeof(unsigned long long) * (size_t)NT * NT, st));
    HIPCHK(hipMemsetAsync(A->d_opstat, 0, sizeof(unsigned long long) * 8, st));
    OpLog o{A->d_tileops, A->d_opstat};

    // workspaces: syrk (Y, Z per row), low-rank x dense gemm (W, X per pair),
    // recompression slots
    RcWs g;
    g.smax = 2 * cap;
    g.stride = rc_slot_doubles(NB, cap);
    size_t lds_fixed = sizeof(double) * (RC_FIXED_SMEM + 2 * (size_t)g.smax + 8);
    size_t lds_avail = ctx->lds_max > lds_fixed ? ctx->lds_max - lds_fixed : 0;
    int jl = (int)floor(sqrt((double)(lds_avail / sizeof(double)) / 2.0));
    if (jl > g.smax) jl = g.smax;
    g.jlds = jl;
    size_t rc_lds = lds_fixed + sizeof(double) * 2 * (size_t)jl * jl;
    long long maxtask = (long long)(NT - 1) * (NT - 2) / 2;
    int ns = nslots_for(ctx, g.stride * sizeof(double), maxtask > 0 ? maxtask : 1);
    size_t syrk_d = (size_t)NT * NB * cap + (size_t)NT * cap * cap;
    size_t lrd_pairs = (b > 1) ? (size_t)NT * (b - 1) : 0;
    size_t lrd_d = lrd_pairs * ((size_t)cap * NB + (size_t)cap * cap);
    // fast recompression path (1024 threads, one row per thread)
    const bool fast = (NB <= FT);
    size_t fstride = rcf_slot_doubles(NB);
    int nsf = fast ? nslots_for(ctx, fstride * sizeof(double), maxtask > 0 ? maxtask : 1) : 0;
    if (nsf > ctx->ncu) nsf = ctx->ncu;
    size_t tasks_d = ((size_t)(maxtask > 0 ? maxtask : 1) + 1) / 2 + 1;  // ints packed in doubles
    size_t winv_d = (size_t)((NB + PBK - 1) / PBK) * PBK * PBK;
    // one-wave recompression (tasks with s <= WV_SMAX): 2 workgroups per CU of
    // the s <= 48 variant (4 waves each) and of the s <= 64 variant (2 waves)
    const bool wave = (NB <= 64 * WV_KMAX);
    const int wgrid = wave ? ctx->ncu : 0;  // times WvCfg<SM>::BPC resident workgroups
    // CUs left to the lookahead stream (syrk(k) + potrf(k+1)) while the
    // recompression kernels of step k run
    const int wres = (wave && ctx->ncu > 64) ? 16 : 0;
    const size_t wstride = wv_slot_doubles(NB);
    const int wcap = (int)(maxtask > 0 ? maxtask : 1);
    // four concurrent variants (32, 48, 64 on side streams), each with up to
    // four resident waves per CU: one quarter of the region each
    size_t wws_d = (size_t)4 * wgrid * 4 * wstride;
    size_t wbs_d = wave ? (size_t)wgrid * wb_slot_doubles(NB) : 0;  // one wide-stack wave per CU
    size_t wlist_d = ((size_t)4 * wcap + 8) / 2 + 1;
    size_t total = g.stride * ns + syrk_d + lrd_d + fstride * nsf + tasks_d + winv_d + wws_d + wlist_d + wbs_d;
    int rc = ensure_ws(ctx, total * sizeof(double));
    if (rc) return rc;
    double *rcws = ctx->ws;
    double *Yws = rcws + g.stride * ns;
    double *Zws = Yws + (size_t)NT * NB * cap;
    double *Wws = Zws + (size_t)NT * cap * cap;
    double *Xws = Wws + lrd_pairs * cap * NB;
    double *fws = Xws + lrd_pairs * cap * cap;
    double *Winv = fws + fstride * nsf;
    int *tasks_s = (int *)(Winv + winv_d);
    double *wws = Winv + winv_d + tasks_d;
    int *wcount = (int *)(wws + wws_d);
    int *wlist = wcount + 8;  // counts [0..3], next-task counters [4..7]
    double *wbs = (double *)wcount + wlist_d;
    if (fast)
        HIPCHK(hipFuncSetAttribute((const void *)k_lr_gemm_fast, hipFuncAttributeMaxDynamicSharedMemorySize,
                                   (int)(RCF_LDS * sizeof(double))));
    HIPCHK(hipFuncSetAttribute((const void *)k_lr_gemm, hipFuncAttributeMaxDynamicSharedMemorySize, (int)rc_lds));
    const size_t wlds32 = sizeof(double) * WvCfg<32>::LDS * WvCfg<32>::WPB;
    const size_t wlds48 = sizeof(double) * WvCfg<48>::LDS * WvCfg<48>::WPB;
    const size_t wlds64 = sizeof(double) * WvCfg<64>::LDS * WvCfg<64>::WPB;
    if (wave) {
        HIPCHK(hipFuncSetAttribute((const void *)k_rc_wave<32>, hipFuncAttributeMaxDynamicSharedMemorySize,
                                   (int)wlds32));
        HIPCHK(hipFuncSetAttribute((const void *)k_rc_wave<48>, hipFuncAttributeMaxDynamicSharedMemorySize,
                                   (int)wlds48));
        HIPCHK(hipFuncSetAttribute((const void *)k_rc_wave<64>, hipFuncAttributeMaxDynamicSharedMemorySize,
                                   (int)wlds64));
        HIPCHK(hipFuncSetAttribute((const void *)k_rc_wide, hipFuncAttributeMaxDynamicSharedMemorySize,
                                   (int)(sizeof(double) * WB_LDS)));
    }

    std::vector<hipEvent_t> ev;
    bool timing = (critical_path_time != nullptr);
    auto mark = [&](void) {
        if (!timing) return;
        hipEvent_t e;
        hipEventCreate(&e);
        hipEventRecord(e, st);
        ev.push_back(e);
    };
    const int ntg = (NB + GT - 1) / GT;
    const int ntl = ntg * (ntg + 1) / 2;
    // Lookahead: syrk(k) and potrf(k+1) touch only diagonal tiles and the gemm
    // updates of step k only off-diagonal ones, so the former run on a side
    // stream while the latter run, and are joined back before trsm(k+1).
    hipStream_t sd = ctx->side[2];
    if (NT > 0 && (rc = launch_potrf(A, 0, Winv, o, st))) return rc;
    for (int k = 0; k < NT; k++) {
        mark();
        mark();
        int nbnd = std::min(b - 1, NT - 1 - k);
        int nlr = NT - k - b;
        if (nbnd > 0)
            hipLaunchKernelGGL(k_trsm_band, dim3(ntg, nbnd), dim3(BS), 0, st, t, k, o);
        if (nlr > 0)
            hipLaunchKernelGGL(k_trsm_lr3, dim3((cap + TRSM_TC - 1) / TRSM_TC, nlr), dim3(BS), 0, st, t, k, (const double *)Winv, o);
        mark();
        HIPCHK(hipEventRecord(ctx->ev_join[2], st));
        HIPCHK(hipStreamWaitEvent(sd, ctx->ev_join[2], 0));
        if (nbnd > 0) hipLaunchKernelGGL(k_syrk_band, dim3(ntl, nbnd), dim3(BS), 0, sd, t, k, o);
        if (nlr > 0) {
            hipLaunchKernelGGL(k_syrk_lr_prep, dim3(nlr), dim3(BS), 0, sd, t, k, Yws, Zws, o);
            hipLaunchKernelGGL(k_syrk_lr_update, dim3(ntl, nlr), dim3(BS), 0, sd, t, k, (const double *)Yws);
        }
        if (k + 1 < NT && (rc = launch_potrf(A, k + 1, Winv, o, sd))) return rc;
        mark();
        if (nbnd >= 2)
            hipLaunchKernelGGL(k_gemm_dense, dim3(ntg * ntg, nbnd * (nbnd - 1) / 2), dim3(BS), 0, st, t, k, o);
        if (b > 1 && nlr > 0) {
            int npair = nlr * (b - 1);
            hipLaunchKernelGGL(k_gemm_lrd_prep, dim3(npair), dim3(BS), 0, st, t, k, Wws, Xws, o);
            hipLaunchKernelGGL(k_gemm_lrd_update, dim3(ntg * ntg, npair), dim3(BS), 0, st, t, k,
                               (const double *)Wws);
        }
        int T = NT - 1 - k;
        long long ntask = (long long)T * (T - 1) / 2;
        if (ntask > 0 && T > b) {
            int grid = (int)std::min<long long>(ns, ntask);
            hipEvent_t e0 = nullptr, e1 = nullptr;
            if (ctx->profile) {
                HIPCHK(hipEventCreate(&e0));
                HIPCHK(hipEventCreate(&e1));
                HIPCHK(hipEventRecord(e0, st));
            }
            hipLaunchKernelGGL(k_classify, dim3((unsigned)((ntask + 255) / 256)), dim3(256), 0, st, t, k, tasks_s);
            if (wave) {
                HIPCHK(hipMemsetAsync(wcount, 0, 8 * sizeof(int), st));
                hipLaunchKernelGGL(k_classify_wave, dim3((unsigned)((ntask + 255) / 256)), dim3(256), 0, st, t, k,
                                   tasks_s, wlist, wcount, wcap);
                // fork: variants 32 and 48 on the context stream, 64 and the
                // wide one on side streams 0 and 1
                HIPCHK(hipEventRecord(ctx->ev_fork, st));
                for (int i = 0; i < 2; i++) HIPCHK(hipStreamWaitEvent(ctx->side[i], ctx->ev_fork, 0));
                int wg = (int)std::min<long long>((wgrid - wres) * WvCfg<32>::BPC, (ntask + WvCfg<32>::WPB - 1) / WvCfg<32>::WPB);
                hipLaunchKernelGGL(k_rc_wave<32>, dim3(wg), dim3(64 * WvCfg<32>::WPB), wlds32, st, t, k,
                                   (const int *)wlist, (const int *)wcount, wcount + 4, wws, acc, fixed_rank, maxrank,
                                   o, ctx->d_ovf);
                wg = (int)std::min<long long>((wgrid - wres) * WvCfg<48>::BPC, (ntask + WvCfg<48>::WPB - 1) / WvCfg<48>::WPB);
                hipLaunchKernelGGL(k_rc_wave<48>, dim3(wg), dim3(64 * WvCfg<48>::WPB), wlds48, st, t, k,
                                   (const int *)(wlist + wcap), (const int *)(wcount + 1), wcount + 5, wws + wws_d / 4,
                                   acc, fixed_rank, maxrank, o, ctx->d_ovf);
                wg = (int)std::min<long long>((wgrid - wres) * WvCfg<64>::BPC, (ntask + WvCfg<64>::WPB - 1) / WvCfg<64>::WPB);
                hipLaunchKernelGGL(k_rc_wave<64>, dim3(wg), dim3(64 * WvCfg<64>::WPB), wlds64, ctx->side[0], t, k,
                                   (const int *)(wlist + 2 * (size_t)wcap), (const int *)(wcount + 2), wcount + 6,
                                   wws + wws_d / 2, acc, fixed_rank, maxrank, o, ctx->d_ovf);
                wg = (int)std::min<long long>(wgrid - wres, ntask);
                hipLaunchKernelGGL(k_rc_wide, dim3(wg), dim3(64 * WB_WAVES), sizeof(double) * WB_LDS, ctx->side[1], t, k,
                                   (const int *)(wlist + 3 * (size_t)wcap), (const int *)(wcount + 3), wcount + 7, wbs,
                                   acc, fixed_rank, maxrank, o, ctx->d_ovf);
                // join
                for (int i = 0; i < 2; i++) {
                    HIPCHK(hipEventRecord(ctx->ev_join[i], ctx->side[i]));
                    HIPCHK(hipStreamWaitEvent(st, ctx->ev_join[i], 0));
                }
            }
            if (fast) {
                int gf = (int)std::min<long long>(nsf, ntask);
                hipLaunchKernelGGL(k_lr_gemm_fast, dim3(gf), dim3(FT), RCF_LDS * sizeof(double), st, t, k,
                                   (const int *)tasks_s, fws, fstride, acc, fixed_rank, maxrank, o, ctx->d_ovf);
            }
            hipLaunchKernelGGL(k_lr_gemm, dim3(grid), dim3(BS), rc_lds, st, t, k, (const int *)tasks_s,
                               fast ? FS : -1, rcws, g, acc, fixed_rank, maxrank, o, ctx->d_ovf);
            if (ctx->profile) {
                HIPCHK(hipEventRecord(e1, st));
                ctx->prof_ev.push_back(e0);
                ctx->prof_ev.push_back(e1);
            }
        }
        HIPCHK(hipEventRecord(ctx->ev_join[2], sd));
        HIPCHK(hipStreamWaitEvent(st, ctx->ev_join[2], 0));
        HIPCHK(hipGetLastError());
    }
    mark();
    HIPCHK(hipStreamSynchronize(st));
    int hinfo[4];
    HIPCHK(hipMemcpy(hinfo, ctx->d_info, 4 * sizeof(int), hipMemcpyDeviceToHost));
    unsigned long long stats[8];
    HIPCHK(hipMemcpy(stats, A->d_opstat, sizeof(stats), hipMemcpyDeviceToHost));
    for (int i = 0; i < 8; i++) A->last_stats[i] = (unsigned long)stats[i];
    A->last_ovf = hinfo[1];
    if (tileopcounters) {
        std::vector<unsigned long long> to((size_t)NT * NT);
        HIPCHK(hipMemcpy(to.data(), A->d_tileops, sizeof(unsigned long long) * to.size(), hipMemcpyDeviceToHost));
        for (size_t i = 0; i < to.size(); i++) tileopcounters[i] = (unsigned long)to[i];
    }
    if (opcounters) opcounters[0] = (unsigned long)(stats[0] + stats[1]);
    if (timing) {
        double ts[3] = {0, 0, 0};
        for (int k = 0; k < NT; k++) {
            for (int ph = 0; ph < 3; ph++) {
                float ms = 0.f;
                hipEventElapsedTime(&ms, ev[(size_t)k * 4 + ph], ev[(size_t)k * 4 + ph + 1]);
                ts[ph] += ms * 1e-3;
            }
        }
        for (int i = 0; i < 6; i++) critical_path_time[i] = i < 3 ? ts[i] : 0.0;
        for (auto e : ev) hipEventDestroy(e);
    }
    if (verbose && hinfo[1])
        fprintf(stderr, "hicma_tlr: %d recompressions capped at maxrank %d\n", hinfo[1], maxrank);
    return hinfo[0];
}

extern "C" int hicma_op_stats(const hicma_tlr_desc *A, unsigned long *stats)
{
    if (!A || !stats) return HICMA_ERR_ARG;
    for (int i = 0; i < 4; i++) stats[i] = A->last_stats[i];
    return 0;
}

extern "C" int hicma_recompress_stats(const hicma_tlr_desc *A, unsigned long *stats)
{
    if (!A || !stats) return HICMA_ERR_ARG;
    for (int i = 0; i < 3; i++) stats[i] = A->last_stats[4 + i];
    return 0;
}

extern "C" int hicma_profile(hicma_ctx *ctx, int enable)
{
    if (!ctx) return HICMA_ERR_ARG;
    for (auto e : ctx->prof_ev) (void)hipEventDestroy(e);
    ctx->prof_ev.clear();
    ctx->profile = enable;
    return 0;
}

extern "C" int hicma_profile_read(hicma_ctx *ctx, double *ms_total, int *launches)
{
    if (!ctx) return HICMA_ERR_ARG;
    HIPCHK(hipStreamSynchronize(ctx->stream));
    double tot = 0.0;
    for (size_t i = 0; i + 1 < ctx->prof_ev.size(); i += 2) {
        float ms = 0.f;
        HIPCHK(hipEventElapsedTime(&ms, ctx->prof_ev[i], ctx->prof_ev[i + 1]));
        tot += ms;
    }
    if (ms_total) *ms_total = tot;
    if (launches) *launches = (int)(ctx->prof_ev.size() / 2);
    return 0;
}

extern "C" int hicma_desc_copy(hicma_tlr_desc *dst, const hicma_tlr_desc *src)
{
    if (!dst || !src) return HICMA_ERR_ARG;
    const TLR &a = src->t;
    TLR &b = dst->t;
    if (a.N != b.N || a.NB != b.NB || a.cap != b.cap || a.kind != b.kind) return HICMA_ERR_ARG;
    hicma_ctx *c = dst->ctx;
    HIPCHK(hipSetDevice(c->device));
    HIPCHK(hipStreamSynchronize(src->ctx->stream));
    int rc = alloc_band(dst, a.band_size);
    if (rc) return rc;
    size_t nb2 = (size_t)a.NB * a.NB, slots = (size_t)a.NT * (a.NT - 1) / 2;
    HIPCHK(hipMemcpyAsync(b.band, a.band, sizeof(double) * nb2 * a.NT * a.band_size, hipMemcpyDeviceToDevice,
                          c->stream));
    if (slots)
        HIPCHK(hipMemcpyAsync(b.lr, a.lr, sizeof(double) * slots * a.NB * 2 * a.cap, hipMemcpyDeviceToDevice,
                              c->stream));
    HIPCHK(hipMemcpyAsync(b.rank, a.rank, sizeof(int) * (size_t)a.NT * a.NT, hipMemcpyDeviceToDevice, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    return 0;
}

extern "C" int hicma_rank_overflows(const hicma_tlr_desc *A) { return A ? A->last_ovf : 0; }

// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(BS) k_uncompress(TLR t, int m, int n, double *out)
{
    __shared__ double sm[BLK_GEMM_SMEM];
    const int NB = t.NB;
    int nt = (NB + GT - 1) / GT;
    int ti = blockIdx.x % nt, tj = blockIdx.x / nt;
    if (m - n < t.band_size) {
        const double *T = band_tile(t, m, n);
        for (int e = threadIdx.x; e < GT * GT; e += BS) {
            int i = ti * GT + e % GT, j = tj * GT + e / GT;
            if (i < NB && j < NB) out[i + (size_t)j * NB] = T[i + (size_t)j * NB];
        }
        return;
    }
    int r = *rank_at(t, m, n);
    const double *U = lr_tile(t, m, n);
    gemm_tile<false, true>(ti * GT, tj * GT, NB, NB, r, 1.0, U, NB, U + (size_t)NB * r, NB, 0.0, out, NB, sm);
}

extern "C" int STARSH_check(hicma_ctx *ctx, int uplo, double *A0, hicma_tlr_desc *A, hicma_tlr_desc *Ar,
                            int band_size, int maxrank, int *info)
{
    (void)Ar;
    (void)maxrank;
    (void)band_size;
    if (!ctx || !A || !A0) return HICMA_ERR_ARG;
    if (uplo != HICMA_LOWER) return HICMA_ERR_UPLO;
    HIPCHK(hipSetDevice(ctx->device));
    TLR &t = A->t;
    const int N = t.N, NB = t.NB, NT = t.NT;
    double *d_tile = nullptr;
    HIPCHK(hipMalloc(&d_tile, sizeof(double) * (size_t)NB * NB));
    std::vector<double> h((size_t)NB * NB);
    memset(A0, 0, sizeof(double) * (size_t)N * N);
    int nt = (NB + GT - 1) / GT;
    for (int n = 0; n < NT; n++)
        for (int m = n; m < NT; m++) {
            hipLaunchKernelGGL(k_uncompress, dim3(nt * nt), dim3(BS), 0, ctx->stream, t, m, n, d_tile);
            HIPCHK(hipGetLastError());
            HIPCHK(hipMemcpyAsync(h.data(), d_tile, sizeof(double) * h.size(), hipMemcpyDeviceToHost, ctx->stream));
            HIPCHK(hipStreamSynchronize(ctx->stream));
            int mm = std::min(NB, N - m * NB), nn = std::min(NB, N - n * NB);
            for (int j = 0; j < nn; j++)
                for (int i = 0; i < mm; i++) {
                    size_t gi = (size_t)m * NB + i, gj = (size_t)n * NB + j;
                    if (gi >= gj) A0[gi + gj * N] = h[i + (size_t)j * NB];
                }
        }
    HIPCHK(hipFree(d_tile));
    if (info) *info = 0;
    return 0;
}

extern "C" int hicma_get_tile(const hicma_tlr_desc *A, int m, int n, double *buf, int *rank)
{
    if (!A || !buf || m < n || m >= A->t.NT || n < 0) return HICMA_ERR_ARG;
    hicma_ctx *c = A->ctx;
    HIPCHK(hipSetDevice(c->device));
    HIPCHK(hipStreamSynchronize(c->stream));
    const TLR &t = A->t;
    size_t nb2 = (size_t)t.NB * t.NB;
    if (m - n < t.band_size) {
        const double *src = t.band + ((size_t)(m - n) * t.NT + m) * nb2;
        HIPCHK(hipMemcpy(buf, src, sizeof(double) * nb2, hipMemcpyDeviceToHost));
        if (rank) *rank = t.NB;
        return 0;
    }
    int r = 0;
    HIPCHK(hipMemcpy(&r, t.rank + (size_t)n * t.NT + m, sizeof(int), hipMemcpyDeviceToHost));
    if (rank) *rank = r;
    if (r > 0) {
        const double *src = t.lr + ((size_t)m * (m - 1) / 2 + n) * (size_t)t.NB * 2 * t.cap;
        HIPCHK(hipMemcpy(buf, src, sizeof(double) * (size_t)t.NB * 2 * r, hipMemcpyDeviceToHost));
    }
    return 0;
}

extern "C" int hicma_set_tile(hicma_tlr_desc *A, int m, int n, const double *buf, int rank)
{
    if (!A || !buf || m < n || m >= A->t.NT || n < 0) return HICMA_ERR_ARG;
    hicma_ctx *c = A->ctx;
    HIPCHK(hipSetDevice(c->device));
    HIPCHK(hipStreamSynchronize(c->stream));
    const TLR &t = A->t;
    size_t nb2 = (size_t)t.NB * t.NB;
    if (m - n < t.band_size) {
        double *dst = t.band + ((size_t)(m - n) * t.NT + m) * nb2;
        HIPCHK(hipMemcpy(dst, buf, sizeof(double) * nb2, hipMemcpyHostToDevice));
        return 0;
    }
    if (rank < 0 || rank > t.cap) return HICMA_ERR_ARG;
    double *dst = t.lr + ((size_t)m * (m - 1) / 2 + n) * (size_t)t.NB * 2 * t.cap;
    if (rank > 0) HIPCHK(hipMemcpy(dst, buf, sizeof(double) * (size_t)t.NB * 2 * rank, hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(t.rank + (size_t)n * t.NT + m, &rank, sizeof(int), hipMemcpyHostToDevice));
    return 0;
}

extern "C" int hicma_logdet(hicma_tlr_desc *A, double *logdet)
{
    if (!A || !logdet) return HICMA_ERR_ARG;
    hicma_ctx *c = A->ctx;
    HIPCHK(hipSetDevice(c->device));
    hipLaunchKernelGGL(k_logdet, dim3(1), dim3(BS), 0, c->stream, A->t, c->d_scalar);
    HIPCHK(hipGetLastError());
    HIPCHK(hipMemcpyAsync(logdet, c->d_scalar, sizeof(double), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    return 0;
}
