// hicma_tlr.hip -- host side of the MI355X TLR Cholesky library (C-ABI).
//
// One HIP stream per context; the PaRSEC task graph of HiCMA_dpotrf_L is
// replaced by a panel-step schedule of batched launches whose grids cover
// every task of a step; tile ranks never leave HBM (kernels read them).
#include "internal.hpp"
#include "tlr_kernels.hpp"
#include "rc_fast.hpp"
#include "rc_wave.hpp"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#include <algorithm>
#include <chrono>

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
// widths of the first two generation sketches (ranks up to 54 and 150 are
// settled by them; wider ranks by the full maxrank + 10)
#define GEN_FIRST_SKETCH 64
#define GEN_SECOND_SKETCH 160

// z ~ N(0, I): the same counter-based normal stream as the sketch
extern "C" int hicma_normal(uint64_t seed, int n, double *z)
{
    if (n < 0 || (!z && n)) return HICMA_ERR_ARG;
    sketch(n, 1, seed, z);
    return 0;
}

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
extern "C" int hicma_init_ex(hicma_ctx **ctx, int device, int flags, int cu_share)
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
    c->ncu_dev = prop.multiProcessorCount;
    c->ncu = (cu_share > 0 && cu_share < c->ncu_dev) ? cu_share : c->ncu_dev;
    c->single = (flags & HICMA_CTX_SINGLE_STREAM) != 0;
    c->pool = (flags & HICMA_CTX_STREAM_ORDERED_ALLOC) != 0;
    c->lds_max = prop.sharedMemPerBlock;
    HIPCHK(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
    for (int i = 0; i < 3; i++) {
        if (c->single) c->side[i] = c->stream;
        else HIPCHK(hipStreamCreateWithFlags(&c->side[i], hipStreamNonBlocking));
        HIPCHK(hipEventCreateWithFlags(&c->ev_join[i], hipEventDisableTiming));
    }
    HIPCHK(hipEventCreateWithFlags(&c->ev_fork, hipEventDisableTiming));
    if (c->pool) {
        // keep freed blocks in the device's pool instead of returning them,
        // and never let one context's allocation reuse another stream's
        // freed block by making its stream wait for that one (the contexts
        // of a server would then serialize)
        hipMemPool_t mp;
        HIPCHK(hipDeviceGetDefaultMemPool(&mp, device));
        uint64_t thr = ~0ULL;
        HIPCHK(hipMemPoolSetAttribute(mp, hipMemPoolAttrReleaseThreshold, &thr));
        int no = 0;
        HIPCHK(hipMemPoolSetAttribute(mp, hipMemPoolReuseAllowInternalDependencies, &no));
    }
    HIPCHK(hipMalloc(&c->d_info, 4 * sizeof(int)));
    HIPCHK(hipMemset(c->d_info, 0, 4 * sizeof(int)));
    c->d_ovf = c->d_info + 1;
    HIPCHK(hipMalloc(&c->d_scalar, 8 * sizeof(double)));
    RCHK(hicma_noalloc_setup(device));
    *ctx = c;
    return 0;
}

extern "C" int hicma_init(hicma_ctx **ctx, int device) { return hicma_init_ex(ctx, device, 0, 0); }

hipError_t ctx_malloc(hicma_ctx *c, void **p, size_t bytes)
{
    if (!c->pool) return hipMalloc(p, bytes);
    hipError_t e = hipMallocAsync(p, bytes, c->stream);
    if (e != hipSuccess) return e;
    return hipStreamSynchronize(c->stream);
}

hipError_t ctx_free(hicma_ctx *c, void *p)
{
    if (!p) return hipSuccess;
    return c->pool ? hipFreeAsync(p, c->stream) : hipFree(p);
}

extern "C" int hicma_fini(hicma_ctx *c)
{
    if (!c) return 0;
    hipSetDevice(c->device);
    if (c->ws) ctx_free(c, c->ws);
    hipStreamSynchronize(c->stream);
    hipFree(c->d_info);
    hipFree(c->d_scalar);
    for (auto e : c->prof_ev) (void)hipEventDestroy(e);
    for (auto e : c->cp_ev) (void)hipEventDestroy(e);
    if (c->blas && c->blas_free) c->blas_free(c->blas);
    for (int i = 0; i < 3; i++) {
        (void)hipEventDestroy(c->ev_join[i]);
        if (!c->single) (void)hipStreamDestroy(c->side[i]);
    }
    (void)hipEventDestroy(c->ev_fork);
    hipStreamDestroy(c->stream);
    delete c;
    return 0;
}

int hicma_ctx_release_ws(hicma_ctx *c)
{
    if (c->ws) HIPCHK(ctx_free(c, c->ws));
    c->ws = nullptr;
    c->ws_bytes = 0;
    return 0;
}

int hicma_ensure_ws(hicma_ctx *c, size_t bytes)
{
    if (c->ws_bytes >= bytes) return 0;
    if (c->ws) HIPCHK(ctx_free(c, c->ws));
    c->ws = nullptr;
    c->ws_bytes = 0;
    HIPCHK(ctx_malloc(c, (void **)&c->ws, bytes));
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
    A->h_xyz0 = (double *)malloc(sizeof(double) * (size_t)t.ndim * N);
    hicma_points(N, t.ndim, seed, A->h_xyz0);
    memcpy(A->h_xyz, A->h_xyz0, sizeof(double) * (size_t)t.ndim * N);
    double *dx = nullptr;
    int NT = t.NT;
    size_t lr_slots = (size_t)NT * (NT - 1) / 2;
    if (ctx_malloc(ctx, (void **)&dx, sizeof(double) * (size_t)t.ndim * N) != hipSuccess ||
        ctx_malloc(ctx, (void **)&t.rank, sizeof(int) * (size_t)NT * NT) != hipSuccess ||
        ctx_malloc(ctx, (void **)&t.prec, (size_t)NT * NT) != hipSuccess ||
        (lr_slots && ctx_malloc(ctx, (void **)&t.lr, sizeof(double) * lr_slots * NB * 2 * maxrank) != hipSuccess) ||
        ctx_malloc(ctx, (void **)&A->d_tileops, sizeof(unsigned long long) * (size_t)NT * NT) != hipSuccess ||
        ctx_malloc(ctx, (void **)&A->d_opstat, sizeof(unsigned long long) * 16) != hipSuccess) {
        (void)hipGetLastError();
        ctx_free(ctx, dx);
        ctx_free(ctx, t.rank);
        ctx_free(ctx, t.prec);
        ctx_free(ctx, t.lr);
        ctx_free(ctx, A->d_tileops);
        free(A->h_xyz);
        free(A->h_xyz0);
        free(A);
        return HICMA_ERR_NOMEM;
    }
    HIPCHK(hipMemcpy(dx, A->h_xyz, sizeof(double) * (size_t)t.ndim * N, hipMemcpyHostToDevice));
    t.xyz = dx;
    HIPCHK(hipMemsetAsync(t.rank, 0xFF, sizeof(int) * (size_t)NT * NT, ctx->stream));
    HIPCHK(hipMemsetAsync(t.prec, LR_F64, (size_t)NT * NT, ctx->stream));
    HIPCHK(hipStreamSynchronize(ctx->stream));
    *out = A;
    return 0;
}

extern "C" int hicma_desc_destroy(hicma_tlr_desc *A)
{
    if (!A) return 0;
    hicma_ctx *c = A->ctx;
    hipSetDevice(c->device);
    ctx_free(c, (void *)A->t.xyz);
    ctx_free(c, A->t.rank);
    ctx_free(c, A->t.prec);
    ctx_free(c, A->t.lr);
    ctx_free(c, A->t.band);
    ctx_free(c, A->d_tileops);
    ctx_free(c, A->d_opstat);
    ctx_free(c, A->rhist);
    if (A->gexec) (void)hipGraphExecDestroy(A->gexec);
    if (A->graph) (void)hipGraphDestroy(A->graph);
    free(A->h_xyz);
    free(A->h_xyz0);
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
        if (t.band) HIPCHK(ctx_free(A->ctx, t.band));
        t.band = nullptr;
        A->band_alloc = 0;
        HIPCHK(ctx_malloc(A->ctx, (void **)&t.band, sizeof(double) * (size_t)band_size * t.NT * t.NB * t.NB));
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

// STARSH_gen's data / kernel / index arguments (the STARS-H problem of the
// reference, hicma_parsec/STARSH_gen.jdf:342-369): NULL keeps the
// descriptor's own kernel parameters, kernel kind and point order; otherwise
// they replace them.  index[i] is the point (in creation order) of matrix row i.
static int apply_problem(hicma_tlr_desc *A, const void *data, const hicma_kernel *kernel, const hicma_index *index)
{
    TLR &t = A->t;
    if (kernel) {
        int k = *kernel;
        if (k != HICMA_KIND_SQEXP_2D && k != HICMA_KIND_SQEXP_3D && k != HICMA_KIND_EXP_3D) return HICMA_ERR_ARG;
        if (kind_ndim(k) != t.ndim) return HICMA_ERR_ARG;  // the point set has a fixed dimension
        t.kind = k;
    }
    if (data) {
        const hicma_kernel_params *p = (const hicma_kernel_params *)data;
        t.sigma2 = p->sigma2;
        t.beta = p->beta;
        t.nugget = p->nugget;
    }
    if (index) {
        const int N = t.N, nd = t.ndim;
        std::vector<char> seen((size_t)N, 0);
        for (int i = 0; i < N; i++) {
            if (index[i] < 0 || index[i] >= N || seen[(size_t)index[i]]) return HICMA_ERR_ARG;
            seen[(size_t)index[i]] = 1;
        }
        for (int d = 0; d < nd; d++)
            for (int i = 0; i < N; i++) A->h_xyz[(size_t)d * N + i] = A->h_xyz0[(size_t)d * N + (size_t)index[i]];
        HIPCHK(hipMemcpy((void *)t.xyz, A->h_xyz, sizeof(double) * (size_t)nd * N, hipMemcpyHostToDevice));
    }
    return 0;
}

// The compression passes first .. last of STARSH_gen over the off-band
// tiles (outside A's current band): pass 0 compresses every tile with the
// narrow sketch, passes 1 and 2 again, at wider sketches up to the full
// width maxrank + GEN_OVS, exactly the tiles the narrower one could not vouch
// for (rank map GEN_REDO).
static int gen_lr(hicma_tlr_desc *A, double tol, int maxrank, int first, int last)
{
    hicma_ctx *ctx = A->ctx;
    TLR &t = A->t;
    int rc = 0;
    int NB = t.NB, p = maxrank + 10;
    if (p > NB) p = NB;
    long long ntask = (long long)t.NT * (t.NT - 1) / 2;
    if (ntask > 0 && t.NT > t.band_size) {
        std::vector<double> om((size_t)NB * p);
        sketch(NB, p, SKETCH_SEED, om.data());
        double *d_om = nullptr;
        HIPCHK(ctx_malloc(ctx, (void **)&d_om, sizeof(double) * om.size()));
        HIPCHK(hipMemcpy(d_om, om.data(), sizeof(double) * om.size(), hipMemcpyHostToDevice));  // stream idle (ctx_malloc)
        GenWs g;
        g.p = p;
        g.stride = gen_slot_doubles(NB, p, t.cap);
        int ns = nslots_for(ctx, g.stride * sizeof(double), ntask);
        if ((rc = hicma_ensure_ws(ctx, g.stride * sizeof(double) * ns))) return rc;
        size_t lds = sizeof(double) * (BLK_GEMM_SMEM + BLK_RED_SMEM + 2 * QRB * QRB + 3 * (size_t)p + 16);
        HIPCHK(hipFuncSetAttribute((const void *)k_compress, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
        // narrow sketch first; tiles whose rank it cannot vouch for again at
        // a wider one, up to the full width
        const int ladder[3] = {GEN_FIRST_SKETCH, GEN_SECOND_SKETCH, p};
        int prev = 0;
        for (int pass = 0; pass <= last; pass++) {
            const int pw = ladder[pass] < p ? ladder[pass] : p;
            if (pw <= prev) continue;
            if (pass >= first)
                hipLaunchKernelGGL(k_compress, dim3(ns), dim3(BS), lds, ctx->stream, t, (const double *)d_om,
                                   ctx->ws, g, pw, pass > 0 ? 1 : 0, tol, maxrank, ctx->d_info);
            prev = pw;
        }
        HIPCHK(hipGetLastError());
        HIPCHK(hipStreamSynchronize(ctx->stream));
        HIPCHK(ctx_free(ctx, d_om));
    }
    return 0;
}

extern "C" int STARSH_gen(hicma_ctx *ctx, int uplo, hicma_tlr_desc *A, hicma_tlr_desc *Ar, void *data,
                          hicma_kernel *kernel, hicma_index *index, double tol, int maxrank, int band_size,
                          int send_full_tile, int *info)
{
    (void)Ar;
    (void)send_full_tile;
    if (!ctx || !A) return HICMA_ERR_ARG;
    if (uplo != HICMA_LOWER && uplo != HICMA_UPPER) return HICMA_ERR_UPLO;
    if (band_size < 1 || maxrank < 1 || maxrank > A->t.cap) return HICMA_ERR_ARG;
    HIPCHK(hipSetDevice(ctx->device));
    RCHK(apply_problem(A, data, kernel, index));
    TLR &t = A->t;
    int rc = alloc_band(A, band_size);
    if (rc) return rc;
    HIPCHK(hipMemsetAsync(t.rank, 0xFF, sizeof(int) * (size_t)t.NT * t.NT, ctx->stream));
    HIPCHK(hipMemsetAsync(t.prec, LR_F64, (size_t)t.NT * t.NT, ctx->stream));
    A->demoted = 0;
    HIPCHK(hipMemsetAsync(ctx->d_info, 0, 4 * sizeof(int), ctx->stream));
    if ((rc = gen_band_tiles(A))) return rc;
    if ((rc = gen_lr(A, tol, maxrank, 0, 2))) return rc;
    HIPCHK(hipStreamSynchronize(ctx->stream));
    int hinfo[4];
    HIPCHK(hipMemcpy(hinfo, ctx->d_info, 4 * sizeof(int), hipMemcpyDeviceToHost));
    if (info) *info = hinfo[0];
    return 0;
}

// STARSH_gen in two stages for a caller that picks the band from the ranks
// (the serving loop): the first compresses every off-band tile of band 1
// with the narrow sketch and leaves the tiles it cannot vouch for at rank
// HICMA_RANK_PENDING; after the caller's parsec_band_regenerate, the second
// runs the wider passes on the pending tiles still outside the band.  The
// off-band tiles end exactly as STARSH_gen leaves them; the pending tiles
// the band took in are never compressed at the wider sketches.
extern "C" int hicma_gen_first(hicma_ctx *ctx, hicma_tlr_desc *A, double tol, int maxrank, int *info)
{
    if (!ctx || !A || maxrank < 1 || maxrank > A->t.cap) return HICMA_ERR_ARG;
    HIPCHK(hipSetDevice(ctx->device));
    TLR &t = A->t;
    int rc = alloc_band(A, 1);
    if (rc) return rc;
    HIPCHK(hipMemsetAsync(t.rank, 0xFF, sizeof(int) * (size_t)t.NT * t.NT, ctx->stream));
    HIPCHK(hipMemsetAsync(t.prec, LR_F64, (size_t)t.NT * t.NT, ctx->stream));
    A->demoted = 0;
    HIPCHK(hipMemsetAsync(ctx->d_info, 0, 4 * sizeof(int), ctx->stream));
    if ((rc = gen_band_tiles(A))) return rc;
    if ((rc = gen_lr(A, tol, maxrank, 0, 0))) return rc;
    HIPCHK(hipStreamSynchronize(ctx->stream));
    int hinfo[4];
    HIPCHK(hipMemcpy(hinfo, ctx->d_info, 4 * sizeof(int), hipMemcpyDeviceToHost));
    if (info) *info = hinfo[0];
    return 0;
}

extern "C" int hicma_gen_finish(hicma_ctx *ctx, hicma_tlr_desc *A, double tol, int maxrank, int *info)
{
    if (!ctx || !A || maxrank < 1 || maxrank > A->t.cap) return HICMA_ERR_ARG;
    HIPCHK(hipSetDevice(ctx->device));
    int rc = gen_lr(A, tol, maxrank, 1, 2);
    if (rc) return rc;
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
extern "C" int hicma_band_autotune(const int *G, int NT, int NB) { return hicma_band_autotune_weighted(G, NT, NB, 1.0); }

extern "C" int hicma_band_autotune_weighted(const int *G, int NT, int NB, double lr_weight)
{
    if (!G || NT <= 0 || NB <= 0 || !(lr_weight > 0)) return HICMA_ERR_ARG;
    const double fluct = 0.66666667;
    int opt_min = 0, ind_min = 0;
    for (int b = 2; b < NT; b++) {
        long long tlr = 0, dense = 0;
        for (int m = b - 1; m < NT; m++) {
            int k = m + 1 - b;
            int ar = G[(size_t)k * NT + m];
            tlr += (long long)(lr_weight * (double)NB * NB * ar);
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
                tlr += (long long)(lr_weight * (double)ops);
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

// Band choice by a time model: every factorization task of a band size b is
// priced at the device rate of the kernel that would run it (dense tiles and
// low-rank products on the matrix cores at dense_tflops; recompressions at
// their reference op counts and lr_tflops, MODEL_WIDE_SLOWDOWN times slower
// for stacks wider than the wave kernels' 128 columns, which go to the
// 256-thread kernel with an HBM-resident SVD), plus MODEL_STEP_2FLOW_S per
// panel step for the dense-band kernels of the 2-flow form, and the b with
// the least total time wins.  A gemm task (m, n, k) changes kind only where
// b passes m - n, n - k or m - k, so each task adds its time to at most four
// ranges of b (a difference array): O(NT^3) in all.  The rank a recompressed
// tile reaches during the factorization is estimated as MODEL_RANK_GROWTH
// times its generated rank.  The constants are fitted to measured
// factorization times of 2D sqexp problems (tile 512, maxrank 256, N 16 k ..
// 64 k, beta 0.02 .. 0.2) on MI355X.
#define MODEL_WIDE_SLOWDOWN 300.0
#define MODEL_STEP_2FLOW_S 5e-4
#define MODEL_RANK_GROWTH 1.25
extern "C" int hicma_band_autotune_model(const int *G, int NT, int NB, double dense_tflops, double lr_tflops)
{
    if (!G || NT <= 0 || NB <= 0 || !(dense_tflops > 0) || !(lr_tflops > 0)) return HICMA_ERR_ARG;
    const double nb = NB, rd = 1.0 / (dense_tflops * 1e12), rl = 1.0 / (lr_tflops * 1e12);
    std::vector<double> D((size_t)NT + 2, 0.0);  // D[b]: time added from band b on
    auto add = [&](int lo, int hi, double t) {     // bands lo .. hi (inclusive, clipped to 1 .. NT)
        lo = std::max(lo, 1);
        hi = std::min(hi, NT);
        if (lo > hi) return;
        D[lo] += t;
        D[hi + 1] -= t;
    };
    auto rk = [&](int m, int n) { return std::max(G[(size_t)n * NT + m], 0); };
    auto lr_case_v = [&](double ra, double rb, double rc) {
        const unsigned long s = (unsigned long)(ra + rc), N = (unsigned long)NB, c = (unsigned long)rc;
        const double ops = (double)(hicma_op_counts('q', N, s, 0, 0) * 2 + hicma_op_counts('m', (unsigned long)ra, (unsigned long)rb, N, 0) +
                                    hicma_op_counts('m', (unsigned long)ra, N, (unsigned long)rb, 0) +
                                    hicma_op_counts('r', s, s, 2, 0) + hicma_op_counts('s', s, 0, 0, 0) + s * c +
                                    hicma_op_counts('o', N, c, s, 1) + hicma_op_counts('o', c, N, s, 2));
        return ops;
    };
    for (int k = 0; k < NT; k++) {
        for (int m = k + 1; m < NT; m++) {
            const double r = rk(m, k), d = m - k;
            // trsm of tile (m, k) and its syrk into (m, m): low rank while b <= d
            add(1, (int)d, (nb * nb * r + 2 * nb * nb * r) * rd);
            add((int)d + 1, NT, (nb * nb * nb + nb * nb * nb) * rd);
        }
        for (int n = k + 1; n < NT; n++)
            for (int m = n + 1; m < NT; m++) {
                const int x = m - n, y = n - k, z = m - k;
                const double ra = rk(m, k), rb = rk(n, k), rc = MODEL_RANK_GROWTH * rk(m, n);
                const double t_all_dense = 2 * nb * nb * nb * rd;
                const double t_cd_bd = 4 * nb * nb * ra * rd;                      // C, B dense; A low rank
                const double t_cd = (4 * nb * ra * rb + 2 * nb * nb * rb) * rd;    // C dense; A, B low rank
                // all low rank, and C, A low rank with B dense: stacks wider
                // than 128 columns (or B of rank > 64) leave the wave
                // kernels for the 256-thread one
                const bool wide = (ra + rc > 128);
                const double t_v = lr_case_v(ra, rb, rc) * ((wide || rb > 64) ? MODEL_WIDE_SLOWDOWN * rl : rl);
                const double t_iv =
                    lr_case_v(ra, 0, rc) * (wide ? MODEL_WIDE_SLOWDOWN * rl : rl) + 2 * nb * nb * ra * rd;
                add(1, std::min(x, y), t_v);
                if (x < y) {
                    add(x + 1, y, t_cd);
                    add(y + 1, z, t_cd_bd);
                } else if (y < x) {
                    add(y + 1, x, t_iv);
                    add(x + 1, z, t_cd_bd);
                } else {
                    add(x + 1, z, t_cd_bd);
                }
                add(z + 1, NT, t_all_dense);
            }
    }
    add(2, NT, MODEL_STEP_2FLOW_S * NT);
    int best = 1;
    double acc = 0.0, tbest = 0.0;
    for (int b = 1; b <= NT; b++) {
        acc += D[b];
        if (b == 1 || acc < tbest) {
            tbest = acc;
            best = b;
        }
    }
    return best;
}

extern "C" int parsec_band_size_auto_tuning(hicma_ctx *ctx, hicma_tlr_desc *Ar, hicma_tlr_desc *Fake,
                                            int *G, int NB)
{
    (void)ctx;
    (void)Fake;
    if (!Ar || !G) return HICMA_ERR_ARG;
    return hicma_band_autotune(G, Ar->t.NT, NB);
}

extern "C" int parsec_band_regenerate(hicma_ctx *ctx, int uplo, hicma_tlr_desc *A, void *data, hicma_kernel *kernel,
                                      hicma_index *index, int band_size)
{
    if (!ctx || !A) return HICMA_ERR_ARG;
    if (uplo != HICMA_LOWER) return HICMA_ERR_UPLO;
    if (band_size < 1 || band_size > A->t.NT) return HICMA_ERR_ARG;
    HIPCHK(hipSetDevice(ctx->device));
    HIPCHK(hipStreamSynchronize(ctx->stream));
    RCHK(apply_problem(A, data, kernel, index));
    int rc = alloc_band(A, band_size);
    if (rc) return rc;
    if ((rc = gen_band_tiles(A))) return rc;
    HIPCHK(hipStreamSynchronize(ctx->stream));
    return 0;
}

// ---------------------------------------------------------------------------
// the factorization

// Workspace of one HiCMA_dpotrf_L call (doubles): syrk (Y, Z per tile row),
// low-rank x dense gemm (W, X per pair), recompression slots of the
// persistent grids (sized for the context's CUs), task lists.
struct PotrfWs {
    RcWs g;
    int ns, nsf, wgrid, wres, wcap;
    bool fast, wave;
    size_t rc_lds, fstride, winv_d, dag_d, linv_d, tasks_d, wws_d, wbs_d, wlist_d, lrd_pairs, syrk_d, lrd_d, qn_d,
        total;
};

static void plan_potrf_ws(hicma_ctx *ctx, int NT, int NB, int cap, int b, PotrfWs &P)
{
    RcWs &g = P.g;
    g.smax = 2 * cap;
    g.stride = rc_slot_doubles(NB, cap);
    size_t lds_fixed = sizeof(double) * (RC_FIXED_SMEM + 2 * (size_t)g.smax + 8);
    size_t lds_avail = ctx->lds_max > lds_fixed ? ctx->lds_max - lds_fixed : 0;
    int jl = (int)floor(sqrt((double)(lds_avail / sizeof(double)) / 2.0));
    if (jl > g.smax) jl = g.smax;
    g.jlds = jl;
    P.rc_lds = lds_fixed + sizeof(double) * 2 * (size_t)jl * jl;
    long long maxtask = (long long)(NT - 1) * (NT - 2) / 2;
    P.ns = nslots_for(ctx, g.stride * sizeof(double), maxtask > 0 ? maxtask : 1);
    P.syrk_d = (size_t)NT * NB * cap + (size_t)NT * cap * cap;
    P.lrd_pairs = (b > 1) ? (size_t)NT * (b - 1) : 0;
    P.lrd_d = P.lrd_pairs * ((size_t)cap * NB + (size_t)cap * cap);
    // fast recompression path (1024 threads, one row per thread)
    P.fast = (NB <= FT);
    P.fstride = rcf_slot_doubles(NB);
    P.nsf = P.fast ? nslots_for(ctx, P.fstride * sizeof(double), maxtask > 0 ? maxtask : 1) : 0;
    if (P.nsf > ctx->ncu) P.nsf = ctx->ncu;
    P.tasks_d = ((size_t)(maxtask > 0 ? maxtask : 1) + 1) / 2 + 1;  // ints packed in doubles
    P.winv_d = (size_t)((NB + PBK - 1) / PBK) * PBK * PBK;
    // the diagonal potrf's task counter and done flags (ints packed in
    // doubles) and the blocks of L_kk^-1
    const size_t nbk = (size_t)((NB + PBK - 1) / PBK);
    P.dag_d = (nbk * nbk + 2 + 1) / 2 + 1;
    P.linv_d = nbk * nbk * PBK * PBK;
    // one-wave recompression (tasks with s <= WV_SMAX): the s <= 32 and
    // s <= 48 variants run 4 waves per workgroup, s <= 64 two
    P.wave = (NB <= 64 * WV_KMAX);
    P.wgrid = P.wave ? ctx->ncu : 0;  // times WvCfg<SM>::BPC resident workgroups
    // CUs left to the lookahead stream (syrk(k) + potrf(k+1)) while the
    // recompression kernels of step k run
    P.wres = (P.wave && ctx->ncu > 64 && ctx->ncu == ctx->ncu_dev) ? 16 : 0;
    P.wcap = (int)(maxtask > 0 ? maxtask : 1);
    // four concurrent variants (32, 48, 64 on side streams), each with up to
    // four resident waves per CU: one quarter of the region each
    P.wws_d = (size_t)4 * P.wgrid * 4 * wv_slot_doubles(NB);
    P.wbs_d = P.wave ? (size_t)P.wgrid * wb_slot_doubles(NB) : 0;  // one wide-stack wave per CU
    P.wlist_d = ((size_t)4 * P.wcap + 8) / 2 + 1;
    // Qn = -B Av of the dense-B tasks the wave kernels take (2-flow only)
    P.qn_d = (P.wave && b > 1) ? (size_t)(b - 1) * (NT - 1) * NB * QN_COLS : 0;
    P.total = g.stride * P.ns + P.syrk_d + P.lrd_d + P.fstride * P.nsf + P.tasks_d + P.winv_d + P.wws_d +
              P.wlist_d + P.wbs_d + P.dag_d + P.linv_d + P.qn_d;
}

// Device bytes a generation + factorization of this shape allocates on the
// context, beyond the descriptor itself (the largest of the two workspaces,
// which share one buffer, plus the generation's sketch).
size_t hicma_work_bytes(hicma_ctx *ctx, int N, int NB, int cap, int band)
{
    const int NT = (N + NB - 1) / NB;
    PotrfWs P;
    plan_potrf_ws(ctx, NT, NB, cap, band, P);
    int p = cap + 10;
    if (p > NB) p = NB;
    const long long ntask = (long long)NT * (NT - 1) / 2;
    const size_t gs = gen_slot_doubles(NB, p, cap);
    const size_t gen = ntask > 0 ? gs * nslots_for(ctx, gs * sizeof(double), ntask) : 0;
    return sizeof(double) * (std::max(P.total, gen) + (size_t)NB * p);
}

size_t hicma_desc_bytes(int N, int NB, int cap, int band, int ndim)
{
    const size_t NT = (size_t)(N + NB - 1) / NB;
    return sizeof(double) * (size_t)ndim * N + sizeof(int) * NT * NT + sizeof(double) * NT * (NT - 1) / 2 * NB * 2 * cap +
           sizeof(unsigned long long) * (NT * NT + 16) + sizeof(double) * (size_t)band * NT * NB * NB + NT * NT;
}

// blocked Cholesky of diagonal tile k: per 64-column block, factor + invert
// the diagonal block in LDS, panel = C W^T, lower trailing update.  The block
// inverses (Winv, NB/64 blocks of 64x64) are reused by the low-rank trsm.
static int launch_potrf(hicma_tlr_desc *A, int k, double *Winv, double *Linv, int *dflags, OpLog o, hipStream_t s)
{
    TLR &t = A->t;
    const int nb = (t.NB + PBK - 1) / PBK;
    hipLaunchKernelGGL(k_fill_i32, dim3((nb * nb + 2 + 255) / 256), dim3(256), 0, s, dflags, nb * nb + 2, 0);
    // one worker per block column (a whole column of tasks can run at once),
    // on the XCD that the kernel elects among DAG_XCD_SPREAD times as many
    // workgroups
    hipLaunchKernelGGL(k_potrf_dag, dim3(nb * DAG_XCD_SPREAD), dim3(BS), sizeof(double) * DAG_SMEM, s, t, k, Winv,
                       Linv, dflags, A->ctx->d_info, A->ctx->d_info + 3, o);
    HIPCHK(hipGetLastError());
    return 0;
}

// columns of V per workgroup of the low-rank trsm: its LDS holds TC columns
// of 64 nb + 1 doubles
static int trsm_tc(int NB)
{
    const size_t rld = (size_t)PBK * ((NB + PBK - 1) / PBK) + 1;
    for (int tc = 16; tc >= 2; tc /= 2)
        if (rld * tc * sizeof(double) <= TRSM_LDS_MAX) return tc;
    return 0;
}

static void launch_trsm_lr(TLR &t, int k, int nlr, int cap, const double *Winv, const double *Linv, OpLog o,
                           hipStream_t s)
{
    const int tc = trsm_tc(t.NB);
    const size_t lds = sizeof(double) * ((size_t)PBK * ((t.NB + PBK - 1) / PBK) + 1) * tc;
    const dim3 grid((cap + tc - 1) / tc, nlr);
    switch (tc) {
    case 16: hipLaunchKernelGGL(k_trsm_lr_inv<16>, grid, dim3(BS), lds, s, t, k, Winv, Linv, o); break;
    case 8: hipLaunchKernelGGL(k_trsm_lr_inv<8>, grid, dim3(BS), lds, s, t, k, Winv, Linv, o); break;
    case 4: hipLaunchKernelGGL(k_trsm_lr_inv<4>, grid, dim3(BS), lds, s, t, k, Winv, Linv, o); break;
    default: hipLaunchKernelGGL(k_trsm_lr_inv<2>, grid, dim3(BS), lds, s, t, k, Winv, Linv, o); break;
    }
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
    // the factorization's kernels read and write FP64 tiles only
    if (A->demoted) return HICMA_ERR_ARG;
    if (lookahead && *lookahead == -1) *lookahead = band_size;
    // 3-flow when band == 1 and not forced 2-flow, else 2-flow (hicma_parsec.c:517, 560-573)
    bool three_flow = (band_size == 1 && !(two_flow && *two_flow));
    if (!three_flow && two_flow) *two_flow = 1;
    if (verbose) fprintf(stderr, "%s version start\n", three_flow ? "3flow" : "2flow");

    HIPCHK(hipSetDevice(ctx->device));
    hipStream_t st = ctx->stream;
    TLR &t = A->t;
    const int NT = t.NT, NB = t.NB, cap = t.cap, b = band_size;
    if (A->rhist_on && !A->rhist) HIPCHK(ctx_malloc(ctx, (void **)&A->rhist, sizeof(int) * 4 * (size_t)NT * NT));
    // one op log per executing stream: [8] context stream, [9] side 0,
    // [10] side 1, [11] the lookahead stream
    OpLog o{A->d_tileops, A->d_opstat, A->d_opstat + 8, A->rhist_on ? A->rhist : nullptr, ctx->d_info + 2};
    OpLog o0 = o, o1 = o, od = o;
    o0.exec += 1;
    o1.exec += 2;
    od.exec += 3;

    PotrfWs P;
    plan_potrf_ws(ctx, NT, NB, cap, b, P);
    RcWs &g = P.g;
    const int ns = P.ns, nsf = P.nsf, wgrid = P.wgrid, wres = P.wres, wcap = P.wcap;
    const bool fast = P.fast, wave = P.wave;
    const size_t rc_lds = P.rc_lds, fstride = P.fstride, winv_d = P.winv_d, tasks_d = P.tasks_d, wws_d = P.wws_d,
                 wlist_d = P.wlist_d, lrd_pairs = P.lrd_pairs, total = P.total;
    int rc = hicma_ensure_ws(ctx, total * sizeof(double));
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
    double *Linv = wbs + P.wbs_d;
    double *qnws = P.qn_d ? Linv + P.linv_d : nullptr;
    int *dflags = (int *)(Linv + P.linv_d + P.qn_d);
    HIPCHK(hipFuncSetAttribute((const void *)k_potrf_dag, hipFuncAttributeMaxDynamicSharedMemorySize,
                               (int)(sizeof(double) * DAG_SMEM)));
    if (!trsm_tc(NB)) return HICMA_ERR_ARG;
    HIPCHK(hipFuncSetAttribute((const void *)k_trsm_lr_inv<16>, hipFuncAttributeMaxDynamicSharedMemorySize,
                               (int)TRSM_LDS_MAX));
    HIPCHK(hipFuncSetAttribute((const void *)k_trsm_lr_inv<8>, hipFuncAttributeMaxDynamicSharedMemorySize,
                               (int)TRSM_LDS_MAX));
    HIPCHK(hipFuncSetAttribute((const void *)k_trsm_lr_inv<4>, hipFuncAttributeMaxDynamicSharedMemorySize,
                               (int)TRSM_LDS_MAX));
    HIPCHK(hipFuncSetAttribute((const void *)k_trsm_lr_inv<2>, hipFuncAttributeMaxDynamicSharedMemorySize,
                               (int)TRSM_LDS_MAX));
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

    // critical-path timing (hicma_parsec/HiCMA_dpotrf_L_2flow_wrapper.c:27-166):
    // potrf(k), trsm(k+1,k) and syrk(k,k+1) durations, each measured by a
    // pair of events on the stream that runs it (trsm of step k is one batched
    // launch on the context stream; potrf(k+1) and syrk(k) run on the
    // lookahead stream).  Events are created once per context and only when
    // the caller asks for the times.
    const bool timing = (critical_path_time != nullptr);
    if (timing) {
        const size_t need = 6 * (size_t)NT;
        while (ctx->cp_ev.size() < need) {
            hipEvent_t e;
            HIPCHK(hipEventCreate(&e));
            ctx->cp_ev.push_back(e);
        }
    }
    auto cp = [&](int k, int slot, hipStream_t s) -> hipError_t {
        return timing ? hipEventRecord(ctx->cp_ev[6 * (size_t)k + slot], s) : hipSuccess;
    };
    // HIP graph: the whole launch sequence below depends only on (NT, band)
    // and on the pointers and scalars in the key, so it is captured once and
    // replayed while they are unchanged (timing, profiling and rank-history
    // runs are always launched directly)
    const bool use_graph = A->use_graph && !timing && !ctx->profile && !A->rhist_on;
    GraphKey key;
    memset(&key, 0, sizeof(key));
    key.band = t.band;
    key.lr = t.lr;
    key.rank = t.rank;
    key.xyz = t.xyz;
    key.ws = ctx->ws;
    key.tileops = A->d_tileops;
    key.opstat = A->d_opstat;
    key.acc = acc;
    key.sigma2 = t.sigma2;
    key.beta = t.beta;
    key.nugget = t.nugget;
    key.fixed_rank = fixed_rank;
    key.maxrank = maxrank;
    key.band_size = b;
    key.N = t.N;
    key.NB = NB;
    key.NT = NT;
    key.cap = cap;
    key.kind = t.kind;
    key.wave_min_tasks = ctx->wave_min_tasks;
    const bool replay = use_graph && A->gexec && memcmp(&key, &A->gkey, sizeof(key)) == 0;
    if (use_graph && !replay) {
        if (A->gexec) (void)hipGraphExecDestroy(A->gexec);
        if (A->graph) (void)hipGraphDestroy(A->graph);
        A->gexec = nullptr;
        A->graph = nullptr;
        HIPCHK(hipStreamBeginCapture(st, hipStreamCaptureModeRelaxed));
    }
    auto enqueue = [&]() -> int {
    int rc = 0;
    hipLaunchKernelGGL(k_fill_i32, dim3(1), dim3(64), 0, st, ctx->d_info, 4, 0);
    hipLaunchKernelGGL(k_fill_u64, dim3((unsigned)(((long)NT * NT + 255) / 256)), dim3(256), 0, st, A->d_tileops,
                       (long)NT * NT);
    hipLaunchKernelGGL(k_fill_u64, dim3(1), dim3(64), 0, st, A->d_opstat, 16L);
    if (A->rhist_on) hipLaunchKernelGGL(k_rank_hist, dim3((NT * NT + 255) / 256), dim3(256), 0, st, t, A->rhist, 0);
    const int ntg = (NB + GT - 1) / GT;
    const int ntl = ntg * (ntg + 1) / 2;
    // Lookahead: syrk(k) and potrf(k+1) touch only diagonal tiles and the gemm
    // updates of step k only off-diagonal ones, so the former run on a side
    // stream while the latter run, and are joined back before trsm(k+1).
    hipStream_t sd = ctx->side[2];
    if (NT > 0) {
        HIPCHK(cp(0, 0, st));
        if ((rc = launch_potrf(A, 0, Winv, Linv, dflags, o, st))) return rc;
        HIPCHK(cp(0, 1, st));
    }
    for (int k = 0; k < NT; k++) {
        int nbnd = std::min(b - 1, NT - 1 - k);
        int nlr = NT - k - b;
        HIPCHK(cp(k, 2, st));
        if (nbnd > 0)
            hipLaunchKernelGGL(k_trsm_band, dim3(ntg, nbnd), dim3(BS), 0, st, t, k, (const double *)Winv,
                               (const double *)Linv, o);
        if (nlr > 0)
            launch_trsm_lr(t, k, nlr, cap, Winv, Linv, o, st);
        HIPCHK(cp(k, 3, st));
        HIPCHK(hipEventRecord(ctx->ev_join[2], st));
        HIPCHK(hipStreamWaitEvent(sd, ctx->ev_join[2], 0));
        HIPCHK(cp(k, 4, sd));
        if (nbnd > 0) hipLaunchKernelGGL(k_syrk_band, dim3(ntl, nbnd), dim3(BS), 0, sd, t, k, od);
        if (nlr > 0) {
            hipLaunchKernelGGL(k_syrk_lr_prep, dim3(nlr), dim3(BS), 0, sd, t, k, Yws, Zws, od);
            hipLaunchKernelGGL(k_syrk_lr_update, dim3(ntl, nlr), dim3(BS), 0, sd, t, k, (const double *)Yws);
        }
        HIPCHK(cp(k, 5, sd));
        if (k + 1 < NT) {
            HIPCHK(cp(k + 1, 0, sd));
            if ((rc = launch_potrf(A, k + 1, Winv, Linv, dflags, od, sd))) return rc;
            HIPCHK(cp(k + 1, 1, sd));
        }
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
                hipLaunchKernelGGL(k_fill_i32, dim3(1), dim3(64), 0, st, wcount, 8, 0);
                const int wide_only = (ntask < ctx->wave_min_tasks) ? 1 : 0;
                if (qnws) {
                    const int nrt = (NB + GT - 1) / GT;
                    hipLaunchKernelGGL(k_qn_dense, dim3(nrt * (QN_COLS / GT), (unsigned)((b - 1) * T)), dim3(BS), 0,
                                       st, t, k, qnws);
                }
                hipLaunchKernelGGL(k_classify_wave, dim3((unsigned)((ntask + 255) / 256)), dim3(256), 0, st, t, k,
                                   tasks_s, wlist, wcount, wcap, wide_only, qnws ? 1 : 0);
                // fork: the wide variant and variant 64 on side streams 1 and 0,
                // variants 48 and 32 on the context stream.  Launched longest
                // tasks first: the dispatcher hands out CUs in launch order, so
                // the long tasks start at once and the short ones fill the
                // step's tail.
                HIPCHK(hipEventRecord(ctx->ev_fork, st));
                for (int i = 0; i < 2; i++) HIPCHK(hipStreamWaitEvent(ctx->side[i], ctx->ev_fork, 0));
                int wg = (int)std::min<long long>(wgrid - wres, ntask);
                hipLaunchKernelGGL(k_rc_wide, dim3(wg), dim3(64 * WB_WAVES), sizeof(double) * WB_LDS, ctx->side[1], t, k,
                                   (const int *)(wlist + 3 * (size_t)wcap), (const int *)(wcount + 3), wcount + 7,
                                   tasks_s, wbs, acc, fixed_rank, maxrank, o1, ctx->d_ovf, (const double *)qnws);
                wg = (int)std::min<long long>((wgrid - wres) * WvCfg<64>::BPC, (ntask + WvCfg<64>::WPB - 1) / WvCfg<64>::WPB);
                hipLaunchKernelGGL(k_rc_wave<64>, dim3(wg), dim3(64 * WvCfg<64>::WPB), wlds64, ctx->side[0], t, k,
                                   (const int *)(wlist + 2 * (size_t)wcap), (const int *)(wcount + 2), wcount + 6,
                                   tasks_s, wws + wws_d / 2, acc, fixed_rank, maxrank, o0, ctx->d_ovf,
                                   (const double *)qnws);
                wg = (int)std::min<long long>((wgrid - wres) * WvCfg<48>::BPC, (ntask + WvCfg<48>::WPB - 1) / WvCfg<48>::WPB);
                hipLaunchKernelGGL(k_rc_wave<48>, dim3(wg), dim3(64 * WvCfg<48>::WPB), wlds48, st, t, k,
                                   (const int *)(wlist + wcap), (const int *)(wcount + 1), wcount + 5, tasks_s,
                                   wws + wws_d / 4, acc, fixed_rank, maxrank, o, ctx->d_ovf, (const double *)qnws);
                wg = (int)std::min<long long>((wgrid - wres) * WvCfg<32>::BPC, (ntask + WvCfg<32>::WPB - 1) / WvCfg<32>::WPB);
                hipLaunchKernelGGL(k_rc_wave<32>, dim3(wg), dim3(64 * WvCfg<32>::WPB), wlds32, st, t, k,
                                   (const int *)wlist, (const int *)wcount, wcount + 4, tasks_s, wws, acc, fixed_rank,
                                   maxrank, o, ctx->d_ovf, (const double *)qnws);
                // join
                for (int i = 0; i < 2; i++) {
                    HIPCHK(hipEventRecord(ctx->ev_join[i], ctx->side[i]));
                    HIPCHK(hipStreamWaitEvent(st, ctx->ev_join[i], 0));
                }
            }
            if (fast) {
                int gf = (int)std::min<long long>(nsf, ntask);
                hipLaunchKernelGGL(k_lr_gemm_fast, dim3(gf), dim3(FT), RCF_LDS * sizeof(double), st, t, k,
                                   tasks_s, fws, fstride, acc, fixed_rank, maxrank, o, ctx->d_ovf);
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
    if (A->rhist_on)
        hipLaunchKernelGGL(k_rank_hist, dim3((NT * NT + 255) / 256), dim3(256), 0, st, t, A->rhist, 1);
    HIPCHK(hipGetLastError());
    return 0;
    };
    if (!replay) {
        rc = enqueue();
        if (use_graph) {
            hipGraph_t gr = nullptr;
            const hipError_t ec = hipStreamEndCapture(st, &gr);
            if (rc || ec != hipSuccess) {
                if (gr) (void)hipGraphDestroy(gr);
                if (rc) return rc;
                HIPCHK(ec);
            }
            A->graph = gr;
            HIPCHK(hipGraphInstantiate(&A->gexec, A->graph, nullptr, nullptr, 0));
            A->gkey = key;
        } else if (rc) {
            return rc;
        }
    }
    if (use_graph) HIPCHK(hipGraphLaunch(A->gexec, st));
    HIPCHK(hipStreamSynchronize(st));
    int hinfo[4];
    HIPCHK(hipMemcpy(hinfo, ctx->d_info, 4 * sizeof(int), hipMemcpyDeviceToHost));
    unsigned long long stats[16];
    HIPCHK(hipMemcpy(stats, A->d_opstat, sizeof(stats), hipMemcpyDeviceToHost));
    for (int i = 0; i < 12; i++) A->last_stats[i] = (unsigned long)stats[i];
    A->last_ovf = hinfo[1];
    A->last_unconv = hinfo[2];
    if (hinfo[3]) {
        fprintf(stderr, "hicma_tlr: a wait of the diagonal-tile schedule ran out; the factor is not valid\n");
        return HICMA_ERR_HIP;
    }
    if (tileopcounters) {
        std::vector<unsigned long long> to((size_t)NT * NT);
        HIPCHK(hipMemcpy(to.data(), A->d_tileops, sizeof(unsigned long long) * to.size(), hipMemcpyDeviceToHost));
        for (size_t i = 0; i < to.size(); i++) tileopcounters[i] = (unsigned long)to[i];
    }
    // one executor (the device): the reference's per-thread counters sum to
    // the total; per-stream totals are in hicma_exec_op_stats
    if (opcounters) opcounters[0] = (unsigned long)(stats[0] + stats[1]);
    if (timing) {
        double ts[3] = {0, 0, 0};
        for (int k = 0; k < NT; k++) {
            for (int ph = 0; ph < 3; ph++) {
                float ms = 0.f;
                HIPCHK(hipEventElapsedTime(&ms, ctx->cp_ev[6 * (size_t)k + 2 * ph], ctx->cp_ev[6 * (size_t)k + 2 * ph + 1]));
                ts[ph] += ms * 1e-3;
            }
        }
        for (int i = 0; i < 6; i++) critical_path_time[i] = i < 3 ? ts[i] : 0.0;
    }
    if (verbose && hinfo[1])
        fprintf(stderr, "hicma_tlr: %d recompressions capped at maxrank %d\n", hinfo[1], maxrank);
    if (verbose && hinfo[2])
        fprintf(stderr, "hicma_tlr: %d recompression SVDs stopped at the sweep limit\n", hinfo[2]);
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
    HIPCHK(hipMemcpyAsync(b.prec, a.prec, (size_t)a.NT * a.NT, hipMemcpyDeviceToDevice, c->stream));
    dst->demoted = src->demoted;
    HIPCHK(hipStreamSynchronize(c->stream));
    return 0;
}

extern "C" int hicma_rank_overflows(const hicma_tlr_desc *A) { return A ? A->last_ovf : 0; }

extern "C" int hicma_jacobi_unconverged(const hicma_tlr_desc *A) { return A ? A->last_unconv : 0; }

extern "C" int hicma_set_wave_min_tasks(hicma_ctx *ctx, int ntask)
{
    if (!ctx || ntask < 0) return HICMA_ERR_ARG;
    ctx->wave_min_tasks = ntask;
    return 0;
}

// Times HiCMA_dpotrf_L on a copy of A for each candidate schedule (steps
// with fewer tasks than the threshold on the four-wave kernel) and keeps the
// fastest in the context.  A is not modified.
extern "C" int hicma_autotune_schedule(hicma_ctx *ctx, hicma_tlr_desc *A, double acc, int maxrank, int reps,
                                       int *best, double *best_seconds)
{
    if (!ctx || !A || A->ctx != ctx || reps < 1) return HICMA_ERR_ARG;
    const TLR &t = A->t;
    const long long maxtask = (long long)(t.NT - 1) * (t.NT - 2) / 2;
    std::vector<int> cand{0};
    for (int c = 50; c < 4 * maxtask && c <= 6400; c *= 2) cand.push_back(c);
    hicma_tlr_desc *W = nullptr;
    RCHK(hicma_desc_create(ctx, &W, t.N, t.NB, t.cap, t.kind, t.sigma2, t.beta, t.nugget, A->seed));
    const int saved = ctx->wave_min_tasks;
    int bi = saved, rc = 0;
    double bs = 1e300;
    for (int c : cand) {
        ctx->wave_min_tasks = c;
        double tmin = 1e300;
        for (int r = 0; r < reps && !rc; r++) {
            if ((rc = hicma_desc_copy(W, A))) break;
            int la = -1, tf = 0;
            const auto t0 = std::chrono::steady_clock::now();
            const int info = HiCMA_dpotrf_L(ctx, HICMA_LOWER, W, W, W, acc, 0, maxrank, &la, t.band_size, 300, maxrank,
                                            0, &tf, NULL, NULL, NULL, 0);
            const double sec = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
            if (info) rc = info;
            tmin = std::min(tmin, sec);
        }
        if (rc) break;
        if (tmin < bs) {
            bs = tmin;
            bi = c;
        }
    }
    hicma_desc_destroy(W);
    ctx->wave_min_tasks = rc ? saved : bi;
    if (rc) return rc;
    if (best) *best = bi;
    if (best_seconds) *best_seconds = bs;
    return 0;
}

extern "C" int hicma_factor_graph(hicma_tlr_desc *A, int enable)
{
    if (!A) return HICMA_ERR_ARG;
    A->use_graph = enable ? 1 : 0;
    return 0;
}

extern "C" int hicma_exec_op_stats(const hicma_tlr_desc *A, unsigned long *ops, int n)
{
    if (!A || !ops || n < 0) return HICMA_ERR_ARG;
    for (int i = 0; i < n; i++) ops[i] = i < 4 ? A->last_stats[8 + i] : 0;
    return 0;
}

extern "C" int hicma_rank_history(hicma_tlr_desc *A, int enable)
{
    if (!A) return HICMA_ERR_ARG;
    A->rhist_on = enable ? 1 : 0;
    return 0;
}

extern "C" int hicma_get_rank_history(const hicma_tlr_desc *A, int *h)
{
    if (!A || !h || !A->rhist) return HICMA_ERR_ARG;
    HIPCHK(hipSetDevice(A->ctx->device));
    HIPCHK(hipStreamSynchronize(A->ctx->stream));
    HIPCHK(hipMemcpy(h, A->rhist, sizeof(int) * 4 * (size_t)A->t.NT * A->t.NT, hipMemcpyDeviceToHost));
    return 0;
}

// parsec_rank_print (hicma_parsec/rank_print.jdf:100-190): the initial,
// minimum, maximum and final rank of every off-band tile over the last
// factorization, and max - min; zero for band tiles, as the reference prints.
extern "C" int parsec_rank_print(hicma_ctx *ctx, hicma_tlr_desc *Rank, int band_size)
{
    if (!ctx || !Rank) return HICMA_ERR_ARG;
    if (!Rank->rhist) return HICMA_ERR_ARG;  // hicma_rank_history(Rank, 1) before HiCMA_dpotrf_L
    HIPCHK(hipSetDevice(ctx->device));
    HIPCHK(hipStreamSynchronize(ctx->stream));
    const int NT = Rank->t.NT;
    std::vector<int> h(4 * (size_t)NT * NT);
    HIPCHK(hipMemcpy(h.data(), Rank->rhist, sizeof(int) * h.size(), hipMemcpyDeviceToHost));
    const char *title[5] = {"Initial Rank", "Minimum Rank", "Maximum Rank", "Final Rank", "Difference Rank, max - min"};
    for (int q = 0; q < 5; q++) {
        printf("\n%s\n", title[q]);
        for (int i = 0; i < NT; i++) {
            for (int j = 0; j < NT; j++) {
                int v = 0;
                if (i - j >= band_size) {
                    const int *p = &h[4 * ((size_t)j * NT + i)];
                    v = q < 4 ? p[q] : p[2] - p[1];
                }
                printf("%3d ", v);
            }
            printf("\n");
        }
        printf("\n\n");
    }
    fflush(stdout);
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
        unsigned char p = LR_F64;
        HIPCHK(hipMemcpy(&p, t.prec + (size_t)n * t.NT + m, 1, hipMemcpyDeviceToHost));
        const size_t ne = (size_t)t.NB * 2 * r, es = p == LR_F64 ? 8 : (p == LR_F32 ? 4 : 2);
        HIPCHK(hipMemcpy(buf, src, es * ne, hipMemcpyDeviceToHost));
        // widen in place from the last element down: double i covers packed
        // elements 2i.. (FP32) or 4i.. (BF16), none below i, so every packed
        // element it replaces has been widened already
        char *raw = (char *)buf;
        for (size_t i = ne; es != 8 && i-- > 0;) {
            float f;
            if (es == 4) {
                memcpy(&f, raw + 4 * i, 4);
            } else {
                unsigned short h;
                memcpy(&h, raw + 2 * i, 2);
                const unsigned w = (unsigned)h << 16;
                memcpy(&f, &w, 4);
            }
            buf[i] = (double)f;
        }
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
    // the tile now holds FP64 values
    unsigned char *pp = t.prec + (size_t)n * t.NT + m, p = LR_F64;
    HIPCHK(hipMemcpy(&p, pp, 1, hipMemcpyDeviceToHost));
    if (p != LR_F64) {
        const unsigned char f64 = LR_F64;
        HIPCHK(hipMemcpy(pp, &f64, 1, hipMemcpyHostToDevice));
        A->demoted--;
    }
    return 0;
}

// 2 sum log L(i, i) into *out (device), asynchronous on s, with NT doubles
// of device workspace in parts; the log-likelihood uses the same kernels,
// so both give the same bits
void logdet_launch(const TLR &t, double *parts, double *out, hipStream_t s)
{
    hipLaunchKernelGGL(k_logdet_parts, dim3(t.NT), dim3(LD_T), 0, s, t, parts);
    hipLaunchKernelGGL(k_logdet_sum, dim3(1), dim3(64), 0, s, (const double *)parts, t.NT, out);
}

extern "C" int hicma_logdet(hicma_tlr_desc *A, double *logdet)
{
    if (!A || !logdet) return HICMA_ERR_ARG;
    hicma_ctx *c = A->ctx;
    HIPCHK(hipSetDevice(c->device));
    RCHK(hicma_ensure_ws(c, sizeof(double) * (size_t)A->t.NT));
    logdet_launch(A->t, c->ws, c->d_scalar, c->stream);
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(c->stream));
    HIPCHK(hipMemcpy(logdet, c->d_scalar, sizeof(double), hipMemcpyDeviceToHost));
    return 0;
}
