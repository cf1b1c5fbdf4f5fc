// hicma_check.hip -- the --check of testing_dpotrf_tlr on the device.
//
// The reference uncompresses the TLR matrix into a dense descriptor
// (STARSH_check, hicma_parsec/STARSH_check.jdf:90-143), factors the
// uncompressed matrix with DPLASMA's dense Cholesky
// (hicma_parsec/testing_dpotrf.c:748-771) and compares
// (check_dpotrf2 / check_diff, hicma_parsec/testing_dpotrf.c:111-230).
// Here every dense matrix is an N x N column-major array in HBM (64.8 GB at
// N = 90,000); the uncompress and the residual ||A - L L^T|| run on the FP64
// matrix cores, the norms are reduced in a fixed order, and the dense
// Cholesky (the checker's reference factorization, as DPLASMA's is in the
// reference) is rocSOLVER's 64-bit-index dpotrf.
#include "internal.hpp"
#include "mfma.hpp"
#include "wave.hpp"

#include <math.h>
#include <rocblas/rocblas.h>
#include <rocsolver/rocsolver.h>
#include <stdlib.h>
#include <string.h>

// ---------------------------------------------------------------------------
// kernels

#define CK_ROWS 256  // rows per workgroup of the element-wise kernels
#define CK_COLS 64   // columns per workgroup of the element-wise kernels

// exact covariance, lower triangle; the strict upper triangle is zeroed
__global__ void __launch_bounds__(CK_ROWS) k_dense_gen(TLR t, double *D, long ld)
{
    const int N = t.N;
    const int i = blockIdx.x * CK_ROWS + threadIdx.x;
    if (i >= N) return;
    const int j0 = blockIdx.y * CK_COLS;
    for (int j = j0; j < j0 + CK_COLS && j < N; j++) D[i + (long)j * ld] = (i >= j) ? kval(t, i, j) : 0.0;
}

// zero the strict upper triangle
__global__ void __launch_bounds__(CK_ROWS) k_dense_zero_upper(double *D, long ld, int N)
{
    const int i = blockIdx.x * CK_ROWS + threadIdx.x;
    if (i >= N) return;
    const int j0 = blockIdx.y * CK_COLS;
    for (int j = j0; j < j0 + CK_COLS && j < N; j++)
        if (i < j) D[i + (long)j * ld] = 0.0;
}

// lower tiles (m, n) of the TLR matrix into D: band tiles copied (the strict
// upper part of diagonal tiles zeroed), low-rank tiles U V^T on the matrix
// cores.  grid.x = 64 x 64 sub-tiles of a tile, grid.y strides over the
// NT (NT + 1) / 2 lower tiles.
__global__ void __launch_bounds__(256) k_dense_uncompress(TLR t, double *D, long ld)
{
    __shared__ double sm[MK_SMEM];
    const int NB = t.NB, NT = t.NT;
    const int nsub = (NB + MK_T - 1) / MK_T;
    const int si = blockIdx.x % nsub, sj = blockIdx.x / nsub;
    const long ntile = (long)NT * (NT + 1) / 2;
    for (long tix = blockIdx.y; tix < ntile; tix += gridDim.y) {
        int hi, lo;
        tri_decode(tix, hi, lo);  // tix = hi (hi - 1) / 2 + lo, lo < hi
        const int m = hi - 1, n = lo;
        if (n > m) continue;
        const int rows = tdim(t, m), cols = tdim(t, n);
        const int i0 = si * MK_T, j0 = sj * MK_T;
        if (i0 >= rows || j0 >= cols) continue;
        double *dst = D + (long)m * NB + (long)n * NB * ld;
        if (m - n < t.band_size) {
            const double *T = band_tile(t, m, n);
            for (int e = threadIdx.x; e < MK_T * MK_T; e += 256) {
                const int i = i0 + (e & (MK_T - 1)), j = j0 + (e >> 6);
                if (i < rows && j < cols) dst[i + (long)j * ld] = (m == n && i < j) ? 0.0 : T[i + (size_t)j * NB];
            }
            continue;
        }
        const int r = *rank_at(t, m, n), p = lr_prec(t, m, n);
        const double *U = lr_tile(t, m, n);
        mk_d4 acc[2][2];
        mk_zero(acc);
        // U V^T from [U | V] in the type the tile is held in
        if (p == LR_F64) {
            const MkOps<false, false> op{U, NB, U + (size_t)NB * r, NB};
            mk_accum(op, rows, cols, i0, j0, 0, r, acc, sm);
        } else if (p == LR_F32) {
            const float *Uf = (const float *)U;
            const MkOps<false, false, float> op{Uf, NB, Uf + (size_t)NB * r, NB};
            mk_accum(op, rows, cols, i0, j0, 0, r, acc, sm);
        } else {
            const lr_bf16 *Uh = (const lr_bf16 *)U;
            const MkOps<false, false, lr_bf16> op{Uh, NB, Uh + (size_t)NB * r, NB};
            mk_accum(op, rows, cols, i0, j0, 0, r, acc, sm);
        }
        mk_store(acc, [&](int i, int j, double v) {
            if (i0 + i < rows && j0 + j < cols) dst[(i0 + i) + (long)(j0 + j) * ld] = v;
        });
    }
}

// fixed-order sum of a block's values: every thread passes its part
__device__ __forceinline__ double ck_block_sum(double v)
{
    __shared__ double red[8];
    v = wsum(v);
    const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
    __syncthreads();
    if (l == 0) red[w] = v;
    __syncthreads();
    double s = 0.0;
    for (int i = 0; i < (int)(blockDim.x >> 6); i++) s += red[i];
    return s;
}

// sum of squares of the symmetric matrix given by the lower triangle of
// X (- Y when Y is given): dlansy('F') squared, strict-lower entries twice
__global__ void __launch_bounds__(CK_ROWS) k_dense_sumsq(const double *X, const double *Y, long ld, int N,
                                                         double *part)
{
    const int i = blockIdx.x * CK_ROWS + threadIdx.x;
    const int j0 = blockIdx.y * CK_COLS;
    double s = 0.0;
    if (i < N)
        for (int j = j0; j < j0 + CK_COLS && j < N && j <= i; j++) {
            double v = X[i + (long)j * ld];
            if (Y) v -= Y[i + (long)j * ld];
            s += (i == j) ? v * v : 2.0 * v * v;
        }
    s = ck_block_sum(s);
    if (threadIdx.x == 0) part[(long)blockIdx.y * gridDim.x + blockIdx.x] = s;
}

// residual L L^T - A0 on the lower triangle, 64 x 64 blocks (bi >= bj) of the
// N x N matrix: sum of squares per block (strict-lower entries twice)
__global__ void __launch_bounds__(256) k_dense_resid(const double *L, const double *A0, long ld, int N, double *part)
{
    __shared__ double sm[MK_SMEM];
    int hi, lo;
    tri_decode((long long)blockIdx.x, hi, lo);
    const int bi = hi - 1, bj = lo;
    const int i0 = bi * MK_T, j0 = bj * MK_T;
    const long kend = (long)(j0 + MK_T) < (long)N ? (long)(j0 + MK_T) : (long)N;
    MkOps<false, false> op{L, ld, L, ld};
    mk_d4 acc[2][2];
    mk_zero(acc);
    mk_accum<false, false, true>(op, N, N, i0, j0, 0, kend, acc, sm);
    double s = 0.0;
    mk_store(acc, [&](int i, int j, double v) {
        const int gi = i0 + i, gj = j0 + j;
        if (gi < N && gj < N && gi >= gj) {
            const double r = v - A0[gi + (long)gj * ld];
            s += (gi == gj) ? r * r : 2.0 * r * r;
        }
    });
    s = ck_block_sum(s);
    if (threadIdx.x == 0) part[blockIdx.x] = s;
}

// 2 sum log L(i, i), per-block parts
__global__ void __launch_bounds__(256) k_dense_logdiag(const double *L, long ld, int N, double *part)
{
    const int i = blockIdx.x * 256 + threadIdx.x;
    double s = (i < N) ? log(L[i + (long)i * ld]) : 0.0;
    s = ck_block_sum(s);
    if (threadIdx.x == 0) part[blockIdx.x] = s;
}

// fixed-order total of n parts (one workgroup)
__global__ void __launch_bounds__(256) k_sum_parts(const double *part, long n, double *out)
{
    double s = 0.0;
    for (long q = threadIdx.x; q < n; q += 256) s += part[q];
    s = ck_block_sum(s);
    if (threadIdx.x == 0) *out = s;
}

// ---------------------------------------------------------------------------
// host side

static int sum_parts(hicma_ctx *c, const double *part, long n, double *out)
{
    hipLaunchKernelGGL(k_sum_parts, dim3(1), dim3(256), 0, c->stream, part, n, c->d_scalar);
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(c->stream));
    HIPCHK(hipMemcpy(out, c->d_scalar, sizeof(double), hipMemcpyDeviceToHost));
    return 0;
}

static dim3 elem_grid(long N) { return dim3((unsigned)((N + CK_ROWS - 1) / CK_ROWS), (unsigned)((N + CK_COLS - 1) / CK_COLS)); }

extern "C" int hicma_dense_create(hicma_ctx *ctx, hicma_dense **out, int N)
{
    if (!ctx || !out || N <= 0) return HICMA_ERR_ARG;
    HIPCHK(hipSetDevice(ctx->device));
    hicma_dense *D = (hicma_dense *)calloc(1, sizeof(hicma_dense));
    D->ctx = ctx;
    D->N = N;
    D->ld = N;
    if (hipMalloc(&D->d, sizeof(double) * (size_t)N * N) != hipSuccess) {
        (void)hipGetLastError();
        free(D);
        return HICMA_ERR_NOMEM;
    }
    *out = D;
    return 0;
}

extern "C" int hicma_dense_destroy(hicma_dense *D)
{
    if (!D) return 0;
    (void)hipSetDevice(D->ctx->device);
    (void)hipFree(D->d);
    free(D);
    return 0;
}

extern "C" int hicma_dense_get(const hicma_dense *D, double *host)
{
    if (!D || !host) return HICMA_ERR_ARG;
    HIPCHK(hipSetDevice(D->ctx->device));
    HIPCHK(hipStreamSynchronize(D->ctx->stream));
    HIPCHK(hipMemcpy(host, D->d, sizeof(double) * (size_t)D->N * D->N, hipMemcpyDeviceToHost));
    return 0;
}

extern "C" int hicma_dense_set(hicma_dense *D, const double *host)
{
    if (!D || !host) return HICMA_ERR_ARG;
    HIPCHK(hipSetDevice(D->ctx->device));
    HIPCHK(hipStreamSynchronize(D->ctx->stream));
    HIPCHK(hipMemcpy(D->d, host, sizeof(double) * (size_t)D->N * D->N, hipMemcpyHostToDevice));
    return 0;
}

// starsh_generate_map of the reference harness (hicma_parsec/testing_dpotrf.c:651-653):
// the exact covariance matrix of the descriptor's problem, without approximation
extern "C" int hicma_dense_generate(hicma_ctx *ctx, int uplo, hicma_dense *D, const hicma_tlr_desc *problem)
{
    if (!ctx || !D || !problem || D->N != problem->t.N) return HICMA_ERR_ARG;
    if (uplo != HICMA_LOWER) return HICMA_ERR_UPLO;
    HIPCHK(hipSetDevice(ctx->device));
    hipLaunchKernelGGL(k_dense_gen, elem_grid(D->N), dim3(CK_ROWS), 0, ctx->stream, problem->t, D->d, D->ld);
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(ctx->stream));
    return 0;
}

// STARSH_check (hicma_parsec/hicma_parsec.h:168-184, STARSH_check.jdf:90-143):
// uncompress the lower triangle of A into the dense device matrix A0; the
// strict upper triangle of A0 is zeroed.
extern "C" int STARSH_check(hicma_ctx *ctx, int uplo, hicma_dense *A0, hicma_tlr_desc *A, hicma_tlr_desc *Ar,
                            int band_size, int maxrank, int *info)
{
    (void)Ar;
    (void)maxrank;
    if (!ctx || !A || !A0 || A0->N != A->t.N) return HICMA_ERR_ARG;
    if (uplo != HICMA_LOWER) return HICMA_ERR_UPLO;
    if (band_size != A->t.band_size) return HICMA_ERR_ARG;
    HIPCHK(hipSetDevice(ctx->device));
    const TLR &t = A->t;
    const int nsub = (t.NB + MK_T - 1) / MK_T;
    const long ntile = (long)t.NT * (t.NT + 1) / 2;
    hipLaunchKernelGGL(k_dense_zero_upper, elem_grid(A0->N), dim3(CK_ROWS), 0, ctx->stream, A0->d, A0->ld, (int)A0->N);
    hipLaunchKernelGGL(k_dense_uncompress, dim3(nsub * nsub, (unsigned)std::min<long>(ntile, 65535)), dim3(256), 0,
                       ctx->stream, t, A0->d, A0->ld);
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(ctx->stream));
    if (info) *info = 0;
    return 0;
}

static void blas_destroy(void *h) { (void)rocblas_destroy_handle((rocblas_handle)h); }

// dense Cholesky of the checker: dplasma_dpotrf of the reference harness
// (hicma_parsec/testing_dpotrf.c:766); returns the LAPACK info in *info
extern "C" int hicma_dense_potrf(hicma_ctx *ctx, int uplo, hicma_dense *D, int *info)
{
    if (!ctx || !D) return HICMA_ERR_ARG;
    if (uplo != HICMA_LOWER) return HICMA_ERR_UPLO;
    HIPCHK(hipSetDevice(ctx->device));
    if (!ctx->blas) {
        rocblas_handle h;
        if (rocblas_create_handle(&h) != rocblas_status_success) return HICMA_ERR_HIP;
        ctx->blas = (void *)h;
        ctx->blas_free = blas_destroy;
    }
    rocblas_handle h = (rocblas_handle)ctx->blas;
    if (rocblas_set_stream(h, ctx->stream) != rocblas_status_success) return HICMA_ERR_HIP;
    int64_t *dinfo = (int64_t *)ctx->d_scalar;
    rocblas_status st = rocsolver_dpotrf_64(h, rocblas_fill_lower, (int64_t)D->N, D->d, (int64_t)D->ld, dinfo);
    if (st != rocblas_status_success) {
        fprintf(stderr, "hicma_tlr: rocsolver_dpotrf_64 failed (%d)\n", (int)st);
        return st == rocblas_status_memory_error ? HICMA_ERR_NOMEM : HICMA_ERR_HIP;
    }
    int64_t hinfo = 0;
    HIPCHK(hipStreamSynchronize(ctx->stream));
    HIPCHK(hipMemcpy(&hinfo, dinfo, sizeof(int64_t), hipMemcpyDeviceToHost));
    if (info) *info = (int)hinfo;
    return 0;
}

// dplasma_dlansy(Frobenius, lower) of D, or of A - B
static int lansy(hicma_ctx *ctx, const hicma_dense *A, const hicma_dense *B, double *norm)
{
    const dim3 g = elem_grid(A->N);
    const long nparts = (long)g.x * g.y;
    RCHK(hicma_ensure_ws(ctx, sizeof(double) * (size_t)nparts));
    hipLaunchKernelGGL(k_dense_sumsq, g, dim3(CK_ROWS), 0, ctx->stream, A->d, B ? B->d : nullptr, A->ld, (int)A->N,
                       ctx->ws);
    HIPCHK(hipGetLastError());
    double s = 0.0;
    RCHK(sum_parts(ctx, ctx->ws, nparts, &s));
    *norm = sqrt(s);
    return 0;
}

extern "C" int hicma_dense_lansy(hicma_ctx *ctx, int uplo, const hicma_dense *D, double *norm)
{
    if (!ctx || !D || !norm) return HICMA_ERR_ARG;
    if (uplo != HICMA_LOWER) return HICMA_ERR_UPLO;
    HIPCHK(hipSetDevice(ctx->device));
    return lansy(ctx, D, nullptr, norm);
}

extern "C" int hicma_dense_diff_lansy(hicma_ctx *ctx, int uplo, const hicma_dense *A, const hicma_dense *B,
                                      double *norm)
{
    if (!ctx || !A || !B || !norm || A->N != B->N) return HICMA_ERR_ARG;
    if (uplo != HICMA_LOWER) return HICMA_ERR_UPLO;
    HIPCHK(hipSetDevice(ctx->device));
    return lansy(ctx, A, B, norm);
}

extern "C" int hicma_dense_logdet(hicma_ctx *ctx, const hicma_dense *L, double *logdet)
{
    if (!ctx || !L || !logdet) return HICMA_ERR_ARG;
    HIPCHK(hipSetDevice(ctx->device));
    const long nb = (L->N + 255) / 256;
    RCHK(hicma_ensure_ws(ctx, sizeof(double) * (size_t)nb));
    hipLaunchKernelGGL(k_dense_logdiag, dim3((unsigned)nb), dim3(256), 0, ctx->stream, L->d, L->ld, (int)L->N, ctx->ws);
    HIPCHK(hipGetLastError());
    double s = 0.0;
    RCHK(sum_parts(ctx, ctx->ws, nb, &s));
    *logdet = 2.0 * s;
    return 0;
}

// check_dpotrf2 (hicma_parsec/testing_dpotrf.c:111-180): ||L L^T - A0|| / ||A0||
// with dlansy Frobenius norms of the lower triangles; L is read as lower
// triangular.  Returns 1 when the result is not finite or above threshold.
extern "C" int hicma_check_dpotrf2(hicma_ctx *ctx, int verbose, int uplo, const hicma_dense *L, const hicma_dense *A0,
                                   double threshold, double *result)
{
    if (!ctx || !L || !A0 || L->N != A0->N) return HICMA_ERR_ARG;
    if (uplo != HICMA_LOWER) return HICMA_ERR_UPLO;
    HIPCHK(hipSetDevice(ctx->device));
    const long nbk = (L->N + MK_T - 1) / MK_T;
    const long nblk = nbk * (nbk + 1) / 2;
    RCHK(hicma_ensure_ws(ctx, sizeof(double) * (size_t)nblk));
    hipLaunchKernelGGL(k_dense_resid, dim3((unsigned)nblk), dim3(256), 0, ctx->stream, L->d, A0->d, L->ld, (int)L->N,
                       ctx->ws);
    HIPCHK(hipGetLastError());
    double rs = 0.0, anorm = 0.0;
    RCHK(sum_parts(ctx, ctx->ws, nblk, &rs));
    RCHK(lansy(ctx, A0, nullptr, &anorm));
    const double rnorm = sqrt(rs), res = rnorm / anorm;
    if (result) *result = res;
    if (verbose) {
        printf("============\n");
        printf("Checking the Cholesky factorization. Threshold is %.2e \n", threshold);
        printf("-- ||A||_F = %e, ||L'L-A||_F = %e\n", anorm, rnorm);
        printf("-- ||L'L-A||_F/(||A||_F) = %e\n", res);
    }
    const int bad = !isfinite(rnorm) || !isfinite(res) || res > threshold;
    if (verbose) printf(bad ? "-- Factorization is suspicious ! \n" : "-- Factorization is CORRECT ! \n");
    return bad;
}

// check_diff (hicma_parsec/testing_dpotrf.c:183-230): ||L - L0|| / ||L|| with
// dlansy Frobenius norms of the lower triangles (L and L0 are not modified)
extern "C" int hicma_check_diff(hicma_ctx *ctx, int verbose, int uplo, const hicma_dense *L, const hicma_dense *L0,
                                double threshold, double *result)
{
    if (!ctx || !L || !L0 || L->N != L0->N) return HICMA_ERR_ARG;
    if (uplo != HICMA_LOWER) return HICMA_ERR_UPLO;
    HIPCHK(hipSetDevice(ctx->device));
    double an = 0.0, a0n = 0.0, rn = 0.0;
    RCHK(lansy(ctx, L, nullptr, &an));
    RCHK(lansy(ctx, L0, nullptr, &a0n));
    RCHK(lansy(ctx, L, L0, &rn));
    const double res = rn / an;
    if (result) *result = res;
    if (verbose) {
        printf("============\n");
        printf("Checking the equality of lower/upper triangular parts of two matrices. Threshold is %.2e\n", threshold);
        printf("-- dpotrf: ||L0||_F = %e, HiCMA: ||L||_F = %e, ||L-L0||_F = %e, ||L-L0||_F/||L||_F = %e \n", a0n, an,
               rn, res);
    }
    const int bad = !isfinite(rn) || !isfinite(res) || res > threshold;
    if (verbose) printf(bad ? "-- DIFFERENT matrices ! \n" : "-- SAME matrices ! \n");
    return bad;
}
