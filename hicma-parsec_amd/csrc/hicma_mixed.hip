// hicma_mixed.hip -- mixed-precision storage of a TLR factor.
//
// The factorization's kernels (potrf, trsm, syrk, the gemm updates and the
// QR/SVD recompression) run in FP64.  Once a factor is complete, every
// off-band tile L_mn with ||L_mn||_F below a threshold (eps ||A||_F for
// FP32, a smaller one for BF16) can be rounded in place to FP32 or BF16:
// its rounding error, 2^-24 ||L_mn||_F (FP32) or 2^-8 ||L_mn||_F (BF16), is
// then below 2^-24 eps ||A||_F, far inside the factorization's own
// truncation error eps.  The solves, the log-likelihood, the dense check
// and hicma_get_tile read such tiles in their stored type and widen them
// to FP64 on load, so their HBM traffic on those tiles halves (FP32) or
// quarters (BF16).  A tile's element type is in the descriptor's precision
// map; the packed [U | V] starts at the slot's first byte, V at element
// NB*r.  The reference holds every tile in FP64 and has no counterpart.
#include "internal.hpp"
#include "mfma.hpp"
#include "wave.hpp"

#include <math.h>
#include <vector>

#define MX_CHUNK 2048  // elements per in-place conversion pass (8 per thread)

// fixed-order sum over a 256-thread block
__device__ __forceinline__ double mx_block_sum(double v, double *red)
{
    v = wsum(v);
    const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
    __syncthreads();
    if (l == 0) red[w] = v;
    __syncthreads();
    return red[0] + red[1] + red[2] + red[3];
}

// ||U V^T||_F^2 = sum_ij (U^T U)_ij (V^T V)_ij: both Gram matrices by 64x64
// blocks on the matrix cores, blocks above the diagonal counted twice.
// Returns this thread's part of the sum.
template <typename TE>
__device__ double lr_frob2_part(const TE *U, int NB, int r, double *sm)
{
    const TE *V = U + (size_t)NB * r;
    const int nbk = (r + MK_T - 1) / MK_T;
    double part = 0.0;
    for (int a = 0; a < nbk; a++)
        for (int b = a; b < nbk; b++) {
            mk_d4 g1[2][2], g2[2][2];
            mk_zero(g1);
            mk_zero(g2);
            // (U^T U)(i, j) = sum_k U(k, i) U(k, j): A k-contiguous, B k-contiguous
            const MkOps<true, true, TE> ou{U, NB, U, NB}, ov{V, NB, V, NB};
            mk_accum(ou, r, r, a * MK_T, b * MK_T, 0, NB, g1, sm);
            mk_accum(ov, r, r, a * MK_T, b * MK_T, 0, NB, g2, sm);
            const double w = (a == b) ? 1.0 : 2.0;
#pragma unroll
            for (int x = 0; x < 2; x++)
#pragma unroll
                for (int y = 0; y < 2; y++)
#pragma unroll
                    for (int z = 0; z < 4; z++) part = fma(w * g1[x][y][z], g2[x][y][z], part);
        }
    return part;
}

// Frobenius norm of every lower tile (m >= n) into nrm[n*NT + m]: band tiles
// over their rows < tdim(m) and columns < tdim(n) (the padding left out),
// low-rank tiles from their Gram matrices in the type they are held in.
// One workgroup per tile, grid-stride over the NT (NT + 1) / 2 tiles.
__global__ void __launch_bounds__(256) k_tile_normf(TLR t, double *nrm)
{
    __shared__ double sm[MK_SMEM];
    __shared__ double red[4];
    const int NB = t.NB, NT = t.NT, tid = threadIdx.x;
    const long ntile = (long)NT * (NT + 1) / 2;
    for (long tix = blockIdx.x; tix < ntile; tix += gridDim.x) {
        int hi, lo;
        tri_decode(tix, hi, lo);  // tix = hi (hi - 1) / 2 + lo, lo < hi
        const int m = hi - 1, n = lo;
        double part = 0.0;
        if (m - n < t.band_size) {
            const double *T = band_tile(t, m, n);
            const int rows = tdim(t, m), cols = tdim(t, n);
            for (long e = tid; e < (long)rows * cols; e += 256) {
                const int j = (int)(e / rows), i = (int)(e - (long)j * rows);
                const double v = T[i + (size_t)j * NB];
                part = fma(v, v, part);
            }
        } else {
            const int r = *rank_at(t, m, n);
            if (r > 0) {
                const double *T = lr_tile(t, m, n);
                const int p = lr_prec(t, m, n);
                if (p == LR_F64) part = lr_frob2_part(T, NB, r, sm);
                else if (p == LR_F32) part = lr_frob2_part((const float *)T, NB, r, sm);
                else part = lr_frob2_part((const lr_bf16 *)T, NB, r, sm);
            }
        }
        const double s = mx_block_sum(part, red);
        if (tid == 0) nrm[(size_t)n * NT + m] = sqrt(fmax(s, 0.0));
        __syncthreads();
    }
}

// round to BF16 through FP32: double -> float (round to nearest even), then
// keep the upper 16 bits of the float with round to nearest even
__device__ __forceinline__ unsigned short mx_to_bf16(double d)
{
    unsigned u = __float_as_uint((float)d);
    u += 0x7FFFu + ((u >> 16) & 1u);
    return (unsigned short)(u >> 16);
}

// In-place demotion of the off-band tiles with nrm < thr32 (FP32) or
// nrm < thr16 (BF16).  One workgroup per tile.  Each pass reads MX_CHUNK
// doubles of [U | V] into LDS (so every load of the pass has returned
// before the barrier) and writes them packed at the same element indices:
// the packed bytes of elements c0.. lie below byte 8 (c0 + MX_CHUNK), so a
// pass overwrites only doubles already read.  counts[p] += tiles now held
// in type p (tiles of rank 0 count as FP64).
__global__ void __launch_bounds__(256) k_lr_demote(TLR t, const double *nrm, double thr32, double thr16,
                                                  int *counts)
{
    __shared__ double buf[MX_CHUNK];
    const int NB = t.NB, NT = t.NT, tid = threadIdx.x;
    const long ntile = (long)NT * (NT - 1) / 2;
    for (long tix = blockIdx.x; tix < ntile; tix += gridDim.x) {
        int m, n;
        tri_decode(tix, m, n);
        if (m - n < t.band_size) continue;
        const int r = *rank_at(t, m, n);
        int p = lr_prec(t, m, n);
        if (r > 0 && p == LR_F64) {
            const double v = nrm[(size_t)n * NT + m];
            p = (v < thr16) ? LR_BF16 : ((v < thr32) ? LR_F32 : LR_F64);
        }
        if (r > 0 && p != LR_F64 && lr_prec(t, m, n) == LR_F64) {
            double *T = lr_tile(t, m, n);
            const size_t ne = (size_t)NB * 2 * r;
            for (size_t c0 = 0; c0 < ne; c0 += MX_CHUNK) {
#pragma unroll
                for (int q = 0; q < MX_CHUNK / 256; q++) {
                    const size_t e = c0 + tid + 256 * q;
                    buf[tid + 256 * q] = (e < ne) ? T[e] : 0.0;
                }
                __syncthreads();
#pragma unroll
                for (int q = 0; q < MX_CHUNK / 256; q++) {
                    const size_t e = c0 + tid + 256 * q;
                    if (e >= ne) continue;
                    const double v = buf[tid + 256 * q];
                    if (p == LR_F32) ((float *)T)[e] = (float)v;
                    else ((unsigned short *)T)[e] = mx_to_bf16(v);
                }
                __syncthreads();
            }
            if (tid == 0) t.prec[(size_t)n * NT + m] = (unsigned char)p;
        }
        if (tid == 0) atomicAdd(counts + (r > 0 ? p : LR_F64), 1);
        __syncthreads();
    }
}

// ---------------------------------------------------------------------------
// C ABI

static int normf_launch(hicma_ctx *ctx, hicma_tlr_desc *A, double **d_nrm)
{
    const TLR &t = A->t;
    const size_t nn = (size_t)t.NT * t.NT;
    RCHK(hicma_ensure_ws(ctx, sizeof(double) * nn + 4 * sizeof(int)));
    *d_nrm = ctx->ws;
    const long ntile = (long)t.NT * (t.NT + 1) / 2;
    const unsigned grid = (unsigned)(ntile < 8192 ? ntile : 8192);
    hipLaunchKernelGGL(k_tile_normf, dim3(grid), dim3(256), 0, ctx->stream, t, ctx->ws);
    HIPCHK(hipGetLastError());
    return 0;
}

extern "C" int hicma_tlr_normf(hicma_tlr_desc *A, double *normF, double *tile_norms)
{
    if (!A || (!normF && !tile_norms)) return HICMA_ERR_ARG;
    hicma_ctx *ctx = A->ctx;
    HIPCHK(hipSetDevice(ctx->device));
    double *d_nrm = nullptr;
    RCHK(normf_launch(ctx, A, &d_nrm));
    const int NT = A->t.NT;
    std::vector<double> h((size_t)NT * NT);
    HIPCHK(hipStreamSynchronize(ctx->stream));
    HIPCHK(hipMemcpy(h.data(), d_nrm, sizeof(double) * h.size(), hipMemcpyDeviceToHost));
    // the symmetric matrix: diagonal tiles once, lower tiles twice; fixed order
    double s = 0.0;
    for (int n = 0; n < NT; n++)
        for (int m = n; m < NT; m++) {
            const double v = h[(size_t)n * NT + m];
            s += (m == n ? 1.0 : 2.0) * v * v;
        }
    if (normF) *normF = sqrt(s);
    if (tile_norms)
        for (int n = 0; n < NT; n++)
            for (int m = 0; m < NT; m++) tile_norms[(size_t)n * NT + m] = (m >= n) ? h[(size_t)n * NT + m] : 0.0;
    return 0;
}

extern "C" int hicma_tlr_demote(hicma_tlr_desc *A, double thr32, double thr16, int *counts)
{
    if (!A || !(thr32 >= 0.0) || thr16 > thr32) return HICMA_ERR_ARG;
    hicma_ctx *ctx = A->ctx;
    HIPCHK(hipSetDevice(ctx->device));
    double *d_nrm = nullptr;
    RCHK(normf_launch(ctx, A, &d_nrm));
    const TLR &t = A->t;
    int *d_cnt = (int *)(d_nrm + (size_t)t.NT * t.NT);
    HIPCHK(hipMemsetAsync(d_cnt, 0, 4 * sizeof(int), ctx->stream));
    const long ntile = (long)t.NT * (t.NT - 1) / 2;
    if (ntile > 0) {
        const unsigned grid = (unsigned)(ntile < 8192 ? ntile : 8192);
        hipLaunchKernelGGL(k_lr_demote, dim3(grid), dim3(256), 0, ctx->stream, t, (const double *)d_nrm, thr32,
                           thr16, d_cnt);
        HIPCHK(hipGetLastError());
    }
    int h[4];
    HIPCHK(hipStreamSynchronize(ctx->stream));
    HIPCHK(hipMemcpy(h, d_cnt, 4 * sizeof(int), hipMemcpyDeviceToHost));
    A->demoted = h[LR_F32] + h[LR_BF16];
    if (counts) {
        counts[0] = h[LR_F64];
        counts[1] = h[LR_F32];
        counts[2] = h[LR_BF16];
    }
    return 0;
}
