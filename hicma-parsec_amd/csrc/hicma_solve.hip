// hicma_solve.hip -- triangular solves with the TLR factor, Sigma^-1 B and
// the Gaussian log-likelihood -1/2 (z^T Sigma^-1 z + log det Sigma + N log 2 pi).
//
// No counterpart in the reference (it stops at the factor); the solves read
// the factor in the layout HiCMA_dpotrf_L leaves (dense band tiles, compact
// [U | V] off-band tiles).  Blocked substitution by tile columns:
//   forward  (B <- L^-1 B):  X_k = L_kk^-1 B_k, then B_m -= L_mk X_k  (m > k)
//   backward (B <- L^-T B):  X_k = L_kk^-T B_k, then B_j -= L_kj^T X_k (j < k)
// with L_mk X_k = U (V^T X_k) for low-rank tiles.  Each diagonal solve uses
// the inverses of the 64 x 64 diagonal blocks of L_kk, formed once per call
// for all tiles in one launch (one wave per block), so a tile's solve is a
// sequence of small matrix-vector products.  All sums run in a fixed order:
// the results do not depend on scheduling.
#include "internal.hpp"
#include "wave.hpp"

#include <math.h>

#define SV_RC 4   // right-hand sides per pass (one wave per right-hand side in the block products)
#define SV_BK 64  // diagonal block edge
#define SV_WL (SV_BK + 1)

// W = L_bb^-1 for every 64 x 64 diagonal block b of every diagonal tile k,
// row-major W(i, c) at Winv[((k nbk + b) 64 + i) 64 + c]; rows and columns
// beyond the tile (NB not a multiple of 64) hold the identity.
__global__ void __launch_bounds__(64) k_tri_inv(TLR t, double *Winv)
{
    __shared__ double Lr[SV_BK * (SV_BK + 2)];
    const int NB = t.NB, nbk = (NB + SV_BK - 1) / SV_BK;
    const int k = blockIdx.x / nbk, b = blockIdx.x % nbk, j0 = SV_BK * b;
    const int jb = min(SV_BK, NB - j0), lane = threadIdx.x;
    const double *Lk = band_tile(t, k, k) + j0 + (size_t)j0 * NB;
    for (int c = 0; c < SV_BK; c++) {
        double v = (lane == c) ? 1.0 : 0.0;
        if (lane < jb && c < jb) v = (c <= lane) ? Lk[lane + (size_t)c * NB] : 0.0;
        Lr[lane * (SV_BK + 2) + c] = v;
    }
    wv_sync();
    // lane = column c of W: right-looking substitution L w = e_c
    double a[SV_BK];
#pragma unroll
    for (int i = 0; i < SV_BK; i++) a[i] = (i == lane) ? 1.0 : 0.0;
#pragma unroll
    for (int l = 0; l < SV_BK; l++) {
        const double w = (lane <= l) ? a[l] / Lr[l * (SV_BK + 2) + l] : 0.0;
        a[l] = w;
#pragma unroll
        for (int i = l + 1; i < SV_BK; i++) a[i] = fma(-Lr[i * (SV_BK + 2) + l], w, a[i]);
    }
    double *W = Winv + (size_t)blockIdx.x * SV_BK * SV_BK;
#pragma unroll
    for (int i = 0; i < SV_BK; i++) W[i * SV_BK + lane] = a[i];
}

// Solve with the diagonal tile k in place on rows k NB .. k NB + NB - 1 of
// the padded right-hand sides P (ld ldp).  One workgroup of 256 threads.
// LDS: x (NBp x SV_RC), W block (64 x 65), partial sums (4 x 64 x SV_RC).
__global__ void __launch_bounds__(256) k_solve_diag(TLR t, int k, const double *Winv, double *P, long ldp, int nrhs,
                                                    int trans)
{
    extern __shared__ double sm[];
    const int NB = t.NB, nbk = (NB + SV_BK - 1) / SV_BK, NBp = nbk * SV_BK;
    double *x = sm;                     // x[l * SV_RC + q]
    double *Ws = x + (size_t)NBp * SV_RC;  // Ws[i * SV_WL + c] = W(i, c)
    double *pt = Ws + SV_BK * SV_WL;    // pt[(g * 64 + i) * SV_RC + q]
    const int tid = threadIdx.x, i = tid & 63, g = tid >> 6, lane = i;
    const double *L = band_tile(t, k, k);
    double *Pk = P + (size_t)k * NB;
    for (int r0 = 0; r0 < nrhs; r0 += SV_RC) {
        const int nq = min(SV_RC, nrhs - r0);
        for (int e = tid; e < NBp * SV_RC; e += 256) {
            const int q = e / NBp, l = e - q * NBp;
            x[l * SV_RC + q] = (l < NB && q < nq) ? Pk[l + (size_t)(r0 + q) * ldp] : 0.0;
        }
        __syncthreads();
        for (int s = 0; s < nbk; s++) {
            const int b = trans ? nbk - 1 - s : s;
            const int c0 = SV_BK * b;
            double acc[SV_RC];
#pragma unroll
            for (int q = 0; q < SV_RC; q++) acc[q] = 0.0;
            if (!trans) {
                // rows of block b times the solved part x(0 : c0): thread (i, g)
                // takes the columns l = g, g + 4, ... (coalesced down a column)
                const int row = c0 + i;
                if (row < NB)
                    for (int l = g; l < c0; l += 4) {
                        const double lv = L[row + (size_t)l * NB];
#pragma unroll
                        for (int q = 0; q < SV_RC; q++) acc[q] = fma(lv, x[l * SV_RC + q], acc[q]);
                    }
#pragma unroll
                for (int q = 0; q < SV_RC; q++) pt[(g * SV_BK + i) * SV_RC + q] = acc[q];
            } else {
                // columns of block b times the solved part x(c0 + 64 : NB):
                // wave g takes the columns c0 + g, c0 + g + 4, ...; its lanes
                // run down the column (coalesced), then a wave sum
                for (int cc = g; cc < SV_BK; cc += 4) {
                    const int col = c0 + cc;
                    double a2[SV_RC];
#pragma unroll
                    for (int q = 0; q < SV_RC; q++) a2[q] = 0.0;
                    if (col < NB)
                        for (int l = c0 + SV_BK + lane; l < NB; l += 64) {
                            const double lv = L[l + (size_t)col * NB];
#pragma unroll
                            for (int q = 0; q < SV_RC; q++) a2[q] = fma(lv, x[l * SV_RC + q], a2[q]);
                        }
#pragma unroll
                    for (int q = 0; q < SV_RC; q++) {
                        const double v = wsum(a2[q]);
                        if (lane == 0) pt[(0 * SV_BK + cc) * SV_RC + q] = v;
                    }
                }
            }
            const double *Wb = Winv + ((size_t)k * nbk + b) * SV_BK * SV_BK;
            for (int e = tid; e < SV_BK * SV_BK; e += 256) Ws[(e >> 6) * SV_WL + (e & 63)] = Wb[e];
            __syncthreads();
            // y = x_b - (partial sums), in pt slot 0
            if (g == 0) {
#pragma unroll
                for (int q = 0; q < SV_RC; q++) {
                    double ssum = pt[i * SV_RC + q];
                    if (!trans)
                        ssum = (ssum + pt[(SV_BK + i) * SV_RC + q]) +
                               (pt[(2 * SV_BK + i) * SV_RC + q] + pt[(3 * SV_BK + i) * SV_RC + q]);
                    pt[i * SV_RC + q] = x[(c0 + i) * SV_RC + q] - ssum;
                }
            }
            __syncthreads();
            // x_b = W y (forward) or W^T y (backward): thread (i, q = g)
            {
                const int q = g;
                double v = 0.0;
                if (!trans)
                    for (int l = 0; l <= i; l++) v = fma(Ws[i * SV_WL + l], pt[l * SV_RC + q], v);
                else
                    for (int l = i; l < SV_BK; l++) v = fma(Ws[l * SV_WL + i], pt[l * SV_RC + q], v);
                x[(c0 + i) * SV_RC + q] = v;
            }
            __syncthreads();
        }
        for (int e = tid; e < NB * SV_RC; e += 256) {
            const int q = e / NB, l = e - q * NB;
            if (q < nq) Pk[l + (size_t)(r0 + q) * ldp] = x[l * SV_RC + q];
        }
        __syncthreads();
    }
}

// The low-rank part of a solve update, [U | V] held as TE (widened on load,
// FP64 arithmetic): w = F^T X_k with F = V (forward) or U (backward), a wave
// per column; then P_o -= G w with G = U (forward) or V (backward).
template <typename TE>
__device__ __forceinline__ void solve_lr(const TE *T, int NB, int r, int trans, const double *xk, double *w,
                                         double *Po, long ldp, int nq)
{
    const int tid = threadIdx.x, lane = tid & 63, g = tid >> 6;
    const TE *U = T, *V = T + (size_t)NB * r;
    const TE *F = trans ? U : V, *Gm = trans ? V : U;
    for (int c = g; c < r; c += 4) {
        double acc[SV_RC];
#pragma unroll
        for (int q = 0; q < SV_RC; q++) acc[q] = 0.0;
        for (int i = lane; i < NB; i += 64) {
            const double fv = (double)F[i + (size_t)c * NB];
#pragma unroll
            for (int q = 0; q < SV_RC; q++) acc[q] = fma(fv, xk[i * SV_RC + q], acc[q]);
        }
#pragma unroll
        for (int q = 0; q < SV_RC; q++) {
            const double v = wsum(acc[q]);
            if (lane == 0) w[c * SV_RC + q] = v;
        }
    }
    __syncthreads();
    for (int i = tid; i < NB; i += 256) {
        double acc[SV_RC];
#pragma unroll
        for (int q = 0; q < SV_RC; q++) acc[q] = 0.0;
        for (int c = 0; c < r; c++) {
            const double gv = (double)Gm[i + (size_t)c * NB];
#pragma unroll
            for (int q = 0; q < SV_RC; q++) acc[q] = fma(gv, w[c * SV_RC + q], acc[q]);
        }
        for (int q = 0; q < nq; q++) Po[i + (size_t)q * ldp] -= acc[q];
    }
}

// The update of one block of right-hand sides by the solved block k:
// forward: P_m -= L_mk X_k for m = k + 1 + blockIdx.x; backward:
// P_j -= L_kj^T X_k for j = blockIdx.x.  LDS: X_k (NB x SV_RC), w (cap x SV_RC).
__global__ void __launch_bounds__(256) k_solve_update(TLR t, int k, double *P, long ldp, int nrhs, int trans)
{
    extern __shared__ double sm[];
    const int NB = t.NB;
    double *xk = sm;                        // xk[l * SV_RC + q]
    double *w = xk + (size_t)NB * SV_RC;    // w[c * SV_RC + q]
    const int tid = threadIdx.x, lane = tid & 63, g = tid >> 6;
    int m, n, o;
    if (!trans) { m = k + 1 + blockIdx.x; n = k; o = m; }
    else { m = k; n = blockIdx.x; o = n; }
    const bool dense = (m - n < t.band_size);
    const int r = dense ? 0 : *rank_at(t, m, n);
    if (!dense && r <= 0) return;
    const double *T = dense ? band_tile(t, m, n) : lr_tile(t, m, n);
    const double *Pk = P + (size_t)k * NB;
    double *Po = P + (size_t)o * NB;
    for (int r0 = 0; r0 < nrhs; r0 += SV_RC) {
        const int nq = min(SV_RC, nrhs - r0);
        for (int e = tid; e < NB * SV_RC; e += 256) {
            const int q = e / NB, l = e - q * NB;
            xk[l * SV_RC + q] = (q < nq) ? Pk[l + (size_t)(r0 + q) * ldp] : 0.0;
        }
        __syncthreads();
        if (dense && !trans) {
            // P_m(i) -= sum_l L_mk(i, l) X_k(l): a thread per row
            for (int i = tid; i < NB; i += 256) {
                double acc[SV_RC];
#pragma unroll
                for (int q = 0; q < SV_RC; q++) acc[q] = 0.0;
                for (int l = 0; l < NB; l++) {
                    const double lv = T[i + (size_t)l * NB];
#pragma unroll
                    for (int q = 0; q < SV_RC; q++) acc[q] = fma(lv, xk[l * SV_RC + q], acc[q]);
                }
                for (int q = 0; q < nq; q++) Po[i + (size_t)(r0 + q) * ldp] -= acc[q];
            }
        } else if (dense) {
            // P_j(l) -= sum_i L_kj(i, l) X_k(i): a wave per column l
            for (int l = g; l < NB; l += 4) {
                double acc[SV_RC];
#pragma unroll
                for (int q = 0; q < SV_RC; q++) acc[q] = 0.0;
                for (int i = lane; i < NB; i += 64) {
                    const double lv = T[i + (size_t)l * NB];
#pragma unroll
                    for (int q = 0; q < SV_RC; q++) acc[q] = fma(lv, xk[i * SV_RC + q], acc[q]);
                }
#pragma unroll
                for (int q = 0; q < SV_RC; q++) {
                    const double v = wsum(acc[q]);
                    if (lane == 0 && q < nq) Po[l + (size_t)(r0 + q) * ldp] -= v;
                }
            }
        } else {
            const int p = lr_prec(t, m, n);
            if (p == LR_F64) solve_lr(T, NB, r, trans, xk, w, Po + (size_t)r0 * ldp, ldp, nq);
            else if (p == LR_F32) solve_lr((const float *)T, NB, r, trans, xk, w, Po + (size_t)r0 * ldp, ldp, nq);
            else solve_lr((const lr_bf16 *)T, NB, r, trans, xk, w, Po + (size_t)r0 * ldp, ldp, nq);
        }
        __syncthreads();
    }
}

// P (ld ldp, NT NB rows) <- B (ld ldb, N rows), padding rows zeroed; or back
__global__ void k_pad_copy(const double *src, long lds, double *dst, long ldd, int rows_src, int rows_dst, int nrhs)
{
    const long e = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= (long)rows_dst * nrhs) return;
    const int q = (int)(e / rows_dst), i = (int)(e - (long)q * rows_dst);
    dst[i + q * ldd] = (i < rows_src) ? src[i + q * lds] : 0.0;
}

// sum of squares of x(0 : n), per-block parts, and their fixed-order total
__global__ void __launch_bounds__(256) k_sumsq_parts(const double *x, int n, double *part)
{
    __shared__ double red[4];
    const int i = blockIdx.x * 256 + threadIdx.x;
    double v = (i < n) ? x[i] * x[i] : 0.0;
    v = wsum(v);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
    __syncthreads();
    if (threadIdx.x == 0) part[blockIdx.x] = (red[0] + red[1]) + (red[2] + red[3]);
}

__global__ void __launch_bounds__(256) k_sum_seq(const double *part, int n, double *out)
{
    __shared__ double red[4];
    double s = 0.0;
    for (int q = threadIdx.x; q < n; q += 256) s += part[q];
    s = wsum(s);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = s;
    __syncthreads();
    if (threadIdx.x == 0) *out = (red[0] + red[1]) + (red[2] + red[3]);
}

// 2 sum log L(i, i) over the true N, in the order of k_logdet (hicma_logdet)
__global__ void __launch_bounds__(256) k_logdet_tiles(TLR t, double *out)
{
    __shared__ double red[4];
    double part = 0.0;
    for (int gi = threadIdx.x; gi < t.N; gi += 256) {
        const int k = gi / t.NB, i = gi % t.NB;
        part += log(band_tile(t, k, k)[i + (size_t)i * t.NB]);
    }
    part = wsum(part);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = part;
    __syncthreads();
    if (threadIdx.x == 0) *out = 2.0 * (((red[0] + red[1]) + red[2]) + red[3]);
}

// ---------------------------------------------------------------------------
// host side

// doubles of workspace a solve needs: block inverses, padded right-hand
// sides, reduction parts, three scalars
size_t solve_ws_doubles(const TLR &t, int nrhs)
{
    const size_t nbk = (size_t)(t.NB + SV_BK - 1) / SV_BK;
    const size_t Np = (size_t)t.NT * t.NB;
    return nbk * t.NT * SV_BK * SV_BK + Np * (size_t)nrhs + (Np + 255) / 256 + 8;
}

size_t hicma_solve_bytes(int N, int NB)
{
    TLR t{};
    t.N = N;
    t.NB = NB;
    t.NT = (N + NB - 1) / NB;
    return sizeof(double) * solve_ws_doubles(t, 1);
}

static size_t diag_lds(const TLR &t)
{
    const size_t nbk = (size_t)(t.NB + SV_BK - 1) / SV_BK;
    return sizeof(double) * (nbk * SV_BK * SV_RC + SV_BK * SV_WL + 4 * SV_BK * SV_RC);
}
static size_t update_lds(const TLR &t) { return sizeof(double) * ((size_t)t.NB + (size_t)t.cap) * SV_RC; }

// dirs: 1 forward, 2 backward, 3 both.  B: device, N x nrhs, ld ldb.
// quad (if not null): ||L^-1 B(:, 0)||^2 after the forward pass.  ws: device
// workspace of solve_ws_doubles() doubles.  Asynchronous on the context stream.
int solve_launch(hicma_ctx *ctx, hicma_tlr_desc *A, int dirs, int nrhs, double *B, long ldb, double *ws,
                 double *d_quad, double *d_logdet)
{
    const TLR &t = A->t;
    const int NT = t.NT, NB = t.NB, nbk = (NB + SV_BK - 1) / SV_BK;
    const long Np = (long)NT * NB;
    hipStream_t st = ctx->stream;
    double *Winv = ws;
    double *P = Winv + (size_t)nbk * NT * SV_BK * SV_BK;
    double *parts = P + (size_t)Np * nrhs;
    const size_t ld1 = diag_lds(t), ld2 = update_lds(t);
    HIPCHK(hipFuncSetAttribute((const void *)k_solve_diag, hipFuncAttributeMaxDynamicSharedMemorySize, (int)ld1));
    HIPCHK(hipFuncSetAttribute((const void *)k_solve_update, hipFuncAttributeMaxDynamicSharedMemorySize, (int)ld2));
    hipLaunchKernelGGL(k_tri_inv, dim3((unsigned)(NT * nbk)), dim3(64), 0, st, t, Winv);
    const long ne = Np * nrhs;
    hipLaunchKernelGGL(k_pad_copy, dim3((unsigned)((ne + 255) / 256)), dim3(256), 0, st, (const double *)B, ldb, P, Np,
                       t.N, (int)Np, nrhs);
    if (dirs & 1) {
        for (int k = 0; k < NT; k++) {
            hipLaunchKernelGGL(k_solve_diag, dim3(1), dim3(256), ld1, st, t, k, (const double *)Winv, P, Np, nrhs, 0);
            if (k + 1 < NT)
                hipLaunchKernelGGL(k_solve_update, dim3(NT - 1 - k), dim3(256), ld2, st, t, k, P, Np, nrhs, 0);
        }
        if (d_quad) {
            const int nb = (int)((t.N + 255) / 256);
            hipLaunchKernelGGL(k_sumsq_parts, dim3(nb), dim3(256), 0, st, (const double *)P, t.N, parts);
            hipLaunchKernelGGL(k_sum_seq, dim3(1), dim3(256), 0, st, (const double *)parts, nb, d_quad);
        }
    }
    if (dirs & 2) {
        for (int k = NT - 1; k >= 0; k--) {
            hipLaunchKernelGGL(k_solve_diag, dim3(1), dim3(256), ld1, st, t, k, (const double *)Winv, P, Np, nrhs, 1);
            if (k > 0) hipLaunchKernelGGL(k_solve_update, dim3(k), dim3(256), ld2, st, t, k, P, Np, nrhs, 1);
        }
    }
    if (d_logdet) hipLaunchKernelGGL(k_logdet_tiles, dim3(1), dim3(256), 0, st, t, d_logdet);
    hipLaunchKernelGGL(k_pad_copy, dim3((unsigned)(((long)t.N * nrhs + 255) / 256)), dim3(256), 0, st,
                       (const double *)P, Np, B, ldb, (int)Np, t.N, nrhs);
    HIPCHK(hipGetLastError());
    return 0;
}

int check_solve_args(hicma_ctx *ctx, hicma_tlr_desc *A, int nrhs, const double *B, long ldb)
{
    if (!ctx || !A || !B || nrhs < 1 || ldb < A->t.N) return HICMA_ERR_ARG;
    if (update_lds(A->t) > ctx->lds_max || diag_lds(A->t) > ctx->lds_max) return HICMA_ERR_ARG;
    return 0;
}

extern "C" int hicma_tlr_trsm(hicma_ctx *ctx, hicma_tlr_desc *A, int trans, int nrhs, double *B, int ldb)
{
    RCHK(check_solve_args(ctx, A, nrhs, B, ldb));
    if (trans != 0 && trans != 1) return HICMA_ERR_ARG;
    HIPCHK(hipSetDevice(ctx->device));
    RCHK(hicma_ensure_ws(ctx, sizeof(double) * solve_ws_doubles(A->t, nrhs)));
    RCHK(solve_launch(ctx, A, trans ? 2 : 1, nrhs, B, ldb, ctx->ws, nullptr, nullptr));
    HIPCHK(hipStreamSynchronize(ctx->stream));
    return 0;
}

extern "C" int hicma_tlr_potrs(hicma_ctx *ctx, hicma_tlr_desc *A, int nrhs, double *B, int ldb)
{
    RCHK(check_solve_args(ctx, A, nrhs, B, ldb));
    HIPCHK(hipSetDevice(ctx->device));
    RCHK(hicma_ensure_ws(ctx, sizeof(double) * solve_ws_doubles(A->t, nrhs)));
    RCHK(solve_launch(ctx, A, 3, nrhs, B, ldb, ctx->ws, nullptr, nullptr));
    HIPCHK(hipStreamSynchronize(ctx->stream));
    return 0;
}

// log-likelihood of one vector z (device, N entries; overwritten by L^-1 z)
int loglik_finish(hicma_ctx *ctx, const TLR &t, double *d_vals, double *logdet, double *quad, double *loglik)
{
    double h[2];
    // wait for the stream first, then copy: a copy into pageable memory that
    // waits for its stream holds the runtime's staging path meanwhile, and
    // other threads' copies queue behind it
    HIPCHK(hipStreamSynchronize(ctx->stream));
    HIPCHK(hipMemcpy(h, d_vals, 2 * sizeof(double), hipMemcpyDeviceToHost));
    if (quad) *quad = h[0];
    if (logdet) *logdet = h[1];
    if (loglik) *loglik = -0.5 * (h[0] + h[1] + (double)t.N * log(2.0 * 3.14159265358979323846));
    return 0;
}

extern "C" int hicma_tlr_loglik(hicma_ctx *ctx, hicma_tlr_desc *A, double *z, double *logdet, double *quad,
                                double *loglik)
{
    RCHK(check_solve_args(ctx, A, 1, z, A ? A->t.N : 0));
    HIPCHK(hipSetDevice(ctx->device));
    const size_t need = solve_ws_doubles(A->t, 1);
    RCHK(hicma_ensure_ws(ctx, sizeof(double) * need));
    double *vals = ctx->ws + need - 8;
    RCHK(solve_launch(ctx, A, 1, 1, z, A->t.N, ctx->ws, vals, vals + 1));
    return loglik_finish(ctx, A->t, vals, logdet, quad, loglik);
}

// ---- device memory plumbing for callers without a GPU runtime of their own ----
extern "C" int hicma_dev_malloc(hicma_ctx *ctx, size_t bytes, void **p)
{
    if (!ctx || !p) return HICMA_ERR_ARG;
    HIPCHK(hipSetDevice(ctx->device));
    HIPCHK(ctx_malloc(ctx, p, bytes ? bytes : 1));
    return 0;
}

extern "C" int hicma_dev_free(hicma_ctx *ctx, void *p)
{
    if (!ctx) return HICMA_ERR_ARG;
    HIPCHK(hipSetDevice(ctx->device));
    if (p) HIPCHK(ctx_free(ctx, p));
    return 0;
}

extern "C" int hicma_memcpy(hicma_ctx *ctx, void *dst, const void *src, size_t bytes, int to_device)
{
    if (!ctx || (!dst && bytes) || (!src && bytes)) return HICMA_ERR_ARG;
    HIPCHK(hipSetDevice(ctx->device));
    HIPCHK(hipStreamSynchronize(ctx->stream));
    HIPCHK(hipMemcpy(dst, src, bytes, to_device ? hipMemcpyHostToDevice : hipMemcpyDeviceToHost));
    return 0;
}
