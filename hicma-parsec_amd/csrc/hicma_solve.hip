// hicma_solve.hip -- triangular solves with the TLR factor, Sigma^-1 B and
// the Gaussian log-likelihood -1/2 (z^T Sigma^-1 z + log det Sigma + N log 2 pi).
//
// No counterpart in the reference (it stops at the factor); the solves read
// the factor in the layout HiCMA_dpotrf_L leaves (dense band tiles, compact
// [U | V] off-band tiles, each in the type the precision map gives).
// Blocked substitution by tile columns:
//   forward  (B <- L^-1 B):  X_k = L_kk^-1 B_k, then B_m -= L_mk X_k  (m > k)
//   backward (B <- L^-T B):  X_k = L_kk^-T B_k, then B_j -= L_kj^T X_k (j < k)
// with L_mk X_k = U (V^T X_k) for low-rank tiles.  Every L_kk^-1 is formed
// once per call (the 64 x 64 diagonal-block inverses one wave each, the
// blocks below on the matrix cores), so the diagonal step is a block
// matrix-vector product over 64-row blocks in parallel instead of a chain
// of dependent 64-row substitutions, and each update runs one workgroup per
// 64-row block of each tile.  The NT diagonal/update launch pairs of a pass
// remain a chain.  A pass reads its right-hand sides from one buffer and
// writes its solution to another, so no launch updates in place what another
// workgroup of it still reads.  All sums run in a fixed order: the results do
// not depend on scheduling.
#include "internal.hpp"
#include "mfma.hpp"
#include "wave.hpp"

#include <math.h>
#include <algorithm>
#include <utility>

#define SV_RC 4                // right-hand sides per pass
#define SV_BK 64               // diagonal block edge
#define SU_T 1024              // threads of the diagonal and update kernels
#define SU_G (SU_T / SV_BK)    // their column groups (one wave each)

// L_kk^-1 is held as nb packed block-column panels per diagonal tile: panel
// c holds block rows c .. nb-1 of block column c, column major with
// ld 64 (nb - c); the panels of tile k start at Xinv + k inv_tile(nb).
__host__ __device__ __forceinline__ size_t inv_tile(int nb) { return (size_t)SV_BK * SV_BK * nb * (nb + 1) / 2; }
__host__ __device__ __forceinline__ size_t inv_poff(int nb, int c)
{
    return (size_t)SV_BK * SV_BK * ((size_t)c * nb - (size_t)c * (c - 1) / 2);
}

// W = L_bb^-1 for every 64 x 64 diagonal block b of every diagonal tile k,
// into block 0 of panel b (rows and columns beyond the tile, NB not a
// multiple of 64, hold the identity).  One wave per block: lane c forms
// column c of W by forward substitution, W(i, c) = (e_c(i) - sum_{c <= l < i}
// L(i, l) W(l, c)) / L(i, i), with L and W in LDS.
__global__ void __launch_bounds__(64) k_tri_inv(TLR t, double *Xinv)
{
    __shared__ double Lr[SV_BK * (SV_BK + 1)];  // L(i, l) at Lr[i * 65 + l]
    __shared__ double Ws[SV_BK * (SV_BK + 1)];  // W(i, c) at Ws[i * 65 + c]
    const int NB = t.NB, nbk = (NB + SV_BK - 1) / SV_BK;
    const int k = blockIdx.x / nbk, b = blockIdx.x % nbk, j0 = SV_BK * b;
    const int jb = min(SV_BK, NB - j0), lane = threadIdx.x;
    const double *Lk = band_tile(t, k, k) + j0 + (size_t)j0 * NB;
    for (int c = 0; c < SV_BK; c++) {
        double v = (lane == c) ? 1.0 : 0.0;
        if (lane < jb && c < jb) v = (c <= lane) ? Lk[lane + (size_t)c * NB] : 0.0;
        Lr[lane * (SV_BK + 1) + c] = v;
    }
    wv_sync();
    const int c = lane;
    for (int i = 0; i < SV_BK; i++) {
        double s = (i == c) ? 1.0 : 0.0;
        for (int l = c; l < i; l++) s = fma(-Lr[i * (SV_BK + 1) + l], Ws[l * (SV_BK + 1) + c], s);
        Ws[i * (SV_BK + 1) + c] = (i >= c) ? s / Lr[i * (SV_BK + 1) + i] : 0.0;
    }
    wv_sync();
    const size_t ld = (size_t)SV_BK * (nbk - b);
    double *W = Xinv + (size_t)k * inv_tile(nbk) + inv_poff(nbk, b);
    for (int j = 0; j < SV_BK; j++) W[lane + (size_t)j * ld] = Ws[lane * (SV_BK + 1) + j];
}

// The off-diagonal blocks of L_kk^-1, block column c of tile k per
// workgroup (grid NT nb): X_bc = -W_b sum_{j=c}^{b-1} L_bj X_jc for
// b = c+1 .. nb-1, each sum one matrix-core product of the row panel
// L(b, c .. b-1) with the panel's blocks c .. b-1, stored into block b's
// slot, then multiplied by W_b from panel b.
__global__ void __launch_bounds__(256) k_tile_inv(TLR t, double *Xinv)
{
    __shared__ double sm[MK_SMEM];
    const int NB = t.NB, nb = (NB + SV_BK - 1) / SV_BK;
    const int k = blockIdx.x / nb, c = blockIdx.x % nb;
    const int ldx = SV_BK * (nb - c);
    double *X0 = Xinv + (size_t)k * inv_tile(nb);
    double *Xp = X0 + inv_poff(nb, c);
    const double *Lk = band_tile(t, k, k);
    for (int b = c + 1; b < nb; b++) {
        mk_d4 acc[2][2];
        mk_zero(acc);
        // S = L(64b .., 64c .. 64b) Xp(0 .. 64(b - c), :)
        const MkOps<false, true> op{Lk + (size_t)SV_BK * c * NB, NB, Xp, ldx};
        mk_accum(op, NB, SV_BK, SV_BK * b, 0, 0, SV_BK * (b - c), acc, sm);
        double *Sb = Xp + SV_BK * (b - c);
        mk_store(acc, [&](int i, int j, double v) { Sb[i + (size_t)j * ldx] = v; });
        __syncthreads();
        // X_bc = -W_b S; W_b is block 0 of panel b
        mk_zero(acc);
        const MkOps<false, true> op2{X0 + inv_poff(nb, b), SV_BK * (nb - b), Sb, ldx};
        mk_accum(op2, SV_BK, SV_BK, 0, 0, 0, SV_BK, acc, sm);  // ends with a barrier: S is read
        mk_store(acc, [&](int i, int j, double v) { Sb[i + (size_t)j * ldx] = -v; });
        __syncthreads();
    }
}

// x_k = L_kk^-1 y_k (forward) or L_kk^-T y_k (backward), block row b per
// workgroup (grid nb): y from the rows of tile k in Pin, x into Pout.
// forward:  x_b = sum_{c <= b} X_bc y_c, thread (i, g) takes columns g + 16u
//           of each block, the 16 groups summed in order;
// backward: x_b = sum_{c >= b} X_cb^T y_c, wave g takes columns g + 16u of
//           panel b, its lanes run down the rows.
// LDS: y (64 nb x SV_RC), partial sums (16 x 64 x SV_RC).
__global__ void __launch_bounds__(SU_T) k_solve_diag(TLR t, int k, const double *Xinv, const double *Pin,
                                                     double *Pout, long ldp, int nrhs, int trans)
{
    extern __shared__ double sm[];
    const int NB = t.NB, nb = (NB + SV_BK - 1) / SV_BK, NBp = nb * SV_BK;
    double *y = sm;                            // y[l * SV_RC + q]
    double *pt = y + (size_t)NBp * SV_RC;      // pt[(g * 64 + i) * SV_RC + q]
    const int tid = threadIdx.x, i = tid & 63, g = tid >> 6, b = blockIdx.x;
    const double *X0 = Xinv + (size_t)k * inv_tile(nb);
    const double *Pk = Pin + (size_t)k * NB;
    double *Po = Pout + (size_t)k * NB;
    for (int r0 = 0; r0 < nrhs; r0 += SV_RC) {
        const int nq = min(SV_RC, nrhs - r0);
        for (int e = tid; e < NBp * SV_RC; e += SU_T) {
            const int q = e / NBp, l = e - q * NBp;
            y[l * SV_RC + q] = (l < NB && q < nq) ? Pk[l + (size_t)(r0 + q) * ldp] : 0.0;
        }
        __syncthreads();
        if (!trans) {
            double acc[SV_RC];
#pragma unroll
            for (int q = 0; q < SV_RC; q++) acc[q] = 0.0;
#pragma unroll 2
            for (int c = 0; c <= b; c++) {
                const size_t ld = (size_t)SV_BK * (nb - c);
                const double *Bp = X0 + inv_poff(nb, c) + SV_BK * (b - c) + i;
                double v[SV_BK / SU_G];
#pragma unroll
                for (int u = 0; u < SV_BK / SU_G; u++) v[u] = Bp[(size_t)(g + SU_G * u) * ld];
#pragma unroll
                for (int u = 0; u < SV_BK / SU_G; u++)
#pragma unroll
                    for (int q = 0; q < SV_RC; q++)
                        acc[q] = fma(v[u], y[(SV_BK * c + g + SU_G * u) * SV_RC + q], acc[q]);
            }
#pragma unroll
            for (int q = 0; q < SV_RC; q++) pt[(g * SV_BK + i) * SV_RC + q] = acc[q];
        } else {
            const size_t ld = (size_t)SV_BK * (nb - b);
            const double *Pb = X0 + inv_poff(nb, b) + i;
            double acc[SV_BK / SU_G][SV_RC];
#pragma unroll
            for (int u = 0; u < SV_BK / SU_G; u++)
#pragma unroll
                for (int q = 0; q < SV_RC; q++) acc[u][q] = 0.0;
#pragma unroll 2
            for (int c = b; c < nb; c++) {
                double v[SV_BK / SU_G];
#pragma unroll
                for (int u = 0; u < SV_BK / SU_G; u++) v[u] = Pb[SV_BK * (c - b) + (size_t)(g + SU_G * u) * ld];
#pragma unroll
                for (int u = 0; u < SV_BK / SU_G; u++)
#pragma unroll
                    for (int q = 0; q < SV_RC; q++) acc[u][q] = fma(v[u], y[(SV_BK * c + i) * SV_RC + q], acc[u][q]);
            }
#pragma unroll
            for (int u = 0; u < SV_BK / SU_G; u++)
#pragma unroll
                for (int q = 0; q < SV_RC; q++) {
                    const double s = wsum(acc[u][q]);
                    if (i == 0) pt[(g + SU_G * u) * SV_RC + q] = s;
                }
        }
        __syncthreads();
        if (g == 0) {
            const int row = SV_BK * b + i;
#pragma unroll
            for (int q = 0; q < SV_RC; q++) {
                double s = pt[i * SV_RC + q];
                if (!trans)
                    for (int gg = 1; gg < SU_G; gg++) s += pt[(gg * SV_BK + i) * SV_RC + q];
                if (row < NB && q < nq) Po[row + (size_t)(r0 + q) * ldp] = s;
            }
        }
        __syncthreads();
    }
}

// w = F^T x (F = NB x r held as TE, x in LDS): wave g takes columns g + 16u,
// its lanes run down the rows; w[c * SV_RC + q] in LDS
template <typename TE>
__device__ __forceinline__ void su_lr_w(const TE *F, int NB, int r, const double *x, double *w)
{
    const int lane = threadIdx.x & 63, g = threadIdx.x >> 6;
    for (int c = g; c < r; c += SU_G) {
        const TE *Fc = F + (size_t)c * NB;
        double acc[SV_RC];
#pragma unroll
        for (int q = 0; q < SV_RC; q++) acc[q] = 0.0;
#pragma unroll 4
        for (int i = lane; i < NB; i += 64) {
            const double fv = (double)Fc[i];
#pragma unroll
            for (int q = 0; q < SV_RC; q++) acc[q] = fma(fv, x[i * SV_RC + q], acc[q]);
        }
#pragma unroll
        for (int q = 0; q < SV_RC; q++) {
            const double v = wsum(acc[q]);
            if (lane == 0) w[c * SV_RC + q] = v;
        }
    }
}

// P_o(row) -= G(row, :) w for every row, one row per thread (G = NB x r
// held as TE)
template <typename TE>
__device__ __forceinline__ void su_lr_rows(const TE *G, int NB, int r, const double *w, double *Po, long ldp, int nq)
{
    for (int row = threadIdx.x; row < NB; row += SU_T) {
        double acc[SV_RC];
#pragma unroll
        for (int q = 0; q < SV_RC; q++) acc[q] = 0.0;
#pragma unroll 4
        for (int c = 0; c < r; c++) {
            const double gv = (double)G[row + (size_t)c * NB];
#pragma unroll
            for (int q = 0; q < SV_RC; q++) acc[q] = fma(gv, w[c * SV_RC + q], acc[q]);
        }
        for (int q = 0; q < nq; q++) Po[row + (size_t)q * ldp] -= acc[q];
    }
}

// The update by the solved block k of one block of the right-hand sides in
// Pupd, x_k read from Psol:
// forward:  P_m -= L_mk x_k for m = k + 1 + tile, tile < ntile;
// backward: P_j -= L_kj^T x_k for j = tile, tile < ntile = k;
// the ndense tiles inside the band (the first ones forward, the last ones
// backward) take one workgroup per 64-row block (rb), each low-rank tile one
// workgroup: w = F^T x_k (F = V forward, U backward), then the rows of G w
// (G = U forward, V backward).  Grid: ndense nb + (ntile - ndense).
// LDS: x_k (NB x SV_RC), w (cap x SV_RC), partial sums (16 x 64 x SV_RC).
__global__ void __launch_bounds__(SU_T) k_solve_update(TLR t, int k, const double *Psol, double *Pupd, long ldp,
                                                       int nrhs, int trans, int ndense)
{
    extern __shared__ double sm[];
    const int NB = t.NB, nb = (NB + SV_BK - 1) / SV_BK;
    int tile, rb = -1;
    if ((int)blockIdx.x < ndense * nb) {
        const int d = blockIdx.x / nb;
        rb = blockIdx.x - d * nb;
        tile = trans ? k - ndense + d : d;
    } else {
        const int e = blockIdx.x - ndense * nb;
        tile = trans ? e : ndense + e;
    }
    double *xk = sm;                            // xk[l * SV_RC + q]
    double *w = xk + (size_t)NB * SV_RC;        // w[c * SV_RC + q]
    double *pt = w + (size_t)t.cap * SV_RC;     // pt[(g * 64 + i) * SV_RC + q]
    const int tid = threadIdx.x, i = tid & 63, g = tid >> 6;
    int m, n, o;
    if (!trans) { m = k + 1 + tile; n = k; o = m; }
    else { m = k; n = tile; o = n; }
    const int r = (rb >= 0) ? 0 : *rank_at(t, m, n);
    if (rb < 0 && r <= 0) return;
    const double *T = (rb >= 0) ? band_tile(t, m, n) : lr_tile(t, m, n);
    const int p = (rb >= 0) ? LR_F64 : lr_prec(t, m, n);
    const double *Pk = Psol + (size_t)k * NB;
    double *Po = Pupd + (size_t)o * NB;
    const int row = SV_BK * rb + i;
    for (int r0 = 0; r0 < nrhs; r0 += SV_RC) {
        const int nq = min(SV_RC, nrhs - r0);
        for (int e = tid; e < NB * SV_RC; e += SU_T) {
            const int q = e / NB, l = e - q * NB;
            xk[l * SV_RC + q] = (q < nq) ? Pk[l + (size_t)(r0 + q) * ldp] : 0.0;
        }
        __syncthreads();
        if (rb < 0) {
            const size_t off = (size_t)NB * r;
            const int fo = trans ? 0 : 1;
            double *Pq = Po + (size_t)r0 * ldp;
            if (p == LR_F64) su_lr_w(T + fo * off, NB, r, xk, w);
            else if (p == LR_F32) su_lr_w((const float *)T + fo * off, NB, r, xk, w);
            else su_lr_w((const lr_bf16 *)T + fo * off, NB, r, xk, w);
            __syncthreads();
            if (p == LR_F64) su_lr_rows(T + (1 - fo) * off, NB, r, w, Pq, ldp, nq);
            else if (p == LR_F32) su_lr_rows((const float *)T + (1 - fo) * off, NB, r, w, Pq, ldp, nq);
            else su_lr_rows((const lr_bf16 *)T + (1 - fo) * off, NB, r, w, Pq, ldp, nq);
            __syncthreads();
            continue;
        }
        if (!trans) {
            // P_m(row) -= sum_l L_mk(row, l) x_k(l): thread (i, g), l = g + 16u
            double acc[SV_RC];
#pragma unroll
            for (int q = 0; q < SV_RC; q++) acc[q] = 0.0;
            if (row < NB) {
                const double *Tr = T + row;
#pragma unroll 8
                for (int l = g; l < NB; l += SU_G) {
                    const double lv = Tr[(size_t)l * NB];
#pragma unroll
                    for (int q = 0; q < SV_RC; q++) acc[q] = fma(lv, xk[l * SV_RC + q], acc[q]);
                }
            }
#pragma unroll
            for (int q = 0; q < SV_RC; q++) pt[(g * SV_BK + i) * SV_RC + q] = acc[q];
        } else {
            // P_j(l) -= sum_i L_kj(i, l) x_k(i) for the 64 columns l of the row
            // block: wave g takes l = 64 rb + g + 16u, its lanes run down i
            for (int u = 0; u < SV_BK / SU_G; u++) {
                const int l = SV_BK * rb + g + SU_G * u;
                double a2[SV_RC];
#pragma unroll
                for (int q = 0; q < SV_RC; q++) a2[q] = 0.0;
                if (l < NB) {
                    const double *Tl = T + (size_t)l * NB;
#pragma unroll 4
                    for (int ii = i; ii < NB; ii += 64) {
                        const double lv = Tl[ii];
#pragma unroll
                        for (int q = 0; q < SV_RC; q++) a2[q] = fma(lv, xk[ii * SV_RC + q], a2[q]);
                    }
                }
#pragma unroll
                for (int q = 0; q < SV_RC; q++) {
                    const double s = wsum(a2[q]);
                    if (i == 0) pt[(g + SU_G * u) * SV_RC + q] = s;
                }
            }
        }
        __syncthreads();
        if (g == 0 && row < NB) {
#pragma unroll
            for (int q = 0; q < SV_RC; q++) {
                double s = pt[i * SV_RC + q];
                if (!trans)
                    for (int gg = 1; gg < SU_G; gg++) s += pt[(gg * SV_BK + i) * SV_RC + q];
                if (q < nq) Po[row + (size_t)(r0 + q) * ldp] -= s;
            }
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

// ---------------------------------------------------------------------------
// host side

// doubles of workspace a solve needs: the panels of every L_kk^-1, two
// padded right-hand-side blocks (a pass reads one and writes the other),
// reduction parts (of the sum of squares, then of the log-det), three scalars
size_t solve_ws_doubles(const TLR &t, int nrhs)
{
    const int nbk = (t.NB + SV_BK - 1) / SV_BK;
    const size_t Np = (size_t)t.NT * t.NB;
    return (size_t)t.NT * inv_tile(nbk) + 2 * Np * (size_t)nrhs + std::max((Np + 255) / 256, (size_t)t.NT) + 8;
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
    return sizeof(double) * (nbk * SV_BK * SV_RC + (size_t)SU_G * SV_BK * SV_RC);
}
static size_t update_lds(const TLR &t)
{
    return sizeof(double) * (((size_t)t.NB + (size_t)t.cap) * SV_RC + (size_t)SU_G * SV_BK * SV_RC);
}

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
    double *Xinv = ws;
    double *P0 = Xinv + (size_t)NT * inv_tile(nbk);
    double *P1 = P0 + (size_t)Np * nrhs;
    double *parts = P1 + (size_t)Np * nrhs;
    const size_t ld1 = diag_lds(t), ld2 = update_lds(t);
    HIPCHK(hipFuncSetAttribute((const void *)k_solve_diag, hipFuncAttributeMaxDynamicSharedMemorySize, (int)ld1));
    HIPCHK(hipFuncSetAttribute((const void *)k_solve_update, hipFuncAttributeMaxDynamicSharedMemorySize, (int)ld2));
    hipLaunchKernelGGL(k_tri_inv, dim3((unsigned)(NT * nbk)), dim3(64), 0, st, t, Xinv);
    if (nbk > 1) hipLaunchKernelGGL(k_tile_inv, dim3((unsigned)(NT * nbk)), dim3(256), 0, st, t, Xinv);
    const long ne = Np * nrhs;
    hipLaunchKernelGGL(k_pad_copy, dim3((unsigned)((ne + 255) / 256)), dim3(256), 0, st, (const double *)B, ldb, P0, Np,
                       t.N, (int)Np, nrhs);
    // each pass reads the right-hand sides from src, leaves its solution in dst
    double *src = P0, *dst = P1;
    auto pass = [&](int trans) {
        for (int s = 0; s < NT; s++) {
            const int k = trans ? NT - 1 - s : s;
            hipLaunchKernelGGL(k_solve_diag, dim3((unsigned)nbk), dim3(SU_T), ld1, st, t, k, (const double *)Xinv,
                               (const double *)src, dst, Np, nrhs, trans);
            const int ntile = trans ? k : NT - 1 - k;
            const int ndense = std::min(t.band_size - 1, ntile);
            if (ntile > 0)
                hipLaunchKernelGGL(k_solve_update, dim3((unsigned)(ndense * nbk + ntile - ndense)), dim3(SU_T), ld2,
                                   st, t, k, (const double *)dst, src, Np, nrhs, trans, ndense);
        }
        std::swap(src, dst);
    };
    if (dirs & 1) {
        pass(0);
        if (d_quad) {
            const int nb = (int)((t.N + 255) / 256);
            hipLaunchKernelGGL(k_sumsq_parts, dim3(nb), dim3(256), 0, st, (const double *)src, t.N, parts);
            hipLaunchKernelGGL(k_sum_seq, dim3(1), dim3(256), 0, st, (const double *)parts, nb, d_quad);
        }
    }
    if (dirs & 2) pass(1);
    if (d_logdet) logdet_launch(t, parts, d_logdet, st);
    hipLaunchKernelGGL(k_pad_copy, dim3((unsigned)(((long)t.N * nrhs + 255) / 256)), dim3(256), 0, st,
                       (const double *)src, Np, B, ldb, (int)Np, t.N, nrhs);
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
