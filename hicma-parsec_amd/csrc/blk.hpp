// blk.hpp -- workgroup-level FP64 building blocks for gfx950 (wave64).
//
// Every routine here is executed by ALL threads of one workgroup of BS
// threads and must be reached uniformly (they contain __syncthreads).
// Matrices are column major.  LDS scratch is passed in explicitly so a
// kernel can overlay the scratch of successive phases.
#pragma once
#include "mfma.hpp"
#include <hip/hip_runtime.h>
#include <stdint.h>
#include "wave.hpp"

#ifndef BS
#define BS 256
#endif
#define NWAVE (BS / 64)
#define QRB 16   // Householder panel width
#define GT 64    // gemm output tile edge
#define GK 16    // gemm k-chunk
#define GLD (GT + 1)

// LDS scratch (doubles) that blk_gemm needs
#define BLK_GEMM_SMEM (2 * GK * GLD)
// sweep limit of the one-sided Jacobi SVDs; reaching it is reported as
// not converged
#define BLK_JACOBI_SWEEPS 60
// LDS scratch for reductions: NWAVE * QRB + QRB
#define BLK_RED_SMEM (NWAVE * QRB + 2 * QRB)

__device__ __forceinline__ double wave_sum(double v) { return wsum(v); }

// sum over aligned groups of g lanes, g a power of two <= 64 known at run time
__device__ __forceinline__ double gsum_rt(double v, int g)
{
    switch (g) {
    case 2: return gsum<2>(v);
    case 4: return gsum<4>(v);
    case 8: return gsum<8>(v);
    case 16: return gsum<16>(v);
    case 32: return gsum<32>(v);
    case 64: return gsum<64>(v);
    default: return v;
    }
}

// Jacobi rotation (c, s) that zeroes the inner product of two columns with
// squared norms al, be and inner product ga: t = tan(theta) with |t| <= 1.
// Returns false when no rotation is needed or the values underflow.
__device__ __forceinline__ bool jacobi_rot(double al, double be, double ga, double tol, double &c, double &s)
{
    if (ga == 0.0 || !(ga * ga > (tol * tol) * (al * be))) return false;  // |ga| > tol sqrt(al be)
    const double diff = be - al;
    const double den = fabs(diff) + sqrt(diff * diff + 4.0 * ga * ga);
    if (!(den > 0.0)) return false;
    // t = sign(zeta) / (|zeta| + sqrt(1 + zeta^2)), zeta = diff / (2 ga)
    const double t = (diff >= 0.0 ? 2.0 * ga : -2.0 * ga) / den;
    c = rsqrt(1.0 + t * t);
    s = c * t;
    return true;
}

// block-wide sum, same result (fixed order) on every thread
__device__ __forceinline__ double blk_sum(double v, double *red)
{
    v = wave_sum(v);
    const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
    __syncthreads();
    if (l == 0) red[w] = v;
    __syncthreads();
    double s = 0;
#pragma unroll
    for (int i = 0; i < NWAVE; i++) s += red[i];
    return s;
}

// block-wide sums of p[0..n-1] (n <= QRB) into out[0..n-1] (LDS)
__device__ __forceinline__ void blk_vsum(const double *p, int n, double *red, double *out)
{
    const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
    __syncthreads();
#pragma unroll
    for (int c = 0; c < QRB; c++) {
        if (c < n) {
            double v = wave_sum(p[c]);
            if (l == 0) red[w * QRB + c] = v;
        }
    }
    __syncthreads();
    if ((int)threadIdx.x < n) {
        double s = 0;
        for (int i = 0; i < NWAVE; i++) s += red[i * QRB + threadIdx.x];
        out[threadIdx.x] = s;
    }
    __syncthreads();
}

// One GT x GT output tile of C = alpha*op(A)*op(B) + beta*C at (i0, j0) on
// the FP64 matrix cores (mk_accum: 4 waves, 2 x 2 blocks of 16 x 16 each,
// K in LDS-staged chunks).  The tile is computed as C^T = op(B)^T op(A)^T:
// the instruction's result layout then puts consecutive rows of C on
// consecutive lanes, so the column-major read-modify-write of C coalesces.
// lower_only: skip elements strictly above the diagonal of C.  All BS
// threads call it.
template <bool TA, bool TB>
__device__ void gemm_tile(int i0, int j0, int M, int N, int K, double alpha,
                          const double *__restrict__ A, int lda,
                          const double *__restrict__ B, int ldb, double beta,
                          double *C, int ldc, double *sm, bool lower_only = false)
{
    static_assert(GT == MK_T && BLK_GEMM_SMEM >= MK_SMEM && BS == 256, "gemm tile shape");
    mk_d4 acc[2][2];
    mk_zero(acc);
    // operand 1 (j, k) = op(B)(k, j), operand 2 (k, i) = op(A)(i, k)
    const MkOps<!TB, TA> op{B, ldb, A, lda};
    mk_accum(op, N, M, j0, i0, 0, K, acc, sm);
    mk_store(acc, [&](int jj, int ii, double v) {
        const int gi = i0 + ii, gj = j0 + jj;
        if (gi >= M || gj >= N || (lower_only && gi < gj)) return;
        double *c = C + gi + (size_t)gj * ldc;
        *c = (beta == 0.0) ? alpha * v : fma(alpha, v, beta * (*c));
    });
}

// whole GEMM by one workgroup
template <bool TA, bool TB>
__device__ void blk_gemm(int M, int N, int K, double alpha, const double *A, int lda,
                         const double *B, int ldb, double beta, double *C, int ldc, double *sm)
{
    for (int j0 = 0; j0 < N; j0 += GT)
        for (int i0 = 0; i0 < M; i0 += GT)
            gemm_tile<TA, TB>(i0, j0, M, N, K, alpha, A, lda, B, ldb, beta, C, ldc, sm);
    __syncthreads();
}

// copy a column-major block (rows x cols) src -> dst, optional scale
__device__ __forceinline__ void blk_copy(int rows, int cols, const double *src, int lds,
                                         double *dst, int ldd, double scale = 1.0)
{
    for (int j = 0; j < cols; j++)
        for (int i = threadIdx.x; i < rows; i += BS)
            dst[i + (size_t)j * ldd] = scale * src[i + (size_t)j * lds];
}

// explicit unit-lower V panel (rows m x jb) from the geqrf storage at A(j0, j0)
__device__ void blk_build_v(int m, int jb, const double *Ap, int lda, double *Vb, int ldv)
{
    for (int c = 0; c < jb; c++)
        for (int i = threadIdx.x; i < m; i += BS) {
            double v = (i < c) ? 0.0 : (i == c ? 1.0 : Ap[i + (size_t)c * lda]);
            Vb[i + (size_t)c * ldv] = v;
        }
    __syncthreads();
}

// T (jb x jb upper, ld QRB) of the compact WY form H = I - V T V^T, from
// G = V^T V (ld QRB) and tau.  Sequential over columns, parallel over rows.
__device__ void blk_build_t(int jb, const double *G, const double *tau, double *T)
{
    const int tid = threadIdx.x;
    for (int i = tid; i < QRB * QRB; i += BS) T[i] = 0.0;
    __syncthreads();
    for (int j = 0; j < jb; j++) {
        double tj = tau[j];
        double v = 0.0;
        if (tid < j) {
            for (int l = tid; l < j; l++) v += T[tid + l * QRB] * G[l + j * QRB];
        }
        __syncthreads();
        if (tid < j) T[tid + j * QRB] = -tj * v;
        if (tid == 0) T[j + j * QRB] = tj;
        __syncthreads();
    }
}

// W(jb x nc, ld QRB) <- op(T) W, T upper (trans: T^T lower)
__device__ void blk_tmul(int jb, int nc, const double *T, bool trans, double *W, double *sm)
{
    // process columns in chunks held in LDS
    for (int c0 = 0; c0 < nc; c0 += BS / QRB) {
        int c = c0 + threadIdx.x / QRB, i = threadIdx.x % QRB;
        double v = 0.0;
        if (c < nc && i < jb) {
            if (!trans) {
                for (int l = i; l < jb; l++) v += T[i + l * QRB] * W[l + (size_t)c * QRB];
            } else {
                for (int l = 0; l <= i; l++) v += T[l + i * QRB] * W[l + (size_t)c * QRB];
            }
        }
        __syncthreads();
        if (c < nc && i < jb) W[i + (size_t)c * QRB] = v;
        __syncthreads();
    }
}

// Householder QR (LAPACK dgeqrf semantics: R on/above diagonal, v below with
// implicit unit, tau) of A (m x n), blocked by QRB with compact-WY trailing
// updates.  Tsave receives the T factor of each panel (QRB*QRB each).
// ws: global scratch of m*QRB (Vb) + QRB*n (W) doubles.  sm: LDS scratch of
// at least BLK_GEMM_SMEM + BLK_RED_SMEM + 2*QRB*QRB doubles.
__device__ void blk_geqrf(int m, int n, double *A, int lda, double *tau, double *Tsave,
                          double *ws, double *sm)
{
    const int tid = threadIdx.x;
    double *red = sm + BLK_GEMM_SMEM;
    double *vs = red + NWAVE * QRB;          // QRB reduced values
    double *G = red + BLK_RED_SMEM;          // QRB*QRB
    double *T = G + QRB * QRB;               // QRB*QRB
    double *Vb = ws;
    double *W = ws + (size_t)m * QRB;
    const int kmax = m < n ? m : n;
    for (int j0 = 0; j0 < kmax; j0 += QRB) {
        const int jb = (kmax - j0) < QRB ? (kmax - j0) : QRB;
        // ---- panel factorization, column by column ----
        for (int jj = 0; jj < jb; jj++) {
            const int j = j0 + jj;
            double *col = A + (size_t)j * lda;
            double part = 0.0;
            for (int i = j + 1 + tid; i < m; i += BS) part += col[i] * col[i];
            double xn2 = blk_sum(part, red);
            double alpha = col[j];
            double tj = 0.0, beta = alpha, scal = 0.0;
            if (xn2 > 0.0) {
                beta = -copysign(sqrt(alpha * alpha + xn2), alpha);
                tj = (beta - alpha) / beta;
                scal = 1.0 / (alpha - beta);
                for (int i = j + 1 + tid; i < m; i += BS) col[i] *= scal;
            }
            __syncthreads();
            if (tid == 0) { col[j] = beta; tau[j] = tj; }
            const int nc = j0 + jb - 1 - j;  // remaining panel columns
            if (tj != 0.0 && nc > 0) {
                double p[QRB];
#pragma unroll
                for (int c = 0; c < QRB; c++) {
                    p[c] = 0.0;
                    if (c < nc) {
                        const double *ac = A + (size_t)(j + 1 + c) * lda;
                        for (int i = j + 1 + tid; i < m; i += BS) p[c] += col[i] * ac[i];
                    }
                }
                blk_vsum(p, nc, red, vs);
                double w[QRB];
#pragma unroll
                for (int c = 0; c < QRB; c++)
                    w[c] = (c < nc) ? tj * (vs[c] + A[j + (size_t)(j + 1 + c) * lda]) : 0.0;
                __syncthreads();
#pragma unroll
                for (int c = 0; c < QRB; c++) {
                    if (c < nc) {
                        double *ac = A + (size_t)(j + 1 + c) * lda;
                        for (int i = j + 1 + tid; i < m; i += BS) ac[i] -= w[c] * col[i];
                        if (tid == 0) ac[j] -= w[c];
                    }
                }
            }
            __syncthreads();
        }
        // ---- T factor of the panel ----
        const int mr = m - j0;
        blk_build_v(mr, jb, A + j0 + (size_t)j0 * lda, lda, Vb, m);
        // G = Vb^T Vb
        for (int e = tid; e < QRB * QRB; e += BS) G[e] = 0.0;
        __syncthreads();
        {
            // jb*jb dot products of length mr
            for (int a = 0; a < jb; a++) {
                double p[QRB];
#pragma unroll
                for (int b = 0; b < QRB; b++) {
                    p[b] = 0.0;
                    if (b < jb && b > a) {
                        for (int i = tid; i < mr; i += BS)
                            p[b] += Vb[i + (size_t)a * m] * Vb[i + (size_t)b * m];
                    }
                }
                blk_vsum(p, jb, red, vs);
                if (tid < jb && tid > a) G[a + tid * QRB] = vs[tid];
                __syncthreads();
            }
        }
        blk_build_t(jb, G, tau + j0, T);
        if (Tsave) {
            for (int e = tid; e < QRB * QRB; e += BS) Tsave[(size_t)(j0 / QRB) * QRB * QRB + e] = T[e];
        }
        __syncthreads();
        // ---- trailing update: A2 <- (I - V T^T V^T) A2 ----
        const int nt = n - (j0 + jb);
        if (nt > 0) {
            double *A2 = A + j0 + (size_t)(j0 + jb) * lda;
            blk_gemm<true, false>(jb, nt, mr, 1.0, Vb, m, A2, lda, 0.0, W, QRB, sm);
            blk_tmul(jb, nt, T, true, W, sm);
            blk_gemm<false, false>(mr, nt, jb, -1.0, Vb, m, W, QRB, 1.0, A2, lda, sm);
        }
    }
}

// C (m x nc) <- Q C with Q = H_0 ... H_{k-1} from blk_geqrf (A, Tsave).
__device__ void blk_ormqr(int m, int nc, int k, const double *A, int lda, const double *Tsave,
                          double *C, int ldc, double *ws, double *sm)
{
    const int tid = threadIdx.x;
    double *T = sm + BLK_GEMM_SMEM + BLK_RED_SMEM + QRB * QRB;
    double *Vb = ws;
    double *W = ws + (size_t)m * QRB;
    if (nc <= 0 || k <= 0) return;
    int np = (k + QRB - 1) / QRB;
    for (int p = np - 1; p >= 0; p--) {
        const int j0 = p * QRB;
        const int jb = (k - j0) < QRB ? (k - j0) : QRB;
        const int mr = m - j0;
        blk_build_v(mr, jb, A + j0 + (size_t)j0 * lda, lda, Vb, m);
        for (int e = tid; e < QRB * QRB; e += BS) T[e] = Tsave[(size_t)p * QRB * QRB + e];
        __syncthreads();
        double *C2 = C + j0;
        blk_gemm<true, false>(jb, nc, mr, 1.0, Vb, m, C2, ldc, 0.0, W, QRB, sm);
        blk_tmul(jb, nc, T, false, W, sm);
        blk_gemm<false, false>(mr, nc, jb, -1.0, Vb, m, W, QRB, 1.0, C2, ldc, sm);
    }
}

// One-sided (Hestenes) Jacobi on the n columns of M (n rows, ld ldm),
// accumulating the right rotations in Vs (n x n, ld ldv).  On return the
// columns of M are u_j*sigma_j, sigma[j] = ||M(:,j)||.  M and Vs may live in
// LDS or global memory.  Round-robin pair ordering; fixed reduction order.
// Returns the number of sweeps (BLK_JACOBI_SWEEPS: not converged).
__device__ int blk_jacobi(int n, double *M, int ldm, double *Vs, int ldv, double *sigma,
                          int *flag, double *red)
{
    const int tid = threadIdx.x;
    for (int j = 0; j < n; j++)
        for (int i = tid; i < n; i += BS) Vs[i + (size_t)j * ldv] = (i == j) ? 1.0 : 0.0;
    __syncthreads();
    if (n <= 1) {
        double p = 0.0;
        for (int i = tid; i < n; i += BS) p += M[i] * M[i];
        double s = blk_sum(p, red);
        if (tid == 0 && n == 1) sigma[0] = sqrt(s);
        __syncthreads();
        return 0;
    }
    const int npad = n + (n & 1);
    const int npairs = npad / 2;
    int g = 64;
    while (g > 1 && g * npairs > BS) g >>= 1;
    const int gpb = BS / g;  // groups (pairs) handled concurrently
    const double tol = 4.0e-16 * sqrt((double)n);
    int sweep = 0;
    for (; sweep < BLK_JACOBI_SWEEPS; sweep++) {
        if (tid == 0) *flag = 0;
        __syncthreads();
        int rotated = 0;
        for (int r = 0; r < npad - 1; r++) {
            for (int pb = 0; pb < npairs; pb += gpb) {
                const int k = pb + tid / g, lane = tid % g;
                int a = -1, b = -1;
                if (k < npairs) {
                    int ia = k, ib = npad - 1 - k;
                    a = (ia == 0) ? 0 : ((ia - 1 + r) % (npad - 1)) + 1;
                    b = (ib == 0) ? 0 : ((ib - 1 + r) % (npad - 1)) + 1;
                    if (a >= n || b >= n) { a = -1; }
                }
                double al = 0, be = 0, ga = 0;
                if (a >= 0) {
                    const double *ca = M + (size_t)a * ldm, *cb = M + (size_t)b * ldm;
                    for (int i = lane; i < n; i += g) {
                        double x = ca[i], y = cb[i];
                        al = fma(x, x, al); be = fma(y, y, be); ga = fma(x, y, ga);
                    }
                }
                al = gsum_rt(al, g);
                be = gsum_rt(be, g);
                ga = gsum_rt(ga, g);
                double c, s;
                if (a >= 0 && jacobi_rot(al, be, ga, tol, c, s)) {
                    double *ca = M + (size_t)a * ldm, *cb = M + (size_t)b * ldm;
                    for (int i = lane; i < n; i += g) {
                        double x = ca[i], y = cb[i];
                        ca[i] = c * x - s * y;
                        cb[i] = s * x + c * y;
                    }
                    double *va = Vs + (size_t)a * ldv, *vb = Vs + (size_t)b * ldv;
                    for (int i = lane; i < n; i += g) {
                        double x = va[i], y = vb[i];
                        va[i] = c * x - s * y;
                        vb[i] = s * x + c * y;
                    }
                    rotated = 1;
                }
                __syncthreads();
            }
        }
        if (rotated) *flag = 1;  // benign same-value race
        __syncthreads();
        int any = *flag;
        __syncthreads();
        if (!any) break;
    }
    for (int j = 0; j < n; j++) {
        double p = 0.0;
        for (int i = tid; i < n; i += BS) p += M[i + (size_t)j * ldm] * M[i + (size_t)j * ldm];
        double s = blk_sum(p, red);
        if (tid == 0) sigma[j] = sqrt(s);
    }
    __syncthreads();
    return sweep;
}

// pos[j] = rank of sigma[j] in descending order (ties by index)
__device__ void blk_sort_pos(int n, const double *sigma, int *pos)
{
    for (int j = threadIdx.x; j < n; j += BS) {
        double sj = sigma[j];
        int p = 0;
        for (int i = 0; i < n; i++) {
            double si = sigma[i];
            p += (si > sj) || (si == sj && i < j);
        }
        pos[j] = p;
    }
    __syncthreads();
}

