// mfma.hpp -- FP64 matrix-core GEMM tile for gfx950 (V_MFMA_F64_16X16X4_F64).
//
// One workgroup of 256 threads (4 waves) computes a 64 x 64 tile of
// C = A B, A is M x K, B is K x N.  Wave w owns rows 32 (w & 1) .. +32 and
// columns 32 (w >> 1) .. +32, i.e. 2 x 2 blocks of 16 x 16.  K runs in chunks
// of MK_KC staged through LDS; the next chunk is loaded into registers while
// the current one is multiplied.
//
// Operand layout of the instruction on a wave: A[i][k] on lane i + 16 k,
// B[k][j] on lane j + 16 k, result C[i][j] on lane j + 16 (i & 3), register
// i >> 2 (each 16 x 16 block holds four doubles per lane).
#pragma once
#include <hip/hip_runtime.h>

#define MK_T 64          // tile edge
#define MK_KC 16         // K chunk
#define MK_LD (MK_T + 1) // LDS row stride (k-major rows of 64 doubles)
#define MK_SMEM (2 * MK_KC * MK_LD)

typedef double mk_d4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ mk_d4 mk_mfma(double a, double b, mk_d4 c)
{
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
}

// Operand addressing.  A element (i, k): AK ? A[k + i lda] : A[i + k lda]
// (AK = "k contiguous", i.e. A given transposed).  B element (k, j):
// BK ? B[k + j ldb] : B[j + k ldb].  TE: the element type in memory, widened
// to double on load (double, float, or a type that converts to double).
template <bool AK, bool BK, typename TE = double>
struct MkOps {
    const TE *A;
    long lda;
    const TE *B;
    long ldb;
};

// Accumulate acc += A(i0:i0+64, k0:k1) B(k0:k1, j0:j0+64), rows >= M and
// columns >= N read as zero.  TRI: both operands are lower triangular in
// their (row, k) / (column, k) indices, A(i, k) = 0 for k > i and
// B(k, j) = 0 for k > j, whatever memory holds there.  All 256 threads must
// call it uniformly.
template <bool AK, bool BK, bool TRI = false, typename TE = double>
__device__ __forceinline__ void mk_accum(const MkOps<AK, BK, TE> &op, int M, int N, int i0, int j0, long k0, long k1,
                                         mk_d4 (&acc)[2][2], double *sm)
{
    double *As = sm, *Bs = sm + MK_KC * MK_LD;
    const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
    const int wi = 32 * (w & 1), wj = 32 * (w >> 1);
    const int i16 = lane & 15, k4 = lane >> 4;
    // each thread stages 4 elements of A and 4 of B per chunk
    double pa[4], pb[4];
    auto load = [&](long kc) {
#pragma unroll
        for (int q = 0; q < 4; q++) {
            const int e = tid + 256 * q;
            int ii, kk;
            if (AK) { kk = e & (MK_KC - 1); ii = e >> 4; }
            else { ii = e & (MK_T - 1); kk = e >> 6; }
            const long gk = kc + kk;
            const int gi = i0 + ii;
            pa[q] = (gi < M && gk < k1 && (!TRI || gk <= gi))
                        ? (double)(AK ? op.A[gk + (long)gi * op.lda] : op.A[gi + gk * op.lda])
                        : 0.0;
            int jj;
            if (BK) { kk = e & (MK_KC - 1); jj = e >> 4; }
            else { jj = e & (MK_T - 1); kk = e >> 6; }
            const long gk2 = kc + kk;
            const int gj = j0 + jj;
            pb[q] = (gj < N && gk2 < k1 && (!TRI || gk2 <= gj))
                        ? (double)(BK ? op.B[gk2 + (long)gj * op.ldb] : op.B[gj + gk2 * op.ldb])
                        : 0.0;
        }
    };
    if (k0 < k1) load(k0);
    for (long kc = k0; kc < k1; kc += MK_KC) {
#pragma unroll
        for (int q = 0; q < 4; q++) {
            const int e = tid + 256 * q;
            int ii, kk, jj;
            if (AK) { kk = e & (MK_KC - 1); ii = e >> 4; }
            else { ii = e & (MK_T - 1); kk = e >> 6; }
            As[kk * MK_LD + ii] = pa[q];
            if (BK) { kk = e & (MK_KC - 1); jj = e >> 4; }
            else { jj = e & (MK_T - 1); kk = e >> 6; }
            Bs[kk * MK_LD + jj] = pb[q];
        }
        __syncthreads();
        if (kc + MK_KC < k1) load(kc + MK_KC);
#pragma unroll
        for (int ks = 0; ks < MK_KC; ks += 4) {
            const double a0 = As[(ks + k4) * MK_LD + wi + i16];
            const double a1 = As[(ks + k4) * MK_LD + wi + 16 + i16];
            const double b0 = Bs[(ks + k4) * MK_LD + wj + i16];
            const double b1 = Bs[(ks + k4) * MK_LD + wj + 16 + i16];
            acc[0][0] = mk_mfma(a0, b0, acc[0][0]);
            acc[0][1] = mk_mfma(a0, b1, acc[0][1]);
            acc[1][0] = mk_mfma(a1, b0, acc[1][0]);
            acc[1][1] = mk_mfma(a1, b1, acc[1][1]);
        }
        __syncthreads();
    }
}

__device__ __forceinline__ void mk_zero(mk_d4 (&acc)[2][2])
{
#pragma unroll
    for (int a = 0; a < 2; a++)
#pragma unroll
        for (int b = 0; b < 2; b++) acc[a][b] = mk_d4{0.0, 0.0, 0.0, 0.0};
}

// Visit the tile's results: f(i, j, v) with (i, j) relative to the tile
// origin, for the 16 values this lane holds.
template <typename F>
__device__ __forceinline__ void mk_store(const mk_d4 (&acc)[2][2], F f)
{
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int wi = 32 * (w & 1), wj = 32 * (w >> 1);
    const int i16 = lane & 15, k4 = lane >> 4;
#pragma unroll
    for (int a = 0; a < 2; a++)
#pragma unroll
        for (int b = 0; b < 2; b++)
#pragma unroll
            for (int r = 0; r < 4; r++) f(wi + 16 * a + k4 + 4 * r, wj + 16 * b + i16, acc[a][b][r]);
}
