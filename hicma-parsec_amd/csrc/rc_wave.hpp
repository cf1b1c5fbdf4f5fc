// rc_wave.hpp -- recompression of a low-rank gemm update by ONE wave.
//
// A task is the update C_mn -= A_mk A_nk^T of three low-rank tiles (the
// reference's hcore_dgemm QR + SVD recompression, hicma_parsec_core.c).  The
// wave owns its task from start to end: every reduction over the NB tile rows
// is a wave reduction (permlane / DPP / readlane), LDS is private to the wave
// and no barrier is shared with other waves, so a CU works on eight tasks at
// once and a step with thousands of tasks fills all CUs.
//
// Mathematics (the reference's up to rounding):
//   X = Av^T Bv,  Qn = -Bu X^T,  P = [Cu Au],  Q = [Cv Qn]        (NB x s)
//   Rp, Rq = R factors of P, Q (Householder, panels of WV_PW columns)
//   M = Rp Rq^T (s x s);  one-sided Jacobi on the rows of M:  M^T J = Vs S
//   -> right singular vectors Vs and singular values S of M;
//   r' = #{S >= acc} (or the fixed rank), capped at maxrank;
//   U' = P (Rq^T Vs_r)   ( = Qp Us_r S_r because P = Qp Rp, M Vs = Us S )
//   V' = Q (Rq^-1 Vs_r)  ( = Qq Vs_r     because Q = Qq Rq )
// so neither orthogonal factor is ever formed.
//
// Layout: lane l holds rows l, l+64, ... of every tall (NB-row) operand, so
// tall data is read and written coalesced and row-locally.
#pragma once
#include "tlr_kernels.hpp"

#define WV_SMAX 64   // largest stack width s = rc + ra, and largest rb
#define WV_KMAX 16   // rows per lane: NB <= 64 * WV_KMAX
#define WV_PW 4      // Householder panel width
#define WV_S2MAX 60  // widest stack of list 2 (k_rc_wave<64>); wider stacks go to the
                     // four-wave k_rc_wide, which balances the two side streams


// Per-variant sizes: SM = largest stack width of the variant (32, 48 or 64).
template <int SM>
struct WvCfg {
    static constexpr int LDJ = SM + 2;             // Jacobi column stride in LDS
    static constexpr int SIG = SM * LDJ;           // offset of the singular values
    static constexpr int LDS = SIG + 2 * SM;       // doubles of LDS per wave
    static constexpr int WPB = SM <= 48 ? 4 : 2;   // waves (tasks in flight) per workgroup
    static constexpr int MINW = 1;                 // one wave per SIMD: the full register file
    static constexpr int BPC = 4 / WPB;            // resident workgroups per CU
};

// per-wave HBM workspace: W (QR copy), Qn, Out (2 blocks), Rp, Rq, Ku
__host__ __device__ inline size_t wv_slot_doubles(int NB)
{
    return (size_t)NB * 4 * WV_SMAX + 4 * (size_t)WV_SMAX * WV_SMAX;
}

__device__ __forceinline__ double rdlane(double v, int l)
{
    return mk64((unsigned)__builtin_amdgcn_readlane((int)lo32(v), l),
                (unsigned)__builtin_amdgcn_readlane((int)hi32(v), l));
}

__device__ __forceinline__ int uni(int v) { return __builtin_amdgcn_readfirstlane(v); }

// FP64 matrix core tile C += A B with A 16x4, B 4x16, C 16x16 (gfx950
// V_MFMA_F64_16X16X4_F64).  Operand layout on the wave: A[i][k] on lane
// i + 16k, B[k][j] on lane j + 16k; result C[i][j] on lane j + 16 (i & 3),
// register i >> 2.
typedef double wv_d4 __attribute__((ext_vector_type(4)));
__device__ __forceinline__ wv_d4 mfma16(double a, double b, wv_d4 c)
{
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
}

// Columns of a tall operand: c0 columns at s0, the rest at s1, column
// stride lda (rows 0 .. NB - 1 of each are used).
struct WvSrc {
    const double *s0;
    int c0;
    const double *s1;
    int lda;
};

// One batch of A rows by buffer loads.  Column group kk (columns 4 kk ..
// 4 kk + 3, one per 16 lanes) is read from the block that holds it; when c0
// is not a multiple of 4 the group that straddles the blocks reads its
// block-0 columns into av and its block-1 columns into ax, and the MFMA
// chain takes both.  A lane outside its block reads 0, so there is neither
// a branch nor a select after a load, and loads are waited for at their use.
template <int SM, int RT>
__device__ __forceinline__ void wv_tall_ld(double (&av)[RT][SM / 4], double (&ax)[RT], int NB, int kd, wv_rsrc r0s,
                                           wv_rsrc r1s, int c0, int lda, int r0, int rstep)
{
    const int lane = threadIdx.x & 63, i16 = lane & 15, k4 = lane >> 4;
    const int cx = 4 * (c0 >> 2) + k4;  // this lane's column of the straddling group
    const bool mix = (c0 & 3) != 0 && cx >= c0 && cx < kd;
#pragma unroll
    for (int t = 0; t < RT; t++) {
        const int row = r0 + rstep * t + i16;
#pragma unroll
        for (int kk = 0; kk < SM / 4; kk++) {
            const int c = 4 * kk + k4;
            const bool b1 = 4 * kk >= c0;  // the whole group is in block 1
            const bool in = row < NB && c < kd && (b1 || c < c0);
            av[t][kk] = wv_ld(b1 ? r1s : r0s, in ? ((b1 ? c - c0 : c) * lda + row) * 8 : WV_OOB);
        }
        ax[t] = wv_ld(r1s, (mix && row < NB) ? ((cx - c0) * lda + row) * 8 : WV_OOB);
    }
}

// This lane's K operands: kv[kk][tn] = K[4 kk + k4][16 tn + i16] and
// kx[tn] = the same for the straddling group's block-1 columns, 0 where
// there is no such row.
template <int SM>
__device__ __forceinline__ void wv_tall_k(double (&kv)[SM / 4][SM / 16], double (&kx)[SM / 16], int kd, int c0,
                                          const double *Kl, int ldk)
{
    const int lane = threadIdx.x & 63, i16 = lane & 15, k4 = lane >> 4;
#pragma unroll
    for (int kk = 0; kk < SM / 4; kk++)
#pragma unroll
        for (int tn = 0; tn < SM / 16; tn++) {
            const int kr = 4 * kk + k4;
            const double v = Kl[(kr < kd ? kr : 0) * ldk + 16 * tn + i16];
            kv[kk][tn] = kr < kd ? v : 0.0;
        }
    const int cx = 4 * (c0 >> 2) + k4;
    const bool mix = (c0 & 3) != 0 && cx >= c0 && cx < kd;
#pragma unroll
    for (int tn = 0; tn < SM / 16; tn++) {
        const double v = Kl[(mix ? cx : 0) * ldk + 16 * tn + i16];
        kx[tn] = mix ? v : 0.0;
    }
}

// D rows of one batch from its loaded A rows and the K operands in
// registers: for each 16-column block of D (a uniform branch on nc), the
// MFMA chain over K runs in chunks of 16 rows (a uniform branch on kd per
// chunk), so the accumulators stay in place between MFMAs.  Stores go
// through a buffer over D's nc columns (those past it are dropped).
template <int SM, int RT>
__device__ __forceinline__ void wv_tall_mul(const double (&av)[RT][SM / 4], const double (&ax)[RT],
                                            const double (&kv)[SM / 4][SM / 16], const double (&kx)[SM / 16], int NB,
                                            int kd, int nc, double alpha, wv_rsrc dr, int r0, int rstep)
{
    const int lane = threadIdx.x & 63, i16 = lane & 15, k4 = lane >> 4;
    constexpr int KS = SM / 4, TN = SM / 16;
#pragma unroll
    for (int tn = 0; tn < TN; tn++) {
        if (16 * tn < nc) {
            wv_d4 acc[RT];
#pragma unroll
            for (int t = 0; t < RT; t++) acc[t] = mfma16(ax[t], kx[tn], wv_d4{0.0, 0.0, 0.0, 0.0});
#pragma unroll
            for (int kc = 0; kc < KS; kc += 4) {
                if (4 * kc < kd) {
#pragma unroll
                    for (int kk = kc; kk < kc + 4; kk++)
#pragma unroll
                        for (int t = 0; t < RT; t++) acc[t] = mfma16(av[t][kk], kv[kk][tn], acc[t]);
                }
            }
            const int j = 16 * tn + i16;
#pragma unroll
            for (int t = 0; t < RT; t++) {
                const int rt = r0 + rstep * t;
#pragma unroll
                for (int r = 0; r < 4; r++) {
                    const int i = rt + k4 + 4 * r;
                    wv_st(alpha * acc[t][r], dr, (j < nc && i < NB) ? (i + j * NB) * 8 : WV_OOB);
                }
            }
        }
    }
}

// The same with K read from LDS, a chunk of 16 K rows at a time (for the
// widest stacks, whose K does not fit in registers).
template <int SM, int RT>
__device__ __forceinline__ void wv_tall_mul_lds(const double (&av)[RT][SM / 4], const double (&ax)[RT], int NB, int kd,
                                                int c0, int nc, double alpha, const double *Kl, int ldk, wv_rsrc dr,
                                                int r0, int rstep)
{
    const int lane = threadIdx.x & 63, i16 = lane & 15, k4 = lane >> 4;
    constexpr int KS = SM / 4, TN = SM / 16;
    const int cx = 4 * (c0 >> 2) + k4;
    const bool mix = (c0 & 3) != 0 && cx >= c0 && cx < kd;
#pragma unroll
    for (int tn = 0; tn < TN; tn++) {
        if (16 * tn < nc) {
            const double kbx = Kl[(mix ? cx : 0) * ldk + 16 * tn + i16];
            wv_d4 acc[RT];
#pragma unroll
            for (int t = 0; t < RT; t++) acc[t] = mfma16(ax[t], mix ? kbx : 0.0, wv_d4{0.0, 0.0, 0.0, 0.0});
#pragma unroll
            for (int kc = 0; kc < KS; kc += 4) {
                if (4 * kc < kd) {
                    double kb[4];
#pragma unroll
                    for (int u = 0; u < 4; u++) {
                        const int kr = 4 * (kc + u) + k4;
                        const double v = Kl[(kr < kd ? kr : 0) * ldk + 16 * tn + i16];
                        kb[u] = kr < kd ? v : 0.0;
                    }
#pragma unroll
                    for (int u = 0; u < 4; u++)
#pragma unroll
                        for (int t = 0; t < RT; t++) acc[t] = mfma16(av[t][kc + u], kb[u], acc[t]);
                }
            }
            const int j = 16 * tn + i16;
#pragma unroll
            for (int t = 0; t < RT; t++) {
                const int rt = r0 + rstep * t;
#pragma unroll
                for (int r = 0; r < 4; r++) {
                    const int i = rt + k4 + 4 * r;
                    wv_st(alpha * acc[t][r], dr, (j < nc && i < NB) ? (i + j * NB) * 8 : WV_OOB);
                }
            }
        }
    }
}

// Tall product D (NB x nc, ld NB) = alpha * A (NB x kd) K (kd x nc) with A's
// columns given by a WvSrc and K in LDS (K[c][j] at Kl[c * ldk + j]).  Work
// goes in batches of RT 16-row tiles (tiles t0, t0 + tstep, ... when several
// waves share the product); a batch's A loads are all issued before its
// first MFMA.  With PP, K is held in registers and batches alternate
// between two register sets: batch i + 1 loads while batch i multiplies and
// stores.  Rows of a batch are read before any of them is written, so D may
// overwrite A's storage row for row.
template <int SM, int RT = (SM <= 32 ? 2 : 1), bool PP = (SM <= 64)>
__device__ __forceinline__ void wv_tall_mm(int NB, int kd, int nc, double alpha, WvSrc A, const double *Kl, int ldk,
                                           double *D, int t0 = 0, int tstep = 1)
{
    const int rstep = 16 * tstep, bstep = rstep * RT;  // rows between tiles, between batches
    int r0 = 16 * t0;
    const wv_rsrc dr = wv_col_rsrc(D, NB * nc * 8);
    const int c0 = A.c0 < kd ? A.c0 : kd;
    const wv_rsrc r0s = wv_col_rsrc(A.s0, c0 * A.lda * 8), r1s = wv_col_rsrc(A.s1, (kd - c0) * A.lda * 8);
    if constexpr (PP) {
        double kv[SM / 4][SM / 16], kx[SM / 16];
        wv_tall_k<SM>(kv, kx, kd, c0, Kl, ldk);
        double a0[RT][SM / 4], x0[RT], a1[RT][SM / 4], x1[RT];
        if (r0 < NB) wv_tall_ld<SM, RT>(a0, x0, NB, kd, r0s, r1s, c0, A.lda, r0, rstep);
#pragma unroll 1
        for (; r0 < NB; r0 += 2 * bstep) {
            if (r0 + bstep < NB) wv_tall_ld<SM, RT>(a1, x1, NB, kd, r0s, r1s, c0, A.lda, r0 + bstep, rstep);
            wv_tall_mul<SM, RT>(a0, x0, kv, kx, NB, kd, nc, alpha, dr, r0, rstep);
            if (r0 + bstep >= NB) break;
            if (r0 + 2 * bstep < NB) wv_tall_ld<SM, RT>(a0, x0, NB, kd, r0s, r1s, c0, A.lda, r0 + 2 * bstep, rstep);
            wv_tall_mul<SM, RT>(a1, x1, kv, kx, NB, kd, nc, alpha, dr, r0 + bstep, rstep);
        }
    } else {
#pragma unroll 1
        for (; r0 < NB; r0 += bstep) {
            double av[RT][SM / 4], ax[RT];
            wv_tall_ld<SM, RT>(av, ax, NB, kd, r0s, r1s, c0, A.lda, r0, rstep);
            wv_tall_mul_lds<SM, RT>(av, ax, NB, kd, c0, nc, alpha, Kl, ldk, dr, r0, rstep);
        }
    }
}

// reduce-scatter of 4 values: lane l returns the wave total of value l >> 4
__device__ __forceinline__ double wrs4(double p0, double p1, double p2, double p3)
{
    double x0 = p0, y0 = p2, x1 = p1, y1 = p3;
    swap32d(x0, y0);  // lanes < 32: values 0, 1; lanes >= 32: values 2, 3
    swap32d(x1, y1);
    double a = x0 + y0, b = x1 + y1;
    swap16d(a, b);    // value index = lane >> 4
    return gsum<16>(a + b);
}

// reduce-scatter of 8 values: lane l returns the wave total of value l >> 3
__device__ __forceinline__ double wrs8(double (&p)[8])
{
    const int lane = threadIdx.x & 63;
    double q[4];
#pragma unroll
    for (int k = 0; k < 4; k++) {
        double x = p[k], y = p[k + 4];
        swap32d(x, y);  // lanes < 32: values 0..3, lanes >= 32: values 4..7
        q[k] = x + y;
    }
    double r[2];
#pragma unroll
    for (int k = 0; k < 2; k++) {
        double x = q[k], y = q[k + 2];
        swap16d(x, y);  // value bit 1 = lane bit 4
        r[k] = x + y;
    }
    const bool up = (lane & 8) != 0;
    const double keep = up ? r[1] : r[0], send = up ? r[0] : r[1];
    return gsum<8>(keep + dppd<DPP_ROR8>(send));  // value bit 0 = lane bit 3
}

// reduce-scatter of N = 4, 8 or 16 values: lane l returns the wave total of
// value l / (64 / N)
template <int N>
__device__ __forceinline__ double wrsN(double (&p)[N])
{
    static_assert(N == 4 || N == 8 || N == 16, "reduce-scatter width");
    if constexpr (N == 4) return wrs4(p[0], p[1], p[2], p[3]);
    else if constexpr (N == 8) return wrs8(p);
    else return wrs16(p);
}

// Squared-norm floor below which a Jacobi pair is not rotated (see k_rc_wave).
__device__ __forceinline__ double wv_skip_floor(double f2, int s, double acc, int fixed_rk)
{
    const double nz = (2.3e-16 * s) * (2.3e-16 * s) * f2;
    const double tz = (fixed_rk > 0 || !(acc > 0.0)) ? 0.0 : (1e-3 * acc) * (1e-3 * acc);
    return fmax(nz, tz);
}

// Buffer resources of the columns of a stack [src0 (c0 columns) | src1] in
// its first pass and of the workspace W afterwards (the fields are plain
// scalars: no indexed access to a structure in private memory).
struct WvRs {
    const double *src0, *src1, *W;
    int c0, NB;
    // column c (no bytes when !valid: loads give 0)
    __device__ __forceinline__ wv_rsrc rd(int c, bool first, bool valid) const
    {
        const double *p = !first ? W + (size_t)c * NB
                                 : (c < c0 ? src0 + (size_t)c * NB : src1 + (size_t)(c - c0) * NB);
        return wv_col_rsrc(valid ? p : W, valid ? NB * 8 : 0);
    }
    // W's column c (no bytes when !valid: stores are dropped)
    __device__ __forceinline__ wv_rsrc wr(int c, bool valid) const
    {
        return wv_col_rsrc(W + (size_t)(valid ? c : 0) * NB, valid ? NB * 8 : 0);
    }
};

// One trailing pair (b0, b1) of a panel: b -= V (T^T (V^T b)), V = the
// panel's reflectors (row block 0 in v0, the rest in a), result to W.
__device__ __forceinline__ void wv_qr_pair(const double (&a)[WV_KMAX][WV_PW], const double (&v0)[WV_PW],
                                           const double (&T)[WV_PW][WV_PW], double (&b0)[WV_KMAX],
                                           double (&b1)[WV_KMAX], wv_rsrc w0, wv_rsrc w1, int lane)
{
    double p[2 * WV_PW];
#pragma unroll
    for (int c = 0; c < WV_PW; c++) {
        double a0 = v0[c] * b0[0], a1 = v0[c] * b1[0];
#pragma unroll
        for (int k = 1; k < WV_KMAX; k++) {
            a0 = fma(a[k][c], b0[k], a0);
            a1 = fma(a[k][c], b1[k], a1);
        }
        p[c] = a0;
        p[c + WV_PW] = a1;
    }
    const double red = wrsN<2 * WV_PW>(p);
    double w0v[WV_PW], w1v[WV_PW];
    {
        double d0[WV_PW], d1[WV_PW];
#pragma unroll
        for (int c = 0; c < WV_PW; c++) {
            d0[c] = rdlane(red, (32 / WV_PW) * c);
            d1[c] = rdlane(red, (32 / WV_PW) * (c + WV_PW));
        }
#pragma unroll
        for (int c = 0; c < WV_PW; c++) {
            double x0 = 0.0, x1 = 0.0;
#pragma unroll
            for (int q = 0; q <= c; q++) {
                x0 = fma(T[q][c], d0[q], x0);
                x1 = fma(T[q][c], d1[q], x1);
            }
            w0v[c] = x0;
            w1v[c] = x1;
        }
    }
    {
        double x0 = b0[0], x1 = b1[0];
#pragma unroll
        for (int c = 0; c < WV_PW; c++) {
            x0 = fma(-v0[c], w0v[c], x0);
            x1 = fma(-v0[c], w1v[c], x1);
        }
        b0[0] = x0;
        b1[0] = x1;
    }
#pragma unroll
    for (int k = 1; k < WV_KMAX; k++) {
        double x0 = b0[k], x1 = b1[k];
#pragma unroll
        for (int c = 0; c < WV_PW; c++) {
            x0 = fma(-a[k][c], w0v[c], x0);
            x1 = fma(-a[k][c], w1v[c], x1);
        }
        b0[k] = x0;
        b1[k] = x1;
    }
    wv_st_col(b0, w0, lane);
    wv_st_col(b1, w1, lane);
}

// R factor of the NB x s stack (s <= SM <= 64, so pivot row j is on lane
// j of the first row block).  Panels of WV_PW columns live in registers,
// trailing columns stream through W in pairs: the pair after the current
// one is loaded while the current one is updated (two register sets used in
// turn, so no load is waited for before its use), and a panel's first pair
// is loaded before the panel is factored.  R (row-major, ld WV_SMAX, zeros
// below the diagonal) is written to Rout.
template <int SM>
__device__ __forceinline__ void wv_geqrf_r(int s, const double *src0, const double *src1, int c0, double *W, int NB,
                                           double *Rout)
{
    const int lane = threadIdx.x & 63;
    const WvRs A{src0, src1, W, c0, NB};
    for (int j0 = 0; j0 < s; j0 += WV_PW) {
        const int jb = min(WV_PW, s - j0);
        const bool first = (j0 == 0);
        double a[WV_KMAX][WV_PW];
#pragma unroll
        for (int c = 0; c < WV_PW; c++) {
            const wv_rsrc r = A.rd(j0 + c, first, c < jb);
#pragma unroll
            for (int k = 0; k < WV_KMAX; k++) a[k][c] = wv_ld(r, (lane + 64 * k) * 8);
        }
        const int t0 = j0 + jb;
        double p0[WV_KMAX], p1[WV_KMAX], q0[WV_KMAX], q1[WV_KMAX];
        wv_ld_col(p0, A.rd(t0, first, t0 < s), lane);
        wv_ld_col(p1, A.rd(t0 + 1, first, t0 + 1 < s), lane);
        double T[WV_PW][WV_PW];
#pragma unroll
        for (int r = 0; r < WV_PW; r++)
#pragma unroll
            for (int c = 0; c < WV_PW; c++) T[r][c] = 0.0;
        // ---- panel factorization ----
#pragma unroll
        for (int jj = 0; jj < WV_PW; jj++) {
            if (jj < jb) {
                const int j = j0 + jj;
                // x = column jj; y = its part strictly below the pivot row j
                const double x0 = a[0][jj];
                const double y0 = (lane > j) ? x0 : 0.0;
                double p[WV_PW];
#pragma unroll
                for (int c = 0; c < WV_PW; c++) p[c] = y0 * (c == jj ? y0 : a[0][c]);
#pragma unroll
                for (int k = 1; k < WV_KMAX; k++)
#pragma unroll
                    for (int c = 0; c < WV_PW; c++) p[c] = fma(a[k][jj], a[k][c], p[c]);
                const double red = wrsN<WV_PW>(p);
                double S[WV_PW], prow[WV_PW];
#pragma unroll
                for (int c = 0; c < WV_PW; c++) {
                    S[c] = rdlane(red, (64 / WV_PW) * c);
                    prow[c] = rdlane(a[0][c], j);
                }
                const double alpha = prow[jj], xn2 = S[jj];
                double tau = 0.0, beta = alpha, scal = 0.0;
                if (xn2 > 0.0) {
                    beta = -copysign(sqrt(alpha * alpha + xn2), alpha);
                    tau = (beta - alpha) / beta;
                    scal = 1.0 / (alpha - beta);
                }
                // reflector v in place of column jj: 1 at the pivot, y * scal below
                a[0][jj] = (lane == j) ? 1.0 : y0 * scal;
#pragma unroll
                for (int k = 1; k < WV_KMAX; k++) a[k][jj] *= scal;
#pragma unroll
                for (int c = 0; c < WV_PW; c++) {
                    if (c > jj) {
                        const double td = tau * (prow[c] + S[c] * scal);  // tau * v^T a_c
#pragma unroll
                        for (int k = 0; k < WV_KMAX; k++) a[k][c] = fma(-td, a[k][jj], a[k][c]);
                    }
                }
                // row block 0 of column jj: beta at the pivot, R entries above it
                a[0][jj] = (lane == j) ? beta : ((lane > j) ? a[0][jj] : x0);
                // T(0:jj, jj) = -tau T(0:jj, 0:jj) w,  w_q = v_q^T v_jj = prow[q] + S[q] * scal
#pragma unroll
                for (int l = 0; l < WV_PW; l++) {
                    if (l < jj) {
                        double acc = 0.0;
#pragma unroll
                        for (int q = 0; q < WV_PW; q++)
                            if (q >= l && q < jj) acc = fma(T[l][q], prow[q] + S[q] * scal, acc);
                        T[l][jj] = -tau * acc;
                    }
                }
                T[jj][jj] = tau;
            }
        }
        // R entries of the panel columns: rows <= j0 + c on lanes of row block 0
#pragma unroll
        for (int c = 0; c < WV_PW; c++)
            if (c < jb && lane < s) Rout[lane * SM + j0 + c] = (lane <= j0 + c) ? a[0][c] : 0.0;
        // reflector entries of row block 0: 1 at the pivot, 0 above
        double v0[WV_PW];
#pragma unroll
        for (int c = 0; c < WV_PW; c++)
            v0[c] = (c < jb) ? ((lane == j0 + c) ? 1.0 : (lane > j0 + c ? a[0][c] : 0.0)) : 0.0;
        // ---- trailing pairs: (t, t + 1) in p while (t + 2, t + 3) loads into q,
        //      then the roles swap ----
        for (int t = t0; t < s; t += 4) {
            wv_ld_col(q0, A.rd(t + 2, first, t + 2 < s), lane);
            wv_ld_col(q1, A.rd(t + 3, first, t + 3 < s), lane);
            wv_qr_pair(a, v0, T, p0, p1, A.wr(t, true), A.wr(t + 1, t + 1 < s), lane);
            if (t + 2 >= s) break;
            wv_ld_col(p0, A.rd(t + 4, first, t + 4 < s), lane);
            wv_ld_col(p1, A.rd(t + 5, first, t + 5 < s), lane);
            wv_qr_pair(a, v0, T, q0, q1, A.wr(t + 2, true), A.wr(t + 3, t + 3 < s), lane);
        }
    }
}

// Dense-B tasks (case iv: tile (n, k) in the band, (m, k) and (m, n) low
// rank) get Qn = -B Av precomputed by k_qn_dense into a slot of QN_COLS
// columns: slot (n - k - 1) T + (m - k - 1), T = NT - 1 - k.
#define QN_COLS 128
__host__ __device__ inline size_t qn_slot(int NT, int NB, int k, int m, int n)
{
    const int T = NT - 1 - k;
    return ((size_t)(n - k - 1) * T + (size_t)(m - k - 1)) * NB * QN_COLS;
}

// Qn = -B Av for the dense-B tasks of step k that the wave kernels take
// (ra > 0, s <= 128); grid (64x64 tiles of an NB x 128 slot, (band - 1) T).
__global__ void __launch_bounds__(BS) k_qn_dense(TLR t, int k, double *qn)
{
    __shared__ double sm[BLK_GEMM_SMEM];
    const int NB = t.NB, NT = t.NT, b = t.band_size, T = NT - 1 - k;
    const int m = k + 1 + (int)(blockIdx.y % T), n = k + 1 + (int)(blockIdx.y / T);
    if (m - n < b) return;
    const int ra = *rank_at(t, m, k), rc = *rank_at(t, m, n);
    if (ra <= 0 || rc + ra > 128) return;
    const int nrt = (NB + GT - 1) / GT;
    const int ti = blockIdx.x % nrt, tj = blockIdx.x / nrt;
    if (tj * GT >= ra) return;
    const double *Av = lr_tile(t, m, k) + (size_t)NB * ra;
    gemm_tile<false, false>(ti * GT, tj * GT, NB, ra, NB, -1.0, band_tile(t, n, k), NB, Av, NB, 0.0,
                            qn + qn_slot(NT, NB, k, m, n), NB, sm);
}

// Tasks of step k whose update runs on one wave: both factors low rank,
// s = rc + ra in [1, 128], rb <= WV_SMAX, NB <= 64 * WV_KMAX.  They are
// appended to list 0 (s, rb <= 32), 1 (<= 48), 2 (s <= WV_S2MAX) or 3 (s <= 128,
// rb <= WV_SMAX: k_rc_wide); the order inside
// a list is irrelevant (tasks of one step are independent).  Listed tasks are
// marked -1 in tasks_s; every other task keeps its k_classify value.  With
// wide_only, every listed task goes to list 3 (four waves per task).
__global__ void k_classify_wave(TLR t, int k, int *tasks_s, int *wlist, int *wcount, int cap, int wide_only,
                                int qn_dense)
{
    const int T = t.NT - 1 - k;
    const long long ntask = (long long)T * (T - 1) / 2;
    long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= ntask) return;
    int hi, lo;
    tri_decode(e, hi, lo);
    const int m = k + 1 + hi, n = k + 1 + lo, b = t.band_size;
    if (m - n < b || t.NB > 64 * WV_KMAX) return;
    // a dense B (band tile) comes with its Qn = -B Av formed by k_qn_dense
    const bool bdense = (n - k < b);
    if (bdense && !qn_dense) return;
    const int ra = *rank_at(t, m, k), rc = *rank_at(t, m, n), rb = bdense ? 0 : *rank_at(t, n, k);
    const int s = rc + ra;
    if (ra > 0 && (bdense || rb > 0) && s <= 128 && rb <= WV_SMAX) {
        const int l = wide_only ? 3 : ((s <= 32 && rb <= 32) ? 0 : ((s <= 48 && rb <= 48) ? 1 : (s <= WV_S2MAX ? 2 : 3)));
        tasks_s[e] = -1;
        wlist[(size_t)l * cap + atomicAdd(wcount + l, 1)] = (int)e;
    }
}

// One-sided Jacobi on the rows of M (lane c owns row c of M in xm, zero
// beyond s), round-robin pairs; a lane reads its partner's row with
// ds_bpermute (no LDS storage, no wave barrier).  Loops cover the first W
// entries (W >= s).  Returns ||M||_F; counts a run out of sweeps in *unconv.
template <int SM, int W>
__device__ __forceinline__ double wv_jacobi(double (&xm)[SM], int s, double acc, int fixed_rk, int *unconv)
{
    const int lane = threadIdx.x & 63;
    const int npad = s + (s & 1);
    const double tol = 4.0e-16 * sqrt((double)s);
    double al = 0.0;
#pragma unroll
    for (int q = 0; q < W; q++) al = fma(xm[q], xm[q], al);
    // A pair is left alone when its smaller column is below the
    // rounding-noise floor (2.3e-16 s)^2 ||M||_F^2, or, when truncating
    // at acc, below (1e-3 acc)^2: such columns stay below the
    // truncation threshold (all of them together hold less than
    // 0.012 acc of energy) and rotating them moves a kept column by
    // less than one rounding of its norm.
    const double f2 = wsum(lane < s ? al : 0.0);
    const double nz = wv_skip_floor(f2, s, acc, fixed_rk);
    // position of column `lane` in round 0 of the round-robin schedule
    const int p0 = (lane == 0) ? 0 : lane;
    int sweep = 0;
    for (; sweep < BLK_JACOBI_SWEEPS; sweep++) {
        int rot_any = 0;
        int pos = p0;  // column at position p in round r: ((p - 1 + r) mod (npad - 1)) + 1
        for (int r = 0; r < npad - 1; r++) {
            const int pp = npad - 1 - pos;
            int pc = 0;
            if (pp != 0) {
                pc = pp + r;
                if (pc > npad - 1) pc -= npad - 1;
            }
            const int src = 4 * pc;  // pc < npad <= 64
            double yv[W];
#pragma unroll
            for (int q = 0; q < W; q++) yv[q] = wv_bperm(src, xm[q]);
            const double be = wv_bperm(src, al);
            if (lane < s && pc < s) {
                double g4[4] = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
                for (int q = 0; q < W; q++) g4[q & 3] = fma(xm[q], yv[q], g4[q & 3]);
                const double ga = (g4[0] + g4[1]) + (g4[2] + g4[3]);
                const bool lo_ = lane < pc;
                double cs, sn;
                if (fmin(al, be) > nz && jacobi_rot(lo_ ? al : be, lo_ ? be : al, ga, tol, cs, sn)) {
                    rot_any = 1;
                    const double ss = lo_ ? -sn : sn;
#pragma unroll
                    for (int q = 0; q < W; q++) xm[q] = fma(ss, yv[q], cs * xm[q]);
                    // squared norm after the rotation
                    al = lo_ ? (cs * cs * al - 2.0 * cs * sn * ga + sn * sn * be)
                             : (sn * sn * be + 2.0 * cs * sn * ga + cs * cs * al);
                }
            }
            // next round: every column but 0 moves one position down
            if (lane > 0) pos = (pos == 1) ? npad - 1 : pos - 1;
        }
        // exact norms once per sweep
        al = 0.0;
#pragma unroll
        for (int q = 0; q < W; q++) al = fma(xm[q], xm[q], al);
        if (__ballot(rot_any) == 0) break;
    }
    if (sweep >= BLK_JACOBI_SWEEPS && lane == 0) atomicAdd(unconv, 1);
    return sqrt(f2);
}

template <int SM>
__global__ void __launch_bounds__(64 * WvCfg<SM>::WPB, WvCfg<SM>::MINW)
k_rc_wave(TLR t, int k, const int *wlist, const int *wcount, int *wnext, int *tasks_s, double *ws, double acc,
          int fixed_rk, int maxrank, OpLog o, int *ovf, const double *qn_dense)
{
    using Cf = WvCfg<SM>;
    extern __shared__ double smem[];
    const int lane = threadIdx.x & 63;
    const int wv = uni(threadIdx.x >> 6);
    double *L = smem + (size_t)wv * Cf::LDS;
    double *sg = L + Cf::SIG;
    const int NB = t.NB, NT = t.NT;
    const int KR = (NB + 63) >> 6;
    const int gw = blockIdx.x * Cf::WPB + wv;
    double *Wq = ws + (size_t)gw * wv_slot_doubles(NB);
    double *Qn = Wq + (size_t)NB * SM;
    double *Out = Qn + (size_t)NB * SM;
    double *Rp = Out + (size_t)NB * 2 * SM;
    double *Rq = Rp + SM * SM;
    double *Ku = Rq + SM * SM;
    const int cnt = uni(*wcount);
    // tasks are taken one at a time from a shared counter (costs differ)
    for (;;) {
        int idx = 0;
        if (lane == 0) idx = atomicAdd(wnext, 1);
        idx = uni(idx);
        if (idx >= cnt) break;
        const int e = uni(wlist[idx]);
        int hi, lo;
        tri_decode(e, hi, lo);
        const int m = uni(k + 1 + hi), n = uni(k + 1 + lo);
        const bool bdense = (n - k < t.band_size);
        const int ra = uni(*rank_at(t, m, k)), rc = uni(*rank_at(t, m, n)), rb = bdense ? 0 : uni(*rank_at(t, n, k));
        const int s = rc + ra;
        const double *Au = lr_tile(t, m, k), *Av = Au + (size_t)NB * ra;
        const double *Bu = bdense ? Au : lr_tile(t, n, k), *Bv = Bu + (size_t)NB * rb;
        double *C = lr_tile(t, m, n);
        // Q's second block: formed below, or precomputed for a dense B
        const double *Qs = bdense ? qn_dense + qn_slot(NT, NB, k, m, n) : Qn;

        // ---- X = Av^T Bv (ra x rb) on the matrix cores: 16x16 tiles of X,
        //      k runs over the NB rows four at a time ----
        if (!bdense) {
            const int i16 = lane & 15, k4 = lane >> 4;
            constexpr int TT = SM / 16, U = SM <= 32 ? 4 : (SM <= 48 ? 2 : 4);
            wv_d4 xacc[TT][TT];
#pragma unroll
            for (int ta = 0; ta < TT; ta++)
#pragma unroll
                for (int tb = 0; tb < TT; tb++) xacc[ta][tb] = wv_d4{0.0, 0.0, 0.0, 0.0};
            // rows r0 .. r0 + 4U - 1 of Av's and Bv's columns by buffer loads;
            // out-of-range lanes read 0 (no branch, no select)
            const wv_rsrc ar = wv_col_rsrc(Av, ra * NB * 8), br = wv_col_rsrc(Bv, rb * NB * 8);
            auto ldx = [&](double (&av)[U][TT], double (&bv)[U][TT], int r0) {
#pragma unroll
                for (int u = 0; u < U; u++) {
                    const int row = r0 + 4 * u + k4;
#pragma unroll
                    for (int tt = 0; tt < TT; tt++) {
                        const int a = 16 * tt + i16;
                        av[u][tt] = wv_ld(ar, (a < ra && row < NB) ? (row + a * NB) * 8 : WV_OOB);
                        bv[u][tt] = wv_ld(br, (a < rb && row < NB) ? (row + a * NB) * 8 : WV_OOB);
                    }
                }
            };
            auto mulx = [&](const double (&av)[U][TT], const double (&bv)[U][TT]) {
#pragma unroll
                for (int u = 0; u < U; u++)
#pragma unroll
                    for (int ta = 0; ta < TT; ta++)
                        if (16 * ta < ra) {
#pragma unroll
                            for (int tb = 0; tb < TT; tb++)
                                if (16 * tb < rb) xacc[ta][tb] = mfma16(av[u][ta], bv[u][tb], xacc[ta][tb]);
                        }
            };
            if constexpr (SM <= 48) {
                // two register sets in turn: rows of the next step load
                // while the current ones multiply
                double av0[U][TT], bv0[U][TT], av1[U][TT], bv1[U][TT];
                ldx(av0, bv0, 0);
#pragma unroll 1
                for (int r0 = 0; r0 < NB; r0 += 8 * U) {
                    if (r0 + 4 * U < NB) ldx(av1, bv1, r0 + 4 * U);
                    mulx(av0, bv0);
                    if (r0 + 4 * U >= NB) break;
                    if (r0 + 8 * U < NB) ldx(av0, bv0, r0 + 8 * U);
                    mulx(av1, bv1);
                }
            } else {  // one set (the registers of the widest class are taken)
#pragma unroll 1
                for (int r0 = 0; r0 < NB; r0 += 4 * U) {
                    double av[U][TT], bv[U][TT];
                    ldx(av, bv, r0);
                    mulx(av, bv);
                }
            }
            // X^T into LDS as Xl[b * SM + a], zero outside ra x rb
            double *Xl = L;
            for (int q = lane; q < SM * SM; q += 64) Xl[q] = 0.0;
            wv_sync();
#pragma unroll
            for (int ta = 0; ta < TT; ta++)
#pragma unroll
                for (int tb = 0; tb < TT; tb++)
                    if (16 * ta < ra && 16 * tb < rb) {
#pragma unroll
                        for (int r = 0; r < 4; r++) {
                            const int a = 16 * ta + k4 + 4 * r, bb = 16 * tb + i16;
                            if (a < ra && bb < rb) Xl[bb * SM + a] = xacc[ta][tb][r];
                        }
                    }
            wv_sync();
        }
        // ---- Qn = -Bu X^T (NB x ra): K = X^T, K[b][a] = Xl[b * SM + a] ----
        if (!bdense) wv_tall_mm<SM>(NB, rb, ra, -1.0, WvSrc{Bu, rb, Bu, NB}, L, SM, Qn);
        wv_sync();
        // ---- R factors of P = [Cu Au] and Q = [Cv Qn] ----
        for (int side = 0; side < 2; side++) {
            wv_geqrf_r<SM>(s, side ? C + (size_t)NB * rc : C, side ? Qs : Au, rc, Wq, NB, side ? Rq : Rp);
        }
        wv_sync();
        // ---- rows of M = Rp Rq^T: lane c holds M(c, :) ----
        double xm[SM];
        {
            for (int q = lane; q < s * SM; q += 64) {
                const int l = q / SM, c = q - l * SM;
                L[q] = (c < s) ? Rq[c * SM + l] : 0.0;  // L[l][c] = Rq(c, l)
            }
            wv_sync();
#pragma unroll
            for (int q = 0; q < SM; q++) xm[q] = 0.0;
            // this lane's row of Rp, 16 entries per batch of loads
            const wv_rsrc rpr = wv_col_rsrc(Rp, SM * SM * 8);
            for (int l0 = 0; l0 < s; l0 += 16) {
                double rv[16];
#pragma unroll
                for (int u = 0; u < 16; u++) {
                    const int l = l0 + u;
                    rv[u] = wv_ld(rpr, (lane < s && l < s && l >= lane) ? (lane * SM + l) * 8 : WV_OOB);
                }
#pragma unroll
                for (int u = 0; u < 16; u++)
                    if (l0 + u < s) {
                        const double *rq = L + (l0 + u) * SM;
#pragma unroll
                        for (int q8 = 0; q8 < SM; q8 += 8)
                            if (q8 < s) {
#pragma unroll
                                for (int q = q8; q < q8 + 8; q++) xm[q] = fma(rv[u], rq[q], xm[q]);
                            }
                    }
            }
            wv_sync();
        }
        // ---- one-sided Jacobi on the columns of M^T (lane c owns column c):
        //      loops run over SM - 8 entries when s fits in them (entries
        //      beyond s are zero, so the width changes no bit) ----
        double fro;
        if constexpr (SM <= 48) {
            constexpr int WJ = SM - 8;
            fro = (s <= WJ) ? wv_jacobi<SM, WJ>(xm, s, acc, fixed_rk, o.unconv)
                            : wv_jacobi<SM, SM>(xm, s, acc, fixed_rk, o.unconv);
        } else {  // one width: a second copy costs the widest class registers
            fro = wv_jacobi<SM, SM>(xm, s, acc, fixed_rk, o.unconv);
        }
        // ---- singular values, order, rank ----
        double sig = 0.0;
#pragma unroll
        for (int q = 0; q < SM; q++) sig = fma(xm[q], xm[q], sig);
        sig = sqrt(sig);
        if (lane < s) sg[lane] = sig;
        wv_sync();
        int pos = 0;
        for (int q = 0; q < s; q++) {
            const double o2 = sg[q];
            pos += (o2 > sig) || (o2 == sig && q < lane);
        }
        int nr;
        if (fixed_rk > 0) nr = fixed_rk < s ? fixed_rk : s;
        else nr = __popcll(__ballot(lane < s && sig >= acc));
        if (nr > maxrank) {
            if (lane == 0) atomicAdd(ovf, 1);
            nr = maxrank;
        }
        nr = uni(nr);
        wv_sync();
        if (nr > 0) {
            // Vs_r into LDS, row-major: Vl[l * SM + pos]; sigma by position
            double *Vl = L;
            double *sgp = sg + SM;
            if (lane < s && pos < nr) {
                const double inv = sig > 0.0 ? 1.0 / sig : 0.0;
#pragma unroll
                for (int q = 0; q < SM; q++)
                    if (q < s) Vl[q * SM + pos] = xm[q] * inv;
                sgp[pos] = sig;
            }
            wv_sync();
            // Ku = Rq^T Vs_r: lane c computes row c -> Ku[c * SM + j]
            {
                double ac[SM];
#pragma unroll
                for (int j = 0; j < SM; j++) ac[j] = 0.0;
                // column `lane` of Rq, 16 entries per batch of loads
                const wv_rsrc rqr = wv_col_rsrc(Rq, SM * SM * 8);
                for (int l0 = 0; l0 < s; l0 += 16) {
                    double rv[16];
#pragma unroll
                    for (int u = 0; u < 16; u++) {
                        const int l = l0 + u;
                        rv[u] = wv_ld(rqr, (lane < s && l < s && l <= lane) ? (l * SM + lane) * 8 : WV_OOB);
                    }
#pragma unroll
                    for (int u = 0; u < 16; u++)
                        if (l0 + u < s) {
                            const double *vr = Vl + (l0 + u) * SM;
#pragma unroll
                            for (int j8 = 0; j8 < SM; j8 += 8)
                                if (j8 < nr) {
#pragma unroll
                                    for (int j = j8; j < j8 + 8; j++) ac[j] = fma(rv[u], vr[j], ac[j]);
                                }
                        }
                }
                if (lane < s) {
#pragma unroll
                    for (int j = 0; j < SM; j++)
                        if (j < nr) Ku[lane * SM + j] = ac[j];
                }
            }
            wv_sync();
            // Kv = Rq^-1 Vs_r in place (lane j solves column j) by plain back
            // substitution.  Q Kv = Qq Vs_r for any Kv with Rq Kv = Vs_r, and
            // the substitution is backward stable, so the tiny pivots of a
            // numerically rank-deficient Q are divided by, not dropped (dropping
            // one breaks the solve when its row carries other entries).  Column
            // j then errs in the tile by at most
            // sigma_j s eps max_q sum_l |Rq(q,l)| |Kv(l,j)|; a task over the
            // tolerance, or with an exactly zero pivot, is left untouched and
            // handed to the Householder path (k_lr_gemm).
            // Row q of Rq is loaded across the lanes (lane l: Rq(q, l), one
            // load per row, the next row's issued ahead) and read by readlane.
            bool bad = false;
            {
                const wv_rsrc rqr = wv_col_rsrc(Rq, SM * SM * 8);
                double gmax = 0.0;
                double rnext = wv_ld(rqr, (lane < s) ? ((s - 1) * SM + lane) * 8 : WV_OOB);
                for (int q = s - 1; q >= 0; q--) {
                    const double rrow = rnext;
                    rnext = wv_ld(rqr, (q > 0 && lane < s) ? ((q - 1) * SM + lane) * 8 : WV_OOB);
                    double v = Vl[q * SM + lane], gq = 0.0;
                    for (int l = q + 1; l < s; l++) {
                        const double rql = rdlane(rrow, l), xl = Vl[l * SM + lane];
                        v = fma(-rql, xl, v);
                        gq = fma(fabs(rql), fabs(xl), gq);
                    }
                    const double d = rdlane(rrow, q);
                    const bool ok = fabs(d) > 1e-300;
                    const double x = ok ? v / d : 0.0;
                    bad = bad || !ok;
                    gmax = fmax(gmax, fma(fabs(d), fabs(x), gq));
                    if (lane < nr) Vl[q * SM + lane] = x;
                }
                bad = lane < nr && (bad || !(sgp[lane] * s * 2.3e-16 * gmax <= rc_solve_tol(acc, fixed_rk, fro)));
            }
            if (__ballot(bad) != 0) {
                if (lane == 0) tasks_s[e] = RC_REDO;
                wv_sync();
                continue;
            }
            wv_sync();
            // ---- outputs, written in place into the tile.  A product reads a
            //      batch of rows of all its sources before writing those rows,
            //      and every column of the tile has stride NB, so an output
            //      may overwrite its own sources.  The two products are
            //      ordered so that neither overwrites the other's sources:
            //      U' = P Ku (columns 0 .. nr-1, sources Cu = columns
            //      0 .. rc-1) and V' = Q Kv (columns nr .. 2nr-1, sources
            //      Cv = columns rc .. 2rc-1): U' first when nr <= rc, V'
            //      first otherwise. ----
            auto v_out = [&]() {  // Kv in LDS
                wv_tall_mm<SM>(NB, s, nr, 1.0, WvSrc{C + (size_t)rc * NB, rc, Qs, NB}, Vl, SM, C + (size_t)nr * NB);
                wv_sync();
            };
            auto u_out = [&]() {  // stages Ku into LDS
                for (int q = lane; q < s * SM; q += 64) {
                    const int c = q / SM, j = q - c * SM;
                    Vl[q] = (j < nr) ? Ku[c * SM + j] : 0.0;
                }
                wv_sync();
                wv_tall_mm<SM>(NB, s, nr, 1.0, WvSrc{C, rc, Au, NB}, Vl, SM, C);
                wv_sync();
            };
            if (nr > rc) {
                v_out();
                u_out();
            } else {
                for (int q = lane; q < s * SM; q += 64) Out[q] = Vl[q];  // Kv aside while Ku uses the LDS
                u_out();
                for (int q = lane; q < s * SM; q += 64) Vl[q] = Out[q];
                wv_sync();
                v_out();
            }
        }
        if (lane == 0) {
            *rank_at(t, m, n) = nr;
            rank_hist(o, NT, m, n, nr);
            unsigned long long mm = tdim(t, m), ss = s;
            unsigned long long qa = opc('q', mm, ss, 0, 0);
            // case iv (dense B) counts the product B Av, case v the products with B's factors
            unsigned long long qb = bdense ? opc('q', mm, ss, 0, 0) + opc('m', ra, mm, mm, 0)
                                           : opc('m', ra, rb, mm, 0) + opc('m', ra, mm, rb, 0) + opc('q', mm, ss, 0, 0);
            unsigned long long sv = opc('r', ss, ss, 2, 0) + opc('s', ss, 0, 0, 0) + ss * nr;
            unsigned long long nu = opc('o', mm, nr, ss, 1), nv = opc('o', nr, mm, ss, 2);
            oplog(o, NT, qa + qb + sv + nu + nv, m, n, false, false);
            atomicAdd(o.stat + 4, qa + qb + sv + nu + nv);
            atomicAdd(o.stat + 5, 1ULL);
            unsigned long long words =
                2ULL * NB * (ra + rc + nr) + (bdense ? (unsigned long long)NB * NB : 2ULL * NB * rb);
            atomicAdd(o.stat + 6, 8ULL * words);
        }
        wv_sync();
    }
}

// ===================== stacks wider than 64 (s <= 128) =====================

#define WB_SM 128                       // largest stack width of the wide variant
#define WB_LD (WB_SM + 1)               // LDS column stride of the Jacobi matrix
#define WB_LDS (WB_SM * WB_LD + 3 * WB_SM)  // doubles of LDS (one workgroup per CU)
#define WB_QR_SH (64 * (WV_KMAX * WV_PW + 1))  // doubles of LDS per QR side (two sides fit in WB_LDS)

__host__ __device__ inline size_t wb_slot_doubles(int NB)
{
    return (size_t)NB * 4 * WB_SM + 3 * (size_t)WB_SM * WB_SM;
}

// wv_geqrf_r for s <= 128 on two waves: pivot row j lives on lane j & 63 of
// row block j >> 6 (0 or 1); rows of block 0 are above every pivot of block 1.
// The lead wave (part 0) factors each panel and passes its reflectors and T
// to the helper (part 1) through `sh` (64 (16 WV_PW + 1) doubles of LDS);
// the two then update alternate trailing columns.  Every wave of the
// workgroup must call it with the same s (the barriers are workgroup-wide).
// R is written column-major (R(i, j) at Rout[i + j * WB_SM]).
__device__ __forceinline__ void wb_geqrf_r(int s, const double *src0, const double *src1, int c0, double *W, int NB,
                                           double *Rout, int part, double *sh)
{
    const int lane = threadIdx.x & 63;
    const WvRs A{src0, src1, W, c0, NB};
    for (int j0 = 0; j0 < s; j0 += WV_PW) {
        const int jb = min(WV_PW, s - j0);
        const bool first = (j0 == 0);
        const int ps = j0 >> 6;  // row block of the panel's pivots (panels never straddle 64)
        double a[WV_KMAX][WV_PW];
        double T[WV_PW][WV_PW];
        double m0[WV_PW], m1[WV_PW];
        // this wave's trailing columns t0, t0 + 2, ... in two register sets
        // used in turn (the next column loads while the current one updates)
        const int t0 = j0 + jb + part;
        double p_[WV_KMAX], q_[WV_KMAX];
        if (part == 0) {
#pragma unroll
            for (int c = 0; c < WV_PW; c++) {
                const wv_rsrc r = A.rd(j0 + c, first, c < jb);
#pragma unroll
                for (int k = 0; k < WV_KMAX; k++) a[k][c] = wv_ld(r, (lane + 64 * k) * 8);
            }
            wv_ld_col(p_, A.rd(t0, first, t0 < s), lane);
#pragma unroll
            for (int r = 0; r < WV_PW; r++)
#pragma unroll
                for (int c = 0; c < WV_PW; c++) T[r][c] = 0.0;
#pragma unroll
            for (int jj = 0; jj < WV_PW; jj++) {
                if (jj < jb) {
                    const int jl = (j0 + jj) & 63;
                    const double x0 = a[0][jj], x1 = a[1][jj];
                    // part of column jj strictly below the pivot, per row block
                    const double y0 = (ps == 0 && lane > jl) ? x0 : 0.0;
                    const double y1 = (ps == 0 || lane > jl) ? x1 : 0.0;
                    double p[WV_PW];
#pragma unroll
                    for (int c = 0; c < WV_PW; c++)
                        p[c] = y0 * (c == jj ? y0 : a[0][c]) + y1 * (c == jj ? y1 : a[1][c]);
#pragma unroll
                    for (int k = 2; k < WV_KMAX; k++)
#pragma unroll
                        for (int c = 0; c < WV_PW; c++) p[c] = fma(a[k][jj], a[k][c], p[c]);
                    const double red = wrsN<WV_PW>(p);
                    double S[WV_PW], prow[WV_PW];
#pragma unroll
                    for (int c = 0; c < WV_PW; c++) {
                        S[c] = rdlane(red, (64 / WV_PW) * c);
                        prow[c] = rdlane(ps ? a[1][c] : a[0][c], jl);
                    }
                    const double alpha = prow[jj], xn2 = S[jj];
                    double tau = 0.0, beta = alpha, scal = 0.0;
                    if (xn2 > 0.0) {
                        beta = -copysign(sqrt(alpha * alpha + xn2), alpha);
                        tau = (beta - alpha) / beta;
                        scal = 1.0 / (alpha - beta);
                    }
                    const double v0 = (ps == 0 && lane == jl) ? 1.0 : y0 * scal;
                    const double v1 = (ps == 1 && lane == jl) ? 1.0 : y1 * scal;
                    a[0][jj] = v0;
                    a[1][jj] = v1;
#pragma unroll
                    for (int k = 2; k < WV_KMAX; k++) a[k][jj] *= scal;
#pragma unroll
                    for (int c = 0; c < WV_PW; c++) {
                        if (c > jj) {
                            const double td = tau * (prow[c] + S[c] * scal);
#pragma unroll
                            for (int k = 0; k < WV_KMAX; k++) a[k][c] = fma(-td, a[k][jj], a[k][c]);
                        }
                    }
                    // R entries above the pivot and beta at the pivot; reflector below
                    if (ps == 0) {
                        a[0][jj] = (lane == jl) ? beta : ((lane > jl) ? v0 : x0);
                    } else {
                        a[0][jj] = x0;
                        a[1][jj] = (lane == jl) ? beta : ((lane > jl) ? v1 : x1);
                    }
#pragma unroll
                    for (int l = 0; l < WV_PW; l++) {
                        if (l < jj) {
                            double acc = 0.0;
#pragma unroll
                            for (int q = 0; q < WV_PW; q++)
                                if (q >= l && q < jj) acc = fma(T[l][q], prow[q] + S[q] * scal, acc);
                            T[l][jj] = -tau * acc;
                        }
                    }
                    T[jj][jj] = tau;
                }
            }
#pragma unroll
            for (int c = 0; c < WV_PW; c++)
                if (c < jb) {
#pragma unroll
                    for (int blk = 0; blk < 2; blk++) {
                        const int i = lane + 64 * blk;
                        if (i < s) Rout[i + (size_t)(j0 + c) * WB_SM] = (i <= j0 + c) ? a[blk][c] : 0.0;
                    }
                }
            // reflector entries of row blocks 0 and 1 (1 at the pivot, 0 above)
#pragma unroll
            for (int c = 0; c < WV_PW; c++) {
                const int gl = (j0 + c) & 63;
                const bool in = c < jb;
                m0[c] = (!in || ps == 1) ? 0.0 : ((lane == gl) ? 1.0 : (lane > gl ? a[0][c] : 0.0));
                m1[c] = !in ? 0.0 : (ps == 0 ? a[1][c] : ((lane == gl) ? 1.0 : (lane > gl ? a[1][c] : 0.0)));
            }

            // the panel's reflectors and T for the helper wave (value-major:
            // consecutive lanes, consecutive banks)
#pragma unroll
            for (int c = 0; c < WV_PW; c++) {
                sh[(0 * WV_PW + c) * 64 + lane] = m0[c];
                sh[(1 * WV_PW + c) * 64 + lane] = m1[c];
#pragma unroll
                for (int k = 2; k < WV_KMAX; k++) sh[(k * WV_PW + c) * 64 + lane] = a[k][c];
            }
            if (lane == 0) {  // T is the same on every lane
#pragma unroll
                for (int r = 0; r < WV_PW; r++)
#pragma unroll
                    for (int c = 0; c < WV_PW; c++) sh[WV_KMAX * WV_PW * 64 + r * WV_PW + c] = T[r][c];
            }
        }
        __syncthreads();
        if (part == 1) {
            wv_ld_col(p_, A.rd(t0, first, t0 < s), lane);
#pragma unroll
            for (int c = 0; c < WV_PW; c++) {
                m0[c] = sh[(0 * WV_PW + c) * 64 + lane];
                m1[c] = sh[(1 * WV_PW + c) * 64 + lane];
#pragma unroll
                for (int k = 2; k < WV_KMAX; k++) a[k][c] = sh[(k * WV_PW + c) * 64 + lane];
            }
#pragma unroll
            for (int r = 0; r < WV_PW; r++)
#pragma unroll
                for (int c = 0; c < WV_PW; c++) T[r][c] = sh[WV_KMAX * WV_PW * 64 + r * WV_PW + c];
        }
        // trailing columns j0 + jb + part, + 2, ...: the two waves of the
        // side take alternate columns
        auto upd = [&](double (&b)[WV_KMAX], int t) {
            double p[WV_PW];
#pragma unroll
            for (int c = 0; c < WV_PW; c++) {
                double acc = fma(m0[c], b[0], m1[c] * b[1]);
#pragma unroll
                for (int k = 2; k < WV_KMAX; k++) acc = fma(a[k][c], b[k], acc);
                p[c] = acc;
            }
            const double red = wrsN<WV_PW>(p);
            double d[WV_PW], w[WV_PW];
#pragma unroll
            for (int c = 0; c < WV_PW; c++) d[c] = rdlane(red, (64 / WV_PW) * c);
#pragma unroll
            for (int c = 0; c < WV_PW; c++) {
                double acc = 0.0;
#pragma unroll
                for (int q = 0; q <= c; q++) acc = fma(T[q][c], d[q], acc);
                w[c] = acc;
            }
#pragma unroll
            for (int c = 0; c < WV_PW; c++) {
                b[0] = fma(-m0[c], w[c], b[0]);
                b[1] = fma(-m1[c], w[c], b[1]);
            }
#pragma unroll
            for (int k = 2; k < WV_KMAX; k++) {
                double acc = b[k];
#pragma unroll
                for (int c = 0; c < WV_PW; c++) acc = fma(-a[k][c], w[c], acc);
                b[k] = acc;
            }
            wv_st_col(b, A.wr(t, true), lane);
        };
        for (int t = t0; t < s; t += 4) {
            wv_ld_col(q_, A.rd(t + 2, first, t + 2 < s), lane);
            upd(p_, t);
            if (t + 2 >= s) break;
            wv_ld_col(p_, A.rd(t + 4, first, t + 4 < s), lane);
            upd(q_, t + 2);
        }
        __syncthreads();  // both waves' trailing columns are in W before the next panel
    }
}

// Tasks of step k with WV_S2MAX < s <= 128 and rb <= 64 (list 3): one workgroup of
// four waves each.  The tall products are split by 16-row tiles over the
// waves, the two R factors run at once on waves 0 and 1, the Jacobi SVD uses
// four lanes per column pair, and the small s x s products and solves take a
// thread per row or column.  The workgroup already holds a whole CU's LDS,
// so the extra waves cost no occupancy.
#define WB_WAVES 4
__global__ void __launch_bounds__(64 * WB_WAVES, 1)
k_rc_wide(TLR t, int k, const int *wlist, const int *wcount, int *wnext, int *tasks_s, double *ws, double acc,
          int fixed_rk, int maxrank, OpLog o, int *ovf, const double *qn_dense)
{
    extern __shared__ double smem[];
    __shared__ int s_idx;
    __shared__ double s_nz, s_fro;
    constexpr int NTH = 64 * WB_WAVES;
    const int tid = threadIdx.x, lane = tid & 63, wid = uni(tid >> 6);
    double *L = smem;
    double *sg = L + WB_SM * WB_LD;
    const int NB = t.NB, NT = t.NT;
    double *Wq = ws + (size_t)blockIdx.x * wb_slot_doubles(NB);
    double *Qn = Wq + (size_t)NB * WB_SM;
    double *Out = Qn + (size_t)NB * WB_SM;
    double *Rp = Out + (size_t)NB * 2 * WB_SM;
    double *Rq = Rp + WB_SM * WB_SM;
    double *Ku = Rq + WB_SM * WB_SM;
    const int cnt = uni(*wcount);
    for (;;) {
        if (tid == 0) s_idx = atomicAdd(wnext, 1);
        __syncthreads();
        const int idx = uni(s_idx);
        __syncthreads();
        if (idx >= cnt) break;
        const int e = uni(wlist[idx]);
        int hi, lo;
        tri_decode(e, hi, lo);
        const int m = uni(k + 1 + hi), n = uni(k + 1 + lo);
        const bool bdense = (n - k < t.band_size);
        const int ra = uni(*rank_at(t, m, k)), rc = uni(*rank_at(t, m, n)), rb = bdense ? 0 : uni(*rank_at(t, n, k));
        const int s = rc + ra;
        const double *Au = lr_tile(t, m, k), *Av = Au + (size_t)NB * ra;
        const double *Bu = bdense ? Au : lr_tile(t, n, k), *Bv = Bu + (size_t)NB * rb;
        double *C = lr_tile(t, m, n);
        // Q's second block: formed below, or precomputed for a dense B
        const double *Qs = bdense ? qn_dense + qn_slot(NT, NB, k, m, n) : Qn;
        const int i16 = lane & 15, k4 = lane >> 4;
        double *J = L;
        // ---- X = Av^T Bv (ra <= 128, rb <= 64): 16-row band ta of X on wave ta % 4 ----
        if (!bdense) {
            double *Xl = L;  // Xl[b * WB_SM + a]
            for (int q = tid; q < 64 * WB_SM; q += NTH) Xl[q] = 0.0;
            __syncthreads();
            for (int ta = wid; 16 * ta < ra; ta += WB_WAVES) {
                wv_d4 xacc[4];
#pragma unroll
                for (int tb = 0; tb < 4; tb++) xacc[tb] = wv_d4{0.0, 0.0, 0.0, 0.0};
                const int a = 16 * ta + i16;
                for (int r0 = 0; r0 < NB; r0 += 16) {
                    double av[4], bv[4][4];
#pragma unroll
                    for (int u = 0; u < 4; u++) {
                        const int row = r0 + 4 * u + k4;
                        av[u] = (a < ra && row < NB) ? Av[row + (size_t)a * NB] : 0.0;
#pragma unroll
                        for (int tb = 0; tb < 4; tb++) {
                            const int bb = 16 * tb + i16;
                            bv[u][tb] = (16 * tb < rb && bb < rb && row < NB) ? Bv[row + (size_t)bb * NB] : 0.0;
                        }
                    }
#pragma unroll
                    for (int u = 0; u < 4; u++)
#pragma unroll
                        for (int tb = 0; tb < 4; tb++)
                            if (16 * tb < rb) xacc[tb] = mfma16(av[u], bv[u][tb], xacc[tb]);
                }
#pragma unroll
                for (int tb = 0; tb < 4; tb++)
                    if (16 * tb < rb) {
#pragma unroll
                        for (int r = 0; r < 4; r++) {
                            const int aa = 16 * ta + k4 + 4 * r, bb = 16 * tb + i16;
                            if (aa < ra && bb < rb) Xl[bb * WB_SM + aa] = xacc[tb][r];
                        }
                    }
            }
            __syncthreads();
        }
        // ---- Qn = -Bu X^T ----
        if (!bdense)
            wv_tall_mm<WB_SM>(NB, rb, ra, -1.0, WvSrc{Bu, rb, Bu, NB}, L, WB_SM, Qn, wid,
                              WB_WAVES);
        __syncthreads();
        // ---- R factors: P = [Cu Au] on waves 0 and 2 (trailing copy in Wq),
        //      Q = [Cv Qn] on waves 1 and 3 (trailing copy in Out, unused until
        //      the outputs); the panels' reflectors pass through the LDS,
        //      free until the Jacobi ----
        {
            const int side = wid & 1;
            wb_geqrf_r(s, side ? C + (size_t)NB * rc : C, side ? Qs : Au, rc, side ? Out : Wq, NB, side ? Rq : Rp,
                       wid >> 1, L + (size_t)side * WB_QR_SH);
        }
        __syncthreads();
        // ---- M^T = Rq Rp^T on the matrix cores (K = Rp^T = Rp's column-major
        //      storage read row-wise, staged in LDS) into W (ld s); column c of
        //      M^T (= row c of M) then goes to the Jacobi matrix ----
        for (int q = tid; q < s * WB_SM; q += NTH) {
            const int c = q % WB_SM;
            L[q] = (c < s) ? Rp[q] : 0.0;  // K[l][c] = Rp(c, l)
        }
        __syncthreads();
        wv_tall_mm<WB_SM>(s, s, s, 1.0, WvSrc{Rq, s, Rq, WB_SM}, L, WB_SM, Wq,
                          wid, WB_WAVES);
        __syncthreads();
        for (int q = tid; q < s * WB_LD; q += NTH) {
            const int c = q / WB_LD, r = q - c * WB_LD;
            J[q] = (r < s) ? Wq[r + (size_t)c * s] : 0.0;
        }
        __syncthreads();
        // squared column norms, kept up to date through the rotations and
        // recomputed exactly after every sweep
        double *nrm = sg + 2 * WB_SM;
        for (int c = tid; c < s; c += NTH) {
            const double *x = J + c * WB_LD;
            double v = 0.0;
            for (int q = 0; q < s; q++) v = fma(x[q], x[q], v);
            nrm[c] = v;
        }
        __syncthreads();
        if (wid == 0) {  // skip floor of the Jacobi rotations
            double f2 = 0.0;
            for (int c = lane; c < s; c += 64) f2 += nrm[c];
            const double f2t = wsum(f2);
            const double nzv = wv_skip_floor(f2t, s, acc, fixed_rk);
            if (lane == 0) {
                s_nz = nzv;
                s_fro = sqrt(f2t);
            }
        }
        __syncthreads();
        // ---- one-sided Jacobi, columns in LDS: column pair tid / 4, the four
        //      lanes of a quad take every fourth entry (the quad sum by DPP
        //      gives all four the same ga, so they agree on the rotation) ----
        {
            const int npad = s + (s & 1), npairs = npad >> 1;
            const double tol = 4.0e-16 * sqrt((double)s);
            const double nz = s_nz;
            const int pk = tid >> 2, sub = tid & 3;
            int sweep = 0;
            for (; sweep < BLK_JACOBI_SWEEPS; sweep++) {
                int rot_any = 0;
                for (int r = 0; r < npad - 1; r++) {
                    if (pk < npairs) {
                        const int ca = (pk == 0) ? 0 : ((pk - 1 + r) % (npad - 1)) + 1;
                        const int pb = npad - 1 - pk;
                        const int cb = (pb == 0) ? 0 : ((pb - 1 + r) % (npad - 1)) + 1;
                        if (ca < s && cb < s) {
                            double *x = J + ca * WB_LD, *y = J + cb * WB_LD;
                            const double al = nrm[ca], be = nrm[cb];
                            double g4[4] = {0, 0, 0, 0};
                            int q = sub;
                            for (; q + 12 < s; q += 16) {
                                double xv[4], yv[4];
#pragma unroll
                                for (int u = 0; u < 4; u++) {
                                    xv[u] = x[q + 4 * u];
                                    yv[u] = y[q + 4 * u];
                                }
#pragma unroll
                                for (int u = 0; u < 4; u++) g4[u] = fma(xv[u], yv[u], g4[u]);
                            }
                            for (; q < s; q += 4) g4[0] = fma(x[q], y[q], g4[0]);
                            const double ga = gsum<4>((g4[0] + g4[1]) + (g4[2] + g4[3]));
                            double cs, sn;
                            if (fmin(al, be) > nz && jacobi_rot(al, be, ga, tol, cs, sn)) {
                                rot_any = 1;
                                if (sub == 0) {  // norms after the rotation
                                    nrm[ca] = cs * cs * al - 2.0 * cs * sn * ga + sn * sn * be;
                                    nrm[cb] = sn * sn * al + 2.0 * cs * sn * ga + cs * cs * be;
                                }
                                int q2 = sub;
                                for (; q2 + 12 < s; q2 += 16) {
                                    double xv[4], yv[4];
#pragma unroll
                                    for (int u = 0; u < 4; u++) {
                                        xv[u] = x[q2 + 4 * u];
                                        yv[u] = y[q2 + 4 * u];
                                    }
#pragma unroll
                                    for (int u = 0; u < 4; u++) {
                                        x[q2 + 4 * u] = cs * xv[u] - sn * yv[u];
                                        y[q2 + 4 * u] = sn * xv[u] + cs * yv[u];
                                    }
                                }
                                for (; q2 < s; q2 += 4) {
                                    const double xv = x[q2], yv = y[q2];
                                    x[q2] = cs * xv - sn * yv;
                                    y[q2] = sn * xv + cs * yv;
                                }
                            }
                        }
                    }
                    __syncthreads();
                }
                for (int c = tid; c < s; c += NTH) {
                    const double *x = J + c * WB_LD;
                    double v = 0.0;
                    for (int q2 = 0; q2 < s; q2++) v = fma(x[q2], x[q2], v);
                    nrm[c] = v;
                }
                if (!__syncthreads_or(rot_any)) break;
            }
            if (sweep >= BLK_JACOBI_SWEEPS && tid == 0) atomicAdd(o.unconv, 1);
        }
        // ---- singular values, order, rank (every wave counts the same nr) ----
        for (int c = tid; c < s; c += NTH) {
            const double *x = J + c * WB_LD;
            double v = 0.0;
            for (int q = 0; q < s; q++) v = fma(x[q], x[q], v);
            sg[c] = sqrt(v);
        }
        __syncthreads();
        int nr = 0;
        {
            int cntv = 0;
            for (int c = lane; c < s; c += 64) cntv += (sg[c] >= acc);
            nr = (fixed_rk > 0) ? (fixed_rk < s ? fixed_rk : s) : (int)wsum((double)cntv);
        }
        if (nr > maxrank) {
            if (tid == 0) atomicAdd(ovf, 1);
            nr = maxrank;
        }
        nr = uni(nr);
        if (nr > 0) {
            // selected columns, normalized, to Out (column pos, ld WB_SM), then
            // row-major into LDS: Vl[l * WB_SM + pos]
            for (int c = tid; c < s; c += NTH) {
                const double sc = sg[c];
                int pos = 0;
                for (int q = 0; q < s; q++) {
                    const double o2 = sg[q];
                    pos += (o2 > sc) || (o2 == sc && q < c);
                }
                if (pos < nr) {
                    const double inv = sc > 0.0 ? 1.0 / sc : 0.0;
                    const double *x = J + c * WB_LD;
                    for (int q = 0; q < s; q++) Out[q + (size_t)pos * WB_SM] = x[q] * inv;
                    sg[WB_SM + pos] = sc;
                }
            }
            __syncthreads();
            double *Vl = L;
            for (int q = tid; q < s * WB_SM; q += NTH) {
                const int l = q / WB_SM, j = q - l * WB_SM;
                Vl[q] = (j < nr) ? Out[l + (size_t)j * WB_SM] : 0.0;
            }
            __syncthreads();
            // Ku = Rq^T Vs_r: thread (row c, half h) takes column blocks 32 h, 32 h + 64
            {
                const int c = tid & (WB_SM - 1), h = tid / WB_SM;
                if (c < s) {
                    for (int j0 = 32 * h; j0 < nr; j0 += 64) {
                        double ac[32];
#pragma unroll
                        for (int j = 0; j < 32; j++) ac[j] = 0.0;
                        for (int l = 0; l <= c; l++) {
                            const double rq = Rq[l + (size_t)c * WB_SM];
                            const double *vr = Vl + l * WB_SM + j0;
#pragma unroll
                            for (int j = 0; j < 32; j++) ac[j] = fma(rq, vr[j], ac[j]);
                        }
#pragma unroll
                        for (int j = 0; j < 32; j++)
                            if (j0 + j < nr) Ku[c * WB_SM + j0 + j] = ac[j];
                    }
                }
            }
            __syncthreads();
            // Kv = Rq^-1 Vs_r in place, thread j solves column j: plain back
            // substitution with the error bound and hand-off of k_rc_wave
            // waves 0 and 1 (thread j = column j, nr <= 128); row q of Rq is
            // loaded across the lanes (lane l: Rq(q, l) and Rq(q, l + 64), the
            // next row's issued ahead) and read by readlane
            int bad = 0;
            if (wid < 2) {
                const double tol = rc_solve_tol(acc, fixed_rk, s_fro);
                const int j = tid;
                const wv_rsrc rqr = wv_col_rsrc(Rq, WB_SM * WB_SM * 8);
                auto ldrow = [&](int q, double &r0, double &r1) {
                    r0 = wv_ld(rqr, (q >= 0 && lane < s) ? (q + lane * WB_SM) * 8 : WV_OOB);
                    r1 = wv_ld(rqr, (q >= 0 && lane + 64 < s) ? (q + (lane + 64) * WB_SM) * 8 : WV_OOB);
                };
                double n0, n1, gmax = 0.0;
                ldrow(s - 1, n0, n1);
                for (int q = s - 1; q >= 0; q--) {
                    const double r0 = n0, r1 = n1;
                    ldrow(q - 1, n0, n1);
                    double v = Vl[q * WB_SM + j], gq = 0.0;
                    for (int l = q + 1; l < s; l++) {
                        const double rql = (l < 64) ? rdlane(r0, l) : rdlane(r1, l - 64), xl = Vl[l * WB_SM + j];
                        v = fma(-rql, xl, v);
                        gq = fma(fabs(rql), fabs(xl), gq);
                    }
                    const double d = (q < 64) ? rdlane(r0, q) : rdlane(r1, q - 64);
                    const bool ok = fabs(d) > 1e-300;
                    const double x = ok ? v / d : 0.0;
                    if (!ok && j < nr) bad = 1;
                    gmax = fmax(gmax, fma(fabs(d), fabs(x), gq));
                    if (j < nr) Vl[q * WB_SM + j] = x;
                }
                if (j < nr && !(sg[WB_SM + j] * s * 2.3e-16 * gmax <= tol)) bad = 1;
            }
            if (__syncthreads_or(bad)) {
                if (tid == 0) tasks_s[e] = RC_REDO;
                __syncthreads();
                continue;
            }
            // ---- outputs in place into the tile, ordered as in k_rc_wave (each
            //      wave reads and writes only its own row tiles) ----
            auto v_out = [&]() {  // Kv in LDS
                wv_tall_mm<WB_SM>(NB, s, nr, 1.0, WvSrc{C + (size_t)rc * NB, rc, Qs, NB}, Vl, WB_SM,
                                  C + (size_t)nr * NB, wid, WB_WAVES);
                __syncthreads();
            };
            auto u_out = [&]() {  // stages Ku into LDS
                for (int q = tid; q < s * WB_SM; q += NTH) {
                    const int c = q / WB_SM, j = q - c * WB_SM;
                    Vl[q] = (j < nr) ? Ku[c * WB_SM + j] : 0.0;
                }
                __syncthreads();
                wv_tall_mm<WB_SM>(NB, s, nr, 1.0, WvSrc{C, rc, Au, NB}, Vl, WB_SM, C, wid, WB_WAVES);
                __syncthreads();
            };
            if (nr > rc) {
                v_out();
                u_out();
            } else {
                for (int q = tid; q < s * WB_SM; q += NTH) Out[q] = Vl[q];  // Kv aside while Ku uses the LDS
                u_out();
                for (int q = tid; q < s * WB_SM; q += NTH) Vl[q] = Out[q];
                __syncthreads();
                v_out();
            }
        }
        if (tid == 0) {
            *rank_at(t, m, n) = nr;
            rank_hist(o, NT, m, n, nr);
            unsigned long long mm = tdim(t, m), ss = s;
            unsigned long long qa = opc('q', mm, ss, 0, 0);
            // case iv (dense B) counts the product B Av, case v the products with B's factors
            unsigned long long qb = bdense ? opc('q', mm, ss, 0, 0) + opc('m', ra, mm, mm, 0)
                                           : opc('m', ra, rb, mm, 0) + opc('m', ra, mm, rb, 0) + opc('q', mm, ss, 0, 0);
            unsigned long long sv = opc('r', ss, ss, 2, 0) + opc('s', ss, 0, 0, 0) + ss * nr;
            unsigned long long nu = opc('o', mm, nr, ss, 1), nv = opc('o', nr, mm, ss, 2);
            oplog(o, NT, qa + qb + sv + nu + nv, m, n, false, false);
            atomicAdd(o.stat + 4, qa + qb + sv + nu + nv);
            atomicAdd(o.stat + 5, 1ULL);
            unsigned long long words =
                2ULL * NB * (ra + rc + nr) + (bdense ? (unsigned long long)NB * NB : 2ULL * NB * rb);
            atomicAdd(o.stat + 6, 8ULL * words);
        }
        __syncthreads();
    }
}
