// rc_fast.hpp -- recompression of a low-rank gemm update by one 1024-thread
// workgroup, for tiles of at most 1024 rows and stacks of at most 64 columns.
//
// Layout: thread i owns row i of every tall (NB-row) operand; tall data stays
// in global memory and is touched row-locally, so only reductions over rows
// need the LDS.  Column reductions use a wave reduce-scatter (a power-of-two
// number of values split across the 64 lanes) followed by a 16-wave sum.
//
// Algorithm (same mathematics as HCORE's QR + SVD recompression):
//   P = [Cu Au], Q = [Cv Qn];  R factors Rp, Rq by Householder QR;
//   M = Rp Rq^T = Us S Vs^T (one-sided Jacobi);  keep r' columns;
//   U' = Qp Us_r S_r = P (Rq^T Vs_r),   V' = Qq Vs_r = Q (Rq^-1 Vs_r).
// The identities follow from P = Qp Rp, Q = Qq Rq and M Vs_r = Us_r S_r, so
// neither Householder basis is ever applied.  The solve with Rq is a plain
// back substitution with an a-posteriori error bound (see k_rc_wave); a task
// that fails it is handed to k_lr_gemm.
#pragma once
#include "tlr_kernels.hpp"

#define FT 1024
#define FW (FT / 64)
#define FS 64     // max stack width s
#define FPB 16    // panel width

// ---- reductions -----------------------------------------------------------

// block sums of p[0..16) into out[0..16) (LDS); red holds FW*16 doubles.
// wrs16 leaves the wave total of value l>>2 on lane l.
__device__ __forceinline__ void f_vsum16(double (&p)[16], double *red, double *out)
{
    const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
    double v = wrs16(p);
    if ((lane & 3) == 0) red[w * 16 + (lane >> 2)] = v;
    __syncthreads();
    if (tid < 16) {
        double s = 0.0;
#pragma unroll
        for (int i = 0; i < FW; i++) s += red[i * 16 + tid];
        out[tid] = s;
    }
    __syncthreads();
}

// block sum of one value; every thread gets it (fixed order)
__device__ __forceinline__ double f_sum(double v, double *red1)
{
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    v = wave_sum(v);
    if (lane == 0) red1[w] = v;
    __syncthreads();
    double s = 0.0;
#pragma unroll
    for (int i = 0; i < FW; i++) s += red1[i];
    return s;
}

// ---- R factor of a tall matrix --------------------------------------------
// Householder QR of A (m x n, m <= FT, n <= FS, ld lda, global, overwritten)
// with panels of FPB columns held in registers (one row per thread) and
// compact-WY trailing updates; the upper triangle R (n x n) is written to
// R (ld ldr, LDS).  LDS scratch sc: FW*FPB + 4*FPB + FPB*FPB + FW + 8 doubles.
__device__ __forceinline__ void f_geqrf_r(int m, int n, double *A, int lda, double *R, int ldr, double *sc)
{
    const int i = threadIdx.x;
    const bool own = i < m;
    double *red = sc;                 // FW*FPB
    double *dv = red + FW * FPB;      // FPB reduced values
    double *wv = dv + FPB;            // FPB (T^T W)
    double *Ts = wv + FPB;            // FPB*FPB, T of the panel
    double *red1 = Ts + FPB * FPB;    // FW
    double *bc = red1 + FW;           // broadcast slots
    for (int j0 = 0; j0 < n; j0 += FPB) {
        const int jb = (n - j0) < FPB ? (n - j0) : FPB;
        double a[FPB];
#pragma unroll
        for (int c = 0; c < FPB; c++) a[c] = (own && c < jb) ? A[i + (size_t)(j0 + c) * lda] : 0.0;
        for (int jj = 0; jj < jb; jj++) {
            const int j = j0 + jj;
            double x = 0.0;
#pragma unroll
            for (int c = 0; c < FPB; c++) x = (c == jj) ? a[c] : x;
            if (i == j) bc[0] = x;
            const double xn2 = f_sum((i > j) ? x * x : 0.0, red1);
            const double alpha = bc[0];
            double tau = 0.0, beta = alpha, scal = 0.0;
            if (xn2 > 0.0) {
                beta = -copysign(sqrt(alpha * alpha + xn2), alpha);
                tau = (beta - alpha) / beta;
                scal = 1.0 / (alpha - beta);
            }
            const double v = (i > j) ? x * scal : (i == j ? 1.0 : 0.0);
            const double nx = (i > j) ? v : (i == j ? beta : x);
#pragma unroll
            for (int c = 0; c < FPB; c++) a[c] = (c == jj) ? nx : a[c];
            double p[FPB];
#pragma unroll
            for (int c = 0; c < FPB; c++) p[c] = (c != jj) ? v * a[c] : 0.0;
            f_vsum16(p, red, dv);
#pragma unroll
            for (int c = 0; c < FPB; c++)
                if (c > jj) a[c] -= tau * dv[c] * v;
            // column jj of T: T[0:jj, jj] = -tau T[0:jj, 0:jj] (V^T v_j)
            if (i < jj) {
                double s = 0.0;
                for (int l = i; l < jj; l++) s += Ts[i + l * FPB] * dv[l];
                Ts[i + jj * FPB] = -tau * s;
            }
            if (i == 0) Ts[jj + jj * FPB] = tau;
        }
        __syncthreads();
        // rows of R held by this panel: rows 0 .. j0+jb-1, columns j0 .. j0+jb-1
        if (i < j0 + jb) {
#pragma unroll
            for (int c = 0; c < FPB; c++)
                if (c < jb && i <= j0 + c) R[i + (size_t)(j0 + c) * ldr] = a[c];
        }
        // Householder vectors of this panel, row i
        double vr[FPB];
#pragma unroll
        for (int c = 0; c < FPB; c++) {
            const int gc = j0 + c;
            vr[c] = (c < jb) ? ((i > gc) ? a[c] : (i == gc ? 1.0 : 0.0)) : 0.0;
        }
        // trailing update, one column at a time: a2 -= V (T^T (V^T a2))
        for (int t = j0 + jb; t < n; t++) {
            double a2 = own ? A[i + (size_t)t * lda] : 0.0;
            double p[FPB];
#pragma unroll
            for (int c = 0; c < FPB; c++) p[c] = vr[c] * a2;
            f_vsum16(p, red, dv);
            if (i < jb) {
                double s = 0.0;
                for (int l = 0; l <= i; l++) s += Ts[l + i * FPB] * dv[l];
                wv[i] = s;
            }
            __syncthreads();
            double s = 0.0;
#pragma unroll
            for (int c = 0; c < FPB; c++) s += vr[c] * wv[c];
            a2 -= s;
            if (own) A[i + (size_t)t * lda] = a2;
        }
        __syncthreads();
    }
}

// ---- one-sided Jacobi on an n x n matrix in LDS ---------------------------
// returns the number of sweeps (BLK_JACOBI_SWEEPS: not converged)
__device__ __forceinline__ int f_jacobi(int n, double *M, int ldm, double *Vs, int ldv, double *sigma, int *flag,
                         double *red1)
{
    int sweep = 0;
    const int tid = threadIdx.x;
    for (int e = tid; e < n * n; e += FT) {
        int r = e % n, c = e / n;
        Vs[r + c * ldv] = (r == c) ? 1.0 : 0.0;
    }
    __syncthreads();
    if (n >= 2) {
        const int npad = n + (n & 1), npairs = npad / 2;
        constexpr int g = 8;  // lanes per column pair (n <= FS: at most 8 rows per lane)
        const double tol = 4.0e-16 * sqrt((double)n);
        for (; sweep < BLK_JACOBI_SWEEPS; sweep++) {
            if (tid == 0) *flag = 0;
            __syncthreads();
            int rotated = 0;
            for (int r = 0; r < npad - 1; r++) {
                const int k = tid / g, lane = tid % g;
                int a = -1, b = -1;
                if (k < npairs) {
                    int ia = k, ib = npad - 1 - k;
                    a = (ia == 0) ? 0 : ((ia - 1 + r) % (npad - 1)) + 1;
                    b = (ib == 0) ? 0 : ((ib - 1 + r) % (npad - 1)) + 1;
                    if (a >= n || b >= n) a = -1;
                }
                double al = 0, be = 0, ga = 0;
                if (a >= 0) {
                    for (int q = lane; q < n; q += g) {
                        double x = M[q + a * ldm], y = M[q + b * ldm];
                        al = fma(x, x, al);
                        be = fma(y, y, be);
                        ga = fma(x, y, ga);
                    }
                }
                al = gsum<g>(al);
                be = gsum<g>(be);
                ga = gsum<g>(ga);
                double cs, sn;
                if (a >= 0 && jacobi_rot(al, be, ga, tol, cs, sn)) {
                    for (int q = lane; q < n; q += g) {
                        double x = M[q + a * ldm], y = M[q + b * ldm];
                        M[q + a * ldm] = cs * x - sn * y;
                        M[q + b * ldm] = sn * x + cs * y;
                        x = Vs[q + a * ldv];
                        y = Vs[q + b * ldv];
                        Vs[q + a * ldv] = cs * x - sn * y;
                        Vs[q + b * ldv] = sn * x + cs * y;
                    }
                    rotated = 1;
                }
                __syncthreads();
            }
            if (rotated) *flag = 1;
            __syncthreads();
            const int any = *flag;
            __syncthreads();
            if (!any) break;
        }
    }
    // column norms, one thread per column
    (void)red1;
    if (tid < n) {
        double p = 0.0;
        for (int q = 0; q < n; q++) p = fma(M[q + tid * ldm], M[q + tid * ldm], p);
        sigma[tid] = sqrt(p);
    }
    __syncthreads();
    return sweep;
}

__host__ __device__ inline size_t rcf_slot_doubles(int NB) { return (size_t)NB * FS * 5 + 64; }

// LDS layout (doubles): R1, R2, M, Vs (FS*FS each), scratch, sigma, pos, misc
#define RCF_SC (FW * FPB + 4 * FPB + FPB * FPB + FW + 8)
#define RCF_LDS (4 * FS * FS + RCF_SC + FS + FS / 2 + 16)

// Per-task path decision for step k: s = rc + ra for case (iv)/(v) tasks that
// do work, -1 for skipped or non-low-rank-result tasks.
__global__ void k_classify(TLR t, int k, int *tasks_s)
{
    const int T = t.NT - 1 - k;
    const long long ntask = (long long)T * (T - 1) / 2;
    long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= ntask) return;
    int hi, lo;
    tri_decode(e, hi, lo);
    const int m = k + 1 + hi, n = k + 1 + lo, b = t.band_size;
    int s = -1;
    if (m - n >= b) {
        const bool bdense = (n - k < b);
        const int ra = *rank_at(t, m, k), rc = *rank_at(t, m, n);
        const int rb = bdense ? 0 : *rank_at(t, n, k);
        if (bdense || (ra > 0 && rb > 0)) s = rc + ra;
        // fast path also needs rb <= FS (X fits the LDS)
        if (!bdense && s >= 0 && s <= FS && rb > FS) s = FS + 1;
    }
    tasks_s[e] = s;
}

// fast recompression: tasks with 0 <= s <= FS and NB <= FT
__global__ void __launch_bounds__(FT) k_lr_gemm_fast(TLR t, int k, int *tasks_s, double *ws, size_t stride,
                                                     double acc, int fixed_rk, int maxrank, OpLog o, int *ovf)
{
    extern __shared__ double sm[];
    double *R1 = sm, *R2 = sm + FS * FS, *M = sm + 2 * FS * FS, *Vs = sm + 3 * FS * FS;
    double *sc = sm + 4 * FS * FS;
    double *sig = sc + RCF_SC;
    int *pos = (int *)(sig + FS);
    int *flag = pos + FS;
    int *nrs = flag + 1;
    const int NB = t.NB, NT = t.NT, b = t.band_size, tid = threadIdx.x;
    const int i = tid;
    const bool own = i < NB;
    double *P = ws + (size_t)blockIdx.x * stride;
    double *Q = P + (size_t)NB * FS;
    double *Qn = Q + (size_t)NB * FS;
    double *Out = Qn + (size_t)NB * FS;
    const int T = NT - 1 - k;
    const long long ntask = (long long)T * (T - 1) / 2;
    for (long long task = blockIdx.x; task < ntask; task += gridDim.x) {
        const int s = tasks_s[task];
        if (s < 0 || s > FS) continue;
        int hi, lo;
        tri_decode(task, hi, lo);
        const int m = k + 1 + hi, n = k + 1 + lo;
        const bool bdense = (n - k < b);
        const int ra = *rank_at(t, m, k), rc = *rank_at(t, m, n);
        const int rb = bdense ? 0 : *rank_at(t, n, k);
        const double *A = lr_tile(t, m, k), *Au = A, *Av = A + (size_t)NB * ra;
        double *C = lr_tile(t, m, n);
        // ---- Qn (NB x ra) ----
        if (!bdense) {
            const double *B = lr_tile(t, n, k), *Bv = B + (size_t)NB * rb;
            // X = Av^T Bv (ra x rb) into R1 (ld FS); rows staged 32 at a time in M
            const int nx = ra * rb;
            double xs[4] = {0.0, 0.0, 0.0, 0.0};
            for (int r0 = 0; r0 < NB; r0 += 32) {
                const int nr = (NB - r0) < 32 ? (NB - r0) : 32;
                for (int e = tid; e < 32 * (ra + rb); e += FT) {
                    int rr = e % 32, c = e / 32;
                    double v = 0.0;
                    if (rr < nr) v = (c < ra) ? Av[r0 + rr + (size_t)c * NB] : Bv[r0 + rr + (size_t)(c - ra) * NB];
                    M[rr + c * 32] = v;
                }
                __syncthreads();
#pragma unroll
                for (int q = 0; q < 4; q++) {
                    int e = tid + q * FT;
                    if (e < nx) {
                        int aa = e % ra, bb = e / ra;
                        double s2 = xs[q];
                        for (int rr = 0; rr < 32; rr++) s2 = fma(M[rr + aa * 32], M[rr + (ra + bb) * 32], s2);
                        xs[q] = s2;
                    }
                }
                __syncthreads();
            }
#pragma unroll
            for (int q = 0; q < 4; q++) {
                int e = tid + q * FT;
                if (e < nx) R1[(e % ra) + (e / ra) * FS] = xs[q];
            }
            __syncthreads();
            // Qn[i][a] = -sum_b Bu[i][b] X[a][b]
            if (own) {
                for (int a0 = 0; a0 < ra; a0 += 16) {
                    double acc16[16];
#pragma unroll
                    for (int q = 0; q < 16; q++) acc16[q] = 0.0;
                    for (int bb = 0; bb < rb; bb++) {
                        double bu = B[i + (size_t)bb * NB];
#pragma unroll
                        for (int q = 0; q < 16; q++)
                            if (a0 + q < ra) acc16[q] = fma(bu, R1[(a0 + q) + bb * FS], acc16[q]);
                    }
#pragma unroll
                    for (int q = 0; q < 16; q++)
                        if (a0 + q < ra) Qn[i + (size_t)(a0 + q) * NB] = -acc16[q];
                }
            }
        } else {
            // Qn = -B Av with B the dense band tile (n,k); Av rows staged in M
            const double *Bd = band_tile(t, n, k);
            for (int a0 = 0; a0 < ra; a0 += 16) {
                double acc16[16];
#pragma unroll
                for (int q = 0; q < 16; q++) acc16[q] = 0.0;
                for (int l0 = 0; l0 < NB; l0 += 64) {
                    const int nl = (NB - l0) < 64 ? (NB - l0) : 64;
                    for (int e = tid; e < 64 * 16; e += FT) {
                        int ll = e % 64, q = e / 64;
                        M[e] = (ll < nl && a0 + q < ra) ? Av[l0 + ll + (size_t)(a0 + q) * NB] : 0.0;
                    }
                    __syncthreads();
                    if (own) {
                        for (int ll = 0; ll < nl; ll++) {
                            double bv = Bd[i + (size_t)(l0 + ll) * NB];
#pragma unroll
                            for (int q = 0; q < 16; q++) acc16[q] = fma(bv, M[ll + q * 64], acc16[q]);
                        }
                    }
                    __syncthreads();
                }
                if (own) {
#pragma unroll
                    for (int q = 0; q < 16; q++)
                        if (a0 + q < ra) Qn[i + (size_t)(a0 + q) * NB] = -acc16[q];
                }
            }
        }
        // ---- stacks P = [Cu Au], Q = [Cv Qn] (copies that QR overwrites) ----
        if (own) {
            for (int c = 0; c < rc; c++) {
                P[i + (size_t)c * NB] = C[i + (size_t)c * NB];
                Q[i + (size_t)c * NB] = C[i + (size_t)(rc + c) * NB];
            }
            for (int c = 0; c < ra; c++) {
                P[i + (size_t)(rc + c) * NB] = Au[i + (size_t)c * NB];
                Q[i + (size_t)(rc + c) * NB] = Qn[i + (size_t)c * NB];
            }
        }
        __syncthreads();
        int nr = 0;
        if (s > 0) {
            for (int side = 0; side < 2; side++) f_geqrf_r(NB, s, side ? Q : P, NB, side ? R2 : R1, FS, sc);
            __syncthreads();
            // M = Rp Rq^T (upper triangles)
            for (int e = tid; e < s * s; e += FT) {
                int r = e % s, c = e / s, l0 = r > c ? r : c;
                double v = 0.0;
                for (int l = l0; l < s; l++) v = fma(R1[r + l * FS], R2[c + l * FS], v);
                M[r + c * FS] = v;
            }
            __syncthreads();
            if (f_jacobi(s, M, FS, Vs, FS, sig, flag, sc + FW * FPB + 2 * FPB + FPB * FPB) >= BLK_JACOBI_SWEEPS &&
                tid == 0)
                atomicAdd(o.unconv, 1);
            for (int j = tid; j < s; j += FT) {
                double sj = sig[j];
                int p = 0;
                for (int q = 0; q < s; q++) p += (sig[q] > sj) || (sig[q] == sj && q < j);
                pos[j] = p;
            }
            __syncthreads();
            if (tid == 0) {
                int c = 0;
                if (fixed_rk > 0) c = fixed_rk < s ? fixed_rk : s;
                else
                    for (int j = 0; j < s; j++) c += (sig[j] >= acc);
                if (c > maxrank) {
                    atomicAdd(ovf, 1);
                    c = maxrank;
                }
                *nrs = c;
            }
            __syncthreads();
            nr = *nrs;
            // Ku = Rq^T Vs_sel (s x nr) into R2's columns once Rq is no longer
            // read; Kv = Rq^-1 Vs_sel by back substitution into R1's columns
            // (Rp is no longer needed), stored transposed: K[pos + c * FS]
            {
                double ku[4];
                int cnt = 0;
                for (int e = tid; e < s * s && cnt < 4; e += FT, cnt++) {
                    int c = e % s, j = e / s;  // output row c, source column j
                    ku[cnt] = 0.0;
                    if (pos[j] < nr) {
                        double su = 0.0;
                        for (int l = 0; l <= c; l++) su = fma(R2[l + c * FS], Vs[l + j * FS], su);
                        ku[cnt] = su;
                    }
                }
                int bad = 0;
                if (tid < s && pos[tid] < nr) {
                    const int j = tid, pj = pos[tid];
                    double fro2 = 0.0;
                    for (int q = 0; q < s; q++) fro2 = fma(sig[q], sig[q], fro2);
                    double gmax = 0.0;
                    for (int q = s - 1; q >= 0; q--) {
                        double v = Vs[q + j * FS], gq = 0.0;
                        for (int l = q + 1; l < s; l++) {
                            const double rql = R2[q + l * FS], xl = R1[pj + l * FS];
                            v = fma(-rql, xl, v);
                            gq = fma(fabs(rql), fabs(xl), gq);
                        }
                        const double d = R2[q + q * FS];
                        const bool ok = fabs(d) > 1e-300;
                        const double x = ok ? v / d : 0.0;
                        if (!ok) bad = 1;
                        gmax = fmax(gmax, fma(fabs(d), fabs(x), gq));
                        R1[pj + q * FS] = x;
                    }
                    if (!(sig[j] * s * 2.3e-16 * gmax <= rc_solve_tol(acc, fixed_rk, sqrt(fro2)))) bad = 1;
                }
                if (__syncthreads_or(bad)) {
                    // C untouched: the Householder path redoes the task
                    if (tid == 0) tasks_s[task] = RC_REDO;
                    __syncthreads();
                    continue;
                }
                cnt = 0;
                for (int e = tid; e < s * s && cnt < 4; e += FT, cnt++) {
                    int c = e % s, j = e / s;
                    // row c, output column pos[j], stored transposed in column c
                    if (pos[j] < nr) R2[pos[j] + c * FS] = ku[cnt];
                }
                __syncthreads();
            }
            // Out = [P Ku | Q Kv], P/Q rows from the original tiles
            if (own && nr > 0) {
                for (int side = 0; side < 2; side++) {
                    const double *K = side ? R1 : R2;
                    for (int b0 = 0; b0 < nr; b0 += 16) {
                        double au[16];
#pragma unroll
                        for (int q = 0; q < 16; q++) au[q] = 0.0;
                        for (int c = 0; c < s; c++) {
                            double pc;
                            if (side == 0) pc = (c < rc) ? C[i + (size_t)c * NB] : Au[i + (size_t)(c - rc) * NB];
                            else pc = (c < rc) ? C[i + (size_t)(rc + c) * NB] : Qn[i + (size_t)(c - rc) * NB];
#pragma unroll
                            for (int q = 0; q < 16; q++) au[q] = fma(pc, K[(b0 + q) + c * FS], au[q]);
                        }
#pragma unroll
                        for (int q = 0; q < 16; q++)
                            if (b0 + q < nr) Out[i + (size_t)(side * nr + b0 + q) * NB] = au[q];
                    }
                }
                for (int c = 0; c < 2 * nr; c++) C[i + (size_t)c * NB] = Out[i + (size_t)c * NB];
            }
        }
        if (tid == 0) {
            *rank_at(t, m, n) = nr;
            rank_hist(o, NT, m, n, nr);
            unsigned long long mm = tdim(t, m), ss = s;
            unsigned long long qa = opc('q', mm, ss, 0, 0), qb;
            if (bdense) qb = opc('q', mm, ss, 0, 0) + opc('m', ra, mm, mm, 0);
            else qb = opc('m', ra, rb, mm, 0) + opc('m', ra, mm, rb, 0) + opc('q', mm, ss, 0, 0);
            unsigned long long sv = opc('r', ss, ss, 2, 0) + opc('s', ss, 0, 0, 0) + ss * nr;
            unsigned long long nu = opc('o', mm, nr, ss, 1), nv = opc('o', nr, mm, ss, 2);
            oplog(o, NT, qa + qb + sv + nu + nv, m, n, false, false);
            atomicAdd(o.stat + 4, qa + qb + sv + nu + nv);
            atomicAdd(o.stat + 5, 1ULL);
            unsigned long long words = 2ULL * NB * (ra + rc + nr) + (bdense ? (unsigned long long)NB * NB : 2ULL * NB * rb);
            atomicAdd(o.stat + 6, 8ULL * words);
        }
        __syncthreads();
    }
}
