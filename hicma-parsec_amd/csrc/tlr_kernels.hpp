// tlr_kernels.hpp -- HIP kernels of the TLR Cholesky path on gfx950.
//
// Device layout (one descriptor, everything resident in HBM):
//   band tile (m,n), m-n < band : NB x NB, ld NB, at band + ((m-n)*NT + m)*NB*NB
//   off-band tile (m,n), m > n   : slot of NB*2*cap doubles at lr + (m(m-1)/2+n)*NB*2*cap,
//                                  compact [U | V]: U = NB x r, V = NB x r at offset NB*r
//   rank map                     : int32 NT x NT column major, rank[n*NT + m]
// The matrix is padded to NT*NB with an identity block, so every tile is
// NB x NB and the factor of the padding is the identity (log 1 = 0).
#pragma once
#include "blk.hpp"

struct TLR {
    double *band;
    double *lr;
    int *rank;
    const double *xyz;
    int N, NB, NT, band_size, cap, kind, ndim;
    double sigma2, beta, nugget;
};

__device__ __forceinline__ double *band_tile(const TLR &t, int m, int n)
{
    return t.band + ((size_t)(m - n) * t.NT + m) * (size_t)t.NB * t.NB;
}
__device__ __forceinline__ double *lr_tile(const TLR &t, int m, int n)
{
    return t.lr + ((size_t)m * (m - 1) / 2 + n) * (size_t)t.NB * 2 * t.cap;
}
__device__ __forceinline__ int *rank_at(const TLR &t, int m, int n)
{
    return t.rank + (size_t)n * t.NT + m;
}
__device__ __forceinline__ int tdim(const TLR &t, int m) { return m == t.NT - 1 ? t.N - m * t.NB : t.NB; }

// covariance entry for global indices (identity on the padding)
__device__ __forceinline__ double kval(const TLR &t, int gi, int gj)
{
    if (gi >= t.N || gj >= t.N) return gi == gj ? 1.0 : 0.0;
    double d2 = 0.0;
    for (int d = 0; d < t.ndim; d++) {
        double df = t.xyz[(size_t)d * t.N + gi] - t.xyz[(size_t)d * t.N + gj];
        d2 += df * df;
    }
    double v = (t.kind == 4) ? t.sigma2 * exp(-sqrt(d2) / t.beta)
                             : t.sigma2 * exp(-d2 / (2.0 * t.beta * t.beta));
    if (gi == gj) v += t.nugget;
    return v;
}

// ---- operation counts: the reference's op_counts formulas (op_counts.h) ----
__device__ __forceinline__ unsigned long long opc(char op, unsigned long long a, unsigned long long b,
                                                  unsigned long long c, unsigned long long d)
{
    unsigned long long r = 0;
    if (op == 'q') {
        r = 2 * a * b * b - (unsigned long long)((float)(2 * b * b * b) / 3.0f) + 2 * a * b +
            (unsigned long long)((float)(17 * b) / 3.0f);
    } else if (op == 'c') {
        r = (unsigned long long)((double)(a * a * a / 3) - (double)(a * a) / 2.0 + (double)(a / 6));
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

struct OpLog {
    unsigned long long *tile;  // NT*NT, index m*NT+n
    unsigned long long *stat;  // band, offband, path, offpath
};

__device__ __forceinline__ void oplog(const OpLog &o, int NT, unsigned long long cnt, int i, int j,
                                      bool band, bool path)
{
    atomicAdd(o.tile + (size_t)i * NT + j, cnt);
    atomicAdd(o.stat + (band ? 0 : 1), cnt);
    atomicAdd(o.stat + (path ? 2 : 3), cnt);
}

// decode a linear index into (hi, lo) with 0 <= lo < hi
__device__ __forceinline__ void tri_decode(long long t, int &hi, int &lo)
{
    int h = (int)((1.0 + sqrt(1.0 + 8.0 * (double)t)) * 0.5);
    while ((long long)h * (h - 1) / 2 > t) h--;
    while ((long long)(h + 1) * h / 2 <= t) h++;
    hi = h;
    lo = (int)(t - (long long)h * (h - 1) / 2);
}

// ======================= generation =======================

// dense band tiles: grid (ceil(NB*NB/BS), band*NT)
__global__ void __launch_bounds__(BS) k_gen_band(TLR t, int band_lo)
{
    int d = blockIdx.y / t.NT + band_lo, m = blockIdx.y % t.NT, n = m - d;
    if (n < 0 || d >= t.band_size) return;
    double *T = band_tile(t, m, n);
    size_t e = (size_t)blockIdx.x * BS + threadIdx.x;
    if (e >= (size_t)t.NB * t.NB) return;
    int i = (int)(e % t.NB), j = (int)(e / t.NB);
    T[e] = kval(t, m * t.NB + i, n * t.NB + j);
}

struct GenWs {
    size_t stride;  // doubles per slot
    int p;          // sketch width
};

__host__ __device__ inline size_t gen_slot_doubles(int NB, int p, int cap)
{
    size_t nt = (size_t)(p + QRB - 1) / QRB;
    return (size_t)NB * NB + 3 * (size_t)NB * p + (size_t)NB * QRB + (size_t)QRB * (p > NB ? p : NB) +
           2 * (size_t)p + 2 * nt * QRB * QRB + 2 * (size_t)p * p + (size_t)p * cap + 64;
}

// STARS-H starsh_dense_dlrrsdd restated as a workgroup pipeline:
// D = kernel tile; Y = D*Omega; Q = qr(Y); X = D^T Q; X = Qx Rx; Jacobi SVD of Rx;
// rank by relative Frobenius tail <= tol; U = Q*Vr, V = Qx*(Ur*S).
__global__ void __launch_bounds__(BS) k_compress(TLR t, const double *Om, double *ws, GenWs g,
                                                 double tol, int maxrank, int *info)
{
    extern __shared__ double sm[];
    double *sig = sm + BLK_GEMM_SMEM + BLK_RED_SMEM + 2 * QRB * QRB;
    double *ssig = sig + g.p;
    int *pos = (int *)(ssig + g.p);
    int *flag = pos + g.p;
    int *rk = flag + 1;
    const int NB = t.NB, NT = t.NT, p = g.p, tid = threadIdx.x;
    double *D = ws + (size_t)blockIdx.x * g.stride;
    double *Y = D + (size_t)NB * NB;
    double *Qm = Y + (size_t)NB * p;
    double *X = Qm + (size_t)NB * p;
    double *bw = X + (size_t)NB * p;
    double *tauY = bw + (size_t)NB * QRB + (size_t)QRB * (p > NB ? p : NB);
    double *tauX = tauY + p;
    size_t ntp = (size_t)(p + QRB - 1) / QRB;
    double *TY = tauX + p;
    double *TX = TY + ntp * QRB * QRB;
    double *Mj = TX + ntp * QRB * QRB;
    double *Vj = Mj + (size_t)p * p;
    double *Vsel = Vj + (size_t)p * p;
    const long long ntask = (long long)NT * (NT - 1) / 2;
    for (long long task = blockIdx.x; task < ntask; task += gridDim.x) {
        int m, n;
        tri_decode(task, m, n);
        if (m - n < t.band_size) continue;
        // D
        for (size_t e = tid; e < (size_t)NB * NB; e += BS) {
            int i = (int)(e % NB), j = (int)(e / NB);
            D[e] = kval(t, m * NB + i, n * NB + j);
        }
        __syncthreads();
        blk_gemm<false, false>(NB, p, NB, 1.0, D, NB, Om, NB, 0.0, Y, NB, sm);
        blk_geqrf(NB, p, Y, NB, tauY, TY, bw, sm);
        for (int j = 0; j < p; j++)
            for (int i = tid; i < NB; i += BS) Qm[i + (size_t)j * NB] = (i == j) ? 1.0 : 0.0;
        __syncthreads();
        blk_ormqr(NB, p, p, Y, NB, TY, Qm, NB, bw, sm);
        blk_gemm<true, false>(NB, p, NB, 1.0, D, NB, Qm, NB, 0.0, X, NB, sm);
        blk_geqrf(NB, p, X, NB, tauX, TX, bw, sm);
        for (int j = 0; j < p; j++)
            for (int i = tid; i < p; i += BS) Mj[i + (size_t)j * p] = (i <= j) ? X[i + (size_t)j * NB] : 0.0;
        __syncthreads();
        blk_jacobi(p, Mj, p, Vj, p, sig, flag, sm + BLK_GEMM_SMEM);
        blk_sort_pos(p, sig, pos);
        for (int j = tid; j < p; j += BS) ssig[pos[j]] = sig[j];
        __syncthreads();
        if (tid == 0) {
            double tot = 0.0, acc = 0.0;
            for (int i = 0; i < p; i++) tot += ssig[i] * ssig[i];
            double stop = tot * tol * tol;
            int r = 0;
            for (int i = p - 1; i >= 0; i--) {
                acc += ssig[i] * ssig[i];
                if (acc > stop) { r = i + 1; break; }
            }
            int rows = tdim(t, m), cols = tdim(t, n);
            int mn = rows < cols ? rows : cols;
            if (!(r < mn / 2 && r <= maxrank)) r = -1;
            *rk = r;
        }
        __syncthreads();
        const int r = *rk;
        if (r < 0) {
            if (tid == 0) { *rank_at(t, m, n) = -1; atomicExch(info, -1); }
            __syncthreads();
            continue;
        }
        double *U = lr_tile(t, m, n), *V = U + (size_t)NB * r;
        // Vsel (p x r): right singular vectors of Rx in descending order
        for (int j = 0; j < p; j++) {
            int q = pos[j];
            if (q >= r) continue;
            for (int i = tid; i < p; i += BS) Vsel[i + (size_t)q * p] = Vj[i + (size_t)j * p];
        }
        // V (NB x r) <- [Ur*S ; 0]
        for (int j = 0; j < p; j++) {
            int q = pos[j];
            if (q >= r) continue;
            for (int i = tid; i < NB; i += BS) V[i + (size_t)q * NB] = (i < p) ? Mj[i + (size_t)j * p] : 0.0;
        }
        __syncthreads();
        blk_gemm<false, false>(NB, r, p, 1.0, Qm, NB, Vsel, p, 0.0, U, NB, sm);
        blk_ormqr(NB, r, p, X, NB, TX, V, NB, bw, sm);
        if (tid == 0) *rank_at(t, m, n) = r;
        __syncthreads();
    }
}

// ======================= dense potrf of a diagonal tile =======================

// trailing lower update S -= P P^T, grid = lower GT tiles of the trailing block
__global__ void __launch_bounds__(BS) k_potrf_update(TLR t, int k, int j0, int jb)
{
    __shared__ double sm[BLK_GEMM_SMEM];
    const int NB = t.NB, j1 = j0 + jb, R = NB - j1;
    int ti, tj;
    tri_decode((long long)blockIdx.x, ti, tj);  // ti > tj
    // map: linear over tiles with ti >= tj; use tri_decode on (ti+1)
    ti -= 1;
    double *A = band_tile(t, k, k);
    const double *P = A + j1 + (size_t)j0 * NB;
    double *S = A + j1 + (size_t)j1 * NB;
    gemm_tile<false, true>(ti * GT, tj * GT, R, R, jb, -1.0, P, NB, P, NB, 1.0, S, NB, sm, true);
}

// ---- LDS-resident diagonal block: Cholesky + inverse ----
// Factors the jb x jb block at A(j0,j0) of tile (k,k) in LDS, writes L back
// and W = L^{-1} (ld PBK) to Winv[j0/PBK].  One workgroup of BS threads.
#define PBK 64
#define PLD (PBK + 1)
__global__ void __launch_bounds__(BS) k_potrf_diag2(TLR t, int k, int j0, int jb, double *Winv, int *info, OpLog o)
{
    // dpotrf of the whole tile counts once, on the band and on the critical path
    if (j0 == 0 && threadIdx.x == 0) oplog(o, t.NT, opc('c', tdim(t, k), 0, 0, 0), k, k, true, true);
    __shared__ double L[PBK * PLD];
    __shared__ double W[PBK * PLD];
    __shared__ int bad;
    const int NB = t.NB, tid = threadIdx.x;
    double *A = band_tile(t, k, k) + j0 + (size_t)j0 * NB;
    for (int e = tid; e < jb * jb; e += BS) {
        int i = e % jb, c = e / jb;
        L[i + c * PLD] = (i >= c) ? A[i + (size_t)c * NB] : 0.0;
    }
    if (tid == 0) bad = 0;
    __syncthreads();
    for (int j = 0; j < jb; j++) {
        const double d2 = L[j + j * PLD];
        if (!(d2 > 0.0)) {
            if (tid == 0) { atomicCAS(info, 0, k * NB + j0 + j + 1); bad = 1; }
            break;
        }
        const double d = sqrt(d2), rd = 1.0 / d;
        __syncthreads();
        if (tid == 0) L[j + j * PLD] = d;
        for (int i = j + 1 + tid; i < jb; i += BS) L[i + j * PLD] *= rd;
        __syncthreads();
        const int r = jb - j - 1;
        for (int e = tid; e < r * r; e += BS) {
            int l = j + 1 + e % r, c = j + 1 + e / r;
            if (l >= c) L[l + c * PLD] -= L[l + j * PLD] * L[c + j * PLD];
        }
        __syncthreads();
    }
    __syncthreads();
    if (bad) return;
    // W = L^{-1}: 4 adjacent lanes per column, rows in sequence; lane q keeps
    // the rows i with i % 4 == q of its column (same-lane LDS reuse)
    {
        const int c = tid >> 2, q = tid & 3;
        const bool act = c < jb;
        for (int i = 0; i < jb; i++) {
            double s = 0.0;
            if (act && i > c)
                for (int l = c + ((q - c) & 3); l < i; l += 4) s = fma(L[i + l * PLD], W[l + c * PLD], s);
            s = gsum<4>(s);
            if (act && (i & 3) == q) {
                double w = (i < c) ? 0.0 : (i == c ? 1.0 / L[c + c * PLD] : -s / L[i + i * PLD]);
                W[i + c * PLD] = w;
            }
        }
    }
    __syncthreads();
    double *Wo = Winv + (size_t)(j0 / PBK) * PBK * PBK;
    for (int e = tid; e < jb * jb; e += BS) {
        int i = e % jb, c = e / jb;
        if (i >= c) A[i + (size_t)c * NB] = L[i + c * PLD];
        Wo[i + c * PBK] = (i >= c) ? W[i + c * PLD] : 0.0;
    }
}

// Same contract as k_potrf_diag2, one wave: left-looking Cholesky with the
// block row-major in LDS (lane i owns row i; the pivot row is read as an LDS
// broadcast and the pivot itself by readlane), then W = L^{-1} with lane c
// owning column c (forward substitution, L rows read as broadcasts).
#define PLR (PBK + 2)
__global__ void __launch_bounds__(64) k_potrf_diag3(TLR t, int k, int j0, int jb, double *Winv, int *info, OpLog o)
{
    if (j0 == 0 && threadIdx.x == 0) oplog(o, t.NT, opc('c', tdim(t, k), 0, 0, 0), k, k, true, true);
    __shared__ double Lr[PBK * PLR];  // Lr[i * PLR + c] = L(i, c)
    __shared__ double Wc[PBK * PLR];  // Wc[c * PLR + i] = W(i, c)
    const int NB = t.NB, lane = threadIdx.x;
    double *A = band_tile(t, k, k) + j0 + (size_t)j0 * NB;
    {
        // all 64 loads in flight before the LDS stores
        double row[PBK];
#pragma unroll
        for (int c = 0; c < PBK; c++) row[c] = (c < jb && lane < jb && lane >= c) ? A[lane + (size_t)c * NB] : 0.0;
#pragma unroll
        for (int c = 0; c < PBK; c++) Lr[lane * PLR + c] = row[c];
    }
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    // right-looking Cholesky, lane = row: a[c] = A(lane, c) is updated in
    // registers; column j of L is broadcast through LDS (Wc[0 .. 63] as a
    // buffer), so each step issues up to 63 independent FMAs
    double a[PBK];
#pragma unroll
    for (int c = 0; c < PBK; c++) a[c] = Lr[lane * PLR + c];
    double *colj = Wc;
#pragma unroll
    for (int j = 0; j < PBK; j++) {
        if (j < jb) {
            const double d2 = mk64((unsigned)__builtin_amdgcn_readlane((int)lo32(a[j]), j),
                                   (unsigned)__builtin_amdgcn_readlane((int)hi32(a[j]), j));
            if (!(d2 > 0.0)) {
                if (lane == 0) atomicCAS(info, 0, k * NB + j0 + j + 1);
                return;
            }
            const double d = sqrt(d2);
            const double lij = (lane == j) ? d : ((lane > j && lane < jb) ? a[j] / d : 0.0);
            a[j] = lij;
            colj[lane] = lij;
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
            __builtin_amdgcn_wave_barrier();
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
#pragma unroll
            for (int c = j + 1; c < PBK; c++) a[c] = fma(-lij, colj[c], a[c]);
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
            __builtin_amdgcn_wave_barrier();
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        }
    }
    // L (lower part of the rows) back to LDS
#pragma unroll
    for (int c = 0; c < PBK; c++) Lr[lane * PLR + c] = (c <= lane && lane < jb) ? a[c] : 0.0;
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    // W = L^-1, right-looking, lane = column c: w(l) = x(l) / L(l, l), then
    // x(i) -= L(i, l) w(l) for i > l (L entries are broadcast LDS reads)
    {
#pragma unroll
        for (int i = 0; i < PBK; i++) a[i] = (i == lane) ? 1.0 : 0.0;
#pragma unroll
        for (int l = 0; l < PBK; l++) {
            if (l < jb) {
                const double w = (lane <= l) ? a[l] / Lr[l * PLR + l] : 0.0;
                a[l] = w;
#pragma unroll
                for (int i = l + 1; i < PBK; i++) a[i] = fma(-Lr[i * PLR + l], w, a[i]);
            }
        }
#pragma unroll
        for (int i = 0; i < PBK; i++) Wc[lane * PLR + i] = (i < jb && lane < jb) ? a[i] : 0.0;
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    }
    double *Wo = Winv + (size_t)(j0 / PBK) * PBK * PBK;
#pragma unroll 8
    for (int c = 0; c < jb; c++) {
        if (lane < jb && lane >= c) A[lane + (size_t)c * NB] = Lr[lane * PLR + c];
        if (lane < jb) Wo[lane + c * PBK] = Wc[c * PLR + lane];
    }
}

// panel rows below the diagonal block: X = C W^T (W = L^{-1}), in place;
// grid = ceil(rows / PBK)
__global__ void __launch_bounds__(BS) k_potrf_panel2(TLR t, int k, int j0, int jb, const double *Winv)
{
    __shared__ double Cs[PBK * PLD];
    __shared__ double Ws[PBK * PLD];
    const int NB = t.NB, tid = threadIdx.x;
    const int r0 = j0 + jb + blockIdx.x * PBK;
    int nr = NB - r0;
    if (nr <= 0) return;
    if (nr > PBK) nr = PBK;
    double *A = band_tile(t, k, k);
    const double *Wi = Winv + (size_t)(j0 / PBK) * PBK * PBK;
    for (int e = tid; e < PBK * jb; e += BS) {
        int i = e % PBK, c = e / PBK;
        Cs[i + c * PLD] = (i < nr) ? A[r0 + i + (size_t)(j0 + c) * NB] : 0.0;
        if (i < jb) Ws[i + c * PLD] = Wi[i + c * PBK];
    }
    __syncthreads();
    // thread: row i = tid % 64, columns c = tid / 64 + 4 q
    const int i = tid % PBK, cg = tid / PBK;
    double x[PBK / (BS / PBK)];
#pragma unroll
    for (int q = 0; q < PBK / (BS / PBK); q++) {
        const int c = cg + q * (BS / PBK);
        double s = 0.0;
        if (c < jb)
            for (int l = 0; l <= c; l++) s = fma(Cs[i + l * PLD], Ws[c + l * PLD], s);
        x[q] = s;
    }
#pragma unroll
    for (int q = 0; q < PBK / (BS / PBK); q++) {
        const int c = cg + q * (BS / PBK);
        if (c < jb && i < nr) A[r0 + i + (size_t)(j0 + c) * NB] = x[q];
    }
}

// low-rank tiles m = k+band .. NT-1: V <- L^{-1} V by left-looking block
// substitution with the inverted diagonal blocks of L; grid (ceil(cap/16), count)
__global__ void __launch_bounds__(BS) k_trsm_lr2(TLR t, int k, const double *Winv, OpLog o)
{
    __shared__ double Vb[PBK * 17];
    const int NB = t.NB, tid = threadIdx.x;
    const int m = k + t.band_size + blockIdx.y;
    const int r = *rank_at(t, m, k);
    if (blockIdx.x == 0 && tid == 0) oplog(o, t.NT, opc('t', tdim(t, m), r, 1, 0), m, k, false, m - k == 1);
    const int c0 = blockIdx.x * 16;
    if (c0 >= r) return;
    const int nc = (r - c0) < 16 ? (r - c0) : 16;
    double *V = lr_tile(t, m, k) + (size_t)NB * r + (size_t)c0 * NB;
    const double *L = band_tile(t, k, k);
    const int i = tid % PBK, cg = tid / PBK;  // 4 column groups of 4
    for (int b0 = 0; b0 < NB; b0 += PBK) {
        const int nb = (NB - b0) < PBK ? (NB - b0) : PBK;
        double acc[4];
#pragma unroll
        for (int q = 0; q < 4; q++) {
            const int c = cg * 4 + q;
            acc[q] = (i < nb && c < nc) ? V[b0 + i + (size_t)c * NB] : 0.0;
        }
        if (i < nb) {
            const double *Lr = L + b0 + i;
#pragma unroll 4
            for (int l = 0; l < b0; l++) {
                const double lv = Lr[(size_t)l * NB];
#pragma unroll
                for (int q = 0; q < 4; q++) {
                    const int c = cg * 4 + q;
                    if (c < nc) acc[q] = fma(-lv, V[l + (size_t)c * NB], acc[q]);
                }
            }
        }
#pragma unroll
        for (int q = 0; q < 4; q++) Vb[i + (cg * 4 + q) * PBK] = acc[q];
        __syncthreads();
        const double *Wb = Winv + (size_t)(b0 / PBK) * PBK * PBK;
        double x[4] = {0.0, 0.0, 0.0, 0.0};
        if (i < nb) {
            for (int l = 0; l <= i; l++) {
                const double w = Wb[i + (size_t)l * PBK];
#pragma unroll
                for (int q = 0; q < 4; q++) x[q] = fma(w, Vb[l + (cg * 4 + q) * PBK], x[q]);
            }
        }
#pragma unroll
        for (int q = 0; q < 4; q++) {
            const int c = cg * 4 + q;
            if (i < nb && c < nc) V[b0 + i + (size_t)c * NB] = x[q];
        }
        __syncthreads();
    }
}

// low-rank tiles m = k+band .. NT-1: V <- L^{-1} V, TRSM_TC columns of one tile
// per workgroup.  Row block b (64 rows) is X_b = W_b (V_b - L_b,0:b X_0:b): the
// sum runs over 64-deep chunks of L and of the solved rows staged in LDS (so
// every global load of a chunk is in flight at once), then the inverted
// diagonal block W_b is applied from LDS.  grid (ceil(cap/TRSM_TC), count).
#define TL3 (PBK + 1)
#define TRSM_TC 8                   // columns of V per workgroup
#define TRSM_XL (TRSM_TC + 1)       // LDS row stride of X
__global__ void __launch_bounds__(BS) k_trsm_lr3(TLR t, int k, const double *Winv, OpLog o)
{
    __shared__ double Ls[PBK * TL3];   // Ls[kk * TL3 + i] = L(b0 + i, l0 + kk); also W_b
    __shared__ double Xs[PBK * TRSM_XL];  // Xs[kk * TRSM_XL + c] = X(l0 + kk, c); also V_b
    const int NB = t.NB, tid = threadIdx.x;
    const int m = k + t.band_size + blockIdx.y;
    const int r = *rank_at(t, m, k);
    if (blockIdx.x == 0 && tid == 0) oplog(o, t.NT, opc('t', tdim(t, m), r, 1, 0), m, k, false, m - k == 1);
    const int c0 = blockIdx.x * TRSM_TC;
    if (c0 >= r) return;
    const int nc = (r - c0) < TRSM_TC ? (r - c0) : TRSM_TC;
    double *V = lr_tile(t, m, k) + (size_t)NB * r + (size_t)c0 * NB;
    const double *L = band_tile(t, k, k);
    constexpr int QC = TRSM_TC / (BS / PBK);  // columns per thread
    const int i = tid % PBK, cg = tid / PBK;  // row i, columns QC cg .. QC cg + QC - 1
    // software pipeline: the next chunk of L and X, and the block's W_b, are
    // loaded into registers while the current chunk is multiplied from LDS
    constexpr int LE = PBK * PBK / BS, XE = PBK * TRSM_TC / BS;
    double pl[LE], px[XE], pw[LE];
    auto load_chunk = [&](int b0, int nb, int l0) {
#pragma unroll
        for (int j = 0; j < LE; j++) {
            const int e = tid + BS * j, ii = e % PBK, kk = e / PBK;
            pl[j] = (ii < nb) ? L[b0 + ii + (size_t)(l0 + kk) * NB] : 0.0;
        }
#pragma unroll
        for (int j = 0; j < XE; j++) {
            const int e = tid + BS * j, kk = e % PBK, c = e / PBK;
            px[j] = (c < nc) ? V[l0 + kk + (size_t)c * NB] : 0.0;
        }
    };
    for (int b0 = 0; b0 < NB; b0 += PBK) {
        const int nb = (NB - b0) < PBK ? (NB - b0) : PBK;
        const double *Wb = Winv + (size_t)(b0 / PBK) * PBK * PBK;
#pragma unroll
        for (int j = 0; j < LE; j++) {
            const int e = tid + BS * j, ii = e % PBK, l = e / PBK;
            pw[j] = Wb[ii + (size_t)l * PBK];
        }
        double acc[QC];
#pragma unroll
        for (int q = 0; q < QC; q++) {
            const int c = cg * QC + q;
            acc[q] = (i < nb && c < nc) ? V[b0 + i + (size_t)c * NB] : 0.0;
        }
        if (b0 > 0) load_chunk(b0, nb, 0);
        for (int l0 = 0; l0 < b0; l0 += PBK) {
            // L(b0:b0+64, l0:l0+64) and X(l0:l0+64, 0:TRSM_TC) from registers to LDS
#pragma unroll
            for (int j = 0; j < LE; j++) {
                const int e = tid + BS * j, ii = e % PBK, kk = e / PBK;
                Ls[kk * TL3 + ii] = pl[j];
            }
#pragma unroll
            for (int j = 0; j < XE; j++) {
                const int e = tid + BS * j, kk = e % PBK, c = e / PBK;
                Xs[kk * TRSM_XL + c] = px[j];
            }
            __syncthreads();
            if (l0 + PBK < b0) load_chunk(b0, nb, l0 + PBK);
#pragma unroll 8
            for (int kk = 0; kk < PBK; kk++) {
                const double lv = Ls[kk * TL3 + i];
#pragma unroll
                for (int q = 0; q < QC; q++) acc[q] = fma(-lv, Xs[kk * TRSM_XL + cg * QC + q], acc[q]);
            }
            __syncthreads();
        }
        // X_b = W_b acc: W_b (lower triangular 64 x 64) into Ls, acc into Xs
#pragma unroll
        for (int j = 0; j < LE; j++) {
            const int e = tid + BS * j, ii = e % PBK, l = e / PBK;
            Ls[l * TL3 + ii] = pw[j];
        }
#pragma unroll
        for (int q = 0; q < QC; q++) Xs[i * TRSM_XL + cg * QC + q] = acc[q];
        __syncthreads();
        double x[QC];
#pragma unroll
        for (int q = 0; q < QC; q++) x[q] = 0.0;
        for (int l = 0; l <= i && l < nb; l++) {
            const double w = Ls[l * TL3 + i];
#pragma unroll
            for (int q = 0; q < QC; q++) x[q] = fma(w, Xs[l * TRSM_XL + cg * QC + q], x[q]);
        }
#pragma unroll
        for (int q = 0; q < QC; q++) {
            const int c = cg * QC + q;
            if (i < nb && c < nc) V[b0 + i + (size_t)c * NB] = x[q];
        }
        __syncthreads();
    }
}

// ======================= trsm =======================

// band tiles m = k+1 .. k+band-1: C <- C L^{-T}; grid (ceil(NB/GT), nb)
__global__ void __launch_bounds__(BS) k_trsm_band(TLR t, int k, OpLog o)
{
    __shared__ double sm[BLK_GEMM_SMEM];
    const int NB = t.NB;
    int m = k + 1 + blockIdx.y;
    int r0 = blockIdx.x * GT;
    int nr = NB - r0 < GT ? NB - r0 : GT;
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        int mm = tdim(t, m);
        oplog(o, t.NT, opc('t', mm, mm, 1, 0), m, k, true, m - k == 1);
    }
    blk_trsm_rltn(nr, NB, band_tile(t, k, k), NB, band_tile(t, m, k) + r0, NB, sm);
}

// ======================= syrk =======================

// band: T_mm -= C C^T for m = k+1..k+band-1; grid (lower tiles, nb)
__global__ void __launch_bounds__(BS) k_syrk_band(TLR t, int k, OpLog o)
{
    __shared__ double sm[BLK_GEMM_SMEM];
    const int NB = t.NB;
    int m = k + 1 + blockIdx.y;
    int ti, tj;
    tri_decode((long long)blockIdx.x, ti, tj);
    ti -= 1;
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        int mm = tdim(t, m);
        oplog(o, t.NT, opc('m', mm, mm, mm, 0), m, m, true, m - k == 1);
    }
    const double *C = band_tile(t, m, k);
    gemm_tile<false, true>(ti * GT, tj * GT, NB, NB, NB, -1.0, C, NB, C, NB, 1.0, band_tile(t, m, m), NB,
                           sm, true);
}

// low rank prep: Y = U (V^T V) for m = k+band..NT-1; grid (count)
__global__ void __launch_bounds__(BS) k_syrk_lr_prep(TLR t, int k, double *Yws, double *Zws, OpLog o)
{
    __shared__ double sm[BLK_GEMM_SMEM];
    const int NB = t.NB, cap = t.cap;
    int m = k + t.band_size + blockIdx.x;
    int r = *rank_at(t, m, k);
    if (threadIdx.x == 0) {
        unsigned long long mm = tdim(t, m);
        unsigned long long c = opc('m', r, r, mm, 0) + opc('m', mm, r, r, 0) + opc('m', mm, mm, r, 0);
        oplog(o, t.NT, c, m, m, true, m - k == 1);
    }
    if (r == 0) return;
    const double *U = lr_tile(t, m, k), *V = U + (size_t)NB * r;
    double *Z = Zws + (size_t)m * cap * cap, *Y = Yws + (size_t)m * NB * cap;
    blk_gemm<true, false>(r, r, NB, 1.0, V, NB, V, NB, 0.0, Z, r, sm);
    blk_gemm<false, false>(NB, r, r, 1.0, U, NB, Z, r, 0.0, Y, NB, sm);
}

// low rank update: T_mm -= Y U^T; grid (lower tiles, count)
__global__ void __launch_bounds__(BS) k_syrk_lr_update(TLR t, int k, const double *Yws)
{
    __shared__ double sm[BLK_GEMM_SMEM];
    const int NB = t.NB, cap = t.cap;
    int m = k + t.band_size + blockIdx.y;
    int r = *rank_at(t, m, k);
    if (r == 0) return;
    int ti, tj;
    tri_decode((long long)blockIdx.x, ti, tj);
    ti -= 1;
    const double *U = lr_tile(t, m, k);
    const double *Y = Yws + (size_t)m * NB * cap;
    gemm_tile<false, true>(ti * GT, tj * GT, NB, NB, r, -1.0, Y, NB, U, NB, 1.0, band_tile(t, m, m), NB, sm,
                           true);
}

// ======================= gemm, dense cases =======================

// (i) all dense: m-k < band; pairs k<n<m<=k+band-1; grid (tiles, pairs)
__global__ void __launch_bounds__(BS) k_gemm_dense(TLR t, int k, OpLog o)
{
    __shared__ double sm[BLK_GEMM_SMEM];
    const int NB = t.NB;
    int hi, lo;
    tri_decode((long long)blockIdx.y, hi, lo);
    int m = k + 1 + hi, n = k + 1 + lo;
    int nt = (NB + GT - 1) / GT;
    int ti = blockIdx.x % nt, tj = blockIdx.x / nt;
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        int mm = tdim(t, m);
        oplog(o, t.NT, opc('m', mm, mm, mm, 0), m, n, true, false);
    }
    gemm_tile<false, true>(ti * GT, tj * GT, NB, NB, NB, -1.0, band_tile(t, m, k), NB, band_tile(t, n, k), NB,
                           1.0, band_tile(t, m, n), NB, sm);
}

// (ii) A LR, B dense, C dense and (iii) A, B LR, C dense: W = Av^T B^T or
// (Av^T Bv) Bu^T (rA x NB).  Pairs: m = k+band+p/(band-1), n = m-1-p%(band-1).
__global__ void __launch_bounds__(BS) k_gemm_lrd_prep(TLR t, int k, double *Wws, double *Xws, OpLog o)
{
    __shared__ double sm[BLK_GEMM_SMEM];
    const int NB = t.NB, cap = t.cap, b = t.band_size;
    int pidx = blockIdx.x;
    int m = k + b + pidx / (b - 1), n = m - 1 - pidx % (b - 1);
    int ra = *rank_at(t, m, k);
    const double *A = lr_tile(t, m, k), *Av = A + (size_t)NB * ra;
    double *W = Wws + (size_t)pidx * cap * NB;
    int mm = tdim(t, m);
    if (n - k < b) {
        if (threadIdx.x == 0) oplog(o, t.NT, opc('m', mm, mm, ra, 0) * 2, m, n, true, false);
        if (ra == 0) return;
        blk_gemm<true, true>(ra, NB, NB, 1.0, Av, NB, band_tile(t, n, k), NB, 0.0, W, cap, sm);
    } else {
        int rb = *rank_at(t, n, k);
        if (threadIdx.x == 0)
            oplog(o, t.NT, opc('m', mm, mm, ra < rb ? ra : rb, 0) + opc('m', mm, ra, rb, 0) * 2, m, n, true,
                  false);
        if (ra == 0 || rb == 0) return;
        const double *B = lr_tile(t, n, k), *Bv = B + (size_t)NB * rb;
        double *X = Xws + (size_t)pidx * cap * cap;
        blk_gemm<true, false>(ra, rb, NB, 1.0, Av, NB, Bv, NB, 0.0, X, ra, sm);
        blk_gemm<false, true>(ra, NB, rb, 1.0, X, ra, B, NB, 0.0, W, cap, sm);
    }
}

// C -= Au W; grid (tiles, pairs)
__global__ void __launch_bounds__(BS) k_gemm_lrd_update(TLR t, int k, const double *Wws)
{
    __shared__ double sm[BLK_GEMM_SMEM];
    const int NB = t.NB, cap = t.cap, b = t.band_size;
    int pidx = blockIdx.y;
    int m = k + b + pidx / (b - 1), n = m - 1 - pidx % (b - 1);
    int ra = *rank_at(t, m, k);
    if (ra == 0) return;
    if (n - k >= b && *rank_at(t, n, k) == 0) return;
    int nt = (NB + GT - 1) / GT;
    int ti = blockIdx.x % nt, tj = blockIdx.x / nt;
    const double *W = Wws + (size_t)pidx * cap * NB;
    gemm_tile<false, false>(ti * GT, tj * GT, NB, NB, ra, -1.0, lr_tile(t, m, k), NB, W, cap, 1.0,
                            band_tile(t, m, n), NB, sm);
}

// ======================= gemm, low-rank result (cases iv, v) =======================

struct RcWs {
    size_t stride;  // doubles per slot
    int smax;       // 2*cap
    int jlds;       // max s for the LDS Jacobi
};

__host__ __device__ inline size_t rc_slot_doubles(int NB, int cap)
{
    size_t smax = 2 * (size_t)cap;
    size_t nt = (smax + QRB - 1) / QRB + 1;
    return 2 * (size_t)NB * smax + (size_t)NB * QRB + QRB * smax + 2 * nt * QRB * QRB + 2 * smax +
           (size_t)cap * cap + 2 * smax * smax + 64;
}

#define RC_FIXED_SMEM (BLK_GEMM_SMEM + BLK_RED_SMEM + 2 * QRB * QRB)

// One workgroup per task (m,n): C' = Cu Cv^T - A B^T recompressed.
// (v)  A, B low rank:  Qn = -Bu (Av^T Bv)^T
// (iv) A low rank, B dense: Qn = -B Av
// P = [Cu Au], Q = [Cv Qn]; QR both; M = Rp Rq^T; Jacobi SVD; keep
// sigma >= acc (or the fixed rank) capped at maxrank; U' = Qp [Us S], V' = Qq [Vs].
__global__ void __launch_bounds__(BS) k_lr_gemm(TLR t, int k, const int *tasks_s, int smin, double *ws, RcWs g,
                                                double acc, int fixed_rk, int maxrank, OpLog o, int *ovf)
{
    extern __shared__ double sm[];
    const int NB = t.NB, NT = t.NT, cap = t.cap, b = t.band_size, tid = threadIdx.x;
    const int smax = g.smax;
    double *sig = sm + RC_FIXED_SMEM;
    int *pos = (int *)(sig + smax);
    int *flag = pos + smax;
    int *nrs = flag + 1;
    double *jl = sm + RC_FIXED_SMEM + 2 * smax + 8;  // LDS Jacobi storage
    double *P = ws + (size_t)blockIdx.x * g.stride;
    double *Q = P + (size_t)NB * smax;
    double *bw = Q + (size_t)NB * smax;
    size_t ntp = (size_t)(smax + QRB - 1) / QRB + 1;
    double *Tp = bw + (size_t)NB * QRB + (size_t)QRB * smax;
    double *Tq = Tp + ntp * QRB * QRB;
    double *taup = Tq + ntp * QRB * QRB;
    double *tauq = taup + smax;
    double *Xr = tauq + smax;
    double *Mg = Xr + (size_t)cap * cap;
    double *Vg = Mg + (size_t)smax * smax;
    const int T = NT - 1 - k;
    const long long ntask = (long long)T * (T - 1) / 2;
    for (long long task = blockIdx.x; task < ntask; task += gridDim.x) {
        // tasks_s: stack width decided before any update of this step; this
        // kernel takes the tasks the fast path leaves (s > smin)
        if (tasks_s[task] <= smin) continue;
        int hi, lo;
        tri_decode(task, hi, lo);
        const int m = k + 1 + hi, n = k + 1 + lo;
        const bool bdense = (n - k < b);
        const int ra = *rank_at(t, m, k), rc = *rank_at(t, m, n);
        const int rb = bdense ? 0 : *rank_at(t, n, k);
        const int s = rc + ra;
        const double *A = lr_tile(t, m, k), *Au = A, *Av = A + (size_t)NB * ra;
        double *C = lr_tile(t, m, n);
        // stacks
        blk_copy(NB, rc, C, NB, P, NB);
        blk_copy(NB, ra, Au, NB, P + (size_t)rc * NB, NB);
        blk_copy(NB, rc, C + (size_t)NB * rc, NB, Q, NB);
        __syncthreads();
        if (bdense) {
            blk_gemm<false, false>(NB, ra, NB, -1.0, band_tile(t, n, k), NB, Av, NB, 0.0, Q + (size_t)rc * NB, NB,
                                   sm);
        } else {
            const double *B = lr_tile(t, n, k), *Bv = B + (size_t)NB * rb;
            blk_gemm<true, false>(ra, rb, NB, 1.0, Av, NB, Bv, NB, 0.0, Xr, ra, sm);
            blk_gemm<false, true>(NB, ra, rb, -1.0, B, NB, Xr, ra, 0.0, Q + (size_t)rc * NB, NB, sm);
        }
        int nr = 0;
        if (s > 0) {
            blk_geqrf(NB, s, P, NB, taup, Tp, bw, sm);
            blk_geqrf(NB, s, Q, NB, tauq, Tq, bw, sm);
            const bool inl = (s <= g.jlds);
            double *M = inl ? jl : Mg;
            double *Vs = inl ? jl + (size_t)s * s : Vg;
            // M = triu(Rp) triu(Rq)^T
            for (int e = tid; e < s * s; e += BS) {
                int i = e % s, j = e / s;
                int l0 = i > j ? i : j;
                double v = 0.0;
                for (int l = l0; l < s; l++) v = fma(P[i + (size_t)l * NB], Q[j + (size_t)l * NB], v);
                M[e] = v;
            }
            __syncthreads();
            blk_jacobi(s, M, s, Vs, s, sig, flag, sm + BLK_GEMM_SMEM);
            blk_sort_pos(s, sig, pos);
            if (tid == 0) {
                int c = 0;
                if (fixed_rk > 0) c = fixed_rk < s ? fixed_rk : s;
                else
                    for (int j = 0; j < s; j++) c += (sig[j] >= acc);
                if (c > maxrank) { atomicAdd(ovf, 1); c = maxrank; }
                *nrs = c;
            }
            __syncthreads();
            nr = *nrs;
            double *Un = C, *Vn = C + (size_t)NB * nr;
            for (int j = 0; j < s; j++) {
                int q = pos[j];
                if (q >= nr) continue;
                for (int i = tid; i < NB; i += BS) {
                    Un[i + (size_t)q * NB] = (i < s) ? M[i + (size_t)j * s] : 0.0;
                    Vn[i + (size_t)q * NB] = (i < s) ? Vs[i + (size_t)j * s] : 0.0;
                }
            }
            __syncthreads();
            blk_ormqr(NB, nr, s, P, NB, Tp, Un, NB, bw, sm);
            blk_ormqr(NB, nr, s, Q, NB, Tq, Vn, NB, bw, sm);
        }
        if (tid == 0) {
            *rank_at(t, m, n) = nr;
            unsigned long long nn = tdim(t, m), ss = s;  // the reference counts with tile m's rows
            unsigned long long qa = opc('q', nn, ss, 0, 0), qb;
            if (bdense) qb = opc('q', nn, ss, 0, 0) + opc('m', ra, nn, nn, 0);
            else qb = opc('m', ra, rb, nn, 0) + opc('m', ra, nn, rb, 0) + opc('q', nn, ss, 0, 0);
            unsigned long long sv = opc('r', ss, ss, 2, 0) + opc('s', ss, 0, 0, 0) + ss * nr;
            unsigned long long nu = opc('o', nn, nr, ss, 1), nv = opc('o', nr, nn, ss, 2);
            oplog(o, NT, qa + qb + sv + nu + nv, m, n, false, false);
            // recompression totals: op count, tasks, algorithmic bytes
            // (8 * (inputs A, B, C + output) words, B dense in case iv)
            atomicAdd(o.stat + 4, qa + qb + sv + nu + nv);
            atomicAdd(o.stat + 5, 1ULL);
            unsigned long long words = 2ULL * NB * (ra + rc + nr) + (bdense ? (unsigned long long)NB * NB : 2ULL * NB * rb);
            atomicAdd(o.stat + 6, 8ULL * words);
        }
        __syncthreads();
    }
}

// ======================= utilities =======================

// 2*sum(log(diag)) over the true N, fixed reduction order; one workgroup
__global__ void __launch_bounds__(BS) k_logdet(TLR t, double *out)
{
    __shared__ double red[NWAVE];
    double part = 0.0;
    for (int g = threadIdx.x; g < t.N; g += BS) {
        int k = g / t.NB, i = g % t.NB;
        part += log(band_tile(t, k, k)[i + (size_t)i * t.NB]);
    }
    double s = blk_sum(part, red);
    if (threadIdx.x == 0) *out = 2.0 * s;
}

// rank map entries of band tiles set to -1 (rank_check.jdf on_band)
__global__ void k_rank_band(TLR t)
{
    int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= t.NT * t.NT) return;
    int m = e % t.NT, n = e / t.NT;
    if (m >= n && m - n < t.band_size) t.rank[e] = -1;
}
