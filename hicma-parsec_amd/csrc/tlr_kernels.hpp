// tlr_kernels.hpp -- HIP kernels of the TLR Cholesky path on gfx950
// (generation, dense and low-rank potrf / trsm / syrk / gemm, utilities).
#pragma once
#include "blk.hpp"
#include "mfma.hpp"
#include "tlr.hpp"

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

// oversampling of the randomized SVD (STARS-H's default)
#define GEN_OVS 10
// rank-map marker of a tile whose first, narrow sketch was too narrow
#define GEN_REDO (-2)

// STARS-H starsh_dense_dlrrsdd restated as a workgroup pipeline:
// D = kernel tile; Y = D*Omega; Q = qr(Y); X = D^T Q; X = Qx Rx; Jacobi SVD of Rx;
// rank by relative Frobenius tail <= tol; U = Q*Vr, V = Qx*(Ur*S).
//
// The sketch width is adaptive: the first pass (redo = 0) uses the first
// pw <= g.p columns of Omega.  A tile whose rank comes out above
// pw - GEN_OVS may have more than the sketch can see, so it is marked
// GEN_REDO and the second pass (redo = 1) compresses exactly the marked
// tiles with the full width g.p = maxrank + GEN_OVS, as STARS-H does for
// every tile.  Ranks far below maxrank (the common case) thus cost a sketch
// of pw columns instead of maxrank + 10.
__global__ void __launch_bounds__(BS) k_compress(TLR t, const double *Om, double *ws, GenWs g, int pw, int redo,
                                                 double tol, int maxrank, int *info)
{
    extern __shared__ double sm[];
    double *sig = sm + BLK_GEMM_SMEM + BLK_RED_SMEM + 2 * QRB * QRB;
    double *ssig = sig + g.p;
    int *pos = (int *)(ssig + g.p);
    int *flag = pos + g.p;
    int *rk = flag + 1;
    // workspace layout for the full width P; this pass computes with p <= P
    const int NB = t.NB, NT = t.NT, P = g.p, p = pw, tid = threadIdx.x;
    double *D = ws + (size_t)blockIdx.x * g.stride;
    double *Y = D + (size_t)NB * NB;
    double *Qm = Y + (size_t)NB * P;
    double *X = Qm + (size_t)NB * P;
    double *bw = X + (size_t)NB * P;
    double *tauY = bw + (size_t)NB * QRB + (size_t)QRB * (P > NB ? P : NB);
    double *tauX = tauY + P;
    size_t ntp = (size_t)(P + QRB - 1) / QRB;
    double *TY = tauX + P;
    double *TX = TY + ntp * QRB * QRB;
    double *Mj = TX + ntp * QRB * QRB;
    double *Vj = Mj + (size_t)P * P;
    double *Vsel = Vj + (size_t)P * P;
    const long long ntask = (long long)NT * (NT - 1) / 2;
    for (long long task = blockIdx.x; task < ntask; task += gridDim.x) {
        int m, n;
        tri_decode(task, m, n);
        if (m - n < t.band_size) continue;
        if (redo && *rank_at(t, m, n) != GEN_REDO) continue;
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
        if (blk_jacobi(p, Mj, p, Vj, p, sig, flag, sm + BLK_GEMM_SMEM) >= BLK_JACOBI_SWEEPS && tid == 0)
            atomicAdd(info + 2, 1);
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
            if (p < P && r > p - GEN_OVS) r = GEN_REDO;
            else if (!(r < mn / 2 && r <= maxrank)) r = -1;
            *rk = r;
        }
        __syncthreads();
        const int r = *rk;
        if (r == GEN_REDO) {
            if (tid == 0) *rank_at(t, m, n) = GEN_REDO;
            __syncthreads();
            continue;
        }
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

#define PBK 64
#define PLR (PBK + 2)

__device__ __forceinline__ void wave_lds_sync()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// ---- 64 x 64 Cholesky and inverse by a workgroup of four waves, blocked by
// 16: per diagonal block b, one wave factors the 16 x 16 block and inverts
// it (16 dependent steps each, pivots and columns broadcast by readlane),
// then the panel L(below, b) = S(below, b) W_b^T and the trailing update
// S -= L L^T run as 16 x 16 MFMA tiles over the four waves; afterwards the
// off-diagonal blocks of W = L^-1, W_IJ = -W_II sum_{l=J}^{I-1} L_Il W_lJ,
// in three rounds of MFMA tiles.
// Lr: the matrix (row-major, ld PLR; lower part, rows of the padding hold
// the identity); on return its lower part is L.  Wc: W(r, c) at
// Wc[c * PLR + r].  Ts: 4 x 16 x 17 doubles of scratch.  Called by all 256
// threads; returns the 1-based index of the first non-positive pivot, or 0.
__device__ __forceinline__ double d16_rd(double v, int l)
{
    return mk64((unsigned)__builtin_amdgcn_readlane((int)lo32(v), l),
                (unsigned)__builtin_amdgcn_readlane((int)hi32(v), l));
}

__device__ int diag_factor_blk(double *Lr, double *Wc, double *Ts, int *s_bad)
{
    const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6, i16 = lane & 15, k4 = lane >> 4;
    if (tid == 0) *s_bad = 0;
    __syncthreads();
    for (int b = 0; b < PBK; b += 16) {
        if (w == 0) {
            // factor: lane i (mod 16) holds row i of the block
            const int i = i16;
            double a[16];
#pragma unroll
            for (int c = 0; c < 16; c++) a[c] = Lr[(b + i) * PLR + b + c];
            int badj = 0;
#pragma unroll
            for (int j = 0; j < 16; j++) {
                const double d2 = d16_rd(a[j], j);
                if (!(d2 > 0.0) && badj == 0) badj = j + 1;
                const double d = sqrt(d2);
                const double lij = (i == j) ? d : (i > j ? a[j] / d : 0.0);
                a[j] = lij;
#pragma unroll
                for (int c = j + 1; c < 16; c++) a[c] = fma(-lij, d16_rd(lij, c), a[c]);
            }
            // inverse: lane c (mod 16) holds column c of W
            double x[16];
#pragma unroll
            for (int r = 0; r < 16; r++) x[r] = (r == i) ? 1.0 : 0.0;
#pragma unroll
            for (int l = 0; l < 16; l++) {
                const double xl = x[l] / d16_rd(a[l], l);
                x[l] = xl;
#pragma unroll
                for (int r = l + 1; r < 16; r++) x[r] = fma(-d16_rd(a[l], r), xl, x[r]);
            }
            if (lane < 16) {
#pragma unroll
                for (int c = 0; c < 16; c++) Lr[(b + i) * PLR + b + c] = (c <= i) ? a[c] : 0.0;
#pragma unroll
                for (int r = 0; r < 16; r++) Wc[(b + i) * PLR + b + r] = x[r];
            }
            if (lane == 0 && badj && *s_bad == 0) *s_bad = b + badj;
        }
        __syncthreads();
        const int nt = (PBK - b - 16) / 16;  // 16-row tiles below the block
        // panel: L(rt + i, b + j) = sum_k S(rt + i, b + k) W(b + j, b + k)
        if (w < nt) {
            const int rt = b + 16 + 16 * w;
            mk_d4 acc = mk_d4{0.0, 0.0, 0.0, 0.0};
#pragma unroll
            for (int k0 = 0; k0 < 16; k0 += 4)
                acc = mk_mfma(Lr[(rt + i16) * PLR + b + k0 + k4], Wc[(b + k0 + k4) * PLR + b + i16], acc);
#pragma unroll
            for (int r = 0; r < 4; r++) Lr[(rt + k4 + 4 * r) * PLR + b + i16] = acc[r];
        }
        __syncthreads();
        // trailing update of the lower block tiles (ti, tj), tj <= ti
        for (int q = w; q < nt * (nt + 1) / 2; q += 4) {
            int ti = 0;
            while ((ti + 1) * (ti + 2) / 2 <= q) ti++;
            const int tj = q - ti * (ti + 1) / 2;
            const int R = b + 16 + 16 * ti, Cc = b + 16 + 16 * tj;
            mk_d4 acc = mk_d4{0.0, 0.0, 0.0, 0.0};
#pragma unroll
            for (int k0 = 0; k0 < 16; k0 += 4)
                acc = mk_mfma(Lr[(R + i16) * PLR + b + k0 + k4], Lr[(Cc + i16) * PLR + b + k0 + k4], acc);
#pragma unroll
            for (int r = 0; r < 4; r++) Lr[(R + k4 + 4 * r) * PLR + Cc + i16] -= acc[r];
        }
        __syncthreads();
    }
    // off-diagonal blocks of W, block row I = 1, 2, 3 (wave J takes block (I, J))
    for (int I = 1; I < PBK / 16; I++) {
        if (w < I) {
            const int J = w;
            double *T = Ts + w * 16 * 17;
            mk_d4 acc = mk_d4{0.0, 0.0, 0.0, 0.0};
            for (int l = J; l < I; l++)
#pragma unroll
                for (int k0 = 0; k0 < 16; k0 += 4)
                    acc = mk_mfma(Lr[(16 * I + i16) * PLR + 16 * l + k0 + k4],
                                  Wc[(16 * J + i16) * PLR + 16 * l + k0 + k4], acc);
#pragma unroll
            for (int r = 0; r < 4; r++) T[(k4 + 4 * r) * 17 + i16] = acc[r];  // T(i, j)
            wave_lds_sync();
            mk_d4 acc2 = mk_d4{0.0, 0.0, 0.0, 0.0};
#pragma unroll
            for (int k0 = 0; k0 < 16; k0 += 4)
                acc2 = mk_mfma(Wc[(16 * I + k0 + k4) * PLR + 16 * I + i16], T[(k0 + k4) * 17 + i16], acc2);
#pragma unroll
            for (int r = 0; r < 4; r++) Wc[(16 * J + i16) * PLR + 16 * I + k4 + 4 * r] = -acc2[r];
        }
        __syncthreads();
    }
    return *s_bad;
}

// ---- the diagonal tile's Cholesky and its inverse as one launch: a
// device-side dataflow schedule over 64 x 64 blocks.
//
// Factor task (i, j), i >= j: S = A_ij - sum_{l<j} L_il L_jl^T (64-deep MFMA
// chunks, chunk l as soon as blocks (i, l) and (j, l) are done); then i == j
// factors S (L_jj and W_j = L_jj^-1, one wave) and i > j forms
// L_ij = S W_j^T once (j, j) is done.
// Inverse task (i, j), i > j: Linv_ij = -W_i sum_{l=j}^{i-1} L_il Linv_lj
// (Linv_jj = W_j), the blocks of L^-1 that the low-rank trsm multiplies by.
//
// Numbering: block column c holds the factor tasks (c .. nb-1, c), then the
// inverse tasks (c, 0 .. c-1); workgroups claim tasks in that order from an
// atomic counter.  Every task waits only for tasks with smaller numbers,
// which started earlier and wait for nothing larger, so the schedule cannot
// deadlock; each wait is bounded all the same (a wait that runs out marks
// *err and the task goes on, so the kernel always ends).
//
// flags: [0] task counter, [1] elected XCD + 1, [2 + task] done; zero at
// launch.  A wait that runs out sets *err (sticky over the factorization; the
// host reports it as an error).
//
// All workers run on ONE XCD, so blocks pass between them through that XCD's
// L2, which all its CUs share: the launch has DAG_XCD_SPREAD workgroups per
// worker, the first workgroup to start elects its XCD (hardware register
// XCC_ID) and workgroups on other XCDs leave at once.  Publishing a block:
// its stores, a wait in every thread until they have completed (the vector L1
// writes through to L2), a barrier and the done flag; reading: the flags,
// one invalidation of this CU's L1 for all dependencies already done, a
// barrier.  No L2 write-back or invalidation is needed (an agent-scope fence
// would write back or drop the whole L2, which the concurrent kernels use).
// The kernel's end publishes the results to other XCDs as every kernel end
// does.
#define DAG_SPIN_LIMIT (1 << 21)
#define DAG_XCD_SPREAD 8
#define DAG_SMEM (2 * PBK * PLR + MK_SMEM)

__device__ __forceinline__ int dag_fac(int i, int j, int nb) { return j * nb + (i - j); }
__device__ __forceinline__ int dag_inv(int i, int j, int nb) { return i * nb + (nb - i) + j; }

// Waits until tasks t0(l) and t1(l) are done, then counts how many of the
// following chunks l+1 .. lend-1 are done as well, and acquires once for all
// of them.  Returns the first chunk not yet acquired.  dep(l, &a, &b) names
// chunk l's two tasks.
template <typename Dep>
__device__ __forceinline__ int dag_acquire(int *flags, int l, int lend, Dep dep, int *err)
{
    __shared__ int s_ready;
    if (threadIdx.x == 0) {
        auto done = [&](int c) {
            int a, b;
            dep(c, a, b);
            return __hip_atomic_load(flags + 2 + a, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) &&
                   __hip_atomic_load(flags + 2 + b, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        };
        int n = 0;
        while (!done(l)) {
            __builtin_amdgcn_s_sleep(2);
            if (++n > DAG_SPIN_LIMIT) {
                __hip_atomic_store(err, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                break;
            }
        }
        int r = l + 1;
        while (r < lend && done(r)) r++;
        // drop this CU's L1 copies: the blocks are read from the XCD's L2
        asm volatile("buffer_inv sc0\n\ts_waitcnt vmcnt(0)" ::: "memory");
        s_ready = r;
    }
    __syncthreads();
    const int r = s_ready;
    __syncthreads();  // s_ready is reused by the next call
    return r;
}

__device__ __forceinline__ void dag_done(int *flags, int task)
{
    // every thread's stores of the block have reached the XCD's L2 before
    // the barrier; then the flag
    asm volatile("s_waitcnt vmcnt(0) lgkmcnt(0)" ::: "memory");
    __syncthreads();
    if (threadIdx.x == 0) __hip_atomic_store(flags + 2 + task, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Linv: nb x nb blocks of 64 x 64 (column major, ld 64) at Linv + (i nb + j) 4096,
// written for i > j; the diagonal blocks are Winv's
__global__ void __launch_bounds__(BS) k_potrf_dag(TLR t, int k, double *Winv, double *Linv, int *flags, int *info,
                                                 int *err, OpLog o)
{
    extern __shared__ double sm[];
    double *Lr = sm, *Wc = sm + PBK * PLR, *Mk = sm + 2 * PBK * PLR;
    __shared__ int s_task, s_bad;
    const int NB = t.NB, nb = (NB + PBK - 1) / PBK, ntask = nb * nb, tid = threadIdx.x;
    double *A = band_tile(t, k, k);
    {
        unsigned xcc;
        asm volatile("s_getreg_b32 %0, hwreg(HW_REG_XCC_ID)" : "=s"(xcc));
        const int mine = (int)(xcc & 15u) + 1;
        if (tid == 0) {
            const int e = atomicCAS(flags + 1, 0, mine);
            s_task = (e == 0) ? mine : e;
        }
        __syncthreads();
        if (s_task != mine) return;  // another XCD runs the schedule
    }
    for (;;) {
        __syncthreads();  // s_task and the LDS buffers are free
        if (tid == 0) s_task = atomicAdd(flags, 1);
        __syncthreads();
        const int task = s_task;
        if (task >= ntask) break;
        const int col = task / nb, pos = task % nb;
        const bool inverse = pos >= nb - col;
        const int i = inverse ? col : col + pos, j = inverse ? pos - (nb - col) : col;
        const int r0 = PBK * i, c0 = PBK * j;
        const int ri = min(PBK, NB - r0), cj = min(PBK, NB - c0);
        if (task == 0 && tid == 0) oplog(o, t.NT, opc('c', tdim(t, k), 0, 0, 0), k, k, true, true);
        mk_d4 acc[2][2];
        mk_zero(acc);
        if (inverse) {
            // S = sum_{l=j}^{i-1} L_il Linv_lj, then Linv_ij = -W_i S
            auto dep = [&](int c, int &a, int &b) {
                a = dag_fac(i, c, nb);
                b = (c == j) ? dag_fac(j, j, nb) : dag_inv(c, j, nb);
            };
            int ready = j;
            for (int l = j; l < i; l++) {
                if (l >= ready) ready = dag_acquire(flags, l, i, dep, err);
                const double *blk = (l == j) ? Winv + (size_t)j * PBK * PBK : Linv + ((size_t)l * nb + j) * PBK * PBK;
                const MkOps<false, true> op{A + r0, NB, blk - (long)PBK * l, PBK};
                mk_accum(op, ri, cj, 0, 0, (long)PBK * l, (long)PBK * l + PBK, acc, Mk);
            }
            mk_store(acc, [&](int ii, int jj, double v) { Lr[ii * PLR + jj] = v; });  // S(ii, jj)
            dag_acquire(flags, 0, 1, [&](int, int &a, int &b) { a = b = dag_fac(i, i, nb); }, err);
            const double *Wi = Winv + (size_t)i * PBK * PBK;
            for (int e = tid; e < PBK * PBK; e += BS) {
                const int ii = e % PBK, l = e / PBK;  // W_i(ii, l), column major
                Wc[ii * PLR + l] = (ii < ri && l < ri) ? Wi[ii + l * PBK] : 0.0;
            }
            __syncthreads();
            // Linv_ij(ii, c) = -sum_{l <= ii} W_i(ii, l) S(l, c); thread: row ii, columns cg + 4 q
            const int ii = tid % PBK, cg = tid / PBK;
            double *Lo = Linv + ((size_t)i * nb + j) * PBK * PBK;
#pragma unroll 4
            for (int q = 0; q < PBK / 4; q++) {
                const int c = cg + 4 * q;
                double x = 0.0;
                for (int l = 0; l <= ii; l++) x = fma(Wc[ii * PLR + l], Lr[l * PLR + c], x);
                Lo[ii + c * PBK] = -x;
            }
        } else {
            // S = A_ij - sum_l L_il L_jl^T
            const MkOps<false, false> op{A + r0, NB, A + c0, NB};
            auto dep = [&](int c, int &a, int &b) {
                a = dag_fac(i, c, nb);
                b = dag_fac(j, c, nb);
            };
            int ready = 0;
            for (int l = 0; l < j; l++) {
                if (l >= ready) ready = dag_acquire(flags, l, j, dep, err);
                mk_accum(op, ri, cj, 0, 0, (long)PBK * l, (long)PBK * l + PBK, acc, Mk);
            }
            if (i == j) {
                // S (lower) row major into Lr, the identity on the padding;
                // factor and invert with the four waves
                mk_store(acc, [&](int ii, int jj, double v) {
                    Lr[ii * PLR + jj] = (ii < cj && jj <= ii) ? A[r0 + ii + (size_t)(c0 + jj) * NB] - v
                                                              : (ii == jj ? 1.0 : 0.0);
                });
                __syncthreads();
                const int bad = diag_factor_blk(Lr, Wc, Mk, &s_bad);
                if (bad) {
                    if (tid == 0) atomicCAS(info, 0, k * NB + c0 + bad);
                } else {
                    // the whole 64 x 64 W_j, zero above the diagonal and outside
                    // cj x cj (the trsm multiplies the padding by zero rows)
                    double *Wo = Winv + (size_t)j * PBK * PBK;
                    for (int e = tid; e < PBK * PBK; e += BS) {
                        const int ii = e % PBK, c = e / PBK;
                        const bool in = c < cj && ii < cj && ii >= c;
                        if (in) A[r0 + ii + (size_t)(c0 + c) * NB] = Lr[ii * PLR + c];
                        Wo[ii + c * PBK] = in ? Wc[c * PLR + ii] : 0.0;
                    }
                }
            } else {
                // S row major into Lr; W_j (written by task (j, j)) into Wc
                mk_store(acc, [&](int ii, int jj, double v) {
                    Lr[ii * PLR + jj] = (ii < ri && jj < cj) ? A[r0 + ii + (size_t)(c0 + jj) * NB] - v : 0.0;
                });
                dag_acquire(flags, 0, 1, [&](int, int &a, int &b) { a = b = dag_fac(j, j, nb); }, err);
                const double *Wj = Winv + (size_t)j * PBK * PBK;
                for (int e = tid; e < PBK * PBK; e += BS) {
                    const int ii = e % PBK, c = e / PBK;  // W(ii, c), column major
                    Wc[c * PLR + ii] = (ii < cj && c < cj) ? Wj[ii + c * PBK] : 0.0;
                }
                __syncthreads();
                // L_ij(ii, c) = sum_{l <= c} S(ii, l) W(c, l); thread: row ii, columns cg + 4 q
                const int ii = tid % PBK, cg = tid / PBK;
#pragma unroll 4
                for (int q = 0; q < PBK / 4; q++) {
                    const int c = cg + 4 * q;
                    double x = 0.0;
                    for (int l = 0; l <= c; l++) x = fma(Lr[ii * PLR + l], Wc[l * PLR + c], x);
                    if (ii < ri && c < cj) A[r0 + ii + (size_t)(c0 + c) * NB] = x;
                }
            }
        }
        dag_done(flags, task);
    }
}

// low-rank tiles m = k+band .. NT-1: V <- L^-1 V as one product with the
// block inverse of L_kk (k_potrf_dag), TC columns of one tile per workgroup.
// The columns are staged in LDS (rows padded to 64 nb, stride 64 nb + 1), so
// the product can overwrite V in place; wave w computes rows 16 w .. 16 w + 15
// of every 64-row block b on the matrix cores: sum over c <= b of
// Linv_bc (16 x 64) V_c (64 x TC), the 16 Linv values of a 64-deep chunk
// loaded before its MFMAs.  grid (ceil(cap / TC), count).
#define TRSM_LDS_MAX (150 * 1024)  // bytes of LDS the staged columns may take
template <int TC>
__global__ void __launch_bounds__(BS) k_trsm_lr_inv(TLR t, int k, const double *Winv, const double *Linv, OpLog o)
{
    extern __shared__ double Vl[];
    const int NB = t.NB, nb = (NB + PBK - 1) / PBK, RLD = PBK * nb + 1, tid = threadIdx.x;
    const int m = k + t.band_size + blockIdx.y;
    const int r = *rank_at(t, m, k);
    if (blockIdx.x == 0 && tid == 0) oplog(o, t.NT, opc('t', tdim(t, m), r, 1, 0), m, k, false, m - k == 1);
    const int c0 = blockIdx.x * TC;
    if (c0 >= r) return;
    const int nc = (r - c0) < TC ? (r - c0) : TC;
    double *V = lr_tile(t, m, k) + (size_t)NB * r + (size_t)c0 * NB;
    for (int e = tid; e < PBK * nb * TC; e += BS) {
        const int row = e % (PBK * nb), c = e / (PBK * nb);
        Vl[row + c * RLD] = (row < NB && c < nc) ? V[row + (size_t)c * NB] : 0.0;
    }
    __syncthreads();
    const int lane = tid & 63, w = tid >> 6, i16 = lane & 15, k4 = lane >> 4;
    const bool colok = i16 < TC;
    for (int b = 0; b < nb; b++) {
        mk_d4 acc = mk_d4{0.0, 0.0, 0.0, 0.0};
        for (int c = 0; c <= b; c++) {
            const double *blk = (c == b) ? Winv + (size_t)b * PBK * PBK : Linv + ((size_t)b * nb + c) * PBK * PBK;
            const double *arow = blk + 16 * w + i16;
            double a[16];
#pragma unroll
            for (int q = 0; q < 16; q++) a[q] = arow[(4 * q + k4) * PBK];
            const double *vcol = Vl + PBK * c + k4 + i16 * RLD;
#pragma unroll
            for (int q = 0; q < 16; q++) acc = mk_mfma(a[q], colok ? vcol[4 * q] : 0.0, acc);
        }
#pragma unroll
        for (int rr = 0; rr < 4; rr++) {
            const int row = PBK * b + 16 * w + k4 + 4 * rr;
            if (row < NB && i16 < nc) V[row + (size_t)i16 * NB] = acc[rr];
        }
    }
}

// ======================= trsm =======================

// band tiles m = k+1 .. k+band-1: C <- C L^{-T} = C (L^{-1})^T on the matrix
// cores with the block inverse of L_kk that k_potrf_dag leaves (Winv: the
// diagonal blocks, Linv: the blocks below, 64 x 64 column major, Linv_bc at
// Linv + (b nb + c) 4096).  One workgroup per 64-row stripe of a tile
// (grid (ceil(NB/64), band-1)), in place: column block jb of the result,
// X(:, jb) = sum_{ib <= jb} C(:, ib) Linv_{jb,ib}^T, reads column blocks
// 0 .. jb of C only, so the blocks run from the last to the first.  For a
// fixed jb the blocks Linv_{jb,0..jb-1} are contiguous, so the whole sum
// over ib < jb is one operand with ld 64.
__global__ void __launch_bounds__(BS) k_trsm_band(TLR t, int k, const double *Winv, const double *Linv, OpLog o)
{
    __shared__ double sm[MK_SMEM];
    const int NB = t.NB, nb = (NB + PBK - 1) / PBK;
    const int m = k + 1 + blockIdx.y, r0 = blockIdx.x * PBK;
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        const int mm = tdim(t, m);
        oplog(o, t.NT, opc('t', mm, mm, 1, 0), m, k, true, m - k == 1);
    }
    double *C = band_tile(t, m, k);
    for (int jb = nb - 1; jb >= 0; jb--) {
        const int kd = min(PBK, NB - PBK * jb);  // columns of block jb inside the tile
        mk_d4 acc[2][2];
        mk_zero(acc);
        if (jb > 0) {
            // B(kk, j) = L^-1(64 jb + j, kk) = Linv[(jb nb) 4096 + j + 64 kk] for kk < 64 jb
            const MkOps<false, false> op{C, NB, Linv + (size_t)jb * nb * PBK * PBK, PBK};
            mk_accum(op, NB, kd, r0, 0, 0, (long)PBK * jb, acc, sm);
        }
        const MkOps<false, false> od{C + (size_t)PBK * jb * NB, NB, Winv + (size_t)jb * PBK * PBK, PBK};
        mk_accum(od, NB, kd, r0, 0, 0, kd, acc, sm);  // ends with a barrier: the block's reads are done
        mk_store(acc, [&](int i, int j, double v) {
            if (r0 + i < NB && j < kd) C[(r0 + i) + (size_t)(PBK * jb + j) * NB] = v;
        });
    }
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

// tasks_s value of a task handed from a fast recompression kernel to the
// Householder path (k_lr_gemm, launched after them on the context stream)
#define RC_REDO (1 << 30)

// Accepted error of the V' = Q (Rq^-1 Vs_r) back-substitution of the fast
// kernels, absolute in the tile: a fraction of the truncation threshold,
// never below rounding of M (fro = ||M||_F).
__device__ __forceinline__ double rc_solve_tol(double acc, int fixed_rk, double fro)
{
    const double floor_ = 1e-13 * fro;
    return (fixed_rk > 0 || !(acc > 0.0)) ? floor_ : fmax(1e-3 * acc, floor_);
}

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
            if (blk_jacobi(s, M, s, Vs, s, sig, flag, sm + BLK_GEMM_SMEM) >= BLK_JACOBI_SWEEPS && tid == 0)
                atomicAdd(o.unconv, 1);
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
            rank_hist(o, NT, m, n, nr);
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

// 2*sum(log(diag)) over the true N in a fixed order: one workgroup per
// diagonal tile sums its part (a wave sum per 64 rows, then the waves in
// order), then k_logdet_sum adds the NT parts in order
#define LD_T 1024
__global__ void __launch_bounds__(LD_T) k_logdet_parts(TLR t, double *parts)
{
    __shared__ double red[LD_T / 64];
    const int k = blockIdx.x, rows = tdim(t, k);
    const double *T = band_tile(t, k, k);
    double part = 0.0;
    for (int i = threadIdx.x; i < rows; i += LD_T) part += log(T[i + (size_t)i * t.NB]);
    part = wsum(part);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = part;
    __syncthreads();
    if (threadIdx.x == 0) {
        double s = 0.0;
        for (int w = 0; w < LD_T / 64; w++) s += red[w];
        parts[k] = s;
    }
}

__global__ void k_logdet_sum(const double *parts, int n, double *out)
{
    if (threadIdx.x == 0) {
        double s = 0.0;
        for (int k = 0; k < n; k++) s += parts[k];
        *out = 2.0 * s;
    }
}

// rank history of the off-band tiles: phase 0 sets initial = min = max =
// final = the current rank, phase 1 the final rank
// (hicma_parsec/HiCMA_dpotrf_L_2flow.jdf:353-361, 1100-1110); 0 elsewhere
__global__ void k_rank_hist(TLR t, int *h, int phase)
{
    int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= t.NT * t.NT) return;
    int m = e % t.NT, n = e / t.NT;
    int *p = h + 4 * (size_t)e;
    int r = (m - n >= t.band_size) ? t.rank[e] : 0;
    if (phase == 0) {
        p[0] = r;
        p[1] = r;
        p[2] = r;
    }
    p[3] = r;
}

// fills that replace hipMemsetAsync in work a HIP graph may capture
__global__ void k_fill_i32(int *p, int n, int v)
{
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e < n) p[e] = v;
}
__global__ void k_fill_u64(unsigned long long *p, long n)
{
    const long e = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (e < n) p[e] = 0ULL;
}

// rank map entries of band tiles set to -1 (rank_check.jdf on_band)
__global__ void k_rank_band(TLR t)
{
    int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= t.NT * t.NT) return;
    int m = e % t.NT, n = e / t.NT;
    if (m >= n && m - n < t.band_size) t.rank[e] = -1;
}
