// tlr.hpp -- the TLR descriptor as the device sees it, and helpers shared by
// every kernel of the library (no kernel definitions: included by each
// translation unit).
//
// Device layout (one descriptor, everything resident in HBM):
//   band tile (m,n), m-n < band : NB x NB, ld NB, at band + ((m-n)*NT + m)*NB*NB
//   off-band tile (m,n), m > n   : slot of NB*2*cap doubles at lr + (m(m-1)/2+n)*NB*2*cap,
//                                  compact [U | V]: U = NB x r, V = NB x r at offset NB*r
//   rank map                     : int32 NT x NT column major, rank[n*NT + m]
//   precision map                : uint8 NT x NT column major, prec[n*NT + m]: the
//                                  element type the off-band tile's [U | V] is held in
//                                  (LR_F64, or after hicma_tlr_demote LR_F32 / LR_BF16,
//                                  packed from the start of the slot, V at element NB*r)
// The matrix is padded to NT*NB with an identity block, so every tile is
// NB x NB and the factor of the padding is the identity (log 1 = 0).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

struct TLR {
    double *band;
    double *lr;
    int *rank;
    unsigned char *prec;
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
// element types of an off-band tile's [U | V]
#define LR_F64 0
#define LR_F32 1
#define LR_BF16 2
// a bfloat16 element: the upper half of an IEEE binary32
struct lr_bf16 {
    unsigned short u;
    __device__ __forceinline__ operator double() const { return (double)__uint_as_float((unsigned)u << 16); }
};
__device__ __forceinline__ int lr_prec(const TLR &t, int m, int n) { return t.prec[(size_t)n * t.NT + m]; }

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
    unsigned long long *exec;  // total of the stream the kernel runs on
    int *rhist;                // rank history {initial, min, max, final} per tile, or null
    int *unconv;               // count of SVDs that reached the sweep limit
};

__device__ __forceinline__ void oplog(const OpLog &o, int NT, unsigned long long cnt, int i, int j,
                                      bool band, bool path)
{
    atomicAdd(o.tile + (size_t)i * NT + j, cnt);
    atomicAdd(o.stat + (band ? 0 : 1), cnt);
    atomicAdd(o.stat + (path ? 2 : 3), cnt);
    atomicAdd(o.exec, cnt);
}

// new rank of the low-rank tile (m, n) after a recompression (the minimum
// and maximum of the reference's rank map, HiCMA_dpotrf_L_2flow.jdf:1100-1110)
__device__ __forceinline__ void rank_hist(const OpLog &o, int NT, int m, int n, int r)
{
    if (o.rhist) {
        int *h = o.rhist + 4 * ((size_t)n * NT + m);
        atomicMin(h + 1, r);
        atomicMax(h + 2, r);
    }
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

