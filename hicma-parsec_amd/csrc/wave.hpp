// wave.hpp -- wave64 cross-lane primitives for FP64 on gfx950.
//
// Built from DPP row operations (quad_perm, row_half_mirror, row_ror) and the
// gfx950 V_PERMLANE16_SWAP / V_PERMLANE32_SWAP instructions instead of
// ds_bpermute: they stay in the VALU and cost a few cycles per step, which
// makes the many small column reductions of QR and Jacobi latency-cheap.
// A double travels as its two 32-bit halves.
#pragma once
#include <hip/hip_runtime.h>

__device__ __forceinline__ unsigned lo32(double v)
{
    return (unsigned)((unsigned long long)__double_as_longlong(v) & 0xffffffffULL);
}
__device__ __forceinline__ unsigned hi32(double v)
{
    return (unsigned)((unsigned long long)__double_as_longlong(v) >> 32);
}
__device__ __forceinline__ double mk64(unsigned l, unsigned h)
{
    return __longlong_as_double((long long)(((unsigned long long)h << 32) | l));
}

// DPP controls (GFX9 encoding)
#define DPP_XOR1 0xB1   // quad_perm [1,0,3,2]
#define DPP_XOR2 0x4E   // quad_perm [2,3,0,1]
#define DPP_REV4 0x1B   // quad_perm [3,2,1,0]  (lane ^ 3 in a quad)
#define DPP_HMIR 0x141  // row_half_mirror      (lane ^ 7 in 8 lanes)
#define DPP_ROR8 0x128  // row_ror:8            (lane ^ 8 in 16 lanes)

template <int CTRL>
__device__ __forceinline__ double dppd(double v)
{
    int l = __builtin_amdgcn_update_dpp(0, (int)lo32(v), CTRL, 0xf, 0xf, false);
    int h = __builtin_amdgcn_update_dpp(0, (int)hi32(v), CTRL, 0xf, 0xf, false);
    return mk64((unsigned)l, (unsigned)h);
}

// value of lane ^ 4
__device__ __forceinline__ double xor4d(double v) { return dppd<DPP_REV4>(dppd<DPP_HMIR>(v)); }

// x <-> y swap of the upper 32 lanes of x with the lower 32 lanes of y
__device__ __forceinline__ void swap32d(double &x, double &y)
{
    auto a = __builtin_amdgcn_permlane32_swap(lo32(x), lo32(y), false, false);
    auto b = __builtin_amdgcn_permlane32_swap(hi32(x), hi32(y), false, false);
    x = mk64(a[0], b[0]);
    y = mk64(a[1], b[1]);
}
// swap of the odd 16-lane rows of x with the even rows of y
__device__ __forceinline__ void swap16d(double &x, double &y)
{
    auto a = __builtin_amdgcn_permlane16_swap(lo32(x), lo32(y), false, false);
    auto b = __builtin_amdgcn_permlane16_swap(hi32(x), hi32(y), false, false);
    x = mk64(a[0], b[0]);
    y = mk64(a[1], b[1]);
}

// sum over groups of G lanes (G = 2, 4, 8, 16, 32, 64), every lane of a
// group gets its group's total
template <int G>
__device__ __forceinline__ double gsum(double v)
{
    if (G >= 2) v += dppd<DPP_XOR1>(v);
    if (G >= 4) v += dppd<DPP_XOR2>(v);
    if (G >= 8) v += xor4d(v);
    if (G >= 16) v += dppd<DPP_ROR8>(v);
    if (G >= 32) {
        double x = v, y = v;
        swap16d(x, y);
        v = x + y;
    }
    if (G >= 64) {
        double x = v, y = v;
        swap32d(x, y);
        v = x + y;
    }
    return v;
}

__device__ __forceinline__ double wsum(double v) { return gsum<64>(v); }

// reduce-scatter of 16 values across the wave: lane l returns the wave total
// of value index l >> 2
__device__ __forceinline__ double wrs16(double (&p)[16])
{
    const int lane = threadIdx.x & 63;
#pragma unroll
    for (int k = 0; k < 8; k++) {
        double x = p[k], y = p[k + 8];
        swap32d(x, y);
        p[k] = x + y;
    }
#pragma unroll
    for (int k = 0; k < 4; k++) {
        double x = p[k], y = p[k + 4];
        swap16d(x, y);
        p[k] = x + y;
    }
    {
        const bool up = (lane & 8) != 0;
#pragma unroll
        for (int k = 0; k < 2; k++) {
            double keep = up ? p[k + 2] : p[k], send = up ? p[k] : p[k + 2];
            p[k] = keep + dppd<DPP_ROR8>(send);
        }
    }
    {
        const bool up = (lane & 4) != 0;
        double keep = up ? p[1] : p[0], send = up ? p[0] : p[1];
        p[0] = keep + xor4d(send);
    }
    double v = p[0];
    v += dppd<DPP_XOR2>(v);
    v += dppd<DPP_XOR1>(v);
    return v;
}

// LDS ordering inside one wave: a code-motion barrier plus wavefront-scope
// fences (LDS operations of one wave complete in order)
__device__ __forceinline__ void wv_sync()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// value of v on lane addr / 4 (ds_bpermute: through the LDS crossbar, no
// LDS storage)
__device__ __forceinline__ double wv_bperm(int addr, double v)
{
    return mk64((unsigned)__builtin_amdgcn_ds_bpermute(addr, (int)lo32(v)),
                (unsigned)__builtin_amdgcn_ds_bpermute(addr, (int)hi32(v)));
}

// Raw buffer access to one column of a tile or workspace.  The resource
// (gfx9 word 3 = 0x00020000: untyped, unswizzled) bounds the column: a load
// at or past `bytes` returns 0 and a store there is dropped, so rows past the
// tile and columns given 0 bytes need no branch.  Buffer loads count only
// against vmcnt and are issued from a scalar base, so loads issued ahead of
// use are waited for at their use, not at the next LDS or scalar access.
typedef unsigned int wv_u2 __attribute__((ext_vector_type(2)));
#define WV_OOB 0x80000000  // a byte offset past every resource: load 0, store dropped
typedef __amdgpu_buffer_rsrc_t wv_rsrc;
__device__ __forceinline__ wv_rsrc wv_col_rsrc(const double *p, int bytes)
{
    // the base and size are wave-uniform; readfirstlane keeps the resource in
    // scalar registers whatever the compiler can prove
    const unsigned long long v = (unsigned long long)p;
    const unsigned lo = __builtin_amdgcn_readfirstlane((unsigned)v);
    const unsigned hi = __builtin_amdgcn_readfirstlane((unsigned)(v >> 32));
    double *u = (double *)(((unsigned long long)hi << 32) | lo);
    return __builtin_amdgcn_make_buffer_rsrc(u, 0, __builtin_amdgcn_readfirstlane(bytes), 0x00020000);
}
__device__ __forceinline__ double wv_ld(wv_rsrc r, int off)
{
    return __builtin_bit_cast(double, __builtin_amdgcn_raw_buffer_load_b64(r, off, 0, 0));
}
__device__ __forceinline__ void wv_st(double v, wv_rsrc r, int off)
{
    __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(wv_u2, v), r, off, 0, 0);
}
// rows lane, lane + 64, ... of a column (KR rows per lane)
template <int KR>
__device__ __forceinline__ void wv_ld_col(double (&x)[KR], wv_rsrc r, int lane)
{
#pragma unroll
    for (int k = 0; k < KR; k++) x[k] = wv_ld(r, (lane + 64 * k) * 8);
}
template <int KR>
__device__ __forceinline__ void wv_st_col(const double (&x)[KR], wv_rsrc r, int lane)
{
#pragma unroll
    for (int k = 0; k < KR; k++) wv_st(x[k], r, (lane + 64 * k) * 8);
}
