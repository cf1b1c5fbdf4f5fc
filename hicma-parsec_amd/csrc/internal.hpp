// internal.hpp -- host-side structures shared by the translation units of
// libhicma_tlr.so (factorization, checks, solve, serving).
#pragma once
#include "tlr.hpp"
#include "../../include/hicma_tlr.h"

#include <stdio.h>
#include <algorithm>
#include <vector>

#define HIPCHK(x)                                                                                   \
    do {                                                                                            \
        hipError_t e_ = (x);                                                                        \
        if (e_ != hipSuccess) {                                                                     \
            fprintf(stderr, "hicma_tlr: HIP error '%s' at %s:%d\n", hipGetErrorString(e_), __FILE__, \
                    __LINE__);                                                                      \
            return e_ == hipErrorOutOfMemory ? HICMA_ERR_NOMEM : HICMA_ERR_HIP;                     \
        }                                                                                           \
    } while (0)

#define RCHK(x)                    \
    do {                           \
        int r_ = (x);              \
        if (r_) return r_;         \
    } while (0)

struct hicma_ctx {
    int device;
    int ncu;       // CUs the persistent grids and their workspaces are sized for
    int ncu_dev;   // CUs of the device
    int single;    // all work on `stream` (side[] alias it)
    int pool;      // stream-ordered allocation (hipMallocAsync / hipFreeAsync on `stream`)
    hipStream_t stream;
    double *ws;
    size_t ws_bytes;
    int *d_info;   // [0] info, [1] rank overflows, [2] unconverged Jacobi SVDs,
                   // [3] a bounded device wait ran out
    int *d_ovf;
    double *d_scalar;
    size_t lds_max;
    // schedule: panel steps with fewer low-rank gemm tasks than this run
    // every task on the four-wave kernel (0: never)
    int wave_min_tasks;
    int profile;                     // record events around recompression launches
    std::vector<hipEvent_t> prof_ev; // pairs
    // the recompression variants of one panel step run concurrently on side
    // streams (default priority), forked from and joined back to `stream`
    // (the device exposes 4 hardware queues per process: 4 streams in all)
    hipStream_t side[3];             // [2]: syrk(k) + potrf(k+1), one step ahead
    hipEvent_t ev_fork, ev_join[3];
    // critical-path timing events (created on first request, reused)
    std::vector<hipEvent_t> cp_ev;
    // dense linear algebra handle of the checks (created on first use)
    void *blas;
    void (*blas_free)(void *);
};

// what a captured factorization graph depends on: every pointer and scalar
// its kernels take (a replay is valid only while all are unchanged)
struct GraphKey {
    const void *band, *lr, *rank, *xyz, *ws, *tileops, *opstat;
    double acc, sigma2, beta, nugget;
    int fixed_rank, maxrank, band_size, N, NB, NT, cap, kind, wave_min_tasks;
};

struct hicma_tlr_desc {
    hicma_ctx *ctx;
    TLR t;  // device pointers + sizes
    int band_alloc;
    uint64_t seed;
    double *h_xyz;                  // points in matrix row order (host copy of t.xyz)
    double *h_xyz0;                 // points as created (before any index permutation)
    unsigned long long *d_tileops;
    unsigned long long *d_opstat;   // [0..3] band/offband/path/offpath, [4..6] recompression,
                                    // [8..11] per executing stream
    unsigned long last_stats[12];
    int last_ovf;
    int last_unconv;
    int *rhist;                     // NT*NT x {initial, min, max, final} rank history, or null
    int rhist_on;
    int demoted;                    // off-band tiles held in FP32 or BF16 (hicma_tlr_demote)
    int use_graph;                  // capture the factorization once, replay it after
    hipGraph_t graph;
    hipGraphExec_t gexec;
    GraphKey gkey;
};

struct hicma_dense {
    hicma_ctx *ctx;
    double *d;
    long N;
    long ld;
};

int hicma_ensure_ws(hicma_ctx *c, size_t bytes);
int hicma_ctx_release_ws(hicma_ctx *c);
// device allocations of a context: plain hipMalloc / hipFree, or, for a
// context created with HICMA_CTX_STREAM_ORDERED_ALLOC, ordered on its stream
// (a free then does not wait for the whole device); the memory is ready for
// any stream when ctx_malloc returns
hipError_t ctx_malloc(hicma_ctx *c, void **p, size_t bytes);
hipError_t ctx_free(hicma_ctx *c, void *p);
// device bytes of a descriptor, and of the workspace a generation plus
// factorization of that shape allocates on a context (hicma_tlr.hip)
size_t hicma_desc_bytes(int N, int NB, int cap, int band, int ndim);
size_t hicma_work_bytes(hicma_ctx *ctx, int N, int NB, int cap, int band);
// device address of the allocation-free entry point's scratch array, fetched
// once per device (hicma_noalloc.hip)
int hicma_noalloc_setup(int device);
// the solve's launch sequence and its helpers (hicma_solve.hip)
size_t solve_ws_doubles(const TLR &t, int nrhs);
void logdet_launch(const TLR &t, double *parts, double *out, hipStream_t s);
int check_solve_args(hicma_ctx *ctx, hicma_tlr_desc *A, int nrhs, const double *B, long ldb);
int solve_launch(hicma_ctx *ctx, hicma_tlr_desc *A, int dirs, int nrhs, double *B, long ldb, double *ws,
                 double *d_quad, double *d_logdet);
int loglik_finish(hicma_ctx *ctx, const TLR &t, double *d_vals, double *logdet, double *quad, double *loglik);
