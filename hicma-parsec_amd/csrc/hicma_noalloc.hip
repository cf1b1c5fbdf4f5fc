// hicma_noalloc.hip -- hicma_tlr_loglik without any device allocation.
#include "internal.hpp"

#include <mutex>

// For callers that cannot allocate (code under graph capture, serving loops):
// the workspace is a file-scope device array of NOALLOC_SCRATCH_DOUBLES
// doubles (64 MiB), which each device holds until the process ends.  Its
// device address is fetched once per device at context set-up
// (hicma_noalloc_setup, called by hicma_init); host code uses only that
// address, never the array's name.  Every call on one device shares the array:
// calls that overlap on a device race.  Nothing in it needs to start at zero
// (every kernel writes what it later reads).
#define NOALLOC_SCRATCH_DOUBLES ((size_t)8 << 20)
#define NOALLOC_MAX_DEVICES 64
static __device__ __align__(256) double scratch[NOALLOC_SCRATCH_DOUBLES];
static double *scratch_dev[NOALLOC_MAX_DEVICES];
static std::mutex scratch_mu;

int hicma_noalloc_setup(int device)
{
    if (device < 0 || device >= NOALLOC_MAX_DEVICES) return HICMA_ERR_ARG;
    std::lock_guard<std::mutex> lk(scratch_mu);
    if (scratch_dev[device]) return 0;
    void *ptr = nullptr;
    HIPCHK(hipSetDevice(device));
    HIPCHK(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
    scratch_dev[device] = (double *)ptr;
    return 0;
}

extern "C" int hicma_tlr_loglik_noalloc(hicma_ctx *ctx, hicma_tlr_desc *A, double *z, double *logdet, double *quad,
                                        double *loglik)
{
    RCHK(check_solve_args(ctx, A, 1, z, A ? A->t.N : 0));
    if (ctx->device < 0 || ctx->device >= NOALLOC_MAX_DEVICES || !scratch_dev[ctx->device]) return HICMA_ERR_ARG;
    const size_t need = solve_ws_doubles(A->t, 1);
    if (need > NOALLOC_SCRATCH_DOUBLES) {
        fprintf(stderr,
                "hicma_tlr_loglik_noalloc: N=%d, NB=%d needs %zu doubles of workspace, above the %zu of the "
                "allocation-free scratch\n",
                A->t.N, A->t.NB, need, NOALLOC_SCRATCH_DOUBLES);
        return HICMA_ERR_ARG;
    }
    HIPCHK(hipSetDevice(ctx->device));
    double *ws = scratch_dev[ctx->device];
    double *vals = ws + need - 8;
    RCHK(solve_launch(ctx, A, 1, 1, z, A->t.N, ws, vals, vals + 1));
    return loglik_finish(ctx, A->t, vals, logdet, quad, loglik);
}

