/*
 * hicma_tlr.h -- C-ABI of the MI355X tile low-rank (TLR) Cholesky library.
 *
 * Drop-in for hicma-parsec's public interface (hicma_parsec/hicma_parsec.h):
 * the same entry-point names, parameter order and meaning, with
 *   parsec_context_t*          -> hicma_ctx*       (HIP device + stream + workspaces)
 *   parsec_tiled_matrix_dc_t*  -> hicma_tlr_desc*  (tiles, rank map, points; in HBM)
 * and no torch types anywhere.  All calls are blocking, like the reference's
 * parsec_context_wait.  One host thread per context; not re-entrant on a
 * context.
 *
 * Storage contract (hicma_parsec/testing_dpotrf.c:419-453, STARSH_gen.jdf:173-198):
 *   band tile (m,n), m-n < band: NB x NB doubles, column major, ld NB;
 *   off-band tile (m,n), m-n >= band: compact [U | V], U = NB x r (ld NB),
 *     V = NB x r (ld NB) at offset NB*r, A(m,n) ~= U V^T, capacity NB*maxrank*2
 *     (the reference's send_full_tile layout);
 *   rank map: int32 NT x NT, column major, G[n*NT + m].
 */
#ifndef HICMA_TLR_H
#define HICMA_TLR_H
#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* PLASMA/DPLASMA enum values for uplo */
#define HICMA_UPPER 121
#define HICMA_LOWER 122

/* return codes (negative: argument / resource errors; the reference's
 * NULL-taskpool path silently returned 0, hicma_parsec/hicma_parsec.c:551,595) */
#define HICMA_SUCCESS 0
#define HICMA_ERR_UPLO (-1)          /* uplo != Lower */
#define HICMA_ERR_MAXRANK_NB (-2)    /* maxrank > NB/2 */
#define HICMA_ERR_MAXRANK_COMP (-3)  /* maxrank > compmaxrank */
#define HICMA_ERR_NOMEM (-4)         /* HBM exhausted */
#define HICMA_ERR_HIP (-5)           /* HIP runtime error */
#define HICMA_ERR_ARG (-6)           /* other invalid argument */

/* problem kinds, numbered as the reference's -D flag
 * (hicma_parsec/testing_dpotrf.c:12, 336-358) */
#define HICMA_KIND_SQEXP_2D 2
#define HICMA_KIND_SQEXP_3D 3
#define HICMA_KIND_EXP_3D 4

typedef struct hicma_ctx hicma_ctx;
typedef struct hicma_tlr_desc hicma_tlr_desc;
/* dense N x N column-major matrix in HBM (the reference's dense dcA0 / dcAd
 * descriptors of the --check, hicma_parsec/testing_dpotrf.c:640-665) */
typedef struct hicma_dense hicma_dense;

/* The STARS-H problem arguments of STARSH_gen / parsec_band_regenerate
 * (hicma_parsec/hicma_parsec.h:108-121, 186-198: void *data,
 * STARSH_kernel *kernel, STARSH_int *index).  Each may be NULL, which keeps
 * what the descriptor was created with:
 *   data   -> const hicma_kernel_params * (sigma2, beta, nugget)
 *   kernel -> one of HICMA_KIND_* with the descriptor's point dimension
 *             (a STARS-H kernel is a host function; here the covariance runs
 *             on the device, so the kernel is named, not passed)
 *   index  -> N point indices: matrix row i uses point index[i] of the point
 *             set as created (a permutation of 0..N-1) */
typedef struct { double sigma2, beta, nugget; } hicma_kernel_params;
typedef int hicma_kernel;
typedef int64_t hicma_index;

/* ---- context and descriptors (replace setup_parsec / descriptor init,
 *      hicma_parsec/hicma_parsec.c:337-427, testing_dpotrf.c:419-453) ---- */
int hicma_init(hicma_ctx **ctx, int device);
/* a context with options, for several factorizations sharing one device:
 * HICMA_CTX_SINGLE_STREAM puts all of its work on one stream (one hardware
 * queue), HICMA_CTX_STREAM_ORDERED_ALLOC allocates and frees in stream order
 * (hipMallocAsync), so freeing does not wait for other contexts' work, and
 * cu_share > 0 sizes its persistent grids and workspaces for that many CUs
 * instead of the whole device (results do not depend on it) */
#define HICMA_CTX_SINGLE_STREAM 1
#define HICMA_CTX_STREAM_ORDERED_ALLOC 2
int hicma_init_ex(hicma_ctx **ctx, int device, int flags, int cu_share);
int hicma_fini(hicma_ctx *ctx);
/* creates the descriptor, generates the point set (jittered grid, Morton
 * order, seeded) and allocates tile slots of capacity maxrank in HBM */
int hicma_desc_create(hicma_ctx *ctx, hicma_tlr_desc **A, int N, int NB, int maxrank, int kind,
                      double sigma2, double beta, double nugget, uint64_t seed);
int hicma_desc_destroy(hicma_tlr_desc *A);
int hicma_desc_info(const hicma_tlr_desc *A, int *N, int *NB, int *NT, int *band_size, int *maxrank);
/* point coordinates, xyz[d*N + i] (host copy) */
int hicma_get_points(const hicma_tlr_desc *A, double *xyz);
/* the point generator itself (same spec), for callers without a descriptor */
int hicma_points(int N, int ndim, uint64_t seed, double *xyz);

/* ---- STARSH_gen (hicma_parsec/hicma_parsec.h:186-198, STARSH_gen.jdf:342-369):
 * dense band tiles + randomized-SVD compression of off-band tiles on the GPU.
 * *info = -1 when a tile is not compressible under maxrank. */
int STARSH_gen(hicma_ctx *ctx, int uplo, hicma_tlr_desc *A, hicma_tlr_desc *Ar, void *data, hicma_kernel *kernel,
               hicma_index *index, double tol, int maxrank, int band_size, int send_full_tile, int *info);

/* ---- rank map (hicma_parsec/hicma_parsec.h:146-166) ---- */
/* parsec_rank_gather: rank_array[n*NT+m] for m-n >= band, 0 elsewhere */
int parsec_rank_gather(hicma_ctx *ctx, hicma_tlr_desc *Ar, int *rank_array, int band_size);
/* parsec_rank_check: returns the number of negative off-band ranks and sets
 * band entries of the device rank map to -1 */
int parsec_rank_check(hicma_ctx *ctx, hicma_tlr_desc *Ar, int band_size);
/* HICMA_get_stat2 (hicma_parsec/hicma_internal.c:52-72) */
void HICMA_get_stat2(const int *G, int lda, int band_size, int *min, int *max, double *avg);

/* ---- band size (hicma_parsec/hicma_parsec.h:121-144) ---- */
/* parsec_band_size_auto_tuning: dcFake of the reference is not needed (pass NULL) */
int parsec_band_size_auto_tuning(hicma_ctx *ctx, hicma_tlr_desc *Ar, hicma_tlr_desc *Fake,
                                 int *rank_array, int NB);
/* the same decision on a host rank map, without a descriptor */
int hicma_band_autotune(const int *rank_array, int NT, int NB);
/* The same choice with the low-rank op counts weighted by lr_weight (1 is
 * the reference's rule): a cost model for a device where the dense band's
 * gemms run on the matrix cores many times faster per op than the low-rank
 * recompressions.  No counterpart in the reference. */
int hicma_band_autotune_weighted(const int *rank_array, int NT, int NB, double lr_weight);

/* Band choice by a time model of this device: the band size with the least
 * estimated factorization time, dense work priced at dense_tflops and the
 * recompressions (the reference's op counts, tile ranks grown by 1.25) at
 * lr_tflops (300 times slower for stacks wider than 128 columns), plus 0.5 ms
 * per panel step for a dense band.  No counterpart in the reference; the
 * serving loop uses it with 60 and 3.5. */
int hicma_band_autotune_model(const int *rank_array, int NT, int NB, double dense_tflops, double lr_tflops);
/* STARSH_gen in two stages, for callers that choose the band from the ranks
 * (the serving loop).  No counterpart in the reference.  hicma_gen_first:
 * STARSH_gen with band 1 and only the narrow first sketch; tiles whose rank it
 * cannot vouch for hold HICMA_RANK_PENDING in the rank map.  Then the caller
 * may call parsec_band_regenerate; hicma_gen_finish compresses the pending
 * tiles still outside the band at the wider sketches.  Off-band tiles end as
 * STARSH_gen leaves them; pending tiles the band took in are never
 * compressed at the wider sketches.  info as for STARSH_gen. */
#define HICMA_RANK_PENDING (-2)
int hicma_gen_first(hicma_ctx *ctx, hicma_tlr_desc *A, double tol, int maxrank, int *info);
int hicma_gen_finish(hicma_ctx *ctx, hicma_tlr_desc *A, double tol, int maxrank, int *info);
/* parsec_band_regenerate (hicma_parsec/hicma_parsec.h:108-121): dense band
 * tiles regenerated in HBM; data / kernel / index as for STARSH_gen */
int parsec_band_regenerate(hicma_ctx *ctx, int uplo, hicma_tlr_desc *A, void *data, hicma_kernel *kernel,
                           hicma_index *index, int band_size);

/* ---- the factorization (hicma_parsec/hicma_parsec.h:38-56) ----
 * A is factored in place: band tiles hold L, off-band tiles hold L = U V^T.
 * Returns info: 0 success, > 0 global 1-based row of the first non-positive
 * pivot (k*NB + iinfo), < 0 argument error (codes above).
 * tileopcounters: NT*NT (index m*NT+n), opcounters: >= 1 entry,
 * critical_path_time: 6 entries: summed device seconds of potrf(k), trsm(k+1,k)
 * and syrk(k,k+1) (each a launch group on its own stream), 3..5 zero; pass
 * NULL unless the times are wanted (the events cost a little);
 * any of the three may be NULL. */
int HiCMA_dpotrf_L(hicma_ctx *ctx, int uplo, hicma_tlr_desc *A, hicma_tlr_desc *Ar, hicma_tlr_desc *Rank,
                   double acc, int fixed_rank, int maxrank, int *lookahead, int band_size, int hmb,
                   int compmaxrank, int send_full_tile, int *two_flow, unsigned long *tileopcounters,
                   unsigned long *opcounters, double *critical_path_time, int verbose);

/* HiCMA_dpotrf_L on this descriptor captures its launch sequence (which
 * depends only on NT and the band, not on the ranks: kernels read them from
 * HBM) in a HIP graph on the first call and replays it on later calls while
 * the descriptor's buffers, workspace and scalar arguments are unchanged,
 * instead of the reference's parsec_context_start / wait
 * (hicma_parsec/hicma_parsec.c:551-558, 595-602).  Calls that ask for
 * critical-path times, rank history or profiling run without the graph. */
int hicma_factor_graph(hicma_tlr_desc *A, int enable);

/* Schedule of the low-rank gemm + recompression tasks: panel steps with
 * fewer tasks than ntask run every task on the four-wave kernel (a task's
 * latency, not the step's throughput, bounds such steps); 0 (the default)
 * sorts each step's tasks by stack width over the one-wave kernels.  No counterpart
 * in the reference (its PaRSEC runtime schedules tasks on CPU cores).
 * Results depend on the schedule only through rounding, and are the same for
 * the same schedule. */
int hicma_set_wave_min_tasks(hicma_ctx *ctx, int ntask);

/* Autotuner of that schedule: factors a copy of A (reps times per candidate:
 * 0, 50, 100, 200, ... tasks) and keeps the fastest in ctx; A is unchanged.
 * Returns the chosen threshold and its time.  Needs memory for one more
 * descriptor of A's shape. */
int hicma_autotune_schedule(hicma_ctx *ctx, hicma_tlr_desc *A, double acc, int maxrank, int reps, int *best,
                            double *best_seconds);

/* op_counts (hicma_parsec/op_counts.h:10-72), for callers that account work */
unsigned long hicma_op_counts(char op, unsigned long a, unsigned long b, unsigned long c, unsigned long d);
/* 4 totals of the last factorization: band, offband, critical path, off path */
int hicma_op_stats(const hicma_tlr_desc *A, unsigned long *stats);
/* number of recompressions whose rank was capped at maxrank in the last call */
int hicma_rank_overflows(const hicma_tlr_desc *A);
/* number of recompression SVDs that stopped at the sweep limit in the last call */
int hicma_jacobi_unconverged(const hicma_tlr_desc *A);
/* op counts of the last call per executing stream (the reference's
 * per-thread opcounters, hicma_parsec/testing_dpotrf.c:920-990): ops[0..3] =
 * context stream, two recompression side streams, lookahead stream; entries
 * n > 4 are 0.  They sum to the total. */
int hicma_exec_op_stats(const hicma_tlr_desc *A, unsigned long *ops, int n);
/* record the initial / minimum / maximum / final rank of every off-band tile
 * in the following factorizations (the reference's dcRank) */
int hicma_rank_history(hicma_tlr_desc *A, int enable);
/* the recorded history: h[4*(n*NT+m) + q], q = initial, min, max, final
 * (0 for band tiles) */
int hicma_get_rank_history(const hicma_tlr_desc *A, int *h);
/* parsec_rank_print (hicma_parsec/hicma_parsec.h:168-177, rank_print.jdf):
 * prints the five rank maps of the last factorization to stdout; needs
 * hicma_rank_history(Rank, 1) before it */
int parsec_rank_print(hicma_ctx *ctx, hicma_tlr_desc *Rank, int band_size);
/* low-rank gemm + recompression totals of the last call: op count (op_counts
 * formulas), number of tasks, algorithmic bytes (8 * words read + written) */
int hicma_recompress_stats(const hicma_tlr_desc *A, unsigned long *stats);
/* kernel timing: when enabled, every recompression launch is bracketed by
 * HIP events on the context stream; read returns the summed milliseconds */
int hicma_profile(hicma_ctx *ctx, int enable);
int hicma_profile_read(hicma_ctx *ctx, double *ms_total, int *launches);
/* copy tiles, band and rank map of src into dst (same N, NB, maxrank, kind) */
int hicma_desc_copy(hicma_tlr_desc *dst, const hicma_tlr_desc *src);

/* ---- the --check of the reference harness, on the device ---- */
int hicma_dense_create(hicma_ctx *ctx, hicma_dense **D, int N);
int hicma_dense_destroy(hicma_dense *D);
/* host copies of the whole N x N array */
int hicma_dense_get(const hicma_dense *D, double *host);
int hicma_dense_set(hicma_dense *D, const double *host);
/* exact covariance of the problem of a TLR descriptor, lower triangle (strict
 * upper zeroed): starsh_generate_map of hicma_parsec/testing_dpotrf.c:651 */
int hicma_dense_generate(hicma_ctx *ctx, int uplo, hicma_dense *D, const hicma_tlr_desc *problem);
/* STARSH_check (hicma_parsec/hicma_parsec.h:168-184, STARSH_check.jdf:90-143):
 * uncompress the lower triangle of A into the dense device matrix A0, strict
 * upper triangle zeroed */
int STARSH_check(hicma_ctx *ctx, int uplo, hicma_dense *A0, hicma_tlr_desc *A, hicma_tlr_desc *Ar, int band_size,
                 int maxrank, int *info);
/* dense Cholesky of the checker (dplasma_dpotrf, hicma_parsec/testing_dpotrf.c:766),
 * rocSOLVER with 64-bit indices; *info = LAPACK info */
int hicma_dense_potrf(hicma_ctx *ctx, int uplo, hicma_dense *D, int *info);
/* dplasma_dlansy(Frobenius, lower) of D, and of A - B */
int hicma_dense_lansy(hicma_ctx *ctx, int uplo, const hicma_dense *D, double *norm);
int hicma_dense_diff_lansy(hicma_ctx *ctx, int uplo, const hicma_dense *A, const hicma_dense *B, double *norm);
/* 2 sum log L(i,i) of a dense factor */
int hicma_dense_logdet(hicma_ctx *ctx, const hicma_dense *L, double *logdet);
/* check_dpotrf2 / check_diff (hicma_parsec/testing_dpotrf.c:111-230): return 0
 * (correct) or 1 (suspicious), the ratio in *result; verbose prints the
 * reference's report */
int hicma_check_dpotrf2(hicma_ctx *ctx, int verbose, int uplo, const hicma_dense *L, const hicma_dense *A0,
                        double threshold, double *result);
int hicma_check_diff(hicma_ctx *ctx, int verbose, int uplo, const hicma_dense *L, const hicma_dense *L0,
                     double threshold, double *result);

/* ---- tile access (host copies) ---- */
/* band tile: NB*NB doubles, returns NB in *rank; off-band: NB*2*r doubles
 * (a tile held in FP32 or BF16 is widened to doubles) */
int hicma_get_tile(const hicma_tlr_desc *A, int m, int n, double *buf, int *rank);
int hicma_set_tile(hicma_tlr_desc *A, int m, int n, const double *buf, int rank);
/* whole rank map, NT*NT, as stored on the device */
int hicma_get_rank_map(const hicma_tlr_desc *A, int *G);

/* ---- Gaussian log-likelihood pieces on the factor ----
 * No counterpart in the reference, which stops at the factor; these consume
 * the factor in the layout HiCMA_dpotrf_L leaves. */
/* 2 * sum(log(diag(L))) */
int hicma_logdet(hicma_tlr_desc *A, double *logdet);
/* triangular solves with the TLR factor, B in DEVICE memory (N x nrhs, column
 * major, ldb >= N), in place: trans 0: B <- L^-1 B, trans 1: B <- L^-T B */
int hicma_tlr_trsm(hicma_ctx *ctx, hicma_tlr_desc *A, int trans, int nrhs, double *B, int ldb);
/* B <- Sigma^-1 B = L^-T L^-1 B (dpotrs) */
int hicma_tlr_potrs(hicma_ctx *ctx, hicma_tlr_desc *A, int nrhs, double *B, int ldb);
/* Gaussian log-likelihood of z (DEVICE, N entries, overwritten by L^-1 z):
 * logdet = log det Sigma, quad = z^T Sigma^-1 z,
 * loglik = -(quad + logdet + N log 2 pi) / 2; any output may be NULL */
int hicma_tlr_loglik(hicma_ctx *ctx, hicma_tlr_desc *A, double *z, double *logdet, double *quad, double *loglik);

/* ---- mixed-precision storage of the factor ----
 * No counterpart in the reference, which holds every tile in FP64.  The
 * factorization computes in FP64; afterwards the off-band tiles whose
 * Frobenius norm is small can be held in FP32 or BF16, and the solves, the
 * log-likelihood, STARSH_check and hicma_get_tile read them in that type.
 * hicma_tlr_normf: ||A||_F of the symmetric matrix the descriptor holds
 * (lower tiles counted twice, the padding left out) and, if tile_norms is
 * not NULL, the Frobenius norm of every lower tile (host, NT*NT doubles,
 * column major, zero above the diagonal).  */
int hicma_tlr_normf(hicma_tlr_desc *A, double *normF, double *tile_norms);
/* Round in place every off-band tile with ||L_mn||_F < thr16 to BF16 and
 * every other one with ||L_mn||_F < thr32 to FP32 (thr16 <= thr32; thr16 <= 0
 * disables BF16); with thr32 = eps ||A||_F the rounding error of each tile is
 * below 2^-24 eps ||A||_F.  counts (may be NULL): off-band tiles now held in
 * FP64, FP32, BF16.  HiCMA_dpotrf_L refuses a descriptor holding demoted
 * tiles (HICMA_ERR_ARG); STARSH_gen, and hicma_set_tile for its tile, make
 * tiles FP64 again. */
int hicma_tlr_demote(hicma_tlr_desc *A, double thr32, double thr16, int *counts);

/* hicma_tlr_loglik without any device allocation.  No counterpart in the
 * reference: for callers that cannot allocate (code under HIP graph capture,
 * serving loops).  Its workspace is a 64 MiB device array of the library
 * (8 Mi doubles: the lower 64 x 64 blocks of every L_kk^-1,
 * NT*4096*nb*(nb+1)/2 doubles with nb = (NB+63)/64, plus 2*NT*NB,
 * max(NT*NB/256 + 1, NT) and 8);
 * larger problems return HICMA_ERR_ARG.  Calls that
 * overlap on one GPU race on that array. */
int hicma_tlr_loglik_noalloc(hicma_ctx *ctx, hicma_tlr_desc *A, double *z, double *logdet, double *quad,
                             double *loglik);

/* ---- serving: continuous batching of Gaussian log-likelihood requests ----
 * No counterpart in the reference.  Each request is a statistics-2d-sqexp
 * problem of its own (N points of the seeded jittered grid, kernel
 * parameters, vector z) answered with log det Sigma and z^T Sigma^-1 z:
 * STARSH_gen, band auto-tuning, HiCMA_dpotrf_L and the forward solve on the
 * device, the factor never leaving HBM.  `slots` workers (one context, one
 * stream and ncu/slots CUs each) run requests concurrently; a request is
 * admitted when a worker is free and its HBM footprint fits the budget
 * (FIFO; a later request may pass a head that does not fit, for at most 1 s).
 * After generation a request widens its footprint to the dense band the time
 * model picks (result `band`), waiting for memory; a request's result does
 * not depend on what else runs, except when every running request is waiting
 * to widen: one then takes the widest band that fits the budget (its result
 * agrees within the factorization's accuracy, not bit for bit).  A request
 * whose allocation fails (HICMA_ERR_NOMEM) runs once more, alone. */
typedef struct hicma_server hicma_server;
typedef struct {
    int N, NB, maxrank;           /* maxrank <= NB/2; 0 selects NB/2 */
    double eps;                   /* accuracy (STARSH_gen and recompression) */
    double sigma2, beta, nugget;  /* C(r) = sigma2 exp(-r^2 / (2 beta^2)) + nugget on the diagonal */
    uint64_t seed;                /* point set */
    const double *z;              /* N values (host, copied at submission), or NULL: z ~ N(0, I) from zseed */
    uint64_t zseed;
} hicma_lik_request;
typedef struct {
    int status;                   /* 0, else the failing step's info or HICMA_ERR_* */
    int band;                     /* auto-tuned band size */
    double logdet, quad, loglik;  /* log det Sigma, z^T Sigma^-1 z, -(quad + logdet + N log 2 pi) / 2 */
    double t_queue, t_run, t_total; /* seconds: submission to start, start to end, submission to end */
    size_t bytes;                 /* HBM reserved */
} hicma_lik_result;
int hicma_server_create(hicma_server **s, int device, int slots, double hbm_budget_gb);
int hicma_server_submit(hicma_server *s, const hicma_lik_request *r, uint64_t *ticket);
/* blocks until the request is done */
int hicma_server_wait(hicma_server *s, uint64_t ticket, hicma_lik_result *res);
/* runs what was submitted to the end, then stops */
int hicma_server_destroy(hicma_server *s);
/* z ~ N(0, I) of the requests' zseed (host) */
int hicma_normal(uint64_t seed, int n, double *z);

/* ---- device memory plumbing for callers without a GPU runtime of their own ---- */
int hicma_dev_malloc(hicma_ctx *ctx, size_t bytes, void **p);
int hicma_dev_free(hicma_ctx *ctx, void *p);
/* blocking copy on the context's device: to_device 1 = host -> device, 0 = device -> host */
int hicma_memcpy(hicma_ctx *ctx, void *dst, const void *src, size_t bytes, int to_device);

#ifdef __cplusplus
}
#endif
#endif
