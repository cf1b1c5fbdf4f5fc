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

/* ---- context and descriptors (replace setup_parsec / descriptor init,
 *      hicma_parsec/hicma_parsec.c:337-427, testing_dpotrf.c:419-453) ---- */
int hicma_init(hicma_ctx **ctx, int device);
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
int STARSH_gen(hicma_ctx *ctx, int uplo, hicma_tlr_desc *A, hicma_tlr_desc *Ar, double tol, int maxrank,
               int band_size, int send_full_tile, int *info);

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
int parsec_band_regenerate(hicma_ctx *ctx, int uplo, hicma_tlr_desc *A, int band_size);

/* ---- the factorization (hicma_parsec/hicma_parsec.h:38-56) ----
 * A is factored in place: band tiles hold L, off-band tiles hold L = U V^T.
 * Returns info: 0 success, > 0 global 1-based row of the first non-positive
 * pivot (k*NB + iinfo), < 0 argument error (codes above).
 * tileopcounters: NT*NT (index m*NT+n), opcounters: >= 1 entry,
 * critical_path_time: 6 entries (potrf, trsm, syrk GPU seconds; 3..5 zero);
 * any of the three may be NULL. */
int HiCMA_dpotrf_L(hicma_ctx *ctx, int uplo, hicma_tlr_desc *A, hicma_tlr_desc *Ar, hicma_tlr_desc *Rank,
                   double acc, int fixed_rank, int maxrank, int *lookahead, int band_size, int hmb,
                   int compmaxrank, int send_full_tile, int *two_flow, unsigned long *tileopcounters,
                   unsigned long *opcounters, double *critical_path_time, int verbose);

/* op_counts (hicma_parsec/op_counts.h:10-72), for callers that account work */
unsigned long hicma_op_counts(char op, unsigned long a, unsigned long b, unsigned long c, unsigned long d);
/* 4 totals of the last factorization: band, offband, critical path, off path */
int hicma_op_stats(const hicma_tlr_desc *A, unsigned long *stats);
/* number of recompressions whose rank was capped at maxrank in the last call */
int hicma_rank_overflows(const hicma_tlr_desc *A);
/* low-rank gemm + recompression totals of the last call: op count (op_counts
 * formulas), number of tasks, algorithmic bytes (8 * words read + written) */
int hicma_recompress_stats(const hicma_tlr_desc *A, unsigned long *stats);
/* kernel timing: when enabled, every recompression launch is bracketed by
 * HIP events on the context stream; read returns the summed milliseconds */
int hicma_profile(hicma_ctx *ctx, int enable);
int hicma_profile_read(hicma_ctx *ctx, double *ms_total, int *launches);
/* copy tiles, band and rank map of src into dst (same N, NB, maxrank, kind) */
int hicma_desc_copy(hicma_tlr_desc *dst, const hicma_tlr_desc *src);

/* ---- STARSH_check (hicma_parsec/hicma_parsec.h:168-184): uncompress the
 * lower triangle into a host N x N column-major array (upper zeroed) ---- */
int STARSH_check(hicma_ctx *ctx, int uplo, double *A0, hicma_tlr_desc *A, hicma_tlr_desc *Ar, int band_size,
                 int maxrank, int *info);

/* ---- tile access (host copies) ---- */
/* band tile: NB*NB doubles, returns NB in *rank; off-band: NB*2*r doubles */
int hicma_get_tile(const hicma_tlr_desc *A, int m, int n, double *buf, int *rank);
int hicma_set_tile(hicma_tlr_desc *A, int m, int n, const double *buf, int rank);
/* whole rank map, NT*NT, as stored on the device */
int hicma_get_rank_map(const hicma_tlr_desc *A, int *G);

/* ---- Gaussian log-likelihood pieces on the factor ---- */
/* 2 * sum(log(diag(L))) */
int hicma_logdet(hicma_tlr_desc *A, double *logdet);

#ifdef __cplusplus
}
#endif
#endif
