/*
 * tlr_oracle.h -- CPU restatement of hicma-parsec's TLR Cholesky path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker: only tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg load it.  The
 * product (libhicma_tlr.so) never links, loads or calls it.
 *
 * Every function follows a reference file:line, cited in tlr_oracle.c.
 * Storage follows the reference descriptor (hicma_parsec/testing_dpotrf.c:419-453):
 *   - band tile (m,n), m-n < band: NB x NB doubles, column major, ld = NB;
 *   - off-band tile (m,n), m-n >= band: compact [U | V], U = NB x r (ld NB),
 *     V = NB x r (ld NB) at offset NB*r, A(m,n) ~= U V^T;
 *   - rank map: int32 NT x NT, column major, G[n*NT + m].
 */
#ifndef TLR_ORACLE_H
#define TLR_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* problem kinds, numbered like the reference's -D flag */
#define ORC_KIND_SQEXP_2D 2
#define ORC_KIND_SQEXP_3D 3
#define ORC_KIND_EXP_3D   4

typedef struct orc_tlr orc_tlr;

/* seeded jittered grid, Morton ordered, coordinates stored xyz[d*N + i] */
int  orc_points(int N, int ndim, uint64_t seed, double *xyz);
/* standard normal sketch matrix, column major ld x ncol */
void orc_sketch(int ld, int ncol, uint64_t seed, double *omega);
/* covariance block rows [row0,row0+nrows) x cols [col0,col0+ncols) */
void orc_kernel_block(int kind, int nrows, int ncols, const double *xyz, int N,
                      int row0, int col0, double sigma2, double beta,
                      double nugget, double *out, int ld);

orc_tlr *orc_create(int N, int NB, int kind, double sigma2, double beta,
                    double nugget, uint64_t seed);
void orc_free(orc_tlr *t);
int  orc_NT(const orc_tlr *t);
const double *orc_xyz(const orc_tlr *t);

/* STARSH_gen: dense band + randomized-SVD off-band; returns info (-1: a tile
 * is not compressible under maxrank) */
int  orc_generate(orc_tlr *t, double tol, int maxrank, int band_size);
/* band regeneration after auto-tuning */
int  orc_band_regenerate(orc_tlr *t, int band_size);
int  orc_band_size(const orc_tlr *t);
/* rank gather / stat / check */
void orc_rank_gather(const orc_tlr *t, int *G);
void orc_rank_stat(const int *G, int NT, int band, int *mn, int *mx, double *avg);
int  orc_rank_check(int *G, int NT, int band);
/* band size auto-tuning from a gathered rank map */
int  orc_band_autotune(const int *G, int NT, int NB);
/* reference op_counts */
unsigned long orc_op_counts(char op, unsigned long a, unsigned long b,
                            unsigned long c, unsigned long d);

/* HiCMA_dpotrf_L: factor in place.  kmax < 0 runs all panels; kmax >= 0 stops
 * after panel kmax-1 (a bounded sample used for CPU timing).  stats[4] =
 * op_band, op_offband, op_path, op_offpath.  tileops: NT*NT (m*NT+n). */
int  orc_dpotrf(orc_tlr *t, double acc, int rk, int maxrank, int kmax,
                unsigned long *tileops, unsigned long *stats, double *seconds);

/* uncompress lower triangle into dense N x N column-major (upper zeroed) */
void orc_uncompress(const orc_tlr *t, double *dense);
/* exact dense covariance, lower triangle (upper zeroed) */
void orc_dense_exact(const orc_tlr *t, double *dense);
double orc_logdet(const orc_tlr *t);
/* ||A0 - L L^T||_F / ||A0||_F with dlansy(F, lower) semantics */
double orc_check_dpotrf2(int N, const double *L, const double *A0);
/* ||L - L0|| / ||L|| with dlansy(F, lower) semantics (check_diff) */
double orc_check_diff(int N, const double *L, const double *L0);
/* dense lower Cholesky in place; returns LAPACK info */
int  orc_dense_potrf(int N, double *A);
double orc_dense_logdet(int N, const double *L);

/* triangular solves with the factor: trans 0: B <- L^{-1} B, 1: B <- L^{-T} B
 * (B: N x nrhs, column major, ld ldb) */
int  orc_tlr_trsm(const orc_tlr *t, int trans, int nrhs, double *B, int ldb);

/* threads of the OpenMP task graph (returns the count in effect) and the
 * BLAS library's version string, for the CPU baseline's report */
int  orc_set_threads(int n);
void orc_blas_version(char *buf, int len);

/* tile access for comparisons / import of a device-generated matrix */
int  orc_get_rank(const orc_tlr *t, int m, int n);
int  orc_get_tile(const orc_tlr *t, int m, int n, double *buf);
int  orc_set_band(orc_tlr *t, int band_size);
int  orc_set_tile(orc_tlr *t, int m, int n, const double *buf, int rank);

#ifdef __cplusplus
}
#endif
#endif
