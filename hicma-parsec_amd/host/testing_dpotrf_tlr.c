/*
 * testing_dpotrf_tlr -- C driver of the MI355X TLR Cholesky, with the flags,
 * phase order and output records of the reference harness
 * (hicma_parsec/testing_dpotrf.c:262-1107, flags hicma_parsec/hicma_parsec.c:92-309).
 *
 * Phases: STARSH_gen -> rank gather -> band_size_auto_tuning ->
 * band_regeneration -> [check: approximation error, STARSH_check + dense
 * potrf of the approximate matrix] -> HiCMA_potrf (timed region) -> [check:
 * check_diff against the dense factor, check_dpotrf2 against the exact
 * matrix, log-det against the dense log-det] -> rank gather -> R-LO record.
 * With --check the exit status ORs in both checks, as the reference's ret.
 * Added flags (reference defaults): --beta 0.1 --nugget 0.1 --sigma2 1 --seed 0;
 * --print-rank prints the rank maps (the reference's PRINT_RANK build option).
 */
#define _POSIX_C_SOURCE 200809L
#include <getopt.h>
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

#include "hicma_tlr.h"

static const char *str_problem[8] = {"randtlr",     "ed-2d-sin",  "st-2d-sqexp", "st-3d-sqexp",
                                     "st-3d-exp",   "ed-3d-sin",  "md-3d-virus", "md-3d-cube"};

static double wtime(void)
{
    struct timespec ts;
    clock_gettime(CLOCK_MONOTONIC, &ts);
    return (double)ts.tv_sec + 1e-9 * (double)ts.tv_nsec;
}

static void usage(void)
{
    fprintf(stderr,
            " -N            : dimension of the matrix (required)\n"
            " -t --NB       : tile size (required)\n"
            " -e --fixedacc : accuracy threshold (default 1e-8)\n"
            " -f --fixedrank: fixed rank in recompression (default 0: accuracy mode)\n"
            " -u --maxrank  : max rank (default NB/2)\n"
            " -G --genmaxrank, -U --compmaxrank : default maxrank\n"
            " -D --kind_of_problem : 2 st-2d-sqexp, 3 st-3d-sqexp, 4 st-3d-exp\n"
            " -Z --band     : band size (default 1)\n"
            " -E --auto-band: auto-tune band size (default 1)\n"
            " -Y --lookahead, -W --two-flow, -F --send_full_tile, -z --HNB, -c --cores, -P, -Q\n"
            " -x --check    : dense checks on the device (check_diff, check_dpotrf2, log-det)\n"
            " --print-rank  : print the initial/min/max/final rank maps\n"
            " --mixed       : after the factorization, hold off-band tiles with ||L_mn||_F < eps ||A||_F in FP32\n"
            "                 (< 1e-3 eps ||A||_F in BF16); --check then checks that factor\n"
            " -v --verbose\n"
            " --beta, --nugget, --sigma2, --seed : kernel parameters\n");
}

int main(int argc, char **argv)
{
    int N = 0, NB = 0, fixedrk = 0, maxrank = 0, genmaxrank = 0, compmaxrank = 0, kind = 2;
    int band_size = 1, auto_band = 1, lookahead = -1, two_flow = 0, send_full_tile = 0, HNB = 300;
    int cores = 1, P = 1, Q = 1, check = 0, verbose = 0, print_rank = 0, mixed = 0;
    double tol = 1e-8, beta = 0.1, nugget = 0.1, sigma2 = 1.0;
    unsigned long long seed = 0;
    static struct option lo[] = {
        {"cores", required_argument, 0, 'c'},    {"c", required_argument, 0, 'c'},
        {"P", required_argument, 0, 'P'},        {"Q", required_argument, 0, 'Q'},
        {"N", required_argument, 0, 'N'},        {"NB", required_argument, 0, 't'},
        {"t", required_argument, 0, 't'},        {"check", no_argument, 0, 'x'},
        {"x", no_argument, 0, 'x'},              {"z", required_argument, 0, 'z'},
        {"HNB", required_argument, 0, 'z'},      {"verbose", optional_argument, 0, 'v'},
        {"v", optional_argument, 0, 'v'},        {"help", no_argument, 0, 'h'},
        {"h", no_argument, 0, 'h'},              {"fixedrank", required_argument, 0, 'f'},
        {"f", required_argument, 0, 'f'},        {"fixedacc", required_argument, 0, 'e'},
        {"e", required_argument, 0, 'e'},        {"maxrank", required_argument, 0, 'u'},
        {"u", required_argument, 0, 'u'},        {"genmaxrank", required_argument, 0, 'G'},
        {"G", required_argument, 0, 'G'},        {"compmaxrank", required_argument, 0, 'U'},
        {"U", required_argument, 0, 'U'},        {"band", required_argument, 0, 'Z'},
        {"Z", required_argument, 0, 'Z'},        {"lookahead", required_argument, 0, 'Y'},
        {"Y", required_argument, 0, 'Y'},        {"kind_of_problem", required_argument, 0, 'D'},
        {"D", required_argument, 0, 'D'},        {"send_full_tile", required_argument, 0, 'F'},
        {"F", required_argument, 0, 'F'},        {"auto-band", required_argument, 0, 'E'},
        {"E", required_argument, 0, 'E'},        {"two-flow", required_argument, 0, 'W'},
        {"W", required_argument, 0, 'W'},        {"beta", required_argument, 0, 1001},
        {"nugget", required_argument, 0, 1002},  {"sigma2", required_argument, 0, 1003},
        {"seed", required_argument, 0, 1004},    {"print-rank", no_argument, 0, 1005},
        {"mixed", no_argument, 0, 1006},         {0, 0, 0, 0}};
    int c, opt = 0;
    while ((c = getopt_long_only(argc, argv, "", lo, &opt)) != -1) {
        switch (c) {
        case 'c': cores = atoi(optarg); break;
        case 'P': P = atoi(optarg); break;
        case 'Q': Q = atoi(optarg); break;
        case 'N': N = atoi(optarg); break;
        case 't': NB = atoi(optarg); break;
        case 'x': check = 1; verbose = 1; break;
        case 'z': HNB = atoi(optarg); break;
        case 'f': fixedrk = atoi(optarg); break;
        case 'u': maxrank = atoi(optarg); break;
        case 'G': genmaxrank = atoi(optarg); break;
        case 'U': compmaxrank = atoi(optarg); break;
        case 'e': tol = atof(optarg); break;
        case 'Z': band_size = atoi(optarg); break;
        case 'Y': lookahead = atoi(optarg); break;
        case 'D': kind = atoi(optarg); break;
        case 'F': send_full_tile = atoi(optarg); break;
        case 'E': auto_band = atoi(optarg); break;
        case 'W': two_flow = atoi(optarg); break;
        case 'v': verbose = optarg ? atoi(optarg) : 1; break;
        case 1001: beta = atof(optarg); break;
        case 1002: nugget = atof(optarg); break;
        case 1003: sigma2 = atof(optarg); break;
        case 1004: seed = strtoull(optarg, NULL, 10); break;
        case 1005: print_rank = 1; break;
        case 1006: mixed = 1; break;
        case 'h': usage(); return 0;
        default: usage(); return 1;
        }
    }
    if (N <= 0 || NB <= 0) {
        fprintf(stderr, "#XXXXX the matrix size (N) and tile size (NB) are required\n");
        return 1;
    }
    if (maxrank <= 0) {
        maxrank = NB / 2;
        fprintf(stderr, "Max rank has not been specified. Forced to the size of a tile %d\n", maxrank);
    }
    if (genmaxrank <= 0) genmaxrank = maxrank;
    if (compmaxrank <= 0) compmaxrank = maxrank;
    if (kind < 2 || kind > 4) {
        fprintf(stderr, "Wrong value of \"kind_of_problem\" parameter (supported: 2, 3, 4)\n");
        return 1;
    }
    int NT = (N + NB - 1) / NB;
    if (verbose) {
        printf("M:%d N:%d NB:%d NB:%d HNB:%d HNB:%d\n", N, N, NB, NB, HNB, HNB);
        printf("nodes:1 P:%d Q:%d cores:%d (GPU: 1)\n", P, Q, cores);
        printf("kind_of_problem:%d %s\n", kind, str_problem[kind]);
        printf("tol:%.1e fixed_rk:%d beta:%g nugget:%g sigma2:%g seed:%llu\n", tol, fixedrk, beta, nugget, sigma2,
               seed);
        printf("send_full_tile:%d lookahead: %d band_size:%d max_rank:%d gen:%d comp:%d\n", send_full_tile,
               lookahead, band_size, maxrank, genmaxrank, compmaxrank);
    }
    hicma_ctx *ctx = NULL;
    if (hicma_init(&ctx, 0)) { fprintf(stderr, "hicma_init failed\n"); return 1; }
    int capacity = genmaxrank > maxrank ? genmaxrank : maxrank;
    hicma_tlr_desc *A = NULL;
    int rc = hicma_desc_create(ctx, &A, N, NB, capacity, kind, sigma2, beta, nugget, seed);
    if (rc) { fprintf(stderr, "hicma_desc_create failed: %d\n", rc); return 1; }
    if (auto_band) band_size = 1;

    int info = 0, ret = 0;
    double t0 = wtime();
    rc = STARSH_gen(ctx, HICMA_LOWER, A, A, NULL, NULL, NULL, tol, genmaxrank, band_size, send_full_tile, &info);
    double time_starsh = wtime() - t0;
    printf("[****] TIME(s) %12.5f : STARSH_gen\tPxQ= %3d %-3d NB= %4d N= %7d\n", time_starsh, P, Q, NB, N);
    if (rc || info != 0) {
        printf("Error in low rank matrix generation, info:%d rc:%d\n", info, rc);
        return 1;
    }
    int *G = (int *)calloc((size_t)NT * NT, sizeof(int));
    int imin = -1, imax = -1, bmin = -1, bmax = -1;
    double iavg = -1, bavg = -1;
    t0 = wtime();
    parsec_rank_gather(ctx, A, G, band_size);
    HICMA_get_stat2(G, NT, band_size, &imin, &imax, &iavg);
    printf("[****] TIME(s) %12.5f : Gather rank\tband_size= %d  NT= %4d avg= %lf min= %d max= %d\n", wtime() - t0,
           band_size, NT, iavg, imin, imax);
    bmin = imin; bmax = imax; bavg = iavg;
    double time_opt_band = 0, time_regenerate = 0;
    if (auto_band) {
        t0 = wtime();
        int opt = parsec_band_size_auto_tuning(ctx, A, NULL, G, NB);
        time_opt_band = wtime() - t0;
        printf("[****] TIME(s) %12.5f : band_size_auto_tuning\tPxQ= %3d %-3d NB= %4d N= %7d maxrank= %d "
               "send_full= %d band_size_opt= %d\n\n",
               time_opt_band, P, Q, NB, N, maxrank, send_full_tile, opt);
        t0 = wtime();
        if (opt != 1) {
            band_size = opt;
            parsec_band_regenerate(ctx, HICMA_LOWER, A, NULL, NULL, NULL, band_size);
        }
        time_regenerate = wtime() - t0;
        printf("[****] TIME(s) %12.5f : band_regeneration\tPxQ= %3d %-3d NB= %4d N= %7d maxrank= %d "
               "send_full= %d band_size= %d\n\n",
               time_regenerate, P, Q, NB, N, maxrank, send_full_tile, band_size);
        if (band_size > 1) {
            parsec_rank_gather(ctx, A, G, band_size);
            HICMA_get_stat2(G, NT, band_size, &bmin, &bmax, &bavg);
            printf("[****] TIME(s) %12.5f : Gather rank\tband_size= %d  NT= %4d avg= %lf min= %d max= %d\n", 0.0,
                   band_size, NT, bavg, bmin, bmax);
        }
    }
    double mem = ((double)(NT - band_size) * (NT - band_size + 1) * NB * bavg +
                  (double)band_size * NT * NB * NB) * 8.0 / 1024 / 1024 / 1024;
    double mem_max = ((double)(NT - band_size) * (NT - band_size + 1) * NB * maxrank +
                      (double)band_size * NT * NB * NB) * 8.0 / 1024 / 1024 / 1024;
    printf("[****] TIME(s) %12.5f : Memory_count\tPxQ= %3d %-3d NB= %4d N= %7d maxrank= %d band_size= %d "
           "send_full= %d memory_per_node= %lf GB memory_per_node_max= %lf GB\n",
           0.0, P, Q, NB, N, maxrank, band_size, send_full_tile, mem, mem_max);

    /* --check before the factorization (hicma_parsec/testing_dpotrf.c:640-771):
     * approximation error, then the dense factor L0 of the approximate matrix */
    hicma_dense *dAd = NULL, *dA0 = NULL;
    if (check) {
        if (hicma_dense_create(ctx, &dAd, N) || hicma_dense_create(ctx, &dA0, N)) {
            fprintf(stderr, "not enough device memory for the dense check matrices (2 x %.1f GB)\n",
                    8.0 * N * (double)N / 1e9);
            return 1;
        }
        hicma_dense_generate(ctx, HICMA_LOWER, dAd, A);
        t0 = wtime();
        STARSH_check(ctx, HICMA_LOWER, dA0, A, A, band_size, maxrank, &info);
        printf("[****] TIME(s) %12.5f : STARSH_check\tPxQ= %3d %-3d NB= %4d N= %7d\n", wtime() - t0, P, Q, NB, N);
        double norm1 = 0, norm2 = 0, diff = 0;
        hicma_dense_lansy(ctx, HICMA_LOWER, dAd, &norm1);
        hicma_dense_lansy(ctx, HICMA_LOWER, dA0, &norm2);
        printf("Check matrix generation:\nNorms: map=%e lr_jdf=%e\n", norm1, norm2);
        hicma_dense_diff_lansy(ctx, HICMA_LOWER, dA0, dAd, &diff);
        printf("Norms: map-lr_jdf=%e\nRelative error of approximation: %e\n\n", diff, diff / norm1);
        double dflops = (double)N * N * N / 3.0;
        t0 = wtime();
        int dinfo = 0;
        hicma_dense_potrf(ctx, HICMA_LOWER, dA0, &dinfo);
        double tp = wtime() - t0;
        printf("[****] TIME(s) %12.5f : Dense_potrf\tPxQ= %3d %-3d NB= %4d N= %7d : %14f gflops\n", tp, P, Q, NB, N,
               dflops / 1e9 / tp);
        if (dinfo != 0) {
            printf("-- dense factorization of the approximate matrix failed (info = %d)\n", dinfo);
            ret |= 1;
        }
    }
    if (print_rank) hicma_rank_history(A, 1);
    double normA = 0;
    if (mixed && hicma_tlr_normf(A, &normA, NULL)) ret |= 1;

    unsigned long *tileops = (unsigned long *)calloc((size_t)NT * NT, sizeof(unsigned long));
    unsigned long *opc = (unsigned long *)calloc((size_t)(cores > 4 ? cores : 4), sizeof(unsigned long));
    double cpt[6] = {0};
    double flops = (double)N * ((1. / 6.) * N + 0.5) * N + (1. / 3.) * N + (double)N * ((1. / 6.) * N) * N - N / 6.0;
    t0 = wtime();
    info = HiCMA_dpotrf_L(ctx, HICMA_LOWER, A, A, A, tol, fixedrk, maxrank, &lookahead, band_size, HNB,
                          compmaxrank, send_full_tile, &two_flow, tileops, opc, cpt, verbose);
    double time_hicma = wtime() - t0;
    printf("[****] TIME(s) %12.5f : HiCMA_potrf\tsend_full_tile= %d band_size= %d lookahead= %d "
           "kind_of_problem= %d HNB= %d PxQ= %3d %-3d NB= %4d N= %7d 2flow= %d : %14f gflops\n",
           time_hicma, send_full_tile, band_size, lookahead, kind, HNB, P, Q, NB, N, two_flow,
           flops / 1e9 / time_hicma);
    if (info != 0) {
        printf("-- Factorization is suspicious (info = %d) ! \n", info);
        ret |= 1;
    }
    if (print_rank) parsec_rank_print(ctx, A, band_size);
    if (mixed) {
        /* mixed-precision storage of the factor; the checks below run on it */
        int cnt[3] = {0, 0, 0};
        t0 = wtime();
        if (hicma_tlr_demote(A, tol * normA, 1e-3 * tol * normA, cnt)) ret |= 1;
        printf("[****] TIME(s) %12.5f : Mixed_precision\t||A||_F= %e thresholds %e %e tiles FP64= %d FP32= %d "
               "BF16= %d\n",
               wtime() - t0, normA, tol * normA, 1e-3 * tol * normA, cnt[0], cnt[1], cnt[2]);
    }
    if (check) {
        printf("\nCHECK RESULT:\n\n");
        /* check_diff: L of the TLR factorization against the dense factor of the approximate matrix */
        STARSH_check(ctx, HICMA_LOWER, dAd, A, A, band_size, maxrank, &info);
        printf("+++++++++++++++++++++++++++\nChecking HiCMA L vs dense L\n");
        double r1 = 0, r2 = 0;
        int bad = hicma_check_diff(ctx, 1, HICMA_LOWER, dAd, dA0, tol, &r1);
        ret |= (bad != 0);
        /* check_dpotrf2: L L^T against the exact matrix */
        printf("++++++++++++++++++++++++++\nChecking HiCMA L vs original A via computing LLT\n");
        hicma_dense_generate(ctx, HICMA_LOWER, dA0, A);
        bad = hicma_check_dpotrf2(ctx, 1, HICMA_LOWER, dAd, dA0, tol, &r2);
        ret |= (bad != 0);
        /* log-det against the dense Cholesky of the exact matrix */
        double ld = 0, ldd = 0;
        int dinfo = 0;
        hicma_logdet(A, &ld);
        hicma_dense_potrf(ctx, HICMA_LOWER, dA0, &dinfo);
        hicma_dense_logdet(ctx, dA0, &ldd);
        printf("log det = %.15e dense = %.15e relative error = %e\n", ld, ldd, fabs(ld - ldd) / fabs(ldd));
        hicma_dense_destroy(dAd);
        hicma_dense_destroy(dA0);
    }
    int fmin = -1, fmax = -1;
    double favg = -1;
    parsec_rank_gather(ctx, A, G, band_size);
    HICMA_get_stat2(G, NT, band_size, &fmin, &fmax, &favg);
    printf("[****] TIME(s) %12.5f : Gather rank\tband_size= %d  NT= %4d avg= %lf min= %d max= %d\n", 0.0,
           band_size, NT, favg, fmin, fmax);
    unsigned long st[4] = {0, 0, 0, 0};
    hicma_op_stats(A, st);
    unsigned long total = st[0] + st[1];
    printf("[****] TIME(s) %12.5f : Operation_counts : band_size= %d total= %le op_band= %le op_offband= %le "
           "ratio_band= %lf op_critical_path= %le op_off_critical_path= %le ratio_critical_path= %lf\n",
           0.0, band_size, (double)total, (double)st[0], (double)st[1], st[1] ? (double)st[0] / st[1] : 0.0,
           (double)st[2], (double)st[3], st[3] ? (double)st[2] / st[3] : 0.0);
    /* the per-executor counters sum to the total (hicma_parsec/testing_dpotrf.c:982-1006) */
    unsigned long ex[4] = {0, 0, 0, 0}, sum_exec = 0, sum_tile = 0;
    hicma_exec_op_stats(A, ex, 4);
    for (int i = 0; i < 4; i++) sum_exec += ex[i];
    for (size_t i = 0; i < (size_t)NT * NT; i++) sum_tile += tileops[i];
    if (sum_exec != total || sum_tile != total || st[2] + st[3] != total) {
        printf("-- operation counts do not add up: streams %lu tiles %lu total %lu\n", sum_exec, sum_tile, total);
        ret |= 1;
    }
    printf("[****] TIME(s) %12.5f : Total_critical_path_time %lf, potrf %lf, trsm %lf, syrk %lf\n\n", 0.0,
           cpt[0] + cpt[1] + cpt[2], cpt[0], cpt[1], cpt[2]);
    printf("R-LO %d %d %d %d   %d %d   %d %d %d %d   %d %s %.1e %g %d %g   %d %d %d %d %d %d %d %d   "
           "%g %d %d   %g %d %d   %g %d %d   %lf %lf   %lf %lf %lf %lf    %lf %lf %lf %lf %lf    "
           "%le %le %le %le %le %lf %lf %d  hicma-mi355x \"",
           N, N, NB, NB, HNB, HNB, 1, P, Q, cores, kind, str_problem[kind], tol, 0.0, fixedrk, 0.0, send_full_tile,
           band_size, lookahead, 0, auto_band, maxrank, genmaxrank, compmaxrank, iavg, imin, imax, bavg, bmin,
           bmax, favg, fmin, fmax, mem, mem_max, cpt[0] + cpt[1] + cpt[2], cpt[0], cpt[1], cpt[2], time_starsh,
           time_hicma, time_opt_band, time_regenerate, 0.0, (double)total, (double)st[0], (double)st[1],
           (double)st[2], (double)st[3], st[1] ? (double)st[0] / st[1] : 0.0, st[3] ? (double)st[2] / st[3] : 0.0,
           two_flow);
    for (int i = 0; i < argc; i++) printf("%s ", argv[i]);
    printf("\"\n");
    free(G);
    free(tileops);
    free(opc);
    hicma_desc_destroy(A);
    hicma_fini(ctx);
    return ret;
}
