/*
 * og_stats.c -- the reference's PageRank rank comparison (-S,
 * graphStats.c:580-1020) with its statistics computed on the GPU
 * (og_stats.hip): ogpu_rank_stats / ogpu_rank_stats_device, and the drop-in
 * collectStatsPageRank_topK and collectStatsPageRank, whose printed tables
 * and .stats file follow the reference's text.
 */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "og_internal.h"
#include "og_pr.h"
#include "og_stats.h"
#include "opengraph_amd.h"
#include "opengraph_amd_ext.h"

_Static_assert(OG_KT_SMALL_K == OG_LEV_SMALL_K, "the small problems share one list");

/* graphStats.c:472-486 from the pair counts. */
static float og_kendall_cor(uint32_t K, long long s, long long m1, long long m2)
{
    const long long nPair = (long long)K * ((long long)K - 1) / 2;
    if (m1 < nPair && m2 < nPair)
        return (float)s / (sqrtf((float)(nPair - m1)) * sqrtf((float)(nPair - m2)));
    return 0.0f;
}

int ogpu_rank_stats_device(const float *d_ref_pr, const uint32_t *d_ref_order, const float *d_pr,
                           const uint32_t *d_order, uint32_t N, const uint32_t *topk,
                           const uint32_t *num_vertices, uint32_t P, struct PageRankCorrelationStats *out)
{
    if (og_ctx_ensure()) return -1;
    if (P == 0) return 0;
    hipStream_t s = og_stream();
    uint32_t *k = (uint32_t *)malloc((size_t)P * 4), *off = (uint32_t *)malloc((size_t)(P + 1) * 4);
    if (!k || !off) { free(k); free(off); og_set_error("out of host memory"); return -1; }
    uint64_t total = 0, groups = 0;
    uint32_t n_small = 0, n_large = 0, n_glob = 0, kmax = 0;
    for (uint32_t p = 0; p < P; p++)
    {
        if (num_vertices[p] > N)
        {
            og_set_error("rank stats problem %u: num_vertices %u above N = %u", p, num_vertices[p], N);
            free(k); free(off);
            return -1;
        }
        k[p] = topk[p] < num_vertices[p] ? topk[p] : num_vertices[p];   /* graphStats.c:588-589 */
        off[p] = (uint32_t)total;
        total += k[p];
        if (k[p] > OG_KT_SMALL_K) groups += (k[p] + 7) / 8;
        if (k[p] <= OG_LEV_SMALL_K) n_small++;
        else if (k[p] <= OG_LEV_LDS_K) n_large++;
        else { n_glob++; if (k[p] > kmax) kmax = k[p]; }
    }
    off[P] = (uint32_t)total;
    if (total > 0xFFFFFFFFull || groups > 0x7FFFFFFFull)
    {
        og_set_error("rank stats: %llu slice entries is too many", (unsigned long long)total);
        free(k); free(off);
        return -1;
    }
    uint32_t *plist = (uint32_t *)malloc((size_t)P * 4);
    uint32_t *grp = (uint32_t *)malloc((size_t)(groups ? groups : 1) * 8);
    uint32_t *lev = (uint32_t *)malloc((size_t)P * 4), *inter = (uint32_t *)malloc((size_t)P * 4);
    uint32_t *mism = (uint32_t *)malloc((size_t)P * 4);
    long long *kt = (long long *)malloc((size_t)P * 48);
    double *ef = (double *)malloc((size_t)P * 8), *er = (double *)malloc((size_t)P * 8);
    og_stats_dev a;
    memset(&a, 0, sizeof(a));
    void *dmem[16] = {0};
    int rc = -1;
    if (!plist || !grp || !lev || !inter || !mism || !kt || !ef || !er) { og_set_error("out of host memory"); goto done; }
    {
        /* problem lists: small LDS, large LDS, global */
        uint32_t is = 0, il = n_small, ig = n_small + n_large;
        for (uint32_t p = 0; p < P; p++)
        {
            if (k[p] <= OG_LEV_SMALL_K) plist[is++] = p;
            else if (k[p] <= OG_LEV_LDS_K) plist[il++] = p;
            else plist[ig++] = p;
        }
        uint64_t gi = 0;
        for (uint32_t p = 0; p < P; p++)
            for (uint32_t i = 0; k[p] > OG_KT_SMALL_K && i < k[p]; i += 8)
            {
                grp[2 * gi] = p;
                grp[2 * gi + 1] = i;
                gi++;
            }
    }
    const size_t T = total ? (size_t)total : 1;
    dmem[0] = og_dalloc((size_t)P * 4);            /* k */
    dmem[1] = og_dalloc((size_t)P * 4);            /* m */
    dmem[2] = og_dalloc((size_t)(P + 1) * 4);      /* off */
    dmem[3] = og_dalloc((size_t)N * 8 + 8);        /* inv, ref_inv */
    dmem[4] = og_dalloc(T * 24);                   /* six slices */
    dmem[5] = og_dalloc((size_t)P * 4);            /* plist */
    dmem[6] = og_dalloc((size_t)(groups ? groups : 1) * 8);
    dmem[7] = og_dalloc((size_t)P * 80);           /* kt 48, lev 4, inter 4, mism 4, pad 4, err 16 */
    dmem[8] = n_glob ? og_dalloc((size_t)3 * (kmax + 1) * 4) : NULL;
    for (int i = 0; i < 8; i++)
        if (!dmem[i]) goto done;
    if (n_glob && !dmem[8]) goto done;
    a.N = N;
    a.P = P;
    a.ref_pr = d_ref_pr;
    a.pr = d_pr;
    a.ref_order = d_ref_order;
    a.order = d_order;
    a.k = (const uint32_t *)dmem[0];
    a.m = (const uint32_t *)dmem[1];
    a.off = (const uint32_t *)dmem[2];
    a.inv = (uint32_t *)dmem[3];
    a.ref_inv = a.inv + N;
    a.ids = (uint32_t *)dmem[4];
    a.ref_ids = a.ids + T;
    a.f = (float *)(a.ref_ids + T);
    a.ref_f = a.f + T;
    a.re = a.ref_f + T;
    a.ref_re = a.re + T;
    a.kt = (long long *)dmem[7];
    a.lev = (uint32_t *)(a.kt + 6 * (size_t)P);
    a.inter = a.lev + P;
    a.mism = a.inter + P;
    a.err_f = (double *)dmem[7] + 8 * (size_t)P;
    a.err_r = a.err_f + P;
    if (og_hip_check(hipMemcpyAsync((void *)a.k, k, (size_t)P * 4, hipMemcpyHostToDevice, s), "H2D k") ||
        og_hip_check(hipMemcpyAsync((void *)a.m, num_vertices, (size_t)P * 4, hipMemcpyHostToDevice, s), "H2D m") ||
        og_hip_check(hipMemcpyAsync((void *)a.off, off, (size_t)(P + 1) * 4, hipMemcpyHostToDevice, s), "H2D off") ||
        og_hip_check(hipMemcpyAsync(dmem[5], plist, (size_t)P * 4, hipMemcpyHostToDevice, s), "H2D plist") ||
        (groups && og_hip_check(hipMemcpyAsync(dmem[6], grp, (size_t)groups * 8, hipMemcpyHostToDevice, s), "H2D groups")) ||
        og_launch_fill_u32((uint32_t *)a.kt, 0, 12ull * P, s) ||
        og_launch_stats_inverse(&a, s) || og_launch_stats_slices(&a, (uint32_t)total, s) ||
        og_launch_lev_small(&a, (const uint32_t *)dmem[5], n_small, s) ||
        og_launch_lev_large(&a, (const uint32_t *)dmem[5] + n_small, n_large, s) ||
        og_launch_kendall_small(&a, (const uint32_t *)dmem[5], n_small, s) ||
        og_launch_kendall(&a, dmem[6], (uint32_t)groups, s) || og_launch_topk_sums(&a, s))
        goto done;
    for (uint32_t g = 0; g < n_glob; g++)
        if (og_launch_lev_global(&a, plist[n_small + n_large + g], (uint32_t *)dmem[8], s)) goto done;
    if (og_hip_check(hipMemcpyAsync(kt, a.kt, (size_t)P * 48, hipMemcpyDeviceToHost, s), "D2H kt") ||
        og_hip_check(hipMemcpyAsync(lev, a.lev, (size_t)P * 4, hipMemcpyDeviceToHost, s), "D2H lev") ||
        og_hip_check(hipMemcpyAsync(inter, a.inter, (size_t)P * 4, hipMemcpyDeviceToHost, s), "D2H inter") ||
        og_hip_check(hipMemcpyAsync(mism, a.mism, (size_t)P * 4, hipMemcpyDeviceToHost, s), "D2H mism") ||
        og_hip_check(hipMemcpyAsync(ef, a.err_f, (size_t)P * 8, hipMemcpyDeviceToHost, s), "D2H ef") ||
        og_hip_check(hipMemcpyAsync(er, a.err_r, (size_t)P * 8, hipMemcpyDeviceToHost, s), "D2H er") ||
        og_hip_check(hipStreamSynchronize(s), "sync rank stats"))
        goto done;
    for (uint32_t p = 0; p < P; p++)
    {
        out[p].levenshtein_distance = lev[p];
        out[p].float_Kendall = og_kendall_cor(k[p], kt[6 * p], kt[6 * p + 1], kt[6 * p + 2]);
        out[p].real_Kendall = og_kendall_cor(k[p], kt[6 * p + 3], kt[6 * p + 4], kt[6 * p + 5]);
        out[p].intersection = inter[p];
        out[p].mismatch = mism[p];
        out[p].avg_error_float = ef[p];
        out[p].avg_error_relative = er[p];
    }
    rc = 0;
done:
    if (rc) (void)hipStreamSynchronize(s);
    for (int i = 0; i < 9; i++) og_dfree(dmem[i]);
    free(k); free(off); free(plist); free(grp); free(lev); free(inter); free(mism); free(kt); free(ef); free(er);
    return rc;
}

int ogpu_rank_stats(const float *ref_pr, const uint32_t *ref_order, const float *pr, const uint32_t *order,
                    uint32_t N, const uint32_t *topk, const uint32_t *num_vertices, uint32_t P,
                    struct PageRankCorrelationStats *out)
{
    if (og_ctx_ensure()) return -1;
    hipStream_t s = og_stream();
    void *d = og_dalloc((size_t)N * 16 + 16);
    if (!d) return -1;
    float *d_ref_pr = (float *)d, *d_pr = d_ref_pr + N;
    uint32_t *d_ref_order = (uint32_t *)(d_pr + N), *d_order = d_ref_order + N;
    int rc = -1;
    if (!og_hip_check(hipMemcpyAsync(d_ref_pr, ref_pr, (size_t)N * 4, hipMemcpyHostToDevice, s), "H2D") &&
        !og_hip_check(hipMemcpyAsync(d_pr, pr, (size_t)N * 4, hipMemcpyHostToDevice, s), "H2D") &&
        !og_hip_check(hipMemcpyAsync(d_ref_order, ref_order, (size_t)N * 4, hipMemcpyHostToDevice, s), "H2D") &&
        !og_hip_check(hipMemcpyAsync(d_order, order, (size_t)N * 4, hipMemcpyHostToDevice, s), "H2D"))
        rc = ogpu_rank_stats_device(d_ref_pr, d_ref_order, d_pr, d_order, N, topk, num_vertices, P, out);
    (void)hipStreamSynchronize(s);
    og_dfree(d);
    return rc;
}

/* ------------------------------------------------------------------ */
/* drop-in entry points (graphStats.c:580-1020)                        */
/* ------------------------------------------------------------------ */

static void og_stats_fatal(const char *where)
{
    fprintf(stderr, "opengraph_amd: %s: %s\n", where, og_get_error());
    exit(EXIT_FAILURE);
}

/* graphStats.c:622-661 */
static void og_print_topk(FILE *o, uint32_t topk, const struct PageRankCorrelationStats *c, int file_form)
{
    fprintf(o, "\n-----------------------------------------------------\n");
    fprintf(o, "topk:         %u \n", topk);
    fprintf(o, "-----------------------------------------------------\n");
    fprintf(o, "levenshtein_distance: %u \n", c->levenshtein_distance);
    fprintf(o, "Rank float Kendall:   %lf\n", c->float_Kendall);
    fprintf(o, "Rank real  Kendall:   %lf\n", c->real_Kendall);
    fprintf(o, "intersection:         %u \n", c->intersection);
    fprintf(o, "mismatch:             %u \n", c->mismatch);
    fprintf(o, file_form ? "avg_error_float:      %lf\n" : "avg_error_float:      %.22lf\n", c->avg_error_float);
    fprintf(o, file_form ? "avg_error_relative:   %lf\n" : "avg_error_relative:   %.22lf\n",
            c->avg_error_relative);
    fprintf(o, "-----------------------------------------------------\n");
}

struct PageRankCorrelationStats collectStatsPageRank_topK(struct PageRankStats *ref_stats, struct PageRankStats *stats,
                                                          uint32_t *ref_rankedVertices_total,
                                                          uint32_t *ref_rankedVertices_inverse,
                                                          uint32_t *rankedVertices_inverse, uint32_t topk,
                                                          uint32_t num_vertices, FILE *fptr, uint32_t verbose)
{
    /* The inverse rankings are recomputed on the device from the two
     * realRanks arrays, as collectStatsPageRank forms them
     * (graphStats.c:746-751). */
    (void)ref_rankedVertices_total;
    (void)ref_rankedVertices_inverse;
    (void)rankedVertices_inverse;
    struct PageRankCorrelationStats c;
    memset(&c, 0, sizeof(c));
    if (topk > num_vertices) topk = num_vertices;
    if (ogpu_rank_stats(ref_stats->pageRanks, ref_stats->realRanks, stats->pageRanks, stats->realRanks,
                        ref_stats->num_vertices, &topk, &num_vertices, 1, &c))
        og_stats_fatal("rank statistics");
    if (verbose > 0)
    {
        og_print_topk(stdout, topk, &c, 0);
        if (fptr) og_print_topk(fptr, topk, &c, 1);
    }
    return c;
}

/* graphStats.c:763-836: one row of the top-k table. */
static void og_table(FILE *o, const uint32_t *ks, uint32_t n, const struct PageRankCorrelationStats *c)
{
    const char *rule = "----------------------------------------------------------------------------------------------------------\n";
    fprintf(o, "%s", rule);
    fprintf(o, "Top K                 ");
    for (uint32_t x = 0; x < n; x++) fprintf(o, "%-14u ", ks[x]);
    fprintf(o, "\n%s", rule);
    fprintf(o, "levenshtein_distance  ");
    for (uint32_t x = 0; x < n; x++) fprintf(o, "%-14u ", c[x].levenshtein_distance);
    fprintf(o, "\nRank float Kendall    ");
    for (uint32_t x = 0; x < n; x++) fprintf(o, "%-14lf ", c[x].float_Kendall);
    fprintf(o, "\nRank real  Kendall    ");
    for (uint32_t x = 0; x < n; x++) fprintf(o, "%-14lf ", c[x].real_Kendall);
    fprintf(o, "\nintersection          ");
    for (uint32_t x = 0; x < n; x++) fprintf(o, "%-14u ", c[x].intersection);
    fprintf(o, "\nmismatch              ");
    for (uint32_t x = 0; x < n; x++) fprintf(o, "%-14u ", c[x].mismatch);
    fprintf(o, "\navg_error_float       ");
    for (uint32_t x = 0; x < n; x++) fprintf(o, "%-14lf ", c[x].avg_error_float);
    fprintf(o, "\navg_error_relative    ");
    for (uint32_t x = 0; x < n; x++) fprintf(o, "%-14lf ", c[x].avg_error_relative);
    fprintf(o, "\n%s", rule);
}

static void og_avg_block(FILE *o, const char *head, const struct PageRankCorrelationStats *a)
{
    fprintf(o, "-----------------------------------------------------\n");
    fprintf(o, "%s", head);
    fprintf(o, "-----------------------------------------------------\n");
    fprintf(o, "levenshtein_distance: %u \n", a->levenshtein_distance);
    fprintf(o, "Rank float Kendall:   %lf\n", a->float_Kendall);
    fprintf(o, "Rank real  Kendall:   %lf\n", a->real_Kendall);
    fprintf(o, "intersection:         %u \n", a->intersection);
    fprintf(o, "mismatch:             %u \n", a->mismatch);
    fprintf(o, "avg_error_float:      %lf\n", a->avg_error_float);
    fprintf(o, "avg_error_relative:   %lf\n", a->avg_error_relative);
    fprintf(o, "-----------------------------------------------------\n");
}

void collectStatsPageRank(struct Arguments *arguments, struct PageRankStats *ref_stats, struct PageRankStats *stats,
                          uint32_t trial)
{
    static const uint32_t topK_array[6] = {30, 100, 300, 1000, 5000, 10000};
    const uint32_t N = ref_stats->num_vertices;
    uint32_t topk = arguments->binSize < N ? arguments->binSize : N;
    uint32_t chunk_x = 1000, chunk_num = (N + chunk_x - 1) / chunk_x;
    const int chunked = arguments->verbosity > 1;
    const int single = chunked && chunk_num == 1;
    if (single) chunk_x = N;

    /* every comparison of the call, as one batch: the top-k table, the
     * chunks (verbosity > 1) and the binSize one */
    uint32_t nk = 0;
    while (nk < 6 && N >= topK_array[nk]) nk++;
    const uint32_t n_chunks = chunked ? (single ? 1 : chunk_num) : 0;
    const uint32_t P = nk + n_chunks + 1;
    uint32_t *ks = (uint32_t *)malloc((size_t)P * 4), *ms = (uint32_t *)malloc((size_t)P * 4);
    struct PageRankCorrelationStats *c = (struct PageRankCorrelationStats *)calloc(P, sizeof(*c));
    if (!ks || !ms || !c)
    {
        og_set_error("out of host memory");
        og_stats_fatal("rank statistics");
    }
    for (uint32_t x = 0; x < nk; x++) { ks[x] = topK_array[x]; ms[x] = N; }
    for (uint32_t x = 0; x < n_chunks; x++)
    {
        ks[nk + x] = chunk_x;
        ms[nk + x] = single ? N : N - chunk_x * x;
    }
    ks[P - 1] = topk;
    ms[P - 1] = N;
    if (ogpu_rank_stats(ref_stats->pageRanks, ref_stats->realRanks, stats->pageRanks, stats->realRanks, N, ks, ms,
                        P, c))
        og_stats_fatal("rank statistics");

    char *fname_txt = (char *)malloc(strlen(arguments->fnameb) + 50);
    sprintf(fname_txt, "%s_%d_%d_%d_%d.%s", arguments->fnameb, arguments->algorithm, arguments->datastructure, trial,
            arguments->pushpull, "stats");
    FILE *fptr = fopen(fname_txt, "a+");
    if (!fptr)
    {
        og_set_error("cannot open %s", fname_txt);
        og_stats_fatal("rank statistics");
    }
    if (arguments->verbosity > 0)
    {
        og_table(stdout, ks, nk, c);
        og_table(fptr, ks, nk, c);
    }
    struct PageRankCorrelationStats sum, avg;
    memset(&sum, 0, sizeof(sum));
    for (uint32_t x = 0; x < n_chunks; x++)
    {
        const struct PageRankCorrelationStats *q = &c[nk + x];
        if (single)
        {
            og_print_topk(stdout, ks[nk + x], q, 0);
            og_print_topk(fptr, ks[nk + x], q, 1);
        }
        sum.levenshtein_distance += q->levenshtein_distance;
        sum.float_Kendall += q->float_Kendall;
        sum.real_Kendall += q->real_Kendall;
        sum.intersection += q->intersection;
        sum.mismatch += q->mismatch;
        sum.avg_error_float += q->avg_error_float;
        sum.avg_error_relative += q->avg_error_relative;
    }
    if (single) chunk_num = 1;
    avg.levenshtein_distance = sum.levenshtein_distance / chunk_num;
    avg.float_Kendall = sum.float_Kendall / chunk_num;
    avg.real_Kendall = sum.real_Kendall / chunk_num;
    avg.intersection = sum.intersection / chunk_num;
    avg.mismatch = sum.mismatch / chunk_num;
    avg.avg_error_float = sum.avg_error_float / chunk_num;
    avg.avg_error_relative = sum.avg_error_relative / chunk_num;

    og_print_topk(stdout, topk, &c[P - 1], 0);
    og_print_topk(fptr, topk, &c[P - 1], 1);

    char head[128];
    snprintf(head, sizeof(head), "Avg (Sum(bin)*n)/n:    (Sum(%u)*%u)/%u \n", chunk_x, chunk_num, chunk_num);
    og_avg_block(stdout, head, &avg);
    fprintf(stdout, "-----------------------------------------------------\n");
    fprintf(stdout, "numThreads:           %u \n", arguments->pre_numThreads);
    fprintf(stdout, "Time (S):             %lf\n", stats->time_total);
    fprintf(stdout, "Iterations:           %u \n", stats->iterations);
    fprintf(stdout, "-----------------------------------------------------\n");
    if (arguments->verbosity > 0)
    {
        snprintf(head, sizeof(head), "Avg (Sum_n(bins))/n:    (Sum_%u(%u)/%u \n", chunk_num, chunk_x, chunk_num);
        og_avg_block(fptr, head, &avg);
        fprintf(fptr, "-----------------------------------------------------\n");
        fprintf(fptr, "numThreads:           %u \n", arguments->pre_numThreads);
        fprintf(fptr, "Time (S):             %lf\n", stats->time_total);
        fprintf(fptr, "Iterations:           %u \n", stats->iterations);
        fprintf(fptr, "-----------------------------------------------------\n");
        if (arguments->verbosity > 2)
        {
            /* graphStats.c:1004-1022 */
            fprintf(fptr, " ----------------------------------------------------- ");
            fprintf(fptr, " -----------------------------------------------------\n");
            fprintf(fptr, "| %-14s | %-14s | %-17s | ", "Ref Rank", "Vertex", "PageRank");
            fprintf(fptr, "| %-14s | %-14s | %-17s | \n", "Rank", "Vertex", "PageRank");
            fprintf(fptr, " ----------------------------------------------------- ");
            fprintf(fptr, " -----------------------------------------------------\n");
            for (uint32_t v = N - topk; v < N; v++)
            {
                fprintf(fptr, "| %-14u | %-14u | %-10.15lf | ", v, ref_stats->realRanks[v],
                        ref_stats->pageRanks[ref_stats->realRanks[v]]);
                fprintf(fptr, "| %-14u | %-14u | %-10.15lf | \n", v, stats->realRanks[v],
                        stats->pageRanks[stats->realRanks[v]]);
            }
            fprintf(fptr, " ----------------------------------------------------- ");
            fprintf(fptr, " -----------------------------------------------------\n");
        }
    }
    fclose(fptr);
    free(fname_txt);
    free(ks);
    free(ms);
    free(c);
}
