/*
 * ref_lib.c -- TEST INFRASTRUCTURE ONLY.
 *
 * Shared-library wrapper around the reference's own compiled PageRank
 * (pageRank.c:871 pageRankGraphCSR), built by oracle/Makefile target "ref"
 * into oracle/_ref/libogref.so.  It wraps caller-owned CSR arrays in the
 * reference's GraphCSR / Vertex / EdgeList structs (graphCSR.h:12-29) and
 * runs one variant with the reference's stdout table silenced, so the
 * reference itself can serve as a timed CPU baseline and as a live parity
 * check.  Nothing here computes PageRank: the reference objects do.
 */
#include <fcntl.h>
#include <omp.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <unistd.h>

#include "edgeList.h"
#include "graphConfig.h"
#include "graphCSR.h"
#include "graphStats.h"
#include "mt19937.h"
#include "myMalloc.h"
#include "pageRank.h"
#include "reorder.h"
#include "SPMV.h"
#include "sortRun.h"

int ogref_pagerank(uint32_t N, uint32_t E, uint32_t *f_out_degree, uint32_t *f_in_degree,
                   uint32_t *f_edges_idx, uint32_t *f_src, uint32_t *f_dst,
                   uint32_t *i_out_degree, uint32_t *i_in_degree, uint32_t *i_edges_idx,
                   uint32_t *i_src, uint32_t *i_dst, uint32_t pushpull, double epsilon,
                   uint32_t iterations, int threads, float *ranks_out, uint32_t *iters_out,
                   double *error_out, double *seconds_out)
{
    struct Vertex fv = {f_out_degree, f_in_degree, f_edges_idx};
    struct Vertex iv = {i_out_degree, i_in_degree, i_edges_idx};
    struct EdgeList fe, ie;
    memset(&fe, 0, sizeof(fe));
    memset(&ie, 0, sizeof(ie));
    fe.num_edges = ie.num_edges = E;
    fe.num_vertices = ie.num_vertices = N;
    fe.avg_degree = ie.avg_degree = N ? E / N : 0;
    fe.edges_array_src = f_src; fe.edges_array_dest = f_dst;
    ie.edges_array_src = i_src; ie.edges_array_dest = i_dst;
    struct GraphCSR g;
    memset(&g, 0, sizeof(g));
    g.num_edges = E;
    g.num_vertices = N;
    g.avg_degree = N ? E / N : 0;
    g.vertices = &fv;
    g.sorted_edges_array = &fe;
    g.inverse_vertices = &iv;
    g.inverse_sorted_edges_array = &ie;

    struct Arguments a;
    memset(&a, 0, sizeof(a));
    a.iterations = iterations;
    a.epsilon = epsilon;
    a.pushpull = pushpull;
    a.algorithm = 1;
    a.pre_numThreads = a.algo_numThreads = a.ker_numThreads = threads;
    initializeMersenneState(&a.mt19937var, 27491095);
    omp_set_num_threads(threads);

    fflush(stdout);
    int saved = dup(1);
    int devnull = open("/dev/null", O_WRONLY);
    if (saved >= 0 && devnull >= 0) dup2(devnull, 1);
    struct PageRankStats *s = pageRankGraphCSR(&a, &g);
    fflush(stdout);
    if (saved >= 0) { dup2(saved, 1); close(saved); }
    if (devnull >= 0) close(devnull);
    if (!s) return -1;
    memcpy(ranks_out, s->pageRanks, (size_t)N * sizeof(float));
    *iters_out = s->iterations;
    *error_out = s->error_total;
    *seconds_out = s->time_total;
    freePageRankStats(s);
    return 0;
}

/* The same run with the reference's stdout (its per-iteration table) sent
 * to stdout_path, and the returned PageRankStats fields (pageRank.h:19-30)
 * in fields_out: damp, base_pr, iterations, num_vertices, error_total,
 * time_total, and 1 if realRanks is the identity (else 0). */
int ogref_pagerank_stdout(uint32_t N, uint32_t E, uint32_t *f_out_degree, uint32_t *f_in_degree,
                          uint32_t *f_edges_idx, uint32_t *f_src, uint32_t *f_dst,
                          uint32_t *i_out_degree, uint32_t *i_in_degree, uint32_t *i_edges_idx,
                          uint32_t *i_src, uint32_t *i_dst, uint32_t pushpull, double epsilon,
                          uint32_t iterations, int threads, const char *stdout_path, float *ranks_out,
                          double *fields_out)
{
    struct Vertex fv = {f_out_degree, f_in_degree, f_edges_idx};
    struct Vertex iv = {i_out_degree, i_in_degree, i_edges_idx};
    struct EdgeList fe, ie;
    memset(&fe, 0, sizeof(fe));
    memset(&ie, 0, sizeof(ie));
    fe.num_edges = ie.num_edges = E;
    fe.num_vertices = ie.num_vertices = N;
    fe.avg_degree = ie.avg_degree = N ? E / N : 0;
    fe.edges_array_src = f_src; fe.edges_array_dest = f_dst;
    ie.edges_array_src = i_src; ie.edges_array_dest = i_dst;
    struct GraphCSR g;
    memset(&g, 0, sizeof(g));
    g.num_edges = E;
    g.num_vertices = N;
    g.avg_degree = N ? E / N : 0;
    g.vertices = &fv;
    g.sorted_edges_array = &fe;
    g.inverse_vertices = &iv;
    g.inverse_sorted_edges_array = &ie;

    struct Arguments a;
    memset(&a, 0, sizeof(a));
    a.iterations = iterations;
    a.epsilon = epsilon;
    a.pushpull = pushpull;
    a.algorithm = 1;
    a.pre_numThreads = a.algo_numThreads = a.ker_numThreads = threads;
    initializeMersenneState(&a.mt19937var, 27491095);
    omp_set_num_threads(threads);

    fflush(stdout);
    int saved = dup(1);
    int out = open(stdout_path, O_WRONLY | O_CREAT | O_TRUNC, 0644);
    if (saved < 0 || out < 0) return -1;
    dup2(out, 1);
    struct PageRankStats *s = pageRankGraphCSR(&a, &g);
    fflush(stdout);
    dup2(saved, 1);
    close(saved);
    close(out);
    if (!s) return -1;
    memcpy(ranks_out, s->pageRanks, (size_t)N * sizeof(float));
    int identity = 1;
    for (uint32_t v = 0; v < N; v++) identity &= s->realRanks[v] == v;
    fields_out[0] = s->damp;
    fields_out[1] = s->base_pr;
    fields_out[2] = s->iterations;
    fields_out[3] = s->num_vertices;
    fields_out[4] = s->error_total;
    fields_out[5] = s->time_total;
    fields_out[6] = identity;
    freePageRankStats(s);
    return 0;
}

/* The reference's own SPMVGraphCSR (SPMV.c:588-1068) with -p pushpull for
 * `iterations` iterations: out gets vector_output, seconds_out time_total. */
int ogref_spmv(uint32_t N, uint32_t E, uint32_t *f_out_degree, uint32_t *f_in_degree,
               uint32_t *f_edges_idx, uint32_t *f_src, uint32_t *f_dst, uint32_t *i_out_degree,
               uint32_t *i_in_degree, uint32_t *i_edges_idx, uint32_t *i_src, uint32_t *i_dst,
               uint32_t pushpull, uint32_t iterations, int threads, float *out, double *seconds_out)
{
    struct Vertex fv = {f_out_degree, f_in_degree, f_edges_idx};
    struct Vertex iv = {i_out_degree, i_in_degree, i_edges_idx};
    struct EdgeList fe, ie;
    memset(&fe, 0, sizeof(fe));
    memset(&ie, 0, sizeof(ie));
    fe.num_edges = ie.num_edges = E;
    fe.num_vertices = ie.num_vertices = N;
    fe.edges_array_src = f_src; fe.edges_array_dest = f_dst;
    ie.edges_array_src = i_src; ie.edges_array_dest = i_dst;
    struct GraphCSR g;
    memset(&g, 0, sizeof(g));
    g.num_edges = E;
    g.num_vertices = N;
    g.vertices = &fv;
    g.sorted_edges_array = &fe;
    g.inverse_vertices = &iv;
    g.inverse_sorted_edges_array = &ie;
    struct Arguments a;
    memset(&a, 0, sizeof(a));
    a.iterations = iterations;
    a.pushpull = pushpull;
    a.pre_numThreads = a.algo_numThreads = a.ker_numThreads = threads;
    omp_set_num_threads(threads);
    fflush(stdout);
    int saved = dup(1);
    int devnull = open("/dev/null", O_WRONLY);
    if (saved >= 0 && devnull >= 0) dup2(devnull, 1);
    struct SPMVStats *s = SPMVGraphCSR(&a, &g);
    fflush(stdout);
    if (saved >= 0) { dup2(saved, 1); close(saved); }
    if (devnull >= 0) close(devnull);
    if (!s) return -1;
    memcpy(out, s->vector_output, (size_t)N * sizeof(float));
    *seconds_out = s->time_total;
    freeSPMVStats(s);
    return 0;
}

/* The reference's own writeToBinFileGraphCSR (graphCSR.c:380-449) on the
 * given CSR arrays: writes <fname>.csr. */
int ogref_write_csr(uint32_t N, uint32_t E, uint32_t *f_out_degree, uint32_t *f_in_degree,
                    uint32_t *f_edges_idx, uint32_t *f_src, uint32_t *f_dst, uint32_t *i_out_degree,
                    uint32_t *i_in_degree, uint32_t *i_edges_idx, uint32_t *i_src, uint32_t *i_dst,
                    const char *fname)
{
    struct Vertex fv = {f_out_degree, f_in_degree, f_edges_idx};
    struct Vertex iv = {i_out_degree, i_in_degree, i_edges_idx};
    struct EdgeList fe, ie;
    memset(&fe, 0, sizeof(fe));
    memset(&ie, 0, sizeof(ie));
    fe.num_edges = ie.num_edges = E;
    fe.num_vertices = ie.num_vertices = N;
    fe.edges_array_src = f_src; fe.edges_array_dest = f_dst;
    ie.edges_array_src = i_src; ie.edges_array_dest = i_dst;
    struct GraphCSR g;
    memset(&g, 0, sizeof(g));
    g.num_edges = E;
    g.num_vertices = N;
    g.vertices = &fv;
    g.sorted_edges_array = &fe;
    g.inverse_vertices = &iv;
    g.inverse_sorted_edges_array = &ie;
    writeToBinFileGraphCSR(fname, &g);
    return 0;
}

/* The reference's own reorderGraphProcess (reorder.c:1556) with -l lmode on
 * a copy of the edge list (src, dst) with N vertices: labels_out gets the
 * edge list's label_array, seconds_out the wall time of the call. */
int ogref_reorder(uint32_t N, uint32_t E, const uint32_t *src, const uint32_t *dst, uint32_t lmode,
                  int threads, uint32_t *labels_out, double *seconds_out)
{
    struct EdgeList *el = newEdgeList(E);
    if (!el) return -1;
    memcpy(el->edges_array_src, src, (size_t)E * 4);
    memcpy(el->edges_array_dest, dst, (size_t)E * 4);
    el->num_vertices = N;
    el->avg_degree = N ? E / N : 0;
    el->mask_array = (uint32_t *)my_malloc((size_t)N * 4);
    el->label_array = (uint32_t *)my_malloc((size_t)N * 4);
    el->inverse_label_array = (uint32_t *)my_malloc((size_t)N * 4);
    for (uint32_t v = 0; v < N; v++)
    {
        el->mask_array[v] = 0;
        el->label_array[v] = v;
        el->inverse_label_array[v] = v;
    }
    struct Arguments a;
    memset(&a, 0, sizeof(a));
    a.lmode = lmode;
    a.pre_numThreads = a.algo_numThreads = a.ker_numThreads = threads;
    omp_set_num_threads(threads);
    fflush(stdout);
    int saved = dup(1);
    int devnull = open("/dev/null", O_WRONLY);
    if (saved >= 0 && devnull >= 0) dup2(devnull, 1);
    double t0 = omp_get_wtime();
    el = reorderGraphProcess(el, &a);
    double t1 = omp_get_wtime();
    fflush(stdout);
    if (saved >= 0) { dup2(saved, 1); close(saved); }
    if (devnull >= 0) close(devnull);
    memcpy(labels_out, el->label_array, (size_t)N * 4);
    *seconds_out = t1 - t0;
    freeEdgeList(el);
    return 0;
}

/* The reference's own collectStatsPageRank_topK (graphStats.c:580-688) on
 * rankings given as arrays, for P problems (topk[p], num_vertices[p]); the
 * inverse rankings are formed once, as collectStatsPageRank does
 * (graphStats.c:746-751).  Per problem: out_u32 gets levenshtein_distance,
 * intersection, mismatch; out_f32 the float and real Kendall; out_f64
 * avg_error_float, avg_error_relative.  seconds_out: the _topK calls. */
int ogref_rank_stats_batch(uint32_t N, float *ref_pr, uint32_t *ref_order, float *pr, uint32_t *order,
                           uint32_t P, const uint32_t *topk, const uint32_t *num_vertices, uint32_t *out_u32,
                           float *out_f32, double *out_f64, double *seconds_out)
{
    struct PageRankStats rs, s;
    memset(&rs, 0, sizeof(rs));
    memset(&s, 0, sizeof(s));
    rs.num_vertices = s.num_vertices = N;
    rs.pageRanks = ref_pr; rs.realRanks = ref_order;
    s.pageRanks = pr; s.realRanks = order;
    uint32_t *inv = (uint32_t *)malloc((size_t)N * 4), *ref_inv = (uint32_t *)malloc((size_t)N * 4);
    if (!inv || !ref_inv) { free(inv); free(ref_inv); return -1; }
    for (uint32_t v = 0; v < N; v++)
    {
        inv[order[v]] = v;
        ref_inv[ref_order[v]] = v;
    }
    double t0 = omp_get_wtime();
    for (uint32_t p = 0; p < P; p++)
    {
        struct PageRankCorrelationStats c =
            collectStatsPageRank_topK(&rs, &s, ref_order, ref_inv, inv, topk[p], num_vertices[p], NULL, 0);
        out_u32[3 * p] = c.levenshtein_distance;
        out_u32[3 * p + 1] = c.intersection;
        out_u32[3 * p + 2] = c.mismatch;
        out_f32[2 * p] = c.float_Kendall;
        out_f32[2 * p + 1] = c.real_Kendall;
        out_f64[2 * p] = c.avg_error_float;
        out_f64[2 * p + 1] = c.avg_error_relative;
    }
    *seconds_out = omp_get_wtime() - t0;
    free(inv);
    free(ref_inv);
    return 0;
}

/* The reference's own collectStatsPageRank (graphStats.c:691-1020): appends
 * to <fnameb>_1_0_<trial>_<pushpull>.stats; stdout goes to stdout_path
 * (NULL: discarded).  seconds_out: wall time of the call. */
int ogref_collect_stats(uint32_t N, float *ref_pr, uint32_t *ref_order, float *pr, uint32_t *order,
                        char *fnameb, uint32_t bin_size, uint32_t verbosity, uint32_t pushpull,
                        uint32_t trial, double time_total, uint32_t iterations, int threads,
                        const char *stdout_path, double *seconds_out)
{
    struct PageRankStats rs, s;
    memset(&rs, 0, sizeof(rs));
    memset(&s, 0, sizeof(s));
    rs.num_vertices = s.num_vertices = N;
    rs.pageRanks = ref_pr; rs.realRanks = ref_order;
    s.pageRanks = pr; s.realRanks = order;
    s.time_total = time_total;
    s.iterations = iterations;
    struct Arguments a;
    memset(&a, 0, sizeof(a));
    a.fnameb = fnameb;
    a.binSize = bin_size;
    a.verbosity = verbosity;
    a.pushpull = pushpull;
    a.algorithm = 1;
    a.datastructure = 0;
    a.pre_numThreads = a.algo_numThreads = a.ker_numThreads = threads;
    fflush(stdout);
    int saved = dup(1);
    int sink = stdout_path ? open(stdout_path, O_WRONLY | O_CREAT | O_TRUNC, 0644) : open("/dev/null", O_WRONLY);
    if (saved >= 0 && sink >= 0) dup2(sink, 1);
    double t0 = omp_get_wtime();
    collectStatsPageRank(&a, &rs, &s, trial);
    double t1 = omp_get_wtime();
    fflush(stdout);
    if (saved >= 0) { dup2(saved, 1); close(saved); }
    if (sink >= 0) close(sink);
    *seconds_out = t1 - t0;
    return 0;
}

/* The reference's CSR preprocessing (graphCSR.c:164-248, without the file
 * read, duplicate removal, relabelling and masks) on caller edge arrays:
 * sortRunAlgorithms (-o sort), graphCSRNew, graphCSRAssignEdgeList (which
 * runs mapVerticesWithInOutDegree), readEdgeListsMem (the inverse list),
 * sortRunAlgorithms and graphCSRAssignEdgeList on it.  seconds_out: wall
 * time of those steps (omp_get_wtime, the edge-list copy-in excluded) on
 * `threads` OpenMP threads, stdout silenced.  out[10] (caller-owned, E or N
 * entries) get f_src, f_dst, f_out_degree, f_in_degree, f_edges_idx, i_src,
 * i_dst, i_out_degree, i_in_degree, i_edges_idx. */
int ogref_csr_build(uint32_t N, uint32_t E, const uint32_t *src, const uint32_t *dst, uint32_t sort,
                    int threads, uint32_t *const out[10], double *seconds_out)
{
    struct EdgeList *el = newEdgeList(E);
    if (!el) return -1;
    memcpy(el->edges_array_src, src, (size_t)E * 4);
    memcpy(el->edges_array_dest, dst, (size_t)E * 4);
    el->num_vertices = N;
    el->avg_degree = N ? E / N : 0;
    el->mask_array = (uint32_t *)my_malloc((size_t)N * 4);
    el->label_array = (uint32_t *)my_malloc((size_t)N * 4);
    el->inverse_label_array = (uint32_t *)my_malloc((size_t)N * 4);
    for (uint32_t v = 0; v < N; v++)
    {
        el->mask_array[v] = 0;
        el->label_array[v] = v;
        el->inverse_label_array[v] = v;
    }
    omp_set_num_threads(threads);
    fflush(stdout);
    int saved = dup(1);
    int devnull = open("/dev/null", O_WRONLY);
    if (saved >= 0 && devnull >= 0) dup2(devnull, 1);
    double t0 = omp_get_wtime();
    el = sortRunAlgorithms(el, sort);
    struct GraphCSR *g = graphCSRNew(el->num_vertices, el->num_edges, 1);
    g = graphCSRAssignEdgeList(g, el, 0);
    struct EdgeList *iel = readEdgeListsMem(el, 1, 0, 0);
    iel = sortRunAlgorithms(iel, sort);
    g = graphCSRAssignEdgeList(g, iel, 1);
    double t1 = omp_get_wtime();
    fflush(stdout);
    if (saved >= 0) { dup2(saved, 1); close(saved); }
    if (devnull >= 0) close(devnull);
    const uint32_t *arr[10] = {
        g->sorted_edges_array->edges_array_src, g->sorted_edges_array->edges_array_dest,
        g->vertices->out_degree, g->vertices->in_degree, g->vertices->edges_idx,
        g->inverse_sorted_edges_array->edges_array_src, g->inverse_sorted_edges_array->edges_array_dest,
        g->inverse_vertices->out_degree, g->inverse_vertices->in_degree, g->inverse_vertices->edges_idx};
    for (int k = 0; k < 10; k++)
        memcpy(out[k], arr[k], (size_t)(k == 0 || k == 1 || k == 5 || k == 6 ? E : N) * 4);
    *seconds_out = t1 - t0;
    graphCSRFree(g);
    return 0;
}
