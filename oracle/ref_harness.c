/*
 * ref_harness.c -- TEST INFRASTRUCTURE ONLY.
 *
 * A small driver, compiled against the reference's own headers and linked
 * with the reference's own PageRank-path objects (oracle/Makefile target
 * "ref"), that writes the reference's outputs for one configuration so the
 * oracle restatement can be pinned against them.  It replaces the
 * reference's CLI main (open-graph.c:267-327) for this purpose: the same
 * default Arguments, graphCSRPreProcessingStep (graphCSR.c:164) and
 * pageRankGraphCSR (pageRank.c:871) calls, then a binary dump instead of
 * the reference's (commented-out) rank printing.
 *
 * usage: ref_harness <graph.bin> <sort> <pushpull> <epsilon> <iterations>
 *                    <threads> <out.bin> [dump_csr 0|1] [lmode]
 * out.bin: u32 magic 'OGRF', u32 N, u32 E, u32 iterations, f64 error_total,
 *          f64 time_total, f32 ranks[N], u32 sorted_labels[N]
 *          [then if dump_csr: fwd src,dst[E]; fwd out,in,idx[N];
 *                             inv src,dst[E]; inv out,in,idx[N]
 *           and if lmode was given: fwd label_array[N]]
 * lmode is the reference's -l reordering (reorder.c:1556-1616) applied by
 * graphCSRPreProcessingStep before the CSR is built.
 * The per-iteration table the reference prints stays on stdout.
 */
#include <omp.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "graphConfig.h"
#include "graphCSR.h"
#include "mt19937.h"
#include "myMalloc.h"
#include "pageRank.h"
#include "reorder.h"

static void put(FILE *f, const void *p, size_t n)
{
    if (n && fwrite(p, 1, n, f) != n) { perror("fwrite"); exit(1); }
}

int main(int argc, char **argv)
{
    if (argc < 8)
    {
        fprintf(stderr, "usage: %s graph.bin sort pushpull eps iters threads out.bin [dump_csr]\n", argv[0]);
        return 2;
    }
    struct Arguments *a = (struct Arguments *)my_malloc(sizeof(struct Arguments));
    memset(a, 0, sizeof(*a));
    /* open-graph.c:270-305 defaults. */
    a->binSize = 512;
    a->iterations = 20;
    a->trials = 1;
    a->epsilon = 0.0001;
    a->algorithm = 1;
    a->cache_size = 32768;
    a->delta = 1;
    a->fnameb_format = 1;
    a->convert_format = 1;
    initializeMersenneState(&(a->mt19937var), 27491095);

    a->fnameb = argv[1];
    a->sort = (uint32_t)atoi(argv[2]);
    a->pushpull = (uint32_t)atoi(argv[3]);
    a->epsilon = atof(argv[4]);
    a->iterations = (uint32_t)atoi(argv[5]);
    int threads = atoi(argv[6]);
    int dump_csr = argc > 8 ? atoi(argv[8]) : 0;
    int with_lmode = argc > 9;
    if (with_lmode) a->lmode = (uint32_t)atoi(argv[9]);
    a->pre_numThreads = a->algo_numThreads = a->ker_numThreads = threads;
    omp_set_nested(1);
    omp_set_num_threads(threads);

    struct GraphCSR *g = graphCSRPreProcessingStep(a);
    struct PageRankStats *s = pageRankGraphCSR(a, g);
    uint32_t N = g->num_vertices, E = g->num_edges;

    uint32_t *labels = (uint32_t *)my_malloc((size_t)N * sizeof(uint32_t));
    for (uint32_t v = 0; v < N; v++) labels[v] = v;
    labels = radixSortEdgesByPageRank(s->pageRanks, labels, N);

    FILE *f = fopen(argv[7], "wb");
    if (!f) { perror("fopen"); return 1; }
    uint32_t hdr[4] = {0x4652474Fu, N, E, s->iterations};
    put(f, hdr, sizeof(hdr));
    put(f, &s->error_total, 8);
    put(f, &s->time_total, 8);
    put(f, s->pageRanks, (size_t)N * 4);
    put(f, labels, (size_t)N * 4);
    if (dump_csr)
    {
        struct EdgeList *fe = g->sorted_edges_array, *ie = g->inverse_sorted_edges_array;
        put(f, fe->edges_array_src, (size_t)E * 4);
        put(f, fe->edges_array_dest, (size_t)E * 4);
        put(f, g->vertices->out_degree, (size_t)N * 4);
        put(f, g->vertices->in_degree, (size_t)N * 4);
        put(f, g->vertices->edges_idx, (size_t)N * 4);
        put(f, ie->edges_array_src, (size_t)E * 4);
        put(f, ie->edges_array_dest, (size_t)E * 4);
        put(f, g->inverse_vertices->out_degree, (size_t)N * 4);
        put(f, g->inverse_vertices->in_degree, (size_t)N * 4);
        put(f, g->inverse_vertices->edges_idx, (size_t)N * 4);
        if (with_lmode) put(f, fe->label_array, (size_t)N * 4);
    }
    fclose(f);
    free(labels);
    freePageRankStats(s);
    graphCSRFree(g);
    free(a);
    return 0;
}
