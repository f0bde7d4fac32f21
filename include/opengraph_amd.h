/*
 * opengraph_amd.h -- drop-in C ABI of the OpenGraph PageRank-over-CSR path,
 * implemented on MI355X (gfx950) by libopengraph_amd.so.
 *
 * Every type below has the exact field order and x86-64 layout of the
 * reference type it replaces (compiled with WEIGHTED=0, DIRECTED=1), so a
 * caller built against the reference headers can link this library instead
 * of the reference objects for this path.  Every function keeps the name,
 * signature, argument meaning and ownership rules of the reference function
 * it replaces; the comment on each one cites it (paths relative to the
 * reference's 00_graph_bench/ tree).
 *
 * Errors: the reference has no error returns (it exits through err()/exit()
 * on failure, e.g. edgeList.c:354-374).  The drop-in entry points keep that
 * convention: on a HIP failure or an unsupported option they print the cause
 * to stderr and exit(EXIT_FAILURE).  They never compute on the CPU instead.
 * The ogpu_* entry points in opengraph_amd_ext.h return status codes.
 */
#ifndef OPENGRAPH_AMD_H
#define OPENGRAPH_AMD_H

#include <stdint.h>
#include <stddef.h>
#include <stdio.h>

#ifdef __cplusplus
extern "C" {
#endif

/* graphConfig.h:7-8 -- the configuration this ABI is laid out for. */
#define WEIGHTED 0
#define DIRECTED 1

/* mt19937.h:59-63 -- embedded in struct Arguments; must keep its size. */
typedef struct mt19937_stateStruct
{
    unsigned long mt[624];
    int mti;
} mt19937state;

/* graphConfig.h:11-44 */
struct Arguments
{
    int wflag;
    int xflag;
    int sflag;
    int Sflag;
    int dflag;
    uint32_t binSize;
    uint32_t verbosity;
    uint32_t iterations;
    uint32_t trials;
    double epsilon;
    int source;
    uint32_t algorithm;
    uint32_t datastructure;
    uint32_t pushpull;
    uint32_t sort;
    uint32_t lmode;
    uint32_t lmode_l2;
    uint32_t lmode_l3;
    uint32_t mmode;
    uint32_t symmetric;
    uint32_t weighted;
    uint32_t delta;
    int pre_numThreads;
    int algo_numThreads;
    int ker_numThreads;
    char *fnameb;
    char *fnamel;
    uint32_t fnameb_format;
    uint32_t convert_format;
    mt19937state mt19937var;
    uint32_t cache_size;
};

/* vertex.h:13-18 */
struct Vertex
{
    uint32_t *out_degree;
    uint32_t *in_degree;
    uint32_t *edges_idx;
};

/* edgeList.h:8-23 */
struct EdgeList
{
    uint32_t num_edges;
    uint32_t num_vertices;
    uint32_t avg_degree;
    uint32_t *edges_array_src;
    uint32_t *edges_array_dest;
    uint32_t *mask_array;
    uint32_t *label_array;
    uint32_t *inverse_label_array;
};

/* graphCSR.h:12-29 */
struct GraphCSR
{
    uint32_t num_edges;
    uint32_t num_vertices;
    uint32_t avg_degree;
    struct Vertex *vertices;
    struct EdgeList *sorted_edges_array;
    struct Vertex *inverse_vertices;
    struct EdgeList *inverse_sorted_edges_array;
};

/* pageRank.h:19-30 */
struct PageRankStats
{
    float damp;
    float base_pr;
    uint32_t iterations;
    uint32_t num_vertices;
    uint32_t activeVertices;
    uint32_t *realRanks;
    float *pageRanks;
    double time_total;
    double error_total;
};

/* ------------------------------------------------------------------------
 * PageRank on CSR.  All variants run on the GPU.  The graph is borrowed and
 * never modified; its device copy is cached by pointer and array identity
 * and rebuilt when any of them changes.  The returned stats and both of its
 * arrays are malloc-family allocations owned by the caller (free() them, or
 * call freePageRankStats).
 * ---------------------------------------------------------------------- */

/* pageRank.c:871-936 (decl pageRank.h:69): dispatch on arguments->pushpull. */
struct PageRankStats *pageRankGraphCSR(struct Arguments *arguments, struct GraphCSR *graph);
/* pageRank.c:939-1062: pull, fp32 ranks, fp64 error sum (-p 0). */
struct PageRankStats *pageRankPullGraphCSR(struct Arguments *arguments, struct GraphCSR *graph);
/* pageRank.c:1064-1186: push with fp32 atomics (-p 1). */
struct PageRankStats *pageRankPushGraphCSR(struct Arguments *arguments, struct GraphCSR *graph);
/* pageRank.c:1190-1324: pull, uint64 fixed point with 32 fraction bits (-p 2). */
struct PageRankStats *pageRankPullFixedPoint64BitGraphCSR(struct Arguments *arguments, struct GraphCSR *graph);
/* pageRank.c:1326-1459: pull, 16 fraction bits in uint32 contributions (-p 9). */
struct PageRankStats *pageRankPullFixedPoint32BitGraphCSR(struct Arguments *arguments, struct GraphCSR *graph);
/* pageRank.c:1461-1591: pull, 14 fraction bits in uint16 contributions (-p 10). */
struct PageRankStats *pageRankPullFixedPoint16BitGraphCSR(struct Arguments *arguments, struct GraphCSR *graph);
/* pageRank.c:1593-1724: pull, 6 fraction bits in uint8 contributions (-p 11). */
struct PageRankStats *pageRankPullFixedPoint8BitGraphCSR(struct Arguments *arguments, struct GraphCSR *graph);
/* pageRank.c:1726-1849: push, uint64 fixed point atomics (-p 3). */
struct PageRankStats *pageRankPushFixedPointGraphCSR(struct Arguments *arguments, struct GraphCSR *graph);
/* pageRank.c:1853-1988: pull, per-iteration 32-bit quantized contributions (-p 4). */
struct PageRankStats *pageRankPullQuant32BitGraphCSR(struct Arguments *arguments, struct GraphCSR *graph);
/* pageRank.c:1990-2128: pull, 16-bit quantized contributions (-p 12). */
struct PageRankStats *pageRankPullQuant16BitGraphCSR(struct Arguments *arguments, struct GraphCSR *graph);
/* pageRank.c:2130-2270: pull, 8-bit quantized contributions (-p 13). */
struct PageRankStats *pageRankPullQuant8BitGraphCSR(struct Arguments *arguments, struct GraphCSR *graph);
/* pageRank.c:2273-2401: push, 32-bit quantized contributions with zero point (-p 5). */
struct PageRankStats *pageRankPushQuantGraphCSR(struct Arguments *arguments, struct GraphCSR *graph);
/* pageRank.c:2405-2564: data-driven pull over byte worklists (-p 6). */
struct PageRankStats *pageRankDataDrivenPullGraphCSR(struct Arguments *arguments, struct GraphCSR *graph);
/* pageRank.c:2566-2720: data-driven residual push (-p 7). */
struct PageRankStats *pageRankDataDrivenPushGraphCSR(struct Arguments *arguments, struct GraphCSR *graph);
/* pageRank.c:2723-2896: data-driven pull-push hybrid (-p 8). */
struct PageRankStats *pageRankDataDrivenPullPushGraphCSR(struct Arguments *arguments, struct GraphCSR *graph);

/* pageRank.c:47-75: damp 0.85f, base_pr 0.15f, pageRanks[v]=0.15f, realRanks[v]=v. */
struct PageRankStats *newPageRankStatsGraphCSR(struct GraphCSR *graph);
/* pageRank.c:167-177 */
void freePageRankStats(struct PageRankStats *stats);

/* SPMV.h:19-28 */
struct SPMVStats
{
    uint32_t iterations;
    struct Arguments *arguments;
    uint32_t num_vertices;
    float *vector_output;
    float *vector_input;
    double time_total;
};

/* ------------------------------------------------------------------------
 * SpMV on CSR (the reference's SpMV benchmark), on the GPU: y = 0, then
 * arguments->iterations passes y += sum over the in-edges of 0.0001f * x,
 * x[v] = 1.0f / out-degree(v) (0.001f without out-edges), as the reference.
 * Same graph caching and ownership rules as the PageRank entry points.
 * ---------------------------------------------------------------------- */

/* SPMV.c:588-614 (decl SPMV.h:51): dispatch on arguments->pushpull. */
struct SPMVStats *SPMVGraphCSR(struct Arguments *arguments, struct GraphCSR *graph);
/* SPMV.c:616-713: pull, fp32 (-p 0). */
struct SPMVStats *SPMVPullGraphCSR(struct Arguments *arguments, struct GraphCSR *graph);
/* SPMV.c:715-810: push, fp32 atomics (-p 1). */
struct SPMVStats *SPMVPushGraphCSR(struct Arguments *arguments, struct GraphCSR *graph);
/* SPMV.c:812-948: pull, Q16.16 fixed point (-p 2). */
struct SPMVStats *SPMVPullFixedPointGraphCSR(struct Arguments *arguments, struct GraphCSR *graph);
/* SPMV.c:950-1068: push, Q32.32 fixed point (-p 3). */
struct SPMVStats *SPMVPushFixedPointGraphCSR(struct Arguments *arguments, struct GraphCSR *graph);
/* SPMV.c:45-66: vectors zeroed, iterations 0. */
struct SPMVStats *newSPMVStatsGraphCSR(struct GraphCSR *graph);
/* SPMV.c:139-152 */
void freeSPMVStats(struct SPMVStats *stats);

/* graphStats.h:28-37 */
struct PageRankCorrelationStats
{
    uint32_t levenshtein_distance;
    float float_Kendall;
    float real_Kendall;
    uint32_t intersection;
    uint32_t mismatch;
    double avg_error_float;
    double avg_error_relative;
};

/* ------------------------------------------------------------------------
 * Rank comparison (-S): a variant's ranking against the pull ranking over
 * the top-k positions of realRanks (ascending by rank, as
 * radixSortEdgesByPageRank leaves them).  The statistics are computed on
 * the GPU; the printed tables and the .stats file are the reference's.
 * ---------------------------------------------------------------------- */

/* graphStats.c:691-1020 (decl graphStats.h:46): top-k table for 30 .. 10000,
 * the 1000-wide chunks when arguments->verbosity > 1, and top binSize;
 * appends to <fnameb>_<algorithm>_<datastructure>_<trial>_<pushpull>.stats. */
void collectStatsPageRank(struct Arguments *arguments, struct PageRankStats *ref_stats, struct PageRankStats *stats,
                          uint32_t trial);
/* graphStats.c:580-688 (decl graphStats.h:47): positions
 * [num_vertices - topk, num_vertices).  The three rank arrays must be what
 * collectStatsPageRank passes (ref_stats->realRanks and the inverses of
 * both realRanks, graphStats.c:746-751): they are formed again on the GPU
 * from the two realRanks. */
struct PageRankCorrelationStats collectStatsPageRank_topK(struct PageRankStats *ref_stats, struct PageRankStats *stats,
                                                          uint32_t *ref_rankedVertices_total,
                                                          uint32_t *ref_rankedVertices_inverse,
                                                          uint32_t *rankedVertices_inverse, uint32_t topk,
                                                          uint32_t num_vertices, FILE *fptr, uint32_t verbose);

/* ------------------------------------------------------------------------
 * Edge list -> CSR builder.  Sorting and degree mapping run on the GPU; the
 * host structures produced are laid out and allocated (128-byte aligned
 * posix_memalign, myMalloc.c:24-67) exactly as the reference's, so the
 * reference's own free functions also release them.
 * ---------------------------------------------------------------------- */

/* graphCSR.c:164-248: read arguments->fnameb (.bin), sort with
 * arguments->sort, map forward and inverse CSR.  Reordering (-l/-L/-O),
 * masking (-M) and duplicate removal (-k) are not supported and exit. */
struct GraphCSR *graphCSRPreProcessingStep(struct Arguments *arguments);
/* graphCSR.c:107-134 */
struct GraphCSR *graphCSRNew(uint32_t V, uint32_t E, uint8_t inverse);
/* graphCSR.c:137-161: attach a sorted list and map its vertices. */
struct GraphCSR *graphCSRAssignEdgeList(struct GraphCSR *graphCSR, struct EdgeList *edgeList, uint8_t inverse);
/* graphCSR.c:42-64 */
void graphCSRFree(struct GraphCSR *graphCSR);
/* graphCSR.c:66-104 */
void graphCSRPrint(struct GraphCSR *graphCSR);
/* vertex.c:204-325: out_degree / edges_idx of the (inverse) list on the GPU;
 * edges_idx[v] is 0 for a vertex with no edges, as in the reference. */
struct GraphCSR *mapVerticesWithInOutDegree(struct GraphCSR *graph, uint8_t inverse);
/* vertex.c:26-49 */
struct Vertex *newVertexArray(uint32_t num_vertices);
/* vertex.c:406-420 */
void freeVertexArray(struct Vertex *vertices);

/* edgeList.c:336-594: mmap a .bin/.wbin file; num_edges = size / record;
 * num_vertices = max id + 1.  Weights (-w) are skipped as the reference
 * does with WEIGHTED=0. */
struct EdgeList *readEdgeListsbin(const char *fname, uint8_t inverse, uint32_t symmetric, uint32_t weighted);
/* edgeList.c:597-745: copy (inverse=1: swap src/dst) of an in-memory list. */
struct EdgeList *readEdgeListsMem(struct EdgeList *edgeListmem, uint8_t inverse, uint32_t symmetric, uint32_t weighted);
/* edgeList.c:98-135 */
struct EdgeList *newEdgeList(uint32_t num_edges);
/* edgeList.c:213-239 */
void freeEdgeList(struct EdgeList *edgeList);

/* sortRun.c:26-69: 0 radix by src, 1 radix by src then dst, 2/3 the same
 * orders by counting sort.  All four are stable, so 0 == 2 and 1 == 3. */
struct EdgeList *sortRunAlgorithms(struct EdgeList *edgeList, uint32_t sort);
/* radixsort.c:231-278 */
struct EdgeList *radixSortEdgesBySource(struct EdgeList *edgeList);
/* radixsort.c:282-335 */
struct EdgeList *radixSortEdgesBySourceAndDestination(struct EdgeList *edgeList);
/* countsort.c:28-141 */
struct EdgeList *countSortEdgesBySource(struct EdgeList *edgeList);
/* countsort.c:262-290 */
struct EdgeList *countSortEdgesBySourceAndDestination(struct EdgeList *edgeList);

/* graphCSR.c:380-449: writes <fname>.csr (E, N, per vertex the forward and
 * inverse out-degree, in-degree, edges_idx, per edge the forward and
 * inverse src, dst; u32, host byte order); exits when it cannot. */
void writeToBinFileGraphCSR(const char *fname, struct GraphCSR *graphCSR);
/* graphCSR.c:452-558: a GraphCSR from such a file (label arrays NULL, as
 * the reference leaves them); exits when the file cannot be read. */
struct GraphCSR *readFromBinFileGraphCSR(const char *fname);

/* reorder.c:165-221: caller-side post-sort of vertex labels by
 * FloatToFixed32SORT(rank) (graphRun.c:1076).  Returns labels, reordered
 * in place, ascending by fixed-point rank, stable. */
uint32_t *radixSortEdgesByPageRank(float *pageRanks, uint32_t *labels, uint32_t num_vertices);

/* ---- degree relabelling before the CSR build (-l 1..5), on the GPU ----
 * graphCSRPreProcessingStep applies -l, -L, -O (graphCSR.c:184-203) when
 * each is 0 or one of these; the other reorderings (HUBSort, HUBCluster,
 * Corder, labels from a file, the thread-dependent random -l 10) exit. */
/* reorder.c:1483-1554: adds the out-degree (lmode 1, 4, 6, 8, default), the
 * in-degree (2, 5, 7, 9) or both (3) of every vertex to degrees. */
uint32_t *reorderGraphGenerateInOutDegrees(uint32_t *degrees, struct EdgeList *edgeList, uint32_t lmode);
/* reorder.c:223-271: stable ascending sort of the (degree, label) pairs;
 * both arrays come back sorted.  Returns labels. */
uint32_t *radixSortEdgesByDegree(uint32_t *degrees, uint32_t *labels, uint32_t num_vertices);
/* reorder.c:1618-1637: src and dst of every edge through labels. */
struct EdgeList *relabelEdgeList(struct EdgeList *edgeList, uint32_t *labels);
/* reorder.c:296-384: labels by decreasing degree (ties by decreasing id),
 * relabel, update label_array / inverse_label_array. */
struct EdgeList *reorderGraphListDegree(struct EdgeList *edgeList, uint32_t *degrees, uint32_t lmode);
/* reorder.c:275-294 */
struct EdgeList *reorderGraphProcessDegree(uint32_t sort, struct EdgeList *edgeList, uint32_t lmode);
/* reorder.c:439-572: DBG -- bucket b = first threshold >= degree; labels
 * over the buckets from the last to the first, in id order within one. */
struct EdgeList *reorderGraphListDBG(struct EdgeList *edgeList, uint32_t *degrees, uint32_t *thresholds,
                                     uint32_t num_buckets, uint32_t lmode);
/* reorder.c:387-437: 11 thresholds from the edge list's average degree. */
struct EdgeList *reorderGraphProcessDBG(uint32_t sort, struct EdgeList *edgeList, uint32_t lmode);
/* reorder.c:1556-1616 for lmode 0..5. */
struct EdgeList *reorderGraphProcess(struct EdgeList *edgeList, struct Arguments *arguments);

#ifdef __cplusplus
}
#endif

#endif /* OPENGRAPH_AMD_H */
