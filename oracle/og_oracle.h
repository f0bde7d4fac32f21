/*
 * og_oracle.h -- CPU oracle for the OpenGraph PageRank-over-CSR path.
 *
 * TEST INFRASTRUCTURE ONLY.  This is a plain C/OpenMP restatement of the
 * reference algorithms, used as the checker of the GPU library by tests/,
 * by __graft_entry__.smoke() and by bench.py's cpu_baseline leg.  The
 * product (libopengraph_amd.so) never links, loads or calls it.
 *
 * Each function cites the reference lines it restates (paths relative to the
 * reference's 00_graph_bench/ tree).  Arithmetic follows the reference's x86-64
 * gcc -O3 build exactly: fp32 division, no FMA contraction (this file set is
 * compiled with -ffp-contract=off), and the x86 results of out-of-range
 * float->integer conversions made explicit by the og_x86_* helpers.
 *
 * Pinning (what checks each part, see tests/test_oracle_*.py):
 *   CSR build   -- golden CSR digests from the compiled reference, numpy
 *                  stable argsort / lexsort / bincount closed forms, 0==2, 1==3.
 *   pull / fixed-point 64/32/16/8 / quant 32/16/8 / data-driven pull
 *               -- bit-exact against golden rank vectors and per-iteration
 *                  active counts written by the compiled reference
 *                  (tests/golden); closed forms (cycle, star, dense solve).
 *   push (fp32 atomics), quant push, data-driven push / pull-push
 *               -- nondeterministic in the reference as well; pinned against
 *                  the FP64 oracle and the golden vectors within tolerances.
 *   FP64 oracle -- dense linear solve on small graphs, mass balance,
 *                  isolated-vertex value.
 */
#ifndef OG_ORACLE_H
#define OG_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Forward and inverse CSR exactly as the reference lays them out:
 * sorted src/dst lists plus the three Vertex arrays of each direction. */
typedef struct
{
    uint32_t num_vertices;
    uint32_t num_edges;
    uint32_t *f_src, *f_dst;              /* graph->sorted_edges_array */
    uint32_t *f_out_degree, *f_in_degree, *f_edges_idx;   /* graph->vertices */
    uint32_t *i_src, *i_dst;              /* graph->inverse_sorted_edges_array */
    uint32_t *i_out_degree, *i_in_degree, *i_edges_idx;   /* graph->inverse_vertices */
} orc_csr;

/* Per-run outputs of a PageRank restatement. */
typedef struct
{
    uint32_t iterations;        /* stats->iterations (off-by-one rule kept) */
    uint32_t rows;              /* number of per-iteration rows written */
    double error_total;         /* last iteration's error_total */
    double pr_sum;              /* "PR Sum" of the normalized ranks */
    double seconds_total;       /* wall time of the timed region */
} orc_pr_result;

/* edgeList.c:559-567: number of vertices = max id + 1. */
uint32_t orc_num_vertices(const uint32_t *src, const uint32_t *dst, uint32_t num_edges);

/* radixsort.c:46-135 / :231-335 and countsort.c: stable sort of (key, val)
 * pairs by key; returns 0. */
int orc_stable_sort_pairs(uint32_t *key, uint32_t *val, uint32_t n, int nthreads);

/* graphCSR.c:164-248 with sortRun.c:26-69, edgeList.c:597-745 and
 * vertex.c:204-325.  sort_mode: 0/2 by src, 1/3 by src then dst.
 * Allocates every array of *out.  Returns 0, or -1 on allocation failure. */
int orc_build_csr(const uint32_t *src, const uint32_t *dst, uint32_t num_edges,
                  uint32_t sort_mode, int nthreads, orc_csr *out);
/* The same with at least min_vertices vertices: a relabelled edge list keeps
 * the vertex count of the list it came from (reorder.c:1618-1637 changes
 * only the ids), even when its largest label has no edge. */
int orc_build_csr_n(const uint32_t *src, const uint32_t *dst, uint32_t num_edges,
                    uint32_t min_vertices, uint32_t sort_mode, int nthreads, orc_csr *out);
void orc_free_csr(orc_csr *g);

/* pageRank.c:871-2896: the -p code variants 0..13 on CSR.
 *   ranks_out        N normalized ranks (pageRanks after r/N)
 *   active_per_iter  optional, max_iter entries: activeVertices per iteration
 *   error_per_iter   optional, max_iter entries: error_total per iteration
 *   init             optional unnormalized initial ranks (NULL: 0.15f,
 *                    pageRank.c:65-70); used by the uniform / learned starts
 *   stop_l1          0: reference rule (stop when activeVertices == 0);
 *                    1: stop when error_total < epsilon (codes 0, 2 only)
 * Returns 0, -1 on allocation failure, -2 on an unknown code. */
int orc_pagerank(const orc_csr *g, uint32_t code, double epsilon, uint32_t max_iter,
                 int nthreads, const float *init, int stop_l1,
                 float *ranks_out, uint32_t *active_per_iter, double *error_per_iter,
                 orc_pr_result *res);

/* FP64 oracle of the same map r' = 0.15 + 0.85 * sum(r_u / outdeg_u) with no
 * dangling redistribution, in double throughout with sequential sums.
 * init: unnormalized start value for every vertex.  Iterates until the
 * normalized L1 change sum|dr|/N < tol or max_iter.  ranks_out is normalized
 * (divided by N).  Returns the number of iterations run. */
uint32_t orc_pagerank_fp64(const orc_csr *g, double init, double tol, uint32_t max_iter,
                           int nthreads, double *ranks_out, double *last_l1);

/* FP64 oracle of personalized PageRank with a one-hot restart at source s:
 * x' = 0.15 * e_s + 0.85 * P^T x (no dangling redistribution), start x = e_s,
 * until sum|dx| < tol.  Returns iterations. */
uint32_t orc_ppr_fp64(const orc_csr *g, uint32_t source, double tol, uint32_t max_iter,
                      double *x_out, double *last_l1);

/* The same iteration in the pull kernel's fp32 arithmetic, the sum of each
 * row in the reference's in-neighbour order (pageRank.c:998-1011 with the
 * restart term (1.0f - 0.85f) only at s): c = x / (float)outdeg,
 * x' = ((v == s) ? 1.0f - 0.85f : 0.0f) + 0.85f * sum, L1 = sum over v of
 * fabs((double)(x' - x)); stops after the first iteration with L1 < tol.
 * l1_hist[i] (may be NULL) gets iteration i's L1.  Returns iterations. */
uint32_t orc_ppr_f32(const orc_csr *g, uint32_t source, double tol, uint32_t max_iter,
                     float *x_out, double *l1_hist);

/* Learned warm start (og_oracle_mlp.c, no counterpart in the reference):
 * features [log1p in-degree, log1p out-degree, mean log1p in-degree of the
 * in-neighbours, 1/N] per vertex (feat: N x 4); the 4 -> 64 -> 64 -> 1
 * ReLU MLP with softplus output in double precision; its mean L1 loss
 * against y and the gradient of that loss (4545 entries). */
void orc_mlp_features(const orc_csr *g, float *feat);
void orc_mlp_forward(const float *params, const float *feat, uint32_t n, double *out);
double orc_mlp_loss(const float *params, const float *feat, const float *y, uint32_t n);
double orc_mlp_grad(const float *params, const float *feat, const float *y, uint32_t n, double *grad);

/* reorder.c:64-221: stable ascending sort of labels by FloatToFixed32SORT(rank). */
void orc_sort_labels_by_rank(const float *ranks, uint32_t *labels, uint32_t n);

/* reorder.c:275-572 and :1483-1554: the new label of every vertex for the
 * reference's -l option lmode (1 out-degree, 2 in-degree, 3 in+out degree:
 * degree sort; 4 out-degree, 5 in-degree: DBG buckets) of the edge list
 * (src, dst) with N vertices.  Returns 0, or -1 for another lmode. */
int orc_reorder_labels(const uint32_t *src, const uint32_t *dst, uint32_t num_edges, uint32_t N,
                       uint32_t lmode, uint32_t *labels);

/* SPMV.c:616-1068: the reference's benchmark SpMV on CSR for `iterations`
 * iterations, -p code 0 pull, 1 push, 2 fixed-point pull, 3 fixed-point
 * push; out = vector_output.  Returns 0, or -1 for another code. */
int orc_spmv(const orc_csr *g, uint32_t code, uint32_t iterations, float *out);

/* graphStats.h:28-37 */
typedef struct
{
    uint32_t levenshtein_distance;
    float float_Kendall;
    float real_Kendall;
    uint32_t intersection;
    uint32_t mismatch;
    double avg_error_float;
    double avg_error_relative;
} orc_corr_stats;

/* graphStats.c:580-688 with the inverse rankings of :746-751: the top-k
 * comparison of a ranking (pr, order = realRanks ascending by rank) against
 * the reference ranking (ref_pr, ref_order) over positions
 * [num_vertices - topk, num_vertices).  Returns 0, -1 on a bad size. */
int orc_rank_stats_topk(const float *ref_pr, const uint32_t *ref_order, const float *pr, const uint32_t *order,
                        uint32_t N, uint32_t topk, uint32_t num_vertices, orc_corr_stats *out);

#ifdef __cplusplus
}
#endif

#endif
