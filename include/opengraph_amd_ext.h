/*
 * opengraph_amd_ext.h -- device-resident entry points of libopengraph_amd.
 *
 * None of these has a counterpart in the reference: they serve callers that
 * keep graphs in HBM across calls (benchmarks, batched and learned-start
 * PageRank, services).  They take plain pointers and sizes, return 0 on
 * success and -1 on failure with the cause in ogpu_last_error(), and never
 * compute on the CPU.  "d_" arguments are device pointers (from ogpu_malloc
 * or any HIP allocation on the same device).
 */
#ifndef OPENGRAPH_AMD_EXT_H
#define OPENGRAPH_AMD_EXT_H

#include <stddef.h>
#include <stdint.h>

#include "opengraph_amd.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct ogpu_graph ogpu_graph;   /* device forward + inverse CSR */
typedef struct ogpu_pr ogpu_pr;         /* a prepared PageRank run on a graph */

/* ---- device, memory ---- */
int ogpu_init(int device);
const char *ogpu_last_error(void);
int ogpu_device_info(char *name, size_t name_len, int *num_cus, size_t *total_mem);
int ogpu_malloc(void **d_ptr, size_t bytes);
int ogpu_free(void *d_ptr);
int ogpu_memcpy_h2d(void *d_dst, const void *src, size_t bytes);
int ogpu_memcpy_d2h(void *dst, const void *d_src, size_t bytes);
int ogpu_synchronize(void);
/* Device-to-device copy by a plain streaming kernel (bytes % 16 == 0); the
 * reference point for memory-traffic counters. */
int ogpu_device_copy(void *d_dst, const void *d_src, size_t bytes);

/* ---- synthetic inputs (the reference reads .bin files only) ----
 * R-MAT: edge e draws one quadrant per level from counter e*64+l of the
 * SplitMix64 stream (seed, 0) with probabilities a, b, c, 1-a-b-c; vertex
 * ids are then permuted by a keyed Feistel bijection of [0, 2^scale) when
 * permute != 0.  Duplicates and self-loops are kept.  Erdos-Renyi: src and
 * dst uniform in [0, n).  The host and device versions produce identical
 * arrays. */
int ogpu_rmat_device(uint32_t scale, uint64_t num_edges, double a, double b, double c,
                     uint64_t seed, int permute, uint32_t *d_src, uint32_t *d_dst);
int ogpu_rmat_host(uint32_t scale, uint64_t num_edges, double a, double b, double c, uint64_t seed,
                   int permute, uint32_t *src, uint32_t *dst);
int ogpu_er_device(uint32_t num_vertices, uint64_t num_edges, uint64_t seed, uint32_t *d_src,
                   uint32_t *d_dst);
int ogpu_er_host(uint32_t num_vertices, uint64_t num_edges, uint64_t seed, uint32_t *src,
                 uint32_t *dst);
/* Write (src,dst) u32 pairs in the reference's raw .bin layout (edgeList.c:383-407). */
int ogpu_write_bin(const char *path, const uint32_t *src, const uint32_t *dst, uint64_t num_edges);

/* ---- device CSR ---- */
/* Build forward + inverse CSR from device edge arrays (graphCSR.c:164-248
 * semantics; sort_mode as -o).  Inputs are not modified. */
int ogpu_graph_build(const uint32_t *d_src, const uint32_t *d_dst, uint32_t num_edges,
                     uint32_t sort_mode, ogpu_graph **out);
/* The reference's degree relabelling before the CSR build (graphCSR.c:184-203
 * with reorder.c:275-572 and :1618-1637; -l / -L / -O): lmodes[i] in 1 out-,
 * 2 in-, 3 in+out-degree sort, 4 / 5 DBG by out- / in-degree (0 skips a
 * level), applied in order; then the CSR of the relabelled edges, keeping
 * N = max id + 1 of the input.  d_labels_out (N entries, nullable) gets the
 * new label of every input vertex (the edge list's label_array).  Inputs
 * are not modified; another lmode is an error. */
int ogpu_graph_build_reordered(const uint32_t *d_src, const uint32_t *d_dst, uint32_t num_edges,
                               uint32_t sort_mode, const uint32_t *lmodes, uint32_t n_lmodes,
                               uint32_t *d_labels_out, ogpu_graph **out);
/* The reference's .csr file (graphCSR.c:380-558, directed, unweighted: E,
 * N, then per vertex forward out-degree, in-degree, edges_idx and the
 * inverse ones, then per edge forward src, dst and inverse src, dst, all
 * u32): ogpu_graph_write_csr writes <fname>.csr (the name the reference
 * gives it), ogpu_graph_read_csr loads such a file into a device graph
 * without preprocessing.  The (de)interleaving runs on the GPU. */
int ogpu_graph_write_csr(const ogpu_graph *g, const char *fname);
int ogpu_graph_read_csr(const char *path, ogpu_graph **out);
/* Upload host arrays in the order f_src, f_dst, f_out_degree, f_in_degree,
 * f_edges_idx, i_src, i_dst, i_out_degree, i_in_degree, i_edges_idx. */
int ogpu_graph_from_arrays(uint32_t num_vertices, uint32_t num_edges,
                           const uint32_t *const host_arrays[10], ogpu_graph **out);
int ogpu_graph_info(const ogpu_graph *g, uint32_t *num_vertices, uint32_t *num_edges);
enum
{
    OGPU_F_SRC = 0, OGPU_F_DST, OGPU_F_OUT_DEGREE, OGPU_F_IN_DEGREE, OGPU_F_EDGES_IDX,
    OGPU_I_SRC, OGPU_I_DST, OGPU_I_OUT_DEGREE, OGPU_I_IN_DEGREE, OGPU_I_EDGES_IDX
};
/* Copy one CSR array to host memory (E or N entries). */
int ogpu_graph_array(const ogpu_graph *g, uint32_t which, uint32_t *host_out);
void ogpu_graph_free(ogpu_graph *g);

/* ---- PageRank on a device graph ---- */
typedef struct
{
    uint32_t variant;       /* the reference's -p code, 0..13 */
    uint32_t max_iter;      /* arguments->iterations */
    double epsilon;         /* arguments->epsilon */
    uint32_t stop_rule;     /* 0: activeVertices == 0 (reference); 1: error_total < epsilon */
    uint32_t init_mode;     /* 0: 0.15 (reference); 1: uniform 1.0 (= 1/N normalized);
                               2: caller's vector (unnormalized, old ids) */
    uint32_t relabel;       /* 1: degree-ordered rows and columns, rows of more than 512
                               in-edges with their in-neighbours sorted by source (default);
                               2: the same without that sort; 0: rows grouped by in-degree
                               class only, original order within a class */
    uint32_t hub_threshold; /* rows with more in-edges are split into 512-edge chunks summed
                               in a fixed tree; rows with at most this many are summed in
                               the reference's order (bit-exact); 0: default 16;
                               0xFFFFFFFF: every row in the reference's order */
    uint32_t ppr;           /* 1: personalized PageRank (variant 0 only): start at the one-hot
                               e_s, x' = (1-d)[v==s] + d * sum, stop when the L1 change
                               sum |x'-x| < epsilon, ranks not divided by N */
    uint32_t ppr_source;    /* s, the caller's vertex id */
    uint32_t hot_cache;     /* pull (variants 0, 2, 4, 9-13): 0 default (on when the graph
                               exceeds 4x the copy and its cached sources send >= 1/5 of the
                               edges), 1 on, 2 off -- one persistent block per CU that
                               serves the gathers of the lowest vertex ids (the hottest
                               after relabelling) from a 157 KB LDS copy; same results.
                               push (variants 1, 3, 5): 0 default: partitioned -- the
                               destinations in ranges of at most 16384 rows, one block
                               per range adds each in-edge's contribution to its
                               destination's accumulator with an LDS atomic (one running
                               sum per destination, as the reference's atomics) and runs
                               the update; 1: atomic scatter over the forward edges with
                               LDS accumulators for the destinations with the lowest ids,
                               flushed once per block (their block partials are more
                               accurate than the reference's running fp32 sums and may
                               differ from them by ~1e-5 relative on rows with thousands
                               of in-edges); 2: plain atomic scatter.  Variant 3 is exact
                               in every form */
    uint32_t exact_above;   /* fp32 pull (variant 0, not personalized, relabel 1 or 2): the
                               rows of more than max(exact_above, 512) in-edges are summed
                               in the reference's order, one wavefront each, instead of in
                               chunks (bit-exact with the reference's running sums on equal
                               inputs); default 4096 (ogpu_pr_default_options), 0 off.
                               Such rows carry the largest fp32 drift of
                               the reference's sums.  With hot_cache's persistent kernel
                               the chains run beside it on a side stream.  fp32 push
                               (variant 1): when not 0, the destinations of more than
                               1024 in-edges are gathered in CSR order the same way
                               (the reference's source-ordered adds) and left out of the
                               atomic adds */
    uint32_t binned;        /* fp32 pull (variant 0, not personalized): 0 default (on),
                               1 on, 2 off -- the binned pull: per iteration a gather pass
                               writes each edge's contribution, source segment by source
                               segment from LDS, into an edge array grouped by destination
                               bins, and a sum pass adds every row from LDS in CSR order
                               (rows longer than a bin by one block each).  Every row is
                               the reference's running sum to the bit, and all global
                               accesses stream.  hub_threshold, exact_above and hot_cache
                               then apply only to the learned start's feature pass */
    uint32_t warm_classes;  /* learned start (ogpu_pr_set_init_learned): the predicted
                               start is corrected over this many in-degree classes (class
                               0 without in-edges, class c > 0 with floor(log2 in-degree)
                               = c - 1, the last one open-ended), each split into the
                               vertices with and without out-edges, so that each class's
                               mass equals that of one application of the iteration to
                               it: one scale factor per class from the classes' mass
                               balance s_i m_i = 0.15 n_i + sum_j A_ij s_j (m_i the start's
                               mass in class i, n_i its vertices, A_ij the mass the
                               iteration moves from class j to class i), one pass over the
                               edges and a (2 nin) x (2 nin) solve.  0: default (16), 1: no
                               correction, 2..16 */
} ogpu_pr_options;

typedef struct
{
    uint32_t iterations;    /* stats->iterations (the reference's off-by-one rule) */
    uint32_t rows;          /* iterations executed */
    uint32_t active;        /* last iteration's activeVertices */
    uint32_t pad;
    double error_total;     /* last iteration's error_total */
    double pr_sum;          /* sum of the normalized ranks */
    double seconds_total;   /* wall time from init to normalized ranks on device */
    double seconds_plan;    /* one-time plan build */
} ogpu_pr_result;

void ogpu_pr_default_options(ogpu_pr_options *opt);
int ogpu_pr_create(ogpu_graph *g, const ogpu_pr_options *opt, ogpu_pr **out);
/* The following runs start from this vector (host pointer, N floats, old
 * ids, unnormalized; init_mode becomes 2).  A vector with a NaN, infinite or
 * negative entry, or with no positive mass, is rejected: the runs then start
 * from the uniform vector (init_mode 1) and ogpu_pr_init_fallback returns 1. */
int ogpu_pr_set_init(ogpu_pr *pr, const float *init);
/* 1 when the last start vector offered (ogpu_pr_set_init or
 * ogpu_pr_set_init_learned) was rejected in favour of the uniform start,
 * 0 when it is used, -1 for a NULL run. */
int ogpu_pr_init_fallback(const ogpu_pr *pr);
/* The start vector the following runs use (N floats, old ids,
 * unnormalized): the one set, or the uniform / reference start. */
int ogpu_pr_get_init(ogpu_pr *pr, float *out);
/* Run to the stop rule or max_iter and normalize; ranks stay on device. */
int ogpu_pr_run(ogpu_pr *pr, ogpu_pr_result *res);
/* Benchmark helpers: reset the state, then enqueue n iterations with the
 * stop rule disabled (every iteration does full work); optionally time the
 * dominant kernel of each iteration with HIP events on the library stream. */
int ogpu_pr_reset(ogpu_pr *pr);
int ogpu_pr_enqueue(ogpu_pr *pr, uint32_t n, int time_kernels);
int ogpu_pr_kernel_ms(ogpu_pr *pr, double *total_ms, uint32_t *count);
/* Normalized ranks (old ids) to host memory. */
int ogpu_pr_ranks(ogpu_pr *pr, float *host_out);
/* Per-iteration activeVertices / error_total rows of the last run. */
int ogpu_pr_rows(ogpu_pr *pr, uint32_t *active, double *error_total, uint32_t max_rows);
/* Algorithmic HBM bytes of one iteration of this variant (pull fp32:
 * 8E + 20N). */
double ogpu_pr_bytes_per_iteration(const ogpu_pr *pr);
void ogpu_pr_free(ogpu_pr *pr);

/* ---- autotuner (no counterpart in the reference) ----
 * Times `iters` plain pull iterations (after two warm-up ones, host clock,
 * launch gaps included) for each candidate hub_threshold (the row-class
 * boundary: rows with more in-edges are chunked) and returns the fastest in
 * *best; ms[i] (may be NULL) gets candidate i's milliseconds per iteration. */
int ogpu_tune_hub_threshold(ogpu_graph *g, const uint32_t *cands, uint32_t n_cands, uint32_t iters,
                            uint32_t *best, double *ms);

/* ---- many graphs in one launch (no counterpart in the reference) ----
 * A plain fp32 pull PageRank (variant 0, reference start and stop rule,
 * opt->max_iter / epsilon) on each of n graphs.  hub_thresholds[g] (or, if
 * NULL, opt->hub_threshold) fixes graph g's row-class boundary; 0 means
 * autotuned (ogpu_tune_hub_threshold over 8..128, once per size class of
 * graphs with the same ceil(log2 N), on the class's largest graph).
 *
 * Each graph's pull also has a tile shape: threads per block (128, 256 or
 * 512), 512-edge chunks one wavefront sums together (1, 2 or 4), and edges
 * a thread loads together on its row (8, or 0: 1, 2, 4 or 8 as the
 * wavefront's longest row needs).  Every shape gives the same ranks;
 * ogpu_multi_tile_shapes lists the compiled ones.  tiles[g].block == 0 (or
 * tiles NULL, as in ogpu_multi_create) means autotuned: once per size class,
 * on all the class's graphs pulled together, timed with HIP events, by
 * coordinate descent over block size, then chunks per wavefront, then row
 * batch from (256, 4, 8).  The graphs of one shape share one pull launch
 * per iteration. */
typedef struct ogpu_multi ogpu_multi;
typedef struct ogpu_multi_tile
{
    uint32_t block;            /* threads per block */
    uint32_t chunks_per_wave;  /* 512-edge chunks a wavefront sums together */
    uint32_t row_batch;        /* edges a thread loads together on its row; 0 = adaptive */
    uint32_t pad;
} ogpu_multi_tile;
int ogpu_multi_create(ogpu_graph *const *graphs, uint32_t n, const ogpu_pr_options *opt,
                      const uint32_t *hub_thresholds, ogpu_multi **out);
int ogpu_multi_create_tiled(ogpu_graph *const *graphs, uint32_t n, const ogpu_pr_options *opt,
                            const uint32_t *hub_thresholds, const ogpu_multi_tile *tiles, ogpu_multi **out);
/* The compiled tile shapes (up to max into out, which may be NULL); returns their number. */
uint32_t ogpu_multi_tile_shapes(ogpu_multi_tile *out, uint32_t max);
/* Graph g's tile shape (given or tuned). */
int ogpu_multi_tile_of(const ogpu_multi *m, uint32_t g, ogpu_multi_tile *t);
/* mode 0: one graph after another, each with its own launches and stop
 *         checks (what per-graph runs cost);
 * mode 1: every iteration of all graphs as three persistent launches over
 *         one work list, max_iter iterations enqueued without host checks
 *         (finished graphs skip their work, launches after the last stop
 *         return at once);
 * mode 2: the same iterations captured once into a hipGraph and launched
 *         as one graph.
 * seconds: host clock over init and iterations (results not normalized). */
int ogpu_multi_run(ogpu_multi *m, int mode, double *seconds);
/* Graph g's result of the last run, as ogpu_pr_run's (ranks divided by N,
 * caller's vertex order; ranks may be NULL). */
int ogpu_multi_result(ogpu_multi *m, uint32_t g, ogpu_pr_result *res, float *ranks);
uint32_t ogpu_multi_hub_threshold(const ogpu_multi *m, uint32_t g);
void ogpu_multi_free(ogpu_multi *m);

/* ---- learned warm start (no counterpart in the reference) ----
 * A per-vertex MLP on degree features, f = [log1p in-degree, log1p
 * out-degree, mean log1p in-degree of the in-neighbours, 1/N], layers
 * 4 -> 64 -> 64 -> 1 (ReLU, softplus output), fp32 parameters packed as
 * W1[64][4] | b1[64] | W2[64][64] | b2[64] | W3[64] | b3 (4545 floats).
 * It predicts the shape of the converged rank vector.  The start vector:
 * vertices without in-edges get their fixed-point rank 1 - d (unnormalized)
 * exactly; the others the predicted shape renormalized to sum 1 and scaled
 * to the mass the reference's iteration keeps at its fixed point (from
 * S = 0.15 + 0.85 (S - mass on vertices without out-edges)).  The unchanged
 * iteration, to the same tolerance, then runs from it; the uniform start is
 * the fallback. */
typedef struct ogpu_mlp ogpu_mlp;
int ogpu_mlp_create(uint64_t seed, ogpu_mlp **out);
/* get (4545 floats) and/or set the parameters (set resets Adam's state). */
int ogpu_mlp_params(ogpu_mlp *m, float *get, const float *set);
/* Adds graph g to the training set: features and the target
 * N * x / sum(x) of its converged ranks x, from a plain pull on the GPU
 * (uniform start, stop at L1 change < tol, max_iter); *l1_out (may be
 * NULL) gets the last L1 change. */
int ogpu_mlp_add_graph(ogpu_mlp *m, ogpu_graph *g, double tol, uint32_t max_iter, double *l1_out);
uint32_t ogpu_mlp_data_size(const ogpu_mlp *m);
/* Adam (lr, 0.9, 0.999, 1e-8) on the mean L1 loss |o - y|, mini-batches
 * of `batch` vertices in a seeded order per epoch; hand-written forward and
 * backward kernels, deterministic gradient sums. */
int ogpu_mlp_train(ogpu_mlp *m, uint32_t epochs, uint32_t batch, float lr, uint64_t seed,
                   double *last_loss, double *seconds);
/* Mean loss and gradient (4545) of the given vertices (host arrays, n x 4
 * features and n targets). */
int ogpu_mlp_grad(ogpu_mlp *m, const float *feat, const float *y, uint32_t n, float *grad,
                  double *loss);
/* No counterpart in the reference.  For callers that cannot allocate device
 * memory (code under HIP graph capture, training loops that forbid
 * hipMalloc): ogpu_mlp_grad over device-resident rows -- d_feat (n x 4) and
 * d_y (n) -- writing the mean gradient to d_grad (4545 floats) and the sum
 * of the n losses |o - y| to *d_loss_sum (one double), all device
 * pointers.  Its block partials live in a static device array of the
 * library (4.4 MiB, held by each GPU the library runs on until the process
 * ends), so the call allocates nothing.  At most OGPU_MLP_GRAD_NOALLOC_MAX
 * rows per call (more is an error); n = 0 does nothing.  Enqueued on the
 * library stream (ogpu_synchronize waits for it).  All calls on a GPU share
 * that array: calls that overlap on one GPU race. */
#define OGPU_MLP_GRAD_NOALLOC_MAX 16384u
int ogpu_mlp_grad_noalloc(ogpu_mlp *m, const float *d_feat, const float *d_y, uint32_t n, float *d_grad,
                          double *d_loss_sum);
/* o for n feature rows (host arrays); mfma = 1: layer 2 on bf16 MFMA. */
int ogpu_mlp_forward(ogpu_mlp *m, const float *feat, uint32_t n, int mfma, float *out);
void ogpu_mlp_free(ogpu_mlp *m);
/* Features of a fp32 pull run's graph, caller's vertex order (N x 4). */
int ogpu_pr_features(ogpu_pr *pr, float *feat);
/* Computes the learned start vector on the device (features, bf16 MFMA
 * inference, mass scaling) and makes the following runs of pr start from
 * it; *seconds (may be NULL): host time of this call.  A prediction with a
 * NaN, infinite or negative value on a vertex with in-edges, or with no
 * mass there, is rejected on the device: the runs then start from the
 * uniform vector (init_mode 1) and ogpu_pr_init_fallback returns 1. */
int ogpu_pr_set_init_learned(ogpu_pr *pr, ogpu_mlp *m, double *seconds);

/* ---- batched personalized PageRank (no counterpart in the reference) ----
 * k <= 64 restart vertices iterated together as a pull-SpMM over an N x k
 * fp32 row-major block: per column exactly the personalized option of
 * ogpu_pr (variant 0, ppr = 1) -- x0 = e_s, x' = (1-d)[v==s] + d * sum of
 * x/outdeg over in-neighbours, column stops after the first iteration with
 * L1 change < epsilon -- with rows of at most 512 in-edges summed in CSR
 * order (bit-exact with the single-source run at hub_threshold 0xFFFFFFFF)
 * and longer rows summed in 512-edge chunks added in order. */
typedef struct ogpu_ppr ogpu_ppr;

typedef struct
{
    uint32_t iterations;      /* iterations executed (the slowest column's count) */
    uint32_t converged;       /* columns that reached L1 < epsilon */
    double seconds_total;     /* init + iterations, host clock */
    double seconds_plan;
    double kernel_ms_per_iter;/* mean duration of the SpMM kernel, HIP events */
} ogpu_ppr_result;

/* sources: k vertex ids of the caller (old ids). */
int ogpu_ppr_create(ogpu_graph *g, const uint32_t *sources, uint32_t k, uint32_t max_iter,
                    double epsilon, ogpu_ppr **out);
int ogpu_ppr_run(ogpu_ppr *p, ogpu_ppr_result *res);
/* x_out: N x k, row-major, caller's vertex order (not divided by N). */
int ogpu_ppr_ranks(ogpu_ppr *p, float *x_out);
/* iters[k]: iterations applied per column; l1_hist[rows x k]: each
 * iteration's L1 change per column (0 after the column stopped).  Returns
 * the rows written (<= max_rows) or -1. */
int ogpu_ppr_columns(ogpu_ppr *p, uint32_t *iters, double *l1_hist, uint32_t max_rows);
/* Algorithmic bytes of one iteration, E (4 + 4k') + N (8 + 12k') with k'
 * the column slots the iteration gathers: after a run, the mean over its
 * executed iterations (slots shrink as columns converge); before, k' = k. */
double ogpu_ppr_bytes_per_iteration(const ogpu_ppr *p);
void ogpu_ppr_free(ogpu_ppr *p);

/* ---- SpMV (SPMV.c:588-1068, the reference's SpMV benchmark on CSR) ----
 * y = 0, then `iterations` passes y += sum over the in-edges (src, v) of
 * w * x[src], with x[v] = 1.0f / out-degree(v) (0.001f without out-edges)
 * and w = 0.0001f, in the arithmetic of -p code:
 *   0 pull, fp32: by default the binned pull, every row continuing y in the
 *     reference's CSR order (bit-exact); with a hub_threshold, the gather
 *     pull: rows of at most hub_threshold in-edges bit-exact, longer rows
 *     add their fixed-tree sum to y (0xFFFFFFFF: every row bit-exact);
 *   1 push, fp32 atomics over the forward edges (unordered, as the
 *     reference's omp atomic), the hottest destinations summed per block
 *     in LDS first;
 *   2 pull, Q16.16 inputs, u64 sums of (7 * x) >> 16 (bit-exact);
 *   3 push, Q32.32, u64 sums of (w * x) >> 32 (bit-exact): gathered over
 *     the in-edges by default, scattered with u64 atomics when scatter = 1. */
typedef struct ogpu_spmv ogpu_spmv;
typedef struct
{
    uint32_t code;           /* the reference's -p, 0..3 */
    uint32_t hub_threshold;  /* 0 (default): -p 0 runs on the binned pull, every row's running
                                y in the reference's order; otherwise -p 0 runs on the gather
                                pull with this row-class threshold, as ogpu_pr_options */
    uint32_t hot_cache;      /* as ogpu_pr_options: the pushes (-p 1, -p 3 scattered) are
                                partitioned by destination range by default (0), 1 scatters
                                with LDS accumulators for the hottest destinations, 2 with
                                plain atomics */
    uint32_t scatter;        /* code 3: 1 scatters with atomics instead of gathering */
} ogpu_spmv_options;
void ogpu_spmv_default_options(ogpu_spmv_options *opt);
int ogpu_spmv_create(ogpu_graph *g, const ogpu_spmv_options *opt, ogpu_spmv **out);
/* Reset y, run `iterations` passes, map vector_output to the caller's ids
 * (kept on device); *seconds = device time of all of it. */
int ogpu_spmv_run(ogpu_spmv *s, uint32_t iterations, double *seconds);
/* vector_output of the last run, N floats, caller's vertex order. */
int ogpu_spmv_output(ogpu_spmv *s, float *host_out);
/* Device time of iteration i of the last run (iteration 0 includes the reset). */
int ogpu_spmv_iteration_seconds(ogpu_spmv *s, uint32_t i, double *seconds);
/* Benchmark helpers as ogpu_pr_reset / _enqueue / _kernel_ms: the timed
 * kernel is the pull gather or the push scatter. */
int ogpu_spmv_reset(ogpu_spmv *s);
int ogpu_spmv_enqueue(ogpu_spmv *s, uint32_t n, int time_kernels);
int ogpu_spmv_kernel_ms(ogpu_spmv *s, double *total_ms, uint32_t *count);
/* Algorithmic HBM bytes of one pass (pull fp32: 8E + 12N). */
double ogpu_spmv_bytes_per_iteration(const ogpu_spmv *s);
void ogpu_spmv_free(ogpu_spmv *s);

/* ---- rank comparison statistics (graphStats.c:580-688 on the GPU) ----
 * P problems: problem p compares the two rankings over positions
 * [num_vertices[p] - k, num_vertices[p]) with k = min(topk[p],
 * num_vertices[p]) and writes out[p] as collectStatsPageRank_topK would.
 * ref_pr / pr: N ranks; ref_order / order: realRanks (permutations of
 * [0, N) ascending by rank).  _device takes device pointers (the rankings
 * stay resident); the other uploads host arrays first. */
int ogpu_rank_stats(const float *ref_pr, const uint32_t *ref_order, const float *pr, const uint32_t *order,
                    uint32_t N, const uint32_t *topk, const uint32_t *num_vertices, uint32_t P,
                    struct PageRankCorrelationStats *out);
int ogpu_rank_stats_device(const float *d_ref_pr, const uint32_t *d_ref_order, const float *d_pr,
                           const uint32_t *d_order, uint32_t N, const uint32_t *topk,
                           const uint32_t *num_vertices, uint32_t P, struct PageRankCorrelationStats *out);

#ifdef __cplusplus
}
#endif

#endif
