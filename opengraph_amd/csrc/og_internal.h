/*
 * og_internal.h -- private declarations of libopengraph_amd shared by the
 * host C files and the HIP kernel files (whose launchers are extern "C").
 */
#ifndef OG_INTERNAL_H
#define OG_INTERNAL_H

#include <stddef.h>
#include <stdint.h>

#include <hip/hip_runtime_api.h>

#include "og_common.h"

#ifdef __cplusplus
extern "C" {
#endif

/* ---------------- errors and device context (og_runtime.c) ---------------- */

void og_set_error(const char *fmt, ...);
const char *og_get_error(void);
/* Makes sure the device and stream exist; returns 0 or -1 with the error set. */
/* The library works on one device per process (ogpu_init), on one stream;
 * its entry points are not reentrant: callers on several host threads
 * serialize their calls, as the reference's single caller does. */
int og_ctx_ensure(void);
hipStream_t og_stream(void);
/* the device the context was created on */
int og_device(void);
/* the second stream and the fork / join events of the pull (created with
 * the context) */
hipStream_t og_side_stream(void);
hipEvent_t og_side_fork(void);
hipEvent_t og_side_join(void);
int og_num_cus(void);
void og_trace_entry(const char *fn, uint32_t n_vertices, uint32_t n_edges);
/* hipMalloc / hipFree with error text; og_dalloc returns NULL on failure. */
void *og_dalloc(size_t bytes);
void og_dfree(void *p);
/* Checks a HIP status; on failure records "<what>: <hip error>" and returns -1. */
int og_hip_check(hipError_t e, const char *what);
/* Kernel launches.  The runtime keeps a failed call's error until
 * hipGetLastError reads it, so a check after a launch could report an
 * earlier call's error that its caller already handled.  OG_LAUNCH clears
 * that first, launches, and keeps the launch's own status: og_launch_status
 * returns the first failing launch's since its last call (and resets it). */
void og_launch_note(hipError_t e);
hipError_t og_launch_status(void);
#define OG_LAUNCH(...)                     \
    do                                     \
    {                                      \
        (void)hipGetLastError();           \
        hipLaunchKernelGGL(__VA_ARGS__);   \
        og_launch_note(hipGetLastError()); \
    } while (0)
/* A copy ordered after all work already queued on the library stream and
 * complete on return (hipMemcpy itself runs on the null stream, which the
 * library's non-blocking stream does not wait for). */
hipError_t og_copy(void *dst, const void *src, size_t bytes, hipMemcpyKind kind);
#define OG_HIP(call) do { if (og_hip_check((call), #call)) return -1; } while (0)
/* 128-byte aligned host allocation, as the reference's my_malloc
 * (myMalloc.c:24-67); exits on failure like the reference. */
void *og_host_alloc(size_t bytes);

/* ---------------- generator / CSR kernels (og_csr.hip) ---------------- */

int og_launch_rmat(uint32_t *d_src, uint32_t *d_dst, uint64_t num_edges, const og_rmat_t *g,
                   hipStream_t s);
int og_launch_er(uint32_t *d_src, uint32_t *d_dst, uint64_t num_edges, uint32_t n, uint64_t key,
                 hipStream_t s);
/* d_out[0] = max over src and dst (d_out zeroed here). */
int og_launch_max_id(const uint32_t *d_src, const uint32_t *d_dst, uint32_t n, uint32_t *d_out,
                     hipStream_t s);
/* The .csr file body (graphCSR.c:380-449): 6 words per vertex, then 4 per
 * edge, from / into the arrays f_src, f_dst, f_out_degree, f_in_degree,
 * f_edges_idx, i_src, i_dst, i_out_degree, i_in_degree, i_edges_idx. */
int og_launch_csr_pack(uint32_t *const arrays[10], uint32_t N, uint32_t E, uint32_t *body, hipStream_t s);
int og_launch_csr_unpack(const uint32_t *body, uint32_t N, uint32_t E, uint32_t *const arrays[10],
                         hipStream_t s);
/* Split interleaved (src,dst) u32 pairs into two arrays. */
int og_launch_deinterleave(const uint32_t *d_pairs, uint32_t n, uint32_t stride_words,
                           uint32_t *d_src, uint32_t *d_dst, hipStream_t s);

/* Stable LSD radix sort of (key, val) u32 pairs over key bits [0, key_bits),
 * 8 bits per pass.  Sorts from (k_in, v_in) using (k_tmp, v_tmp) as the
 * other buffer; *result_in_tmp tells which pair holds the output.  k_in may
 * equal a different array than the output lists (e.g. the swapped
 * forward arrays for the inverse list): the input arrays are only read in
 * the first pass when first_out_is_tmp... see og_csr.hip. */
size_t og_radix_scratch_bytes(uint32_t n);
int og_radix_sort_pairs(const uint32_t *k_in, const uint32_t *v_in, uint32_t *k_a, uint32_t *v_a,
                        uint32_t *k_b, uint32_t *v_b, uint32_t n, uint32_t key_bits,
                        void *d_scratch, hipStream_t s, int *result_in_b);

/* out_degree / edges_idx of a sorted key list (vertex.c:262-289). */
int og_launch_map_vertices(const uint32_t *d_sorted_src, uint32_t n_edges, uint32_t n_vertices,
                           uint32_t *d_out_degree, uint32_t *d_edges_idx, hipStream_t s);

/* ---------------- device graph (og_graph.c) ---------------- */

typedef struct og_graph
{
    uint32_t N, E;
    /* forward (sorted by src) and inverse (sorted by dst) CSR on device */
    uint32_t *f_src, *f_dst, *f_out_degree, *f_in_degree, *f_edges_idx;
    uint32_t *i_src, *i_dst, *i_out_degree, *i_in_degree, *i_edges_idx;
    /* host identity of the GraphCSR this copy was uploaded from */
    const void *host_key[8];
    uint32_t sort_mode;
} og_graph;

int og_graph_build_device(const uint32_t *d_src, const uint32_t *d_dst, uint32_t E,
                          uint32_t sort_mode, int keep_inputs, og_graph **out);
/* The same with N = max(max id + 1, min_vertices): a relabelled edge list
 * keeps its vertex count even when its last label has no edge. */
int og_graph_build_device_n(const uint32_t *d_src, const uint32_t *d_dst, uint32_t E,
                            uint32_t min_vertices, uint32_t sort_mode, og_graph **out);

/* ---------------- degree relabelling (og_reorder.hip, og_graph.c) ---------------- */

/* DBG thresholds (reorder.c:404-414): n buckets, t[n-1] = UINT32_MAX. */
typedef struct og_dbg_thresholds
{
    uint32_t n;
    uint32_t t[16];
} og_dbg_thresholds;

int og_launch_reorder_degrees(const uint32_t *src, const uint32_t *dst, uint32_t E, uint32_t count_src,
                              uint32_t count_dst, uint32_t *deg, hipStream_t s);
int og_launch_reorder_dbg_keys(const uint32_t *deg, uint32_t N, og_dbg_thresholds th, uint32_t *key,
                               uint32_t *ids, hipStream_t s);
int og_launch_iota(uint32_t *ids, uint32_t N, hipStream_t s);
int og_launch_reorder_labels(const uint32_t *sorted, uint32_t N, uint32_t reverse, uint32_t *labels,
                             hipStream_t s);
int og_launch_relabel(uint32_t *src, uint32_t *dst, uint32_t E, const uint32_t *labels, hipStream_t s);
int og_launch_compose_labels(uint32_t *label_array, uint32_t *inverse, uint32_t N, const uint32_t *labels,
                             hipStream_t s);
/* The reference's thresholds from the integer average degree (reorder.c:404-414). */
og_dbg_thresholds og_dbg_default_thresholds(uint32_t avg_degree);
/* d_labels[v] = new label of v for -l lmode (1 out-, 2 in-, 3 in+out-degree
 * sort, 4 / 5 DBG by out- / in-degree; reorder.c:275-572) of the device
 * edge list (d_src, d_dst) with N vertices; avg_degree sets the DBG
 * thresholds.  Returns 0, or -1 (error set) for another lmode. */
int og_reorder_labels_device(const uint32_t *d_src, const uint32_t *d_dst, uint32_t E, uint32_t N,
                             uint32_t avg_degree, uint32_t lmode, uint32_t *d_labels);
/* The relabellings lmodes[0..n) (0 skips) applied in turn to the device
 * edge list in place; *labels_out (device, N entries, the caller frees it
 * with og_dfree) gets the composed label of every input vertex. */
int og_graph_build_reordered_inplace(uint32_t *src, uint32_t *dst, uint32_t E, uint32_t N,
                                     const uint32_t *lmodes, uint32_t n_lmodes, uint32_t **labels_out);
/* The same from given degrees (reorder.c:296-384 / :439-572): degree sort
 * (lmode 1..3) or DBG buckets with thresholds th (lmode 4, 5). */
int og_reorder_labels_from_degrees(const uint32_t *d_deg, uint32_t N, uint32_t dbg, og_dbg_thresholds th,
                                   uint32_t *d_labels);
void og_graph_destroy(og_graph *g);

/* ---------------- PageRank plan / run (og_pagerank.c, og_pr.hip) ---------------- */

/* Pull plan.  Rows (new ids) fall in four classes, in this order:
 *   [0, H)  more than OG_CH in-edges: several chunks of OG_CH edges, each
 *           summed by one wavefront; the finalize kernel adds the chunk
 *           sums of a row in chunk order.
 *   [H, C)  more than seq_t and at most OG_CH in-edges: one chunk, whose
 *           wavefront also runs the row's epilogue.
 *   [C, Z)  1..seq_t in-edges: one thread sums the row in CSR order.
 *   [Z, N)  no in-edges: epilogue only.
 * Rows with at most seq_t in-edges therefore add in exactly the reference's
 * order; chunked rows add in a fixed tree (deterministic).  With chain_t
 * (>= OG_CH, relabel 1 or 2), the rows of more than chain_t in-edges come
 * first, [0, L), longest first: the fp32 pull sums each of them in the
 * reference's order with one wavefront and skips their chunks. */
#define OG_CH 512u            /* edges per wavefront chunk: 64 lanes x 8 */
#define OG_CH_PER_WAVE 4u     /* consecutive chunks one wavefront processes */
#define OG_SEQ_T 32u          /* seq_t of the batched-PPR plan */
#define OG_CH_MULTI 0x80000000u /* ch_row flag: the row has several chunks */

typedef struct og_plan
{
    uint32_t N, E;
    uint32_t relabel;
    uint32_t seq_t;              /* rows with more in-edges are chunked */
    uint32_t *new2old, *old2new; /* N each */
    uint32_t *row_ptr;           /* N+1, rows in new order (inverse CSR) */
    uint32_t *cols;              /* E, in-neighbours as new ids, reference order kept */
    uint32_t *outdeg;            /* N, forward out-degree in new order */
    uint32_t H, C, Z;            /* class boundaries, see above */
    uint32_t L;                  /* rows [0, L) of class 0: more than chain_t in-edges, in-degree
                                    descending, reference order kept (chain_t 0: L = 0) */
    uint32_t ch_first;           /* first chunk of row L */
    uint32_t n_chunks;           /* chunks of rows [0, C) */
    uint32_t *chunk_base;        /* H+1: first chunk of multi-chunk row h */
    uint32_t *ch_row;            /* n_chunks: row of the chunk | OG_CH_MULTI */
    uint32_t *ch_edge;           /* n_chunks+1: first edge of the chunk */
    uint32_t n_chunk_blocks;     /* pull blocks over the chunks [ch_first, n_chunks) */
    uint32_t n_row_blocks;       /* pull blocks over rows [C, N), one row per thread */
    /* forward side for push: source and destination of each forward edge, new ids */
    uint32_t *push_src, *push_dst;
    /* forward CSR in new ids for the data-driven activation */
    uint32_t *f_row_ptr;         /* N+1 */
    uint32_t *f_cols;            /* E */
} og_plan;

int og_plan_build(const og_graph *g, uint32_t relabel, uint32_t seq_t, uint32_t chain_t, int need_push,
                  og_plan **out);
void og_plan_destroy(og_plan *p);

/* Per-iteration statistics written by the device. */
typedef struct
{
    double error_total;
    uint32_t active;
    uint32_t pad;
} og_iter_stat;

#ifdef __cplusplus
}
#endif

#endif /* OG_INTERNAL_H */
