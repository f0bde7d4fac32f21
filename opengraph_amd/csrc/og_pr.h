/*
 * og_pr.h -- argument block and launchers of the PageRank kernels (og_pr.hip),
 * shared with the host orchestration in og_pagerank.c.
 */
#ifndef OG_PR_H
#define OG_PR_H

#include <stdint.h>

#include <hip/hip_runtime_api.h>

#include "og_internal.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Arithmetic of the gathered contributions. */
#define OG_MODE_F32 0        /* -p 0, 1 */
#define OG_MODE_FX64 1       /* -p 2, 3 */
#define OG_MODE_FX32 2       /* -p 9 */
#define OG_MODE_FX16 3       /* -p 10 */
#define OG_MODE_FX8 4        /* -p 11 */
#define OG_MODE_Q32 5        /* -p 4 */
#define OG_MODE_Q16 6        /* -p 12 */
#define OG_MODE_Q8 7         /* -p 13 */
#define OG_MODE_QPUSH_EPI 8  /* -p 5 */
/* graphs of at most this many vertices run the data-driven push and
 * pull-push in vertex order on one thread: the reference's one OpenMP
 * chunk (schedule(dynamic, 1024), pageRank.c:2645, :2808) */
#define OG_DD_SERIAL_N 1024u

/* Iteration index that writes no statistics row (the init-time reduction). */
#define OG_NO_STATS 0xFFFFFFFFu


typedef struct
{
    double scale_d;
    float scale_f;
    uint32_t zero;
} og_qparams;

/* Everything a PageRank kernel needs, passed by value at each launch. */
typedef struct og_pr_dev
{
    uint32_t N, E, H, C, Z;  /* row classes of the plan (og_internal.h) */
    uint32_t n_chunks, n_chunk_blocks, n_row_blocks;
    uint32_t row_end;        /* pull rows [C, row_end) one per thread: N, or Z when */
    uint32_t z_active;       /* rows [Z, N) are left out; this is added to the active count */
    uint32_t n_parts;        /* valid first-level partials for the next final reduction */
    uint32_t n_final_blocks; /* blocks of the final reduction */
    uint32_t stop_l1;        /* 0: stop when active == 0; 1: when error_total < epsilon */
    uint32_t qmode;          /* 0 or the quantized mode whose parameters the final step computes */
    uint32_t qzero;          /* compute a zero point (quantized push) */
    uint32_t allow_stop;     /* final reduction may raise *done */
    uint32_t hot_grid;       /* > 0: pull runs as og_pull_hot_kernel on this many blocks */
    uint32_t raw;            /* fp32 pull writes the plain in-neighbour sum to r (features) */
    uint32_t ppr;            /* personalized: restart only at ppr_src, error = L1 (no 1/N) */
    uint32_t ppr_src;        /* new id of the restart vertex */
    uint32_t spmv;           /* pull adds the row sums onto y (SpMV): fp32 y in r, u64 in next */
    /* fp32 pull rows summed in the reference's order by one wavefront each
     * (chains): n_chains of them, in the first n_chain_blocks pull blocks
     * (with hot_grid: the n_chain_blocks one-wavefront blocks of
     * og_pull_chains_kernel, launched beside the persistent kernel).
     * Every row exact: chain_rows lists rows [0, C) longest first (H = 0, no
     * chunks); otherwise (chain_rows NULL) the chains are rows [0, n_chains)
     * of the multi-chunk class, whose chunks [0, ch_first) are skipped. */
    const uint32_t *chain_rows;
    uint32_t n_chains, n_chain_blocks, ch_first;
    /* first statistics partial of og_pull_chains_kernel's blocks */
    uint32_t chain_part_base;
    /* fp32 push: destinations [0, push_skip) are the chains, summed by
     * gathering in CSR order; the scatter and its epilogue leave them out */
    uint32_t push_skip;
    /* partitioned push (pp_v0 set): destination ranges [pp_v0[p],
     * pp_v0[p + 1]) of at most OG_PP_ROWS rows, ranges [pp_first, n_pp)
     * pushed (pp_v0[pp_first] = push_skip); pp_dloc[e] = the destination of
     * inverse-CSR edge e minus its range's first row */
    const uint32_t *pp_v0;
    const uint16_t *pp_dloc;
    uint32_t n_pp, pp_first;
    double epsilon;
    /* plan */
    const uint32_t *row_ptr, *cols, *outdeg, *chunk_base, *ch_row, *ch_edge;
    const uint32_t *new2old, *old2new;
    const uint32_t *push_src, *push_dst;
    /* state */
    float *r;
    const void *g_in;        /* gathered contributions of this iteration */
    void *c_out;             /* contributions written by this iteration's epilogue */
    void *next;              /* push accumulators */
    const void *cq_src;      /* quantize: fp32 contributions */
    void *cq_dst;            /* quantize: quantized contributions */
    void *chunk_sum;         /* chunk sums of the rows [0, H) (acc type) */
    double *part_err;
    uint32_t *part_act;
    float *part_min, *part_max;
    double *fin_err;
    uint32_t *fin_act;
    float *fin_min, *fin_max;
    uint32_t *ticket;
    uint32_t *done;
    og_iter_stat *stats;
    og_qparams *qp, *qp_next;
    /* data-driven variants (pageRank.c:2405-2896) */
    uint32_t dd;             /* 6: pull over worklists; 7: residual push; 8: pull-push */
    const uint32_t *f_row_ptr, *f_cols;
    uint8_t *wl_cur, *wl_next, *changed;
    float *resid;
} og_pr_dev;

int og_launch_pull(const og_pr_dev *a, int mode, uint32_t it, hipStream_t s, hipEvent_t ev0,
                   hipEvent_t ev1);

/* The persistent pull runs one block of OG_HOT_SUBS * 256 threads per CU
 * (16 wavefronts, the occupancy its registers allow) and keeps
 * OG_HOT_BYTES of contributions in LDS -- most of a CU's 160 KB. */
#define OG_HOT_SUBS 4u
#define OG_HOT_BYTES 160768u

/* First-level statistics partials one pull launch writes. */
static inline uint32_t og_pull_parts(const og_pr_dev *a)
{
    return a->hot_grid ? a->hot_grid + a->n_chain_blocks : a->n_chain_blocks + a->n_chunk_blocks + a->n_row_blocks;
}

/* ---- many graphs in one launch (fp32 pull) ----
 * Work item: block `first` of graph g's pull (kind 0 chunk block, 1 row
 * block), of its long-row finalize (kind 2, rows [first*4, first*4+4)) or
 * of its init (kind 3, vertices [first*256, first*256+256)).  Items are
 * grouped by graph; the kernels are persistent (block b takes items
 * b, b + grid, ...) and write one statistics partial per item. */
typedef struct
{
    uint32_t g, kind, first, pad;
} og_multi_work;

typedef struct og_multi_dev
{
    const og_pr_dev *devs;       /* 2 per graph: [2g + parity of the iteration] */
    const og_multi_work *work;   /* pull items */
    const og_multi_work *lwork;  /* long-row items */
    const og_multi_work *iwork;  /* init items */
    uint32_t n_work, n_lwork, n_iwork, n_graphs;
    const uint32_t *wr;          /* 2 n_graphs: [2g, 2g + 1) = graph g's pull items (those of a
                                  * tile shape are together, so graphs are not in order) */
    const uint32_t *lr;          /* n_graphs + 1: long-row item ranges per graph */
    double *w_err, *l_err;
    uint32_t *w_act, *l_act;
    uint32_t *live;              /* graphs not yet stopped */
    uint32_t *all_done;          /* set when live reaches 0 */
} og_multi_dev;

/* Tile shape of the multi-graph pull: threads per block, chunks per
 * wavefront, edges per load batch of a thread's row (0: adapted to the
 * wavefront's longest row). */
typedef struct og_multi_tile
{
    uint32_t block, chunks, row_batch;
} og_multi_tile;
/* The items of one pull launch, split by XCD: blocks b with b % 8 == x
 * (the XCD the hardware places them on) take items [o[x], o[x + 1]). */
#define OG_XCDS 8u
typedef struct og_multi_xr
{
    uint32_t o[OG_XCDS + 1];
} og_multi_xr;
/* Pull items xr of one tile shape, on grid blocks (a multiple of 8). */
typedef struct og_multi_group
{
    uint32_t tile, grid;
    og_multi_xr xr;
} og_multi_group;
uint32_t og_multi_tile_count(void);
int og_multi_tile_shape(uint32_t id, og_multi_tile *t);
int og_launch_multi_init(const og_multi_dev *m, uint32_t grid, hipStream_t s);
/* The pull of iteration `it` over the items xr with tile shape `tile`. */
int og_launch_multi_pull(const og_multi_dev *m, uint32_t tile, const og_multi_xr *xr, uint32_t it,
                         uint32_t grid, hipStream_t s);
/* One iteration: the pull of every group, the long rows (grid blocks), the
 * per-graph statistics and stop rule. */
int og_launch_multi_iter(const og_multi_dev *m, const og_multi_group *groups, uint32_t n_groups,
                         uint32_t it, uint32_t grid, hipStream_t s);
int og_launch_push(const og_pr_dev *a, int pmode, uint32_t it, hipStream_t s, hipEvent_t ev0,
                   hipEvent_t ev1);
/* Destination rows per range of the partitioned push (its LDS accumulators). */
#define OG_PP_ROWS 16384u
int og_launch_push_dloc(const uint32_t *row_ptr, const uint32_t *pp_v0, uint32_t n_pp, uint16_t *dloc,
                        hipStream_t s);
/* Only the scatter of a push (next[dst] += g_in[src] over the forward
 * edges), without the epilogue: SpMV's push accumulates into next. */
int og_launch_push_scatter(const og_pr_dev *a, int pmode, hipStream_t s, hipEvent_t ev0,
                           hipEvent_t ev1);
/* SpMV (og_spmv.hip): the per-vertex terms w * x of -p code 0..3 in new ids
 * into `terms` (fp32, u32 or u64), and vector_output in old ids from r
 * (codes 0, 1) or the u64 accumulators in next (codes 2, 3). */
int og_launch_spmv_terms(const og_pr_dev *a, uint32_t code, void *terms, hipStream_t s);
int og_launch_spmv_out(const og_pr_dev *a, uint32_t code, float *out_old, hipStream_t s);
int og_launch_quantize(const og_pr_dev *a, int mode, hipStream_t s);
int og_launch_init(const og_pr_dev *a, int mode, const float *init_old, float init_value,
                   hipStream_t s);
int og_launch_final(const og_pr_dev *a, uint32_t it, hipStream_t s);
int og_launch_normalize(const og_pr_dev *a, float *d_out_old, double *d_sum, hipStream_t s);
int og_launch_fill_u32(uint32_t *p, uint32_t v, uint64_t n, hipStream_t s);
/* data-driven variants */
int og_launch_dd_contrib(const og_pr_dev *a, hipStream_t s);
int og_launch_dd_activate(const og_pr_dev *a, hipStream_t s);
int og_launch_dd_resid_init(const og_pr_dev *a, hipStream_t s);
int og_launch_dd_push(const og_pr_dev *a, uint32_t code, uint32_t it, hipStream_t s);

#ifdef __cplusplus
}
#endif

#endif
