/*
 * og_pr_run.h -- the state of one PageRank run on the GPU (ogpu_pr), shared
 * by its host orchestration (og_pagerank.c) and the SpMV entry points
 * (og_spmv.c), which drive the same pull and push kernels.
 */
#ifndef OG_PR_RUN_H
#define OG_PR_RUN_H

#include "og_internal.h"
#include "og_pr.h"
#include "opengraph_amd_ext.h"

#ifdef __cplusplus
extern "C" {
#endif

struct ogpu_pr
{
    og_graph *g;
    og_plan *plan;
    int own_plan;
    ogpu_pr_options opt;
    int mode, push, quant, dd;
    og_pr_dev dev;
    uint32_t gv;              /* grid of the per-vertex kernels */
    uint32_t parts_cap;       /* entries of the first-level partial arrays */
    uint32_t n_final_pull, n_final_vert;
    float *r;
    void *c[2];
    void *cq;
    void *next;
    void *chunk_sum;
    float *init_d;
    float *out_old;
    double *d_sum;
    double *part_err, *fin_err;
    uint32_t *part_act, *fin_act;
    float *part_min, *part_max, *fin_min, *fin_max;
    uint32_t *ticket, *done;
    og_iter_stat *stats;
    uint32_t stats_cap;
    og_qparams *qp;
    uint32_t *done_host;
    uint32_t it;
    /* per-iteration end events of a run, kernel timing pairs of enqueue */
    hipEvent_t *iter_ev;
    uint32_t n_iter_ev;
    hipEvent_t *kev;
    uint32_t n_kev, kev_used;
    double seconds_plan;
    uint32_t rows;
    og_iter_stat *rows_host;
    /* data-driven state: two byte worklists, changed flags, residuals */
    uint8_t *wl[2];
    uint8_t *changed;
    float *resid;
    /* learned start: features (N x 4) and reduction scratch */
    float *feat;
    double *lscratch;
    /* every-row-exact fp32 pull: rows [0, C) longest first (device) */
    uint32_t *chain_rows;
    /* binned fp32 pull (og_bin.h): layout, grids of its passes, blocks of
     * its final reduction */
    struct og_bin_plan *bins;
    uint32_t bin_sum_grid, bin_long_grid, n_final_bin;
    /* partitioned push (og_push_part_kernel): destination ranges (device
     * n_pp + 1 row bounds), each inverse-CSR edge's offset in its range,
     * the first range pushed, blocks of its final reduction */
    uint32_t *pp_v0;
    uint16_t *pp_dloc;
    uint32_t n_pp, pp_first, n_final_pp;
    /* the last start vector offered (ogpu_pr_set_init, the learned start)
     * was rejected and the runs start from the uniform vector */
    uint32_t init_fallback;
    /* drop-in runs print the reference's per-iteration table rows as the
     * iterations complete: rows [0, printed) are out, print_done after the
     * stopping row */
    int print_rows, print_done;
    uint32_t printed;
};

/* Create a run of opt->variant on graph g; plan may be supplied (borrowed)
 * or is built here.  Buffers are allocated, the state is not initialized. */
int og_pr_create_with(og_graph *g, og_plan *plan, const ogpu_pr_options *opt, ogpu_pr **out);
/* Grow an event array to `need` created events. */
int og_pr_events(hipEvent_t **arr, uint32_t *have, uint32_t need);
/* The drop-ins' device copy of a caller's GraphCSR and its default plan,
 * cached by pointer and array identity (forward edges when need_forward). */
struct GraphCSR;
int og_cache_graph(struct GraphCSR *graph, int need_forward, og_graph **g, og_plan **plan);
/* An SpMV run on graph g over a borrowed plan (NULL: built here). */
int og_spmv_create_with(og_graph *g, og_plan *plan, const ogpu_spmv_options *opt, ogpu_spmv **out);

#ifdef __cplusplus
}
#endif

#endif
