/*
 * og_pagerank.c -- host orchestration of PageRank on the GPU and the
 * drop-in pageRank*GraphCSR entry points (pageRank.c:47-2896).
 *
 * A run is: init kernel (ranks 0.15f or the caller's start, first
 * contributions), then per iteration the variant's kernels (og_pr.hip),
 * enqueued back to back without host round trips; every 8 iterations the
 * host reads the device stop flag.  Kernels of iterations after the stop
 * return at once, so the result is that of the reference's loop
 * (pageRank.c:984-1035).  Finally ranks are divided by N and mapped back to
 * the caller's vertex ids (pageRank.c:1038-1044).
 */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

#include "og_bin.h"
#include "og_internal.h"
#include "og_learn.h"
#include "og_pr.h"
#include "og_pr_run.h"
#include "opengraph_amd_ext.h"

#define OG_CHECK_EVERY 8u
/* Rows with more in-edges than this go to the chunk waves: 12..16 is the
 * fastest on R-MAT 24 (1.38-1.40 ms against 1.47 at 32 and 1.50 at 4) and
 * on Erdos-Renyi 24 (4.80-4.87 ms against 5.08 at 32). */
#define OG_DEFAULT_SEQ_T 16u
/* fp32 pull: rows of more in-edges are summed in the reference's order by
 * default (exact_above) */
#define OG_DEFAULT_EXACT_ABOVE 4096u
/* fp32 push: destinations with more in-edges are gathered in CSR order
 * (when exact_above is not 0) */
#define OG_PUSH_EXACT_ABOVE 1024u
/* every-row-exact fp32 pull: rows of at most this many in-edges are summed
 * one per thread, the longer ones as wavefront chains */
#ifndef OG_EXACT_SEQ_T
#define OG_EXACT_SEQ_T 16u
#endif

/* Blocks of a final reduction over n_parts first-level partials that also
 * finalizes `rows` long rows (a wavefront each): enough that each block
 * folds about 256 partials, at most 1024 so that the last block's pass
 * over the block results stays short. */
static uint32_t og_final_blocks(uint32_t n_parts, uint32_t rows)
{
    uint64_t b = ((uint64_t)n_parts + 255) / 256;
    uint64_t r = ((uint64_t)rows + 3) / 4;
    if (r > b) b = r;
    if (b > 1024) b = 1024;
    if (b < 1) b = 1;
    return (uint32_t)b;
}

static int og_dd_run(ogpu_pr *p, ogpu_pr_result *res);

static double og_now(void)
{
    struct timespec ts;
    clock_gettime(CLOCK_MONOTONIC, &ts);
    return ts.tv_sec + ts.tv_nsec * 1e-9;
}

void ogpu_pr_default_options(ogpu_pr_options *o)
{
    memset(o, 0, sizeof(*o));
    o->variant = 0;
    o->max_iter = 20;          /* open-graph.c:278 */
    o->epsilon = 0.0001;       /* open-graph.c:280 */
    o->relabel = 1;
    o->hub_threshold = 0;
    o->exact_above = OG_DEFAULT_EXACT_ABOVE;
}

static int og_variant_mode(uint32_t v, int *mode, int *push, int *quant, int *dd)
{
    *push = *quant = *dd = 0;
    switch (v)
    {
    case 0: *mode = OG_MODE_F32; break;
    case 1: *mode = OG_MODE_F32; *push = 1; break;
    case 2: *mode = OG_MODE_FX64; break;
    case 3: *mode = OG_MODE_FX64; *push = 1; break;
    case 4: *mode = OG_MODE_Q32; *quant = 1; break;
    case 5: *mode = OG_MODE_QPUSH_EPI; *push = 1; *quant = 1; break;
    case 6: case 7: case 8: *mode = OG_MODE_F32; *dd = 1; break;
    case 9: *mode = OG_MODE_FX32; break;
    case 10: *mode = OG_MODE_FX16; break;
    case 11: *mode = OG_MODE_FX8; break;
    case 12: *mode = OG_MODE_Q16; *quant = 1; break;
    case 13: *mode = OG_MODE_Q8; *quant = 1; break;
    default: *mode = OG_MODE_F32; break;   /* pageRank.c:929-931: default is pull */
    }
    return 0;
}

static size_t og_elem_bytes(int mode)
{
    switch (mode)
    {
    case OG_MODE_FX64: return 8;
    case OG_MODE_FX16: case OG_MODE_Q16: return 2;
    case OG_MODE_FX8: case OG_MODE_Q8: return 1;
    default: return 4;
    }
}

void ogpu_pr_free(ogpu_pr *p)
{
    if (!p) return;
    if (p->own_plan) og_plan_destroy(p->plan);
    og_dfree(p->r); og_dfree(p->c[0]); og_dfree(p->c[1]); og_dfree(p->cq); og_dfree(p->next);
    og_dfree(p->chunk_sum); og_dfree(p->init_d); og_dfree(p->out_old); og_dfree(p->d_sum);
    og_dfree(p->part_err); og_dfree(p->fin_err); og_dfree(p->part_act); og_dfree(p->fin_act);
    og_dfree(p->part_min); og_dfree(p->part_max); og_dfree(p->fin_min); og_dfree(p->fin_max);
    og_dfree(p->ticket); og_dfree(p->done); og_dfree(p->stats); og_dfree(p->qp);
    if (p->done_host) (void)hipHostFree(p->done_host);
    for (uint32_t i = 0; i < p->n_iter_ev; i++) (void)hipEventDestroy(p->iter_ev[i]);
    for (uint32_t i = 0; i < p->n_kev; i++) (void)hipEventDestroy(p->kev[i]);
    free(p->iter_ev);
    free(p->kev);
    free(p->rows_host);
    og_dfree(p->wl[0]); og_dfree(p->wl[1]); og_dfree(p->changed); og_dfree(p->resid);
    og_dfree(p->feat); og_dfree(p->lscratch);
    og_dfree(p->chain_rows);
    og_bin_plan_destroy(p->bins);
    og_dfree(p->pp_v0); og_dfree(p->pp_dloc);
    free(p);
}

/* Buffers the learned start (og_learn.c) borrows from a fp32 pull run: its
 * kernel arguments and plan, c[1] as scratch, the start-vector buffer and
 * the feature array (allocated on first use). */
int og_pr_learn_parts(ogpu_pr *p, const og_pr_dev **dev, const og_plan **plan, uint32_t *n_final,
                      float **c1, float **init_d, float **feat, double **scratch)
{
    if (p->opt.warm_classes > OG_WC_IN_MAX)
    {
        og_set_error("learned start: warm_classes %u (0, 1 or 2..%u)", p->opt.warm_classes, OG_WC_IN_MAX);
        return -1;
    }
    if (p->mode != OG_MODE_F32 || p->push || p->dd)
    {
        og_set_error("learned start: fp32 pull runs (variant 0) only");
        return -1;
    }
    const size_t N = p->g->N;
    if (!p->init_d && !(p->init_d = og_dalloc(N * 4 + 4))) return -1;
    if (!p->feat && !(p->feat = og_dalloc(N * 16 + 16))) return -1;
    if (!p->lscratch && !(p->lscratch = og_dalloc(OG_WARM_SCRATCH * 8))) return -1;
    *dev = &p->dev;
    *plan = p->plan;
    *n_final = p->n_final_pull;
    *c1 = (float *)p->c[1];
    *init_d = p->init_d;
    *feat = p->feat;
    *scratch = p->lscratch;
    return 0;
}

/* The graph, two N-float buffers free until the next run's init (r, c[0])
 * and the number of in-degree classes of a learned start's correction
 * (1: none). */
int og_pr_learn_graph(ogpu_pr *p, og_graph **g, float **tmp_a, float **tmp_b, uint32_t *classes)
{
    *g = p->g;
    *tmp_a = p->r;
    *tmp_b = (float *)p->c[0];
    *classes = p->opt.warm_classes ? p->opt.warm_classes : OG_WC_IN_MAX;
    return 0;
}

/* The next runs start from the vector in init_d (accepted) or, when the
 * start offered was rejected, from the uniform start. */
void og_pr_use_init(ogpu_pr *p, int accepted)
{
    p->opt.init_mode = accepted ? 2 : 1;
    p->init_fallback = !accepted;
}

int ogpu_pr_init_fallback(const ogpu_pr *p)
{
    return p ? (int)p->init_fallback : -1;
}

int ogpu_pr_get_init(ogpu_pr *p, float *out)
{
    const uint32_t N = p->g->N;
    if (p->opt.init_mode == 2 && p->init_d)
        return og_hip_check(og_copy(out, p->init_d, (size_t)N * 4, hipMemcpyDeviceToHost), "D2H start");
    const float x = p->opt.init_mode == 1 ? 1.0f : OG_BASE_PR;
    for (uint32_t v = 0; v < N; v++) out[v] = x;
    return 0;
}

/* The rows [0, C) of a plan ordered by in-degree, longest first (stable):
 * the order the every-row-exact pull starts its wavefront chains in, so the
 * longest one, which bounds the launch, starts at once. */
static int og_pr_chain_rows(const og_plan *pl, uint32_t **out)
{
    *out = NULL;
    const uint32_t C = pl->C;
    uint32_t *rp = (uint32_t *)malloc((size_t)(C + 1) * 4), *rows = (uint32_t *)malloc((size_t)(C ? C : 1) * 4);
    uint32_t *d = og_dalloc((size_t)(C ? C : 1) * 4);
    if (!rp || !rows || !d)
    {
        if (!rp || !rows) og_set_error("out of host memory");
        free(rp); free(rows); og_dfree(d);
        return -1;
    }
    int rc = -1;
    if (!og_hip_check(og_copy(rp, pl->row_ptr, (size_t)(C + 1) * 4, hipMemcpyDeviceToHost), "D2H row_ptr"))
    {
        /* counting sort by length, longest first; ties keep row order */
        uint32_t maxlen = 0;
        for (uint32_t r = 0; r < C; r++) if (rp[r + 1] - rp[r] > maxlen) maxlen = rp[r + 1] - rp[r];
        uint32_t *cnt = (uint32_t *)calloc((size_t)maxlen + 2, 4);
        if (cnt)
        {
            for (uint32_t r = 0; r < C; r++) cnt[maxlen - (rp[r + 1] - rp[r])]++;
            uint32_t acc = 0;
            for (uint32_t k = 0; k <= maxlen; k++) { uint32_t t = cnt[k]; cnt[k] = acc; acc += t; }
            for (uint32_t r = 0; r < C; r++) rows[cnt[maxlen - (rp[r + 1] - rp[r])]++] = r;
            free(cnt);
            rc = og_hip_check(og_copy(d, rows, (size_t)C * 4, hipMemcpyHostToDevice), "H2D chain rows");
        }
        else og_set_error("out of host memory");
    }
    free(rp);
    free(rows);
    if (rc) { og_dfree(d); return -1; }
    *out = d;
    return 0;
}

/* Destination ranges of the partitioned push: consecutive rows of the plan,
 * at most OG_PP_ROWS (its LDS accumulators) and about E / (8 CUs) in-edges
 * each (a single row may have more), with a boundary at row first (the
 * chains before it are gathered instead); each inverse-CSR edge's offset in
 * its range on the device. */
static int og_push_parts(ogpu_pr *p, uint32_t first)
{
    const og_plan *pl = p->plan;
    const uint32_t N = pl->N, E = pl->E;
    uint32_t *rp = (uint32_t *)malloc(((size_t)N + 1) * 4);
    uint32_t *v0 = (uint32_t *)malloc(((size_t)N + 2) * 4);   /* at most N ranges */
    if (!rp || !v0)
    {
        free(rp); free(v0);
        og_set_error("out of host memory");
        return -1;
    }
    int rc = -1;
    if (og_hip_check(og_copy(rp, pl->row_ptr, ((size_t)N + 1) * 4, hipMemcpyDeviceToHost), "D2H row_ptr")) goto out;
    {
        const uint32_t cus = og_num_cus() > 0 ? (uint32_t)og_num_cus() : 256u;
        uint64_t cap = (uint64_t)E / (8ull * cus);
        if (cap < 4096) cap = 4096;
        uint32_t n = 0, v = 0, pf = 0;
        v0[n++] = 0;
        while (v < N)
        {
            const uint32_t start = v;
            uint64_t e = 0;
            if (v == first) pf = n - 1;
            while (v < N && v - start < OG_PP_ROWS)
            {
                if (v == first && v != start) break;
                const uint32_t d = rp[v + 1] - rp[v];
                if (v != start && e + d > cap) break;
                e += d;
                v++;
            }
            v0[n++] = v;
        }
        if (first >= N) pf = n - 1;   /* every row a chain: nothing pushed */
        p->n_pp = n - 1;
        p->pp_first = pf;
        p->pp_v0 = og_dalloc((size_t)n * 4);
        p->pp_dloc = og_dalloc((size_t)E * 2 + 2);
        if (!p->pp_v0 || !p->pp_dloc) goto out;
        hipStream_t s = og_stream();
        if (og_hip_check(hipMemcpyAsync(p->pp_v0, v0, (size_t)n * 4, hipMemcpyHostToDevice, s), "H2D") ||
            og_launch_push_dloc(pl->row_ptr, p->pp_v0, p->n_pp, p->pp_dloc, s) ||
            og_hip_check(hipStreamSynchronize(s), "sync push ranges"))
            goto out;
    }
    rc = 0;
out:
    free(rp);
    free(v0);
    return rc;
}

/* Create a run on graph g.  plan may be supplied (borrowed) or built here. */
int og_pr_create_with(og_graph *g, og_plan *plan, const ogpu_pr_options *opt, ogpu_pr **out)
{
    *out = NULL;
    if (og_ctx_ensure()) return -1;
    if (opt->variant > 13)
    {
        og_set_error("unknown PageRank variant %u", opt->variant);
        return -1;
    }
    if (opt->ppr && (opt->variant != 0 || opt->ppr_source >= g->N))
    {
        og_set_error("personalized PageRank needs variant 0 and a source below %u (got variant %u, "
                     "source %u)", g->N, opt->variant, opt->ppr_source);
        return -1;
    }
    ogpu_pr *p = (ogpu_pr *)calloc(1, sizeof(ogpu_pr));
    if (!p) { og_set_error("out of host memory"); return -1; }
    p->g = g;
    p->opt = *opt;
    if (p->opt.hub_threshold == 0) p->opt.hub_threshold = OG_DEFAULT_SEQ_T;
    og_variant_mode(opt->variant, &p->mode, &p->push, &p->quant, &p->dd);
    const uint32_t N = g->N, E = g->E;
    /* the binned fp32 pull (og_bin.h) on a plan that keeps every row's
     * in-neighbours in CSR order; a borrowed plan that sorts long rows by
     * source rules it out */
    int binned = opt->variant == 0 && !opt->ppr && opt->binned != 2 && g->E > 0 &&
                 !(plan && plan->relabel == 1 && plan->H);
    if (opt->binned == 1 && !binned)
    {
        og_set_error("binned pull: fp32 pull (variant 0, not personalized) on a graph with edges and a "
                     "plan in CSR order only");
        free(p);
        return -1;
    }
    /* fp32 pull with every row in the reference's order: the rows of more
     * than OG_EXACT_SEQ_T in-edges become wavefront chains over CSR-order
     * rows (relabel 1's sorted long rows are kept in CSR order instead) */
    const int chain = !binned && opt->variant == 0 && opt->hub_threshold >= 0x80000000u;
    /* otherwise, with exact_above, only the longest rows become chains; the
     * fp32 push gathers its destinations of more than OG_PUSH_EXACT_ABOVE
     * in-edges the same way (their atomic sums would stray furthest from the
     * reference's source-ordered adds) */
    /* (binned: every row of more than 512 in-edges counts as a chain, so the
     * plan orders all rows by in-degree, longest first) */
    const uint32_t chain_t = binned ? OG_CH
                           : !chain && opt->variant == 0 && !opt->ppr ? opt->exact_above
                           : opt->variant == 1 && opt->exact_above ? OG_PUSH_EXACT_ABOVE : 0;
    double t0 = og_now();
    if (plan)
    {
        p->plan = plan;
        if (chain && plan->relabel == 1 && plan->H)
        {
            og_set_error("every-row-exact pull needs a plan that keeps long rows in CSR order");
            free(p);
            return -1;
        }
    }
    else
    {
        const uint32_t rel = (chain || binned) && p->opt.relabel == 1 ? 2u : p->opt.relabel;
        if (og_plan_build(g, rel, chain ? OG_EXACT_SEQ_T : p->opt.hub_threshold, chain_t, p->push || p->dd,
                          &p->plan))
        {
            free(p);
            return -1;
        }
        p->own_plan = 1;
    }
    if (chain && og_pr_chain_rows(p->plan, &p->chain_rows))
    {
        ogpu_pr_free(p);
        return -1;
    }
    if (binned && og_bin_plan_build(p->plan, &p->bins))
    {
        if (opt->binned == 1)
        {
            ogpu_pr_free(p);
            return -1;
        }
        p->bins = NULL;   /* default: the gather pull instead */
    }
    if (p->bins)
    {
        p->bin_sum_grid = og_bin_sum_grid(p->bins);
        p->bin_long_grid = og_bin_long_grid(p->bins);
        p->n_final_bin = og_final_blocks(p->bin_sum_grid + p->bin_long_grid, 0);
    }
    p->seconds_plan = og_now() - t0;
    og_plan *pl = p->plan;
    if ((p->push || p->dd) && !pl->push_src)
    {
        og_set_error("plan has no forward edges for variant %u", opt->variant);
        ogpu_pr_free(p);
        return -1;
    }
    p->gv = (N + 255) / 256;
    if (p->gv > 2048) p->gv = 2048;
    if (p->gv == 0) p->gv = 1;
    /* pull as the persistent kernel with an LDS copy of the hottest
     * contributions: by default when the graph is several times larger
     * than that copy */
    uint32_t hot_grid = 0;
    /* (the every-row-exact pull stages its chains in LDS instead; the
     * chains of exact_above run beside it, in a launch of their own) */
    if (!p->push && !p->dd && opt->hot_cache != 2 && !p->chain_rows)
    {
        int on = opt->hot_cache == 1;
        const uint32_t nh = (uint32_t)(OG_HOT_BYTES / og_elem_bytes(p->mode));
        if (!on && (uint64_t)N * og_elem_bytes(p->mode) > 4ull * OG_HOT_BYTES && E)
        {
            /* worth it when the cached sources send at least a fifth of the edges */
            uint32_t *deg = (uint32_t *)malloc((size_t)nh * 4);
            if (!deg) { og_set_error("out of host memory"); ogpu_pr_free(p); return -1; }
            if (og_hip_check(og_copy(deg, pl->outdeg, (size_t)nh * 4, hipMemcpyDeviceToHost), "D2H"))
            {
                free(deg);
                ogpu_pr_free(p);
                return -1;
            }
            uint64_t hot_edges = 0;
            for (uint32_t v = 0; v < nh; v++) hot_edges += deg[v];
            free(deg);
            on = hot_edges * 5 >= (uint64_t)E;
        }
        if (on) hot_grid = (uint32_t)og_num_cus();
    }
    else if (p->push && !p->dd && opt->hot_cache == 1 && E)
    {
        /* the scatter over the forward edges with LDS accumulators for the
         * destinations with the most in-edges (the smallest ids), on
         * request: their block partials are more accurate than the
         * reference's running fp32 sums, which drop the small terms of a
         * long in-edge list, and can differ from them by a little over
         * 1e-5 relative on such rows (exact for the u64 push) */
        hot_grid = (uint32_t)og_num_cus();
    }
    /* chain blocks (four rows each) and chunk-wave blocks: every row exact,
     * chains [0, C) and no chunks; fp32 pull on a plan with chains [0, L),
     * their chunks skipped; otherwise every chunk */
    uint32_t n_chains = 0, ch_first = 0, chunk_blocks = pl->n_chunk_blocks;
    /* (a borrowed plan may carry chains the options did not ask for) */
    const int push_chains = opt->variant == 1 && opt->exact_above && pl->L;
    if (p->chain_rows)
    {
        n_chains = pl->C;
        chunk_blocks = 0;
    }
    else if (pl->L && opt->variant == 0 && !opt->ppr && !binned)
    {
        n_chains = pl->L;
        ch_first = pl->ch_first;
    }
    else if (push_chains)
        n_chains = pl->L;
    else if (pl->L)
        chunk_blocks = ((pl->n_chunks + OG_CH_PER_WAVE - 1) / OG_CH_PER_WAVE + 3) / 4;
    /* four chains per 256-thread pull block; beside the persistent kernel
     * (or the push scatter), one wavefront per CU taking every grid-th chain */
    const uint32_t cus = (uint32_t)og_num_cus();
    /* (the partitioned push leaves its range blocks room for eight chain
     * wavefronts per CU beside them; the scatter, one) */
    const uint32_t push_chain_cap = opt->hot_cache == 0 ? 8 * cus : cus;
    const uint32_t chain_blocks = push_chains ? (n_chains < push_chain_cap ? n_chains : push_chain_cap)
                                : hot_grid && !p->chain_rows ? (n_chains < hot_grid ? n_chains : hot_grid)
                                : (n_chains + 3) / 4;
    const uint32_t pull_blocks = hot_grid ? hot_grid + chain_blocks : chain_blocks + chunk_blocks + pl->n_row_blocks;
    p->n_final_pull = og_final_blocks(pull_blocks, pl->H);
    p->n_final_vert = og_final_blocks(p->gv + (push_chains ? chain_blocks : 0), 0);
    /* the default push (hot_cache 0): partitioned by destination range */
    if (p->push && !p->dd && opt->hot_cache == 0 && og_push_parts(p, push_chains ? n_chains : 0))
    {
        ogpu_pr_free(p);
        return -1;
    }
    const uint32_t pp_blocks = p->pp_v0 ? p->n_pp - p->pp_first + (push_chains ? chain_blocks : 0) : 0;
    p->n_final_pp = og_final_blocks(pp_blocks, 0);
    uint32_t parts = pull_blocks > p->gv ? pull_blocks : p->gv;
    if (parts < pp_blocks) parts = pp_blocks;
    if (push_chains && parts < p->gv + chain_blocks) parts = p->gv + chain_blocks;
    if (p->bins && parts < p->bin_sum_grid + p->bin_long_grid) parts = p->bin_sum_grid + p->bin_long_grid;
    p->parts_cap = parts;
    uint32_t fins = p->n_final_pull > p->n_final_vert ? p->n_final_pull : p->n_final_vert;
    if (fins < p->n_final_bin) fins = p->n_final_bin;
    if (fins < p->n_final_pp) fins = p->n_final_pp;
    p->stats_cap = (opt->max_iter ? opt->max_iter : 1) + 1;

    size_t eb = og_elem_bytes(p->mode);
    size_t cbytes = p->quant ? 4 : eb;
    p->r = og_dalloc((size_t)N * 4);
    p->c[0] = og_dalloc((size_t)N * cbytes);
    p->c[1] = og_dalloc((size_t)N * cbytes);
    if (p->quant) p->cq = og_dalloc((size_t)N * eb);
    if (p->push || p->dd) p->next = og_dalloc((size_t)N * 8);
    p->chunk_sum = og_dalloc((size_t)(pl->n_chunks + 1) * 8);
    p->out_old = og_dalloc((size_t)N * 4);
    p->d_sum = og_dalloc(8);
    p->part_err = og_dalloc((size_t)parts * 8);
    p->part_act = og_dalloc((size_t)parts * 4);
    p->part_min = og_dalloc((size_t)parts * 4);
    p->part_max = og_dalloc((size_t)parts * 4);
    p->fin_err = og_dalloc((size_t)fins * 8);
    p->fin_act = og_dalloc((size_t)fins * 4);
    p->fin_min = og_dalloc((size_t)fins * 4);
    p->fin_max = og_dalloc((size_t)fins * 4);
    p->ticket = og_dalloc(4);
    p->done = og_dalloc(4);
    p->stats = og_dalloc((size_t)p->stats_cap * sizeof(og_iter_stat));
    p->qp = og_dalloc(2 * sizeof(og_qparams));
    if (!p->r || !p->c[0] || !p->c[1] || (p->quant && !p->cq) || ((p->push || p->dd) && !p->next) ||
        !p->chunk_sum || !p->out_old || !p->d_sum || !p->part_err || !p->part_act || !p->part_min ||
        !p->part_max || !p->fin_err || !p->fin_act || !p->fin_min || !p->fin_max || !p->ticket ||
        !p->done || !p->stats || !p->qp)
    {
        ogpu_pr_free(p);
        return -1;
    }
    if (p->dd)
    {
        size_t wb = ((size_t)N + 3) & ~(size_t)3;
        p->wl[0] = og_dalloc(wb);
        p->wl[1] = og_dalloc(wb);
        p->changed = og_dalloc(wb);
        p->resid = og_dalloc((size_t)N * 4);
        if (!p->wl[0] || !p->wl[1] || !p->changed || !p->resid)
        {
            ogpu_pr_free(p);
            return -1;
        }
    }
    if (og_hip_check(hipHostMalloc((void **)&p->done_host, 4, hipHostMallocDefault), "hipHostMalloc"))
    {
        ogpu_pr_free(p);
        return -1;
    }

    og_pr_dev *a = &p->dev;
    memset(a, 0, sizeof(*a));
    a->N = N;
    a->E = E;
    a->H = pl->H;
    a->C = pl->C;
    a->Z = pl->Z;
    a->row_end = N;
    a->n_chunks = pl->n_chunks;
    a->n_chunk_blocks = chunk_blocks;
    a->n_chains = n_chains;
    a->n_chain_blocks = chain_blocks;
    a->ch_first = ch_first;
    if (p->chain_rows)
    {
        /* rows [0, C) are chains, four per 256-thread block; no chunk sums */
        a->chain_rows = p->chain_rows;
        a->H = 0;
        a->n_chunks = 0;
    }
    a->n_row_blocks = pl->n_row_blocks;
    a->push_skip = push_chains ? n_chains : 0;
    a->hot_grid = hot_grid;
    a->stop_l1 = opt->stop_rule == 1 || opt->ppr;
    a->ppr = opt->ppr ? 1 : 0;
    a->ppr_src = 0;
    a->qmode = p->quant ? (uint32_t)p->mode : 0;
    a->qzero = p->mode == OG_MODE_QPUSH_EPI;
    a->allow_stop = 1;
    a->epsilon = opt->epsilon;
    a->row_ptr = pl->row_ptr;
    a->cols = pl->cols;
    a->outdeg = pl->outdeg;
    a->chunk_base = pl->chunk_base;
    a->ch_row = pl->ch_row;
    a->ch_edge = pl->ch_edge;
    a->new2old = pl->new2old;
    a->old2new = pl->old2new;
    a->push_src = pl->push_src;
    a->push_dst = pl->push_dst;
    a->pp_v0 = p->pp_v0;
    a->pp_dloc = p->pp_dloc;
    a->n_pp = p->n_pp;
    a->pp_first = p->pp_first;
    a->r = p->r;
    a->next = p->next;
    a->chunk_sum = p->chunk_sum;
    a->part_err = p->part_err;
    a->part_act = p->part_act;
    a->part_min = p->part_min;
    a->part_max = p->part_max;
    a->fin_err = p->fin_err;
    a->fin_act = p->fin_act;
    a->fin_min = p->fin_min;
    a->fin_max = p->fin_max;
    a->ticket = p->ticket;
    a->done = p->done;
    a->stats = p->stats;
    a->qp = p->qp;
    a->qp_next = p->qp + 1;
    a->dd = p->dd ? opt->variant : 0;
    a->f_row_ptr = pl->f_row_ptr;
    a->f_cols = pl->f_cols;
    a->changed = p->changed;
    a->resid = p->resid;
    if (opt->ppr &&
        (og_hip_check(hipMemcpyAsync(&a->ppr_src, pl->old2new + opt->ppr_source, 4,
                                     hipMemcpyDeviceToHost, og_stream()), "D2H source") ||
         og_hip_check(hipStreamSynchronize(og_stream()), "sync source")))
    {
        ogpu_pr_free(p);
        return -1;
    }
    if (og_launch_fill_u32(p->ticket, 0, 1, og_stream())) { ogpu_pr_free(p); return -1; }
    *out = p;
    return 0;
}

int ogpu_pr_create(ogpu_graph *g, const ogpu_pr_options *opt, ogpu_pr **out)
{
    ogpu_pr_options o;
    if (opt) o = *opt; else ogpu_pr_default_options(&o);
    return og_pr_create_with((og_graph *)g, NULL, &o, out);
}

int ogpu_pr_set_init(ogpu_pr *p, const float *init)
{
    const uint32_t N = p->g->N;
    /* a start the iteration can use: finite, non-negative, some mass */
    double sum = 0.0;
    int ok = 1;
    for (uint32_t v = 0; v < N && ok; v++)
    {
        const float x = init[v];
        if (!(x >= 0.0f && x <= 3.402823466e38f)) ok = 0;
        else sum += x;
    }
    if (!(sum > 0.0)) ok = 0;
    if (!ok)
    {
        og_pr_use_init(p, 0);
        return 0;
    }
    if (!p->init_d)
    {
        p->init_d = og_dalloc((size_t)N * 4);
        if (!p->init_d) return -1;
    }
    if (og_hip_check(hipMemcpyAsync(p->init_d, init, (size_t)N * 4, hipMemcpyHostToDevice, og_stream()),
                     "H2D init") ||
        og_hip_check(hipStreamSynchronize(og_stream()), "sync init"))
        return -1;
    og_pr_use_init(p, 1);
    return 0;
}

/* Initial ranks and contributions (and, for the quantized variants, the
 * first quantization parameters and quantized contributions). */
static int og_pr_init_state(ogpu_pr *p)
{
    hipStream_t s = og_stream();
    p->it = 0;
    p->kev_used = 0;
    if (og_launch_fill_u32(p->done, 0, 1, s) || og_launch_fill_u32(p->ticket, 0, 1, s)) return -1;
    if (p->next && og_launch_fill_u32((uint32_t *)p->next, 0, (uint64_t)p->g->N * 2, s)) return -1;
    og_pr_dev a = p->dev;
    a.c_out = p->c[0];
    a.n_parts = p->gv;
    const float *init_old = NULL;
    float init_value = OG_BASE_PR;
    if (p->opt.init_mode == 1) init_value = 1.0f;
    else if (p->opt.init_mode == 2)
    {
        if (!p->init_d) { og_set_error("init_mode 2 needs ogpu_pr_set_init"); return -1; }
        init_old = p->init_d;
    }
    if (og_launch_init(&a, p->mode, init_old, init_value, s)) return -1;
    if (p->quant)
    {
        a.n_final_blocks = p->n_final_vert;
        a.allow_stop = 0;
        a.qp_next = p->qp;
        if (og_launch_final(&a, OG_NO_STATS, s)) return -1;
        a.cq_src = p->c[0];
        a.cq_dst = p->cq;
        a.qp = p->qp;
        if (og_launch_quantize(&a, p->mode, s)) return -1;
    }
    return 0;
}

static int og_pr_launch_iter(ogpu_pr *p, uint32_t it, int allow_stop, hipEvent_t e0, hipEvent_t e1)
{
    hipStream_t s = og_stream();
    og_pr_dev a = p->dev;
    a.allow_stop = allow_stop;
    a.qp = p->qp + (it & 1);
    a.qp_next = p->qp + ((it + 1) & 1);
    a.g_in = p->quant ? p->cq : p->c[it & 1];
    a.c_out = p->quant ? p->c[0] : p->c[(it + 1) & 1];
    uint32_t row = it % p->stats_cap;
    int rc;
    if (p->bins)
    {
        /* binned pull: gather, sum (rows without in-edges only in the first
         * two iterations, as below), long rows, final reduction */
        og_bin_dev b = p->bins->d;
        if (it >= 2)
        {
            b.n_bins_live = b.n_bins - p->bins->n_zero_bins;
            a.z_active = 0.0 >= a.epsilon ? a.N - a.Z : 0;
        }
        a.n_parts = p->bin_sum_grid + p->bin_long_grid;
        a.n_final_blocks = p->n_final_bin;
        if (a.n_parts > p->parts_cap)
        {
            og_set_error("binned pull with %u partials, %u allocated", a.n_parts, p->parts_cap);
            return -1;
        }
        if (og_launch_bin_pull(&b, &a, row, p->bin_sum_grid, p->bin_long_grid, s, e0, e1) ||
            og_launch_final(&a, row, s))
            return -1;
        return 0;
    }
    if (!p->push && !p->quant && !p->dd && it >= 2)
    {
        /* Rows [Z, N) have no in-edges: from iteration 1 on their rank is
         * (1-d) + d*0 every time (|delta| = 0) and both contribution buffers
         * already hold (1-d)/outdeg, so the launch leaves them out and only
         * their activeVertices count (0 >= epsilon) is added. */
        a.row_end = a.Z;
        a.n_row_blocks = (uint32_t)(((uint64_t)(a.Z - a.C) + 255) / 256);
        a.z_active = 0.0 >= a.epsilon ? a.N - a.Z : 0;
    }
    if (!p->push)
    {
        a.n_parts = og_pull_parts(&a);
        a.n_final_blocks = p->n_final_pull;
        if (a.n_parts > p->parts_cap)
        {
            og_set_error("pull launch with %u partials, %u allocated", a.n_parts, p->parts_cap);
            return -1;
        }
        rc = og_launch_pull(&a, p->mode, row, s, e0, e1);
    }
    else if (a.pp_v0)
    {
        a.n_parts = a.n_pp - a.pp_first + (p->mode == OG_MODE_F32 && a.push_skip ? a.n_chain_blocks : 0);
        a.n_final_blocks = p->n_final_pp;
        rc = og_launch_push(&a, p->mode, row, s, e0, e1);
    }
    else
    {
        a.n_parts = p->gv + (a.push_skip ? a.n_chain_blocks : 0);
        a.n_final_blocks = p->n_final_vert;
        rc = og_launch_push(&a, p->mode, row, s, e0, e1);
    }
    if (rc) return -1;
    if (p->quant)
    {
        a.cq_src = p->c[0];
        a.cq_dst = p->cq;
        a.qp = p->qp + ((it + 1) & 1);
        if (og_launch_quantize(&a, p->mode, s)) return -1;
    }
    return 0;
}

int og_pr_events(hipEvent_t **arr, uint32_t *have, uint32_t need)
{
    if (*have >= need) return 0;
    hipEvent_t *n = (hipEvent_t *)realloc(*arr, need * sizeof(hipEvent_t));
    if (!n) { og_set_error("out of host memory"); return -1; }
    *arr = n;
    for (uint32_t i = *have; i < need; i++)
    {
        if (og_hip_check(hipEventCreate(&n[i]), "hipEventCreate")) return -1;
        *have = i + 1;
    }
    return 0;
}

static int og_pr_finish(ogpu_pr *p, double *pr_sum)
{
    og_pr_dev a = p->dev;
    a.n_parts = p->gv;
    if (og_launch_normalize(&a, p->out_old, p->d_sum, og_stream())) return -1;
    if (og_hip_check(hipMemcpyAsync(pr_sum, p->d_sum, 8, hipMemcpyDeviceToHost, og_stream()), "D2H sum"))
        return -1;
    return og_hip_check(hipStreamSynchronize(og_stream()), "sync finish");
}

static double og_pr_row_seconds(ogpu_pr *p, uint32_t i);

/* The stop rule applied to one statistics row. */
static int og_pr_row_stops(const ogpu_pr *p, const og_iter_stat *r)
{
    if (!p->dd && (p->opt.stop_rule == 1 || p->opt.ppr)) return r->error_total < p->opt.epsilon;
    return r->active == 0;
}

/* Drop-in runs: the reference's per-iteration rows (pageRank.c:1031 and
 * the same line of every variant) for the iterations [printed, upto) that
 * have completed, up to and including the one that stops the run; the
 * data-driven variants first print their "Init" row (pageRank.c:2468,
 * :2636, :2800). */
static int og_pr_print_rows(ogpu_pr *p, uint32_t upto)
{
    if (!p->print_rows || p->print_done || upto <= p->printed) return 0;
    hipStream_t s = og_stream();
    const uint32_t from = p->printed, n = upto - from;
    og_iter_stat *rows = (og_iter_stat *)malloc((size_t)n * sizeof(og_iter_stat));
    if (!rows) { og_set_error("out of host memory"); return -1; }
    if (og_hip_check(hipMemcpyAsync(rows, p->stats + from, (size_t)n * sizeof(og_iter_stat),
                                    hipMemcpyDeviceToHost, s), "D2H stats") ||
        og_hip_check(hipStreamSynchronize(s), "sync stats"))
    {
        free(rows);
        return -1;
    }
    if (p->dd && from == 0)
    {
        float ms = 0.0f;
        if (p->n_iter_ev > 1) (void)hipEventElapsedTime(&ms, p->iter_ev[0], p->iter_ev[1]);
        printf("| %-10s | %-8u | %-15.13lf | %-9f | \n", "Init", p->g->N, 0.0, ms * 1e-3);
    }
    for (uint32_t i = 0; i < n && !p->print_done; i++)
    {
        printf("| %-10u | %-8u | %-15.13lf | %-9f | \n", from + i, rows[i].active, rows[i].error_total,
               og_pr_row_seconds(p, from + i));
        if (og_pr_row_stops(p, &rows[i])) p->print_done = 1;
    }
    p->printed = upto;
    free(rows);
    return 0;
}

/* Init and iterations until the device stop flag or max_iter; returns the
 * iterations launched (or -1).  The host reads the flag every
 * OG_CHECK_EVERY iterations. */
static int64_t og_pr_solve(ogpu_pr *p)
{
    hipStream_t s = og_stream();
    const uint32_t max_iter = p->opt.max_iter;
    if (og_pr_events(&p->iter_ev, &p->n_iter_ev, max_iter + 1)) return -1;
    if (og_hip_check(hipEventRecord(p->iter_ev[0], s), "hipEventRecord")) return -1;
    if (og_pr_init_state(p)) return -1;
    uint32_t launched = 0;
    while (launched < max_iter)
    {
        uint32_t n = max_iter - launched < OG_CHECK_EVERY ? max_iter - launched : OG_CHECK_EVERY;
        for (uint32_t i = 0; i < n; i++)
        {
            if (og_pr_launch_iter(p, launched + i, 1, NULL, NULL)) return -1;
            if (og_hip_check(hipEventRecord(p->iter_ev[launched + i + 1], s), "hipEventRecord")) return -1;
        }
        launched += n;
        if (og_hip_check(hipMemcpyAsync(p->done_host, p->done, 4, hipMemcpyDeviceToHost, s), "D2H done") ||
            og_hip_check(hipStreamSynchronize(s), "sync iterations") || og_pr_print_rows(p, launched))
            return -1;
        if (*p->done_host) break;
    }
    return launched;
}

static int og_pr_collect(ogpu_pr *p, uint32_t launched, double pr_sum, ogpu_pr_result *res);

int ogpu_pr_run(ogpu_pr *p, ogpu_pr_result *res)
{
    memset(res, 0, sizeof(*res));
    res->seconds_plan = p->seconds_plan;
    if (p->dd) return og_dd_run(p, res);
    double t0 = og_now();
    int64_t launched = og_pr_solve(p);
    if (launched < 0) return -1;
    double pr_sum = 0.0;
    if (og_pr_finish(p, &pr_sum)) return -1;
    res->seconds_total = og_now() - t0;
    return og_pr_collect(p, (uint32_t)launched, pr_sum, res);
}

/* Statistics rows of the last solve: iterations by the stop rule, the
 * reference's iteration count, last error / active. */
static int og_pr_collect(ogpu_pr *p, uint32_t launched, double pr_sum, ogpu_pr_result *res)
{
    hipStream_t s = og_stream();
    const uint32_t max_iter = p->opt.max_iter;
    free(p->rows_host);
    p->rows_host = (og_iter_stat *)calloc(launched ? launched : 1, sizeof(og_iter_stat));
    if (launched &&
        (og_hip_check(hipMemcpyAsync(p->rows_host, p->stats, (size_t)launched * sizeof(og_iter_stat),
                                     hipMemcpyDeviceToHost, s), "D2H stats") ||
         og_hip_check(hipStreamSynchronize(s), "sync stats")))
        return -1;
    uint32_t rows = launched, iterations = max_iter;
    for (uint32_t i = 0; i < launched; i++)
    {
        if (og_pr_row_stops(p, &p->rows_host[i]))
        {
            rows = i + 1;
            iterations = i;
            break;
        }
    }
    p->rows = rows;
    res->iterations = iterations;
    res->rows = rows;
    res->active = rows ? p->rows_host[rows - 1].active : 0;
    res->error_total = rows ? p->rows_host[rows - 1].error_total : 0.0;
    res->pr_sum = pr_sum;
    return 0;
}

int ogpu_pr_rows(ogpu_pr *p, uint32_t *active, double *error_total, uint32_t max_rows)
{
    uint32_t n = p->rows < max_rows ? p->rows : max_rows;
    for (uint32_t i = 0; i < n; i++)
    {
        if (active) active[i] = p->rows_host[i].active;
        if (error_total) error_total[i] = p->rows_host[i].error_total;
    }
    return (int)n;
}

/* Seconds of iteration i (for the reference's per-iteration table). */
static double og_pr_row_seconds(ogpu_pr *p, uint32_t i)
{
    float ms = 0.0f;
    uint32_t k = p->dd ? i + 1 : i;    /* data-driven runs record an init event first */
    if (k + 1 >= p->n_iter_ev) return 0.0;
    if (hipEventElapsedTime(&ms, p->iter_ev[k], p->iter_ev[k + 1]) != hipSuccess) return 0.0;
    return ms * 1e-3;
}

int ogpu_pr_reset(ogpu_pr *p)
{
    if (og_pr_init_state(p)) return -1;
    return og_hip_check(hipStreamSynchronize(og_stream()), "sync reset");
}

int ogpu_pr_enqueue(ogpu_pr *p, uint32_t n, int time_kernels)
{
    if (p->dd) { og_set_error("ogpu_pr_enqueue: not available for data-driven variants"); return -1; }
    if (time_kernels && og_pr_events(&p->kev, &p->n_kev, p->kev_used + 2 * n)) return -1;
    for (uint32_t i = 0; i < n; i++)
    {
        hipEvent_t e0 = NULL, e1 = NULL;
        if (time_kernels)
        {
            e0 = p->kev[p->kev_used++];
            e1 = p->kev[p->kev_used++];
        }
        if (og_pr_launch_iter(p, p->it++, 0, e0, e1)) return -1;
    }
    return 0;
}

int ogpu_pr_kernel_ms(ogpu_pr *p, double *total_ms, uint32_t *count)
{
    if (og_hip_check(hipStreamSynchronize(og_stream()), "sync")) return -1;
    double t = 0.0;
    for (uint32_t i = 0; i + 1 < p->kev_used; i += 2)
    {
        float ms = 0.0f;
        if (og_hip_check(hipEventElapsedTime(&ms, p->kev[i], p->kev[i + 1]), "hipEventElapsedTime"))
            return -1;
        t += ms;
    }
    *total_ms = t;
    *count = p->kev_used / 2;
    p->kev_used = 0;
    return 0;
}

int ogpu_pr_ranks(ogpu_pr *p, float *host_out)
{
    if (og_hip_check(hipMemcpyAsync(host_out, p->out_old, (size_t)p->g->N * 4, hipMemcpyDeviceToHost,
                                    og_stream()), "D2H ranks"))
        return -1;
    return og_hip_check(hipStreamSynchronize(og_stream()), "sync ranks");
}

double ogpu_pr_bytes_per_iteration(const ogpu_pr *p)
{
    double N = p->g->N, E = p->g->E;
    switch (p->opt.variant)
    {
    /* the partitioned pushes: source id, u16 destination offset and the
     * gathered contribution per edge; rank read and written, out-degree and
     * the new contribution per vertex.  The scatters: index pair + atomic
     * read-modify-write per edge. */
    case 1: return p->pp_v0 ? 10.0 * E + 16.0 * N : 12.0 * E + 20.0 * N;
    case 2: return 12.0 * E + 24.0 * N;           /* 8-byte contributions */
    case 3: return p->pp_v0 ? 14.0 * E + 20.0 * N : 20.0 * E + 24.0 * N;
    case 9: case 4: return 8.0 * E + 20.0 * N;
    case 10: case 12: return 6.0 * E + 18.0 * N;
    case 11: case 13: return 5.0 * E + 17.0 * N;
    default: return 8.0 * E + 20.0 * N;           /* index + fp32 contribution per edge */
    }
}

/* Data-driven variants (pageRank.c:2405-2896).  Per iteration:
 *   6: contributions -> pull over the worklist (chunks + short rows, fused
 *      update) -> long-row finalize + reduction -> activation of the
 *      out-neighbours of moved vertices;
 *   7/8: one residual-push kernel over the worklist + reduction;
 * then the worklists swap and the old one is cleared
 * (pageRank.c:2520-2522). */
static int og_dd_run(ogpu_pr *p, ogpu_pr_result *res)
{
    hipStream_t s = og_stream();
    const uint32_t max_iter = p->opt.max_iter;
    const uint32_t code = p->opt.variant;
    const uint32_t N = p->g->N;
    const uint64_t wwords = ((uint64_t)N + 3) / 4;
    if (og_pr_events(&p->iter_ev, &p->n_iter_ev, max_iter + 2)) return -1;
    double t0 = og_now();
    if (og_hip_check(hipEventRecord(p->iter_ev[0], s), "hipEventRecord")) return -1;
    p->it = 0;
    if (og_launch_fill_u32(p->done, 0, 1, s) || og_launch_fill_u32(p->ticket, 0, 1, s)) return -1;
    /* pageRank.c:2455-2466 / :2612-2633: every vertex active, ranks 0.15 */
    if (og_launch_fill_u32((uint32_t *)p->wl[0], 0x01010101u, wwords, s) ||
        og_launch_fill_u32((uint32_t *)p->wl[1], 0, wwords, s))
        return -1;
    og_pr_dev a = p->dev;
    a.c_out = p->c[0];
    a.n_parts = p->gv;
    float init_value = p->opt.init_mode == 1 ? 1.0f : OG_BASE_PR;
    const float *init_old = p->opt.init_mode == 2 ? p->init_d : NULL;
    if (p->opt.init_mode == 2 && !init_old) { og_set_error("init_mode 2 needs ogpu_pr_set_init"); return -1; }
    if (og_launch_init(&a, OG_MODE_F32, init_old, init_value, s)) return -1;
    if (code != 6 && og_launch_dd_resid_init(&a, s)) return -1;
    if (og_hip_check(hipEventRecord(p->iter_ev[1], s), "hipEventRecord")) return -1;
    uint32_t launched = 0;
    while (launched < max_iter)
    {
        uint32_t n = max_iter - launched < OG_CHECK_EVERY ? max_iter - launched : OG_CHECK_EVERY;
        for (uint32_t i = 0; i < n; i++)
        {
            uint32_t it = launched + i;
            og_pr_dev b = p->dev;
            b.wl_cur = p->wl[it & 1];
            b.wl_next = p->wl[(it + 1) & 1];
            b.allow_stop = 1;
            if (code == 6)
            {
                b.c_out = p->c[0];
                b.g_in = p->c[0];
                if (og_launch_dd_contrib(&b, s)) return -1;
                b.n_parts = og_pull_parts(&b);
                b.n_final_blocks = p->n_final_pull;
                if (og_launch_pull(&b, OG_MODE_F32, it, s, NULL, NULL)) return -1;
                if (og_launch_dd_activate(&b, s)) return -1;
            }
            else
            {
                b.n_parts = p->gv;
                b.n_final_blocks = p->n_final_vert;
                if (og_launch_dd_push(&b, code, it, s)) return -1;
            }
            if (og_launch_fill_u32((uint32_t *)p->wl[it & 1], 0, wwords, s)) return -1;
            if (og_hip_check(hipEventRecord(p->iter_ev[it + 2], s), "hipEventRecord")) return -1;
        }
        launched += n;
        if (og_hip_check(hipMemcpyAsync(p->done_host, p->done, 4, hipMemcpyDeviceToHost, s), "D2H done") ||
            og_hip_check(hipStreamSynchronize(s), "sync iterations") || og_pr_print_rows(p, launched))
            return -1;
        if (*p->done_host) break;
    }
    double pr_sum = 0.0;
    if (og_pr_finish(p, &pr_sum)) return -1;
    double t1 = og_now();
    free(p->rows_host);
    p->rows_host = (og_iter_stat *)calloc(launched ? launched : 1, sizeof(og_iter_stat));
    if (launched &&
        (og_hip_check(hipMemcpyAsync(p->rows_host, p->stats, (size_t)launched * sizeof(og_iter_stat),
                                     hipMemcpyDeviceToHost, s), "D2H stats") ||
         og_hip_check(hipStreamSynchronize(s), "sync stats")))
        return -1;
    uint32_t rows = launched, iterations = max_iter;
    for (uint32_t i = 0; i < launched; i++)
        if (p->rows_host[i].active == 0)
        {
            rows = i + 1;
            iterations = i;
            break;
        }
    p->rows = rows;
    res->iterations = iterations;
    res->rows = rows;
    res->active = rows ? p->rows_host[rows - 1].active : 0;
    res->error_total = rows ? p->rows_host[rows - 1].error_total : 0.0;
    res->pr_sum = pr_sum;
    res->seconds_total = t1 - t0;
    return 0;
}

/* ===================================================================== */
/* drop-in entry points                                                  */
/* ===================================================================== */

/* One cached device copy of the caller's GraphCSR (keyed by the struct and
 * array pointers, sizes and the sort mode) and its plan. */
static struct
{
    const void *key[12];
    uint32_t N, E;
    og_graph *g;
    og_plan *plan;
} g_cache;

static void og_fatal(const char *where)
{
    fprintf(stderr, "opengraph_amd: %s: %s\n", where, og_get_error());
    exit(EXIT_FAILURE);
}

int og_cache_graph(struct GraphCSR *graph, int need_forward, og_graph **g, og_plan **plan)
{
    const void *key[12] = {graph,
                           graph->vertices ? graph->vertices->out_degree : NULL,
                           graph->vertices ? graph->vertices->in_degree : NULL,
                           graph->vertices ? graph->vertices->edges_idx : NULL,
                           graph->sorted_edges_array ? graph->sorted_edges_array->edges_array_src : NULL,
                           graph->sorted_edges_array ? graph->sorted_edges_array->edges_array_dest : NULL,
                           graph->inverse_vertices ? graph->inverse_vertices->out_degree : NULL,
                           graph->inverse_vertices ? graph->inverse_vertices->in_degree : NULL,
                           graph->inverse_vertices ? graph->inverse_vertices->edges_idx : NULL,
                           graph->inverse_sorted_edges_array ? graph->inverse_sorted_edges_array->edges_array_src : NULL,
                           graph->inverse_sorted_edges_array ? graph->inverse_sorted_edges_array->edges_array_dest : NULL,
                           NULL};
    if (g_cache.g && memcmp(g_cache.key, key, sizeof(key)) == 0 && g_cache.N == graph->num_vertices &&
        g_cache.E == graph->num_edges)
    {
        if (need_forward && !g_cache.plan->push_src)
        {
            og_plan_destroy(g_cache.plan);
            g_cache.plan = NULL;
            if (og_plan_build(g_cache.g, 1, OG_DEFAULT_SEQ_T, OG_DEFAULT_EXACT_ABOVE, 1, &g_cache.plan)) return -1;
        }
        *g = g_cache.g;
        *plan = g_cache.plan;
        return 0;
    }
    og_plan_destroy(g_cache.plan);
    og_graph_destroy(g_cache.g);
    memset(&g_cache, 0, sizeof(g_cache));
    if (!graph->vertices || !graph->inverse_vertices || !graph->sorted_edges_array ||
        !graph->inverse_sorted_edges_array)
    {
        og_set_error("GraphCSR without forward and inverse arrays (DIRECTED build expected)");
        return -1;
    }
    const uint32_t *arrays[10] = {
        graph->sorted_edges_array->edges_array_src, graph->sorted_edges_array->edges_array_dest,
        graph->vertices->out_degree, graph->vertices->in_degree, graph->vertices->edges_idx,
        graph->inverse_sorted_edges_array->edges_array_src,
        graph->inverse_sorted_edges_array->edges_array_dest, graph->inverse_vertices->out_degree,
        graph->inverse_vertices->in_degree, graph->inverse_vertices->edges_idx};
    ogpu_graph *gg = NULL;
    if (ogpu_graph_from_arrays(graph->num_vertices, graph->num_edges, arrays, &gg)) return -1;
    og_plan *pl = NULL;
    if (og_plan_build((og_graph *)gg, 1, OG_DEFAULT_SEQ_T, 0, need_forward, &pl))
    {
        ogpu_graph_free(gg);
        return -1;
    }
    memcpy(g_cache.key, key, sizeof(key));
    g_cache.N = graph->num_vertices;
    g_cache.E = graph->num_edges;
    g_cache.g = (og_graph *)gg;
    g_cache.plan = pl;
    *g = g_cache.g;
    *plan = pl;
    return 0;
}

struct PageRankStats *newPageRankStatsGraphCSR(struct GraphCSR *graph)
{
    struct PageRankStats *stats = (struct PageRankStats *)og_host_alloc(sizeof(struct PageRankStats));
    stats->damp = OG_DAMP;
    stats->base_pr = 1.0f - stats->damp;
    stats->iterations = 0;
    stats->num_vertices = graph->num_vertices;
    stats->time_total = 0.0;
    stats->error_total = 0.0;
    stats->activeVertices = 0;
    stats->realRanks = (uint32_t *)og_host_alloc((size_t)graph->num_vertices * sizeof(uint32_t));
    stats->pageRanks = (float *)og_host_alloc((size_t)graph->num_vertices * sizeof(float));
    for (uint32_t v = 0; v < stats->num_vertices; v++)
    {
        stats->pageRanks[v] = stats->base_pr;
        stats->realRanks[v] = v;
    }
    return stats;
}

void freePageRankStats(struct PageRankStats *stats)
{
    if (stats)
    {
        free(stats->realRanks);
        free(stats->pageRanks);
        free(stats);
    }
}

static const char *og_variant_title(uint32_t v)
{
    /* the titles each reference variant prints (pageRank.c:968-974, ...) */
    switch (v)
    {
    case 1: return "Starting Page Rank Push (tolerance/epsilon)";
    case 2: return "Starting Page Rank Pull FP_64 (tolerance/epsilon)";
    case 3: return "Starting Page Rank Push FP (tolerance/epsilon)";
    case 4: return "Starting Page Rank Pull Quant_32";
    case 5: return "Starting Page Rank Push Quant_32";
    case 6: return "Starting Page Rank Pull DD (tolerance/epsilon)";
    case 7: return "Starting Page Rank Push DD (tolerance/epsilon)";
    case 8: return "Starting Page Rank Pull-Push DD (tolerance/epsilon)";
    case 9: return "Starting Page Rank Pull FP_32 (tolerance/epsilon)";
    case 10: return "Starting Page Rank Pull FP_16 (tolerance/epsilon)";
    case 11: return "Starting Page Rank Pull FP_8 (tolerance/epsilon)";
    case 12: return "Starting Page Rank Pull Quant_16";
    case 13: return "Starting Page Rank Pull Quant_8";
    default: return "Starting Page Rank Pull (tolerance/epsilon)";
    }
}

static struct PageRankStats *og_dropin(struct Arguments *arguments, struct GraphCSR *graph, uint32_t variant)
{
    if (og_ctx_ensure()) og_fatal("device");
    ogpu_pr_options o;
    ogpu_pr_default_options(&o);
    o.variant = variant;
    o.max_iter = arguments->iterations;
    o.epsilon = arguments->epsilon;
    int mode, push, quant, dd;
    og_variant_mode(variant, &mode, &push, &quant, &dd);
    og_graph *g = NULL;
    og_plan *plan = NULL;
    if (og_cache_graph(graph, push || dd, &g, &plan)) og_fatal("graph upload");
    og_trace_entry("pageRankGraphCSR", g->N, g->E);
    ogpu_pr *pr = NULL;
    /* the fp32 pull runs binned, on a plan of its own that keeps every row
     * in CSR order (the cached plan sorts long rows for the gather pulls) */
    if (og_pr_create_with(g, variant == 0 ? NULL : plan, &o, &pr)) og_fatal("pagerank setup");

    struct PageRankStats *stats = newPageRankStatsGraphCSR(graph);
    const char *title = og_variant_title(variant);
    printf(" -----------------------------------------------------\n");
    if (quant)
        printf("| %-30s %-19s| \n", title, "(tolerance/epsilon)");
    else
        printf("| %-51s | \n", title);
    printf(" -----------------------------------------------------\n");
    printf("| %-51.13lf | \n", arguments->epsilon);
    printf(" -----------------------------------------------------\n");
    printf("| %-10s | %-8s | %-15s | %-9s | \n", "Iteration", "Active", "Error", "Time (S)");
    printf(" -----------------------------------------------------\n");

    /* the rows are printed as the iterations complete (read back with each
     * stop-flag check) */
    pr->print_rows = 1;
    pr->print_done = 0;
    pr->printed = 0;
    ogpu_pr_result res;
    if (ogpu_pr_run(pr, &res)) og_fatal("pagerank run");
    if (ogpu_pr_ranks(pr, stats->pageRanks)) og_fatal("pagerank ranks");
    stats->iterations = res.iterations;
    stats->activeVertices = res.active;
    stats->error_total = res.error_total;
    stats->time_total = res.seconds_total;
    printf(" -----------------------------------------------------\n");
    printf("| %-10s | %-8s | %-15s | %-9s | \n", "Iterations", "PR Sum", "Error", "Time (S)");
    printf(" -----------------------------------------------------\n");
    printf("| %-10u | %-8lf | %-15.13lf | %-9f | \n", stats->iterations, res.pr_sum, stats->error_total,
           stats->time_total);
    printf(" -----------------------------------------------------\n");
    fflush(stdout);
    ogpu_pr_free(pr);
    return stats;
}

struct PageRankStats *pageRankGraphCSR(struct Arguments *arguments, struct GraphCSR *graph)
{
    uint32_t v = arguments->pushpull;
    return og_dropin(arguments, graph, v <= 13 ? v : 0);
}

struct PageRankStats *pageRankPullGraphCSR(struct Arguments *a, struct GraphCSR *g) { return og_dropin(a, g, 0); }
struct PageRankStats *pageRankPushGraphCSR(struct Arguments *a, struct GraphCSR *g) { return og_dropin(a, g, 1); }
struct PageRankStats *pageRankPullFixedPoint64BitGraphCSR(struct Arguments *a, struct GraphCSR *g) { return og_dropin(a, g, 2); }
struct PageRankStats *pageRankPushFixedPointGraphCSR(struct Arguments *a, struct GraphCSR *g) { return og_dropin(a, g, 3); }
struct PageRankStats *pageRankPullQuant32BitGraphCSR(struct Arguments *a, struct GraphCSR *g) { return og_dropin(a, g, 4); }
struct PageRankStats *pageRankPushQuantGraphCSR(struct Arguments *a, struct GraphCSR *g) { return og_dropin(a, g, 5); }
struct PageRankStats *pageRankDataDrivenPullGraphCSR(struct Arguments *a, struct GraphCSR *g) { return og_dropin(a, g, 6); }
struct PageRankStats *pageRankDataDrivenPushGraphCSR(struct Arguments *a, struct GraphCSR *g) { return og_dropin(a, g, 7); }
struct PageRankStats *pageRankDataDrivenPullPushGraphCSR(struct Arguments *a, struct GraphCSR *g) { return og_dropin(a, g, 8); }
struct PageRankStats *pageRankPullFixedPoint32BitGraphCSR(struct Arguments *a, struct GraphCSR *g) { return og_dropin(a, g, 9); }
struct PageRankStats *pageRankPullFixedPoint16BitGraphCSR(struct Arguments *a, struct GraphCSR *g) { return og_dropin(a, g, 10); }
struct PageRankStats *pageRankPullFixedPoint8BitGraphCSR(struct Arguments *a, struct GraphCSR *g) { return og_dropin(a, g, 11); }
struct PageRankStats *pageRankPullQuant16BitGraphCSR(struct Arguments *a, struct GraphCSR *g) { return og_dropin(a, g, 12); }
struct PageRankStats *pageRankPullQuant8BitGraphCSR(struct Arguments *a, struct GraphCSR *g) { return og_dropin(a, g, 13); }

/* ------------------------------------------------------------------ */
/* autotuner of the row-class threshold                                */
/* ------------------------------------------------------------------ */

int ogpu_tune_hub_threshold(ogpu_graph *gg, const uint32_t *cands, uint32_t n_cands, uint32_t iters,
                            uint32_t *best, double *ms)
{
    og_graph *g = (og_graph *)gg;
    if (!g || !n_cands || !iters) { og_set_error("ogpu_tune_hub_threshold: bad arguments"); return -1; }
    double best_ms = 0.0;
    *best = cands[0];
    for (uint32_t i = 0; i < n_cands; i++)
    {
        ogpu_pr_options o;
        ogpu_pr_default_options(&o);
        o.max_iter = iters + 2;
        o.hub_threshold = cands[i] ? cands[i] : OG_DEFAULT_SEQ_T;
        o.binned = 2;   /* the threshold belongs to the gather pull */
        ogpu_pr *p = NULL;
        if (og_pr_create_with(g, NULL, &o, &p)) return -1;
        /* two warm-up iterations, then `iters` timed on the host clock
         * (pull, finalize and launch gaps: what a run pays) */
        double t = 0.0;
        int bad = ogpu_pr_reset(p) || ogpu_pr_enqueue(p, 2, 0) ||
                  og_hip_check(hipStreamSynchronize(og_stream()), "sync");
        if (!bad)
        {
            double t0 = og_now();
            bad = ogpu_pr_enqueue(p, iters, 0) || og_hip_check(hipStreamSynchronize(og_stream()), "sync");
            t = (og_now() - t0) * 1e3 / iters;
        }
        ogpu_pr_free(p);
        if (bad) return -1;
        if (ms) ms[i] = t;
        if (i == 0 || t < best_ms)
        {
            best_ms = t;
            *best = cands[i];
        }
    }
    return 0;
}

/* ------------------------------------------------------------------ */
/* many graphs                                                         */
/* ------------------------------------------------------------------ */

#define OG_MULTI_TUNE_ITERS 5u
#define OG_MULTI_WAVES 4u      /* wavefronts per block of the long-row kernel */
#define OG_MULTI_TILE_WARM 2u  /* untimed pulls before each tile shape's timed ones */
#define OG_MULTI_TILE_ITERS 6u /* timed pulls per tile shape */

struct ogpu_multi
{
    uint32_t n;
    ogpu_pr **pr;
    uint32_t *hub;              /* chosen threshold per graph */
    uint32_t *tile;             /* chosen tile shape per graph */
    og_multi_group *groups;     /* pull launches of an iteration, one per tile shape in use */
    uint32_t n_groups;
    og_multi_dev dev;
    og_pr_dev *d_devs;
    og_multi_work *d_work, *d_lwork, *d_iwork;
    uint32_t *d_wr, *d_lr;
    double *w_err, *l_err;
    uint32_t *w_act, *l_act, *d_live, *d_all_done;
    uint32_t grid, max_iter;
    int64_t *launched;          /* per graph, of the last run */
    hipGraphExec_t exec;
    int have_exec;
};

void ogpu_multi_free(ogpu_multi *m)
{
    if (!m) return;
    if (m->have_exec) (void)hipGraphExecDestroy(m->exec);
    for (uint32_t g = 0; m->pr && g < m->n; g++) ogpu_pr_free(m->pr[g]);
    free(m->pr); free(m->hub); free(m->tile); free(m->groups); free(m->launched);
    og_dfree(m->d_devs); og_dfree(m->d_work); og_dfree(m->d_lwork); og_dfree(m->d_iwork);
    og_dfree(m->d_wr); og_dfree(m->d_lr); og_dfree(m->w_err); og_dfree(m->l_err);
    og_dfree(m->w_act); og_dfree(m->l_act); og_dfree(m->d_live); og_dfree(m->d_all_done);
    free(m);
}

uint32_t ogpu_multi_hub_threshold(const ogpu_multi *m, uint32_t g)
{
    return g < m->n ? m->hub[g] : 0;
}

uint32_t ogpu_multi_tile_shapes(ogpu_multi_tile *out, uint32_t max)
{
    const uint32_t n = og_multi_tile_count();
    for (uint32_t i = 0; out && i < n && i < max; i++)
    {
        og_multi_tile t;
        (void)og_multi_tile_shape(i, &t);
        out[i] = (ogpu_multi_tile){t.block, t.chunks, t.row_batch, 0};
    }
    return n;
}

int ogpu_multi_tile_of(const ogpu_multi *m, uint32_t g, ogpu_multi_tile *t)
{
    if (!m || g >= m->n) { og_set_error("ogpu_multi_tile_of: graph %u of %u", g, m ? m->n : 0); return -1; }
    og_multi_tile s;
    (void)og_multi_tile_shape(m->tile[g], &s);
    *t = (ogpu_multi_tile){s.block, s.chunks, s.row_batch, 0};
    return 0;
}

/* The compiled shape with these parameters, or -1. */
static int og_multi_tile_id(uint32_t block, uint32_t chunks, uint32_t row_batch)
{
    for (uint32_t i = 0; i < og_multi_tile_count(); i++)
    {
        og_multi_tile t;
        (void)og_multi_tile_shape(i, &t);
        if (t.block == block && t.chunks == chunks && t.row_batch == row_batch) return (int)i;
    }
    return -1;
}

static uint32_t og_class_of(uint32_t N)
{
    uint32_t c = 0;
    while (c < 32 && (1u << c) < N) c++;
    return c;
}

/* Threshold per graph: given, or tuned once per size class (graphs whose
 * vertex counts share the same power of two) on the class's largest
 * graph. */
static int og_multi_pick(ogpu_multi *m, og_graph *const *graphs, const ogpu_pr_options *opt,
                         const uint32_t *hub)
{
    static const uint32_t cands[] = {8, 16, 32, 64, 128};
    int32_t cls_best[33];
    for (int i = 0; i < 33; i++) cls_best[i] = -1;
    for (uint32_t g = 0; g < m->n; g++)
    {
        uint32_t h = hub ? hub[g] : opt->hub_threshold;
        if (h) { m->hub[g] = h; continue; }
        const uint32_t cls = og_class_of(graphs[g]->N);
        if (cls_best[cls] < 0)
        {
            uint32_t big = g;
            for (uint32_t q = 0; q < m->n; q++)
                if (og_class_of(graphs[q]->N) == cls && graphs[q]->E > graphs[big]->E) big = q;
            uint32_t b = 0;
            if (ogpu_tune_hub_threshold((ogpu_graph *)graphs[big], cands, 5, OG_MULTI_TUNE_ITERS, &b, NULL))
                return -1;
            cls_best[cls] = (int32_t)b;
        }
        m->hub[g] = (uint32_t)cls_best[cls];
    }
    return 0;
}

/* Graph g's pull items for tile shape t (into out when not NULL): blocks of
 * t.block / 64 wavefronts of t.chunks chunks each over the chunks
 * [ch_first, n_chunks), then blocks of t.block rows over [C, N). */
static size_t og_multi_items(const og_pr_dev *a, uint32_t g, const og_multi_tile *t, og_multi_work *out)
{
    const uint32_t nw = t->block / 64;
    const uint32_t waves = a->n_chunks > a->ch_first ? (a->n_chunks - a->ch_first + t->chunks - 1) / t->chunks : 0;
    const uint32_t cb = (waves + nw - 1) / nw;
    const uint32_t rb = a->N > a->C ? (uint32_t)(((uint64_t)(a->N - a->C) + t->block - 1) / t->block) : 0;
    if (out)
    {
        size_t k = 0;
        for (uint32_t b = 0; b < cb; b++) out[k++] = (og_multi_work){g, 0, b, 0};
        for (uint32_t b = 0; b < rb; b++) out[k++] = (og_multi_work){g, 1, b, 0};
    }
    return (size_t)cb + rb;
}

/* Blocks of one pull launch: enough for every CU to hold 2048 threads, no
 * more than there are items, a multiple of the 8 XCDs. */
static uint32_t og_multi_grid(uint32_t block, size_t items)
{
    uint64_t g = (uint64_t)(og_num_cus() > 0 ? og_num_cus() : 256) * (2048u / block);
    if (items < g) g = items;
    g = (g + OG_XCDS - 1) / OG_XCDS * OG_XCDS;
    return g ? (uint32_t)g : OG_XCDS;
}

/* The items [w0, w0 + sum cnt) of graphs gs[0..ng) (cnt[i] items each, in
 * that order) cut into 8 contiguous ranges of about equal work, one per
 * XCD: a graph's work is its edges and vertices, spread evenly over its
 * items.  Graphs stay whole except at the cuts, so each XCD's L2 holds the
 * contributions of its own graphs. */
static void og_multi_split(const ogpu_multi *m, const uint32_t *gs, const size_t *cnt, uint32_t ng, uint32_t w0,
                           og_multi_xr *xr)
{
    double tot = 0.0;
    for (uint32_t i = 0; i < ng; i++) tot += (double)m->pr[gs[i]]->dev.E + m->pr[gs[i]]->dev.N;
    uint32_t x = 0, pos = w0;
    double acc = 0.0;
    xr->o[0] = w0;
    for (uint32_t i = 0; i < ng; i++)
    {
        const double w = cnt[i] ? ((double)m->pr[gs[i]]->dev.E + m->pr[gs[i]]->dev.N) / (double)cnt[i] : 0.0;
        for (size_t j = 0; j < cnt[i]; j++, pos++)
        {
            while (x + 1 < OG_XCDS && acc >= tot * (double)(x + 1) / OG_XCDS) xr->o[++x] = pos;
            acc += w;
        }
    }
    while (x < OG_XCDS) xr->o[++x] = pos;
}

/* Milliseconds per pull of the graphs gs[0..ng) with tile shape `tile`:
 * their items for that shape go to the front of the work list and are
 * pulled OG_MULTI_TILE_WARM times untimed, then OG_MULTI_TILE_ITERS times
 * between HIP events.  The graphs must be initialised (a pull iterates
 * their ranks as a run does; the run initialises them again). */
static int og_multi_time_tile(ogpu_multi *m, const uint32_t *gs, uint32_t ng, uint32_t tile,
                              og_multi_work *host, size_t *cnt, hipEvent_t e0, hipEvent_t e1, float *ms)
{
    og_multi_tile t;
    (void)og_multi_tile_shape(tile, &t);
    size_t k = 0;
    for (uint32_t i = 0; i < ng; i++)
    {
        cnt[i] = og_multi_items(&m->pr[gs[i]]->dev, gs[i], &t, host + k);
        k += cnt[i];
    }
    og_multi_xr xr;
    og_multi_split(m, gs, cnt, ng, 0, &xr);
    hipStream_t s = og_stream();
    if (og_hip_check(hipMemcpyAsync(m->d_work, host, k * sizeof(og_multi_work), hipMemcpyHostToDevice, s), "H2D"))
        return -1;
    const uint32_t grid = og_multi_grid(t.block, k);
    for (uint32_t it = 0; it < OG_MULTI_TILE_WARM; it++)
        if (og_launch_multi_pull(&m->dev, tile, &xr, 2 + it, grid, s)) return -1;
    if (og_hip_check(hipEventRecord(e0, s), "hipEventRecord")) return -1;
    for (uint32_t it = 0; it < OG_MULTI_TILE_ITERS; it++)
        if (og_launch_multi_pull(&m->dev, tile, &xr, 2 + OG_MULTI_TILE_WARM + it, grid, s)) return -1;
    if (og_hip_check(hipEventRecord(e1, s), "hipEventRecord") ||
        og_hip_check(hipEventSynchronize(e1), "hipEventSynchronize") ||
        og_hip_check(hipEventElapsedTime(ms, e0, e1), "hipEventElapsedTime"))
        return -1;
    *ms /= OG_MULTI_TILE_ITERS;
    return 0;
}

/* Tile shape per graph: given (tiles[g].block != 0), or tuned once per size
 * class on all the class's graphs together, by coordinate descent from 256
 * threads, 4 chunks per wavefront, row batches of 8: the block size first,
 * then the chunks per wavefront, then the row batch, each the fastest of
 * its candidates with the others held. */
static int og_multi_pick_tiles(ogpu_multi *m, const ogpu_multi_tile *tiles, og_multi_work *host)
{
    static const uint32_t blocks[] = {128, 256, 512}, chunks[] = {1, 2, 4}, batches[] = {0, 8};
    uint32_t *gs = (uint32_t *)malloc((m->n ? m->n : 1) * sizeof(uint32_t));
    size_t *cnt = (size_t *)malloc((m->n ? m->n : 1) * sizeof(size_t));
    int8_t done[33] = {0};
    hipEvent_t e0 = NULL, e1 = NULL;
    int rc = -1;
    if (!gs || !cnt) { free(gs); free(cnt); og_set_error("out of host memory"); return -1; }
    for (uint32_t g = 0; g < m->n; g++)
    {
        if (tiles && tiles[g].block)
        {
            const int id = og_multi_tile_id(tiles[g].block, tiles[g].chunks_per_wave, tiles[g].row_batch);
            if (id < 0)
            {
                og_set_error("ogpu_multi_create: no compiled tile shape (block %u, chunks %u, row batch %u)",
                             tiles[g].block, tiles[g].chunks_per_wave, tiles[g].row_batch);
                goto out;
            }
            m->tile[g] = (uint32_t)id;
        }
        else
            m->tile[g] = UINT32_MAX;
    }
    for (uint32_t g = 0; g < m->n; g++)
    {
        if (m->tile[g] != UINT32_MAX) continue;
        const uint32_t cls = og_class_of(m->pr[g]->dev.N);
        if (done[cls]) continue;
        done[cls] = 1;
        uint32_t ng = 0;
        for (uint32_t q = g; q < m->n; q++)
            if (m->tile[q] == UINT32_MAX && og_class_of(m->pr[q]->dev.N) == cls) gs[ng++] = q;
        if (!e0)
        {
            if (og_hip_check(hipEventCreate(&e0), "hipEventCreate") || og_hip_check(hipEventCreate(&e1), "hipEventCreate") ||
                og_launch_multi_init(&m->dev, m->grid, og_stream()))
                goto out;
        }
        float t_ms[64];
        for (int i = 0; i < 64; i++) t_ms[i] = -1.0f;
        uint32_t blk = 256, ch = 4, rb = 8;
        for (int step = 0; step < 3; step++)
        {
            const uint32_t *c = step == 0 ? blocks : step == 1 ? chunks : batches;
            const uint32_t nc = step == 2 ? 2 : 3;
            uint32_t best = step == 0 ? blk : step == 1 ? ch : rb;
            float best_ms = -1.0f;
            for (uint32_t i = 0; i < nc; i++)
            {
                const int id = og_multi_tile_id(step == 0 ? c[i] : blk, step == 1 ? c[i] : ch, step == 2 ? c[i] : rb);
                if (id < 0 || id >= 64) continue;
                if (t_ms[id] < 0.0f && og_multi_time_tile(m, gs, ng, (uint32_t)id, host, cnt, e0, e1, &t_ms[id])) goto out;
                if (best_ms < 0.0f || t_ms[id] < best_ms)
                {
                    best_ms = t_ms[id];
                    best = c[i];
                }
            }
            if (step == 0) blk = best;
            else if (step == 1) ch = best;
            else rb = best;
        }
        const int id = og_multi_tile_id(blk, ch, rb);
        for (uint32_t i = 0; i < ng; i++) m->tile[gs[i]] = (uint32_t)id;
    }
    rc = 0;
out:
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    free(gs);
    free(cnt);
    return rc;
}

int ogpu_multi_create(ogpu_graph *const *graphs_in, uint32_t n, const ogpu_pr_options *opt_in,
                      const uint32_t *hub, ogpu_multi **out)
{
    return ogpu_multi_create_tiled(graphs_in, n, opt_in, hub, NULL, out);
}

int ogpu_multi_create_tiled(ogpu_graph *const *graphs_in, uint32_t n, const ogpu_pr_options *opt_in,
                            const uint32_t *hub, const ogpu_multi_tile *tiles, ogpu_multi **out)
{
    *out = NULL;
    og_graph *const *graphs = (og_graph *const *)graphs_in;
    ogpu_pr_options opt;
    if (opt_in) opt = *opt_in; else ogpu_pr_default_options(&opt);
    if (opt.variant != 0 || opt.ppr || opt.init_mode != 0)
    {
        og_set_error("ogpu_multi_create: plain pull (variant 0, reference start) only");
        return -1;
    }
    if (og_ctx_ensure()) return -1;
    if (og_multi_tile_count() > 64) { og_set_error("ogpu_multi_create: too many tile shapes"); return -1; }
    ogpu_multi *m = (ogpu_multi *)calloc(1, sizeof(ogpu_multi));
    if (!m) { og_set_error("out of host memory"); return -1; }
    m->n = n;
    m->max_iter = opt.max_iter;
    m->pr = (ogpu_pr **)calloc(n ? n : 1, sizeof(ogpu_pr *));
    m->hub = (uint32_t *)calloc(n ? n : 1, 4);
    m->tile = (uint32_t *)calloc(n ? n : 1, 4);
    m->launched = (int64_t *)calloc(n ? n : 1, 8);
    if (!m->pr || !m->hub || !m->tile || !m->launched) { og_set_error("out of host memory"); ogpu_multi_free(m); return -1; }
    if (og_multi_pick(m, graphs, &opt, hub)) { ogpu_multi_free(m); return -1; }
    /* per-graph runs, device table and work lists */
    og_pr_dev *devs = (og_pr_dev *)calloc(2 * (size_t)(n ? n : 1), sizeof(og_pr_dev));
    uint32_t *wr = (uint32_t *)calloc(2 * (size_t)n + 1, 4), *lr = (uint32_t *)calloc(n + 1, 4);
    size_t max_work = 0, n_work = 0, n_lwork = 0, n_iwork = 0;
    int rc = -1;
    og_multi_work *work = NULL, *lwork = NULL, *iwork = NULL;
    og_multi_tile smallest = {128, 1, 0};   /* the shape with the most items */
    if (!devs || !wr || !lr) { og_set_error("out of host memory"); goto fail; }
    for (uint32_t g = 0; g < n; g++)
    {
        ogpu_pr_options o = opt;
        /* every row in CSR order by thread-per-row rows: the multi-graph
         * kernels have no wavefront-chain items */
        o.hub_threshold = m->hub[g] >= 0x80000000u ? 0x7FFFFFFFu : m->hub[g];
        o.exact_above = 0;
        o.binned = 2;
        if (og_pr_create_with(graphs[g], NULL, &o, &m->pr[g])) goto fail;
        ogpu_pr *p = m->pr[g];
        for (int q = 0; q < 2; q++)
        {
            og_pr_dev a = p->dev;
            a.g_in = p->c[q];
            a.c_out = p->c[q ^ 1];
            a.allow_stop = 1;
            devs[2 * g + q] = a;
        }
        lr[g] = (uint32_t)n_lwork;
        max_work += og_multi_items(&p->dev, g, &smallest, NULL);
        n_lwork += (p->plan->H + OG_MULTI_WAVES - 1) / OG_MULTI_WAVES;
        n_iwork += (p->g->N + 255) / 256;
    }
    lr[n] = (uint32_t)n_lwork;
    work = (og_multi_work *)calloc(max_work ? max_work : 1, sizeof(og_multi_work));
    lwork = (og_multi_work *)calloc(n_lwork ? n_lwork : 1, sizeof(og_multi_work));
    iwork = (og_multi_work *)calloc(n_iwork ? n_iwork : 1, sizeof(og_multi_work));
    m->groups = (og_multi_group *)calloc(og_multi_tile_count(), sizeof(og_multi_group));
    if (!work || !lwork || !iwork || !m->groups) { og_set_error("out of host memory"); goto fail; }
    {
        size_t l = 0, i = 0;
        for (uint32_t g = 0; g < n; g++)
        {
            og_plan *pl = m->pr[g]->plan;
            for (uint32_t b = 0; b < (pl->H + OG_MULTI_WAVES - 1) / OG_MULTI_WAVES; b++)
                lwork[l++] = (og_multi_work){g, 2, b, 0};
            for (uint32_t b = 0; b < (m->pr[g]->g->N + 255) / 256; b++) iwork[i++] = (og_multi_work){g, 3, b, 0};
        }
    }
    m->d_devs = og_dalloc(2 * (size_t)(n ? n : 1) * sizeof(og_pr_dev));
    m->d_work = og_dalloc((max_work ? max_work : 1) * sizeof(og_multi_work));
    m->d_lwork = og_dalloc((n_lwork ? n_lwork : 1) * sizeof(og_multi_work));
    m->d_iwork = og_dalloc((n_iwork ? n_iwork : 1) * sizeof(og_multi_work));
    m->d_wr = og_dalloc(2 * (size_t)n * 4 + 4);
    m->d_lr = og_dalloc((size_t)(n + 1) * 4);
    m->w_err = og_dalloc((max_work ? max_work : 1) * 8);
    m->w_act = og_dalloc((max_work ? max_work : 1) * 4);
    m->l_err = og_dalloc((n_lwork ? n_lwork : 1) * 8);
    m->l_act = og_dalloc((n_lwork ? n_lwork : 1) * 4);
    m->d_live = og_dalloc(4);
    m->d_all_done = og_dalloc(4);
    if (!m->d_devs || !m->d_work || !m->d_lwork || !m->d_iwork || !m->d_wr || !m->d_lr || !m->w_err ||
        !m->w_act || !m->l_err || !m->l_act || !m->d_live || !m->d_all_done)
        goto fail;
    {
        hipStream_t s = og_stream();
        if (og_hip_check(hipMemcpyAsync(m->d_devs, devs, 2 * (size_t)n * sizeof(og_pr_dev), hipMemcpyHostToDevice, s), "H2D") ||
            og_hip_check(hipMemcpyAsync(m->d_lwork, lwork, n_lwork * sizeof(og_multi_work), hipMemcpyHostToDevice, s), "H2D") ||
            og_hip_check(hipMemcpyAsync(m->d_iwork, iwork, n_iwork * sizeof(og_multi_work), hipMemcpyHostToDevice, s), "H2D") ||
            og_hip_check(hipMemcpyAsync(m->d_lr, lr, (size_t)(n + 1) * 4, hipMemcpyHostToDevice, s), "H2D") ||
            og_hip_check(hipStreamSynchronize(s), "sync"))
            goto fail;
    }
    og_multi_dev *d = &m->dev;
    d->devs = m->d_devs;
    d->work = m->d_work;
    d->lwork = m->d_lwork;
    d->iwork = m->d_iwork;
    d->n_lwork = (uint32_t)n_lwork;
    d->n_iwork = (uint32_t)n_iwork;
    d->n_graphs = n;
    d->wr = m->d_wr;
    d->lr = m->d_lr;
    d->w_err = m->w_err;
    d->l_err = m->l_err;
    d->w_act = m->w_act;
    d->l_act = m->l_act;
    d->live = m->d_live;
    d->all_done = m->d_all_done;
    m->grid = (uint32_t)og_num_cus() * 8;
    if (m->grid == 0) m->grid = 2048;
    if (og_multi_pick_tiles(m, tiles, work)) goto fail;
    /* the work list: the items of every graph of tile shape 0, then of
     * shape 1, ...; one pull launch per shape in use, its items cut into
     * one range per XCD */
    {
        uint32_t *gs = (uint32_t *)malloc((n ? n : 1) * sizeof(uint32_t));
        size_t *cnt = (size_t *)malloc((n ? n : 1) * sizeof(size_t));
        if (!gs || !cnt) { free(gs); free(cnt); og_set_error("out of host memory"); goto fail; }
        for (uint32_t t = 0; t < og_multi_tile_count(); t++)
        {
            og_multi_tile sh;
            (void)og_multi_tile_shape(t, &sh);
            const size_t w0 = n_work;
            uint32_t ng = 0;
            for (uint32_t g = 0; g < n; g++)
            {
                if (m->tile[g] != t) continue;
                wr[2 * g] = (uint32_t)n_work;
                cnt[ng] = og_multi_items(&m->pr[g]->dev, g, &sh, work + n_work);
                n_work += cnt[ng];
                gs[ng++] = g;
                wr[2 * g + 1] = (uint32_t)n_work;
            }
            if (n_work > w0)
            {
                og_multi_group *q = &m->groups[m->n_groups++];
                q->tile = t;
                q->grid = og_multi_grid(sh.block, n_work - w0);
                og_multi_split(m, gs, cnt, ng, (uint32_t)w0, &q->xr);
            }
        }
        free(gs);
        free(cnt);
    }
    d->n_work = (uint32_t)n_work;
    {
        hipStream_t s = og_stream();
        if (og_hip_check(hipMemcpyAsync(m->d_work, work, (n_work ? n_work : 1) * sizeof(og_multi_work), hipMemcpyHostToDevice, s), "H2D") ||
            og_hip_check(hipMemcpyAsync(m->d_wr, wr, 2 * (size_t)n * 4 + 4, hipMemcpyHostToDevice, s), "H2D") ||
            og_hip_check(hipStreamSynchronize(s), "sync"))
            goto fail;
    }
    rc = 0;
fail:
    free(devs); free(wr); free(lr); free(work); free(lwork); free(iwork);
    if (rc) { ogpu_multi_free(m); return -1; }
    *out = m;
    return 0;
}

static int og_multi_enqueue_all(ogpu_multi *m, hipStream_t s)
{
    for (uint32_t it = 0; it < m->max_iter; it++)
        if (og_launch_multi_iter(&m->dev, m->groups, m->n_groups, it, m->grid, s)) return -1;
    return 0;
}

int ogpu_multi_run(ogpu_multi *m, int mode, double *seconds)
{
    hipStream_t s = og_stream();
    double t0 = og_now();
    if (mode == 0)
    {
        for (uint32_t g = 0; g < m->n; g++)
        {
            m->launched[g] = og_pr_solve(m->pr[g]);
            if (m->launched[g] < 0) return -1;
        }
    }
    else
    {
        if (mode == 2 && !m->have_exec)
        {
            /* capture the max_iter iterations once; the init stays outside
             * so that a replay starts from a fresh state */
            hipGraph_t graph;
            if (og_hip_check(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal), "hipStreamBeginCapture"))
                return -1;
            int bad = og_multi_enqueue_all(m, s);
            hipError_t e = hipStreamEndCapture(s, &graph);
            if (bad || og_hip_check(e, "hipStreamEndCapture")) return -1;
            e = hipGraphInstantiate(&m->exec, graph, NULL, NULL, 0);
            (void)hipGraphDestroy(graph);
            if (og_hip_check(e, "hipGraphInstantiate")) return -1;
            m->have_exec = 1;
            t0 = og_now();
        }
        if (og_launch_multi_init(&m->dev, m->grid, s)) return -1;
        if (mode == 2)
        {
            if (og_hip_check(hipGraphLaunch(m->exec, s), "hipGraphLaunch")) return -1;
        }
        else if (og_multi_enqueue_all(m, s))
            return -1;
        if (og_hip_check(hipStreamSynchronize(s), "sync multi")) return -1;
        for (uint32_t g = 0; g < m->n; g++) m->launched[g] = m->max_iter;
    }
    if (seconds) *seconds = og_now() - t0;
    return 0;
}

int ogpu_multi_result(ogpu_multi *m, uint32_t g, ogpu_pr_result *res, float *ranks)
{
    if (g >= m->n) { og_set_error("ogpu_multi_result: graph %u of %u", g, m->n); return -1; }
    memset(res, 0, sizeof(*res));
    ogpu_pr *p = m->pr[g];
    double pr_sum = 0.0;
    if (og_pr_finish(p, &pr_sum)) return -1;
    uint32_t launched = m->launched[g] > 0 ? (uint32_t)m->launched[g] : 0;
    if (og_pr_collect(p, launched, pr_sum, res)) return -1;
    res->seconds_plan = p->seconds_plan;
    return ranks ? ogpu_pr_ranks(p, ranks) : 0;
}
