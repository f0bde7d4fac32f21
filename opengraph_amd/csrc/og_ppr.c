/*
 * og_ppr.c -- host side of batched personalized PageRank (og_ppr.hip): k <= 64
 * restart vertices iterated together on one plan of the graph.
 */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

#include "og_internal.h"
#include "og_plan.h"
#include "og_ppr.h"
#include "opengraph_amd_ext.h"

#define OG_PPR_CHECK_EVERY 8u

struct ogpu_ppr
{
    og_graph *g;
    og_plan *plan;
    og_ppr_dev dev;
    uint32_t k, max_iter;
    float *x, *c[2], *chunk_part, *out;
    double *part_l1, *fin_l1, *l1;
    uint64_t *active;
    uint32_t *ticket, *done, *iters, *src, *colmap;
    uint32_t *done_host;
    uint32_t src_col[OG_PPR_KMAX];   /* restart vertex (new id) of each column */
    hipEvent_t *ev;              /* 2 per iteration around the main kernel, + start/end */
    uint32_t n_ev;
    uint32_t launched, executed;
    double seconds_plan, seconds_total, kernel_ms;
    uint32_t *kact_it;           /* max_iter: column slots each launched iteration gathered */
    double bytes_run;            /* algorithmic bytes of the last run's executed iterations */
};

static double og_ppr_now(void)
{
    struct timespec ts;
    clock_gettime(CLOCK_MONOTONIC, &ts);
    return ts.tv_sec + ts.tv_nsec * 1e-9;
}

void ogpu_ppr_free(ogpu_ppr *p)
{
    if (!p) return;
    og_plan_destroy(p->plan);
    og_dfree(p->x); og_dfree(p->c[0]); og_dfree(p->c[1]); og_dfree(p->chunk_part); og_dfree(p->out);
    og_dfree(p->part_l1); og_dfree(p->fin_l1); og_dfree(p->l1); og_dfree(p->active);
    og_dfree(p->ticket); og_dfree(p->done); og_dfree(p->iters); og_dfree(p->src); og_dfree(p->colmap);
    if (p->done_host) (void)hipHostFree(p->done_host);
    for (uint32_t i = 0; i < p->n_ev; i++) (void)hipEventDestroy(p->ev[i]);
    free(p->ev);
    free(p->kact_it);
    free(p);
}

int ogpu_ppr_create(ogpu_graph *gg, const uint32_t *sources, uint32_t k, uint32_t max_iter,
                    double epsilon, ogpu_ppr **out)
{
    *out = NULL;
    og_graph *g = (og_graph *)gg;
    if (og_ctx_ensure()) return -1;
    if (!g || k == 0 || k > OG_PPR_KMAX)
    {
        og_set_error("ogpu_ppr_create: need a graph and 1 <= k <= %u (got k = %u)", OG_PPR_KMAX, k);
        return -1;
    }
    for (uint32_t j = 0; j < k; j++)
        if (sources[j] >= g->N)
        {
            og_set_error("ogpu_ppr_create: source %u of column %u is not below N = %u", sources[j], j, g->N);
            return -1;
        }
    ogpu_ppr *p = (ogpu_ppr *)calloc(1, sizeof(ogpu_ppr));
    if (!p) { og_set_error("out of host memory"); return -1; }
    p->g = g;
    p->k = k;
    p->max_iter = max_iter;
    hipStream_t s = og_stream();
    double t0 = og_ppr_now();
    if (og_plan_build(g, 1, OG_SEQ_T, 0, 0, &p->plan)) { free(p); return -1; }
    p->seconds_plan = og_ppr_now() - t0;
    og_plan *pl = p->plan;
    const uint32_t N = g->N;
    uint32_t n_multi = 0;
    if (og_hip_check(hipMemcpyAsync(&n_multi, pl->chunk_base + pl->H, 4, hipMemcpyDeviceToHost, s), "D2H") ||
        og_hip_check(hipStreamSynchronize(s), "sync"))
    {
        ogpu_ppr_free(p);
        return -1;
    }
    const size_t blk = (size_t)N * k * 4;
    const uint32_t iters_cap = max_iter ? max_iter : 1;
    p->x = og_dalloc(blk);
    p->c[0] = og_dalloc(blk);
    p->c[1] = og_dalloc(blk);
    p->chunk_part = og_dalloc((size_t)(n_multi ? n_multi : 1) * k * 4);
    p->part_l1 = og_dalloc((size_t)OG_PPR_MAIN_BLOCKS * k * 8);
    p->fin_l1 = og_dalloc((size_t)OG_PPR_FIN_BLOCKS * k * 8);
    p->l1 = og_dalloc((size_t)iters_cap * k * 8);
    p->active = og_dalloc(8);
    p->ticket = og_dalloc(4);
    p->done = og_dalloc(4);
    p->iters = og_dalloc((size_t)k * 4);
    p->src = og_dalloc((size_t)k * 4);
    p->colmap = og_dalloc((size_t)k * 4);
    uint32_t *src_old = og_dalloc((size_t)k * 4);
    int bad = !p->x || !p->c[0] || !p->c[1] || !p->chunk_part || !p->part_l1 || !p->fin_l1 || !p->l1 ||
              !p->active || !p->ticket || !p->done || !p->iters || !p->src || !p->colmap || !src_old ||
              og_hip_check(hipHostMalloc((void **)&p->done_host, 4, hipHostMallocDefault), "hipHostMalloc") ||
              og_hip_check(hipMemcpyAsync(src_old, sources, (size_t)k * 4, hipMemcpyHostToDevice, s), "H2D") ||
              og_launch_plan_map_pairs(k, src_old, src_old, pl->old2new, p->src, p->src, s) ||
              og_hip_check(hipMemcpyAsync(p->src_col, p->src, (size_t)k * 4, hipMemcpyDeviceToHost, s), "D2H") ||
              og_hip_check(hipStreamSynchronize(s), "sync");
    og_dfree(src_old);
    if (bad) { ogpu_ppr_free(p); return -1; }
    p->n_ev = 0;
    p->ev = (hipEvent_t *)calloc((size_t)2 * iters_cap + 2, sizeof(hipEvent_t));
    p->kact_it = (uint32_t *)calloc(iters_cap, sizeof(uint32_t));
    if (!p->ev || !p->kact_it) { og_set_error("out of host memory"); ogpu_ppr_free(p); return -1; }
    for (uint32_t i = 0; i < 2 * iters_cap + 2; i++)
    {
        if (og_hip_check(hipEventCreate(&p->ev[i]), "hipEventCreate")) { ogpu_ppr_free(p); return -1; }
        p->n_ev = i + 1;
    }
    og_ppr_dev *a = &p->dev;
    memset(a, 0, sizeof(*a));
    a->N = N;
    a->E = pl->E;
    a->k = k;
    a->H = pl->H;
    a->C = pl->C;
    a->n_chunks = pl->n_chunks;
    a->max_iter = max_iter;
    a->epsilon = epsilon;
    a->row_ptr = pl->row_ptr;
    a->cols = pl->cols;
    a->outdeg = pl->outdeg;
    a->chunk_base = pl->chunk_base;
    a->ch_row = pl->ch_row;
    a->ch_edge = pl->ch_edge;
    a->old2new = pl->old2new;
    a->src = p->src;
    a->colmap = p->colmap;
    a->kact = k;
    a->x = p->x;
    a->chunk_part = p->chunk_part;
    a->part_l1 = p->part_l1;
    a->fin_l1 = p->fin_l1;
    a->active = p->active;
    a->ticket = p->ticket;
    a->done = p->done;
    a->iters = p->iters;
    a->l1 = p->l1;
    *out = p;
    return 0;
}

int ogpu_ppr_run(ogpu_ppr *p, ogpu_ppr_result *res)
{
    memset(res, 0, sizeof(*res));
    hipStream_t s = og_stream();
    og_ppr_dev a = p->dev;
    a.c_in = p->c[0];
    /* slots start in column order */
    uint32_t colmap[OG_PPR_KMAX], src_slot[OG_PPR_KMAX];
    for (uint32_t j = 0; j < p->k; j++)
    {
        colmap[j] = j;
        src_slot[j] = p->src_col[j];
    }
    a.kact = p->k;
    if (og_hip_check(og_copy(p->colmap, colmap, (size_t)p->k * 4, hipMemcpyHostToDevice), "H2D") ||
        og_hip_check(og_copy(p->src, src_slot, (size_t)p->k * 4, hipMemcpyHostToDevice), "H2D"))
        return -1;
    double t0 = og_ppr_now();
    if (og_hip_check(hipEventRecord(p->ev[0], s), "hipEventRecord")) return -1;
    if (og_launch_ppr_init(&a, s)) return -1;
    p->launched = 0;
    while (p->launched < p->max_iter)
    {
        uint32_t n = p->max_iter - p->launched;
        if (n > OG_PPR_CHECK_EVERY) n = OG_PPR_CHECK_EVERY;
        for (uint32_t i = 0; i < n; i++)
        {
            const uint32_t it = p->launched + i;
            a.it = it;
            a.c_in = p->c[it & 1];
            a.c_out = p->c[(it + 1) & 1];
            if (og_launch_ppr_iter(&a, s, p->ev[2 + 2 * it], p->ev[3 + 2 * it])) return -1;
            p->kact_it[it] = a.kact;
        }
        p->launched += n;
        uint64_t act = 0;
        if (og_hip_check(hipMemcpyAsync(p->done_host, p->done, 4, hipMemcpyDeviceToHost, s), "D2H done") ||
            og_hip_check(hipMemcpyAsync(&act, p->active, 8, hipMemcpyDeviceToHost, s), "D2H active") ||
            og_hip_check(hipStreamSynchronize(s), "sync iterations"))
            return -1;
        if (*p->done_host) break;
        /* Once a quarter of the slots in use hold converged columns, move the
         * iterating ones to the lowest slots (stable), so the gathered rows
         * shrink to their first kact floats.  Each column's arithmetic is
         * unchanged; only its slot moves. */
        uint32_t na = 0;
        for (uint32_t j = 0; j < a.kact; j++) na += (uint32_t)((act >> j) & 1u);
        if (na > 0 && 4 * na <= 3 * a.kact)
        {
            og_ppr_perm perm;
            uint32_t m = 0;
            for (uint32_t j = 0; j < a.kact; j++)
                if ((act >> j) & 1u) perm.p[m++] = (uint8_t)j;
            for (uint32_t j = 0; j < p->k; j++)
                if (j >= a.kact || !((act >> j) & 1u)) perm.p[m++] = (uint8_t)j;
            for (uint32_t j = p->k; j < OG_PPR_KMAX; j++) perm.p[j] = (uint8_t)j;
            uint32_t ncol[OG_PPR_KMAX], nsrc[OG_PPR_KMAX];
            for (uint32_t j = 0; j < p->k; j++)
            {
                ncol[j] = colmap[perm.p[j]];
                nsrc[j] = src_slot[perm.p[j]];
            }
            memcpy(colmap, ncol, (size_t)p->k * 4);
            memcpy(src_slot, nsrc, (size_t)p->k * 4);
            const uint64_t nact = ((uint64_t)1 << na) - 1;   /* na < kact <= 64 */
            if (og_launch_ppr_permute(&a, p->c[p->launched & 1], perm, s) ||
                og_hip_check(hipMemcpyAsync(p->colmap, colmap, (size_t)p->k * 4, hipMemcpyHostToDevice, s), "H2D") ||
                og_hip_check(hipMemcpyAsync(p->src, src_slot, (size_t)p->k * 4, hipMemcpyHostToDevice, s), "H2D") ||
                og_hip_check(hipMemcpyAsync(p->active, &nact, 8, hipMemcpyHostToDevice, s), "H2D") ||
                og_hip_check(hipStreamSynchronize(s), "sync permute"))
                return -1;
            a.kact = na;
        }
    }
    if (og_hip_check(hipEventRecord(p->ev[1], s), "hipEventRecord") ||
        og_hip_check(hipEventSynchronize(p->ev[1]), "hipEventSynchronize"))
        return -1;
    p->seconds_total = og_ppr_now() - t0;
    uint32_t it_col[OG_PPR_KMAX];
    if (og_hip_check(og_copy(it_col, p->iters, (size_t)p->k * 4, hipMemcpyDeviceToHost), "D2H iters"))
        return -1;
    uint32_t executed = 0, converged = 0;
    for (uint32_t j = 0; j < p->k; j++)
        if (it_col[j] > executed) executed = it_col[j];
    uint64_t act = 0;
    if (og_hip_check(og_copy(&act, p->active, 8, hipMemcpyDeviceToHost), "D2H active")) return -1;
    for (uint32_t j = 0; j < p->k; j++) converged += !((act >> j) & 1u);
    p->executed = executed;
    double ms = 0.0;
    for (uint32_t i = 0; i < executed; i++)
    {
        float e = 0.0f;
        if (og_hip_check(hipEventElapsedTime(&e, p->ev[2 + 2 * i], p->ev[3 + 2 * i]), "hipEventElapsedTime"))
            return -1;
        ms += e;
    }
    p->kernel_ms = ms;
    p->bytes_run = 0.0;
    for (uint32_t i = 0; i < executed; i++)
        p->bytes_run += (double)p->g->E * (4.0 + 4.0 * p->kact_it[i]) +
                        (double)p->g->N * (8.0 + 12.0 * p->kact_it[i]);
    res->iterations = executed;
    res->converged = converged;
    res->seconds_total = p->seconds_total;
    res->seconds_plan = p->seconds_plan;
    res->kernel_ms_per_iter = executed ? ms / executed : 0.0;
    return 0;
}

int ogpu_ppr_ranks(ogpu_ppr *p, float *x_out)
{
    hipStream_t s = og_stream();
    const size_t blk = (size_t)p->g->N * p->k * 4;
    if (!p->out && !(p->out = og_dalloc(blk))) return -1;
    if (og_launch_ppr_out(&p->dev, p->out, s) ||
        og_hip_check(hipMemcpyAsync(x_out, p->out, blk, hipMemcpyDeviceToHost, s), "D2H ranks") ||
        og_hip_check(hipStreamSynchronize(s), "sync ranks"))
        return -1;
    return 0;
}

int ogpu_ppr_columns(ogpu_ppr *p, uint32_t *iters, double *l1_hist, uint32_t max_rows)
{
    if (iters && og_hip_check(og_copy(iters, p->iters, (size_t)p->k * 4, hipMemcpyDeviceToHost), "D2H"))
        return -1;
    uint32_t rows = p->executed < max_rows ? p->executed : max_rows;
    if (l1_hist && rows &&
        og_hip_check(og_copy(l1_hist, p->l1, (size_t)rows * p->k * 8, hipMemcpyDeviceToHost), "D2H"))
        return -1;
    return (int)rows;
}

double ogpu_ppr_bytes_per_iteration(const ogpu_ppr *p)
{
    /* per edge: 4 B index + k' * 4 B gathered row; per vertex: row offset,
     * out-degree, and k' * 4 B each of old x, new x and new contribution */
    if (p->executed && p->bytes_run > 0.0) return p->bytes_run / p->executed;
    return (double)p->g->E * (4.0 + 4.0 * p->k) + (double)p->g->N * (8.0 + 12.0 * p->k);
}
