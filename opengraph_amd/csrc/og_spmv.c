/*
 * og_spmv.c -- the reference's benchmark SpMV on CSR (SPMV.c:588-1068) on
 * the GPU: the ogpu_spmv runs and the drop-in SPMV*GraphCSR entry points.
 *
 * y = 0; `iterations` times y += A^T (w x) with the fixed x and w of the
 * reference (og_spmv.hip).  The passes are PageRank's kernels with the SpMV
 * store.  -p 0 runs on the binned pull by default: every row continues y
 * in the reference's order (bit-exact); with a hub_threshold it runs on the
 * degree-binned gather pull, where rows of at most hub_threshold in-edges
 * continue y in order and longer rows add their fixed-tree sum.  -p 2 (and
 * 3 by default) gather over the in-edges, -p 1 (and 3 on request) push over
 * the forward edges, partitioned by destination range.  The integer sums of
 * -p 2 and 3 are exact in any order, so -p 3 (a push in the reference)
 * gathers by default: the same bits without atomics.
 */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

#include "og_bin.h"
#include "og_internal.h"
#include "og_pr.h"
#include "og_pr_run.h"
#include "opengraph_amd.h"
#include "opengraph_amd_ext.h"

struct ogpu_spmv
{
    ogpu_pr *p;               /* plan, y (r), partials of the pull's reduction */
    ogpu_spmv_options opt;
    int mode;                 /* arithmetic of the gathered / scattered terms */
    int scatter;              /* 1: push scatter over the forward edges */
    void *terms;              /* w * x per vertex, new ids */
    uint64_t *acc;            /* u64 accumulators of -p 2, 3 */
    uint32_t it;              /* iterations enqueued since the reset */
    hipEvent_t *ev;           /* ev[0]: reset, ev[i + 1]: end of iteration i */
    uint32_t n_ev;
    double seconds;           /* last run: iterations and output, on device */
};

void ogpu_spmv_default_options(ogpu_spmv_options *o)
{
    memset(o, 0, sizeof(*o));
}

void ogpu_spmv_free(ogpu_spmv *s)
{
    if (!s) return;
    ogpu_pr_free(s->p);
    og_dfree(s->acc);
    for (uint32_t i = 0; i < s->n_ev; i++) (void)hipEventDestroy(s->ev[i]);
    free(s->ev);
    free(s);
}

int ogpu_spmv_create(ogpu_graph *gg, const ogpu_spmv_options *opt_in, ogpu_spmv **out)
{
    return og_spmv_create_with((og_graph *)gg, NULL, opt_in, out);
}

int og_spmv_create_with(og_graph *g, og_plan *plan, const ogpu_spmv_options *opt_in, ogpu_spmv **out)
{
    *out = NULL;
    ogpu_spmv_options o;
    if (opt_in) o = *opt_in; else ogpu_spmv_default_options(&o);
    if (o.code > 3)
    {
        og_set_error("SpMV -p code %u is not 0..3 (pull, push, fixed-point pull, fixed-point push)", o.code);
        return -1;
    }
    ogpu_spmv *s = (ogpu_spmv *)calloc(1, sizeof(ogpu_spmv));
    if (!s) { og_set_error("out of host memory"); return -1; }
    s->opt = o;
    s->scatter = o.code == 1 || (o.code == 3 && o.scatter);
    s->mode = o.code < 2 ? OG_MODE_F32 : o.code == 2 ? OG_MODE_FX32 : OG_MODE_FX64;
    ogpu_pr_options po;
    ogpu_pr_default_options(&po);
    /* the PageRank variant whose plan, buffers and kernels the passes use */
    po.variant = o.code == 0 ? 0 : o.code == 1 ? 1 : o.code == 2 ? 9 : (s->scatter ? 3 : 2);
    po.max_iter = 1;
    po.hub_threshold = o.hub_threshold;
    /* the pushes (-p 1, -p 3 scattered) are partitioned by destination
     * range by default (hot_cache 0): LDS atomics onto each destination's
     * running y, as the reference's atomics (its sums are unordered too) */
    po.hot_cache = o.hot_cache;
    /* every destination of the -p 1 scatter takes its adds as atomics (the
     * PageRank push would gather its longest rows instead) */
    if (o.code == 1) po.exact_above = 0;
    /* -p 0 with the default row split runs on the binned pull (every row's
     * running y in the reference's order); a given hub_threshold, and the
     * other codes, on the gather pull's kernels */
    po.binned = o.code == 0 && o.hub_threshold == 0 ? 0 : 2;
    if (og_pr_create_with(g, plan, &po, &s->p))
    {
        free(s);
        return -1;
    }
    s->terms = s->p->c[0];
    if (o.code >= 2)
    {
        s->acc = (uint64_t *)og_dalloc((size_t)g->N * 8);
        if (!s->acc && g->N) { ogpu_spmv_free(s); return -1; }
    }
    *out = s;
    return 0;
}

int ogpu_spmv_reset(ogpu_spmv *s)
{
    hipStream_t st = og_stream();
    ogpu_pr *p = s->p;
    const uint32_t N = p->g->N;
    if (og_pr_events(&s->ev, &s->n_ev, 1) || og_hip_check(hipEventRecord(s->ev[0], st), "hipEventRecord"))
        return -1;
    s->it = 0;
    og_pr_dev a = p->dev;
    if (og_launch_fill_u32(p->done, 0, 1, st) || og_launch_fill_u32(p->ticket, 0, 1, st) ||
        og_launch_fill_u32((uint32_t *)p->r, 0, N, st) ||
        (s->acc && og_launch_fill_u32((uint32_t *)s->acc, 0, (uint64_t)N * 2, st)) ||
        og_launch_spmv_terms(&a, s->opt.code, s->terms, st))
        return -1;
    return 0;
}

/* One pass y += A^T t. */
static int og_spmv_launch_iter(ogpu_spmv *s, hipEvent_t e0, hipEvent_t e1)
{
    hipStream_t st = og_stream();
    ogpu_pr *p = s->p;
    og_pr_dev a = p->dev;
    a.allow_stop = 0;
    a.g_in = s->terms;
    a.c_out = p->c[1];
    if (s->scatter)
    {
        a.next = s->opt.code == 1 ? (void *)p->r : (void *)s->acc;
        return og_launch_push_scatter(&a, s->mode, st, e0, e1);
    }
    a.spmv = 1;
    a.next = s->acc;
    if (p->bins)
    {
        /* binned: gather of the terms, then the sum and long-row passes
         * adding onto each row's y; rows without in-edges keep theirs */
        og_bin_dev b = p->bins->d;
        b.n_bins_live = b.n_bins - p->bins->n_zero_bins;
        a.n_parts = p->bin_sum_grid + p->bin_long_grid;
        if (a.n_parts > p->parts_cap)
        {
            og_set_error("SpMV binned pass with %u partials, %u allocated", a.n_parts, p->parts_cap);
            return -1;
        }
        return og_launch_bin_pull(&b, &a, 0, p->bin_sum_grid, p->bin_long_grid, st, e0, e1);
    }
    /* rows [Z, N) have no in-edges: their y stays 0 */
    a.row_end = a.Z;
    a.n_row_blocks = (uint32_t)(((uint64_t)(a.Z - a.C) + 255) / 256);
    a.z_active = 0;
    a.n_parts = og_pull_parts(&a);
    a.n_final_blocks = p->n_final_pull;
    if (a.n_parts > p->parts_cap)
    {
        og_set_error("SpMV pull launch with %u partials, %u allocated", a.n_parts, p->parts_cap);
        return -1;
    }
    return og_launch_pull(&a, s->mode, OG_NO_STATS, st, e0, e1);
}

int ogpu_spmv_enqueue(ogpu_spmv *s, uint32_t n, int time_kernels)
{
    ogpu_pr *p = s->p;
    if (time_kernels && og_pr_events(&p->kev, &p->n_kev, p->kev_used + 2 * n)) return -1;
    for (uint32_t i = 0; i < n; i++)
    {
        hipEvent_t e0 = NULL, e1 = NULL;
        if (time_kernels)
        {
            e0 = p->kev[p->kev_used++];
            e1 = p->kev[p->kev_used++];
        }
        if (og_spmv_launch_iter(s, e0, e1)) return -1;
        s->it++;
    }
    return 0;
}

int ogpu_spmv_kernel_ms(ogpu_spmv *s, double *total_ms, uint32_t *count)
{
    return ogpu_pr_kernel_ms(s->p, total_ms, count);
}

int ogpu_spmv_run(ogpu_spmv *s, uint32_t iterations, double *seconds)
{
    hipStream_t st = og_stream();
    if (ogpu_spmv_reset(s) || og_pr_events(&s->ev, &s->n_ev, iterations + 2)) return -1;
    for (uint32_t i = 0; i < iterations; i++)
        if (ogpu_spmv_enqueue(s, 1, 0) || og_hip_check(hipEventRecord(s->ev[i + 1], st), "hipEventRecord"))
            return -1;
    og_pr_dev a = s->p->dev;
    a.next = s->acc;
    if (og_launch_spmv_out(&a, s->opt.code, s->p->out_old, st) ||
        og_hip_check(hipEventRecord(s->ev[iterations + 1], st), "hipEventRecord") ||
        og_hip_check(hipEventSynchronize(s->ev[iterations + 1]), "sync spmv"))
        return -1;
    float ms = 0.0f;
    if (og_hip_check(hipEventElapsedTime(&ms, s->ev[0], s->ev[iterations + 1]), "hipEventElapsedTime"))
        return -1;
    s->seconds = ms * 1e-3;
    if (seconds) *seconds = s->seconds;
    return 0;
}

int ogpu_spmv_iteration_seconds(ogpu_spmv *s, uint32_t i, double *seconds)
{
    float ms = 0.0f;
    if (i + 1 >= s->n_ev || i >= s->it)
    {
        og_set_error("ogpu_spmv_iteration_seconds: iteration %u was not run", i);
        return -1;
    }
    if (og_hip_check(hipEventElapsedTime(&ms, s->ev[i], s->ev[i + 1]), "hipEventElapsedTime")) return -1;
    *seconds = ms * 1e-3;
    return 0;
}

int ogpu_spmv_output(ogpu_spmv *s, float *host_out)
{
    if (og_hip_check(hipMemcpyAsync(host_out, s->p->out_old, (size_t)s->p->g->N * 4, hipMemcpyDeviceToHost,
                                    og_stream()), "D2H vector_output"))
        return -1;
    return og_hip_check(hipStreamSynchronize(og_stream()), "sync vector_output");
}

double ogpu_spmv_bytes_per_iteration(const ogpu_spmv *s)
{
    const double N = s->p->g->N, E = s->p->g->E;
    switch (s->opt.code)
    {
    /* partitioned pushes: source id, u16 destination offset, gathered term
     * per edge; y read and written per vertex.  Scatters: src, dst and the
     * atomic read + write per edge. */
    case 1: return s->p->pp_v0 ? 10.0 * E + 8.0 * N : 16.0 * E + 4.0 * N;
    case 2: return 8.0 * E + 20.0 * N;   /* index + u32 term per edge; row_ptr, u64 y read + write */
    case 3:
        return !s->scatter ? 12.0 * E + 20.0 * N   /* index + u64 term per edge */
               : s->p->pp_v0 ? 14.0 * E + 16.0 * N
                             : 24.0 * E + 8.0 * N;
    default: return 8.0 * E + 12.0 * N;  /* index + fp32 term per edge; row_ptr, y read + write */
    }
}

/* ------------------------------------------------------------------ */
/* drop-in entry points (SPMV.c:45-66, :588-1068)                      */
/* ------------------------------------------------------------------ */

static void og_spmv_fatal(const char *where)
{
    fprintf(stderr, "opengraph_amd: %s: %s\n", where, og_get_error());
    exit(EXIT_FAILURE);
}

struct SPMVStats *newSPMVStatsGraphCSR(struct GraphCSR *graph)
{
    struct SPMVStats *stats = (struct SPMVStats *)og_host_alloc(sizeof(struct SPMVStats));
    stats->iterations = 0;
    stats->arguments = NULL;
    stats->num_vertices = graph->num_vertices;
    stats->time_total = 0.0;
    stats->vector_output = (float *)og_host_alloc((size_t)graph->num_vertices * sizeof(float));
    stats->vector_input = (float *)og_host_alloc((size_t)graph->num_vertices * sizeof(float));
    for (uint32_t v = 0; v < stats->num_vertices; v++)
    {
        stats->vector_output[v] = 0.0f;
        stats->vector_input[v] = 0.0f;
    }
    return stats;
}

void freeSPMVStats(struct SPMVStats *stats)
{
    if (!stats) return;
    free(stats->vector_output);
    free(stats->vector_input);
    free(stats);
}

static const char *og_spmv_title(uint32_t code)
{
    switch (code)
    {
    case 1: return "Starting SPMV-PUSH";
    case 2: return "Starting SPMV-PULL Fixed-Point";
    case 3: return "Starting SPMV-PUSH Fixed-Point";
    default: return "Starting SPMV-PULL";
    }
}

static struct SPMVStats *og_spmv_dropin(struct Arguments *arguments, struct GraphCSR *graph, uint32_t code)
{
    og_graph *g = NULL;
    og_plan *plan = NULL;
    if (og_cache_graph(graph, code == 1, &g, &plan)) og_spmv_fatal("graph upload");
    og_trace_entry("SPMVGraphCSR", g->N, g->E);
    ogpu_spmv_options o;
    ogpu_spmv_default_options(&o);
    o.code = code;
    ogpu_spmv *s = NULL;
    /* -p 0 runs binned, on a plan of its own that keeps every row in CSR
     * order (the cached plan sorts long rows for the gather pulls) */
    if (og_spmv_create_with(g, code == 0 ? NULL : plan, &o, &s)) og_spmv_fatal("spmv setup");
    struct SPMVStats *stats = newSPMVStatsGraphCSR(graph);
    stats->arguments = arguments;
    /* SPMV.c:663-669: the input vector the passes use */
    for (uint32_t v = 0; v < graph->num_vertices; v++)
    {
        const uint32_t d = graph->vertices->out_degree[v];
        stats->vector_input[v] = d ? 1.0f / (float)d : 0.001f;
    }
    printf(" -----------------------------------------------------\n");
    printf("| %-51s | \n", og_spmv_title(code));
    printf(" -----------------------------------------------------\n");
    printf("| %-21s | %-27s | \n", "Iteration", "Time (S)");
    printf(" -----------------------------------------------------\n");
    double secs = 0.0;
    if (ogpu_spmv_run(s, arguments->iterations, &secs)) og_spmv_fatal("spmv run");
    if (ogpu_spmv_output(s, stats->vector_output)) og_spmv_fatal("spmv output");
    for (uint32_t i = 0; i < arguments->iterations; i++)
    {
        double t = 0.0;
        if (ogpu_spmv_iteration_seconds(s, i, &t)) og_spmv_fatal("spmv timing");
        printf("| %-21u | %-27f | \n", i, t);
    }
    stats->iterations = arguments->iterations;
    stats->time_total = secs;
    /* SPMV.c:697-701: the rounded sum */
    double sum = 0.0;
    for (uint32_t v = 0; v < graph->num_vertices; v++)
        sum += ((int)(stats->vector_output[v] * 100 + .5) / 100.0);
    printf(" -----------------------------------------------------\n");
    printf("| %-15s | %-15s | %-15s | \n", "Iterations", "Sum", "Time (S)");
    printf(" -----------------------------------------------------\n");
    printf("| %-15u | %-15lf | %-15f | \n", stats->iterations, sum, stats->time_total);
    printf(" -----------------------------------------------------\n");
    ogpu_spmv_free(s);
    return stats;
}

struct SPMVStats *SPMVGraphCSR(struct Arguments *arguments, struct GraphCSR *graph)
{
    const uint32_t c = arguments->pushpull;
    return og_spmv_dropin(arguments, graph, c <= 3 ? c : 0);   /* SPMV.c:609-611: default is pull */
}

struct SPMVStats *SPMVPullGraphCSR(struct Arguments *a, struct GraphCSR *g) { return og_spmv_dropin(a, g, 0); }
struct SPMVStats *SPMVPushGraphCSR(struct Arguments *a, struct GraphCSR *g) { return og_spmv_dropin(a, g, 1); }
struct SPMVStats *SPMVPullFixedPointGraphCSR(struct Arguments *a, struct GraphCSR *g) { return og_spmv_dropin(a, g, 2); }
struct SPMVStats *SPMVPushFixedPointGraphCSR(struct Arguments *a, struct GraphCSR *g) { return og_spmv_dropin(a, g, 3); }
