/*
 * og_csr_host.c -- drop-in edge-list and CSR entry points
 * (graphCSR.c, vertex.c, edgeList.c, sortRun.c, radixsort.c, countsort.c,
 * reorder.c:165-221).  Host structures are allocated exactly as the
 * reference allocates them; sorting, vertex mapping and the rank post-sort
 * run on the GPU.
 */
#include <errno.h>
#include <fcntl.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <sys/mman.h>
#include <sys/stat.h>
#include <time.h>
#include <unistd.h>

#include "og_internal.h"
#include "opengraph_amd_ext.h"

int og_launch_rank_keys(const float *d_ranks, uint32_t n, uint32_t *d_keys, hipStream_t s);

static void og_die(const char *what)
{
    fprintf(stderr, "opengraph_amd: %s: %s\n", what, og_get_error());
    exit(EXIT_FAILURE);
}

static double og_wall(void)
{
    struct timespec ts;
    clock_gettime(CLOCK_MONOTONIC, &ts);
    return ts.tv_sec + ts.tv_nsec * 1e-9;
}

static void og_print_time(const char *msg, double t)
{
    /* graphCSR.c:560-569 */
    printf(" -----------------------------------------------------\n");
    printf("| %-51s | \n", msg);
    printf(" -----------------------------------------------------\n");
    printf("| %-51f | \n", t);
    printf(" -----------------------------------------------------\n");
}

/* --------------------------- structures --------------------------- */

struct EdgeList *newEdgeList(uint32_t num_edges)
{
    /* edgeList.c:98-135 */
    struct EdgeList *el = (struct EdgeList *)og_host_alloc(sizeof(struct EdgeList));
    el->edges_array_src = (uint32_t *)og_host_alloc((size_t)num_edges * sizeof(uint32_t));
    el->edges_array_dest = (uint32_t *)og_host_alloc((size_t)num_edges * sizeof(uint32_t));
    memset(el->edges_array_src, 0, (size_t)num_edges * sizeof(uint32_t));
    memset(el->edges_array_dest, 0, (size_t)num_edges * sizeof(uint32_t));
    el->mask_array = NULL;
    el->label_array = NULL;
    el->inverse_label_array = NULL;
    el->num_edges = num_edges;
    el->num_vertices = 0;
    el->avg_degree = 0;
    return el;
}

void freeEdgeList(struct EdgeList *el)
{
    if (!el) return;
    free(el->edges_array_src);
    free(el->edges_array_dest);
    free(el->mask_array);
    free(el->label_array);
    free(el->inverse_label_array);
    free(el);
}

struct Vertex *newVertexArray(uint32_t num_vertices)
{
    /* vertex.c:26-49 */
    struct Vertex *v = (struct Vertex *)og_host_alloc(sizeof(struct Vertex));
    size_t b = (size_t)num_vertices * sizeof(uint32_t);
    v->out_degree = (uint32_t *)og_host_alloc(b);
    v->in_degree = (uint32_t *)og_host_alloc(b);
    v->edges_idx = (uint32_t *)og_host_alloc(b);
    memset(v->out_degree, 0, b);
    memset(v->in_degree, 0, b);
    memset(v->edges_idx, 0, b);
    return v;
}

void freeVertexArray(struct Vertex *v)
{
    if (!v) return;
    free(v->edges_idx);
    free(v->out_degree);
    free(v->in_degree);
    free(v);
}

struct GraphCSR *graphCSRNew(uint32_t V, uint32_t E, uint8_t inverse)
{
    /* graphCSR.c:107-134 */
    struct GraphCSR *g = (struct GraphCSR *)og_host_alloc(sizeof(struct GraphCSR));
    g->num_vertices = V;
    g->num_edges = E;
    g->avg_degree = V ? E / V : 0;
    g->sorted_edges_array = NULL;
    g->inverse_sorted_edges_array = NULL;
    g->vertices = newVertexArray(V);
    g->inverse_vertices = inverse ? newVertexArray(V) : NULL;
    return g;
}

void graphCSRFree(struct GraphCSR *g)
{
    if (!g) return;
    freeVertexArray(g->vertices);
    freeEdgeList(g->sorted_edges_array);
    freeVertexArray(g->inverse_vertices);
    freeEdgeList(g->inverse_sorted_edges_array);
    free(g);
}

static void og_print_max_degree(const char *c2, const char *c3, const struct Vertex *v, uint32_t n,
                                int use_out)
{
    /* vertex.c:327-386: the last vertex holding the maximum */
    uint32_t best = 0, idx = 0;
    for (uint32_t i = 0; i < n; i++)
    {
        uint32_t d = use_out ? v->out_degree[i] : v->in_degree[i];
        if (d >= best) { best = d; idx = i; }
    }
    printf(" -----------------------------------------------------\n");
    printf("| %-15s | %-15s | %-15s | \n", "Node", c2, c3);
    printf(" -----------------------------------------------------\n");
    if (n)
        printf("| %-15u | %-15u | %-15u | \n", idx, v->out_degree[idx], v->in_degree[idx]);
    printf(" -----------------------------------------------------\n");
}

void graphCSRPrint(struct GraphCSR *g)
{
    /* graphCSR.c:66-104 */
    printf(" -----------------------------------------------------\n");
    printf("| %-51s | \n", "GraphCSR Properties");
    printf(" -----------------------------------------------------\n");
    printf("| %-51s | \n", "UN-WEIGHTED");
    printf("| %-51s | \n", "DIRECTED");
    printf(" -----------------------------------------------------\n");
    printf("| %-51s | \n", "Average Degree (D)");
    printf("| %-51u | \n", g->avg_degree);
    printf(" -----------------------------------------------------\n");
    printf("| %-51s | \n", "Number of Vertices (V)");
    printf("| %-51u | \n", g->num_vertices);
    printf(" -----------------------------------------------------\n");
    printf("| %-51s | \n", "Number of Edges (E)");
    printf("| %-51u | \n", g->num_edges);
    printf(" -----------------------------------------------------\n");
    if (g->vertices) og_print_max_degree("*max_out_degree", "in_degree", g->vertices, g->num_vertices, 1);
    if (g->inverse_vertices)
        og_print_max_degree("out_degree", "*max_in_degree", g->inverse_vertices, g->num_vertices, 0);
}

/* ------------------------------ I/O ------------------------------ */

static void og_identity_labels(struct EdgeList *el)
{
    /* edgeList.c:577-588 */
    uint32_t n = el->num_vertices;
    el->mask_array = (uint32_t *)og_host_alloc((size_t)n * 4);
    el->label_array = (uint32_t *)og_host_alloc((size_t)n * 4);
    el->inverse_label_array = (uint32_t *)og_host_alloc((size_t)n * 4);
    for (uint32_t i = 0; i < n; i++)
    {
        el->mask_array[i] = 0;
        el->label_array[i] = i;
        el->inverse_label_array[i] = i;
    }
}

/* mmap a .bin / .wbin file: returns the u32 records and count. */
static const uint32_t *og_map_bin(const char *fname, uint32_t weighted, uint32_t *n_edges,
                                  uint32_t *stride, size_t *map_len, int *fd_out)
{
    int fd = open(fname, O_RDONLY);
    if (fd == -1) { fprintf(stderr, "open: %s: %s\n", fname, strerror(errno)); exit(1); }
    struct stat fs;
    if (fstat(fd, &fs) == -1) { fprintf(stderr, "stat: %s: %s\n", fname, strerror(errno)); exit(1); }
    /* edgeList.c:380-407: 8-byte records, 12 with weights; the count is
     * truncated to whole records. */
    *stride = weighted ? 3 : 2;
    *n_edges = (uint32_t)((uint64_t)fs.st_size / (*stride * 4));
    *map_len = (size_t)fs.st_size;
    *fd_out = fd;
    if (fs.st_size == 0) return NULL;
    void *p = mmap(NULL, (size_t)fs.st_size, PROT_READ, MAP_PRIVATE, fd, 0);
    if (p == MAP_FAILED) { fprintf(stderr, "mmap: %s: %s\n", fname, strerror(errno)); close(fd); exit(1); }
    return (const uint32_t *)p;
}

struct EdgeList *readEdgeListsbin(const char *fname, uint8_t inverse, uint32_t symmetric, uint32_t weighted)
{
    uint32_t n, stride;
    size_t len;
    int fd;
    const uint32_t *buf = og_map_bin(fname, weighted, &n, &stride, &len, &fd);
    struct EdgeList *el = newEdgeList(symmetric ? n * 2 : n);
    uint32_t maxid = 0;
    for (uint32_t i = 0; i < n; i++)
    {
        uint32_t s = buf[(size_t)stride * i], d = buf[(size_t)stride * i + 1];
        if (inverse) { uint32_t t = s; s = d; d = t; }
        el->edges_array_src[i] = s;
        el->edges_array_dest[i] = d;
        if (symmetric)
        {
            el->edges_array_src[i + n] = d;
            el->edges_array_dest[i + n] = s;
        }
        if (s > maxid) maxid = s;
        if (d > maxid) maxid = d;
    }
    if (buf) munmap((void *)buf, len);
    close(fd);
    el->num_vertices = maxid + 1;     /* edgeList.c:567 */
    el->avg_degree = el->num_edges / el->num_vertices;
    og_identity_labels(el);
    return el;
}

struct EdgeList *readEdgeListsMem(struct EdgeList *m, uint8_t inverse, uint32_t symmetric, uint32_t weighted)
{
    /* edgeList.c:597-745 (with WEIGHTED=0 the weight branches vanish and
     * symmetric only keeps the orientation) */
    (void)symmetric;
    (void)weighted;
    uint32_t n = m->num_edges, nv = m->num_vertices;
    struct EdgeList *el = newEdgeList(n);
    if (m->mask_array)
    {
        el->mask_array = (uint32_t *)og_host_alloc((size_t)nv * 4);
        memcpy(el->mask_array, m->mask_array, (size_t)nv * 4);
    }
    if (m->label_array)
    {
        el->label_array = (uint32_t *)og_host_alloc((size_t)nv * 4);
        memcpy(el->label_array, m->label_array, (size_t)nv * 4);
    }
    if (m->inverse_label_array)
    {
        el->inverse_label_array = (uint32_t *)og_host_alloc((size_t)nv * 4);
        memcpy(el->inverse_label_array, m->inverse_label_array, (size_t)nv * 4);
    }
    memcpy(el->edges_array_src, inverse ? m->edges_array_dest : m->edges_array_src, (size_t)n * 4);
    memcpy(el->edges_array_dest, inverse ? m->edges_array_src : m->edges_array_dest, (size_t)n * 4);
    el->num_vertices = nv;
    if (nv) el->avg_degree = n / nv;
    return el;
}

/* ----------------------------- sorting ----------------------------- */

static uint32_t og_bits(uint32_t m)
{
    uint32_t b = 0;
    while (b < 32 && (m >> b)) b++;
    return b ? b : 1;
}

/* Stable sort of a host edge list on the GPU: by_dst_first selects the
 * (src, dst) lexicographic order of radixsort.c:282-335. */
static int og_sort_host_list(struct EdgeList *el, int by_dst_first)
{
    if (og_ctx_ensure()) return -1;
    uint32_t n = el->num_edges;
    if (n < 2) return 0;
    hipStream_t s = og_stream();
    size_t b = (size_t)n * 4;
    uint32_t *k0 = og_dalloc(b), *v0 = og_dalloc(b), *k1 = og_dalloc(b), *v1 = og_dalloc(b);
    uint32_t *k2 = og_dalloc(b), *v2 = og_dalloc(b), *dmax = og_dalloc(4);
    void *scr = og_dalloc(og_radix_scratch_bytes(n));
    int rc = -1;
    uint32_t maxid = 0;
    if (!k0 || !v0 || !k1 || !v1 || !k2 || !v2 || !dmax || !scr) goto out;
    if (og_hip_check(hipMemcpyAsync(k0, el->edges_array_src, b, hipMemcpyHostToDevice, s), "H2D") ||
        og_hip_check(hipMemcpyAsync(v0, el->edges_array_dest, b, hipMemcpyHostToDevice, s), "H2D"))
        goto out;
    if (og_launch_max_id(k0, v0, n, dmax, s) ||
        og_hip_check(hipMemcpyAsync(&maxid, dmax, 4, hipMemcpyDeviceToHost, s), "D2H") ||
        og_hip_check(hipStreamSynchronize(s), "sync"))
        goto out;
    {
        uint32_t bits = og_bits(maxid);
        const uint32_t *rk = k0, *rv = v0;   /* current (src, dst) arrays */
        int in_b = 0;
        if (by_dst_first)
        {
            /* by dst: (dst, src) pairs; then by src, reading the src column */
            if (og_radix_sort_pairs(v0, k0, k1, v1, k2, v2, n, bits, scr, s, &in_b)) goto out;
            uint32_t *rk1 = in_b ? k2 : k1, *rv1 = in_b ? v2 : v1;
            if (og_radix_sort_pairs(rv1, rk1, k0, v0, rk1, rv1, n, bits, scr, s, &in_b)) goto out;
            rk = in_b ? rk1 : k0;
            rv = in_b ? rv1 : v0;
        }
        else
        {
            if (og_radix_sort_pairs(k0, v0, k1, v1, k2, v2, n, bits, scr, s, &in_b)) goto out;
            rk = in_b ? k2 : k1;
            rv = in_b ? v2 : v1;
        }
        if (og_hip_check(hipMemcpyAsync(el->edges_array_src, rk, b, hipMemcpyDeviceToHost, s), "D2H") ||
            og_hip_check(hipMemcpyAsync(el->edges_array_dest, rv, b, hipMemcpyDeviceToHost, s), "D2H") ||
            og_hip_check(hipStreamSynchronize(s), "sync"))
            goto out;
    }
    rc = 0;
out:
    og_dfree(k0); og_dfree(v0); og_dfree(k1); og_dfree(v1); og_dfree(k2); og_dfree(v2);
    og_dfree(dmax); og_dfree(scr);
    return rc;
}

struct EdgeList *radixSortEdgesBySource(struct EdgeList *el)
{
    if (og_sort_host_list(el, 0)) og_die("radixSortEdgesBySource");
    return el;
}

struct EdgeList *radixSortEdgesBySourceAndDestination(struct EdgeList *el)
{
    if (og_sort_host_list(el, 1)) og_die("radixSortEdgesBySourceAndDestination");
    return el;
}

/* countsort.c: the same stable orders. */
struct EdgeList *countSortEdgesBySource(struct EdgeList *el) { return radixSortEdgesBySource(el); }
struct EdgeList *countSortEdgesBySourceAndDestination(struct EdgeList *el)
{
    return radixSortEdgesBySourceAndDestination(el);
}

struct EdgeList *sortRunAlgorithms(struct EdgeList *el, uint32_t sort)
{
    /* sortRun.c:26-69 */
    static const char *msg[4] = {"Radix Sort Edges By Source (Seconds)",
                                 "Radix Sort Edges By Source/Dest (Seconds)",
                                 "Count Sort Edges By Source (Seconds)",
                                 "Count Sort Edges By Source/Dest (Seconds)"};
    uint32_t m = sort <= 3 ? sort : 0;
    double t0 = og_wall();
    if (og_sort_host_list(el, m == 1 || m == 3)) og_die("sortRunAlgorithms");
    og_print_time(msg[m], og_wall() - t0);
    return el;
}

/* --------------------------- mapping --------------------------- */

struct GraphCSR *mapVerticesWithInOutDegree(struct GraphCSR *graph, uint8_t inverse)
{
    /* vertex.c:204-325 on the GPU */
    if (og_ctx_ensure()) og_die("mapVerticesWithInOutDegree");
    struct EdgeList *el = inverse ? graph->inverse_sorted_edges_array : graph->sorted_edges_array;
    struct Vertex *v = inverse ? graph->inverse_vertices : graph->vertices;
    struct Vertex *other = inverse ? graph->vertices : graph->inverse_vertices;
    uint32_t E = graph->num_edges, N = graph->num_vertices;
    hipStream_t s = og_stream();
    uint32_t *dk = og_dalloc((size_t)E * 4), *dd = og_dalloc((size_t)N * 4), *di = og_dalloc((size_t)N * 4);
    if (!dk || !dd || !di) og_die("mapVerticesWithInOutDegree");
    if ((E && og_hip_check(hipMemcpyAsync(dk, el->edges_array_src, (size_t)E * 4, hipMemcpyHostToDevice, s), "H2D")) ||
        og_launch_map_vertices(dk, E, N, dd, di, s) ||
        og_hip_check(hipMemcpyAsync(v->out_degree, dd, (size_t)N * 4, hipMemcpyDeviceToHost, s), "D2H") ||
        og_hip_check(hipMemcpyAsync(v->edges_idx, di, (size_t)N * 4, hipMemcpyDeviceToHost, s), "D2H") ||
        og_hip_check(hipStreamSynchronize(s), "sync"))
        og_die("mapVerticesWithInOutDegree");
    og_dfree(dk); og_dfree(dd); og_dfree(di);
    if (other) memcpy(other->in_degree, v->out_degree, (size_t)N * 4);   /* vertex.c:291-310 */
    return graph;
}

struct GraphCSR *graphCSRAssignEdgeList(struct GraphCSR *g, struct EdgeList *el, uint8_t inverse)
{
    /* graphCSR.c:137-161 */
    if (inverse) g->inverse_sorted_edges_array = el;
    else g->sorted_edges_array = el;
    return mapVerticesWithInOutDegree(g, inverse);
}

/* ------------------------- whole pipeline ------------------------- */

static struct EdgeList *og_download_list(const uint32_t *d_src, const uint32_t *d_dst, uint32_t E,
                                         uint32_t N)
{
    struct EdgeList *el = newEdgeList(E);
    if (E && (og_hip_check(hipMemcpyAsync(el->edges_array_src, d_src, (size_t)E * 4, hipMemcpyDeviceToHost, og_stream()), "D2H") ||
              og_hip_check(hipMemcpyAsync(el->edges_array_dest, d_dst, (size_t)E * 4, hipMemcpyDeviceToHost, og_stream()), "D2H")))
        og_die("download edge list");
    el->num_vertices = N;
    el->avg_degree = N ? E / N : 0;
    og_identity_labels(el);
    return el;
}

static void og_download_vertex(struct Vertex *v, const uint32_t *od, const uint32_t *id,
                               const uint32_t *ix, uint32_t N)
{
    if (og_hip_check(hipMemcpyAsync(v->out_degree, od, (size_t)N * 4, hipMemcpyDeviceToHost, og_stream()), "D2H") ||
        og_hip_check(hipMemcpyAsync(v->in_degree, id, (size_t)N * 4, hipMemcpyDeviceToHost, og_stream()), "D2H") ||
        og_hip_check(hipMemcpyAsync(v->edges_idx, ix, (size_t)N * 4, hipMemcpyDeviceToHost, og_stream()), "D2H"))
        og_die("download vertices");
}

int ogpu_graph_to_host(const ogpu_graph *gg, struct GraphCSR **out)
{
    const og_graph *g = (const og_graph *)gg;
    struct GraphCSR *h = graphCSRNew(g->N, g->E, 1);
    h->sorted_edges_array = og_download_list(g->f_src, g->f_dst, g->E, g->N);
    h->inverse_sorted_edges_array = og_download_list(g->i_src, g->i_dst, g->E, g->N);
    og_download_vertex(h->vertices, g->f_out_degree, g->f_in_degree, g->f_edges_idx, g->N);
    og_download_vertex(h->inverse_vertices, g->i_out_degree, g->i_in_degree, g->i_edges_idx, g->N);
    if (og_hip_check(hipStreamSynchronize(og_stream()), "sync download")) return -1;
    *out = h;
    return 0;
}

struct GraphCSR *graphCSRPreProcessingStep(struct Arguments *arguments)
{
    /* graphCSR.c:164-248 */
    const uint32_t lm[3] = {arguments->lmode, arguments->lmode_l2, arguments->lmode_l3};
    int any_lmode = 0;
    for (int i = 0; i < 3; i++)
    {
        if (lm[i] > 5)
        {
            fprintf(stderr, "opengraph_amd: -l/-L/-O %u: only the degree (1, 2, 3) and DBG (4, 5) "
                            "relabellings are built\n", lm[i]);
            exit(EXIT_FAILURE);
        }
        any_lmode |= lm[i] != 0;
    }
    if (arguments->dflag || arguments->mmode)
    {
        fprintf(stderr, "opengraph_amd: duplicate removal (-k) and masking (-M) are not supported by "
                        "this CSR builder\n");
        exit(EXIT_FAILURE);
    }
    if (og_ctx_ensure()) og_die("device");
    hipStream_t s = og_stream();
    double t0 = og_wall();
    uint32_t n, stride;
    size_t len;
    int fd;
    const uint32_t *buf = og_map_bin(arguments->fnameb, arguments->weighted, &n, &stride, &len, &fd);
    uint32_t E = arguments->symmetric ? 2 * n : n;
    uint32_t *d_pairs = og_dalloc(len ? len : 4);
    uint32_t *d_src = og_dalloc((size_t)E * 4), *d_dst = og_dalloc((size_t)E * 4);
    if (!d_pairs || !d_src || !d_dst) og_die("graphCSRPreProcessingStep");
    if (len && (og_hip_check(hipMemcpyAsync(d_pairs, buf, (size_t)n * stride * 4, hipMemcpyHostToDevice, s), "H2D edges") ||
                og_hip_check(hipStreamSynchronize(s), "sync")))
        og_die("graphCSRPreProcessingStep");
    if (buf) munmap((void *)buf, len);
    close(fd);
    if (og_launch_deinterleave(d_pairs, n, stride, d_src, d_dst, s)) og_die("graphCSRPreProcessingStep");
    if (arguments->symmetric)
    {
        /* edgeList.c:445-449: edge i + n is edge i reversed */
        if (og_hip_check(hipMemcpyAsync(d_src + n, d_dst, (size_t)n * 4, hipMemcpyDeviceToDevice, s), "D2D") ||
            og_hip_check(hipMemcpyAsync(d_dst + n, d_src, (size_t)n * 4, hipMemcpyDeviceToDevice, s), "D2D"))
            og_die("graphCSRPreProcessingStep");
    }
    if (og_hip_check(hipStreamSynchronize(s), "sync")) og_die("graphCSRPreProcessingStep");
    og_dfree(d_pairs);
    og_print_time("Read Edge List From File (Seconds)", og_wall() - t0);

    t0 = og_wall();
    og_graph *g = NULL;
    uint32_t *d_labels = NULL;
    if (any_lmode)
    {
        /* graphCSR.c:184-203: the -l, -L, -O relabellings in turn, then the
         * CSR of the relabelled list with the input's vertex count */
        double t1 = og_wall();
        uint32_t N0 = 0;
        uint32_t *dmax = og_dalloc(4);
        if (!dmax || og_launch_max_id(d_src, d_dst, E, dmax, s) ||
            og_hip_check(hipMemcpyAsync(&N0, dmax, 4, hipMemcpyDeviceToHost, s), "D2H") ||
            og_hip_check(hipStreamSynchronize(s), "sync"))
            og_die("graphCSRPreProcessingStep");
        og_dfree(dmax);
        N0 = E ? N0 + 1 : 1;
        if (og_graph_build_reordered_inplace(d_src, d_dst, E, N0, lm, 3, &d_labels))
            og_die("graphCSRPreProcessingStep");
        og_print_time("Total Reorder Complete (Seconds)", og_wall() - t1);
        if (og_graph_build_device_n(d_src, d_dst, E, N0, arguments->sort, &g)) og_die("graphCSRPreProcessingStep");
    }
    else if (og_graph_build_device(d_src, d_dst, E, arguments->sort, 0, &g))
        og_die("graphCSRPreProcessingStep");
    og_dfree(d_src);
    og_dfree(d_dst);
    static const char *msg[4] = {"Radix Sort Edges By Source (Seconds)",
                                 "Radix Sort Edges By Source/Dest (Seconds)",
                                 "Count Sort Edges By Source (Seconds)",
                                 "Count Sort Edges By Source/Dest (Seconds)"};
    og_print_time(msg[arguments->sort <= 3 ? arguments->sort : 0], og_wall() - t0);
    t0 = og_wall();
    struct GraphCSR *h = NULL;
    if (ogpu_graph_to_host((ogpu_graph *)g, &h)) og_die("graphCSRPreProcessingStep");
    og_graph_destroy(g);
    if (d_labels)
    {
        /* reorder.c:368-373: the edge lists' label arrays carry the new
         * label of every input vertex (readEdgeListsMem copies them to the
         * inverse list, edgeList.c:718-727) */
        struct EdgeList *fe = h->sorted_edges_array, *ie = h->inverse_sorted_edges_array;
        if (og_hip_check(og_copy(fe->label_array, d_labels, (size_t)h->num_vertices * 4, hipMemcpyDeviceToHost),
                         "D2H labels"))
            og_die("graphCSRPreProcessingStep");
        og_dfree(d_labels);
        for (uint32_t v = 0; v < h->num_vertices; v++) fe->inverse_label_array[fe->label_array[v]] = v;
        memcpy(ie->label_array, fe->label_array, (size_t)h->num_vertices * 4);
        memcpy(ie->inverse_label_array, fe->inverse_label_array, (size_t)h->num_vertices * 4);
    }
    og_print_time("Mappign Vertices to CSR (Seconds)", og_wall() - t0);
    og_trace_entry("graphCSRPreProcessingStep", h->num_vertices, h->num_edges);
    graphCSRPrint(h);
    return h;
}

/* --------------------------- .csr files --------------------------- */

void writeToBinFileGraphCSR(const char *fname, struct GraphCSR *g)
{
    /* graphCSR.c:380-449: <fname>.csr, E, N, 6 words per vertex, 4 per edge */
    const uint32_t N = g->num_vertices, E = g->num_edges;
    const int inv = g->inverse_vertices != NULL;
    const size_t per_v = inv ? 6 : 3, per_e = inv ? 4 : 2;
    const size_t words = 2 + per_v * N + per_e * E;
    uint32_t *h = (uint32_t *)og_host_alloc(words * 4);
    h[0] = E;
    h[1] = N;
    uint32_t *o = h + 2;
    for (uint32_t v = 0; v < N; v++, o += per_v)
    {
        o[0] = g->vertices->out_degree[v]; o[1] = g->vertices->in_degree[v]; o[2] = g->vertices->edges_idx[v];
        if (inv)
        {
            o[3] = g->inverse_vertices->out_degree[v];
            o[4] = g->inverse_vertices->in_degree[v];
            o[5] = g->inverse_vertices->edges_idx[v];
        }
    }
    for (uint32_t e = 0; e < E; e++, o += per_e)
    {
        o[0] = g->sorted_edges_array->edges_array_src[e];
        o[1] = g->sorted_edges_array->edges_array_dest[e];
        if (inv)
        {
            o[2] = g->inverse_sorted_edges_array->edges_array_src[e];
            o[3] = g->inverse_sorted_edges_array->edges_array_dest[e];
        }
    }
    char *path = (char *)og_host_alloc(strlen(fname) + 5);
    strcpy(path, fname);
    strcat(path, ".csr");
    FILE *f = fopen(path, "wb");
    if (!f) { fprintf(stderr, "open: %s: %s\n", path, strerror(errno)); exit(1); }
    size_t wrote = fwrite(h, 4, words, f);
    if (fclose(f) != 0 || wrote != words) { fprintf(stderr, "write: %s\n", path); exit(1); }
    free(path);
    free(h);
}

struct GraphCSR *readFromBinFileGraphCSR(const char *fname)
{
    /* graphCSR.c:452-558 */
    FILE *f = fopen(fname, "rb");
    if (!f) { fprintf(stderr, "open: %s: %s\n", fname, strerror(errno)); exit(1); }
    uint32_t hdr[2] = {0, 0};
    size_t ret = fread(hdr, 4, 2, f);
    const uint32_t E = hdr[0], N = hdr[1];
    struct GraphCSR *g = graphCSRNew(N, E, 1);
    g->sorted_edges_array = newEdgeList(E);
    g->sorted_edges_array->num_vertices = N;
    g->inverse_sorted_edges_array = newEdgeList(E);
    g->inverse_sorted_edges_array->num_vertices = N;
    const size_t words = 6 * (size_t)N + 4 * (size_t)E;
    uint32_t *h = (uint32_t *)og_host_alloc(words * 4 + 4);
    ret = words ? fread(h, 4, words, f) : ret;
    fclose(f);
    if (ret != (words ? words : 2))
    {
        fprintf(stderr, "read: %s: truncated .csr file\n", fname);
        exit(1);
    }
    const uint32_t *o = h;
    for (uint32_t v = 0; v < N; v++, o += 6)
    {
        g->vertices->out_degree[v] = o[0]; g->vertices->in_degree[v] = o[1]; g->vertices->edges_idx[v] = o[2];
        g->inverse_vertices->out_degree[v] = o[3];
        g->inverse_vertices->in_degree[v] = o[4];
        g->inverse_vertices->edges_idx[v] = o[5];
    }
    for (uint32_t e = 0; e < E; e++, o += 4)
    {
        g->sorted_edges_array->edges_array_src[e] = o[0];
        g->sorted_edges_array->edges_array_dest[e] = o[1];
        g->inverse_sorted_edges_array->edges_array_src[e] = o[2];
        g->inverse_sorted_edges_array->edges_array_dest[e] = o[3];
    }
    free(h);
    graphCSRPrint(g);
    return g;
}

/* ------------------------ rank post-sort ------------------------ */

uint32_t *radixSortEdgesByPageRank(float *pageRanks, uint32_t *labels, uint32_t n)
{
    /* reorder.c:165-221: stable ascending by FloatToFixed32SORT(rank) */
    if (n < 2) return labels;
    if (og_ctx_ensure()) og_die("radixSortEdgesByPageRank");
    hipStream_t s = og_stream();
    size_t b = (size_t)n * 4;
    float *dr = og_dalloc(b);
    uint32_t *k0 = og_dalloc(b), *v0 = og_dalloc(b), *k1 = og_dalloc(b), *v1 = og_dalloc(b);
    uint32_t *k2 = og_dalloc(b), *v2 = og_dalloc(b);
    void *scr = og_dalloc(og_radix_scratch_bytes(n));
    int in_b = 0;
    if (!dr || !k0 || !v0 || !k1 || !v1 || !k2 || !v2 || !scr ||
        og_hip_check(hipMemcpyAsync(dr, pageRanks, b, hipMemcpyHostToDevice, s), "H2D") ||
        og_hip_check(hipMemcpyAsync(v0, labels, b, hipMemcpyHostToDevice, s), "H2D") ||
        og_launch_rank_keys(dr, n, k0, s) ||
        og_radix_sort_pairs(k0, v0, k1, v1, k2, v2, n, 32, scr, s, &in_b) ||
        og_hip_check(hipMemcpyAsync(labels, in_b ? v2 : v1, b, hipMemcpyDeviceToHost, s), "D2H") ||
        og_hip_check(hipStreamSynchronize(s), "sync"))
        og_die("radixSortEdgesByPageRank");
    og_dfree(dr); og_dfree(k0); og_dfree(v0); og_dfree(k1); og_dfree(v1); og_dfree(k2); og_dfree(v2);
    og_dfree(scr);
    return labels;
}

/* ----------------------- degree relabelling ----------------------- */

/* Upload n u32 to a new device buffer (dies on failure). */
static uint32_t *og_up(const uint32_t *h, uint32_t n)
{
    uint32_t *d = og_dalloc((size_t)(n ? n : 1) * 4);
    if (!d || (n && og_hip_check(og_copy(d, h, (size_t)n * 4, hipMemcpyHostToDevice), "H2D")))
        og_die("upload");
    return d;
}

static void og_down(uint32_t *h, const uint32_t *d, uint32_t n)
{
    if (n && og_hip_check(og_copy(h, d, (size_t)n * 4, hipMemcpyDeviceToHost), "D2H")) og_die("download");
}

static void og_reorder_banner(const char *what)
{
    printf(" -----------------------------------------------------\n");
    printf("| %-51s | \n", what);
    printf(" -----------------------------------------------------\n");
}

uint32_t *reorderGraphGenerateInOutDegrees(uint32_t *degrees, struct EdgeList *el, uint32_t lmode)
{
    /* reorder.c:1483-1554: the counts are added to degrees (the callers
     * pass it zeroed).  lmode 10 (random keys drawn per OpenMP thread) has
     * no thread-independent result and is not built. */
    if (lmode == 10)
    {
        fprintf(stderr, "opengraph_amd: -l 10 (random relabelling) is not supported\n");
        exit(EXIT_FAILURE);
    }
    if (og_ctx_ensure()) og_die("reorderGraphGenerateInOutDegrees");
    const uint32_t N = el->num_vertices, E = el->num_edges;
    const uint32_t count_src = !(lmode == 2 || lmode == 5 || lmode == 7 || lmode == 9);
    const uint32_t count_dst = lmode == 2 || lmode == 3 || lmode == 5 || lmode == 7 || lmode == 9;
    uint32_t *ds = og_up(el->edges_array_src, E), *dd = og_up(el->edges_array_dest, E);
    uint32_t *deg = og_up(degrees, N);
    if (og_launch_reorder_degrees(ds, dd, E, count_src, count_dst, deg, og_stream()) ||
        og_hip_check(hipStreamSynchronize(og_stream()), "sync"))
        og_die("reorderGraphGenerateInOutDegrees");
    og_down(degrees, deg, N);
    og_dfree(ds); og_dfree(dd); og_dfree(deg);
    return degrees;
}

uint32_t *radixSortEdgesByDegree(uint32_t *degrees, uint32_t *labels, uint32_t n)
{
    /* reorder.c:223-271: stable ascending sort of (degree, label) pairs;
     * both arrays come back sorted */
    if (n < 2) return labels;
    if (og_ctx_ensure()) og_die("radixSortEdgesByDegree");
    hipStream_t s = og_stream();
    size_t b = (size_t)n * 4;
    uint32_t *k0 = og_up(degrees, n), *v0 = og_up(labels, n);
    uint32_t *k1 = og_dalloc(b), *v1 = og_dalloc(b), *k2 = og_dalloc(b), *v2 = og_dalloc(b);
    void *scr = og_dalloc(og_radix_scratch_bytes(n));
    int in_b = 0;
    if (!k1 || !v1 || !k2 || !v2 || !scr || og_radix_sort_pairs(k0, v0, k1, v1, k2, v2, n, 32, scr, s, &in_b) ||
        og_hip_check(hipStreamSynchronize(s), "sync"))
        og_die("radixSortEdgesByDegree");
    og_down(degrees, in_b ? k2 : k1, n);
    og_down(labels, in_b ? v2 : v1, n);
    og_dfree(k0); og_dfree(v0); og_dfree(k1); og_dfree(v1); og_dfree(k2); og_dfree(v2); og_dfree(scr);
    return labels;
}

struct EdgeList *relabelEdgeList(struct EdgeList *el, uint32_t *labels)
{
    /* reorder.c:1618-1637 */
    if (og_ctx_ensure()) og_die("relabelEdgeList");
    const uint32_t E = el->num_edges;
    uint32_t *ds = og_up(el->edges_array_src, E), *dd = og_up(el->edges_array_dest, E);
    uint32_t *dl = og_up(labels, el->num_vertices);
    if (og_launch_relabel(ds, dd, E, dl, og_stream()) || og_hip_check(hipStreamSynchronize(og_stream()), "sync"))
        og_die("relabelEdgeList");
    og_down(el->edges_array_src, ds, E);
    og_down(el->edges_array_dest, dd, E);
    og_dfree(ds); og_dfree(dd); og_dfree(dl);
    return el;
}

/* reorder.c:368-373 / :547-552: carry the new labels into the edge list's
 * label arrays. */
static void og_update_label_arrays(struct EdgeList *el, const uint32_t *labels)
{
    if (!el->label_array || !el->inverse_label_array) return;
    for (uint32_t v = 0; v < el->num_vertices; v++)
    {
        el->label_array[v] = labels[el->label_array[v]];
        el->inverse_label_array[el->label_array[v]] = v;
    }
}

struct EdgeList *reorderGraphListDegree(struct EdgeList *el, uint32_t *degrees, uint32_t lmode)
{
    /* reorder.c:296-384 */
    static const char *names[4] = {"OUT-DEGREE", "OUT-DEGREE", "IN-DEGREE", "(IN+OUT)-DEGREE"};
    const uint32_t N = el->num_vertices;
    og_reorder_banner("Starting Degree Reordering/Relabeling");
    printf("| %-51s | \n", lmode <= 3 ? names[lmode] : "OUT-DEGREE");
    double t0 = og_wall();
    uint32_t *inv = (uint32_t *)og_host_alloc((size_t)(N ? N : 1) * 4);
    uint32_t *labels = (uint32_t *)og_host_alloc((size_t)(N ? N : 1) * 4);
    for (uint32_t v = 0; v < N; v++) inv[v] = v;
    inv = radixSortEdgesByDegree(degrees, inv, N);
    for (uint32_t v = 0; v < N; v++) labels[inv[v]] = N - 1 - v;
    og_reorder_banner("Reordering Complete");
    printf("| %-51f | \n", og_wall() - t0);
    t0 = og_wall();
    el = relabelEdgeList(el, labels);
    og_reorder_banner("Relabeling Complete");
    printf("| %-51f | \n", og_wall() - t0);
    printf(" -----------------------------------------------------\n");
    og_update_label_arrays(el, labels);
    free(inv);
    free(labels);
    return el;
}

struct EdgeList *reorderGraphProcessDegree(uint32_t sort, struct EdgeList *el, uint32_t lmode)
{
    /* reorder.c:275-294 */
    (void)sort;
    uint32_t *degrees = (uint32_t *)og_host_alloc((size_t)(el->num_vertices ? el->num_vertices : 1) * 4);
    memset(degrees, 0, (size_t)el->num_vertices * 4);
    degrees = reorderGraphGenerateInOutDegrees(degrees, el, lmode);
    el = reorderGraphListDegree(el, degrees, lmode);
    free(degrees);
    return el;
}

struct EdgeList *reorderGraphListDBG(struct EdgeList *el, uint32_t *degrees, uint32_t *thresholds,
                                     uint32_t num_buckets, uint32_t lmode)
{
    /* reorder.c:439-572: buckets from the last to the first, each in id order */
    (void)lmode;
    if (num_buckets < 1 || num_buckets > 16)
    {
        fprintf(stderr, "opengraph_amd: DBG with %u buckets (1..16 supported)\n", num_buckets);
        exit(EXIT_FAILURE);
    }
    if (og_ctx_ensure()) og_die("reorderGraphListDBG");
    const uint32_t N = el->num_vertices;
    og_dbg_thresholds th;
    memset(&th, 0, sizeof(th));
    th.n = num_buckets;
    memcpy(th.t, thresholds, (size_t)num_buckets * 4);
    double t0 = og_wall();
    uint32_t *dd = og_up(degrees, N), *dl = og_dalloc((size_t)(N ? N : 1) * 4);
    uint32_t *labels = (uint32_t *)og_host_alloc((size_t)(N ? N : 1) * 4);
    if (!dl || (N && og_reorder_labels_from_degrees(dd, N, 1, th, dl))) og_die("reorderGraphListDBG");
    og_down(labels, dl, N);
    og_dfree(dd); og_dfree(dl);
    og_reorder_banner("Reordering Complete");
    printf("| %-51f | \n", og_wall() - t0);
    t0 = og_wall();
    el = relabelEdgeList(el, labels);
    og_reorder_banner("Relabeling Complete");
    printf("| %-51f | \n", og_wall() - t0);
    printf(" -----------------------------------------------------\n");
    og_update_label_arrays(el, labels);
    free(labels);
    return el;
}

struct EdgeList *reorderGraphProcessDBG(uint32_t sort, struct EdgeList *el, uint32_t lmode)
{
    /* reorder.c:387-437 */
    (void)sort;
    og_dbg_thresholds th = og_dbg_default_thresholds(el->avg_degree);
    printf("| %-51s | \n", lmode == 5 ? "DBG IN-DEGREE" : "DBG OUT-DEGREE");
    uint32_t *degrees = (uint32_t *)og_host_alloc((size_t)(el->num_vertices ? el->num_vertices : 1) * 4);
    memset(degrees, 0, (size_t)el->num_vertices * 4);
    degrees = reorderGraphGenerateInOutDegrees(degrees, el, lmode);
    el = reorderGraphListDBG(el, degrees, th.t, th.n, lmode);
    free(degrees);
    return el;
}

struct EdgeList *reorderGraphProcess(struct EdgeList *el, struct Arguments *arguments)
{
    /* reorder.c:1556-1616 for the degree and DBG relabellings */
    og_reorder_banner("Reorder Process");
    double t0 = og_wall();
    switch (arguments->lmode)
    {
    case 0: case 1: case 2: case 3:
        el = reorderGraphProcessDegree(arguments->sort, el, arguments->lmode);
        break;
    case 4: case 5:
        el = reorderGraphProcessDBG(arguments->sort, el, arguments->lmode);
        break;
    default:
        fprintf(stderr, "opengraph_amd: -l %u: only the degree (1, 2, 3) and DBG (4, 5) relabellings "
                        "are built\n", arguments->lmode);
        exit(EXIT_FAILURE);
    }
    og_reorder_banner("Total Reorder Complete");
    printf("| %-51f | \n", og_wall() - t0);
    printf(" -----------------------------------------------------\n");
    return el;
}
