/*
 * og_graph.c -- device-resident CSR (forward + inverse, the reference's
 * GraphCSR content) and the PageRank plan built from it.
 *
 * The CSR build follows graphCSRPreProcessingStep (graphCSR.c:164-248):
 * N = max id + 1, stable sort of the forward list (sortRun.c:26-69), vertex
 * mapping (vertex.c:204-325), inverse list = swapped sorted forward list
 * sorted again, mapping.  Because the forward list is already sorted by
 * source, the inverse list's first (by-destination) pass of sort modes 1/3
 * is the identity and only the by-source sort of (dst, src) runs; the
 * result is the same stable order.
 */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "og_internal.h"
#include "og_plan.h"
#include "og_pr.h"
#include "opengraph_amd_ext.h"

static uint32_t og_bits_for(uint32_t maxval)
{
    uint32_t b = 0;
    while (b < 32 && (maxval >> b)) b++;
    return b ? b : 1;
}

void og_graph_destroy(og_graph *g)
{
    if (!g) return;
    og_dfree(g->f_src); og_dfree(g->f_dst); og_dfree(g->f_out_degree);
    og_dfree(g->f_in_degree); og_dfree(g->f_edges_idx);
    og_dfree(g->i_src); og_dfree(g->i_dst); og_dfree(g->i_out_degree);
    og_dfree(g->i_in_degree); og_dfree(g->i_edges_idx);
    free(g);
}

static og_graph *og_graph_alloc(uint32_t N, uint32_t E)
{
    og_graph *g = (og_graph *)calloc(1, sizeof(og_graph));
    if (!g) { og_set_error("out of host memory"); return NULL; }
    g->N = N;
    g->E = E;
    size_t eb = (size_t)E * 4, nb = (size_t)N * 4;
    g->f_src = og_dalloc(eb); g->f_dst = og_dalloc(eb);
    g->i_src = og_dalloc(eb); g->i_dst = og_dalloc(eb);
    g->f_out_degree = og_dalloc(nb); g->f_in_degree = og_dalloc(nb); g->f_edges_idx = og_dalloc(nb);
    g->i_out_degree = og_dalloc(nb); g->i_in_degree = og_dalloc(nb); g->i_edges_idx = og_dalloc(nb);
    if (!g->f_src || !g->f_dst || !g->i_src || !g->i_dst || !g->f_out_degree || !g->f_in_degree ||
        !g->f_edges_idx || !g->i_out_degree || !g->i_in_degree || !g->i_edges_idx)
    {
        og_graph_destroy(g);
        return NULL;
    }
    return g;
}

/* Sort (k_in, v_in) stably by key into the pool pair (*k_res, *v_res): the
 * two pairs (ka, va), (kb, vb) are the ping-pong buffers and must not alias
 * the inputs; the pair not holding the result is returned as spare. */
static int og_sort_into(const uint32_t *k_in, const uint32_t *v_in, uint32_t **ka, uint32_t **va,
                        uint32_t **kb, uint32_t **vb, uint32_t n, uint32_t key_bits, void *scratch)
{
    int in_b = 0;
    if (og_radix_sort_pairs(k_in, v_in, *ka, *va, *kb, *vb, n, key_bits, scratch, og_stream(), &in_b))
        return -1;
    if (in_b)
    {
        uint32_t *t = *ka; *ka = *kb; *kb = t;
        t = *va; *va = *vb; *vb = t;
    }
    return 0;   /* result now in (*ka, *va) */
}

int og_graph_build_device(const uint32_t *d_src, const uint32_t *d_dst, uint32_t E,
                          uint32_t sort_mode, int keep_inputs, og_graph **out)
{
    (void)keep_inputs;
    return og_graph_build_device_n(d_src, d_dst, E, 0, sort_mode, out);
}

int og_graph_build_device_n(const uint32_t *d_src, const uint32_t *d_dst, uint32_t E,
                            uint32_t min_vertices, uint32_t sort_mode, og_graph **out)
{
    *out = NULL;
    if (og_ctx_ensure()) return -1;
    if (sort_mode > 3) sort_mode = 0;     /* sortRun.c:60-66: default is radix by source */
    hipStream_t s = og_stream();
    uint32_t *d_max = og_dalloc(4);
    if (!d_max) return -1;
    uint32_t maxid = 0;
    if (og_launch_max_id(d_src, d_dst, E, d_max, s) ||
        og_hip_check(hipMemcpyAsync(&maxid, d_max, 4, hipMemcpyDeviceToHost, s), "copy max id") ||
        og_hip_check(hipStreamSynchronize(s), "sync max id"))
    {
        og_dfree(d_max);
        return -1;
    }
    og_dfree(d_max);
    if (E && maxid > OG_ID_MASK)
    {
        og_set_error("vertex id %u does not fit the 30-bit id space (reorder.h:18-27)", maxid);
        return -1;
    }
    uint32_t N = E ? maxid + 1 : 1;
    if (N < min_vertices) N = min_vertices;
    if (N - 1 > OG_ID_MASK)
    {
        og_set_error("%u vertices do not fit the 30-bit id space (reorder.h:18-27)", N);
        return -1;
    }
    og_graph *g = og_graph_alloc(N, E);
    if (!g) return -1;
    g->sort_mode = sort_mode;
    uint32_t kb_bits = og_bits_for(N - 1);
    size_t eb = (size_t)E * 4;
    uint32_t *t0 = og_dalloc(eb), *t1 = og_dalloc(eb);
    void *scratch = og_dalloc(og_radix_scratch_bytes(E));
    int rc = -1;
    if (!t0 || !t1 || !scratch) goto done;

    if (E)
    {
        /* forward list */
        uint32_t *ka = g->f_src, *va = g->f_dst, *kb = t0, *vb = t1;
        if (sort_mode == 1 || sort_mode == 3)
        {
            /* by destination, then by source (radixsort.c:282-335) */
            if (og_sort_into(d_dst, d_src, &ka, &va, &kb, &vb, E, kb_bits, scratch)) goto done;
            /* result (dst, src) in (ka, va); sort (src, dst) by src into the other pair */
            uint32_t *k2 = kb, *v2 = vb, *k3 = ka, *v3 = va;
            if (og_sort_into(va, ka, &k2, &v2, &k3, &v3, E, kb_bits, scratch)) goto done;
            ka = k2; va = v2; kb = k3; vb = v3;
        }
        else
        {
            if (og_sort_into(d_src, d_dst, &ka, &va, &kb, &vb, E, kb_bits, scratch)) goto done;
        }
        /* make the graph own the result pair, keep the other as temp */
        g->f_src = ka; g->f_dst = va; t0 = kb; t1 = vb;

        /* inverse list: (dst, src) of the sorted forward list sorted by dst */
        uint32_t *ia = g->i_src, *iva = g->i_dst, *ib = t0, *ivb = t1;
        if (og_sort_into(g->f_dst, g->f_src, &ia, &iva, &ib, &ivb, E, kb_bits, scratch)) goto done;
        g->i_src = ia; g->i_dst = iva; t0 = ib; t1 = ivb;
    }
    if (og_launch_map_vertices(g->f_src, E, N, g->f_out_degree, g->f_edges_idx, s)) goto done;
    if (og_launch_map_vertices(g->i_src, E, N, g->i_out_degree, g->i_edges_idx, s)) goto done;
    /* vertex.c:291-310 */
    if (og_hip_check(hipMemcpyAsync(g->f_in_degree, g->i_out_degree, (size_t)N * 4,
                                    hipMemcpyDeviceToDevice, s), "copy in_degree") ||
        og_hip_check(hipMemcpyAsync(g->i_in_degree, g->f_out_degree, (size_t)N * 4,
                                    hipMemcpyDeviceToDevice, s), "copy in_degree") ||
        og_hip_check(hipStreamSynchronize(s), "sync csr build"))
        goto done;
    rc = 0;
done:
    og_dfree(t0);
    og_dfree(t1);
    og_dfree(scratch);
    if (rc) og_graph_destroy(g);
    else *out = g;
    return rc;
}

/* ----------------------------- plan ----------------------------- */

void og_plan_destroy(og_plan *p)
{
    if (!p) return;
    og_dfree(p->new2old); og_dfree(p->old2new); og_dfree(p->row_ptr); og_dfree(p->cols);
    og_dfree(p->outdeg); og_dfree(p->chunk_base); og_dfree(p->ch_row); og_dfree(p->ch_edge);
    og_dfree(p->push_src); og_dfree(p->push_dst); og_dfree(p->f_row_ptr); og_dfree(p->f_cols);
    free(p);
}

int og_plan_build(const og_graph *g, uint32_t relabel, uint32_t seq_t, uint32_t chain_t, int need_forward,
                  og_plan **out)
{
    /* chains are class-0 rows at the front, which needs the relabelling */
    if (!relabel) chain_t = 0;
    if (chain_t && chain_t < OG_CH) chain_t = OG_CH;
    *out = NULL;
    if (og_ctx_ensure()) return -1;
    hipStream_t s = og_stream();
    const uint32_t N = g->N, E = g->E;
    og_plan *p = (og_plan *)calloc(1, sizeof(og_plan));
    if (!p) { og_set_error("out of host memory"); return -1; }
    p->N = N;
    p->E = E;
    p->relabel = relabel;
    p->seq_t = seq_t;
    size_t nb = (size_t)N * 4;
    uint32_t *key = og_dalloc(nb), *key2 = og_dalloc(nb), *val2 = og_dalloc(nb);
    uint32_t *indeg_new = og_dalloc(nb + 4), *d_cnt = og_dalloc(32);
    size_t scr = og_radix_scratch_bytes(N);
    size_t scan_scr = og_scan_scratch_bytes(N + 1);
    void *scratch = og_dalloc(scr > scan_scr ? scr : scan_scr);
    p->new2old = og_dalloc(nb);
    p->old2new = og_dalloc(nb);
    p->outdeg = og_dalloc(nb);
    p->row_ptr = og_dalloc(nb + 4);
    p->cols = og_dalloc((size_t)E * 4);
    int rc = -1;
    if (!key || !key2 || !val2 || !indeg_new || !d_cnt || !scratch || !p->new2old || !p->old2new ||
        !p->outdeg || !p->row_ptr || !p->cols)
        goto done;

    /* 1. new order: by row class, then out-degree (in-degree for the rows
     *    one thread sums) descending, stable */
    if (og_launch_plan_keys(N, g->f_out_degree, g->i_out_degree, seq_t, chain_t, relabel, key, p->new2old, s))
        goto done;
    {
        uint32_t *ka = key2, *va = val2, *kb = key, *vb = p->new2old;
        int in_b = 0;
        if (og_radix_sort_pairs(key, p->new2old, ka, va, kb, vb, N, relabel ? 32 : 2, scratch, s, &in_b))
            goto done;
        if (!in_b)
        {
            /* result in (key2, val2): swap so new2old holds the permutation */
            uint32_t *t = p->new2old; p->new2old = val2; val2 = t;
        }
    }
    /* 2. inverse permutation, degrees in new order, class sizes */
    if (og_launch_fill_u32(d_cnt, 0, 5, s)) goto done;
    if (og_launch_plan_perm(N, p->new2old, g->f_out_degree, g->i_out_degree, seq_t, chain_t, p->old2new,
                            p->outdeg, indeg_new, d_cnt, s))
        goto done;
    /* 3. row pointers of the relabelled inverse CSR */
    if (og_scan_exclusive(indeg_new, p->row_ptr, N, scratch, s)) goto done;
    uint32_t cls[5] = {0, 0, 0, 0, 0};
    if (og_hip_check(hipMemcpyAsync(cls, d_cnt, 20, hipMemcpyDeviceToHost, s), "copy class sizes") ||
        og_hip_check(hipStreamSynchronize(s), "sync plan"))
        goto done;
    const uint32_t H = cls[0], C = H + cls[1];
    p->H = H;
    p->C = C;
    p->Z = C + cls[2];
    p->L = cls[4];
    /* 4. relabelled in-neighbour lists, reference order kept */
    if (E && og_launch_plan_copy(H, N, p->new2old, p->old2new, g->i_edges_idx, g->i_dst, p->row_ptr,
                                 p->cols, s))
        goto done;
    /* 4b. rows [L, H) (several chunks; their sums are a fixed tree, not the
     *     CSR order): in-neighbours sorted by source id, so the 64 lanes of
     *     a chunk gather neighbouring contributions -- with the degree
     *     order, often the same cache line.  relabel 2 keeps CSR order, and
     *     so do the chains [0, L).  Edges [EL, EH) are sorted in place. */
    if (relabel == 1 && H > p->L)
    {
        const uint32_t L = p->L;
        uint32_t EH = 0, EL = 0;
        if (og_hip_check(hipMemcpyAsync(&EH, p->row_ptr + H, 4, hipMemcpyDeviceToHost, s), "copy") ||
            og_hip_check(hipMemcpyAsync(&EL, p->row_ptr + L, 4, hipMemcpyDeviceToHost, s), "copy") ||
            og_hip_check(hipStreamSynchronize(s), "sync plan"))
            goto done;
        uint32_t *cols_h = p->cols + EL;
        EH -= EL;
        uint32_t *rid = og_dalloc((size_t)EH * 4 + 4), *ka = og_dalloc((size_t)EH * 4 + 4),
                 *va = og_dalloc((size_t)EH * 4 + 4), *kb = og_dalloc((size_t)EH * 4 + 4),
                 *vb = og_dalloc((size_t)EH * 4 + 4);
        void *rscr = og_dalloc(og_radix_scratch_bytes(EH));
        int bad = !rid || !ka || !va || !kb || !vb || !rscr;
        uint32_t nbits = 1, hbits = 1;
        while (nbits < 32 && (1ull << nbits) < N) nbits++;
        while (hbits < 32 && (1ull << hbits) < H - L) hbits++;
        int in_b = 0;
        if (!bad)
        {
            /* by source (stable), then by row (stable): (row, source) order;
             * row ids are relative to L, at edge offsets relative to EL */
            bad = og_launch_plan_row_ids_at(H - L, p->row_ptr + L, EL, rid, s) ||
                  og_radix_sort_pairs(cols_h, rid, ka, va, kb, vb, EH, nbits, rscr, s, &in_b);
        }
        if (!bad)
        {
            uint32_t *k1 = in_b ? kb : ka, *v1 = in_b ? vb : va;
            /* second sort reads (row, source) from the first's result; its
             * outputs alternate (ka|kb not holding that result) and rid/cols */
            uint32_t *oa_k = in_b ? ka : kb, *oa_v = in_b ? va : vb;
            int in_b2 = 0;
            bad = og_radix_sort_pairs(v1, k1, oa_k, oa_v, rid, k1, EH, hbits, rscr, s, &in_b2);
            if (!bad)
            {
                const uint32_t *src_sorted = in_b2 ? k1 : oa_v;
                bad = og_hip_check(hipMemcpyAsync(cols_h, src_sorted, (size_t)EH * 4, hipMemcpyDeviceToDevice, s),
                                   "copy sorted rows") ||
                      og_hip_check(hipStreamSynchronize(s), "sync plan");
            }
        }
        og_dfree(rid); og_dfree(ka); og_dfree(va); og_dfree(kb); og_dfree(vb); og_dfree(rscr);
        if (bad) goto done;
    }
    /* 5. chunks of the rows [0, C): key = chunks per row, chunk_first = its prefix */
    {
        uint32_t *first = og_dalloc((size_t)(C + 1) * 4);
        if (!first) goto done;
        uint32_t n_chunks = 0;
        if (og_launch_plan_chunks(C, p->row_ptr, OG_CH, key, s) ||
            og_scan_exclusive(key, first, C, scratch, s) ||
            og_hip_check(hipMemcpyAsync(&n_chunks, first + C, 4, hipMemcpyDeviceToHost, s), "copy") ||
            og_hip_check(hipStreamSynchronize(s), "sync plan"))
        {
            og_dfree(first);
            goto done;
        }
        p->n_chunks = n_chunks;
        if (p->L && og_hip_check(hipMemcpyAsync(&p->ch_first, first + p->L, 4, hipMemcpyDeviceToHost, s),
                                 "copy") == 0)
            og_hip_check(hipStreamSynchronize(s), "sync plan");
        p->chunk_base = og_dalloc((size_t)(H + 1) * 4);
        p->ch_row = og_dalloc((size_t)n_chunks * 4 + 4);
        p->ch_edge = og_dalloc((size_t)(n_chunks + 1) * 4);
        int bad = !p->chunk_base || !p->ch_row || !p->ch_edge ||
                  og_hip_check(hipMemcpyAsync(p->chunk_base, first, (size_t)(H + 1) * 4,
                                              hipMemcpyDeviceToDevice, s), "copy chunk_base") ||
                  og_launch_plan_chunk_desc(H, C, p->row_ptr, first, OG_CH, p->ch_row, p->ch_edge, s) ||
                  og_hip_check(hipStreamSynchronize(s), "sync plan");
        og_dfree(first);
        if (bad) goto done;
        uint32_t waves = (n_chunks - p->ch_first + OG_CH_PER_WAVE - 1) / OG_CH_PER_WAVE;
        p->n_chunk_blocks = (waves + 3) / 4;
        p->n_row_blocks = (uint32_t)(((uint64_t)(N - C) + 255) / 256);
    }
    /* 7. forward edges in new ids (push) and the forward CSR in new ids
     *    (data-driven activation of out-neighbours) */
    if (need_forward)
    {
        p->push_src = og_dalloc((size_t)E * 4);
        p->push_dst = og_dalloc((size_t)E * 4);
        p->f_row_ptr = og_dalloc(nb + 4);
        p->f_cols = og_dalloc((size_t)E * 4);
        if (!p->push_src || !p->push_dst || !p->f_row_ptr || !p->f_cols) goto done;
        if (og_launch_plan_map_pairs(E, g->f_src, g->f_dst, p->old2new, p->push_src, p->push_dst, s))
            goto done;
        if (og_scan_exclusive(p->outdeg, p->f_row_ptr, N, scratch, s)) goto done;
        if (E && og_launch_plan_copy(0, N, p->new2old, p->old2new, g->f_edges_idx, g->f_dst,
                                     p->f_row_ptr, p->f_cols, s))
            goto done;
    }
    if (og_hip_check(hipStreamSynchronize(s), "sync plan")) goto done;
    rc = 0;
done:
    og_dfree(key); og_dfree(key2); og_dfree(val2); og_dfree(indeg_new); og_dfree(d_cnt);
    og_dfree(scratch);
    if (rc) og_plan_destroy(p);
    else *out = p;
    return rc;
}

/* ------------------------- public graph API ------------------------- */

int ogpu_graph_build(const uint32_t *d_src, const uint32_t *d_dst, uint32_t num_edges,
                     uint32_t sort_mode, ogpu_graph **out)
{
    return og_graph_build_device(d_src, d_dst, num_edges, sort_mode, 1, (og_graph **)out);
}

int ogpu_graph_info(const ogpu_graph *gg, uint32_t *num_vertices, uint32_t *num_edges)
{
    const og_graph *g = (const og_graph *)gg;
    if (!g) { og_set_error("null graph"); return -1; }
    if (num_vertices) *num_vertices = g->N;
    if (num_edges) *num_edges = g->E;
    return 0;
}

int ogpu_graph_array(const ogpu_graph *gg, uint32_t which, uint32_t *host_out)
{
    const og_graph *g = (const og_graph *)gg;
    if (!g) { og_set_error("null graph"); return -1; }
    const uint32_t *src;
    size_t n;
    switch (which)
    {
    case OGPU_F_SRC: src = g->f_src; n = g->E; break;
    case OGPU_F_DST: src = g->f_dst; n = g->E; break;
    case OGPU_F_OUT_DEGREE: src = g->f_out_degree; n = g->N; break;
    case OGPU_F_IN_DEGREE: src = g->f_in_degree; n = g->N; break;
    case OGPU_F_EDGES_IDX: src = g->f_edges_idx; n = g->N; break;
    case OGPU_I_SRC: src = g->i_src; n = g->E; break;
    case OGPU_I_DST: src = g->i_dst; n = g->E; break;
    case OGPU_I_OUT_DEGREE: src = g->i_out_degree; n = g->N; break;
    case OGPU_I_IN_DEGREE: src = g->i_in_degree; n = g->N; break;
    case OGPU_I_EDGES_IDX: src = g->i_edges_idx; n = g->N; break;
    default: og_set_error("unknown graph array %u", which); return -1;
    }
    if (n == 0) return 0;
    if (og_hip_check(hipMemcpyAsync(host_out, src, n * 4, hipMemcpyDeviceToHost, og_stream()), "D2H"))
        return -1;
    return og_hip_check(hipStreamSynchronize(og_stream()), "sync");
}

/* A CSR pair is well formed when, for both the forward and the inverse list
 * (graphCSR.c:107-161, vertex.c:204-325): the vertex count fits the 30-bit
 * id space (reorder.h:18-27); every src/dst id is < N; the out-degrees sum
 * to E and vertex v's edges [edges_idx[v], edges_idx[v] + out_degree[v])
 * lie inside [0, E) and all have src == v; and each list's in-degree is the
 * other list's out-degree.  The plan and pull kernels index with these
 * arrays unchecked, so inputs from files or callers are checked here, on
 * the host, before anything is uploaded.  Arrays are read with a stride
 * (words between consecutive entries) so the interleaved .csr image can be
 * checked in place. */
typedef struct
{
    const uint32_t *a;
    size_t stride;
} og_strided;

#define OG_AT(x, i) ((x).a[(size_t)(i) * (x).stride])

static int og_csr_check_list(uint32_t N, uint32_t E, og_strided src, og_strided dst, og_strided deg,
                             og_strided idx, og_strided other_indeg, const char *what)
{
    uint64_t total = 0;
    int64_t bad_v = -1, bad_e = -1;
#pragma omp parallel for reduction(+ : total) schedule(static)
    for (uint32_t v = 0; v < N; v++)
    {
        const uint32_t d = OG_AT(deg, v), b = OG_AT(idx, v);
        total += d;
        int ok = (uint64_t)b + d <= E && OG_AT(other_indeg, v) == d;
        for (uint32_t j = b; ok && j < b + d; j++) ok = OG_AT(src, j) == v;
        if (!ok)
        {
#pragma omp critical
            if (bad_v < 0 || v < bad_v) bad_v = v;
        }
    }
#pragma omp parallel for schedule(static)
    for (uint32_t e = 0; e < E; e++)
        if (OG_AT(src, e) >= N || OG_AT(dst, e) >= N)
        {
#pragma omp critical
            if (bad_e < 0 || e < bad_e) bad_e = e;
        }
    if (bad_e >= 0)
    {
        og_set_error("%s edge %lld: id (%u, %u) not below N = %u", what, (long long)bad_e,
                     OG_AT(src, bad_e), OG_AT(dst, bad_e), N);
        return -1;
    }
    if (bad_v >= 0)
    {
        og_set_error("%s vertex %lld: degree %u at edges_idx %u is not a run of its edges in [0, %u) "
                     "(or the other list's in-degree differs)",
                     what, (long long)bad_v, OG_AT(deg, bad_v), OG_AT(idx, bad_v), E);
        return -1;
    }
    if (total != E)
    {
        og_set_error("%s out-degrees sum to %llu, not E = %u", what, (unsigned long long)total, E);
        return -1;
    }
    return 0;
}

static int og_csr_check(uint32_t N, uint32_t E, const og_strided a[10])
{
    if (N == 0 && E != 0) { og_set_error("a CSR with %u edges and no vertices", E); return -1; }
    if (N && N - 1 > OG_ID_MASK)
    {
        og_set_error("%u vertices do not fit the 30-bit id space (reorder.h:18-27)", N);
        return -1;
    }
    return og_csr_check_list(N, E, a[OGPU_F_SRC], a[OGPU_F_DST], a[OGPU_F_OUT_DEGREE], a[OGPU_F_EDGES_IDX],
                             a[OGPU_I_IN_DEGREE], "forward") ||
           og_csr_check_list(N, E, a[OGPU_I_SRC], a[OGPU_I_DST], a[OGPU_I_OUT_DEGREE], a[OGPU_I_EDGES_IDX],
                             a[OGPU_F_IN_DEGREE], "inverse");
}

int ogpu_graph_from_arrays(uint32_t N, uint32_t E, const uint32_t *const host_arrays[10],
                           ogpu_graph **out)
{
    *out = NULL;
    og_strided chk[10];
    for (int i = 0; i < 10; i++)
    {
        size_t n = (i == 0 || i == 1 || i == 5 || i == 6) ? E : N;
        if (n && !host_arrays[i])
        {
            og_set_error("graph array %d missing", i);
            return -1;
        }
        chk[i].a = host_arrays[i];
        chk[i].stride = 1;
    }
    if (og_csr_check(N, E, chk) || og_ctx_ensure()) return -1;
    og_graph *g = og_graph_alloc(N, E);
    if (!g) return -1;
    uint32_t *dst[10] = {g->f_src, g->f_dst, g->f_out_degree, g->f_in_degree, g->f_edges_idx,
                         g->i_src, g->i_dst, g->i_out_degree, g->i_in_degree, g->i_edges_idx};
    for (int i = 0; i < 10; i++)
    {
        size_t n = (i == 0 || i == 1 || i == 5 || i == 6) ? E : N;
        if (!n) continue;
        if (!host_arrays[i])
        {
            og_set_error("graph array %d missing", i);
            og_graph_destroy(g);
            return -1;
        }
        if (og_hip_check(hipMemcpyAsync(dst[i], host_arrays[i], n * 4, hipMemcpyHostToDevice, og_stream()),
                         "H2D graph array"))
        {
            og_graph_destroy(g);
            return -1;
        }
    }
    if (og_hip_check(hipStreamSynchronize(og_stream()), "sync upload"))
    {
        og_graph_destroy(g);
        return -1;
    }
    *out = (ogpu_graph *)g;
    return 0;
}

void ogpu_graph_free(ogpu_graph *g) { og_graph_destroy((og_graph *)g); }

/* ---------------------------- .csr files ---------------------------- */

static void og_graph_array_list(og_graph *g, uint32_t *a[10])
{
    a[0] = g->f_src; a[1] = g->f_dst; a[2] = g->f_out_degree; a[3] = g->f_in_degree; a[4] = g->f_edges_idx;
    a[5] = g->i_src; a[6] = g->i_dst; a[7] = g->i_out_degree; a[8] = g->i_in_degree; a[9] = g->i_edges_idx;
}

int ogpu_graph_write_csr(const ogpu_graph *gg, const char *fname)
{
    /* graphCSR.c:380-449: <fname>.csr = E, N, 6 words per vertex, 4 per
     * edge; interleaved on the GPU, written with one fwrite */
    og_graph *g = (og_graph *)gg;
    if (og_ctx_ensure()) return -1;
    const uint32_t N = g->N, E = g->E;
    const size_t words = 2 + 6 * (size_t)N + 4 * (size_t)E;
    uint32_t *d_body = og_dalloc((words - 2) * 4 + 4);
    uint32_t *h = (uint32_t *)malloc(words * 4);
    char *path = (char *)malloc(strlen(fname) + 5);
    int rc = -1;
    if (!d_body || !h || !path) { og_set_error("out of memory for the .csr image"); goto done; }
    uint32_t *arrays[10];
    og_graph_array_list(g, arrays);
    h[0] = E;
    h[1] = N;
    if (og_launch_csr_pack(arrays, N, E, d_body, og_stream()) ||
        og_hip_check(og_copy(h + 2, d_body, (words - 2) * 4, hipMemcpyDeviceToHost), "D2H .csr"))
        goto done;
    strcpy(path, fname);
    strcat(path, ".csr");
    FILE *f = fopen(path, "wb");
    if (!f) { og_set_error("open: %s", path); goto done; }
    size_t wrote = fwrite(h, 4, words, f);
    if (fclose(f) != 0 || wrote != words) { og_set_error("write: %s", path); goto done; }
    rc = 0;
done:
    og_dfree(d_body);
    free(h);
    free(path);
    return rc;
}

int ogpu_graph_read_csr(const char *path, ogpu_graph **out)
{
    /* graphCSR.c:452-558 into a device graph; the file is checked on the
     * host before the device is touched */
    *out = NULL;
    FILE *f = fopen(path, "rb");
    if (!f) { og_set_error("open: %s", path); return -1; }
    uint32_t hdr[2];
    if (fread(hdr, 4, 2, f) != 2) { fclose(f); og_set_error("%s: no .csr header", path); return -1; }
    const uint32_t E = hdr[0], N = hdr[1];
    const size_t words = 6 * (size_t)N + 4 * (size_t)E;
    uint32_t *h = (uint32_t *)malloc(words * 4 + 4);
    if (!h) { fclose(f); og_set_error("out of host memory"); return -1; }
    size_t got = fread(h, 4, words, f);
    const int trailing = fgetc(f) != EOF;
    fclose(f);
    if (got != words || trailing)
    {
        free(h);
        og_set_error("%s: %s: the file must hold exactly 2 + 6 N + 4 E = %zu words for E = %u, N = %u "
                     "(directed, unweighted)", path, trailing ? "bytes after the last edge" : "truncated",
                     words + 2, E, N);
        return -1;
    }
    /* file layout: 6 words per vertex, then 4 per edge */
    const uint32_t *ev = h + 6 * (size_t)N;
    const og_strided chk[10] = {{ev + 0, 4}, {ev + 1, 4}, {h + 0, 6}, {h + 1, 6}, {h + 2, 6},
                                {ev + 2, 4}, {ev + 3, 4}, {h + 3, 6}, {h + 4, 6}, {h + 5, 6}};
    if (og_csr_check(N, E, chk))
    {
        free(h);
        char msg[512];
        snprintf(msg, sizeof(msg), "%s: %s", path, ogpu_last_error());
        og_set_error("%s", msg);
        return -1;
    }
    if (og_ctx_ensure())
    {
        free(h);
        return -1;
    }
    og_graph *g = og_graph_alloc(N, E);
    uint32_t *d_body = og_dalloc(words * 4 + 4);
    int rc = -1;
    if (!g || !d_body) goto done;
    uint32_t *arrays[10];
    og_graph_array_list(g, arrays);
    if (og_hip_check(og_copy(d_body, h, words * 4, hipMemcpyHostToDevice), "H2D .csr") ||
        og_launch_csr_unpack(d_body, N, E, arrays, og_stream()) ||
        og_hip_check(hipStreamSynchronize(og_stream()), "sync .csr"))
        goto done;
    rc = 0;
done:
    free(h);
    og_dfree(d_body);
    if (rc) og_graph_destroy(g);
    else *out = (ogpu_graph *)g;
    return rc;
}

/* ------------------------- degree relabelling ------------------------- */

og_dbg_thresholds og_dbg_default_thresholds(uint32_t avg_degree)
{
    /* reorder.c:404-414: 11 buckets, the first at half the average degree
     * (1 when that is at most 1), doubling, the last unbounded */
    og_dbg_thresholds th;
    memset(&th, 0, sizeof(th));
    th.n = 11;
    th.t[0] = avg_degree <= 1 ? 1 : avg_degree / 2;
    for (uint32_t i = 1; i < th.n - 1; i++) th.t[i] = th.t[i - 1] * 2;
    th.t[th.n - 1] = 0xFFFFFFFFu;
    return th;
}

int og_reorder_labels_from_degrees(const uint32_t *d_deg, uint32_t N, uint32_t dbg, og_dbg_thresholds th,
                                   uint32_t *d_labels)
{
    hipStream_t s = og_stream();
    size_t nb = (size_t)N * 4;
    uint32_t *key = og_dalloc(nb), *ids = og_dalloc(nb);
    uint32_t *ka = og_dalloc(nb), *va = og_dalloc(nb), *kb = og_dalloc(nb), *vb = og_dalloc(nb);
    void *scratch = og_dalloc(og_radix_scratch_bytes(N));
    int rc = -1, in_b = 0;
    if (!key || !ids || !ka || !va || !kb || !vb || !scratch) goto done;
    if (dbg)
    {
        /* reorder.c:439-523: buckets from the last to the first, each in id
         * order: a stable sort of the ids by (n - 1 - bucket) */
        if (th.n < 1 || th.n > 16)
        {
            og_set_error("DBG needs 1..16 buckets (got %u)", th.n);
            goto done;
        }
        if (og_launch_reorder_dbg_keys(d_deg, N, th, key, ids, s) ||
            og_radix_sort_pairs(key, ids, ka, va, kb, vb, N, 8, scratch, s, &in_b) ||
            og_launch_reorder_labels(in_b ? vb : va, N, 0, d_labels, s))
            goto done;
    }
    else
    {
        /* reorder.c:223-271 and :337-347: stable ascending sort of the ids by
         * degree, then labels[sorted[i]] = N - 1 - i */
        if (og_hip_check(hipMemcpyAsync(key, d_deg, nb, hipMemcpyDeviceToDevice, s), "D2D") ||
            og_launch_iota(ids, N, s) ||
            og_radix_sort_pairs(key, ids, ka, va, kb, vb, N, 32, scratch, s, &in_b) ||
            og_launch_reorder_labels(in_b ? vb : va, N, 1, d_labels, s))
            goto done;
    }
    rc = og_hip_check(hipStreamSynchronize(s), "sync relabel");
done:
    og_dfree(key); og_dfree(ids); og_dfree(ka); og_dfree(va); og_dfree(kb); og_dfree(vb);
    og_dfree(scratch);
    return rc;
}

int og_reorder_labels_device(const uint32_t *d_src, const uint32_t *d_dst, uint32_t E, uint32_t N,
                             uint32_t avg_degree, uint32_t lmode, uint32_t *d_labels)
{
    if (lmode < 1 || lmode > 5)
    {
        og_set_error("-l %u: only the degree (1, 2, 3) and DBG (4, 5) relabellings are built", lmode);
        return -1;
    }
    if (og_ctx_ensure()) return -1;
    hipStream_t s = og_stream();
    uint32_t *deg = og_dalloc((size_t)N * 4);
    if (!deg) return -1;
    /* reorder.c:1492-1530: out-degree for 1 and 4, in-degree for 2 and 5, both for 3 */
    const uint32_t count_src = lmode != 2 && lmode != 5, count_dst = lmode == 2 || lmode == 3 || lmode == 5;
    int rc = og_launch_fill_u32(deg, 0, N, s) ||
             og_launch_reorder_degrees(d_src, d_dst, E, count_src, count_dst, deg, s) ||
             og_reorder_labels_from_degrees(deg, N, lmode >= 4, og_dbg_default_thresholds(avg_degree), d_labels);
    og_dfree(deg);
    return rc ? -1 : 0;
}

int og_graph_build_reordered_inplace(uint32_t *src, uint32_t *dst, uint32_t E, uint32_t N,
                                     const uint32_t *lmodes, uint32_t n_lmodes, uint32_t **labels_out)
{
    *labels_out = NULL;
    hipStream_t s = og_stream();
    /* edgeList.c:569: the integer average degree the DBG thresholds use */
    const uint32_t avg = E / N;
    size_t nb = (size_t)N * 4;
    uint32_t *lab = og_dalloc(nb), *acc = og_dalloc(nb);
    if (!lab || !acc || og_launch_iota(acc, N, s)) goto fail;
    /* graphCSR.c:184-203: each level relabels the (sorted) list and sorts it
     * again; the labels and the final stable sort do not depend on the
     * order the edges had, so the relabelled list in input order, sorted
     * once at the end, gives the same CSR. */
    for (uint32_t i = 0; i < n_lmodes; i++)
    {
        if (!lmodes[i]) continue;
        if (og_reorder_labels_device(src, dst, E, N, avg, lmodes[i], lab) ||
            og_launch_relabel(src, dst, E, lab, s) || og_launch_compose_labels(acc, NULL, N, lab, s))
            goto fail;
    }
    if (og_hip_check(hipStreamSynchronize(s), "sync relabel")) goto fail;
    og_dfree(lab);
    *labels_out = acc;
    return 0;
fail:
    og_dfree(lab);
    og_dfree(acc);
    return -1;
}

int ogpu_graph_build_reordered(const uint32_t *d_src, const uint32_t *d_dst, uint32_t num_edges,
                               uint32_t sort_mode, const uint32_t *lmodes, uint32_t n_lmodes,
                               uint32_t *d_labels_out, ogpu_graph **out)
{
    *out = NULL;
    if (og_ctx_ensure()) return -1;
    hipStream_t s = og_stream();
    const uint32_t E = num_edges;
    uint32_t maxid = 0;
    uint32_t *d_max = og_dalloc(4);
    if (!d_max) return -1;
    if (og_launch_max_id(d_src, d_dst, E, d_max, s) ||
        og_hip_check(hipMemcpyAsync(&maxid, d_max, 4, hipMemcpyDeviceToHost, s), "copy max id") ||
        og_hip_check(hipStreamSynchronize(s), "sync max id"))
    {
        og_dfree(d_max);
        return -1;
    }
    og_dfree(d_max);
    const uint32_t N = E ? maxid + 1 : 1;
    size_t eb = (size_t)E * 4;
    uint32_t *src = og_dalloc(eb ? eb : 4), *dst = og_dalloc(eb ? eb : 4), *acc = NULL;
    int rc = -1;
    if (!src || !dst) goto done;
    if (E && (og_hip_check(hipMemcpyAsync(src, d_src, eb, hipMemcpyDeviceToDevice, s), "D2D") ||
              og_hip_check(hipMemcpyAsync(dst, d_dst, eb, hipMemcpyDeviceToDevice, s), "D2D")))
        goto done;
    if (og_graph_build_reordered_inplace(src, dst, E, N, lmodes, n_lmodes, &acc)) goto done;
    if (d_labels_out &&
        og_hip_check(hipMemcpyAsync(d_labels_out, acc, (size_t)N * 4, hipMemcpyDeviceToDevice, s), "D2D"))
        goto done;
    rc = og_graph_build_device_n(src, dst, E, N, sort_mode, (og_graph **)out);
done:
    og_dfree(src); og_dfree(dst); og_dfree(acc);
    return rc;
}
