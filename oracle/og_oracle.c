/*
 * og_oracle.c -- CPU restatement of the reference CSR build and PageRank
 * variants.  TEST INFRASTRUCTURE ONLY (see og_oracle.h): never linked into
 * the product library.
 */
#include "og_oracle.h"

#include <math.h>
#include <omp.h>
#include <stdlib.h>
#include <string.h>

#define ORC_ID_MASK 0x3FFFFFFFu           /* reorder.h:19 VERTEX_CACHE_MASK_U32 */
#define ORC_DAMP 0.85f                    /* pageRank.h:13 */

/* ---------------------------------------------------------------------
 * x86-64 results of the float->integer conversions gcc emits for the
 * reference's casts (cvttss2si / cvttsd2si; out-of-range gives the
 * "integer indefinite" value INT_MIN of the destination width).
 * ------------------------------------------------------------------- */
static inline int32_t og_x86_f2i32(float x)
{
    if (!(x > -2147483904.0f && x < 2147483648.0f)) return INT32_MIN;
    return (int32_t)x;
}
static inline int64_t og_x86_f2i64(float x)
{
    if (!(x > -9223373136366403584.0f && x < 9223372036854775808.0f)) return INT64_MIN;
    return (int64_t)x;
}
static inline int64_t og_x86_d2i64(double x)
{
    if (!(x > -9223372036854777856.0 && x < 9223372036854775808.0)) return INT64_MIN;
    return (int64_t)x;
}
/* (uint64_t)double as gcc emits it: split at 2^63. */
static inline uint64_t og_x86_d2u64(double x)
{
    if (x >= 9223372036854775808.0)
        return (uint64_t)og_x86_d2i64(x - 9223372036854775808.0) ^ 0x8000000000000000ull;
    return (uint64_t)og_x86_d2i64(x);
}

static void *orc_alloc(size_t bytes)
{
    void *p = NULL;
    if (bytes == 0) bytes = 1;
    if (posix_memalign(&p, 128, bytes) != 0) return NULL;
    return p;
}

/* --------------------------------------------------------------------- */
/* CSR build                                                              */
/* --------------------------------------------------------------------- */

uint32_t orc_num_vertices(const uint32_t *src, const uint32_t *dst, uint32_t num_edges)
{
    /* edgeList.c:559-567: N = max(src, dst) + 1 over every edge. */
    uint32_t m = 0;
    for (uint32_t i = 0; i < num_edges; i++)
    {
        if (src[i] > m) m = src[i];
        if (dst[i] > m) m = dst[i];
    }
    return num_edges ? m + 1 : 1;
}

/* One LSD pass over 8 key bits: per-thread histogram of contiguous chunks,
 * exclusive scan in (digit, thread) order, stable scatter
 * (radixsort.c:46-135). */
static void orc_radix_pass(const uint32_t *k_in, const uint32_t *v_in, uint32_t *k_out,
                           uint32_t *v_out, uint32_t n, uint32_t shift, int P, uint32_t *cnt)
{
    #pragma omp parallel num_threads(P)
    {
        int t = omp_get_thread_num();
        int T = omp_get_num_threads();
        uint32_t chunk = n / (uint32_t)T;
        uint32_t b = (uint32_t)t * chunk;
        uint32_t e = (t == T - 1) ? n : b + chunk;
        uint32_t *h = cnt + (size_t)t * 256;
        memset(h, 0, 256 * sizeof(uint32_t));
        for (uint32_t i = b; i < e; i++) h[(k_in[i] >> shift) & 0xFF]++;
        #pragma omp barrier
        #pragma omp single
        {
            uint32_t base = 0;
            for (uint32_t d = 0; d < 256; d++)
                for (int j = 0; j < T; j++)
                {
                    uint32_t c = cnt[(size_t)j * 256 + d];
                    cnt[(size_t)j * 256 + d] = base;
                    base += c;
                }
        }
        for (uint32_t i = b; i < e; i++)
        {
            uint32_t o = h[(k_in[i] >> shift) & 0xFF]++;
            k_out[o] = k_in[i];
            v_out[o] = v_in[i];
        }
    }
}

int orc_stable_sort_pairs(uint32_t *key, uint32_t *val, uint32_t n, int nthreads)
{
    int P = nthreads > 0 ? nthreads : omp_get_max_threads();
    uint32_t *kt = orc_alloc((size_t)n * 4), *vt = orc_alloc((size_t)n * 4);
    uint32_t *cnt = orc_alloc((size_t)P * 256 * 4);
    if (!kt || !vt || !cnt) { free(kt); free(vt); free(cnt); return -1; }
    /* Four 8-bit digits (radixsort.c:246-270): an even number of passes, so
     * the result ends in the caller's arrays. */
    for (uint32_t pass = 0; pass < 4; pass += 2)
    {
        orc_radix_pass(key, val, kt, vt, n, pass * 8, P, cnt);
        orc_radix_pass(kt, vt, key, val, n, (pass + 1) * 8, P, cnt);
    }
    free(kt); free(vt); free(cnt);
    return 0;
}

static int orc_sort_list(uint32_t *src, uint32_t *dst, uint32_t n, uint32_t mode, int P)
{
    /* sortRun.c:26-69: modes 1 and 3 sort by dst first, then by src. */
    if (mode == 1 || mode == 3)
        if (orc_stable_sort_pairs(dst, src, n, P)) return -1;
    return orc_stable_sort_pairs(src, dst, n, P);
}

static void orc_map(const uint32_t *sorted_src, uint32_t E, uint32_t N,
                    uint32_t *out_degree, uint32_t *edges_idx)
{
    /* vertex.c:262-289: count edges per source and record the first index
     * of each source's run; vertices without edges keep 0 (vertex.c:26-49). */
    memset(out_degree, 0, (size_t)N * 4);
    memset(edges_idx, 0, (size_t)N * 4);
    for (uint32_t i = 0; i < E; i++)
    {
        uint32_t v = sorted_src[i] & ORC_ID_MASK;
        out_degree[v]++;
        if (i == 0 || sorted_src[i] != sorted_src[i - 1]) edges_idx[v] = i;
    }
}

int orc_build_csr(const uint32_t *src, const uint32_t *dst, uint32_t E,
                  uint32_t sort_mode, int nthreads, orc_csr *g)
{
    return orc_build_csr_n(src, dst, E, 0, sort_mode, nthreads, g);
}

int orc_build_csr_n(const uint32_t *src, const uint32_t *dst, uint32_t E, uint32_t min_vertices,
                    uint32_t sort_mode, int nthreads, orc_csr *g)
{
    memset(g, 0, sizeof(*g));
    uint32_t N = orc_num_vertices(src, dst, E);
    if (N < min_vertices) N = min_vertices;
    g->num_vertices = N;
    g->num_edges = E;
    size_t eb = (size_t)E * 4, nb = (size_t)N * 4;
    g->f_src = orc_alloc(eb); g->f_dst = orc_alloc(eb);
    g->i_src = orc_alloc(eb); g->i_dst = orc_alloc(eb);
    g->f_out_degree = orc_alloc(nb); g->f_in_degree = orc_alloc(nb); g->f_edges_idx = orc_alloc(nb);
    g->i_out_degree = orc_alloc(nb); g->i_in_degree = orc_alloc(nb); g->i_edges_idx = orc_alloc(nb);
    if (!g->f_src || !g->f_dst || !g->i_src || !g->i_dst || !g->f_out_degree || !g->f_in_degree ||
        !g->f_edges_idx || !g->i_out_degree || !g->i_in_degree || !g->i_edges_idx)
    {
        orc_free_csr(g);
        return -1;
    }
    memcpy(g->f_src, src, eb);
    memcpy(g->f_dst, dst, eb);
    /* graphCSR.c:178: sort the forward list. */
    if (orc_sort_list(g->f_src, g->f_dst, E, sort_mode, nthreads)) { orc_free_csr(g); return -1; }
    orc_map(g->f_src, E, N, g->f_out_degree, g->f_edges_idx);
    /* graphCSR.c:225: inverse list = the sorted forward list with src and
     * dst swapped (edgeList.c:705-711), then the same sort. */
    memcpy(g->i_src, g->f_dst, eb);
    memcpy(g->i_dst, g->f_src, eb);
    if (orc_sort_list(g->i_src, g->i_dst, E, sort_mode, nthreads)) { orc_free_csr(g); return -1; }
    orc_map(g->i_src, E, N, g->i_out_degree, g->i_edges_idx);
    /* vertex.c:291-310: each direction's in-degree is the other's out-degree. */
    memcpy(g->i_in_degree, g->f_out_degree, nb);
    memcpy(g->f_in_degree, g->i_out_degree, nb);
    return 0;
}

void orc_free_csr(orc_csr *g)
{
    free(g->f_src); free(g->f_dst); free(g->i_src); free(g->i_dst);
    free(g->f_out_degree); free(g->f_in_degree); free(g->f_edges_idx);
    free(g->i_out_degree); free(g->i_in_degree); free(g->i_edges_idx);
    memset(g, 0, sizeof(*g));
}

/* --------------------------------------------------------------------- */
/* PageRank                                                               */
/* --------------------------------------------------------------------- */

/* pageRank.c:989-996 (and the same loop in every variant): the fp32
 * contribution r/outdeg, 0 for a vertex without out-edges. */
static inline float orc_contrib(float r, uint32_t outdeg)
{
    return outdeg ? r / (float)outdeg : 0.0f;
}

/* pageRank.c:1014-1028: fp32 update, fp64 error accounting. */
#define ORC_EPILOGUE(NEXT_EXPR)                                                     \
    do {                                                                            \
        float prev_ = r[v];                                                         \
        float next_ = (NEXT_EXPR);                                                  \
        r[v] = next_;                                                               \
        double err_ = fabs(next_ - prev_);                                          \
        error_total += err_ / N;                                                    \
        if (err_ >= epsilon) active++;                                              \
    } while (0)

/* pageRank.c:1038-1047: normalize by N, PR sum in double. */
static double orc_normalize(float *r, uint32_t N)
{
    double sum = 0.0;
    for (uint32_t v = 0; v < N; v++)
    {
        r[v] = r[v] / (float)N;
        sum += r[v];
    }
    return sum;
}

static inline void orc_record(uint32_t it, uint32_t active, double err, uint32_t *act_out,
                              double *err_out, orc_pr_result *res)
{
    if (act_out) act_out[it] = active;
    if (err_out) err_out[it] = err;
    res->rows = it + 1;
}

int orc_pagerank(const orc_csr *g, uint32_t code, double epsilon, uint32_t max_iter,
                 int nthreads, const float *init, int stop_l1,
                 float *ranks_out, uint32_t *act_out, double *err_out, orc_pr_result *res)
{
    const uint32_t N = g->num_vertices;
    const uint32_t *f_deg = g->f_out_degree, *f_idx = g->f_edges_idx, *f_dst = g->f_dst;
    const uint32_t *i_deg = g->i_out_degree, *i_idx = g->i_edges_idx, *i_dst = g->i_dst;
    const float base_pr = 1.0f - ORC_DAMP, damp = ORC_DAMP;
    int P = nthreads > 0 ? nthreads : omp_get_max_threads();
    float *r = ranks_out;
    memset(res, 0, sizeof(*res));
    if (code > 13) return -2;

    for (uint32_t v = 0; v < N; v++) r[v] = init ? init[v] : base_pr;   /* pageRank.c:65-70 */

    float *cf = orc_alloc((size_t)N * 4);          /* riDividedOnDiClause (float) */
    float *nf = orc_alloc((size_t)N * 4);          /* pageRanksNext (float) */
    uint64_t *c64 = orc_alloc((size_t)N * 8);      /* fixed-point contributions */
    uint64_t *n64 = orc_alloc((size_t)N * 8);      /* fixed-point sums */
    uint8_t *wl_cur = orc_alloc(N), *wl_next = orc_alloc(N);
    if (!cf || !nf || !c64 || !n64 || !wl_cur || !wl_next)
    {
        free(cf); free(nf); free(c64); free(n64); free(wl_cur); free(wl_next);
        return -1;
    }
    memset(nf, 0, (size_t)N * 4);
    memset(n64, 0, (size_t)N * 8);

    double t0 = omp_get_wtime();
    double error_total = 0.0;
    uint32_t it = 0;

    if (code == 6 || code == 7 || code == 8)
    {
        /* Data-driven variants: pageRank.c:2405-2896. */
        float *resid = cf;
        for (uint32_t v = 0; v < N; v++) { wl_cur[v] = 1; wl_next[v] = 0; }
        if (code != 6)
        {
            /* pageRank.c:2616-2633: initial residual
             * 0.15 * 0.85 * sum over in-neighbours of 1/outdeg. */
            #pragma omp parallel for schedule(dynamic, 1024) num_threads(P)
            for (uint32_t v = 0; v < N; v++)
            {
                float a = 0.0f;
                for (uint32_t j = i_idx[v]; j < i_idx[v] + i_deg[v]; j++)
                {
                    uint32_t u = i_dst[j] & ORC_ID_MASK;
                    if (f_deg[u]) a += 1.0f / (float)f_deg[u];
                }
                resid[v] = (1.0f - damp) * damp * a;
            }
        }
        for (it = 0; it < max_iter; it++)
        {
            error_total = 0.0;
            uint32_t active = 0;
            if (code == 6)
            {
                for (uint32_t v = 0; v < N; v++) cf[v] = orc_contrib(r[v], f_deg[v]);
                /* pageRank.c:2474-2519: recompute active vertices, activate
                 * out-neighbours of the ones that moved by >= epsilon. */
                #pragma omp parallel for schedule(dynamic, 1024) num_threads(P) reduction(+ : error_total, active)
                for (uint32_t v = 0; v < N; v++)
                {
                    if (!wl_cur[v]) continue;
                    float s = 0.0f;
                    for (uint32_t j = i_idx[v]; j < i_idx[v] + i_deg[v]; j++)
                        s += cf[i_dst[j] & ORC_ID_MASK];
                    float old = r[v];
                    float nw = base_pr + (damp * s);
                    double err = fabs(nw - old);
                    error_total += err / N;
                    if (err >= epsilon)
                    {
                        r[v] = nw;
                        for (uint32_t j = f_idx[v]; j < f_idx[v] + f_deg[v]; j++)
                        {
                            #pragma omp atomic write
                            wl_next[f_dst[j] & ORC_ID_MASK] = 1;
                        }
                        active++;
                    }
                }
            }
            else
            {
                /* pageRank.c:2640-2685 (push) and :2796-2852 (pull-push):
                 * residual propagation with racy reads, as in the reference. */
                #pragma omp parallel for schedule(dynamic, 1024) num_threads(P) reduction(+ : error_total, active)
                for (uint32_t v = 0; v < N; v++)
                {
                    if (!wl_cur[v]) continue;
                    float old = r[v];
                    float nw;
                    if (code == 7)
                        nw = resid[v] + r[v];
                    else
                    {
                        float s = 0.0f;
                        for (uint32_t j = i_idx[v]; j < i_idx[v] + i_deg[v]; j++)
                        {
                            uint32_t u = i_dst[j] & ORC_ID_MASK;
                            s += r[u] / (float)f_deg[u];
                        }
                        nw = base_pr + (damp * s);
                    }
                    error_total += fabs(nw / N - old / N);
                    #pragma omp atomic write
                    r[v] = nw;
                    uint32_t deg = f_deg[v];
                    float delta = damp * (resid[v] / (float)deg);
                    for (uint32_t j = f_idx[v]; j < f_idx[v] + deg; j++)
                    {
                        uint32_t u = f_dst[j] & ORC_ID_MASK;
                        float prev;
                        #pragma omp atomic read
                        prev = resid[u];
                        #pragma omp atomic update
                        resid[u] += delta;
                        if ((fabs(prev + delta) >= epsilon) && (prev <= epsilon))
                        {
                            active++;
                            if (code == 8)
                            {
                                #pragma omp atomic update
                                resid[u] += delta;
                            }
                            wl_next[u] = 1;
                        }
                    }
                    resid[v] = 0.0f;
                }
            }
            uint8_t *tmp = wl_cur; wl_cur = wl_next; wl_next = tmp;
            memset(wl_next, 0, N);
            orc_record(it, active, error_total, act_out, err_out, res);
            if (active == 0) break;
        }
    }
    else
    {
        for (it = 0; it < max_iter; it++)
        {
            error_total = 0.0;
            uint32_t active = 0;
            double qscale_d = 1.0;
            float qscale_f = 1.0f;
            uint32_t qzero = 0;

            /* Contribution step. */
            switch (code)
            {
            case 0: case 1: case 4: case 5: case 12: case 13:
                for (uint32_t v = 0; v < N; v++) cf[v] = orc_contrib(r[v], f_deg[v]);
                break;
            case 2: case 3:
                /* pageRank.c:1251-1256: DoubleToFixed64 of the fp32 quotient
                 * (fixedPoint.h:36). */
                for (uint32_t v = 0; v < N; v++)
                {
                    if (!f_deg[v]) { c64[v] = 0; continue; }
                    float q = r[v] / (float)f_deg[v];
                    c64[v] = og_x86_d2u64((double)q * 4294967296.0 + (q >= 0 ? 0.5 : -0.5));
                }
                break;
            case 9: case 10: case 11:
            {
                /* FloatToFixed32/16/8 (fixedPoint.h:29-31): fp32 multiply by
                 * 2^16 / 2^14 / 2^6, fp32 add of +-0.5, x86 truncation. */
                float sc = code == 9 ? 65536.0f : (code == 10 ? 16384.0f : 64.0f);
                for (uint32_t v = 0; v < N; v++)
                {
                    if (!f_deg[v]) { c64[v] = 0; continue; }
                    float q = r[v] / (float)f_deg[v];
                    float x = q * sc + (q >= 0 ? 0.5f : -0.5f);
                    if (code == 9) c64[v] = (uint32_t)og_x86_f2i64(x);
                    else if (code == 10) c64[v] = (uint16_t)og_x86_f2i32(x);
                    else c64[v] = (uint8_t)og_x86_f2i32(x);
                }
                break;
            }
            default:
                break;
            }

            if (code == 4 || code == 5 || code == 12 || code == 13)
            {
                /* quantization.c:24-157: min and max of the contributions
                 * (the if / else-if scan yields the plain min and max). */
                float mn = cf[0], mx = cf[0];
                for (uint32_t v = 1; v < N; v++)
                {
                    if (cf[v] > mx) mx = cf[v];
                    if (cf[v] < mn) mn = cf[v];
                }
                double range = fabs((double)mx - (double)mn);
                if (code == 4)
                {
                    /* quantization.h:133-137 (GetScale_32 / quantize_32, zero 0). */
                    qscale_d = (mn == mx) ? 1.0 : range / 4294967295.0;
                    for (uint32_t v = 0; v < N; v++)
                        c64[v] = (uint32_t)og_x86_d2i64((double)cf[v] / qscale_d + 0.5);
                }
                else if (code == 5)
                {
                    /* quantization.h:79-87 (GetScale / GetZeroPoint / quantize). */
                    qscale_d = (mn == mx) ? 1.0 : range / 4294967295.0;
                    uint32_t rmax = (uint32_t)og_x86_d2i64((double)mx / qscale_d + 0.5);
                    qzero = 4294967295u - rmax;
                    for (uint32_t v = 0; v < N; v++)
                        c64[v] = (uint32_t)((uint32_t)og_x86_d2i64((double)cf[v] / qscale_d + 0.5) + qzero);
                }
                else
                {
                    /* quantization.h:93-125: float scale, clamp to the range. */
                    double rng = code == 12 ? 65535.0 : 255.0;
                    qscale_f = (float)((mn == mx) ? 1.0 : range / rng);
                    uint32_t qmax = code == 12 ? 65535u : 255u;
                    for (uint32_t v = 0; v < N; v++)
                    {
                        uint32_t q = (uint32_t)og_x86_d2i64((double)cf[v] / (double)qscale_f + 0.5);
                        c64[v] = q > qmax ? qmax : q;
                    }
                }
            }

            /* Gather / scatter step. */
            if (code == 1 || code == 3 || code == 5)
            {
                /* pageRank.c:1117-1134, :1785-1799, :2338-2352: push over the
                 * forward CSR with atomic adds. */
                #pragma omp parallel for schedule(dynamic, 1024) num_threads(P)
                for (uint32_t v = 0; v < N; v++)
                {
                    for (uint32_t j = f_idx[v]; j < f_idx[v] + f_deg[v]; j++)
                    {
                        uint32_t u = f_dst[j] & ORC_ID_MASK;
                        if (code == 1)
                        {
                            #pragma omp atomic update
                            nf[u] += cf[v];
                        }
                        else if (code == 3)
                        {
                            #pragma omp atomic update
                            n64[u] += c64[v];
                        }
                        else
                        {
                            double add = qscale_d * (double)(uint32_t)((uint32_t)c64[v] - qzero);
                            #pragma omp atomic update
                            nf[u] += add;
                        }
                    }
                }
            }
            else
            {
                /* pageRank.c:998-1011 and the same gather in every pull
                 * variant: in-neighbours of v in inverse-CSR order. */
                #pragma omp parallel for schedule(dynamic, 1024) num_threads(P)
                for (uint32_t v = 0; v < N; v++)
                {
                    uint32_t b = i_idx[v], e = i_idx[v] + i_deg[v];
                    if (code == 0)
                    {
                        float s = 0.0f;
                        for (uint32_t j = b; j < e; j++) s += cf[i_dst[j] & ORC_ID_MASK];
                        nf[v] = s;
                    }
                    else
                    {
                        uint64_t s = 0;
                        for (uint32_t j = b; j < e; j++) s += c64[i_dst[j] & ORC_ID_MASK];
                        if (code == 4) nf[v] = (float)(qscale_d * (double)s);
                        else if (code == 12 || code == 13) nf[v] = qscale_f * (float)s;
                        else n64[v] += s;
                    }
                }
            }

            /* Update step (pageRank.c:1014-1028 and its twins). */
            {
                #pragma omp parallel for num_threads(P) reduction(+ : error_total, active)
                for (uint32_t v = 0; v < N; v++)
                {
                    switch (code)
                    {
                    case 2: case 3:
                        /* Fixed64ToDouble, double arithmetic, float store. */
                        ORC_EPILOGUE((float)((double)base_pr + ((double)damp * ((double)n64[v] / 4294967296.0))));
                        n64[v] = 0;
                        break;
                    case 9:
                        ORC_EPILOGUE(base_pr + (damp * ((float)n64[v] / 65536.0f)));
                        n64[v] = 0;
                        break;
                    case 10:
                        ORC_EPILOGUE(base_pr + (damp * ((float)n64[v] / 16384.0f)));
                        n64[v] = 0;
                        break;
                    case 11:
                        ORC_EPILOGUE(base_pr + (damp * ((float)n64[v] / 64.0f)));
                        n64[v] = 0;
                        break;
                    default:
                        ORC_EPILOGUE(base_pr + (damp * nf[v]));
                        nf[v] = 0.0f;
                        break;
                    }
                }
            }
            orc_record(it, active, error_total, act_out, err_out, res);
            if (stop_l1 ? (error_total < epsilon) : (active == 0)) break;
        }
    }
    res->iterations = it;
    res->error_total = error_total;
    res->pr_sum = orc_normalize(r, N);
    res->seconds_total = omp_get_wtime() - t0;
    free(cf); free(nf); free(c64); free(n64); free(wl_cur); free(wl_next);
    return 0;
}

/* --------------------------------------------------------------------- */
/* FP64 oracles                                                           */
/* --------------------------------------------------------------------- */

uint32_t orc_pagerank_fp64(const orc_csr *g, double init, double tol, uint32_t max_iter,
                           int nthreads, double *out, double *last_l1)
{
    const uint32_t N = g->num_vertices;
    int P = nthreads > 0 ? nthreads : omp_get_max_threads();
    double *r = orc_alloc((size_t)N * 8), *c = orc_alloc((size_t)N * 8);
    if (!r || !c) { free(r); free(c); return 0; }
    for (uint32_t v = 0; v < N; v++) r[v] = init;
    uint32_t it;
    double l1 = 0.0;
    for (it = 0; it < max_iter; it++)
    {
        for (uint32_t v = 0; v < N; v++)
            c[v] = g->f_out_degree[v] ? r[v] / (double)g->f_out_degree[v] : 0.0;
        l1 = 0.0;
        #pragma omp parallel for schedule(dynamic, 1024) num_threads(P) reduction(+ : l1)
        for (uint32_t v = 0; v < N; v++)
        {
            double s = 0.0;
            for (uint32_t j = g->i_edges_idx[v]; j < g->i_edges_idx[v] + g->i_out_degree[v]; j++)
                s += c[g->i_dst[j] & ORC_ID_MASK];
            double nr = 0.15 + 0.85 * s;
            l1 += fabs(nr - r[v]);
            r[v] = nr;
        }
        l1 /= N;
        if (l1 < tol) { it++; break; }
    }
    for (uint32_t v = 0; v < N; v++) out[v] = r[v] / N;
    if (last_l1) *last_l1 = l1;
    free(r); free(c);
    return it;
}

uint32_t orc_ppr_fp64(const orc_csr *g, uint32_t s, double tol, uint32_t max_iter,
                      double *x, double *last_l1)
{
    const uint32_t N = g->num_vertices;
    double *c = orc_alloc((size_t)N * 8), *nx = orc_alloc((size_t)N * 8);
    if (!c || !nx) { free(c); free(nx); return 0; }
    for (uint32_t v = 0; v < N; v++) x[v] = 0.0;
    x[s] = 1.0;
    uint32_t it;
    double l1 = 0.0;
    for (it = 0; it < max_iter; it++)
    {
#pragma omp parallel for schedule(static)
        for (uint32_t v = 0; v < N; v++)
            c[v] = g->f_out_degree[v] ? x[v] / (double)g->f_out_degree[v] : 0.0;
        l1 = 0.0;
        /* rows in parallel (each row's sum in CSR order), the L1 change
         * reduced in any order: an fp64 checker, not a bit-exact one */
#pragma omp parallel for schedule(dynamic, 1024) reduction(+ : l1)
        for (uint32_t v = 0; v < N; v++)
        {
            double acc = 0.0;
            for (uint32_t j = g->i_edges_idx[v]; j < g->i_edges_idx[v] + g->i_out_degree[v]; j++)
                acc += c[g->i_dst[j] & ORC_ID_MASK];
            nx[v] = 0.85 * acc + (v == s ? 0.15 : 0.0);
            l1 += fabs(nx[v] - x[v]);
        }
        memcpy(x, nx, (size_t)N * 8);
        if (l1 < tol) { it++; break; }
    }
    if (last_l1) *last_l1 = l1;
    free(c); free(nx);
    return it;
}

uint32_t orc_ppr_f32(const orc_csr *g, uint32_t s, double tol, uint32_t max_iter,
                     float *x, double *l1_hist)
{
    const uint32_t N = g->num_vertices;
    const float base = 1.0f - 0.85f, damp = 0.85f;
    float *c = orc_alloc((size_t)N * 4), *nx = orc_alloc((size_t)N * 4);
    if (!c || !nx) { free(c); free(nx); return 0; }
    for (uint32_t v = 0; v < N; v++) x[v] = v == s ? 1.0f : 0.0f;
    uint32_t it;
    for (it = 0; it < max_iter; it++)
    {
        for (uint32_t v = 0; v < N; v++)
            c[v] = g->f_out_degree[v] ? x[v] / (float)g->f_out_degree[v] : 0.0f;
        double l1 = 0.0;
        for (uint32_t v = 0; v < N; v++)
        {
            float acc = 0.0f;
            for (uint32_t j = g->i_edges_idx[v]; j < g->i_edges_idx[v] + g->i_out_degree[v]; j++)
                acc += c[g->i_dst[j] & ORC_ID_MASK];
            nx[v] = (v == s ? base : 0.0f) + damp * acc;
            l1 += fabs((double)(nx[v] - x[v]));
        }
        memcpy(x, nx, (size_t)N * 4);
        if (l1_hist) l1_hist[it] = l1;
        if (l1 < tol) { it++; break; }
    }
    free(c); free(nx);
    return it;
}

void orc_sort_labels_by_rank(const float *ranks, uint32_t *labels, uint32_t n)
{
    /* reorder.c:165-221: key FloatToFixed32SORT(rank) = (uint32)(rank*2^20 +
     * 0.5f) (fixedPoint.h:33), stable ascending; labels start as given. */
    uint32_t *key = orc_alloc((size_t)n * 4);
    if (!key) return;
    for (uint32_t v = 0; v < n; v++)
    {
        float x = ranks[v] * 1048576.0f + (ranks[v] >= 0 ? 0.5f : -0.5f);
        key[v] = (uint32_t)og_x86_f2i64(x);
    }
    orc_stable_sort_pairs(key, labels, n, 1);
    free(key);
}

int orc_reorder_labels(const uint32_t *src, const uint32_t *dst, uint32_t num_edges, uint32_t N,
                       uint32_t lmode, uint32_t *labels)
{
    if (lmode < 1 || lmode > 5) return -1;
    uint32_t *deg = orc_alloc((size_t)N * 4), *key = orc_alloc((size_t)N * 4);
    if (!deg || !key) { free(deg); free(key); return -1; }
    /* reorder.c:1483-1530: out-degree (1, 4), in-degree (2, 5), both (3) */
    memset(deg, 0, (size_t)N * 4);
    for (uint32_t i = 0; i < num_edges; i++)
    {
        if (lmode != 2 && lmode != 5) deg[src[i]]++;
        if (lmode == 2 || lmode == 3 || lmode == 5) deg[dst[i]]++;
    }
    if (lmode <= 3)
    {
        /* reorder.c:296-358: stable ascending sort of the vertex ids by
         * degree (radixSortEdgesByDegree), then labels[sorted[i]] = N-1-i:
         * the largest degree gets label 0, equal degrees in decreasing id
         * order. */
        for (uint32_t v = 0; v < N; v++) { key[v] = deg[v]; labels[v] = v; }
        orc_stable_sort_pairs(key, labels, N, 1);
        memcpy(key, labels, (size_t)N * 4);
        for (uint32_t i = 0; i < N; i++) labels[key[i]] = N - 1 - i;
    }
    else
    {
        /* reorder.c:387-437 thresholds from the integer average degree,
         * :439-520: bucket = first threshold >= degree; labels run over the
         * buckets from the last to the first, within a bucket in id order. */
        const uint32_t nb = 11;
        uint32_t th[11];
        const uint32_t avg = num_edges / N;
        th[0] = avg <= 1 ? 1 : avg / 2;
        for (uint32_t i = 1; i < nb - 1; i++) th[i] = th[i - 1] * 2;
        th[nb - 1] = 0xFFFFFFFFu;
        uint32_t count[11] = {0}, start[11];
        for (uint32_t v = 0; v < N; v++)
        {
            uint32_t b = 0;
            while (deg[v] > th[b]) b++;
            key[v] = b;
            count[b]++;
        }
        uint32_t at = 0;
        for (int b = (int)nb - 1; b >= 0; b--) { start[b] = at; at += count[b]; }
        for (uint32_t v = 0; v < N; v++) labels[v] = start[key[v]]++;
    }
    free(deg);
    free(key);
    return 0;
}

int orc_spmv(const orc_csr *g, uint32_t code, uint32_t iterations, float *out)
{
    /* SPMV.c:616-1068, unweighted (every weight 0.0001f), run sequentially
     * (the push variants in source order) */
    const uint32_t N = g->num_vertices;
    float *x = orc_alloc((size_t)N * 4);
    float *y = orc_alloc((size_t)N * 4);
    uint64_t *acc = orc_alloc((size_t)N * 8), *xq = orc_alloc((size_t)N * 8);
    if (!x || !y || !acc || !xq) { free(x); free(y); free(acc); free(xq); return -1; }
    /* SPMV.c:663-669: x = 1/out-degree, 0.001 without out-edges */
    for (uint32_t v = 0; v < N; v++)
    {
        x[v] = g->f_out_degree[v] ? 1.0f / (float)g->f_out_degree[v] : 0.001f;
        y[v] = 0.0f;
        acc[v] = 0;
    }
    const float w = 0.0001f;
    if (code == 0)
    {
        /* SPMV.c:673-693: y[v] += w * x[src] over the inverse row, every iteration */
        for (uint32_t it = 0; it < iterations; it++)
            for (uint32_t v = 0; v < N; v++)
                for (uint32_t j = g->i_edges_idx[v]; j < g->i_edges_idx[v] + g->i_out_degree[v]; j++)
                {
                    const float t = w * x[g->i_dst[j] & ORC_ID_MASK];
                    y[v] = y[v] + t;
                }
    }
    else if (code == 1)
    {
        /* SPMV.c:763-783: y[dst] += w * x[v] over the forward row */
        for (uint32_t it = 0; it < iterations; it++)
            for (uint32_t v = 0; v < N; v++)
                for (uint32_t j = g->f_edges_idx[v]; j < g->f_edges_idx[v] + g->f_out_degree[v]; j++)
                {
                    const uint32_t d = g->f_dst[j] & ORC_ID_MASK;
                    const float t = w * x[v];
                    y[d] = y[d] + t;
                }
    }
    else if (code == 2)
    {
        /* SPMV.c:857-908: Q16.16 inputs, the weight FloatToFixed32(0.0001f) = 7
         * held in a float, u64 sums of (7 * x) >> 16, Fixed32ToFloat */
        const float wq = (float)(uint32_t)og_x86_f2i64(w * 65536.0f + 0.5f);
        for (uint32_t v = 0; v < N; v++) xq[v] = (uint32_t)og_x86_f2i64(x[v] * 65536.0f + 0.5f);
        const uint64_t wu = (uint64_t)og_x86_f2i64(wq);
        for (uint32_t it = 0; it < iterations; it++)
            for (uint32_t v = 0; v < N; v++)
                for (uint32_t j = g->i_edges_idx[v]; j < g->i_edges_idx[v] + g->i_out_degree[v]; j++)
                    acc[v] += (wu * xq[g->i_dst[j] & ORC_ID_MASK]) >> 16;
        for (uint32_t v = 0; v < N; v++) y[v] = (float)acc[v] / 65536.0f;
    }
    else if (code == 3)
    {
        /* SPMV.c:995-1034: Q32.32 inputs and weight DoubleToFixed64(0.0001f),
         * u64 sums of the 128-bit product >> 32, Fixed64ToDouble */
        const uint64_t wu = og_x86_d2u64((double)w * 4294967296.0 + 0.5);
        for (uint32_t v = 0; v < N; v++) xq[v] = og_x86_d2u64((double)x[v] * 4294967296.0 + 0.5);
        for (uint32_t it = 0; it < iterations; it++)
            for (uint32_t v = 0; v < N; v++)
                for (uint32_t j = g->f_edges_idx[v]; j < g->f_edges_idx[v] + g->f_out_degree[v]; j++)
                    acc[g->f_dst[j] & ORC_ID_MASK] += (uint64_t)(((unsigned __int128)wu * xq[v]) >> 32);
        for (uint32_t v = 0; v < N; v++) y[v] = (float)((double)acc[v] / 4294967296.0);
    }
    else
    {
        free(x); free(y); free(acc); free(xq);
        return -1;
    }
    memcpy(out, y, (size_t)N * 4);
    free(x); free(y); free(acc); free(xq);
    return 0;
}

/* graphStats.c:37-50: the smallest of three, except that a tie between the
 * first two below the third yields the third (as the reference's min3). */
static int orc_min3_ref(int a, int b, int c)
{
    if (a < b && a < c) return a;
    if (b < a && b < c) return b;
    return c;
}

/* graphStats.c:52-186: the score of the edit-distance matrix of a1 against
 * a2 (deletion, insertion, substitution cost 1), two rows at a time. */
static uint32_t orc_levenshtein(const uint32_t *a1, uint32_t n1, const uint32_t *a2, uint32_t n2)
{
    if (n1 == 0) return n2;
    if (n2 == 0) return n1;
    /* column j of the reference's mat[i][j] kept for all i: prev = j - 1 */
    uint32_t *prev = orc_alloc((size_t)(n1 + 1) * 4), *cur = orc_alloc((size_t)(n1 + 1) * 4);
    if (!prev || !cur) { free(prev); free(cur); return 0; }
    for (uint32_t i = 0; i <= n1; i++) prev[i] = i;
    for (uint32_t j = 1; j <= n2; j++)
    {
        cur[0] = j;
        for (uint32_t i = 1; i <= n1; i++)
        {
            const uint32_t cost = a1[i - 1] == a2[j - 1] ? 0u : 1u;
            const uint32_t del = prev[i] + 1, ins = cur[i - 1] + 1, sub = prev[i - 1] + cost;
            cur[i] = (uint32_t)orc_min3_ref((int)del, (int)ins, (int)sub);
        }
        uint32_t *t = prev; prev = cur; cur = t;
    }
    const uint32_t d = prev[n1];
    free(prev);
    free(cur);
    return d;
}

/* graphStats.c:427-489: Kendall tau-b by pair counting. */
static float orc_kendall_small(const float *arr1, const float *arr2, int len)
{
    int m1 = 0, m2 = 0, s = 0;
    for (int i = 0; i < len; i++)
        for (int j = i + 1; j < len; j++)
        {
            if (arr2[i] > arr2[j])
            {
                if (arr1[i] > arr1[j]) s++;
                else if (arr1[i] < arr1[j]) s--;
                else m1++;
            }
            else if (arr2[i] < arr2[j])
            {
                if (arr1[i] > arr1[j]) s--;
                else if (arr1[i] < arr1[j]) s++;
                else m1++;
            }
            else
            {
                m2++;
                if (arr1[i] == arr1[j]) m1++;
            }
        }
    const int nPair = len * (len - 1) / 2;
    if (m1 < nPair && m2 < nPair) return s / (sqrtf((float)(nPair - m1)) * sqrtf((float)(nPair - m2)));
    return 0.0f;
}

int orc_rank_stats_topk(const float *ref_pr, const uint32_t *ref_order, const float *pr, const uint32_t *order,
                        uint32_t N, uint32_t topk, uint32_t num_vertices, orc_corr_stats *out)
{
    /* graphStats.c:746-751 (inverse rankings) and :580-688 */
    if (num_vertices > N) return -1;
    if (topk > num_vertices) topk = num_vertices;
    const uint32_t K = topk, M = num_vertices;
    uint32_t *inv = orc_alloc((size_t)N * 4), *ref_inv = orc_alloc((size_t)N * 4);
    uint32_t *ids = orc_alloc((size_t)K * 4 + 4), *ref_ids = orc_alloc((size_t)K * 4 + 4);
    float *f = orc_alloc((size_t)K * 4 + 4), *ref_f = orc_alloc((size_t)K * 4 + 4);
    float *re = orc_alloc((size_t)K * 4 + 4), *ref_re = orc_alloc((size_t)K * 4 + 4);
    int rc = -1;
    if (!inv || !ref_inv || !ids || !ref_ids || !f || !ref_f || !re || !ref_re) goto done;
    for (uint32_t v = 0; v < N; v++)
    {
        inv[order[v]] = v;
        ref_inv[ref_order[v]] = v;
    }
    for (uint32_t u = 0, v = M - K; v < M; v++, u++)
    {
        ids[u] = order[v];
        f[u] = pr[order[v]];
        re[u] = (float)order[v] / (float)1.0;
        ref_ids[u] = ref_order[v];
        ref_f[u] = ref_pr[order[v]];
        ref_re[u] = (float)ref_order[v] / (float)1.0;
    }
    out->levenshtein_distance = orc_levenshtein(ref_ids, K, ids, K);
    out->float_Kendall = orc_kendall_small(ref_f, f, (int)K);
    out->real_Kendall = orc_kendall_small(ref_re, re, (int)K);
    /* graphStats.c:560-578: intersection_topK(ref order, inverse, M, K) */
    uint32_t inter = 0, mism = 0;
    double ef = 0.0, er = 0.0;
    for (uint32_t v = M - K; v < M; v++)
        if (inv[ref_order[v]] >= M - K) inter++;
    for (uint32_t v = M - K; v < M; v++)
    {
        const uint32_t w = ref_order[v];
        /* graphStats.c:539-557 */
        if ((double)ref_pr[w] > 0.0f) ef += fabs((double)pr[w] - (double)ref_pr[w]) / (double)ref_pr[w];
        /* graphStats.c:518-536: abs() of the unsigned difference taken as int */
        er += abs((int)(inv[w] - ref_inv[w])) / ((double)ref_inv[w] + 1);
        /* graphStats.c:500-516 */
        if (inv[w] != ref_inv[w]) mism++;
    }
    out->intersection = inter;
    out->mismatch = mism;
    out->avg_error_float = ef / K;
    out->avg_error_relative = er / K;
    rc = 0;
done:
    free(inv); free(ref_inv); free(ids); free(ref_ids); free(f); free(ref_f); free(re); free(ref_re);
    return rc;
}
