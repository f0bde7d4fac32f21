/*
 * og_pr.hip -- gfx950 kernels of the PageRank iteration (pageRank.c:939-2401)
 * and of the per-graph plan they run on.
 *
 * Data layout in HBM (all u32 indices, new vertex ids):
 *   row_ptr[N+1], cols[E]  inverse CSR with rows relabelled so that rows with
 *                          more than hub_t in-edges ("hubs") come first and
 *                          the rest follow by forward out-degree, descending
 *                          (so the contributions most often gathered share
 *                          cache lines).  Each row keeps the reference's
 *                          in-neighbour order, so every sum runs in the
 *                          reference's order.
 *   outdeg[N]              forward out-degree in new order.
 *   r[N]                   fp32 ranks (unnormalized, as the reference).
 *   c[2][N]                contributions r/outdeg in the variant's type,
 *                          ping-ponged between iterations; computed in the
 *                          epilogue of the iteration that produced r.
 *
 * One pull iteration is two kernels:
 *   og_pull_kernel<MODE>     blocks [0, n_chunks) sum one TE-edge chunk of a
 *                            hub row each; the remaining blocks each take
 *                            WIN_PER_BLOCK windows of the non-hub rows: the
 *                            window's edges are gathered (coalesced index
 *                            stream, parallel gathers) into LDS, then each
 *                            thread sums whole rows from LDS sequentially in
 *                            CSR order (bit-exact with pageRank.c:998-1011)
 *                            and runs the fused epilogue
 *                            (pageRank.c:1014-1028 + the next iteration's
 *                            contribution, pageRank.c:989-996).
 *   og_pull_finalize<MODE>   sums each hub's chunk partials in chunk order,
 *                            runs its epilogue, and reduces the per-block
 *                            error / active partials in a fixed tree (the
 *                            last block, by ticket, writes the iteration's
 *                            stats and the stop flag).
 * The fp64 error sum is therefore deterministic; the active count is exact.
 */
#include <hip/hip_runtime.h>

#include "og_internal.h"
#include "og_pr.h"

#define OG_BLOCK 256
#define OG_WAVES (OG_BLOCK / 64)
#define OG_LDS_EDGES (2 * OG_TE)

/* ------------------------------------------------------------------ */
/* per-mode arithmetic                                                 */
/* ------------------------------------------------------------------ */

template <int MODE> struct og_mode;
template <> struct og_mode<OG_MODE_F32>  { typedef float    elem; typedef float    acc; };
template <> struct og_mode<OG_MODE_FX64> { typedef uint64_t elem; typedef uint64_t acc; };
template <> struct og_mode<OG_MODE_FX32> { typedef uint32_t elem; typedef uint64_t acc; };
template <> struct og_mode<OG_MODE_FX16> { typedef uint16_t elem; typedef uint64_t acc; };
template <> struct og_mode<OG_MODE_FX8>  { typedef uint8_t  elem; typedef uint64_t acc; };
template <> struct og_mode<OG_MODE_Q32>  { typedef uint32_t elem; typedef uint64_t acc; };
template <> struct og_mode<OG_MODE_Q16>  { typedef uint16_t elem; typedef uint64_t acc; };
template <> struct og_mode<OG_MODE_Q8>   { typedef uint8_t  elem; typedef uint64_t acc; };

/* The reference's next-rank expression for each variant, from the gathered
 * sum (pageRank.c:1016, :1278, :1412, :1544, :1676, :1944, :2083, :2222). */
template <int MODE>
__device__ __forceinline__ float og_next_rank(typename og_mode<MODE>::acc s, const og_qparams *qp)
{
    const float base = OG_BASE_PR, damp = OG_DAMP;
    if constexpr (MODE == OG_MODE_F32) return base + (damp * s);
    else if constexpr (MODE == OG_MODE_FX64)
        return (float)((double)base + ((double)damp * ((double)s / 4294967296.0)));
    else if constexpr (MODE == OG_MODE_FX32) return base + (damp * ((float)s / 65536.0f));
    else if constexpr (MODE == OG_MODE_FX16) return base + (damp * ((float)s / 16384.0f));
    else if constexpr (MODE == OG_MODE_FX8) return base + (damp * ((float)s / 64.0f));
    else if constexpr (MODE == OG_MODE_Q32) return base + damp * (float)(qp->scale_d * (double)s);
    else return base + damp * (qp->scale_f * (float)s);
}

/* The contribution the next iteration gathers (pageRank.c:989-996 and the
 * fixed-point conversions of fixedPoint.h:29-36).  Quantized modes store
 * the fp32 quotient; og_quantize_kernel converts it once the scale is known. */
template <int MODE>
__device__ __forceinline__ void og_store_contrib(void *c_out, uint32_t v, float r, uint32_t deg)
{
    /* Stores are non-temporal: the next iteration reads them back through
     * L2, this one must not let them evict the lines being gathered. */
    if constexpr (MODE == OG_MODE_F32 || MODE >= OG_MODE_Q32)
        __builtin_nontemporal_store(deg ? r / (float)deg : 0.0f, &((float *)c_out)[v]);
    else
    {
        typedef typename og_mode<MODE>::elem E;
        E x = 0;
        if (deg)
        {
            float q = r / (float)deg;
            if constexpr (MODE == OG_MODE_FX64) x = og_fixed64_of(q);
            else if constexpr (MODE == OG_MODE_FX32)
                x = (uint32_t)og_x86_f2i64(q * 65536.0f + (q >= 0 ? 0.5f : -0.5f));
            else if constexpr (MODE == OG_MODE_FX16)
                x = (uint16_t)og_x86_f2i32(q * 16384.0f + (q >= 0 ? 0.5f : -0.5f));
            else
                x = (uint8_t)og_x86_f2i32(q * 64.0f + (q >= 0 ? 0.5f : -0.5f));
        }
        __builtin_nontemporal_store(x, &((E *)c_out)[v]);
    }
}

/* Running per-thread statistics of the epilogue. */
struct og_stat
{
    double err;
    uint32_t active;
    float cmin, cmax;
};

__device__ __forceinline__ void og_stat_init(og_stat &s)
{
    s.err = 0.0;
    s.active = 0;
    s.cmin = 3.402823466e38f;
    s.cmax = -3.402823466e38f;
}

template <int MODE>
__device__ __forceinline__ void og_epilogue_pf(const og_pr_dev &a, uint32_t v,
                                               typename og_mode<MODE>::acc sum, float prev,
                                               uint32_t deg, og_stat &st);

/* pageRank.c:1016-1027: update, |delta| in double, error_total += err/N,
 * activeVertices += err >= epsilon; then the next contribution. */
template <int MODE>
__device__ __forceinline__ void og_epilogue(const og_pr_dev &a, uint32_t v,
                                            typename og_mode<MODE>::acc sum, og_stat &st)
{
    if constexpr (MODE == OG_MODE_F32)
    {
        if (a.dd)
        {
            /* data-driven pull, pageRank.c:2483-2516: only worklist vertices;
             * the rank moves only when it changed by at least epsilon, and
             * then its out-neighbours are activated (og_dd_activate_kernel). */
            if (!a.wl_cur[v]) return;
            float old = a.r[v];
            float nr = og_next_rank<MODE>(sum, a.qp);
            double err = fabs((double)(nr - old));
            st.err += err / (double)a.N;
            if (err >= a.epsilon)
            {
                a.r[v] = nr;
                a.changed[v] = 1;
                st.active++;
            }
            return;
        }
    }
    /* per-vertex arrays stream through once: non-temporal accesses */
    float prev = __builtin_nontemporal_load(&a.r[v]);
    uint32_t deg = __builtin_nontemporal_load(&a.outdeg[v]);
    og_epilogue_pf<MODE>(a, v, sum, prev, deg, st);
}

/* The same with the old rank and out-degree already loaded (the windowed
 * pull prefetches them with the window's gathers). */
template <int MODE>
__device__ __forceinline__ void og_epilogue_pf(const og_pr_dev &a, uint32_t v,
                                               typename og_mode<MODE>::acc sum, float prev,
                                               uint32_t deg, og_stat &st)
{
    float nr = og_next_rank<MODE>(sum, a.qp);
    __builtin_nontemporal_store(nr, &a.r[v]);
    double err = fabs((double)(nr - prev));
    st.err += err / (double)a.N;
    st.active += err >= a.epsilon;
    og_store_contrib<MODE>(a.c_out, v, nr, deg);
    if constexpr (MODE >= OG_MODE_Q32)
    {
        float c = deg ? nr / (float)deg : 0.0f;
        st.cmin = fminf(st.cmin, c);
        st.cmax = fmaxf(st.cmax, c);
    }
}

/* ------------------------------------------------------------------ */
/* block reductions in a fixed order                                   */
/* ------------------------------------------------------------------ */

template <typename T> __device__ __forceinline__ T og_wave_sum(T v)
{
    for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o, 64);
    return v;
}

/* Reduce og_stat over the block; result valid in thread 0. */
__device__ og_stat og_block_stat(og_stat s)
{
    __shared__ double sh_err[OG_WAVES];
    __shared__ uint32_t sh_act[OG_WAVES];
    __shared__ float sh_min[OG_WAVES], sh_max[OG_WAVES];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    s.err = og_wave_sum(s.err);
    s.active = og_wave_sum(s.active);
    for (int o = 32; o > 0; o >>= 1)
    {
        s.cmin = fminf(s.cmin, __shfl_down(s.cmin, o, 64));
        s.cmax = fmaxf(s.cmax, __shfl_down(s.cmax, o, 64));
    }
    if (lane == 0)
    {
        sh_err[w] = s.err;
        sh_act[w] = s.active;
        sh_min[w] = s.cmin;
        sh_max[w] = s.cmax;
    }
    __syncthreads();
    og_stat r;
    og_stat_init(r);
    if (threadIdx.x == 0)
        for (int i = 0; i < OG_WAVES; i++)
        {
            r.err += sh_err[i];
            r.active += sh_act[i];
            r.cmin = fminf(r.cmin, sh_min[i]);
            r.cmax = fmaxf(r.cmax, sh_max[i]);
        }
    __syncthreads();
    return r;
}

template <typename T> __device__ T og_block_sum(T v)
{
    __shared__ T sh[OG_WAVES];
    v = og_wave_sum(v);
    if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = v;
    __syncthreads();
    T r = 0;
    if (threadIdx.x == 0)
        for (int i = 0; i < OG_WAVES; i++) r += sh[i];
    __syncthreads();
    return r;
}

/* ------------------------------------------------------------------ */
/* pull iteration                                                      */
/* ------------------------------------------------------------------ */

/* Issue the index loads, then the gathers, of a window of ne <= LDS_EDGES
 * edges starting at eb: slot u of thread t holds edge u*BLOCK + t.  The
 * slot guards are block-uniform, lanes past the end re-read the last edge,
 * so all loads of a thread are in flight together. */
template <typename E>
__device__ __forceinline__ void og_window_issue(const uint32_t *__restrict__ cols,
                                                const E *__restrict__ g, uint32_t eb, uint32_t ne,
                                                uint32_t t, E (&val)[OG_LDS_EDGES / OG_BLOCK])
{
    constexpr uint32_t SLOTS = OG_LDS_EDGES / OG_BLOCK;
    uint32_t idx[SLOTS];
#pragma unroll
    for (uint32_t u = 0; u < SLOTS; u++)
        if (u * OG_BLOCK < ne) idx[u] = __builtin_nontemporal_load(&cols[eb + min(u * OG_BLOCK + t, ne - 1)]);
#pragma unroll
    for (uint32_t u = 0; u < SLOTS; u++)
        if (u * OG_BLOCK < ne) val[u] = g[idx[u]];
}

template <int MODE>
__global__ __launch_bounds__(OG_BLOCK) void og_pull_kernel(og_pr_dev a, uint32_t it)
{
    typedef typename og_mode<MODE>::elem E;
    typedef typename og_mode<MODE>::acc A;
    if (*a.done) return;
    const E *__restrict__ g = (const E *)a.g_in;
    const uint32_t t = threadIdx.x;

    if (blockIdx.x < a.n_chunks)
    {
        /* One TE-edge chunk of a hub row: each thread sums a fixed strided
         * subset, then a fixed tree over the block. */
        uint32_t c = blockIdx.x;
        uint32_t lo = 0, hi = a.H;             /* chunk_base[h] <= c < chunk_base[h+1] */
        while (hi - lo > 1)
        {
            uint32_t mid = (lo + hi) >> 1;
            if (a.chunk_base[mid] <= c) lo = mid; else hi = mid;
        }
        uint32_t h = lo;
        if (a.dd && !a.wl_cur[h])
        {
            if (t == 0) ((A *)a.chunk_sum)[c] = 0;
            return;
        }
        uint32_t eb = a.row_ptr[h] + (c - a.chunk_base[h]) * OG_TE;
        uint32_t ee = min(eb + OG_TE, a.row_ptr[h + 1]);
        /* Branch-free: out-of-range lanes re-read the chunk's last edge and
         * drop it, so all index loads and then all gathers are in flight
         * together. */
        A s = 0;
        uint32_t idx[OG_TE / OG_BLOCK];
#pragma unroll
        for (uint32_t u = 0; u < OG_TE / OG_BLOCK; u++)
            idx[u] = __builtin_nontemporal_load(&a.cols[min(eb + u * OG_BLOCK + t, ee - 1)]);
        E val[OG_TE / OG_BLOCK];
#pragma unroll
        for (uint32_t u = 0; u < OG_TE / OG_BLOCK; u++) val[u] = g[idx[u]];
#pragma unroll
        for (uint32_t u = 0; u < OG_TE / OG_BLOCK; u++)
            if (eb + u * OG_BLOCK + t < ee) s += (A)val[u];
        s = og_block_sum<A>(s);
        if (t == 0) ((A *)a.chunk_sum)[c] = s;
        return;
    }

    /* Windows are software-pipelined through two LDS buffers: the index
     * loads and gathers of window w+1 -- and the row offsets, old ranks and
     * out-degrees of its (at most 2 per thread) rows -- are issued before
     * window w is summed, so the sums and epilogues of w run on LDS and
     * registers while w+1's traffic is in flight. */
    constexpr uint32_t SLOTS = OG_LDS_EDGES / OG_BLOCK;
    constexpr uint32_t RSLOTS = OG_TE / OG_ROW_W / OG_BLOCK;
    __shared__ E lds[2][OG_LDS_EDGES];
    __shared__ uint32_t s_wrow[OG_WIN_PER_BLOCK + 1], s_wedge[OG_WIN_PER_BLOCK + 1];
    og_stat st;
    og_stat_init(st);
    const uint32_t tb = blockIdx.x - a.n_chunks;
    const uint32_t w0 = tb * OG_WIN_PER_BLOCK;
    const uint32_t nw = min(OG_WIN_PER_BLOCK, a.n_windows - w0);
    if (t <= nw)
    {
        s_wrow[t] = a.win_row[w0 + t];
        s_wedge[t] = a.win_edge[w0 + t];
    }
    __syncthreads();
    E val[SLOTS];
    uint32_t nj[RSLOTS], ne_[RSLOTS], nd[RSLOTS];
    float nr[RSLOTS];
    auto issue = [&](uint32_t k) {
        uint32_t rb = s_wrow[k], re = s_wrow[k + 1];
        uint32_t eb = s_wedge[k], ne = s_wedge[k + 1] - eb;
        if (ne <= OG_LDS_EDGES) og_window_issue<E>(a.cols, g, eb, ne, t, val);
#pragma unroll
        for (uint32_t q = 0; q < RSLOTS; q++)
        {
            uint32_t v = rb + q * OG_BLOCK + t;
            if (v < re)
            {
                nj[q] = __builtin_nontemporal_load(&a.row_ptr[v]);
                ne_[q] = __builtin_nontemporal_load(&a.row_ptr[v + 1]);
                nr[q] = __builtin_nontemporal_load(&a.r[v]);
                nd[q] = __builtin_nontemporal_load(&a.outdeg[v]);
            }
        }
    };
    issue(0);
    uint32_t buf = 0;
    for (uint32_t k = 0; k < nw; k++)
    {
        const uint32_t rb = s_wrow[k], re = s_wrow[k + 1];
        const uint32_t eb = s_wedge[k], ne = s_wedge[k + 1] - eb;
        const bool in_lds = ne <= OG_LDS_EDGES;
        if (in_lds)
        {
#pragma unroll
            for (uint32_t u = 0; u < SLOTS; u++)
                if (u * OG_BLOCK + t < ne) lds[buf][u * OG_BLOCK + t] = val[u];
        }
        uint32_t cj[RSLOTS], ce[RSLOTS], cd[RSLOTS];
        float cr[RSLOTS];
#pragma unroll
        for (uint32_t q = 0; q < RSLOTS; q++)
        {
            cj[q] = nj[q];
            ce[q] = ne_[q];
            cr[q] = nr[q];
            cd[q] = nd[q];
        }
        __syncthreads();
        if (k + 1 < nw) issue(k + 1);
        if (in_lds)
        {
            /* Each thread sums whole rows in CSR order; LDS reads are issued
             * eight at a time ahead of the (sequential) adds. */
            const E *L = lds[buf];
#pragma unroll
            for (uint32_t q = 0; q < RSLOTS; q++)
            {
                uint32_t v = rb + q * OG_BLOCK + t;
                if (v >= re) break;
                uint32_t j = cj[q] - eb, e = ce[q] - eb;
                A s = 0;
                for (; j + 8 <= e; j += 8)
                {
                    E x[8];
#pragma unroll
                    for (uint32_t u = 0; u < 8; u++) x[u] = L[j + u];
#pragma unroll
                    for (uint32_t u = 0; u < 8; u++) s += (A)x[u];
                }
                for (; j < e; j++) s += (A)L[j];
                if (a.dd) og_epilogue<MODE>(a, v, s, st);
                else og_epilogue_pf<MODE>(a, v, s, cr[q], cd[q], st);
            }
        }
        else
        {
            /* Window larger than LDS (only when hub_t > TE): direct sequential sums. */
            for (uint32_t v = rb + t; v < re; v += OG_BLOCK)
            {
                A s = 0;
                for (uint32_t j = a.row_ptr[v]; j < a.row_ptr[v + 1]; j++) s += (A)g[a.cols[j]];
                og_epilogue<MODE>(a, v, s, st);
            }
        }
        buf ^= 1;
    }
    og_stat r = og_block_stat(st);
    if (t == 0)
    {
        a.part_err[tb] = r.err;
        a.part_act[tb] = r.active;
        a.part_min[tb] = r.cmin;
        a.part_max[tb] = r.cmax;
    }
}

/* Final reduction shared by pull and push: block b folds a fixed slice of
 * the first-level partials (plus, for pull, the hubs it finalizes); the
 * last block to arrive folds the block results in index order and writes
 * the iteration's statistics, the quantization parameters of the next
 * iteration and the stop flag. */
__device__ void og_reduce_tail(const og_pr_dev &a, uint32_t it, og_stat mine)
{
    const uint32_t t = threadIdx.x;
    const uint32_t per = OG_PARTS_PER_FINAL_BLOCK;
    uint32_t pb = blockIdx.x * per, pe = min(pb + per, a.n_parts);
    for (uint32_t i = pb + t; i < pe; i += OG_BLOCK)
    {
        mine.err += a.part_err[i];
        mine.active += a.part_act[i];
        mine.cmin = fminf(mine.cmin, a.part_min[i]);
        mine.cmax = fmaxf(mine.cmax, a.part_max[i]);
    }
    og_stat r = og_block_stat(mine);
    __shared__ uint32_t s_last;
    if (t == 0)
    {
        a.fin_err[blockIdx.x] = r.err;
        a.fin_act[blockIdx.x] = r.active;
        a.fin_min[blockIdx.x] = r.cmin;
        a.fin_max[blockIdx.x] = r.cmax;
        __threadfence();
        uint32_t ticket = atomicAdd(a.ticket, 1u);
        s_last = ticket == gridDim.x - 1;
    }
    __syncthreads();
    if (!s_last) return;
    __threadfence();
    og_stat acc;
    og_stat_init(acc);
    for (uint32_t i = t; i < gridDim.x; i += OG_BLOCK)
    {
        acc.err += __hip_atomic_load(&a.fin_err[i], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        acc.active += __hip_atomic_load(&a.fin_act[i], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        acc.cmin = fminf(acc.cmin, __hip_atomic_load(&a.fin_min[i], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT));
        acc.cmax = fmaxf(acc.cmax, __hip_atomic_load(&a.fin_max[i], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT));
    }
    og_stat tot = og_block_stat(acc);
    if (t == 0)
    {
        if (it != OG_NO_STATS)
        {
            a.stats[it].error_total = tot.err;
            a.stats[it].active = tot.active;
        }
        bool stop = a.stop_l1 ? (tot.err < a.epsilon) : (tot.active == 0);
        if (stop && a.allow_stop) *a.done = 1;
        if (a.qmode)
        {
            /* quantization.h:79-137 and quantization.c:24-157 (the
             * reference's if/else-if scan yields the plain min and max). */
            float mn = tot.cmin, mx = tot.cmax;
            double range = fabs((double)mx - (double)mn);
            og_qparams q;
            q.scale_d = (mn == mx) ? 1.0 : range / 4294967295.0;
            q.zero = 0;
            if (a.qmode == OG_MODE_Q16) q.scale_f = (float)((mn == mx) ? 1.0 : range / 65535.0);
            else if (a.qmode == OG_MODE_Q8) q.scale_f = (float)((mn == mx) ? 1.0 : range / 255.0);
            else q.scale_f = (float)q.scale_d;
            if (a.qzero)
                q.zero = 4294967295u - (uint32_t)og_x86_d2i64((double)mx / q.scale_d + 0.5);
            *a.qp_next = q;
        }
        *a.ticket = 0;
    }
}

template <int MODE>
__global__ __launch_bounds__(OG_BLOCK) void og_pull_finalize_kernel(og_pr_dev a, uint32_t it)
{
    typedef typename og_mode<MODE>::acc A;
    if (*a.done) return;
    og_stat st;
    og_stat_init(st);
    uint32_t h = blockIdx.x * OG_BLOCK + threadIdx.x;
    if (h < a.H)
    {
        A s = 0;
        for (uint32_t c = a.chunk_base[h]; c < a.chunk_base[h + 1]; c++) s += ((const A *)a.chunk_sum)[c];
        og_epilogue<MODE>(a, h, s, st);
    }
    og_reduce_tail(a, it, st);
}

/* Quantize the fp32 contributions with this iteration's parameters
 * (quantization.h:85, :121, :105, :137). */
template <int MODE>
__global__ void og_quantize_kernel(og_pr_dev a)
{
    if (*a.done) return;
    const og_qparams q = *a.qp;
    const float *cf = (const float *)a.cq_src;
    for (uint32_t v = blockIdx.x * blockDim.x + threadIdx.x; v < a.N; v += gridDim.x * blockDim.x)
    {
        double x = (double)cf[v];
        if constexpr (MODE == OG_MODE_Q32)
            ((uint32_t *)a.cq_dst)[v] = (uint32_t)og_x86_d2i64(x / q.scale_d + 0.5);
        else if constexpr (MODE == OG_MODE_Q16)
        {
            uint32_t u = (uint32_t)og_x86_d2i64(x / (double)q.scale_f + 0.5);
            ((uint16_t *)a.cq_dst)[v] = (uint16_t)(u > 65535u ? 65535u : u);
        }
        else if constexpr (MODE == OG_MODE_Q8)
        {
            uint32_t u = (uint32_t)og_x86_d2i64(x / (double)q.scale_f + 0.5);
            ((uint8_t *)a.cq_dst)[v] = (uint8_t)(u > 255u ? 255u : u);
        }
        else
            ((uint32_t *)a.cq_dst)[v] = (uint32_t)og_x86_d2i64(x / q.scale_d + 0.5) + q.zero;
    }
}

/* ------------------------------------------------------------------ */
/* push iteration (pageRank.c:1117-1134, :1785-1799, :2338-2352)      */
/* ------------------------------------------------------------------ */

template <int PMODE>
__global__ __launch_bounds__(OG_BLOCK) void og_push_scatter_kernel(og_pr_dev a)
{
    if (*a.done) return;
    const uint32_t E = a.E;
    for (uint32_t e = blockIdx.x * OG_BLOCK + threadIdx.x; e < E; e += gridDim.x * OG_BLOCK)
    {
        uint32_t s = __builtin_nontemporal_load(&a.push_src[e]);
        uint32_t d = __builtin_nontemporal_load(&a.push_dst[e]);
        if constexpr (PMODE == OG_MODE_F32)
        {
            unsafeAtomicAdd(&((float *)a.next)[d], ((const float *)a.g_in)[s]);
        }
        else if constexpr (PMODE == OG_MODE_FX64)
        {
            atomicAdd((unsigned long long *)&((uint64_t *)a.next)[d],
                      (unsigned long long)((const uint64_t *)a.g_in)[s]);
        }
        else
        {
            uint32_t qv = ((const uint32_t *)a.g_in)[s];
            double add = a.qp->scale_d * (double)(uint32_t)(qv - a.qp->zero);
            unsafeAtomicAdd(&((float *)a.next)[d], (float)add);
        }
    }
}

template <int PMODE>
__global__ __launch_bounds__(OG_BLOCK) void og_push_epilogue_kernel(og_pr_dev a)
{
    if (*a.done) return;
    og_stat st;
    og_stat_init(st);
    const int EMODE = PMODE == OG_MODE_FX64 ? OG_MODE_FX64 : (PMODE == OG_MODE_F32 ? OG_MODE_F32 : OG_MODE_QPUSH_EPI);
    for (uint32_t v = blockIdx.x * OG_BLOCK + threadIdx.x; v < a.N; v += gridDim.x * OG_BLOCK)
    {
        if constexpr (EMODE == OG_MODE_FX64)
        {
            uint64_t s = ((uint64_t *)a.next)[v];
            ((uint64_t *)a.next)[v] = 0;
            og_epilogue<OG_MODE_FX64>(a, v, s, st);
        }
        else
        {
            float s = ((float *)a.next)[v];
            ((float *)a.next)[v] = 0.0f;
            if constexpr (EMODE == OG_MODE_F32) og_epilogue<OG_MODE_F32>(a, v, s, st);
            else
            {
                /* quantized push: fp32 update, fp32 contribution + min/max */
                float prev = a.r[v];
                float nr = OG_BASE_PR + OG_DAMP * s;
                a.r[v] = nr;
                double err = fabs((double)(nr - prev));
                st.err += err / (double)a.N;
                st.active += err >= a.epsilon;
                uint32_t deg = a.outdeg[v];
                float c = deg ? nr / (float)deg : 0.0f;
                ((float *)a.c_out)[v] = c;
                st.cmin = fminf(st.cmin, c);
                st.cmax = fmaxf(st.cmax, c);
            }
        }
    }
    og_stat r = og_block_stat(st);
    if (threadIdx.x == 0)
    {
        a.part_err[blockIdx.x] = r.err;
        a.part_act[blockIdx.x] = r.active;
        a.part_min[blockIdx.x] = r.cmin;
        a.part_max[blockIdx.x] = r.cmax;
    }
}

__global__ __launch_bounds__(OG_BLOCK) void og_final_kernel(og_pr_dev a, uint32_t it)
{
    if (*a.done) return;
    og_stat st;
    og_stat_init(st);
    og_reduce_tail(a, it, st);
}

/* ------------------------------------------------------------------ */
/* init / normalize                                                    */
/* ------------------------------------------------------------------ */

/* r = initial ranks (pageRank.c:65-70 or the caller's start vector, given
 * in old ids), c = contributions of r, min/max partials for quantization. */
template <int MODE>
__global__ __launch_bounds__(OG_BLOCK) void og_init_kernel(og_pr_dev a, const float *init_old,
                                                          float init_value)
{
    og_stat st;
    og_stat_init(st);
    for (uint32_t v = blockIdx.x * OG_BLOCK + threadIdx.x; v < a.N; v += gridDim.x * OG_BLOCK)
    {
        float r = init_old ? init_old[a.new2old[v]] : init_value;
        a.r[v] = r;
        uint32_t deg = a.outdeg[v];
        og_store_contrib<MODE>(a.c_out, v, r, deg);
        if constexpr (MODE >= OG_MODE_Q32)
        {
            float c = deg ? r / (float)deg : 0.0f;
            st.cmin = fminf(st.cmin, c);
            st.cmax = fmaxf(st.cmax, c);
        }
    }
    og_stat r = og_block_stat(st);
    if (threadIdx.x == 0)
    {
        a.part_err[blockIdx.x] = 0.0;
        a.part_act[blockIdx.x] = 0;
        a.part_min[blockIdx.x] = r.cmin;
        a.part_max[blockIdx.x] = r.cmax;
    }
}

/* pageRank.c:1038-1044: normalized ranks back in old ids, PR sum partials. */
__global__ __launch_bounds__(OG_BLOCK) void og_normalize_kernel(og_pr_dev a, float *out_old)
{
    double s = 0.0;
    const float fn = (float)a.N;
    for (uint32_t v = blockIdx.x * OG_BLOCK + threadIdx.x; v < a.N; v += gridDim.x * OG_BLOCK)
    {
        float x = a.r[a.old2new[v]] / fn;
        out_old[v] = x;
        s += x;
    }
    s = og_block_sum<double>(s);
    if (threadIdx.x == 0) a.part_err[blockIdx.x] = s;
}

__global__ void og_sum_parts_kernel(const double *p, uint32_t n, double *out)
{
    double s = 0.0;
    for (uint32_t i = threadIdx.x; i < n; i += OG_BLOCK) s += p[i];
    s = og_block_sum<double>(s);
    if (threadIdx.x == 0) *out = s;
}

/* ------------------------------------------------------------------ */
/* launchers                                                           */
/* ------------------------------------------------------------------ */

template <int MODE> static void og_launch_pull_t(const og_pr_dev *a, uint32_t it, hipStream_t s,
                                                 hipEvent_t ev0, hipEvent_t ev1)
{
    uint32_t grid = a->n_chunks + a->n_tile_blocks;
    if (ev0) hipEventRecord(ev0, s);
    if (grid) hipLaunchKernelGGL(og_pull_kernel<MODE>, dim3(grid), dim3(OG_BLOCK), 0, s, *a, it);
    if (ev1) hipEventRecord(ev1, s);
    hipLaunchKernelGGL(og_pull_finalize_kernel<MODE>, dim3(a->n_final_blocks), dim3(OG_BLOCK), 0, s,
                       *a, it);
}

extern "C" int og_launch_pull(const og_pr_dev *a, int mode, uint32_t it, hipStream_t s,
                              hipEvent_t ev0, hipEvent_t ev1)
{
    switch (mode)
    {
    case OG_MODE_F32: og_launch_pull_t<OG_MODE_F32>(a, it, s, ev0, ev1); break;
    case OG_MODE_FX64: og_launch_pull_t<OG_MODE_FX64>(a, it, s, ev0, ev1); break;
    case OG_MODE_FX32: og_launch_pull_t<OG_MODE_FX32>(a, it, s, ev0, ev1); break;
    case OG_MODE_FX16: og_launch_pull_t<OG_MODE_FX16>(a, it, s, ev0, ev1); break;
    case OG_MODE_FX8: og_launch_pull_t<OG_MODE_FX8>(a, it, s, ev0, ev1); break;
    case OG_MODE_Q32: og_launch_pull_t<OG_MODE_Q32>(a, it, s, ev0, ev1); break;
    case OG_MODE_Q16: og_launch_pull_t<OG_MODE_Q16>(a, it, s, ev0, ev1); break;
    case OG_MODE_Q8: og_launch_pull_t<OG_MODE_Q8>(a, it, s, ev0, ev1); break;
    default: og_set_error("og_launch_pull: bad mode %d", mode); return -1;
    }
    return og_hip_check(hipGetLastError(), "og_pull_kernel");
}

extern "C" int og_launch_quantize(const og_pr_dev *a, int mode, hipStream_t s)
{
    unsigned grid = (a->N + OG_BLOCK - 1) / OG_BLOCK;
    if (grid > 4096) grid = 4096;
    if (grid == 0) grid = 1;
    switch (mode)
    {
    case OG_MODE_Q32: hipLaunchKernelGGL(og_quantize_kernel<OG_MODE_Q32>, dim3(grid), dim3(OG_BLOCK), 0, s, *a); break;
    case OG_MODE_Q16: hipLaunchKernelGGL(og_quantize_kernel<OG_MODE_Q16>, dim3(grid), dim3(OG_BLOCK), 0, s, *a); break;
    case OG_MODE_Q8: hipLaunchKernelGGL(og_quantize_kernel<OG_MODE_Q8>, dim3(grid), dim3(OG_BLOCK), 0, s, *a); break;
    default: hipLaunchKernelGGL(og_quantize_kernel<OG_MODE_QPUSH_EPI>, dim3(grid), dim3(OG_BLOCK), 0, s, *a); break;
    }
    return og_hip_check(hipGetLastError(), "og_quantize_kernel");
}

extern "C" int og_launch_push(const og_pr_dev *a, int pmode, uint32_t it, hipStream_t s,
                              hipEvent_t ev0, hipEvent_t ev1)
{
    unsigned gs = (a->E + OG_BLOCK - 1) / OG_BLOCK;
    if (gs > 16384) gs = 16384;
    if (gs == 0) gs = 1;
    if (ev0) hipEventRecord(ev0, s);
    if (pmode == OG_MODE_F32)
    {
        hipLaunchKernelGGL(og_push_scatter_kernel<OG_MODE_F32>, dim3(gs), dim3(OG_BLOCK), 0, s, *a);
        if (ev1) hipEventRecord(ev1, s);
        hipLaunchKernelGGL(og_push_epilogue_kernel<OG_MODE_F32>, dim3(a->n_parts), dim3(OG_BLOCK), 0, s, *a);
    }
    else if (pmode == OG_MODE_FX64)
    {
        hipLaunchKernelGGL(og_push_scatter_kernel<OG_MODE_FX64>, dim3(gs), dim3(OG_BLOCK), 0, s, *a);
        if (ev1) hipEventRecord(ev1, s);
        hipLaunchKernelGGL(og_push_epilogue_kernel<OG_MODE_FX64>, dim3(a->n_parts), dim3(OG_BLOCK), 0, s, *a);
    }
    else
    {
        hipLaunchKernelGGL(og_push_scatter_kernel<OG_MODE_QPUSH_EPI>, dim3(gs), dim3(OG_BLOCK), 0, s, *a);
        if (ev1) hipEventRecord(ev1, s);
        hipLaunchKernelGGL(og_push_epilogue_kernel<OG_MODE_QPUSH_EPI>, dim3(a->n_parts), dim3(OG_BLOCK), 0, s, *a);
    }
    hipLaunchKernelGGL(og_final_kernel, dim3(a->n_final_blocks), dim3(OG_BLOCK), 0, s, *a, it);
    return og_hip_check(hipGetLastError(), "og_push");
}

extern "C" int og_launch_init(const og_pr_dev *a, int mode, const float *init_old, float init_value,
                              hipStream_t s)
{
    unsigned grid = a->n_parts;
    switch (mode)
    {
    case OG_MODE_F32: hipLaunchKernelGGL(og_init_kernel<OG_MODE_F32>, dim3(grid), dim3(OG_BLOCK), 0, s, *a, init_old, init_value); break;
    case OG_MODE_FX64: hipLaunchKernelGGL(og_init_kernel<OG_MODE_FX64>, dim3(grid), dim3(OG_BLOCK), 0, s, *a, init_old, init_value); break;
    case OG_MODE_FX32: hipLaunchKernelGGL(og_init_kernel<OG_MODE_FX32>, dim3(grid), dim3(OG_BLOCK), 0, s, *a, init_old, init_value); break;
    case OG_MODE_FX16: hipLaunchKernelGGL(og_init_kernel<OG_MODE_FX16>, dim3(grid), dim3(OG_BLOCK), 0, s, *a, init_old, init_value); break;
    case OG_MODE_FX8: hipLaunchKernelGGL(og_init_kernel<OG_MODE_FX8>, dim3(grid), dim3(OG_BLOCK), 0, s, *a, init_old, init_value); break;
    default: hipLaunchKernelGGL(og_init_kernel<OG_MODE_Q32>, dim3(grid), dim3(OG_BLOCK), 0, s, *a, init_old, init_value); break;
    }
    return og_hip_check(hipGetLastError(), "og_init_kernel");
}

extern "C" int og_launch_final(const og_pr_dev *a, uint32_t it, hipStream_t s)
{
    hipLaunchKernelGGL(og_final_kernel, dim3(a->n_final_blocks), dim3(OG_BLOCK), 0, s, *a, it);
    return og_hip_check(hipGetLastError(), "og_final_kernel");
}

extern "C" int og_launch_normalize(const og_pr_dev *a, float *d_out_old, double *d_sum,
                                   hipStream_t s)
{
    hipLaunchKernelGGL(og_normalize_kernel, dim3(a->n_parts), dim3(OG_BLOCK), 0, s, *a, d_out_old);
    hipLaunchKernelGGL(og_sum_parts_kernel, dim3(1), dim3(OG_BLOCK), 0, s, a->part_err, a->n_parts, d_sum);
    return og_hip_check(hipGetLastError(), "og_normalize_kernel");
}

/* ------------------------------------------------------------------ */
/* data-driven variants (pageRank.c:2405-2896)                         */
/* ------------------------------------------------------------------ */

/* pageRank.c:2469-2476: contributions of the current ranks; clears the
 * changed flags of the previous iteration. */
__global__ void og_dd_contrib_kernel(og_pr_dev a)
{
    if (*a.done) return;
    for (uint32_t v = blockIdx.x * blockDim.x + threadIdx.x; v < a.N; v += gridDim.x * blockDim.x)
    {
        uint32_t deg = a.outdeg[v];
        ((float *)a.c_out)[v] = deg ? a.r[v] / (float)deg : 0.0f;
        a.changed[v] = 0;
    }
}

/* pageRank.c:2504-2513: workListNext[u] = 1 for every out-neighbour u of a
 * vertex whose rank moved (edge-parallel over the forward edges). */
__global__ void og_dd_activate_kernel(og_pr_dev a)
{
    if (*a.done) return;
    for (uint32_t e = blockIdx.x * blockDim.x + threadIdx.x; e < a.E; e += gridDim.x * blockDim.x)
        if (a.changed[a.push_src[e]]) a.wl_next[a.push_dst[e]] = 1;
}

/* pageRank.c:2616-2633: residual = (1-d) * d * sum over in-neighbours of
 * 1/outdeg, in CSR order. */
__global__ void og_dd_resid_init_kernel(og_pr_dev a)
{
    for (uint32_t v = blockIdx.x * blockDim.x + threadIdx.x; v < a.N; v += gridDim.x * blockDim.x)
    {
        float acc = 0.0f;
        for (uint32_t j = a.row_ptr[v]; j < a.row_ptr[v + 1]; j++)
        {
            uint32_t d = a.outdeg[a.cols[j]];
            if (d) acc += 1.0f / (float)d;
        }
        a.resid[v] = (1.0f - OG_DAMP) * OG_DAMP * acc;
    }
}

/* pageRank.c:2645-2685 (push) and :2796-2852 (pull-push): one thread per
 * worklist vertex, residuals moved to out-neighbours with atomic adds.  The
 * reference's reads of neighbours' residuals and ranks race with their
 * updates; here they race the same way, so results are nondeterministic. */
__global__ __launch_bounds__(OG_BLOCK) void og_dd_push_kernel(og_pr_dev a, uint32_t code)
{
    if (*a.done) return;
    og_stat st;
    og_stat_init(st);
    const float fn = (float)a.N;
    for (uint32_t v = blockIdx.x * OG_BLOCK + threadIdx.x; v < a.N; v += gridDim.x * OG_BLOCK)
    {
        if (!a.wl_cur[v]) continue;
        float old = a.r[v];
        float nr;
        if (code == 7)
            nr = a.resid[v] + old;
        else
        {
            float s = 0.0f;
            for (uint32_t j = a.row_ptr[v]; j < a.row_ptr[v + 1]; j++)
            {
                uint32_t u = a.cols[j];
                s += __hip_atomic_load(&a.r[u], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) / (float)a.outdeg[u];
            }
            nr = OG_BASE_PR + (OG_DAMP * s);
        }
        st.err += fabs((double)(nr / fn - old / fn));
        __hip_atomic_store(&a.r[v], nr, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        uint32_t deg = a.outdeg[v];
        float delta = OG_DAMP * (a.resid[v] / (float)deg);
        for (uint32_t j = a.f_row_ptr[v]; j < a.f_row_ptr[v + 1]; j++)
        {
            uint32_t u = a.f_cols[j];
            float prev = atomicAdd(&a.resid[u], delta);
            if ((fabs((double)(prev + delta)) >= a.epsilon) && ((double)prev <= a.epsilon))
            {
                st.active++;
                if (code == 8) atomicAdd(&a.resid[u], delta);
                a.wl_next[u] = 1;
            }
        }
        a.resid[v] = 0.0f;
    }
    og_stat r = og_block_stat(st);
    if (threadIdx.x == 0)
    {
        a.part_err[blockIdx.x] = r.err;
        a.part_act[blockIdx.x] = r.active;
        a.part_min[blockIdx.x] = r.cmin;
        a.part_max[blockIdx.x] = r.cmax;
    }
}

static unsigned og_vgrid(uint64_t n)
{
    uint64_t g = (n + OG_BLOCK - 1) / OG_BLOCK;
    if (g > 4096) g = 4096;
    return g ? (unsigned)g : 1u;
}

extern "C" int og_launch_dd_contrib(const og_pr_dev *a, hipStream_t s)
{
    hipLaunchKernelGGL(og_dd_contrib_kernel, dim3(og_vgrid(a->N)), dim3(OG_BLOCK), 0, s, *a);
    return og_hip_check(hipGetLastError(), "og_dd_contrib_kernel");
}

extern "C" int og_launch_dd_activate(const og_pr_dev *a, hipStream_t s)
{
    if (!a->E) return 0;
    unsigned g = (a->E + OG_BLOCK - 1) / OG_BLOCK;
    if (g > 16384) g = 16384;
    hipLaunchKernelGGL(og_dd_activate_kernel, dim3(g), dim3(OG_BLOCK), 0, s, *a);
    return og_hip_check(hipGetLastError(), "og_dd_activate_kernel");
}

extern "C" int og_launch_dd_resid_init(const og_pr_dev *a, hipStream_t s)
{
    hipLaunchKernelGGL(og_dd_resid_init_kernel, dim3(og_vgrid(a->N)), dim3(OG_BLOCK), 0, s, *a);
    return og_hip_check(hipGetLastError(), "og_dd_resid_init_kernel");
}

extern "C" int og_launch_dd_push(const og_pr_dev *a, uint32_t code, uint32_t it, hipStream_t s)
{
    hipLaunchKernelGGL(og_dd_push_kernel, dim3(a->n_parts), dim3(OG_BLOCK), 0, s, *a, code);
    hipLaunchKernelGGL(og_final_kernel, dim3(a->n_final_blocks), dim3(OG_BLOCK), 0, s, *a, it);
    return og_hip_check(hipGetLastError(), "og_dd_push_kernel");
}
