/*
 * og_pr.hip -- gfx950 kernels of the PageRank iteration (pageRank.c:939-2401)
 * and of the per-graph plan they run on.
 *
 * Data layout in HBM (all u32 indices, new vertex ids):
 *   row_ptr[N+1], cols[E]  inverse CSR with rows relabelled by in-degree
 *                          class (several chunks, one chunk, short, none),
 *                          within a class by forward out-degree descending
 *                          (the contributions gathered most often share
 *                          cache lines) or, for the short rows, by
 *                          in-degree descending.  Each row keeps the
 *                          reference's in-neighbour order.
 *   ch_row[], ch_edge[]    chunks of OG_CH edges tiling the long rows.
 *   outdeg[N]              forward out-degree in new order.
 *   r[N]                   fp32 ranks (unnormalized, as the reference).
 *   c[2][N]                contributions r/outdeg in the variant's type,
 *                          ping-ponged between iterations; computed in the
 *                          epilogue of the iteration that produced r.
 *
 * One pull iteration (a degree-binned gather SpMV plus the update) is two
 * kernels:
 *   og_pull_kernel<MODE>     blocks [0, n_chain_blocks) (fp32 only, with
 *                            chains): one wavefront per chain row adds its
 *                            in-edges in CSR order, bit-exact with the
 *                            reference's running sum; the next
 *                            n_chunk_blocks: each wavefront sums
 *                            OG_CH_PER_WAVE chunks (64 lanes x 8 edges,
 *                            fixed tree); a row of one chunk gets its fused
 *                            epilogue there (pageRank.c:1014-1028 plus the
 *                            next contribution, pageRank.c:989-996), a
 *                            longer row stores the chunk sum.  The other
 *                            blocks take the short rows, one per thread,
 *                            summed in CSR order (bit-exact with
 *                            pageRank.c:998-1011), and their epilogues.
 *   og_pull_finalize<MODE>   one wavefront per long row adds its chunk sums
 *                            and runs its epilogue; then the per-block
 *                            error / active partials are reduced in a fixed
 *                            tree (the last block, by ticket, writes the
 *                            iteration's stats and the stop flag).
 * The fp64 error sum is therefore deterministic; the active count is exact.
 */
#include <hip/hip_runtime.h>

#include "og_internal.h"
#include "og_pr.h"

#define OG_BLOCK 256
#define OG_WAVES (OG_BLOCK / 64)
#define OG_CH_SLOTS (OG_CH / 64) /* edges per lane of a chunk */
#define OG_ROW_BATCH 8           /* edges of a short row loaded together */

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

/* SpMV (SPMV.c:673-693, :857-908): y[v] += the row's sum of the terms
 * w * x[src].  fp32: y lives in r; `chained` when the sum already started
 * from y and added the terms in CSR order, as the reference's running
 * `vector_output[dest] +=` does (bit-exact), else y + the row's sum.
 * Integer modes: u64 accumulators in next, exact in any order. */
template <int MODE>
__device__ __forceinline__ void og_spmv_store(const og_pr_dev &a, uint32_t v,
                                              typename og_mode<MODE>::acc s, float prev, bool chained)
{
    if constexpr (MODE == OG_MODE_F32) a.r[v] = chained ? s : prev + s;
    else ((uint64_t *)a.next)[v] += (uint64_t)s;
}

/* pageRank.c:1016-1027: update, |delta| in double, error_total += err/N,
 * activeVertices += err >= epsilon; then the next contribution. */
template <int MODE>
__device__ __forceinline__ void og_epilogue(const og_pr_dev &a, uint32_t v,
                                            typename og_mode<MODE>::acc sum, og_stat &st)
{
    if (a.spmv)
    {
        og_spmv_store<MODE>(a, v, sum, MODE == OG_MODE_F32 ? a.r[v] : 0.0f, false);
        return;
    }
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

/* The same with the old rank and out-degree already loaded (the pull loads
 * them together with the row's first edges). */
template <int MODE>
__device__ __forceinline__ void og_epilogue_pf(const og_pr_dev &a, uint32_t v,
                                               typename og_mode<MODE>::acc sum, float prev,
                                               uint32_t deg, og_stat &st)
{
    float nr;
    if constexpr (MODE == OG_MODE_F32)
    {
        if (a.raw)
        {
            a.r[v] = sum;            /* sum over in-neighbours of g_in (feature pass) */
            return;
        }
        /* personalized: the restart term (1 - d) only at the source */
        const float base = !a.ppr || v == a.ppr_src ? OG_BASE_PR : 0.0f;
        nr = base + OG_DAMP * sum;
    }
    else
        nr = og_next_rank<MODE>(sum, a.qp);
    __builtin_nontemporal_store(nr, &a.r[v]);
    double err = fabs((double)(nr - prev));
    st.err += a.ppr ? err : err / (double)a.N;
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

/* Reduce og_stat over a block of NW wavefronts; result valid in thread 0. */
template <int NW = OG_WAVES> __device__ og_stat og_block_stat(og_stat s)
{
    __shared__ double sh_err[NW];
    __shared__ uint32_t sh_act[NW];
    __shared__ float sh_min[NW], sh_max[NW];
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
        for (int i = 0; i < NW; i++)
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

/* Sum over the wavefront by an xor butterfly: every lane ends with the same
 * value (each tree node adds the same two operands, in either order). */
template <typename T> __device__ __forceinline__ T og_wave_allsum(T v)
{
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

/* Gather g[i].  With a hot cache (HOT) the first nh entries come from LDS
 * and those lanes' global load is pointed at g[0] -- one line, an L1 hit --
 * so only the other lanes send requests to L2, whose request rate bounds
 * the scattered gathers.  Branch-free: both loads of every slot stay in
 * flight together. */
template <typename E, bool HOT>
__device__ __forceinline__ E og_gather(const E *__restrict__ g, const E *hot, uint32_t nh, uint32_t i)
{
    if constexpr (!HOT) return g[i];
    else
    {
        const bool h = i < nh;
        const E xg = g[h ? 0u : i];
        const E xh = hot[h ? i : 0u];
        return h ? xh : xg;
    }
}

/* The loads, sums and outputs of a wavefront's K chunks with S slots per
 * lane (S * 64 >= the longest of them). */
template <int MODE, bool HOT, uint32_t S, uint32_t K>
__device__ __forceinline__ void og_chunk_body(const og_pr_dev &a, og_stat &st, uint32_t c0, uint32_t nk,
                                              const uint32_t (&eb)[K], const uint32_t (&ee)[K], uint32_t my_tag,
                                              const typename og_mode<MODE>::elem *hot, uint32_t nh)
{
    typedef typename og_mode<MODE>::elem E;
    typedef typename og_mode<MODE>::acc A;
    const E *__restrict__ g = (const E *)a.g_in;
    const uint32_t lane = threadIdx.x & 63;
    uint32_t idx[K][S];
#pragma unroll
    for (uint32_t k = 0; k < K; k++)
#pragma unroll
        for (uint32_t u = 0; u < S; u++)
            idx[k][u] = __builtin_nontemporal_load(&a.cols[min(eb[k] + u * 64 + lane, ee[k] - 1)]);
    /* lane k: old rank and out-degree of chunk k's row, for a one-chunk row */
    const uint32_t my_row = my_tag & ~OG_CH_MULTI;
    const float prev = a.r[my_row];
    const uint32_t deg = a.outdeg[my_row];
    E val[K][S];
#pragma unroll
    for (uint32_t k = 0; k < K; k++)
#pragma unroll
        for (uint32_t u = 0; u < S; u++) val[k][u] = og_gather<E, HOT>(g, hot, nh, idx[k][u]);
    A s[K];
#pragma unroll
    for (uint32_t k = 0; k < K; k++)
    {
        s[k] = 0;
#pragma unroll
        for (uint32_t u = 0; u < S; u++) s[k] += eb[k] + u * 64 + lane < ee[k] ? (A)val[k][u] : (A)0;
    }
#pragma unroll
    for (uint32_t k = 0; k < K; k++) s[k] = og_wave_allsum(s[k]);
    A mine = s[0];
#pragma unroll
    for (uint32_t k = 1; k < K; k++)
        if (lane == k) mine = s[k];
    if (lane < nk)
    {
        const uint32_t c = c0 + lane;
        if (a.dd && !a.wl_cur[my_row])
        {
            if (my_tag & OG_CH_MULTI) ((A *)a.chunk_sum)[c] = 0;
        }
        else if (my_tag & OG_CH_MULTI) ((A *)a.chunk_sum)[c] = mine;
        else if (a.spmv) og_spmv_store<MODE>(a, my_row, mine, prev, false);
        else if (a.dd) og_epilogue<MODE>(a, my_row, mine, st);
        else og_epilogue_pf<MODE>(a, my_row, mine, prev, deg, st);
    }
}

/* Chunks [c0, c0 + K) of one wavefront (K = OG_CH_PER_WAVE but in the
 * multi-graph kernels, whose tile shape is tuned; a block has NW
 * wavefronts), all in flight together: lane l loads edges l, l+64, ... of each chunk (the index loads
 * coalesce), gathers them and adds them in that order; the wavefront then
 * adds its 64 lane sums in a fixed tree.  Lanes past a chunk's end re-read
 * its last edge and add zero, so no branch separates the loads and the
 * gathers of a lane overlap; the number of slots per lane follows the
 * longest of the K chunks (the one-chunk rows are sorted by length, so
 * neighbouring chunks are about as long).  Lane k then stores chunk k's
 * sum (row with several chunks) or runs the row's epilogue. */
template <int MODE, bool HOT = false, uint32_t NW = OG_WAVES, uint32_t K = OG_CH_PER_WAVE>
__device__ __forceinline__ void og_pull_chunks(const og_pr_dev &a, og_stat &st, uint32_t cblock,
                                               const typename og_mode<MODE>::elem *hot = nullptr,
                                               uint32_t nh = 0)
{
    const uint32_t lane = threadIdx.x & 63;
    const uint32_t wave = __builtin_amdgcn_readfirstlane((threadIdx.x & (NW * 64 - 1)) >> 6);
    const uint32_t c0 = a.ch_first + (cblock * NW + wave) * K;
    if (c0 >= a.n_chunks) return;
    const uint32_t nk = min(K, a.n_chunks - c0);
    /* descriptors: lane k holds chunk k's row tag and first edge (lane nk
     * the end of the last chunk) */
    const uint32_t my_tag = a.ch_row[c0 + min(lane, nk - 1)];
    const uint32_t my_edge = a.ch_edge[c0 + min(lane, nk)];
    uint32_t eb[K], ee[K];
    uint32_t longest = 0;
#pragma unroll
    for (uint32_t k = 0; k < K; k++)
    {
        const uint32_t kk = min(k, nk - 1);   /* chunks past nk repeat the last one */
        eb[k] = __builtin_amdgcn_readlane(my_edge, kk);
        ee[k] = __builtin_amdgcn_readlane(my_edge, kk + 1);
        longest = max(longest, ee[k] - eb[k]);
    }
    if (longest <= 64) og_chunk_body<MODE, HOT, 1, K>(a, st, c0, nk, eb, ee, my_tag, hot, nh);
    else if (longest <= 128) og_chunk_body<MODE, HOT, 2, K>(a, st, c0, nk, eb, ee, my_tag, hot, nh);
    else if (longest <= 256) og_chunk_body<MODE, HOT, 4, K>(a, st, c0, nk, eb, ee, my_tag, hot, nh);
    else og_chunk_body<MODE, HOT, OG_CH_SLOTS, K>(a, st, c0, nk, eb, ee, my_tag, hot, nh);
}

/* Rows [0, C) in the reference's order (fp32 pull with every row exact):
 * one wavefront per row, the longest rows first (chain_rows).  The lanes
 * load and gather the row's in-edges OG_CHAIN_U x 64 at a time and stage
 * each group in LDS in edge order; lane 0 then adds the group's values one
 * after the other onto its running sum, four per LDS read: the reference's
 * sequential `+=` over the row (pageRank.c:998-1011, SPMV.c:673-693) to the
 * bit.  Two groups are in flight while one is added: the gathers of the
 * next group and the indices of the one after it.  Groups over which the
 * sum stays in one binade skip the one-by-one adds (og_chain_fast). */
#define OG_CHAIN_U 8u
#define OG_CHAIN_G (64u * OG_CHAIN_U)
#define OG_CHAIN_PRIO_EDGES 8192u

__device__ __forceinline__ void og_chain_idx(const og_pr_dev &a, uint32_t base, uint32_t e,
                                             uint32_t (&idx)[OG_CHAIN_U])
{
    const uint32_t lane = threadIdx.x & 63;
#pragma unroll
    for (uint32_t u = 0; u < OG_CHAIN_U; u++)
        idx[u] = __builtin_nontemporal_load(&a.cols[min(base + u * 64 + lane, e - 1)]);
}

__device__ __forceinline__ void og_chain_gather(const og_pr_dev &a, const uint32_t (&idx)[OG_CHAIN_U],
                                                float (&val)[OG_CHAIN_U])
{
    const float *__restrict__ g = (const float *)a.g_in;
#pragma unroll
    for (uint32_t u = 0; u < OG_CHAIN_U; u++) val[u] = g[idx[u]];
}

/* the wavefront's LDS writes visible to all its lanes, and its reads done
 * before the next writes */
__device__ __forceinline__ void og_wave_sync()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

/* Group values g (edges [base, base + OG_CHAIN_G) of a row ending at e)
 * staged in the wavefront's LDS slice sw and added by lane 0 onto y in
 * edge order. */
__device__ __forceinline__ void og_chain_add(float *sw, const float (&g)[OG_CHAIN_U], uint32_t base, uint32_t e,
                                             uint32_t lane, float &y)
{
#pragma unroll
    for (uint32_t u = 0; u < OG_CHAIN_U; u++) sw[64 * u + lane] = g[u];
    og_wave_sync();
    if (lane == 0)
    {
        const float4 *s4 = (const float4 *)sw;
        if (e - base >= OG_CHAIN_G)
        {
            /* reads of the next 16 values issued before the adds of these
             * 16: the LDS latency stays off the chain of adds */
            float4 c[4], n[4];
#pragma unroll
            for (uint32_t i = 0; i < 4; i++) c[i] = s4[i];
#pragma unroll
            for (uint32_t k = 4; k <= OG_CHAIN_G / 4; k += 4)
            {
                if (k < OG_CHAIN_G / 4)
#pragma unroll
                    for (uint32_t i = 0; i < 4; i++) n[i] = s4[k + i];
#pragma unroll
                for (uint32_t i = 0; i < 4; i++)
                {
                    y = y + c[i].x;
                    y = y + c[i].y;
                    y = y + c[i].z;
                    y = y + c[i].w;
                }
#pragma unroll
                for (uint32_t i = 0; i < 4; i++) c[i] = n[i];
            }
        }
        else
        {
            const uint32_t n = e - base;
            for (uint32_t k = 0; k < n; k++) y = y + sw[k];
        }
    }
    og_wave_sync();
}

/* Exact fast-forward of a chain over one group.  While the running sum y
 * stays in one binade [2^(ey-1), 2^ey), every representable value is a
 * multiple of u = 2^(ey-24), so for a term a >= 0 the rounded sum is
 * fl(y + a) = y + u * m with m the integer nearest a / u -- unless a / u
 * ends in exactly one half (round-half-even then depends on y).  The
 * group's sequential sums therefore equal y + u * (m_1 + ... + m_k), an
 * integer sum in any order, as long as no term is a tie and the last sum
 * stays below 2^ey.  Those conditions are checked here (y uniform across
 * the wavefront); when one fails, nothing changes and the caller adds the
 * group one term at a time. */
__device__ __forceinline__ bool og_chain_fast(const float (&g)[OG_CHAIN_U], uint32_t base, uint32_t e,
                                              uint32_t lane, float &y)
{
    if (!(y >= 0x1p-100f && y < 0x1p100f)) return false;
    int ey;
    frexpf(y, &ey);
    const int sc = 24 - ey;
    uint32_t msum = 0;
    bool bad = false;
#pragma unroll
    for (uint32_t u = 0; u < OG_CHAIN_U; u++)
    {
        const float t = ldexpf(g[u], sc);          /* a / u, exact */
        const float fl = floorf(t);
        const float fr = t - fl;                   /* exact */
        const bool in = base + 64 * u + lane < e;
        bad |= in && (!(t >= 0.0f && t < 0x1p24f) || fr == 0.5f);
        msum += in ? (uint32_t)fl + (fr > 0.5f ? 1u : 0u) : 0u;
    }
    bad |= msum >= (1u << 24);
    if (__any(bad)) return false;
    const uint32_t M = og_wave_allsum(msum);      /* < 2^30 */
    const uint32_t Y = (uint32_t)ldexpf(y, sc);   /* in [2^23, 2^24) */
    if (Y + M >= (1u << 24)) return false;
    y = ldexpf((float)(Y + M), -sc);
    return true;
}

__device__ __forceinline__ void og_pull_chain(const og_pr_dev &a, og_stat &st, uint32_t cblock)
{
    __shared__ __align__(16) float stage[OG_WAVES][OG_CHAIN_G];
    const uint32_t lane = threadIdx.x & 63;
    const uint32_t wave = __builtin_amdgcn_readfirstlane((threadIdx.x & (OG_BLOCK - 1)) >> 6);
    const uint32_t q = cblock * OG_WAVES + wave;
    if (q >= a.n_chains) return;
    const uint32_t v = a.chain_rows ? __builtin_amdgcn_readfirstlane(a.chain_rows[q]) : q;
    const uint32_t j = __builtin_amdgcn_readfirstlane(a.row_ptr[v]);
    const uint32_t e = __builtin_amdgcn_readfirstlane(a.row_ptr[v + 1]);
    const float prev = a.r[v];
    float y = a.spmv ? prev : 0.0f;
    float *sw = stage[wave];
    /* a chain thousands of edges long bounds the whole launch: its
     * wavefront takes the SIMD's issue slots before the others */
    const bool lead = e - j > OG_CHAIN_PRIO_EDGES;
    if (lead) __builtin_amdgcn_s_setprio(3);
    /* groups past the row's end re-read its last edge; they are never added */
    uint32_t i1[OG_CHAIN_U], i2[OG_CHAIN_U];
    float g0[OG_CHAIN_U], g1[OG_CHAIN_U];
    og_chain_idx(a, j, e, i1);
    og_chain_gather(a, i1, g0);
    og_chain_idx(a, j + OG_CHAIN_G, e, i1);
    for (uint32_t base = j; base < e; base += OG_CHAIN_G)
    {
        og_chain_gather(a, i1, g1);                     /* group + 1 */
        og_chain_idx(a, base + 2 * OG_CHAIN_G, e, i2);  /* group + 2 */
        if (!og_chain_fast(g0, base, e, lane, y))
        {
            og_chain_add(sw, g0, base, e, lane, y);
            y = __int_as_float(__builtin_amdgcn_readfirstlane(__float_as_int(y)));   /* lane 0's sum */
        }
#pragma unroll
        for (uint32_t u = 0; u < OG_CHAIN_U; u++)
        {
            g0[u] = g1[u];
            i1[u] = i2[u];
        }
    }
    if (lead) __builtin_amdgcn_s_setprio(0);
    if (lane == 0)
    {
        if (a.spmv) og_spmv_store<OG_MODE_F32>(a, v, y, prev, true);
        else og_epilogue_pf<OG_MODE_F32>(a, v, y, prev, a.outdeg[v], st);
    }
}

/* One group of a lane-read chain: fast-forwarded when exact, else its
 * values reach the running sum (uniform across the wavefront) through
 * lane reads -- eight reads into separate scalars, then their eight adds,
 * so no read-after-write wait separates each pair. */
__device__ __forceinline__ void og_chain_lanes_add(const float (&g)[OG_CHAIN_U], uint32_t base, uint32_t e,
                                                   uint32_t lane, float &y)
{
    if (og_chain_fast(g, base, e, lane, y)) return;
#pragma unroll
    for (uint32_t u = 0; u < OG_CHAIN_U; u++)
    {
        const uint32_t s = base + 64 * u;
        if (s >= e) break;
        const int x = __float_as_int(g[u]);
        if (e - s >= 64)
        {
#pragma unroll
            for (int k0 = 0; k0 < 64; k0 += 8)
            {
                int t[8];
#pragma unroll
                for (int k = 0; k < 8; k++) t[k] = __builtin_amdgcn_readlane(x, k0 + k);
#pragma unroll
                for (int k = 0; k < 8; k++) y = y + __int_as_float(t[k]);
            }
        }
        else
        {
            const int n = (int)(e - s);
            for (int k = 0; k < n; k++) y = y + __int_as_float(__builtin_amdgcn_readlane(x, k));
        }
    }
}

/* Gathers of one group into gn (indices in idx), then the indices of the
 * group `ahead` groups later into idx. */
template <bool HOT>
__device__ __forceinline__ void og_chain_issue(const og_pr_dev &a, const float *hot, uint32_t nh,
                                               uint32_t next_base, uint32_t e, uint32_t (&idx)[OG_CHAIN_U],
                                               float (&gn)[OG_CHAIN_U])
{
    const float *__restrict__ g = (const float *)a.g_in;
#pragma unroll
    for (uint32_t u = 0; u < OG_CHAIN_U; u++) gn[u] = og_gather<float, HOT>(g, hot, nh, idx[u]);
    og_chain_idx(a, next_base, e, idx);
}

/* The chains of a pull that also has chunks (exact_above): the longest rows
 * only, so the time of the longest bounds the launch.  Most of its groups
 * take the exact fast-forward, which leaves the gathers as the limit: three
 * groups rotate through three register sets (no copies, so no step waits
 * for loads younger than its own group's), two in flight while one is
 * added. */
template <bool HOT>
__device__ __forceinline__ void og_pull_chain_lanes(const og_pr_dev &a, og_stat &st, uint32_t q,
                                                    const float *hot, uint32_t nh)
{
    constexpr uint32_t G = OG_CHAIN_G;
    const uint32_t lane = threadIdx.x & 63;
    if (q >= a.n_chains) return;
    const uint32_t v = a.chain_rows ? __builtin_amdgcn_readfirstlane(a.chain_rows[q]) : q;
    const uint32_t j = __builtin_amdgcn_readfirstlane(a.row_ptr[v]);
    const uint32_t e = __builtin_amdgcn_readfirstlane(a.row_ptr[v + 1]);
    const float prev = a.r[v];
    float y = a.spmv ? prev : 0.0f;
    const bool lead = e - j > OG_CHAIN_PRIO_EDGES;
    if (lead) __builtin_amdgcn_s_setprio(3);
    /* groups past the row's end re-read its last edge; they are never added */
    uint32_t idx[OG_CHAIN_U];
    float g0[OG_CHAIN_U], g1[OG_CHAIN_U], g2[OG_CHAIN_U];
    og_chain_idx(a, j, e, idx);
    og_chain_issue<HOT>(a, hot, nh, j + G, e, idx, g0);
    og_chain_issue<HOT>(a, hot, nh, j + 2 * G, e, idx, g1);
    for (uint32_t base = j; base < e; base += 3 * G)
    {
        og_chain_issue<HOT>(a, hot, nh, base + 3 * G, e, idx, g2);
        og_chain_lanes_add(g0, base, e, lane, y);
        if (base + G >= e) break;
        og_chain_issue<HOT>(a, hot, nh, base + 4 * G, e, idx, g0);
        og_chain_lanes_add(g1, base + G, e, lane, y);
        if (base + 2 * G >= e) break;
        og_chain_issue<HOT>(a, hot, nh, base + 5 * G, e, idx, g1);
        og_chain_lanes_add(g2, base + 2 * G, e, lane, y);
    }
    if (lead) __builtin_amdgcn_s_setprio(0);
    if (lane == 0)
    {
        if (a.spmv) og_spmv_store<OG_MODE_F32>(a, v, y, prev, true);
        else og_epilogue_pf<OG_MODE_F32>(a, v, y, prev, a.outdeg[v], st);
    }
}

/* In-edges [j, e) of one row added in CSR order onto s, B loaded and
 * gathered at a time.  Lanes past their row's end re-read its last edge and
 * add zero: x + 0 == x for every partial sum (PageRank's start at +0; the
 * SpMV's running y is never -0, its terms being >= 0). */
template <int MODE, bool HOT, uint32_t B>
__device__ __forceinline__ typename og_mode<MODE>::acc og_row_sum(const og_pr_dev &a, uint32_t j,
                                                                  uint32_t e,
                                                                  const typename og_mode<MODE>::elem *hot,
                                                                  uint32_t nh,
                                                                  typename og_mode<MODE>::acc s)
{
    typedef typename og_mode<MODE>::elem E;
    typedef typename og_mode<MODE>::acc A;
    const E *__restrict__ g = (const E *)a.g_in;
    for (; j < e; j += B)
    {
        uint32_t idx[B];
#pragma unroll
        for (uint32_t u = 0; u < B; u++) idx[u] = __builtin_nontemporal_load(&a.cols[min(j + u, e - 1)]);
        E val[B];
#pragma unroll
        for (uint32_t u = 0; u < B; u++) val[u] = og_gather<E, HOT>(g, hot, nh, idx[u]);
#pragma unroll
        for (uint32_t u = 0; u < B; u++) s += j + u < e ? (A)val[u] : (A)0;
    }
    return s;
}

/* Rows [C, row_end), one per thread, each summed in CSR order
 * (pageRank.c:998-1011), so the sum is the reference's to the bit.  The
 * loads per batch follow the wavefront's longest row (1, 2, 4 or
 * OG_ROW_BATCH): most short rows of a power-law graph have one to four
 * in-edges, and rows are sorted by in-degree, so the 64 rows of a
 * wavefront are about as long.  ADAPT = false keeps RB (fewer registers,
 * for the multi-graph kernels; the sums are the same).  A block has BLK
 * threads, one row each. */
template <int MODE, bool HOT = false, bool ADAPT = true, uint32_t BLK = OG_BLOCK, uint32_t RB = OG_ROW_BATCH>
__device__ __forceinline__ void og_pull_rows(const og_pr_dev &a, og_stat &st, uint32_t rblock,
                                             const typename og_mode<MODE>::elem *hot = nullptr,
                                             uint32_t nh = 0)
{
    typedef typename og_mode<MODE>::acc A;
    const uint32_t v = a.C + rblock * BLK + (threadIdx.x & (BLK - 1));
    const bool mine = v < a.row_end && !(a.dd && !a.wl_cur[v]);
    uint32_t j = 0, e = 0;
    float prev = 0.0f;
    uint32_t deg = 0;
    if (mine)
    {
        j = __builtin_nontemporal_load(&a.row_ptr[v]);
        e = __builtin_nontemporal_load(&a.row_ptr[v + 1]);
        prev = __builtin_nontemporal_load(&a.r[v]);
        if (!a.spmv) deg = __builtin_nontemporal_load(&a.outdeg[v]);
    }
    /* longest row of the wavefront (every lane still active here) */
    uint32_t m = RB;
    if constexpr (ADAPT)
    {
        m = e - j;
        for (int o = 32; o > 0; o >>= 1) m = max(m, (uint32_t)__shfl_xor(m, o, 64));
    }
    if (!mine) return;
    /* SpMV continues the running fp32 y of the row, in the reference's order */
    A s = 0;
    if constexpr (MODE == OG_MODE_F32)
        if (a.spmv) s = prev;
    if (m <= 1) s = og_row_sum<MODE, HOT, 1>(a, j, e, hot, nh, s);
    else if (m <= 2) s = og_row_sum<MODE, HOT, 2>(a, j, e, hot, nh, s);
    else if (m <= 4) s = og_row_sum<MODE, HOT, 4>(a, j, e, hot, nh, s);
    else s = og_row_sum<MODE, HOT, ADAPT ? OG_ROW_BATCH : RB>(a, j, e, hot, nh, s);
    if (a.spmv) og_spmv_store<MODE>(a, v, s, prev, true);
    else if (a.dd) og_epilogue<MODE>(a, v, s, st);
    else og_epilogue_pf<MODE>(a, v, s, prev, deg, st);
}

template <int MODE>
__global__ __launch_bounds__(OG_BLOCK) void og_pull_kernel(og_pr_dev a, uint32_t it)
{
    if (*a.done) return;
    og_stat st;
    og_stat_init(st);
    const uint32_t nc = a.n_chain_blocks;
    if (blockIdx.x >= nc + a.n_chunk_blocks) og_pull_rows<MODE>(a, st, blockIdx.x - nc - a.n_chunk_blocks);
    else if (blockIdx.x >= nc) og_pull_chunks<MODE>(a, st, blockIdx.x - nc);
    else if constexpr (MODE == OG_MODE_F32)
        og_pull_chain_lanes<false>(a, st, blockIdx.x * OG_WAVES + __builtin_amdgcn_readfirstlane(threadIdx.x >> 6),
                                   nullptr, 0);
    og_stat r = og_block_stat(st);
    if (threadIdx.x == 0)
    {
        a.part_err[blockIdx.x] = r.err;
        a.part_act[blockIdx.x] = r.active;
        a.part_min[blockIdx.x] = r.cmin;
        a.part_max[blockIdx.x] = r.cmax;
    }
}

/* The chains of a pull whose other rows run in the persistent kernel (or
 * of a push whose other destinations are scattered): a launch of its own on
 * a side stream, concurrent with it.  One wavefront per block, at most one
 * block per CU (grid = n_chain_blocks <= CUs), each taking chains q = b,
 * b + grid, ...; no LDS and at most 96 registers, so a wavefront fits on a
 * SIMD beside the persistent block's four.  Its statistics partials start
 * at chain_part_base, after the other kernel's. */
__global__ __launch_bounds__(64, 5) void og_pull_chains_kernel(og_pr_dev a, uint32_t it)
{
    if (*a.done) return;
    og_stat st;
    og_stat_init(st);
    for (uint32_t q = blockIdx.x; q < a.n_chains; q += gridDim.x) og_pull_chain_lanes<false>(a, st, q, nullptr, 0);
    og_stat r = og_block_stat<1>(st);
    if (threadIdx.x == 0)
    {
        const uint32_t b = a.chain_part_base + blockIdx.x;
        a.part_err[b] = r.err;
        a.part_act[b] = r.active;
        a.part_min[b] = r.cmin;
        a.part_max[b] = r.cmax;
    }
}

/* fp32 pull with every row exact: chains and rows only (no chunk path, so
 * fewer registers and more wavefronts in flight per SIMD) */
__global__ __launch_bounds__(OG_BLOCK) void og_pull_exact_kernel(og_pr_dev a, uint32_t it)
{
    if (*a.done) return;
    og_stat st;
    og_stat_init(st);
    if (blockIdx.x >= a.n_chain_blocks) og_pull_rows<OG_MODE_F32>(a, st, blockIdx.x - a.n_chain_blocks);
    else og_pull_chain(a, st, blockIdx.x);
    og_stat r = og_block_stat(st);
    if (threadIdx.x == 0)
    {
        a.part_err[blockIdx.x] = r.err;
        a.part_act[blockIdx.x] = r.active;
        a.part_min[blockIdx.x] = r.cmin;
        a.part_max[blockIdx.x] = r.cmax;
    }
}

/* The same pull as a persistent kernel, one block per CU, whose blocks
 * first copy the OG_HOT_BYTES of contributions with the smallest ids --
 * after the degree relabelling, those of the vertices with the most in- and
 * out-edges -- into LDS and serve those gathers from there.  A block is
 * OG_HOT_SUBS groups of 256 threads; group q of block b takes the work
 * blocks S*b + q, S*b + q + S*grid, ... (S = OG_HOT_SUBS) with no barrier
 * between groups; statistics partial per block. */
template <int MODE>
__global__ __launch_bounds__(OG_HOT_SUBS *OG_BLOCK) void og_pull_hot_kernel(og_pr_dev a, uint32_t it)
{
    typedef typename og_mode<MODE>::elem E;
    constexpr uint32_t NH = OG_HOT_BYTES / sizeof(E);
    __shared__ E hot[NH];
    if (*a.done) return;
    const E *__restrict__ g = (const E *)a.g_in;
    const uint32_t nh = min(NH, a.N);
    for (uint32_t i = threadIdx.x; i < nh; i += OG_HOT_SUBS * OG_BLOCK) hot[i] = g[i];
    __syncthreads();
    og_stat st;
    og_stat_init(st);
    /* (chains run beside this kernel, in og_pull_chains_kernel) */
    const uint32_t items = a.n_chunk_blocks + a.n_row_blocks;
    const uint32_t q = threadIdx.x / OG_BLOCK;
    for (uint32_t w = blockIdx.x * OG_HOT_SUBS + q; w < items; w += gridDim.x * OG_HOT_SUBS)
    {
        if (w >= a.n_chunk_blocks) og_pull_rows<MODE, true>(a, st, w - a.n_chunk_blocks, hot, nh);
        else og_pull_chunks<MODE, true>(a, st, w, hot, nh);
    }
    og_stat r = og_block_stat<OG_HOT_SUBS * OG_WAVES>(st);
    if (threadIdx.x == 0)
    {
        a.part_err[blockIdx.x] = r.err;
        a.part_act[blockIdx.x] = r.active;
        a.part_min[blockIdx.x] = r.cmin;
        a.part_max[blockIdx.x] = r.cmax;
    }
}

/* Final reduction shared by pull and push: block b folds a fixed slice of
 * the first-level partials (plus, for pull, the long rows it finalizes);
 * the last block to arrive folds the block results in index order and
 * writes the iteration's statistics, the quantization parameters of the
 * next iteration and the stop flag. */
__device__ void og_reduce_tail(const og_pr_dev &a, uint32_t it, og_stat mine)
{
    const uint32_t t = threadIdx.x;
    const uint32_t per = (a.n_parts + gridDim.x - 1) / gridDim.x;
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
        tot.active += a.z_active;
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

/* Row h < H of one wavefront: lane l adds chunks l, l+64, ... of the row in
 * order (four loads in flight; lanes past the end re-read the last chunk
 * and add zero), then a fixed tree over the wavefront; lane 0 runs the
 * epilogue. */
template <int MODE>
__device__ __forceinline__ void og_finalize_long_row(const og_pr_dev &a, uint32_t h, og_stat &st)
{
    typedef typename og_mode<MODE>::acc A;
    const A *__restrict__ cs = (const A *)a.chunk_sum;
    const uint32_t lane = threadIdx.x & 63;
    if (a.dd && !a.wl_cur[h]) return;
    const uint32_t cb = a.chunk_base[h], ce = a.chunk_base[h + 1];
    A s = 0;
    for (uint32_t c = cb + lane; c < ce; c += 256)
    {
        A x[4];
#pragma unroll
        for (uint32_t u = 0; u < 4; u++) x[u] = cs[min(c + 64 * u, ce - 1)];
#pragma unroll
        for (uint32_t u = 0; u < 4; u++) s += c + 64 * u < ce ? x[u] : (A)0;
    }
    s = og_wave_sum(s);
    if (lane == 0) og_epilogue<MODE>(a, h, s, st);
}

template <int MODE>
__global__ __launch_bounds__(OG_BLOCK) void og_pull_finalize_kernel(og_pr_dev a, uint32_t it)
{
    if (*a.done) return;
    og_stat st;
    og_stat_init(st);
    const uint32_t wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    /* (rows [0, n_chains) of the multi-chunk class are chains, already done) */
    for (uint32_t h = a.n_chains + blockIdx.x * OG_WAVES + wave; h < a.H; h += gridDim.x * OG_WAVES)
        og_finalize_long_row<MODE>(a, h, st);
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

/* Scatter of the contributions over the forward edges, E atomic adds into
 * next[].  With HOT, one 1024-thread block per CU keeps the OG_HOT_BYTES of
 * accumulators of the destinations with the smallest ids -- after the
 * relabelling, those with the most in-edges -- in LDS, adds their edges
 * there and flushes each nonzero accumulator with one global atomic at the
 * end: the hottest destinations then see one global atomic per block
 * instead of one per edge.  fp32 adds are unordered either way (as the
 * reference's omp atomic); the u64 fixed-point sum is exact in any order. */
template <int PMODE, bool HOT>
__global__ __launch_bounds__(HOT ? OG_HOT_SUBS * OG_BLOCK : OG_BLOCK) void og_push_scatter_kernel(og_pr_dev a)
{
    typedef typename std::conditional<PMODE == OG_MODE_FX64, unsigned long long, float>::type T;
    constexpr uint32_t NH = HOT ? OG_HOT_BYTES / sizeof(T) : 1;
    __shared__ T acc[NH];
    if (*a.done) return;
    const uint32_t E = a.E, nb = blockDim.x;
    const uint32_t nh = HOT ? min(NH, a.N) : 0;
    T *__restrict__ next = (T *)a.next;
    if constexpr (HOT)
    {
        for (uint32_t i = threadIdx.x; i < nh; i += nb) acc[i] = 0;
        __syncthreads();
    }
    /* four edges per thread in flight: loads first (lanes past the end
     * re-read the last edge), then the adds of the valid ones */
    constexpr uint32_t B = 4;
    const uint32_t stride = gridDim.x * nb;
    for (uint32_t e0 = blockIdx.x * nb + threadIdx.x; e0 < E; e0 += B * stride)
    {
        uint32_t s[B], d[B];
#pragma unroll
        for (uint32_t u = 0; u < B; u++)
        {
            const uint32_t e = min(e0 + u * stride, E - 1);
            s[u] = __builtin_nontemporal_load(&a.push_src[e]);
            d[u] = __builtin_nontemporal_load(&a.push_dst[e]);
        }
        T add[B];
#pragma unroll
        for (uint32_t u = 0; u < B; u++)
        {
            if constexpr (PMODE == OG_MODE_F32) add[u] = ((const float *)a.g_in)[s[u]];
            else if constexpr (PMODE == OG_MODE_FX64) add[u] = (T)((const uint64_t *)a.g_in)[s[u]];
            else
            {
                uint32_t qv = ((const uint32_t *)a.g_in)[s[u]];
                add[u] = (float)(a.qp->scale_d * (double)(uint32_t)(qv - a.qp->zero));
            }
        }
#pragma unroll
        for (uint32_t u = 0; u < B; u++)
        {
            if (e0 + u * stride >= E) break;
            if (d[u] < a.push_skip) continue;
            if (HOT && d[u] < nh) atomicAdd(&acc[d[u]], add[u]);
            else if constexpr (PMODE == OG_MODE_FX64) atomicAdd(&next[d[u]], add[u]);
            else unsafeAtomicAdd(&next[d[u]], add[u]);
        }
    }
    if constexpr (HOT)
    {
        __syncthreads();
        for (uint32_t i = threadIdx.x; i < nh; i += nb)
        {
            const T x = acc[i];
            if (x == (T)0) continue;   /* contributions are >= 0: nothing to add */
            if constexpr (PMODE == OG_MODE_FX64) atomicAdd(&next[i], x);
            else unsafeAtomicAdd(&next[i], x);
        }
    }
}

/* The update of vertex v from its push sum s (pageRank.c:1117-1134 and
 * the next contribution): u64 fixed point, fp32, or the quantized push's
 * fp32 update with the fp32 contribution and its min / max. */
template <int PMODE, typename T>
__device__ __forceinline__ void og_push_vertex(const og_pr_dev &a, uint32_t v, T s, og_stat &st)
{
    if constexpr (PMODE == OG_MODE_FX64) og_epilogue<OG_MODE_FX64>(a, v, (uint64_t)s, st);
    else if constexpr (PMODE == OG_MODE_F32) og_epilogue<OG_MODE_F32>(a, v, (float)s, st);
    else
    {
        float prev = a.r[v];
        float nr = OG_BASE_PR + OG_DAMP * (float)s;
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

template <int PMODE>
__global__ __launch_bounds__(OG_BLOCK) void og_push_epilogue_kernel(og_pr_dev a)
{
    if (*a.done) return;
    og_stat st;
    og_stat_init(st);
    for (uint32_t v = a.push_skip + blockIdx.x * OG_BLOCK + threadIdx.x; v < a.N; v += gridDim.x * OG_BLOCK)
    {
        if constexpr (PMODE == OG_MODE_FX64)
        {
            uint64_t s = ((uint64_t *)a.next)[v];
            ((uint64_t *)a.next)[v] = 0;
            og_push_vertex<PMODE>(a, v, s, st);
        }
        else
        {
            float s = ((float *)a.next)[v];
            ((float *)a.next)[v] = 0.0f;
            og_push_vertex<PMODE>(a, v, s, st);
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

/* The partitioned push: block b takes destination range p = pp_first + b,
 * rows [pp_v0[p], pp_v0[p + 1]) of the plan, with one accumulator per
 * destination in LDS.  Its in-edges -- the inverse CSR's edges of those
 * rows, so a contiguous range -- are visited OG_PP_B per thread at a time,
 * and each adds its source's contribution to its destination's accumulator
 * with an LDS atomic (pp_dloc[e]: the destination's offset in the range):
 * the reference's atomic scatter (pageRank.c:1091-1112), with the atomics
 * kept in LDS.  Each destination still has one running sum to which its
 * terms are added one at a time, in the order the atomics land, as the
 * reference's OpenMP atomics do; the u64 fixed-point sums are exact in any
 * order.  The block then runs the update of its rows (as
 * og_push_epilogue_kernel) and writes its statistics partial.  SPMV: one
 * SpMV push pass (SPMV.c): the accumulators start from the running output
 * next[v] (fp32 y, or the u64 sums), the terms are added onto it as the
 * reference's atomics add them onto y, and it is stored back; no update. */
#define OG_PP_THREADS 512u
#define OG_PP_B 4u
template <int PMODE, bool SPMV>
__global__ __launch_bounds__(OG_PP_THREADS) void og_push_part_kernel(og_pr_dev a)
{
    typedef typename std::conditional<PMODE == OG_MODE_FX64, unsigned long long, float>::type T;
    __shared__ T acc[OG_PP_ROWS];
    if (*a.done) return;
    const uint32_t p = a.pp_first + blockIdx.x;
    const uint32_t v0 = a.pp_v0[p], nv = a.pp_v0[p + 1] - v0;
    const uint32_t e0 = a.row_ptr[v0], e1 = a.row_ptr[v0 + nv];
    for (uint32_t i = threadIdx.x; i < nv; i += OG_PP_THREADS) acc[i] = SPMV ? ((const T *)a.next)[v0 + i] : (T)0;
    __syncthreads();
    for (uint32_t e = e0 + threadIdx.x; e < e1; e += OG_PP_B * OG_PP_THREADS)
    {
        uint32_t src[OG_PP_B], dl[OG_PP_B];
#pragma unroll
        for (uint32_t k = 0; k < OG_PP_B; k++)
        {
            const uint32_t ek = min(e + k * OG_PP_THREADS, e1 - 1);   /* past the end: re-read, not added */
            src[k] = __builtin_nontemporal_load(&a.cols[ek]);
            dl[k] = __builtin_nontemporal_load(&a.pp_dloc[ek]);
        }
        T add[OG_PP_B];
#pragma unroll
        for (uint32_t k = 0; k < OG_PP_B; k++)
        {
            if constexpr (PMODE == OG_MODE_F32) add[k] = ((const float *)a.g_in)[src[k]];
            else if constexpr (PMODE == OG_MODE_FX64) add[k] = (T)((const uint64_t *)a.g_in)[src[k]];
            else
            {
                const uint32_t qv = ((const uint32_t *)a.g_in)[src[k]];
                add[k] = (float)(a.qp->scale_d * (double)(uint32_t)(qv - a.qp->zero));
            }
        }
#pragma unroll
        for (uint32_t k = 0; k < OG_PP_B; k++)
            if (e + k * OG_PP_THREADS < e1) atomicAdd(&acc[dl[k]], add[k]);
    }
    __syncthreads();
    if constexpr (SPMV)
    {
        for (uint32_t i = threadIdx.x; i < nv; i += OG_PP_THREADS) ((T *)a.next)[v0 + i] = acc[i];
        return;
    }
    og_stat st;
    og_stat_init(st);
    for (uint32_t i = threadIdx.x; i < nv; i += OG_PP_THREADS) og_push_vertex<PMODE>(a, v0 + i, acc[i], st);
    og_stat r = og_block_stat<OG_PP_THREADS / 64>(st);
    if (threadIdx.x == 0)
    {
        a.part_err[blockIdx.x] = r.err;
        a.part_act[blockIdx.x] = r.active;
        a.part_min[blockIdx.x] = r.cmin;
        a.part_max[blockIdx.x] = r.cmax;
    }
}

/* pp_dloc[e] = (row of edge e) - pp_v0[p] for the edges of range p (one
 * block per range; the row by binary search over its row pointers). */
__global__ __launch_bounds__(OG_BLOCK) void og_push_dloc_kernel(const uint32_t *__restrict__ row_ptr,
                                                               const uint32_t *__restrict__ pp_v0,
                                                               uint16_t *__restrict__ dloc)
{
    const uint32_t v0 = pp_v0[blockIdx.x], v1 = pp_v0[blockIdx.x + 1];
    const uint32_t e0 = row_ptr[v0], e1 = row_ptr[v1];
    for (uint32_t e = e0 + threadIdx.x; e < e1; e += OG_BLOCK)
    {
        uint32_t lo = v0, hi = v1;   /* row_ptr[lo] <= e < row_ptr[hi] */
        while (hi - lo > 1)
        {
            const uint32_t mid = (lo + hi) / 2;
            if (row_ptr[mid] <= e) lo = mid;
            else hi = mid;
        }
        dloc[e] = (uint16_t)(lo - v0);
    }
}

extern "C" int og_launch_push_dloc(const uint32_t *row_ptr, const uint32_t *pp_v0, uint32_t n_pp, uint16_t *dloc,
                                   hipStream_t s)
{
    if (n_pp) OG_LAUNCH(og_push_dloc_kernel, dim3(n_pp), dim3(OG_BLOCK), 0, s, row_ptr, pp_v0, dloc);
    return og_hip_check(og_launch_status(), "og_push_dloc_kernel");
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
        float r = init_old ? init_old[a.new2old[v]] : a.ppr ? (v == a.ppr_src ? 1.0f : 0.0f) : init_value;
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
    const float fn = a.ppr ? 1.0f : (float)a.N;   /* personalized: not divided by N */
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

template <int MODE> static int og_launch_pull_t(const og_pr_dev *a, uint32_t it, hipStream_t s,
                                                hipEvent_t ev0, hipEvent_t ev1)
{
    uint32_t grid = a->n_chain_blocks + a->n_chunk_blocks + a->n_row_blocks;
    if (ev0) hipEventRecord(ev0, s);
    if (a->hot_grid && a->n_chain_blocks)
    {
        /* fork: chains on the side stream beside the persistent kernel, joined
         * before the finalize */
        hipStream_t side = og_side_stream();
        if (og_hip_check(hipEventRecord(og_side_fork(), s), "hipEventRecord") ||
            og_hip_check(hipStreamWaitEvent(side, og_side_fork(), 0), "hipStreamWaitEvent"))
            return -1;
        og_pr_dev b = *a;
        b.chain_part_base = a->hot_grid;
        OG_LAUNCH(og_pull_chains_kernel, dim3(a->n_chain_blocks), dim3(64), 0, side, b, it);
        OG_LAUNCH(og_pull_hot_kernel<MODE>, dim3(a->hot_grid), dim3(OG_HOT_SUBS * OG_BLOCK), 0,
                           s, *a, it);
        if (og_hip_check(hipEventRecord(og_side_join(), side), "hipEventRecord") ||
            og_hip_check(hipStreamWaitEvent(s, og_side_join(), 0), "hipStreamWaitEvent"))
            return -1;
    }
    else if (a->hot_grid)
        OG_LAUNCH(og_pull_hot_kernel<MODE>, dim3(a->hot_grid), dim3(OG_HOT_SUBS * OG_BLOCK), 0,
                           s, *a, it);
    else if (MODE == OG_MODE_F32 && a->chain_rows && !a->n_chunk_blocks && grid)
        OG_LAUNCH(og_pull_exact_kernel, dim3(grid), dim3(OG_BLOCK), 0, s, *a, it);
    else if (grid)
        OG_LAUNCH(og_pull_kernel<MODE>, dim3(grid), dim3(OG_BLOCK), 0, s, *a, it);
    if (ev1) hipEventRecord(ev1, s);
    OG_LAUNCH(og_pull_finalize_kernel<MODE>, dim3(a->n_final_blocks), dim3(OG_BLOCK), 0, s,
                       *a, it);
    return 0;
}

extern "C" int og_launch_pull(const og_pr_dev *a, int mode, uint32_t it, hipStream_t s,
                              hipEvent_t ev0, hipEvent_t ev1)
{
    int rc;
    switch (mode)
    {
    case OG_MODE_F32: rc = og_launch_pull_t<OG_MODE_F32>(a, it, s, ev0, ev1); break;
    case OG_MODE_FX64: rc = og_launch_pull_t<OG_MODE_FX64>(a, it, s, ev0, ev1); break;
    case OG_MODE_FX32: rc = og_launch_pull_t<OG_MODE_FX32>(a, it, s, ev0, ev1); break;
    case OG_MODE_FX16: rc = og_launch_pull_t<OG_MODE_FX16>(a, it, s, ev0, ev1); break;
    case OG_MODE_FX8: rc = og_launch_pull_t<OG_MODE_FX8>(a, it, s, ev0, ev1); break;
    case OG_MODE_Q32: rc = og_launch_pull_t<OG_MODE_Q32>(a, it, s, ev0, ev1); break;
    case OG_MODE_Q16: rc = og_launch_pull_t<OG_MODE_Q16>(a, it, s, ev0, ev1); break;
    case OG_MODE_Q8: rc = og_launch_pull_t<OG_MODE_Q8>(a, it, s, ev0, ev1); break;
    default: og_set_error("og_launch_pull: bad mode %d", mode); return -1;
    }
    if (rc) return rc;
    return og_hip_check(og_launch_status(), "og_pull_kernel");
}

extern "C" int og_launch_quantize(const og_pr_dev *a, int mode, hipStream_t s)
{
    unsigned grid = (a->N + OG_BLOCK - 1) / OG_BLOCK;
    if (grid > 4096) grid = 4096;
    if (grid == 0) grid = 1;
    switch (mode)
    {
    case OG_MODE_Q32: OG_LAUNCH(og_quantize_kernel<OG_MODE_Q32>, dim3(grid), dim3(OG_BLOCK), 0, s, *a); break;
    case OG_MODE_Q16: OG_LAUNCH(og_quantize_kernel<OG_MODE_Q16>, dim3(grid), dim3(OG_BLOCK), 0, s, *a); break;
    case OG_MODE_Q8: OG_LAUNCH(og_quantize_kernel<OG_MODE_Q8>, dim3(grid), dim3(OG_BLOCK), 0, s, *a); break;
    default: OG_LAUNCH(og_quantize_kernel<OG_MODE_QPUSH_EPI>, dim3(grid), dim3(OG_BLOCK), 0, s, *a); break;
    }
    return og_hip_check(og_launch_status(), "og_quantize_kernel");
}

extern "C" int og_launch_push(const og_pr_dev *a, int pmode, uint32_t it, hipStream_t s,
                              hipEvent_t ev0, hipEvent_t ev1)
{
    unsigned gs = (a->E + OG_BLOCK - 1) / OG_BLOCK;
    if (gs > 16384) gs = 16384;
    if (gs == 0) gs = 1;
    if (ev0) hipEventRecord(ev0, s);
    if (a->pp_v0)
    {
        /* partitioned push; fp32 with chains: the chain destinations gathered
         * on the side stream, their partials after the ranges' */
        const uint32_t nr = a->n_pp - a->pp_first;
        const uint32_t nc = pmode == OG_MODE_F32 && a->push_skip ? a->n_chain_blocks : 0;
        hipStream_t side = og_side_stream();
        if (nc)
        {
            og_pr_dev b = *a;
            b.chain_part_base = nr;
            if (og_hip_check(hipEventRecord(og_side_fork(), s), "hipEventRecord") ||
                og_hip_check(hipStreamWaitEvent(side, og_side_fork(), 0), "hipStreamWaitEvent"))
                return -1;
            OG_LAUNCH(og_pull_chains_kernel, dim3(nc), dim3(64), 0, side, b, it);
        }
        if (nr)
        {
            if (pmode == OG_MODE_F32)
                OG_LAUNCH((og_push_part_kernel<OG_MODE_F32, false>), dim3(nr), dim3(OG_PP_THREADS), 0, s, *a);
            else if (pmode == OG_MODE_FX64)
                OG_LAUNCH((og_push_part_kernel<OG_MODE_FX64, false>), dim3(nr), dim3(OG_PP_THREADS), 0, s, *a);
            else
                OG_LAUNCH((og_push_part_kernel<OG_MODE_QPUSH_EPI, false>), dim3(nr), dim3(OG_PP_THREADS), 0, s, *a);
        }
        if (nc && (og_hip_check(hipEventRecord(og_side_join(), side), "hipEventRecord") ||
                   og_hip_check(hipStreamWaitEvent(s, og_side_join(), 0), "hipStreamWaitEvent")))
            return -1;
        if (ev1) hipEventRecord(ev1, s);
        OG_LAUNCH(og_final_kernel, dim3(a->n_final_blocks), dim3(OG_BLOCK), 0, s, *a, it);
        return og_hip_check(og_launch_status(), "og_push_part_kernel");
    }
    if (pmode == OG_MODE_F32)
    {
        /* the chain destinations gathered on the side stream while the
         * others are scattered; their partials follow the epilogue's */
        const uint32_t nc = a->push_skip ? a->n_chain_blocks : 0;
        hipStream_t side = og_side_stream();
        if (nc)
        {
            og_pr_dev b = *a;
            b.chain_part_base = a->n_parts - nc;
            if (og_hip_check(hipEventRecord(og_side_fork(), s), "hipEventRecord") ||
                og_hip_check(hipStreamWaitEvent(side, og_side_fork(), 0), "hipStreamWaitEvent"))
                return -1;
            OG_LAUNCH(og_pull_chains_kernel, dim3(nc), dim3(64), 0, side, b, it);
        }
        if (a->hot_grid)
            OG_LAUNCH((og_push_scatter_kernel<OG_MODE_F32, true>), dim3(a->hot_grid),
                               dim3(OG_HOT_SUBS * OG_BLOCK), 0, s, *a);
        else
            OG_LAUNCH((og_push_scatter_kernel<OG_MODE_F32, false>), dim3(gs), dim3(OG_BLOCK), 0, s, *a);
        if (ev1) hipEventRecord(ev1, s);
        OG_LAUNCH(og_push_epilogue_kernel<OG_MODE_F32>, dim3(a->n_parts - nc), dim3(OG_BLOCK), 0, s, *a);
        if (nc && (og_hip_check(hipEventRecord(og_side_join(), side), "hipEventRecord") ||
                   og_hip_check(hipStreamWaitEvent(s, og_side_join(), 0), "hipStreamWaitEvent")))
            return -1;
    }
    else if (pmode == OG_MODE_FX64)
    {
        if (a->hot_grid)
            OG_LAUNCH((og_push_scatter_kernel<OG_MODE_FX64, true>), dim3(a->hot_grid),
                               dim3(OG_HOT_SUBS * OG_BLOCK), 0, s, *a);
        else
            OG_LAUNCH((og_push_scatter_kernel<OG_MODE_FX64, false>), dim3(gs), dim3(OG_BLOCK), 0, s, *a);
        if (ev1) hipEventRecord(ev1, s);
        OG_LAUNCH(og_push_epilogue_kernel<OG_MODE_FX64>, dim3(a->n_parts), dim3(OG_BLOCK), 0, s, *a);
    }
    else
    {
        if (a->hot_grid)
            OG_LAUNCH((og_push_scatter_kernel<OG_MODE_QPUSH_EPI, true>), dim3(a->hot_grid),
                               dim3(OG_HOT_SUBS * OG_BLOCK), 0, s, *a);
        else
            OG_LAUNCH((og_push_scatter_kernel<OG_MODE_QPUSH_EPI, false>), dim3(gs), dim3(OG_BLOCK), 0,
                               s, *a);
        if (ev1) hipEventRecord(ev1, s);
        OG_LAUNCH(og_push_epilogue_kernel<OG_MODE_QPUSH_EPI>, dim3(a->n_parts), dim3(OG_BLOCK), 0, s, *a);
    }
    OG_LAUNCH(og_final_kernel, dim3(a->n_final_blocks), dim3(OG_BLOCK), 0, s, *a, it);
    return og_hip_check(og_launch_status(), "og_push");
}

extern "C" int og_launch_push_scatter(const og_pr_dev *a, int pmode, hipStream_t s, hipEvent_t ev0,
                                      hipEvent_t ev1)
{
    unsigned gs = (a->E + OG_BLOCK - 1) / OG_BLOCK;
    if (gs > 16384) gs = 16384;
    if (gs == 0) gs = 1;
    if (ev0) hipEventRecord(ev0, s);
    if (a->pp_v0 && (pmode == OG_MODE_F32 || pmode == OG_MODE_FX64))
    {
        /* partitioned: every destination range, onto the running output */
        const uint32_t nr = a->n_pp - a->pp_first;
        if (nr && pmode == OG_MODE_F32)
            OG_LAUNCH((og_push_part_kernel<OG_MODE_F32, true>), dim3(nr), dim3(OG_PP_THREADS), 0, s, *a);
        else if (nr)
            OG_LAUNCH((og_push_part_kernel<OG_MODE_FX64, true>), dim3(nr), dim3(OG_PP_THREADS), 0, s, *a);
    }
    else if (pmode == OG_MODE_F32)
    {
        if (a->hot_grid)
            OG_LAUNCH((og_push_scatter_kernel<OG_MODE_F32, true>), dim3(a->hot_grid),
                               dim3(OG_HOT_SUBS * OG_BLOCK), 0, s, *a);
        else
            OG_LAUNCH((og_push_scatter_kernel<OG_MODE_F32, false>), dim3(gs), dim3(OG_BLOCK), 0, s, *a);
    }
    else if (pmode == OG_MODE_FX64)
    {
        if (a->hot_grid)
            OG_LAUNCH((og_push_scatter_kernel<OG_MODE_FX64, true>), dim3(a->hot_grid),
                               dim3(OG_HOT_SUBS * OG_BLOCK), 0, s, *a);
        else
            OG_LAUNCH((og_push_scatter_kernel<OG_MODE_FX64, false>), dim3(gs), dim3(OG_BLOCK), 0, s, *a);
    }
    else
    {
        og_set_error("og_launch_push_scatter: bad mode %d", pmode);
        return -1;
    }
    if (ev1) hipEventRecord(ev1, s);
    return og_hip_check(og_launch_status(), "og_push_scatter_kernel");
}

extern "C" int og_launch_init(const og_pr_dev *a, int mode, const float *init_old, float init_value,
                              hipStream_t s)
{
    unsigned grid = a->n_parts;
    switch (mode)
    {
    case OG_MODE_F32: OG_LAUNCH(og_init_kernel<OG_MODE_F32>, dim3(grid), dim3(OG_BLOCK), 0, s, *a, init_old, init_value); break;
    case OG_MODE_FX64: OG_LAUNCH(og_init_kernel<OG_MODE_FX64>, dim3(grid), dim3(OG_BLOCK), 0, s, *a, init_old, init_value); break;
    case OG_MODE_FX32: OG_LAUNCH(og_init_kernel<OG_MODE_FX32>, dim3(grid), dim3(OG_BLOCK), 0, s, *a, init_old, init_value); break;
    case OG_MODE_FX16: OG_LAUNCH(og_init_kernel<OG_MODE_FX16>, dim3(grid), dim3(OG_BLOCK), 0, s, *a, init_old, init_value); break;
    case OG_MODE_FX8: OG_LAUNCH(og_init_kernel<OG_MODE_FX8>, dim3(grid), dim3(OG_BLOCK), 0, s, *a, init_old, init_value); break;
    default: OG_LAUNCH(og_init_kernel<OG_MODE_Q32>, dim3(grid), dim3(OG_BLOCK), 0, s, *a, init_old, init_value); break;
    }
    return og_hip_check(og_launch_status(), "og_init_kernel");
}

extern "C" int og_launch_final(const og_pr_dev *a, uint32_t it, hipStream_t s)
{
    OG_LAUNCH(og_final_kernel, dim3(a->n_final_blocks), dim3(OG_BLOCK), 0, s, *a, it);
    return og_hip_check(og_launch_status(), "og_final_kernel");
}

extern "C" int og_launch_normalize(const og_pr_dev *a, float *d_out_old, double *d_sum,
                                   hipStream_t s)
{
    OG_LAUNCH(og_normalize_kernel, dim3(a->n_parts), dim3(OG_BLOCK), 0, s, *a, d_out_old);
    OG_LAUNCH(og_sum_parts_kernel, dim3(1), dim3(OG_BLOCK), 0, s, a->part_err, a->n_parts, d_sum);
    return og_hip_check(og_launch_status(), "og_normalize_kernel");
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

/* One worklist vertex v of the data-driven push (pageRank.c:2648-2677) or
 * pull-push (:2810-2850): its new rank, then its residual moved to its
 * out-neighbours with atomic adds.
 *
 * The reference reads v's residual, pushes it and then stores 0 into it:
 * whatever another thread adds to it in between is lost.  With a few CPU
 * threads that window rarely matters; with every worklist vertex in flight
 * at once it would drop much of the mass.  So the concurrent push (`take`)
 * takes the residual with one atomic exchange instead: what arrives later
 * stays for the next iteration, as it does in the reference's
 * one-thread order.  Its self-loops push nothing, since in that order the
 * final store drops what a self-loop added.  The serial form (one-chunk
 * graphs) keeps the reference's read / push / store-0 sequence. */
__device__ __forceinline__ void og_dd_push_vertex(const og_pr_dev &a, uint32_t code, uint32_t v, float fn,
                                                  og_stat &st, bool take = false)
{
    float old = a.r[v];
    float nr;
    const float res = take ? atomicExch(&a.resid[v], 0.0f) : a.resid[v];
    if (code == 7)
        nr = res + old;
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
    float delta = OG_DAMP * ((take ? res : a.resid[v]) / (float)deg);
    for (uint32_t j = a.f_row_ptr[v]; j < a.f_row_ptr[v + 1]; j++)
    {
        uint32_t u = a.f_cols[j];
        if (take && u == v) continue;
        float prev = atomicAdd(&a.resid[u], delta);
        if ((fabs((double)(prev + delta)) >= a.epsilon) && ((double)prev <= a.epsilon))
        {
            st.active++;
            if (code == 8) atomicAdd(&a.resid[u], delta);
            a.wl_next[u] = 1;
        }
    }
    if (!take) a.resid[v] = 0.0f;
}

/* pageRank.c:2645-2685 (push) and :2796-2852 (pull-push).  The reference
 * runs its vertex loop with schedule(dynamic, 1024): a graph of at most
 * 1024 vertices is one chunk, taken by one thread in vertex order, so its
 * result is deterministic; such graphs run here the same way, one thread
 * over the caller's vertex order (old ids), every read seeing the updates
 * of the vertices before it.  Larger graphs: one thread per worklist
 * vertex; the reads of neighbours' residuals and ranks race with their
 * updates, as the reference's threads do, so results are nondeterministic
 * (in the reference as well). */
__global__ __launch_bounds__(OG_BLOCK) void og_dd_push_kernel(og_pr_dev a, uint32_t code)
{
    if (*a.done) return;
    og_stat st;
    og_stat_init(st);
    const float fn = (float)a.N;
    if (a.N <= OG_DD_SERIAL_N)
    {
        if (blockIdx.x == 0 && threadIdx.x == 0)
            for (uint32_t vo = 0; vo < a.N; vo++)
            {
                const uint32_t v = a.old2new[vo];
                if (a.wl_cur[v]) og_dd_push_vertex(a, code, v, fn, st);
            }
    }
    else
        for (uint32_t v = blockIdx.x * OG_BLOCK + threadIdx.x; v < a.N; v += gridDim.x * OG_BLOCK)
            if (a.wl_cur[v]) og_dd_push_vertex(a, code, v, fn, st, code == 7);
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
    OG_LAUNCH(og_dd_contrib_kernel, dim3(og_vgrid(a->N)), dim3(OG_BLOCK), 0, s, *a);
    return og_hip_check(og_launch_status(), "og_dd_contrib_kernel");
}

extern "C" int og_launch_dd_activate(const og_pr_dev *a, hipStream_t s)
{
    if (!a->E) return 0;
    unsigned g = (a->E + OG_BLOCK - 1) / OG_BLOCK;
    if (g > 16384) g = 16384;
    OG_LAUNCH(og_dd_activate_kernel, dim3(g), dim3(OG_BLOCK), 0, s, *a);
    return og_hip_check(og_launch_status(), "og_dd_activate_kernel");
}

extern "C" int og_launch_dd_resid_init(const og_pr_dev *a, hipStream_t s)
{
    OG_LAUNCH(og_dd_resid_init_kernel, dim3(og_vgrid(a->N)), dim3(OG_BLOCK), 0, s, *a);
    return og_hip_check(og_launch_status(), "og_dd_resid_init_kernel");
}

extern "C" int og_launch_dd_push(const og_pr_dev *a, uint32_t code, uint32_t it, hipStream_t s)
{
    OG_LAUNCH(og_dd_push_kernel, dim3(a->n_parts), dim3(OG_BLOCK), 0, s, *a, code);
    OG_LAUNCH(og_final_kernel, dim3(a->n_final_blocks), dim3(OG_BLOCK), 0, s, *a, it);
    return og_hip_check(og_launch_status(), "og_dd_push_kernel");
}

/* ------------------------------------------------------------------ */
/* many graphs in one launch (fp32 pull)                               */
/* ------------------------------------------------------------------ */

/* Each graph keeps its own plan and buffers (an og_pr_dev per graph and
 * iteration parity); the kernels below walk one work list that holds the
 * blocks of every graph, so an iteration of all graphs is three launches
 * whatever their number and sizes, and a run is capturable in one hipGraph.
 * A graph that has stopped skips its items; once every graph has stopped
 * the kernels return at once. */

__global__ __launch_bounds__(OG_BLOCK) void og_multi_init_kernel(og_multi_dev m)
{
    for (uint32_t i = blockIdx.x; i < m.n_iwork; i += gridDim.x)
    {
        const og_multi_work w = m.iwork[i];
        const og_pr_dev &a = m.devs[2 * w.g];
        const uint32_t v = w.first * OG_BLOCK + threadIdx.x;
        if (v < a.N)
        {
            a.r[v] = OG_BASE_PR;
            og_store_contrib<OG_MODE_F32>((void *)a.g_in, v, OG_BASE_PR, a.outdeg[v]);
        }
        if (w.first == 0 && threadIdx.x == 0) *a.done = 0;
    }
    if (blockIdx.x == 0 && threadIdx.x == 0)
    {
        *m.live = m.n_graphs;
        *m.all_done = m.n_graphs == 0;
    }
}

/* The items xr of the work list: the graphs whose tile shape is BLK
 * threads per block, K chunks per wavefront and RB edges per load batch of
 * a thread's row (0: as many as the wavefront's longest row needs, up to
 * OG_ROW_BATCH).  The shape changes neither a row's sum nor a chunk's, so
 * every shape gives the same ranks.  settled != 0 from the third iteration
 * on: row blocks entirely in [Z, N) (no in-edges; rank and both
 * contribution buffers already final) only report their activeVertices
 * count (0 >= epsilon).  Block b runs on XCD b % 8 and takes that XCD's
 * items (contiguous ranges of whole graphs, so a graph's contributions are
 * gathered through one XCD's L2). */
template <uint32_t BLK, uint32_t K, uint32_t RB>
__global__ __launch_bounds__(BLK) void og_multi_pull_kernel(og_multi_dev m, og_multi_xr xr,
                                                           uint32_t parity, uint32_t settled)
{
    if (*m.all_done) return;
    const uint32_t x = blockIdx.x % OG_XCDS, nb = gridDim.x / OG_XCDS;
    for (uint32_t i = xr.o[x] + blockIdx.x / OG_XCDS; i < xr.o[x + 1]; i += nb)
    {
        const og_multi_work w = m.work[i];
        const og_pr_dev &a = m.devs[2 * w.g + parity];
        if (*a.done) continue;
        const uint32_t v0 = a.C + w.first * BLK;
        if (w.kind == 1 && settled && v0 >= a.Z)
        {
            if (threadIdx.x == 0)
            {
                m.w_err[i] = 0.0;
                m.w_act[i] = 0.0 >= a.epsilon ? min(BLK, a.N - v0) : 0u;
            }
            continue;
        }
        og_stat st;
        og_stat_init(st);
        if (w.kind == 0) og_pull_chunks<OG_MODE_F32, false, BLK / 64, K>(a, st, w.first);
        else og_pull_rows<OG_MODE_F32, false, RB == 0, BLK, (RB ? RB : OG_ROW_BATCH)>(a, st, w.first);
        og_stat r = og_block_stat<BLK / 64>(st);
        if (threadIdx.x == 0)
        {
            m.w_err[i] = r.err;
            m.w_act[i] = r.active;
        }
    }
}

__global__ __launch_bounds__(OG_BLOCK) void og_multi_long_kernel(og_multi_dev m, uint32_t parity)
{
    if (*m.all_done) return;
    const uint32_t wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    for (uint32_t i = blockIdx.x; i < m.n_lwork; i += gridDim.x)
    {
        const og_multi_work w = m.lwork[i];
        const og_pr_dev &a = m.devs[2 * w.g + parity];
        if (*a.done) continue;
        og_stat st;
        og_stat_init(st);
        const uint32_t h = w.first * OG_WAVES + wave;
        if (h < a.H) og_finalize_long_row<OG_MODE_F32>(a, h, st);
        og_stat r = og_block_stat(st);
        if (threadIdx.x == 0)
        {
            m.l_err[i] = r.err;
            m.l_act[i] = r.active;
        }
    }
}

/* Block g: graph g's statistics of iteration `it` from its items' partials
 * in item order, its stop rule (pageRank.c:1030-1033), and the live count. */
__global__ __launch_bounds__(OG_BLOCK) void og_multi_reduce_kernel(og_multi_dev m, uint32_t parity,
                                                                  uint32_t it)
{
    if (*m.all_done) return;
    const uint32_t g = blockIdx.x;
    const og_pr_dev &a = m.devs[2 * g + parity];
    if (*a.done) return;
    og_stat st;
    og_stat_init(st);
    for (uint32_t i = m.wr[2 * g] + threadIdx.x; i < m.wr[2 * g + 1]; i += OG_BLOCK)
    {
        st.err += m.w_err[i];
        st.active += m.w_act[i];
    }
    for (uint32_t i = m.lr[g] + threadIdx.x; i < m.lr[g + 1]; i += OG_BLOCK)
    {
        st.err += m.l_err[i];
        st.active += m.l_act[i];
    }
    og_stat r = og_block_stat(st);
    if (threadIdx.x == 0)
    {
        a.stats[it].error_total = r.err;
        a.stats[it].active = r.active;
        const bool stop = a.stop_l1 ? (r.err < a.epsilon) : (r.active == 0);
        if (stop)
        {
            *a.done = 1;
            if (atomicSub(m.live, 1u) == 1u) *m.all_done = 1;
        }
    }
}

extern "C" int og_launch_multi_init(const og_multi_dev *m, uint32_t grid, hipStream_t s)
{
    OG_LAUNCH(og_multi_init_kernel, dim3(grid), dim3(OG_BLOCK), 0, s, *m);
    return og_hip_check(og_launch_status(), "og_multi_init_kernel");
}

/* The compiled tile shapes: blocks of 128, 256 or 512 threads, 1, 2 or 4
 * chunks per wavefront, row batches adapted to the wavefront or fixed at
 * OG_ROW_BATCH. */
#define OG_MULTI_TILES(X)                                                                        \
    X(0, 128, 1, 0) X(1, 128, 2, 0) X(2, 128, 4, 0) X(3, 256, 1, 0) X(4, 256, 2, 0) X(5, 256, 4, 0) \
    X(6, 512, 1, 0) X(7, 512, 2, 0) X(8, 512, 4, 0) X(9, 128, 1, 8) X(10, 128, 2, 8)              \
    X(11, 128, 4, 8) X(12, 256, 1, 8) X(13, 256, 2, 8) X(14, 256, 4, 8) X(15, 512, 1, 8)          \
    X(16, 512, 2, 8) X(17, 512, 4, 8)

extern "C" uint32_t og_multi_tile_count(void)
{
#define OG_X(id, blk, k, rb) +1
    return 0 OG_MULTI_TILES(OG_X);
#undef OG_X
}

extern "C" int og_multi_tile_shape(uint32_t id, og_multi_tile *t)
{
    switch (id)
    {
#define OG_X(i, blk, k, rb) \
    case i: t->block = blk; t->chunks = k; t->row_batch = rb; return 0;
        OG_MULTI_TILES(OG_X)
#undef OG_X
    default: return -1;
    }
}

extern "C" int og_launch_multi_pull(const og_multi_dev *m, uint32_t tile, const og_multi_xr *xr,
                                    uint32_t it, uint32_t grid, hipStream_t s)
{
    const uint32_t parity = it & 1, settled = it >= 2;
    if (xr->o[OG_XCDS] <= xr->o[0]) return 0;
    if (grid < OG_XCDS || grid % OG_XCDS)
    {
        og_set_error("og_launch_multi_pull: grid %u is not a positive multiple of %u", grid, OG_XCDS);
        return -1;
    }
    switch (tile)
    {
#define OG_X(i, blk, k, rb)                                                                      \
    case i:                                                                                      \
        OG_LAUNCH((og_multi_pull_kernel<blk, k, rb>), dim3(grid), dim3(blk), 0, s, *m, *xr, parity, \
                  settled);                                                                      \
        break;
        OG_MULTI_TILES(OG_X)
#undef OG_X
    default: og_set_error("og_launch_multi_pull: no tile shape %u", tile); return -1;
    }
    return og_hip_check(og_launch_status(), "og_multi_pull_kernel");
}

extern "C" int og_launch_multi_iter(const og_multi_dev *m, const og_multi_group *groups, uint32_t n_groups,
                                    uint32_t it, uint32_t grid, hipStream_t s)
{
    const uint32_t parity = it & 1;
    for (uint32_t q = 0; q < n_groups; q++)
        if (og_launch_multi_pull(m, groups[q].tile, &groups[q].xr, it, groups[q].grid, s)) return -1;
    if (m->n_lwork)
        OG_LAUNCH(og_multi_long_kernel, dim3(grid), dim3(OG_BLOCK), 0, s, *m, parity);
    OG_LAUNCH(og_multi_reduce_kernel, dim3(m->n_graphs), dim3(OG_BLOCK), 0, s, *m, parity, it);
    return og_hip_check(og_launch_status(), "og_multi_pull_kernel");
}
