/*
 * og_stats.hip -- gfx950 kernels of the reference's PageRank rank comparison
 * (-S, graphStats.c:37-688 and :746-751): how far a variant's ranking is
 * from the pull ranking over the top k positions.
 *
 * A batch holds P problems (k_p, m_p): the positions [m_p - k_p, m_p) of
 * both rankings (order = realRanks, ascending by rank).  Kernels:
 *   - inverse rankings (a scatter);
 *   - per problem the top-k slices (vertex ids, the ranks of the compared
 *     ranking's vertices in both runs, the ids as floats) packed at
 *     off_p = k_0 + ... + k_{p-1};
 *   - the edit distance of the two id lists: one block per problem sweeps
 *     the (k+1) x (k+1) matrix column by column, each thread holding a
 *     strip of rows in registers, skewed by one step per thread (above
 *     OG_LEV_LDS_K: anti-diagonals in global memory); the reference's min3
 *     tie rule is kept (a tie of deletion and insertion below substitution
 *     yields substitution);
 *   - Kendall pair counts (s, ties in arr1, ties in arr2) for the float and
 *     the id arrays: one block per small problem with its arrays in LDS,
 *     or blocks over groups of rows i, threads over j > i;
 *   - intersection and mismatch counts and the two average errors, whose
 *     double sums run in position order (as the reference's loops).
 */
#include <hip/hip_runtime.h>

#include "og_internal.h"
#include "og_stats.h"

#define SB 256u
#define OG_LEV_BLOCK 512u
#define OG_KT_ROWS 8u
#define OG_SUM_TILE 2048u

static unsigned og_grid(uint64_t n, unsigned bs)
{
    uint64_t g = (n + bs - 1) / bs;
    if (g > 65536) g = 65536;
    return g ? (unsigned)g : 1u;
}

/* graphStats.c:746-751 */
__global__ __launch_bounds__(SB) void og_stats_inverse_kernel(const uint32_t *__restrict__ order,
                                                              const uint32_t *__restrict__ ref_order, uint32_t N,
                                                              uint32_t *__restrict__ inv, uint32_t *__restrict__ ref_inv)
{
    /* realRanks are permutations of [0, N); an id outside is skipped */
    for (uint32_t v = blockIdx.x * SB + threadIdx.x; v < N; v += gridDim.x * SB)
    {
        const uint32_t w = order[v], rw = ref_order[v];
        if (w < N) inv[w] = v;
        if (rw < N) ref_inv[rw] = v;
    }
}

__device__ __forceinline__ uint32_t og_problem_of(const uint32_t *__restrict__ off, uint32_t P, uint32_t t)
{
    uint32_t lo = 0, hi = P;   /* off[lo] <= t < off[hi] */
    while (hi - lo > 1)
    {
        const uint32_t mid = (lo + hi) >> 1;
        if (off[mid] <= t) lo = mid; else hi = mid;
    }
    return lo;
}

/* graphStats.c:604-619: the top-k slices of problem p. */
__global__ __launch_bounds__(SB) void og_stats_slices_kernel(og_stats_dev a)
{
    const uint32_t T = a.off[a.P];
    for (uint32_t t = blockIdx.x * SB + threadIdx.x; t < T; t += gridDim.x * SB)
    {
        const uint32_t p = og_problem_of(a.off, a.P, t);
        const uint32_t u = t - a.off[p];
        const uint32_t v = a.m[p] - a.k[p] + u;
        const uint32_t w = a.order[v], rw = a.ref_order[v];
        a.ids[t] = w;
        a.ref_ids[t] = rw;
        a.f[t] = w < a.N ? a.pr[w] : 0.0f;
        a.ref_f[t] = w < a.N ? a.ref_pr[w] : 0.0f;
        a.re[t] = (float)w / 1.0f;
        a.ref_re[t] = (float)rw / 1.0f;
    }
}

/* graphStats.c:37-50 */
__device__ __forceinline__ uint32_t og_min3_ref(uint32_t a, uint32_t b, uint32_t c)
{
    const int x = (int)a, y = (int)b, z = (int)c;
    if (x < y && x < z) return a;
    if (y < x && y < z) return b;
    return c;
}

/* One anti-diagonal d of the edit-distance matrix: cur[i] = mat[i][d - i]
 * from the two diagonals before it (graphStats.c:52-100). */
template <typename S>
__device__ __forceinline__ void og_lev_diag(uint32_t d, uint32_t K, const uint32_t *a1, const uint32_t *a2,
                                            S *cur, const S *p1, const S *p2)
{
    const uint32_t ilo = d > K ? d - K : 0u, ihi = d < K ? d : K;
    for (uint32_t i = ilo + threadIdx.x; i <= ihi; i += blockDim.x)
    {
        const uint32_t j = d - i;
        uint32_t v;
        if (i == 0) v = j;
        else if (j == 0) v = i;
        else
        {
            const uint32_t cost = a1[i - 1] == a2[j - 1] ? 0u : 1u;
            v = og_min3_ref((uint32_t)p1[i - 1] + 1u, (uint32_t)p1[i] + 1u, (uint32_t)p2[i - 1] + cost);
        }
        cur[i] = (S)v;
    }
}

/* Problems with k <= R * 64 * W: one block of W wavefronts; thread t owns
 * the R rows t*R+1 .. t*R+R of the matrix and keeps their current column
 * in registers.  Thread t computes column j at step j + t - 1 (a skewed
 * sweep): the cell above its strip comes from thread t - 1's last row of
 * the step before, by a lane shift inside a wavefront and through LDS (one
 * barrier per step) across wavefronts; the compared list sits in LDS, the
 * strip's own list entries in registers. */
template <uint32_t R, uint32_t W>
__global__ __launch_bounds__(64 * W) void og_lev_strip_kernel(og_stats_dev a, const uint32_t *__restrict__ plist)
{
    constexpr uint32_t T = 64 * W;
    __shared__ uint32_t cols[R * T];
    __shared__ uint32_t bot[2][W];
    const uint32_t p = plist[blockIdx.x];
    const uint32_t K = a.k[p], o = a.off[p];
    const uint32_t t = threadIdx.x, lane = t & 63, w = t >> 6;
    for (uint32_t i = t; i < K; i += T) cols[i] = a.ids[o + i];
    uint32_t rid[R], col[R];
#pragma unroll
    for (uint32_t r = 0; r < R; r++)
    {
        const uint32_t i = t * R + r + 1;              /* 1-based row */
        rid[r] = i <= K ? a.ref_ids[o + i - 1] : 0xFFFFFFFFu;
        col[r] = i;                                    /* column 0: mat[i][0] = i */
    }
    uint32_t top_prev = t * R;                         /* mat[t*R][0] */
    uint32_t bottom = col[R - 1];
    __syncthreads();
    const uint32_t steps = K ? K + T - 1 : 0;
    for (uint32_t s = 0; s < steps; s++)
    {
        uint32_t left = (uint32_t)__shfl_up((int)bottom, 1, 64);
        if constexpr (W > 1)
        {
            if (lane == 63) bot[s & 1][w] = bottom;
            __syncthreads();
            if (lane == 0 && w > 0) left = bot[s & 1][w - 1];
        }
        const int j = (int)s - (int)t + 1;
        if (j >= 1 && j <= (int)K)
        {
            const uint32_t top = t == 0 ? (uint32_t)j : left;   /* mat[t*R][j] */
            const uint32_t cj = cols[j - 1];
            uint32_t up = top, diag = top_prev;
#pragma unroll
            for (uint32_t r = 0; r < R; r++)
            {
                const uint32_t v = og_min3_ref(up + 1u, col[r] + 1u, diag + (rid[r] == cj ? 0u : 1u));
                diag = col[r];
                col[r] = v;
                up = v;
            }
            top_prev = top;
            bottom = col[R - 1];
        }
    }
    /* after the sweep every strip holds column K; row K's owner writes it */
    if (K && t == (K - 1) / R)
    {
        const uint32_t rk = (K - 1) % R;
        uint32_t d = 0;
#pragma unroll
        for (uint32_t r = 0; r < R; r++)
            if (r == rk) d = col[r];
        a.lev[p] = d;
    }
    else if (!K && t == 0) a.lev[p] = 0;
}

/* Larger problems: u32 diagonals in global scratch (3 * (k + 1) entries). */
__global__ __launch_bounds__(OG_LEV_BLOCK) void og_lev_global_kernel(og_stats_dev a, uint32_t p, uint32_t *__restrict__ dg)
{
    const uint32_t K = a.k[p], o = a.off[p];
    const uint32_t *a1 = a.ref_ids + o, *a2 = a.ids + o;
    uint32_t *D[3] = {dg, dg + (K + 1), dg + 2 * (K + 1)};
    for (uint32_t d = 0; d <= 2 * K; d++)
    {
        og_lev_diag<uint32_t>(d, K, a1, a2, D[d % 3], D[(d + 2) % 3], D[(d + 1) % 3]);
        __syncthreads();
    }
    if (threadIdx.x == 0) a.lev[p] = D[(2 * K) % 3][K];
}

/* graphStats.c:427-489 for one pair (i, j): s += +-1, ties counted. */
__device__ __forceinline__ void og_kt_pair(float x1i, float x1j, float x2i, float x2j, int &s, int &m1, int &m2)
{
    if (x2i > x2j)
    {
        if (x1i > x1j) s++;
        else if (x1i < x1j) s--;
        else m1++;
    }
    else if (x2i < x2j)
    {
        if (x1i > x1j) s--;
        else if (x1i < x1j) s++;
        else m1++;
    }
    else
    {
        m2++;
        if (x1i == x1j) m1++;
    }
}

template <typename T> __device__ __forceinline__ T og_bsum(T v, T *sh)
{
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    const uint32_t w = threadIdx.x >> 6, nw = blockDim.x >> 6;
    __syncthreads();
    if ((threadIdx.x & 63) == 0) sh[w] = v;
    __syncthreads();
    T r = 0;
    for (uint32_t i = 0; i < nw; i++) r += sh[i];
    return r;
}

__device__ __forceinline__ void og_kt_flush(long long *dst, int v, long long *sh)
{
    const long long t = og_bsum<long long>((long long)v, sh);
    if (threadIdx.x == 0 && t) atomicAdd((unsigned long long *)dst, (unsigned long long)t);
}

/* Row group g = (problem, first row): rows [i0, i0 + OG_KT_ROWS) against
 * every later position; counts of both arrays added to the problem's six
 * counters. */
__global__ __launch_bounds__(SB) void og_kendall_kernel(og_stats_dev a, const uint2 *__restrict__ groups)
{
    __shared__ long long sh[SB / 64];
    const uint2 gdesc = groups[blockIdx.x];
    const uint32_t p = gdesc.x, i0 = gdesc.y;
    const uint32_t K = a.k[p], o = a.off[p];
    const float *f1 = a.ref_f + o, *f2 = a.f + o, *r1 = a.ref_re + o, *r2 = a.re + o;
    int sf = 0, m1f = 0, m2f = 0, sr = 0, m1r = 0, m2r = 0;
    const uint32_t iend = min(i0 + OG_KT_ROWS, K);
    for (uint32_t i = i0; i < iend; i++)
    {
        const float x1 = f1[i], x2 = f2[i], y1 = r1[i], y2 = r2[i];
        for (uint32_t j = i + 1 + threadIdx.x; j < K; j += SB)
        {
            og_kt_pair(x1, f1[j], x2, f2[j], sf, m1f, m2f);
            og_kt_pair(y1, r1[j], y2, r2[j], sr, m1r, m2r);
        }
    }
    og_kt_flush(a.kt + 6 * p + 0, sf, sh);
    og_kt_flush(a.kt + 6 * p + 1, m1f, sh);
    og_kt_flush(a.kt + 6 * p + 2, m2f, sh);
    og_kt_flush(a.kt + 6 * p + 3, sr, sh);
    og_kt_flush(a.kt + 6 * p + 4, m1r, sh);
    og_kt_flush(a.kt + 6 * p + 5, m2r, sh);
}

/* Problems with k <= OG_KT_SMALL_K: one block each, the four arrays staged
 * in LDS; thread t takes the rows t, t + 256, ... against every later
 * position. */
__global__ __launch_bounds__(SB) void og_kendall_small_kernel(og_stats_dev a, const uint32_t *__restrict__ plist)
{
    __shared__ float s1[OG_KT_SMALL_K], s2[OG_KT_SMALL_K], q1[OG_KT_SMALL_K], q2[OG_KT_SMALL_K];
    __shared__ long long sh[SB / 64];
    const uint32_t p = plist[blockIdx.x];
    const uint32_t K = a.k[p], o = a.off[p];
    for (uint32_t i = threadIdx.x; i < K; i += SB)
    {
        s1[i] = a.ref_f[o + i];
        s2[i] = a.f[o + i];
        q1[i] = a.ref_re[o + i];
        q2[i] = a.re[o + i];
    }
    __syncthreads();
    int sf = 0, m1f = 0, m2f = 0, sr = 0, m1r = 0, m2r = 0;
    for (uint32_t i = threadIdx.x; i < K; i += SB)
    {
        const float x1 = s1[i], x2 = s2[i], y1 = q1[i], y2 = q2[i];
        for (uint32_t j = i + 1; j < K; j++)
        {
            og_kt_pair(x1, s1[j], x2, s2[j], sf, m1f, m2f);
            og_kt_pair(y1, q1[j], y2, q2[j], sr, m1r, m2r);
        }
    }
    og_kt_flush(a.kt + 6 * p + 0, sf, sh);
    og_kt_flush(a.kt + 6 * p + 1, m1f, sh);
    og_kt_flush(a.kt + 6 * p + 2, m2f, sh);
    og_kt_flush(a.kt + 6 * p + 3, sr, sh);
    og_kt_flush(a.kt + 6 * p + 4, m1r, sh);
    og_kt_flush(a.kt + 6 * p + 5, m2r, sh);
}

/* graphStats.c:500-578: per problem, one block.  Counts reduce in any
 * order; the two error sums are formed in position order: the block writes
 * a tile of terms to LDS, thread 0 adds them in sequence. */
__global__ __launch_bounds__(SB) void og_topk_sums_kernel(og_stats_dev a)
{
    __shared__ double tf[OG_SUM_TILE], tr[OG_SUM_TILE];
    __shared__ unsigned long long sh[SB / 64];
    const uint32_t p = blockIdx.x;
    const uint32_t K = a.k[p], M = a.m[p];
    unsigned long long inter = 0, mism = 0;
    double ef = 0.0, er = 0.0;
    for (uint32_t base = M - K; base < M; base += OG_SUM_TILE)
    {
        const uint32_t n = min(OG_SUM_TILE, M - base);
        for (uint32_t t = threadIdx.x; t < n; t += SB)
        {
            const uint32_t v = base + t;
            /* intersection_topK (graphStats.c:560-578) */
            const uint32_t w = min(a.ref_order[v], a.N - 1);
            inter += a.inv[w] >= M - K;
            const uint32_t iw = a.inv[w], rw = a.ref_inv[w];
            mism += iw != rw;
            const float rp = a.ref_pr[w];
            tf[t] = (double)rp > 0.0f ? fabs((double)a.pr[w] - (double)rp) / (double)rp : 0.0;
            tr[t] = abs((int)(iw - rw)) / ((double)rw + 1);
        }
        __syncthreads();
        if (threadIdx.x == 0)
            for (uint32_t t = 0; t < n; t++)
            {
                ef += tf[t];
                er += tr[t];
            }
        __syncthreads();
    }
    inter = og_bsum<unsigned long long>(inter, sh);
    mism = og_bsum<unsigned long long>(mism, sh);
    if (threadIdx.x == 0)
    {
        a.inter[p] = (uint32_t)inter;
        a.mism[p] = (uint32_t)mism;
        a.err_f[p] = ef / K;
        a.err_r[p] = er / K;
    }
}

extern "C" int og_launch_stats_inverse(const og_stats_dev *a, hipStream_t s)
{
    OG_LAUNCH(og_stats_inverse_kernel, dim3(og_grid(a->N, SB)), dim3(SB), 0, s, a->order, a->ref_order,
                       a->N, a->inv, a->ref_inv);
    return og_hip_check(og_launch_status(), "og_stats_inverse_kernel");
}

extern "C" int og_launch_stats_slices(const og_stats_dev *a, uint32_t total, hipStream_t s)
{
    OG_LAUNCH(og_stats_slices_kernel, dim3(og_grid(total, SB)), dim3(SB), 0, s, *a);
    return og_hip_check(og_launch_status(), "og_stats_slices_kernel");
}

extern "C" int og_launch_lev_small(const og_stats_dev *a, const uint32_t *plist, uint32_t n, hipStream_t s)
{
    static_assert(16 * 64 * 1 == OG_LEV_SMALL_K, "small edit-distance strips cover OG_LEV_SMALL_K");
    if (n) OG_LAUNCH((og_lev_strip_kernel<16, 1>), dim3(n), dim3(64), 0, s, *a, plist);
    return og_hip_check(og_launch_status(), "og_lev_strip_kernel<16, 1>");
}

extern "C" int og_launch_lev_large(const og_stats_dev *a, const uint32_t *plist, uint32_t n, hipStream_t s)
{
    static_assert(24 * 64 * 8 == OG_LEV_LDS_K, "large edit-distance strips cover OG_LEV_LDS_K");
    if (n) OG_LAUNCH((og_lev_strip_kernel<24, 8>), dim3(n), dim3(512), 0, s, *a, plist);
    return og_hip_check(og_launch_status(), "og_lev_strip_kernel<24, 8>");
}

extern "C" int og_launch_lev_global(const og_stats_dev *a, uint32_t p, uint32_t *scratch, hipStream_t s)
{
    OG_LAUNCH(og_lev_global_kernel, dim3(1), dim3(OG_LEV_BLOCK), 0, s, *a, p, scratch);
    return og_hip_check(og_launch_status(), "og_lev_global_kernel");
}

extern "C" int og_launch_kendall_small(const og_stats_dev *a, const uint32_t *plist, uint32_t n, hipStream_t s)
{
    if (n) OG_LAUNCH(og_kendall_small_kernel, dim3(n), dim3(SB), 0, s, *a, plist);
    return og_hip_check(og_launch_status(), "og_kendall_small_kernel");
}

extern "C" int og_launch_kendall(const og_stats_dev *a, const void *groups, uint32_t n_groups, hipStream_t s)
{
    if (n_groups)
        OG_LAUNCH(og_kendall_kernel, dim3(n_groups), dim3(SB), 0, s, *a, (const uint2 *)groups);
    return og_hip_check(og_launch_status(), "og_kendall_kernel");
}

extern "C" int og_launch_topk_sums(const og_stats_dev *a, hipStream_t s)
{
    if (a->P) OG_LAUNCH(og_topk_sums_kernel, dim3(a->P), dim3(SB), 0, s, *a);
    return og_hip_check(og_launch_status(), "og_topk_sums_kernel");
}
