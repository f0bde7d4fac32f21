/*
 * og_csr.hip -- gfx950 kernels of the edge-list -> CSR builder and of the
 * synthetic graph generators.
 *
 *  - og_rmat_kernel / og_er_kernel: one thread per edge of the counter-based
 *    generators in og_common.h (the reference has no generator; its inputs
 *    are .bin files, edgeList.c:336-594).
 *  - og_max_id_kernel: N = max id + 1 (edgeList.c:559-567).
 *  - Stable LSD radix sort of (key, value) u32 pairs, 8-bit digits
 *    (radixsort.c:46-135): a per-block digit histogram over a contiguous
 *    range, one exclusive scan in (digit, block) order, and a stable scatter
 *    in which every 4096-item tile is ranked with wavefront ballots, placed
 *    digit-contiguous in LDS and written out in runs.
 *  - og_map_* : out_degree and edges_idx of the sorted list
 *    (vertex.c:262-289), with edges_idx left 0 for vertices without edges.
 */
#include <hip/hip_runtime.h>

#include "og_internal.h"
#include "og_plan.h"

#define OG_BLOCK 256
#define RS_TILE 4096u
#define RS_ROUNDS (RS_TILE / OG_BLOCK)
#define RS_MAX_BLOCKS 2048u

/* ------------------------------------------------------------------ */
/* generators                                                          */
/* ------------------------------------------------------------------ */

__global__ void og_rmat_kernel(uint32_t *__restrict__ src, uint32_t *__restrict__ dst,
                               uint64_t n, og_rmat_t g)
{
    for (uint64_t e = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; e < n;
         e += (uint64_t)gridDim.x * blockDim.x)
    {
        uint32_t s, d;
        og_rmat_edge(&g, e, &s, &d);
        src[e] = s;
        dst[e] = d;
    }
}

__global__ void og_er_kernel(uint32_t *__restrict__ src, uint32_t *__restrict__ dst, uint64_t n,
                             uint32_t nv, uint64_t key)
{
    for (uint64_t e = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; e < n;
         e += (uint64_t)gridDim.x * blockDim.x)
    {
        uint32_t s, d;
        og_er_edge(key, nv, e, &s, &d);
        src[e] = s;
        dst[e] = d;
    }
}

static unsigned og_grid_for(uint64_t n, unsigned per_block, unsigned cap)
{
    uint64_t b = (n + per_block - 1) / per_block;
    if (b < 1) b = 1;
    if (b > cap) b = cap;
    return (unsigned)b;
}

extern "C" int og_launch_rmat(uint32_t *d_src, uint32_t *d_dst, uint64_t n, const og_rmat_t *g,
                              hipStream_t s)
{
    if (n == 0) return 0;
    unsigned grid = og_grid_for(n, OG_BLOCK, 65536);
    OG_LAUNCH(og_rmat_kernel, dim3(grid), dim3(OG_BLOCK), 0, s, d_src, d_dst, n, *g);
    return og_hip_check(og_launch_status(), "og_rmat_kernel");
}

extern "C" int og_launch_er(uint32_t *d_src, uint32_t *d_dst, uint64_t n, uint32_t nv,
                            uint64_t key, hipStream_t s)
{
    if (n == 0) return 0;
    unsigned grid = og_grid_for(n, OG_BLOCK, 65536);
    OG_LAUNCH(og_er_kernel, dim3(grid), dim3(OG_BLOCK), 0, s, d_src, d_dst, n, nv, key);
    return og_hip_check(og_launch_status(), "og_er_kernel");
}

/* ------------------------------------------------------------------ */
/* small utility kernels                                               */
/* ------------------------------------------------------------------ */

__device__ __forceinline__ uint32_t og_wave_max_u32(uint32_t v)
{
    for (int o = 32; o > 0; o >>= 1) v = max(v, (uint32_t)__shfl_xor(v, o, 64));
    return v;
}

__global__ void og_max_id_kernel(const uint32_t *__restrict__ a, const uint32_t *__restrict__ b,
                                 uint32_t n, uint32_t *out)
{
    __shared__ uint32_t s[OG_BLOCK / 64];
    uint32_t m = 0;
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
        m = max(m, max(a[i], b[i]));
    m = og_wave_max_u32(m);
    if ((threadIdx.x & 63) == 0) s[threadIdx.x >> 6] = m;
    __syncthreads();
    if (threadIdx.x == 0)
    {
        for (int w = 1; w < OG_BLOCK / 64; w++) m = max(m, s[w]);
        atomicMax(out, m);
    }
}

__global__ void og_fill_u32_kernel(uint32_t *p, uint32_t v, uint64_t n)
{
    for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n;
         i += (uint64_t)gridDim.x * blockDim.x)
        p[i] = v;
}

extern "C" int og_launch_fill_u32(uint32_t *p, uint32_t v, uint64_t n, hipStream_t s)
{
    if (n == 0) return 0;
    OG_LAUNCH(og_fill_u32_kernel, dim3(og_grid_for(n, OG_BLOCK, 16384)), dim3(OG_BLOCK), 0,
                       s, p, v, n);
    return og_hip_check(og_launch_status(), "og_fill_u32_kernel");
}

extern "C" int og_launch_max_id(const uint32_t *d_src, const uint32_t *d_dst, uint32_t n,
                                uint32_t *d_out, hipStream_t s)
{
    if (og_launch_fill_u32(d_out, 0, 1, s)) return -1;
    if (n == 0) return 0;
    OG_LAUNCH(og_max_id_kernel, dim3(og_grid_for(n, OG_BLOCK * 8, 4096)), dim3(OG_BLOCK),
                       0, s, d_src, d_dst, n, d_out);
    return og_hip_check(og_launch_status(), "og_max_id_kernel");
}

__global__ void og_deinterleave_kernel(const uint32_t *__restrict__ pairs, uint32_t n,
                                       uint32_t stride, uint32_t *__restrict__ src,
                                       uint32_t *__restrict__ dst)
{
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
    {
        src[i] = pairs[(uint64_t)i * stride];
        dst[i] = pairs[(uint64_t)i * stride + 1];
    }
}

extern "C" int og_launch_deinterleave(const uint32_t *d_pairs, uint32_t n, uint32_t stride,
                                      uint32_t *d_src, uint32_t *d_dst, hipStream_t s)
{
    if (n == 0) return 0;
    OG_LAUNCH(og_deinterleave_kernel, dim3(og_grid_for(n, OG_BLOCK, 65536)),
                       dim3(OG_BLOCK), 0, s, d_pairs, n, stride, d_src, d_dst);
    return og_hip_check(og_launch_status(), "og_deinterleave_kernel");
}

/* The .csr file body (graphCSR.c:380-449, directed, unweighted): per
 * vertex forward out-degree, in-degree, edges_idx, then the inverse ones
 * (6 words); then per edge forward src, dst, inverse src, dst (4 words).
 * a[10] = f_src, f_dst, f_out, f_in, f_idx, i_src, i_dst, i_out, i_in, i_idx. */
struct og_csr_arrays
{
    uint32_t *a[10];
};

__global__ __launch_bounds__(OG_BLOCK) void og_csr_pack_kernel(og_csr_arrays g, uint32_t N, uint32_t E,
                                                              uint32_t *__restrict__ body)
{
    const uint64_t n = N > E ? N : E;
    for (uint64_t i = blockIdx.x * (uint64_t)OG_BLOCK + threadIdx.x; i < n; i += (uint64_t)gridDim.x * OG_BLOCK)
    {
        if (i < N)
        {
            uint32_t *o = body + 6 * i;
            o[0] = g.a[2][i]; o[1] = g.a[3][i]; o[2] = g.a[4][i];
            o[3] = g.a[7][i]; o[4] = g.a[8][i]; o[5] = g.a[9][i];
        }
        if (i < E)
        {
            uint32_t *o = body + 6 * (uint64_t)N + 4 * i;
            o[0] = g.a[0][i]; o[1] = g.a[1][i]; o[2] = g.a[5][i]; o[3] = g.a[6][i];
        }
    }
}

__global__ __launch_bounds__(OG_BLOCK) void og_csr_unpack_kernel(const uint32_t *__restrict__ body, uint32_t N,
                                                                uint32_t E, og_csr_arrays g)
{
    const uint64_t n = N > E ? N : E;
    for (uint64_t i = blockIdx.x * (uint64_t)OG_BLOCK + threadIdx.x; i < n; i += (uint64_t)gridDim.x * OG_BLOCK)
    {
        if (i < N)
        {
            const uint32_t *o = body + 6 * i;
            g.a[2][i] = o[0]; g.a[3][i] = o[1]; g.a[4][i] = o[2];
            g.a[7][i] = o[3]; g.a[8][i] = o[4]; g.a[9][i] = o[5];
        }
        if (i < E)
        {
            const uint32_t *o = body + 6 * (uint64_t)N + 4 * i;
            g.a[0][i] = o[0]; g.a[1][i] = o[1]; g.a[5][i] = o[2]; g.a[6][i] = o[3];
        }
    }
}

extern "C" int og_launch_csr_pack(uint32_t *const arrays[10], uint32_t N, uint32_t E, uint32_t *body,
                                  hipStream_t s)
{
    og_csr_arrays g;
    for (int k = 0; k < 10; k++) g.a[k] = arrays[k];
    const uint64_t n = N > E ? N : E;
    if (n) OG_LAUNCH(og_csr_pack_kernel, dim3(og_grid_for(n, OG_BLOCK, 65536)), dim3(OG_BLOCK), 0, s,
                              g, N, E, body);
    return og_hip_check(og_launch_status(), "og_csr_pack_kernel");
}

extern "C" int og_launch_csr_unpack(const uint32_t *body, uint32_t N, uint32_t E, uint32_t *const arrays[10],
                                    hipStream_t s)
{
    og_csr_arrays g;
    for (int k = 0; k < 10; k++) g.a[k] = arrays[k];
    const uint64_t n = N > E ? N : E;
    if (n) OG_LAUNCH(og_csr_unpack_kernel, dim3(og_grid_for(n, OG_BLOCK, 65536)), dim3(OG_BLOCK), 0, s,
                              body, N, E, g);
    return og_hip_check(og_launch_status(), "og_csr_unpack_kernel");
}

/* ------------------------------------------------------------------ */
/* stable LSD radix sort                                               */
/* ------------------------------------------------------------------ */

/* Block b owns items [b*chunk, min((b+1)*chunk, n)); chunk is a multiple of
 * RS_TILE.  counts[d * nb + b] = number of its items with digit d. */
__global__ __launch_bounds__(OG_BLOCK) void og_rs_hist_kernel(const uint32_t *__restrict__ keys,
                                                              uint32_t n, uint32_t chunk,
                                                              uint32_t shift,
                                                              uint32_t *__restrict__ counts)
{
    __shared__ uint32_t h[256];
    const uint32_t nb = gridDim.x;
    h[threadIdx.x] = 0;
    __syncthreads();
    uint64_t beg = (uint64_t)blockIdx.x * chunk;
    uint64_t end = beg + chunk < n ? beg + chunk : n;
    for (uint64_t i = beg + threadIdx.x; i < end; i += OG_BLOCK)
        atomicAdd(&h[(keys[i] >> shift) & 0xFF], 1u);
    __syncthreads();
    counts[threadIdx.x * nb + blockIdx.x] = h[threadIdx.x];
}

__device__ __forceinline__ uint64_t og_lanemask_lt()
{
    uint32_t lane = threadIdx.x & 63;
    return lane ? (~0ull >> (64 - lane)) : 0ull;
}

/* Stable scatter.  For each tile of RS_TILE consecutive items:
 *  1. rounds of 256 items (item r*256 + t on thread t): lanes with the same
 *     digit are found with 8 ballots; the rank of an item among the tile's
 *     items with its digit is (earlier rounds) + (earlier waves this round)
 *     + (lower lanes in its wave), which is exactly its index order;
 *  2. the tile is placed digit-contiguous in LDS at tile_start[d] + rank;
 *  3. LDS slot i goes to base[d] + (i - tile_start[d]), so consecutive
 *     threads write consecutive addresses of one digit run. */
__global__ __launch_bounds__(OG_BLOCK) void og_rs_scatter_kernel(
    const uint32_t *__restrict__ k_in, const uint32_t *__restrict__ v_in,
    uint32_t *__restrict__ k_out, uint32_t *__restrict__ v_out, uint32_t n, uint32_t chunk,
    uint32_t shift, const uint32_t *__restrict__ offsets)
{
    __shared__ uint32_t s_base[256];         /* global write position of each digit */
    __shared__ uint32_t s_run[256];          /* items of each digit in this tile so far */
    __shared__ uint32_t s_start[256];        /* tile-local start of each digit */
    __shared__ uint32_t s_wcnt[OG_BLOCK / 64][256];
    __shared__ uint32_t s_key[RS_TILE];
    __shared__ uint32_t s_val[RS_TILE];
    __shared__ uint8_t s_dig[RS_TILE];

    const uint32_t t = threadIdx.x, w = t >> 6;
    const uint32_t nb = gridDim.x;
    s_base[t] = offsets[t * nb + blockIdx.x];
    const uint64_t lt = og_lanemask_lt();

    uint64_t beg = (uint64_t)blockIdx.x * chunk;
    uint64_t end = beg + chunk < n ? beg + chunk : n;
    for (uint64_t tile = beg; tile < end; tile += RS_TILE)
    {
        uint32_t cnt_tile = (uint32_t)(end - tile < RS_TILE ? end - tile : RS_TILE);
        uint32_t my_key[RS_ROUNDS], my_val[RS_ROUNDS], my_rank[RS_ROUNDS];
        uint32_t my_dig[RS_ROUNDS];
        s_run[t] = 0;
        for (uint32_t ww = 0; ww < OG_BLOCK / 64; ww++) s_wcnt[ww][t] = 0;
        __syncthreads();
#pragma unroll
        for (uint32_t r = 0; r < RS_ROUNDS; r++)
        {
            uint32_t li = r * OG_BLOCK + t;
            bool valid = li < cnt_tile;
            uint32_t k = 0, v = 0, d = 0;
            if (valid)
            {
                k = k_in[tile + li];
                v = v_in[tile + li];
                d = (k >> shift) & 0xFF;
            }
            uint64_t m = __ballot(valid);
#pragma unroll
            for (int bit = 0; bit < 8; bit++)
            {
                uint64_t bb = __ballot((d >> bit) & 1);
                m &= ((d >> bit) & 1) ? bb : ~bb;
            }
            uint32_t rank_w = __popcll(m & lt);
            if (valid && rank_w == 0) s_wcnt[w][d] = __popcll(m);
            __syncthreads();
            uint32_t rank = 0;
            if (valid)
            {
                rank = s_run[d] + rank_w;
                for (uint32_t ww = 0; ww < w; ww++) rank += s_wcnt[ww][d];
            }
            __syncthreads();
            uint32_t add = 0;
            for (uint32_t ww = 0; ww < OG_BLOCK / 64; ww++)
            {
                add += s_wcnt[ww][t];
                s_wcnt[ww][t] = 0;
            }
            s_run[t] += add;
            __syncthreads();
            my_key[r] = k;
            my_val[r] = v;
            my_dig[r] = d;
            my_rank[r] = valid ? rank : 0xFFFFFFFFu;
        }
        /* exclusive scan of the 256 digit counts of this tile */
        {
            uint32_t c = s_run[t];
            uint32_t x = c;
            /* wave-level inclusive scan then cross-wave fixup */
            for (int o = 1; o < 64; o <<= 1)
            {
                uint32_t y = __shfl_up(x, o, 64);
                if ((t & 63) >= (uint32_t)o) x += y;
            }
            if ((t & 63) == 63) s_wcnt[0][w] = x;
            __syncthreads();
            uint32_t pre = 0;
            for (uint32_t ww = 0; ww < w; ww++) pre += s_wcnt[0][ww];
            s_start[t] = pre + x - c;
            __syncthreads();
            if (t < OG_BLOCK / 64) s_wcnt[0][t] = 0;
        }
        __syncthreads();
#pragma unroll
        for (uint32_t r = 0; r < RS_ROUNDS; r++)
        {
            if (my_rank[r] != 0xFFFFFFFFu)
            {
                uint32_t pos = s_start[my_dig[r]] + my_rank[r];
                s_key[pos] = my_key[r];
                s_val[pos] = my_val[r];
                s_dig[pos] = (uint8_t)my_dig[r];
            }
        }
        __syncthreads();
        for (uint32_t i = t; i < cnt_tile; i += OG_BLOCK)
        {
            uint32_t d = s_dig[i];
            uint32_t g = s_base[d] + (i - s_start[d]);
            k_out[g] = s_key[i];
            v_out[g] = s_val[i];
        }
        __syncthreads();
        s_base[t] += s_run[t];
        __syncthreads();
    }
}

/* Scratch: the (digit, block) counts plus their total, then the multi-block
 * scan's block sums. */
#define RS_COUNTS_WORDS (256u * RS_MAX_BLOCKS + 4u)

extern "C" size_t og_radix_scratch_bytes(uint32_t n)
{
    (void)n;
    return (size_t)RS_COUNTS_WORDS * sizeof(uint32_t) + og_scan_scratch_bytes(256u * RS_MAX_BLOCKS);
}

extern "C" int og_radix_sort_pairs(const uint32_t *k_in, const uint32_t *v_in, uint32_t *k_a,
                                   uint32_t *v_a, uint32_t *k_b, uint32_t *v_b, uint32_t n,
                                   uint32_t key_bits, void *d_scratch, hipStream_t s,
                                   int *result_in_b)
{
    /* Pass p reads the previous pass's output; the first pass reads
     * (k_in, v_in), which may be any arrays (e.g. the forward list read as
     * (dst, src) to form the inverse list).  Outputs alternate a, b, a, ... */
    uint32_t passes = (key_bits + 7) / 8;
    if (passes == 0) passes = 1;
    uint32_t *counts = (uint32_t *)d_scratch;
    uint32_t tiles = (n + RS_TILE - 1) / RS_TILE;
    uint32_t nb = tiles < RS_MAX_BLOCKS ? tiles : RS_MAX_BLOCKS;
    if (nb == 0) nb = 1;
    uint32_t tiles_per_block = (tiles + nb - 1) / nb;
    uint32_t chunk = tiles_per_block * RS_TILE;
    nb = (n + chunk - 1) / chunk;
    if (nb == 0) nb = 1;
    const uint32_t *ki = k_in, *vi = v_in;
    for (uint32_t p = 0; p < passes; p++)
    {
        uint32_t *ko = (p & 1) ? k_b : k_a;
        uint32_t *vo = (p & 1) ? v_b : v_a;
        OG_LAUNCH(og_rs_hist_kernel, dim3(nb), dim3(OG_BLOCK), 0, s, ki, n, chunk, p * 8,
                           counts);
        if (og_scan_exclusive(counts, counts, 256 * nb, counts + RS_COUNTS_WORDS, s)) return -1;
        OG_LAUNCH(og_rs_scatter_kernel, dim3(nb), dim3(OG_BLOCK), 0, s, ki, vi, ko, vo, n,
                           chunk, p * 8, counts);
        if (og_hip_check(og_launch_status(), "og_radix_sort_pairs")) return -1;
        ki = ko;
        vi = vo;
    }
    *result_in_b = ((passes - 1) & 1) ? 1 : 0;
    return 0;
}

/* ------------------------------------------------------------------ */
/* vertex mapping                                                      */
/* ------------------------------------------------------------------ */

__global__ void og_map_start_kernel(const uint32_t *__restrict__ key, uint32_t n,
                                    uint32_t *__restrict__ edges_idx)
{
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
    {
        uint32_t k = key[i];
        if (i == 0 || key[i - 1] != k) edges_idx[k & OG_ID_MASK] = i;
    }
}

__global__ void og_map_end_kernel(const uint32_t *__restrict__ key, uint32_t n,
                                  const uint32_t *__restrict__ edges_idx,
                                  uint32_t *__restrict__ out_degree)
{
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
    {
        uint32_t k = key[i];
        if (i == n - 1 || key[i + 1] != k)
        {
            uint32_t v = k & OG_ID_MASK;
            out_degree[v] = i + 1 - edges_idx[v];
        }
    }
}

extern "C" int og_launch_map_vertices(const uint32_t *d_key, uint32_t n_edges, uint32_t n_vertices,
                                      uint32_t *d_out_degree, uint32_t *d_edges_idx, hipStream_t s)
{
    if (og_launch_fill_u32(d_out_degree, 0, n_vertices, s)) return -1;
    if (og_launch_fill_u32(d_edges_idx, 0, n_vertices, s)) return -1;
    if (n_edges == 0) return 0;
    unsigned grid = og_grid_for(n_edges, OG_BLOCK, 65536);
    OG_LAUNCH(og_map_start_kernel, dim3(grid), dim3(OG_BLOCK), 0, s, d_key, n_edges,
                       d_edges_idx);
    OG_LAUNCH(og_map_end_kernel, dim3(grid), dim3(OG_BLOCK), 0, s, d_key, n_edges,
                       d_edges_idx, d_out_degree);
    return og_hip_check(og_launch_status(), "og_map_vertices");
}

/* ------------------------------------------------------------------ */
/* rank post-sort keys                                                 */
/* ------------------------------------------------------------------ */

/* fixedPoint.h:33 FloatToFixed32SORT: (uint32)(rank * 2^20 +- 0.5f). */
__global__ void og_rank_keys_kernel(const float *__restrict__ r, uint32_t n, uint32_t *__restrict__ k)
{
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
    {
        float x = r[i];
        k[i] = (uint32_t)og_x86_f2i64(x * 1048576.0f + (x >= 0 ? 0.5f : -0.5f));
    }
}

extern "C" int og_launch_rank_keys(const float *d_ranks, uint32_t n, uint32_t *d_keys, hipStream_t s)
{
    if (n == 0) return 0;
    OG_LAUNCH(og_rank_keys_kernel, dim3(og_grid_for(n, OG_BLOCK, 16384)), dim3(OG_BLOCK), 0, s,
                       d_ranks, n, d_keys);
    return og_hip_check(og_launch_status(), "og_rank_keys_kernel");
}

/* ------------------------------------------------------------------ */
/* plain copy (reference point for memory counters)                    */
/* ------------------------------------------------------------------ */

__global__ void og_copy_kernel(const uint4 *__restrict__ src, uint4 *__restrict__ dst, uint64_t n16)
{
    for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n16;
         i += (uint64_t)gridDim.x * blockDim.x)
        dst[i] = src[i];
}

extern "C" int og_launch_copy(const void *src, void *dst, uint64_t bytes, hipStream_t s)
{
    uint64_t n16 = bytes / 16;
    if (n16 == 0) return 0;
    OG_LAUNCH(og_copy_kernel, dim3(og_grid_for(n16, OG_BLOCK * 4, 65536)), dim3(OG_BLOCK), 0, s,
                       (const uint4 *)src, (uint4 *)dst, n16);
    return og_hip_check(og_launch_status(), "og_copy_kernel");
}
