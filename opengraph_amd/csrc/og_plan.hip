/*
 * og_plan.hip -- gfx950 kernels that build the per-graph PageRank plan from
 * the device CSR: the vertex relabelling, the relabelled inverse CSR (rows
 * keep the reference's in-neighbour order), the chunk descriptors of the
 * long rows, the relabelled forward edge list for push and the forward CSR for the
 * data-driven variants.  Also a general exclusive scan.
 */
#include <hip/hip_runtime.h>

#include "og_internal.h"
#include "og_plan.h"

#define OG_BLOCK 256
#define SCAN_ITEMS 16
#define SCAN_TILE (OG_BLOCK * SCAN_ITEMS)

/* ------------------------------------------------------------------ */
/* exclusive scan of u32 (out[n] = total)                              */
/* ------------------------------------------------------------------ */

__global__ __launch_bounds__(OG_BLOCK) void og_scan_tile_kernel(const uint32_t *__restrict__ in,
                                                               uint32_t *__restrict__ out, uint32_t n,
                                                               uint32_t *__restrict__ bsum)
{
    __shared__ uint32_t s[OG_BLOCK];
    const uint32_t t = threadIdx.x;
    const uint64_t base = (uint64_t)blockIdx.x * SCAN_TILE + (uint64_t)t * SCAN_ITEMS;
    uint32_t v[SCAN_ITEMS];
    uint32_t tot = 0;
#pragma unroll
    for (int i = 0; i < SCAN_ITEMS; i++)
    {
        v[i] = base + i < n ? in[base + i] : 0;
        tot += v[i];
    }
    s[t] = tot;
    __syncthreads();
    for (uint32_t o = 1; o < OG_BLOCK; o <<= 1)
    {
        uint32_t x = t >= o ? s[t - o] : 0;
        __syncthreads();
        s[t] += x;
        __syncthreads();
    }
    uint32_t run = s[t] - tot;
#pragma unroll
    for (int i = 0; i < SCAN_ITEMS; i++)
    {
        if (base + i < n) out[base + i] = run;
        run += v[i];
    }
    if (t == OG_BLOCK - 1) bsum[blockIdx.x] = s[t];
}

__global__ __launch_bounds__(1024) void og_scan_single_kernel(uint32_t *__restrict__ c, uint32_t total,
                                                             uint32_t *__restrict__ grand)
{
    __shared__ uint32_t part[1024];
    const uint32_t t = threadIdx.x;
    const uint32_t per = (total + 1023) / 1024;
    uint32_t b = t * per, e = min(b + per, total);
    uint32_t s = 0;
    for (uint32_t i = b; i < e; i++) s += c[i];
    part[t] = s;
    __syncthreads();
    for (uint32_t o = 1; o < 1024; o <<= 1)
    {
        uint32_t v = t >= o ? part[t - o] : 0;
        __syncthreads();
        part[t] += v;
        __syncthreads();
    }
    uint32_t run = part[t] - s;
    for (uint32_t i = b; i < e; i++)
    {
        uint32_t x = c[i];
        c[i] = run;
        run += x;
    }
    if (t == 1023 && grand) *grand = part[1023];
}

__global__ void og_scan_add_kernel(uint32_t *__restrict__ out, uint32_t n,
                                   const uint32_t *__restrict__ bsum)
{
    for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n;
         i += (uint64_t)gridDim.x * blockDim.x)
        out[i] += bsum[i / SCAN_TILE];
}

extern "C" size_t og_scan_scratch_bytes(uint32_t n)
{
    return ((size_t)(n / SCAN_TILE) + 2) * sizeof(uint32_t);
}

extern "C" int og_scan_exclusive(const uint32_t *d_in, uint32_t *d_out, uint32_t n, void *d_scratch,
                                 hipStream_t s)
{
    uint32_t nb = (n + SCAN_TILE - 1) / SCAN_TILE;
    uint32_t *bsum = (uint32_t *)d_scratch;
    if (n == 0)
    {
        OG_LAUNCH(og_scan_single_kernel, dim3(1), dim3(1024), 0, s, bsum, 0u, d_out);
        return og_hip_check(og_launch_status(), "og_scan_exclusive");
    }
    OG_LAUNCH(og_scan_tile_kernel, dim3(nb), dim3(OG_BLOCK), 0, s, d_in, d_out, n, bsum);
    OG_LAUNCH(og_scan_single_kernel, dim3(1), dim3(1024), 0, s, bsum, nb, d_out + n);
    unsigned g = nb < 65535 ? nb : 65535;
    OG_LAUNCH(og_scan_add_kernel, dim3(g), dim3(OG_BLOCK), 0, s, d_out, n, bsum);
    return og_hip_check(og_launch_status(), "og_scan_exclusive");
}

/* ------------------------------------------------------------------ */
/* relabelling                                                         */
/* ------------------------------------------------------------------ */

/* Row class of a vertex with in-degree d: 0 several chunks, 1 one chunk,
 * 2 summed by one thread, 3 no in-edges. */
__device__ __forceinline__ uint32_t og_row_class(uint32_t d, uint32_t seq_t)
{
    if (d > seq_t) return d > OG_CH ? 0u : 1u;
    return d ? 2u : 3u;
}

/* Sort key of old vertex v: its class; with relabel, then (classes 0, 3)
 * forward out-degree descending, so the contributions gathered most often
 * share cache lines, or (classes 1, 2) in-degree descending, so the rows or
 * chunks one wavefront takes have about the same length.  With chain_t,
 * the class-0 rows of more than chain_t in-edges go first, in-degree
 * descending.  Ties keep id order (the sort is stable). */
__global__ void og_plan_keys_kernel(uint32_t N, const uint32_t *__restrict__ outdeg,
                                    const uint32_t *__restrict__ indeg, uint32_t seq_t,
                                    uint32_t chain_t, uint32_t relabel, uint32_t *__restrict__ key,
                                    uint32_t *__restrict__ val)
{
    for (uint32_t v = blockIdx.x * blockDim.x + threadIdx.x; v < N; v += gridDim.x * blockDim.x)
    {
        uint32_t d = indeg[v];
        uint32_t cls = og_row_class(d, seq_t);
        uint32_t k;
        if (relabel && chain_t && cls == 0)
        {
            /* bit 29 clear: chains by in-degree; set: the rest by out-degree */
            const bool ch = d > chain_t;
            const uint32_t w = min(ch ? d : outdeg[v], 0x1FFFFFFFu);
            k = (ch ? 0u : 0x20000000u) | (0x1FFFFFFFu - w);
        }
        else if (relabel)
        {
            uint32_t w = min(cls == 1 || cls == 2 ? d : outdeg[v], 0x3FFFFFFFu);
            k = (cls << 30) | (0x3FFFFFFFu - w);
        }
        else
            k = cls;
        key[v] = k;
        val[v] = v;
    }
}

__global__ void og_plan_perm_kernel(uint32_t N, const uint32_t *__restrict__ new2old,
                                    const uint32_t *__restrict__ outdeg_old,
                                    const uint32_t *__restrict__ indeg_old, uint32_t seq_t,
                                    uint32_t chain_t, uint32_t *__restrict__ old2new,
                                    uint32_t *__restrict__ outdeg_new, uint32_t *__restrict__ indeg_new,
                                    uint32_t *__restrict__ class_cnt)
{
    uint32_t cnt[5] = {0, 0, 0, 0, 0};
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < N; i += gridDim.x * blockDim.x)
    {
        uint32_t o = new2old[i];
        old2new[o] = i;
        outdeg_new[i] = outdeg_old[o];
        uint32_t d = indeg_old[o];
        indeg_new[i] = d;
        cnt[og_row_class(d, seq_t)]++;
        if (chain_t && d > chain_t) cnt[4]++;
    }
    for (int k = 0; k < 5; k++)
        if (cnt[k]) atomicAdd(&class_cnt[k], cnt[k]);
}

/* cols[row_ptr[i] + k] = old2new[src_list[idx_old[o] + k]], o = new2old[i]:
 * one wavefront per row for rows [r0, r1). */
__global__ __launch_bounds__(OG_BLOCK) void og_plan_copy_rows_kernel(
    uint32_t r0, uint32_t r1, const uint32_t *__restrict__ new2old,
    const uint32_t *__restrict__ old2new, const uint32_t *__restrict__ idx_old,
    const uint32_t *__restrict__ lists, const uint32_t *__restrict__ row_ptr,
    uint32_t *__restrict__ cols)
{
    const uint32_t lane = threadIdx.x & 63;
    const uint32_t wave = (blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    const uint32_t nwaves = (gridDim.x * blockDim.x) >> 6;
    for (uint32_t i = r0 + wave; i < r1; i += nwaves)
    {
        uint32_t o = new2old[i];
        uint32_t src = idx_old[o];
        uint32_t dst = row_ptr[i], n = row_ptr[i + 1] - dst;
        for (uint32_t k = lane; k < n; k += 64) cols[dst + k] = old2new[lists[src + k] & OG_ID_MASK];
    }
}

/* The same for the rows [0, H) (more than one chunk), one block per row. */
__global__ __launch_bounds__(OG_BLOCK) void og_plan_copy_hubs_kernel(
    const uint32_t *__restrict__ new2old, const uint32_t *__restrict__ old2new,
    const uint32_t *__restrict__ idx_old, const uint32_t *__restrict__ lists,
    const uint32_t *__restrict__ row_ptr, uint32_t *__restrict__ cols)
{
    uint32_t i = blockIdx.x;
    uint32_t o = new2old[i];
    uint32_t src = idx_old[o];
    uint32_t dst = row_ptr[i], n = row_ptr[i + 1] - dst;
    for (uint32_t k = threadIdx.x; k < n; k += OG_BLOCK) cols[dst + k] = old2new[lists[src + k] & OG_ID_MASK];
}

/* rowid[e - base] = r for the edges e of rows [0, R) of row_ptr, one
 * wavefront per row. */
__global__ __launch_bounds__(OG_BLOCK) void og_plan_row_ids_kernel(uint32_t R, const uint32_t *__restrict__ row_ptr,
                                                                  uint32_t base, uint32_t *__restrict__ rowid)
{
    const uint32_t lane = threadIdx.x & 63;
    const uint32_t wave = (blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    const uint32_t nwaves = (gridDim.x * blockDim.x) >> 6;
    for (uint32_t r = wave; r < R; r += nwaves)
        for (uint32_t e = row_ptr[r] + lane; e < row_ptr[r + 1]; e += 64) rowid[e - base] = r;
}

extern "C" int og_launch_plan_row_ids_at(uint32_t R, const uint32_t *row_ptr, uint32_t base, uint32_t *rowid,
                                         hipStream_t s)
{
    if (!R) return 0;
    uint64_t waves = R;
    unsigned grid = (unsigned)((waves + 3) / 4 < 65536 ? (waves + 3) / 4 : 65536);
    OG_LAUNCH(og_plan_row_ids_kernel, dim3(grid), dim3(OG_BLOCK), 0, s, R, row_ptr, base, rowid);
    return og_hip_check(og_launch_status(), "og_plan_row_ids_kernel");
}

__global__ void og_plan_chunk_count_kernel(uint32_t C, const uint32_t *__restrict__ row_ptr,
                                           uint32_t ch, uint32_t *__restrict__ cnt)
{
    for (uint32_t r = blockIdx.x * blockDim.x + threadIdx.x; r < C; r += gridDim.x * blockDim.x)
        cnt[r] = (row_ptr[r + 1] - row_ptr[r] + ch - 1) / ch;
}

/* Chunk k of row r covers edges [row_ptr[r] + k*ch, min(+ch, row_ptr[r+1])):
 * chunks tile the edges of rows [0, C) in order, so ch_edge[c+1] ends
 * chunk c.  Rows below H have several chunks and carry OG_CH_MULTI. */
__global__ void og_plan_chunk_desc_kernel(uint32_t H, uint32_t C, const uint32_t *__restrict__ row_ptr,
                                          const uint32_t *__restrict__ first, uint32_t ch,
                                          uint32_t *__restrict__ ch_row, uint32_t *__restrict__ ch_edge)
{
    for (uint32_t r = blockIdx.x * blockDim.x + threadIdx.x; r <= C; r += gridDim.x * blockDim.x)
    {
        if (r == C)
        {
            ch_edge[first[C]] = row_ptr[C];
            continue;
        }
        uint32_t c0 = first[r], n = first[r + 1] - c0, e0 = row_ptr[r];
        uint32_t tag = r | (r < H ? OG_CH_MULTI : 0u);
        for (uint32_t k = 0; k < n; k++)
        {
            ch_row[c0 + k] = tag;
            ch_edge[c0 + k] = e0 + k * ch;
        }
    }
}

__global__ void og_plan_map_pairs_kernel(uint32_t E, const uint32_t *__restrict__ a,
                                         const uint32_t *__restrict__ b,
                                         const uint32_t *__restrict__ old2new,
                                         uint32_t *__restrict__ oa, uint32_t *__restrict__ ob)
{
    for (uint32_t e = blockIdx.x * blockDim.x + threadIdx.x; e < E; e += gridDim.x * blockDim.x)
    {
        oa[e] = old2new[a[e] & OG_ID_MASK];
        ob[e] = old2new[b[e] & OG_ID_MASK];
    }
}

static unsigned og_grid(uint64_t n, unsigned cap)
{
    uint64_t g = (n + OG_BLOCK - 1) / OG_BLOCK;
    if (g < 1) g = 1;
    if (g > cap) g = cap;
    return (unsigned)g;
}

extern "C" int og_launch_plan_keys(uint32_t N, const uint32_t *outdeg, const uint32_t *indeg,
                                   uint32_t seq_t, uint32_t chain_t, uint32_t relabel, uint32_t *key,
                                   uint32_t *val, hipStream_t s)
{
    OG_LAUNCH(og_plan_keys_kernel, dim3(og_grid(N, 16384)), dim3(OG_BLOCK), 0, s, N, outdeg,
                       indeg, seq_t, chain_t, relabel, key, val);
    return og_hip_check(og_launch_status(), "og_plan_keys_kernel");
}

extern "C" int og_launch_plan_perm(uint32_t N, const uint32_t *new2old, const uint32_t *outdeg_old,
                                   const uint32_t *indeg_old, uint32_t seq_t, uint32_t chain_t,
                                   uint32_t *old2new,
                                   uint32_t *outdeg_new, uint32_t *indeg_new, uint32_t *class_cnt,
                                   hipStream_t s)
{
    OG_LAUNCH(og_plan_perm_kernel, dim3(og_grid(N, 16384)), dim3(OG_BLOCK), 0, s, N, new2old,
                       outdeg_old, indeg_old, seq_t, chain_t, old2new, outdeg_new, indeg_new, class_cnt);
    return og_hip_check(og_launch_status(), "og_plan_perm_kernel");
}

extern "C" int og_launch_plan_copy(uint32_t H, uint32_t N, const uint32_t *new2old,
                                   const uint32_t *old2new, const uint32_t *idx_old,
                                   const uint32_t *lists, const uint32_t *row_ptr, uint32_t *cols,
                                   hipStream_t s)
{
    if (H) OG_LAUNCH(og_plan_copy_hubs_kernel, dim3(H), dim3(OG_BLOCK), 0, s, new2old, old2new,
                              idx_old, lists, row_ptr, cols);
    if (N > H)
    {
        uint64_t waves = N - H;
        unsigned grid = (unsigned)((waves + 3) / 4 < 65536 ? (waves + 3) / 4 : 65536);
        OG_LAUNCH(og_plan_copy_rows_kernel, dim3(grid), dim3(OG_BLOCK), 0, s, H, N, new2old,
                           old2new, idx_old, lists, row_ptr, cols);
    }
    return og_hip_check(og_launch_status(), "og_plan_copy");
}

extern "C" int og_launch_plan_chunks(uint32_t C, const uint32_t *row_ptr, uint32_t ch, uint32_t *cnt,
                                     hipStream_t s)
{
    if (!C) return 0;
    OG_LAUNCH(og_plan_chunk_count_kernel, dim3(og_grid(C, 16384)), dim3(OG_BLOCK), 0, s, C,
                       row_ptr, ch, cnt);
    return og_hip_check(og_launch_status(), "og_plan_chunk_count_kernel");
}

extern "C" int og_launch_plan_chunk_desc(uint32_t H, uint32_t C, const uint32_t *row_ptr,
                                         const uint32_t *first, uint32_t ch, uint32_t *ch_row,
                                         uint32_t *ch_edge, hipStream_t s)
{
    OG_LAUNCH(og_plan_chunk_desc_kernel, dim3(og_grid((uint64_t)C + 1, 16384)), dim3(OG_BLOCK),
                       0, s, H, C, row_ptr, first, ch, ch_row, ch_edge);
    return og_hip_check(og_launch_status(), "og_plan_chunk_desc_kernel");
}

extern "C" int og_launch_plan_map_pairs(uint32_t E, const uint32_t *a, const uint32_t *b,
                                        const uint32_t *old2new, uint32_t *oa, uint32_t *ob,
                                        hipStream_t s)
{
    if (!E) return 0;
    OG_LAUNCH(og_plan_map_pairs_kernel, dim3(og_grid(E, 65536)), dim3(OG_BLOCK), 0, s, E, a, b,
                       old2new, oa, ob);
    return og_hip_check(og_launch_status(), "og_plan_map_pairs_kernel");
}
