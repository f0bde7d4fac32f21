/*
 * og_ppr.hip -- gfx950 kernels of batched personalized PageRank: k <= 64
 * restart vertices iterated together as a pull-SpMM over an N x k block.
 *
 * For column j with restart vertex s_j, one iteration is
 *     x'[v][j] = (1 - d) [v == s_j] + d * sum_{u -> v} c[u][j],
 *     c[u][j]  = x[u][j] / outdeg(u)  (0 for a vertex without out-edges),
 * d = 0.85, the same arithmetic as the single-source pull (og_pr.hip with
 * the personalized option), starting from x = e_{s_j}.  Column j stops
 * (keeps its x) after the first iteration whose L1 change sum_v |x'-x| is
 * below epsilon.
 *
 * Layout: x, c[2] are N x k fp32, row-major (row v = k consecutive floats,
 * 256 B for k = 64), in the plan's vertex order.  An iteration is one pass
 * per 32 column slots in use; in a pass 32 lanes own the 32 slots of a work
 * item, so every gathered piece c[u][32 cb .. 32 cb + 32) is one coalesced
 * 128 B read, and a row's sum runs over its in-edges in CSR order per column
 * -- for rows of at most OG_CH in-edges bit-exactly the single-source
 * sequential sum.  Rows with more in-edges are cut into OG_CH-edge chunks
 * (the plan's chunks) whose sums the finalize kernel adds in chunk order.
 *
 * Work items of the main kernel: the plan's chunks (rows [0, C)), then the
 * rows [C, N) one each; wavefront w of the fixed grid takes items
 * w, w + W, w + 2W, ... so the per-column L1 partial sums are deterministic.
 */
#include <hip/hip_runtime.h>

#include "og_internal.h"
#include "og_ppr.h"

#define PB 256           /* block */
#define PWAVES (PB / 64)
#define PG 16            /* gathers in flight per lane */

/* New x and c of row v in slot `col` of an iterating column, and its L1.  A
 * frozen column keeps its x; its c is never gathered again. */
__device__ __forceinline__ void og_ppr_epilogue(const og_ppr_dev &a, uint32_t v, float s,
                                                uint32_t col, uint32_t my_src, bool on,
                                                double &l1)
{
    if (!on) return;
    const size_t i = (size_t)v * a.k + col;
    const float xo = a.x[i];
    const float base = v == my_src ? OG_BASE_PR : 0.0f;
    const float xn = base + OG_DAMP * s;
    a.x[i] = xn;
    l1 += fabs((double)(xn - xo));
    const uint32_t deg = a.outdeg[v];
    a.c_out[i] = deg ? xn / (float)deg : 0.0f;
}

/* Per-column sum of a per-lane double over the block, in wave order;
 * valid in threads [0, 64). */
__device__ __forceinline__ double og_ppr_block_col(double v)
{
    __shared__ double sh[PWAVES][64];
    const uint32_t lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    sh[w][lane] = v;
    __syncthreads();
    double r = 0.0;
    if (threadIdx.x < 64)
        for (uint32_t i = 0; i < PWAVES; i++) r += sh[i][threadIdx.x];
    __syncthreads();
    return r;
}

/* One pass of an iteration over column slots [32 cb, 32 cb + 32): a
 * wavefront takes PRW = 2 consecutive work items at a time, 32 lanes each,
 * lane 32 r + j owning slot 32 cb + j of item r.  Every gathered piece
 * c[u][32 cb .. 32 cb + 32) is one whole 128-byte line, and a pass's
 * working set is N x 128 B of contributions instead of N x 256 B (16
 * columns per pass read half lines and took 3.8 ms per iteration at
 * R-MAT-22, k = 64; 32 columns 3.0 ms; the whole 256-byte row 4.6 ms).
 * Each lane adds its column's terms over the item's in-edges in CSR order
 * (+0 past the end), as before, so the results do not depend on the
 * blocking. */
#define PCB 32u          /* column slots per pass: a 128-byte line of a row */
#define PRW (64u / PCB)  /* work items per wavefront */
__global__ __launch_bounds__(PB) void og_ppr_main_kernel(og_ppr_dev a, uint32_t cb)
{
    if (*a.done) return;
    const uint32_t lane = threadIdx.x & 63;
    const uint32_t rr = lane / PCB, jj = lane % PCB;
    const uint32_t slot = cb * PCB + jj;
    const uint32_t col = min(slot, a.k - 1);
    const bool live = slot < a.kact;
    const uint64_t act = *a.active;
    const bool on = live && ((act >> slot) & 1u);
    const uint32_t my_src = a.src[col];
    const uint32_t wave = blockIdx.x * PWAVES + __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const uint32_t W = gridDim.x * PWAVES;
    const uint32_t items = a.n_chunks + (a.N - a.C);
    double l1 = 0.0;
    for (uint32_t it0 = wave * PRW; it0 < items; it0 += W * PRW)
    {
        const uint32_t it = it0 + rr;
        const bool have = it < items;
        uint32_t eb = 0, ee = 0, row = 0, tag = 0;
        if (have)
        {
            if (it < a.n_chunks)
            {
                tag = a.ch_row[it];
                row = tag & ~OG_CH_MULTI;
                eb = a.ch_edge[it];
                ee = a.ch_edge[it + 1];
            }
            else
            {
                row = a.C + (it - a.n_chunks);
                eb = a.row_ptr[row];
                ee = a.row_ptr[row + 1];
            }
        }
        const uint32_t len = ee - eb;
        uint32_t maxlen = len;   /* the wavefront's longest item */
        for (int o = PCB; o < 64; o <<= 1) maxlen = max(maxlen, (uint32_t)__shfl_xor(maxlen, o, 64));
        float s = 0.0f;
        if (on)
            for (uint32_t b0 = 0; b0 < maxlen; b0 += PG)
            {
                /* branch-free: slots past the item's end re-read an edge in
                 * bounds and add +0 */
                float v[PG];
#pragma unroll
                for (uint32_t u = 0; u < PG; u++)
                {
                    const uint32_t e = min(eb + min(b0 + u, len - 1u), a.E - 1u);
                    const uint32_t src = a.cols[e];
                    const float t = a.c_in[(size_t)src * a.k + slot];
                    v[u] = b0 + u < len ? t : 0.0f;
                }
#pragma unroll
                for (uint32_t u = 0; u < PG; u++) s += v[u];
            }
        if (!have || !live) continue;
        if (it < a.n_chunks && (tag & OG_CH_MULTI)) a.chunk_part[(size_t)it * a.k + slot] = s;
        else og_ppr_epilogue(a, row, s, slot, my_src, on, l1);
    }
    /* per slot: the PRW lane groups of each wavefront, then the wavefronts,
     * in a fixed order */
    __shared__ double sh[PWAVES][64];
    const uint32_t w = threadIdx.x >> 6;
    sh[w][lane] = l1;
    __syncthreads();
    if (threadIdx.x < PCB && cb * PCB + threadIdx.x < a.k)
    {
        double r = 0.0;
        for (uint32_t i = 0; i < PWAVES; i++)
            for (uint32_t g = 0; g < PRW; g++) r += sh[i][g * PCB + threadIdx.x];
        a.part_l1[(size_t)blockIdx.x * a.k + cb * PCB + threadIdx.x] = r;
    }
}

/* Rows [0, H): add the chunk sums in chunk order, epilogue.  Then the
 * per-column L1 of the iteration in a fixed order: block b folds main
 * blocks [b*R, (b+1)*R), the last block (by ticket) folds the block results
 * and freezes the columns whose L1 fell below epsilon. */
__global__ __launch_bounds__(PB) void og_ppr_finalize_kernel(og_ppr_dev a)
{
    if (*a.done) return;
    const uint32_t lane = threadIdx.x & 63;
    const uint32_t col = min(lane, a.k - 1);
    const bool live = lane < a.kact;
    const uint64_t act = *a.active;
    const bool on = live && ((act >> lane) & 1u);
    const uint32_t my_src = a.src[col];
    const uint32_t wave = blockIdx.x * PWAVES + __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    double l1 = 0.0;
    for (uint32_t h = wave; h < a.H; h += gridDim.x * PWAVES)
    {
        const uint32_t cb = a.chunk_base[h], ce = a.chunk_base[h + 1];
        float s = 0.0f;
        for (uint32_t c0 = cb; c0 < ce; c0 += PG)
        {
            /* PG chunk sums in flight, added in chunk order (+0 past the end) */
            float v[PG];
#pragma unroll
            for (uint32_t u = 0; u < PG; u++)
                v[u] = c0 + u < ce ? a.chunk_part[(size_t)(c0 + u) * a.k + col] : 0.0f;
#pragma unroll
            for (uint32_t u = 0; u < PG; u++) s += v[u];
        }
        if (live) og_ppr_epilogue(a, h, s, col, my_src, on, l1);
    }
    /* fold this block's share of the main kernel's partials */
    const uint32_t per = OG_PPR_MAIN_BLOCKS / OG_PPR_FIN_BLOCKS;
    const uint32_t w = threadIdx.x >> 6;
    for (uint32_t i = w; i < per; i += PWAVES)
        if (live) l1 += a.part_l1[(size_t)(blockIdx.x * per + i) * a.k + col];
    const double b = og_ppr_block_col(l1);
    __shared__ uint32_t s_last;
    if (threadIdx.x < a.k) a.fin_l1[(size_t)blockIdx.x * a.k + threadIdx.x] = b;
    __threadfence();
    __syncthreads();
    if (threadIdx.x == 0) s_last = atomicAdd(a.ticket, 1u) == gridDim.x - 1;
    __syncthreads();
    if (!s_last) return;
    __threadfence();
    /* four threads per column, each a fixed stride of the block results */
    const uint32_t q = threadIdx.x >> 6;
    double t = 0.0;
    if (live)
    {
        double vals[OG_PPR_FIN_BLOCKS / PWAVES];
#pragma unroll
        for (uint32_t i = 0; i < OG_PPR_FIN_BLOCKS / PWAVES; i++)
            vals[i] = __hip_atomic_load(&a.fin_l1[(size_t)(q + i * PWAVES) * a.k + col], __ATOMIC_RELAXED,
                                        __HIP_MEMORY_SCOPE_AGENT);
#pragma unroll
        for (uint32_t i = 0; i < OG_PPR_FIN_BLOCKS / PWAVES; i++) t += vals[i];
    }
    const double tot = og_ppr_block_col(t);
    __shared__ uint32_t s_frozen[2];
    if (threadIdx.x == 0) s_frozen[0] = s_frozen[1] = 0;
    __syncthreads();
    if (threadIdx.x < a.k)
    {
        const uint32_t j = threadIdx.x;
        const uint32_t column = a.colmap[j];
        const bool was = (act >> j) & 1u;
        a.l1[(size_t)a.it * a.k + column] = was ? tot : 0.0;
        if (was)
        {
            a.iters[column] = a.it + 1;
            if (tot < a.epsilon) atomicOr(&s_frozen[j >> 5], 1u << (j & 31));
        }
    }
    __syncthreads();
    if (threadIdx.x == 0)
    {
        const uint64_t frozen = (uint64_t)s_frozen[0] | ((uint64_t)s_frozen[1] << 32);
        const uint64_t next = act & ~frozen;
        *a.active = next;
        if (next == 0 || a.it + 1 >= a.max_iter) *a.done = 1;
        *a.ticket = 0;
    }
}

/* x = e_{s_j} per column, c = x / outdeg; every column active. */
__global__ void og_ppr_init_kernel(og_ppr_dev a)
{
    const size_t n = (size_t)a.N * a.k;
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
    {
        const uint32_t v = (uint32_t)(i / a.k), j = (uint32_t)(i % a.k);
        const float x = v == a.src[j] ? 1.0f : 0.0f;
        a.x[i] = x;
        const uint32_t deg = a.outdeg[v];
        ((float *)a.c_in)[i] = deg ? x / (float)deg : 0.0f;
    }
    if (blockIdx.x == 0 && threadIdx.x == 0)
    {
        *a.active = a.k == 64 ? ~(uint64_t)0 : (((uint64_t)1 << a.k) - 1);
        *a.done = a.max_iter == 0;
        *a.ticket = 0;
    }
    if (blockIdx.x == 0 && threadIdx.x < a.k) a.iters[threadIdx.x] = 0;
}

/* out[old][colmap[j]] = x[old2new[old]][j]: the caller's vertex order and
 * column order. */
__global__ void og_ppr_out_kernel(og_ppr_dev a, float *out)
{
    const size_t n = (size_t)a.N * a.k;
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
    {
        const uint32_t o = (uint32_t)(i / a.k), j = (uint32_t)(i % a.k);
        out[(size_t)o * a.k + a.colmap[j]] = a.x[(size_t)a.old2new[o] * a.k + j];
    }
}

/* Slot j of every row of x and c takes old slot perm[j]: a wavefront per
 * row, every lane reads before any writes, so the move is in place. */
__global__ __launch_bounds__(PB) void og_ppr_permute_kernel(og_ppr_dev a, float *c, og_ppr_perm perm)
{
    const uint32_t lane = threadIdx.x & 63;
    const uint32_t j = min(lane, a.k - 1);
    const uint32_t from = perm.p[j];
    const uint32_t W = gridDim.x * PWAVES;
    for (uint32_t v = blockIdx.x * PWAVES + (threadIdx.x >> 6); v < a.N; v += W)
    {
        const size_t row = (size_t)v * a.k;
        const float xv = a.x[row + from], cv = c[row + from];
        __builtin_amdgcn_wave_barrier();
        if (lane < a.k)
        {
            a.x[row + lane] = xv;
            c[row + lane] = cv;
        }
    }
}

extern "C" int og_launch_ppr_permute(const og_ppr_dev *a, float *c, og_ppr_perm perm, hipStream_t s)
{
    OG_LAUNCH(og_ppr_permute_kernel, dim3(4096), dim3(PB), 0, s, *a, c, perm);
    return og_hip_check(og_launch_status(), "og_ppr_permute_kernel");
}

extern "C" int og_launch_ppr_init(const og_ppr_dev *a, hipStream_t s)
{
    OG_LAUNCH(og_ppr_init_kernel, dim3(4096), dim3(PB), 0, s, *a);
    return og_hip_check(og_launch_status(), "og_ppr_init_kernel");
}

extern "C" int og_launch_ppr_iter(const og_ppr_dev *a, hipStream_t s, hipEvent_t ev0, hipEvent_t ev1)
{
    if (ev0) hipEventRecord(ev0, s);
    /* one pass per 32 column slots in use */
    for (uint32_t cb = 0; cb * PCB < a->kact; cb++)
        OG_LAUNCH(og_ppr_main_kernel, dim3(OG_PPR_MAIN_BLOCKS), dim3(PB), 0, s, *a, cb);
    if (ev1) hipEventRecord(ev1, s);
    OG_LAUNCH(og_ppr_finalize_kernel, dim3(OG_PPR_FIN_BLOCKS), dim3(PB), 0, s, *a);
    return og_hip_check(og_launch_status(), "og_ppr_main_kernel");
}

extern "C" int og_launch_ppr_out(const og_ppr_dev *a, float *out_old, hipStream_t s)
{
    OG_LAUNCH(og_ppr_out_kernel, dim3(4096), dim3(PB), 0, s, *a, out_old);
    return og_hip_check(og_launch_status(), "og_ppr_out_kernel");
}
