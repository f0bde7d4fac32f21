/*
 * og_reorder.hip -- gfx950 kernels of the reference's degree relabelling
 * (-l 1..5, reorder.c:275-572 and :1483-1637): per-vertex degree counts,
 * the DBG bucket keys, label assignment from a stable sort, and the
 * relabelling of an edge list.  The stable sort itself is the CSR builder's
 * LSD radix sort (og_csr.hip).
 */
#include <hip/hip_runtime.h>

#include "og_internal.h"

#define RB 256

static unsigned og_rgrid(uint64_t n)
{
    uint64_t g = (n + RB - 1) / RB;
    if (g > 65536) g = 65536;
    return g ? (unsigned)g : 1u;
}

/* reorder.c:1483-1530: out-degree (src) and / or in-degree (dst) counts;
 * the sums are order-free. */
__global__ __launch_bounds__(RB) void og_reorder_degree_kernel(const uint32_t *__restrict__ src,
                                                               const uint32_t *__restrict__ dst, uint32_t E,
                                                               uint32_t count_src, uint32_t count_dst,
                                                               uint32_t *__restrict__ deg)
{
    for (uint32_t i = blockIdx.x * RB + threadIdx.x; i < E; i += gridDim.x * RB)
    {
        if (count_src) atomicAdd(&deg[src[i]], 1u);
        if (count_dst) atomicAdd(&deg[dst[i]], 1u);
    }
}

/* reorder.c:488-500: DBG bucket of each vertex, the first threshold at or
 * above its degree; the key (nb - 1 - bucket) sorts the last bucket first. */
__global__ __launch_bounds__(RB) void og_reorder_dbg_key_kernel(const uint32_t *__restrict__ deg, uint32_t N,
                                                                og_dbg_thresholds th,
                                                                uint32_t *__restrict__ key,
                                                                uint32_t *__restrict__ ids)
{
    for (uint32_t v = blockIdx.x * RB + threadIdx.x; v < N; v += gridDim.x * RB)
    {
        const uint32_t d = deg[v];
        uint32_t b = 0;
        while (b + 1 < th.n && d > th.t[b]) b++;
        key[v] = th.n - 1 - b;
        ids[v] = v;
    }
}

__global__ __launch_bounds__(RB) void og_reorder_iota_kernel(uint32_t *__restrict__ ids, uint32_t N)
{
    for (uint32_t v = blockIdx.x * RB + threadIdx.x; v < N; v += gridDim.x * RB) ids[v] = v;
}

/* labels[sorted[i]] = N - 1 - i (degree sort, reorder.c:343-347) or i
 * (DBG, reorder.c:515-523). */
__global__ __launch_bounds__(RB) void og_reorder_label_kernel(const uint32_t *__restrict__ sorted, uint32_t N,
                                                              uint32_t reverse, uint32_t *__restrict__ labels)
{
    for (uint32_t i = blockIdx.x * RB + threadIdx.x; i < N; i += gridDim.x * RB)
        labels[sorted[i]] = reverse ? N - 1 - i : i;
}

/* reorder.c:1618-1637 */
__global__ __launch_bounds__(RB) void og_relabel_kernel(uint32_t *__restrict__ src, uint32_t *__restrict__ dst,
                                                        uint32_t E, const uint32_t *__restrict__ labels)
{
    for (uint32_t i = blockIdx.x * RB + threadIdx.x; i < E; i += gridDim.x * RB)
    {
        src[i] = labels[src[i]];
        dst[i] = labels[dst[i]];
    }
}

/* label_array[v] = labels[label_array[v]], inverse[label_array[v]] = v
 * (reorder.c:368-373). */
__global__ __launch_bounds__(RB) void og_compose_labels_kernel(uint32_t *__restrict__ label_array,
                                                               uint32_t *__restrict__ inverse, uint32_t N,
                                                               const uint32_t *__restrict__ labels)
{
    for (uint32_t v = blockIdx.x * RB + threadIdx.x; v < N; v += gridDim.x * RB)
    {
        const uint32_t l = labels[label_array[v]];
        label_array[v] = l;
        if (inverse) inverse[l] = v;
    }
}

extern "C" int og_launch_reorder_degrees(const uint32_t *src, const uint32_t *dst, uint32_t E, uint32_t count_src,
                                         uint32_t count_dst, uint32_t *deg, hipStream_t s)
{
    if (E) OG_LAUNCH(og_reorder_degree_kernel, dim3(og_rgrid(E)), dim3(RB), 0, s, src, dst, E, count_src,
                              count_dst, deg);
    return og_hip_check(og_launch_status(), "og_reorder_degree_kernel");
}

extern "C" int og_launch_reorder_dbg_keys(const uint32_t *deg, uint32_t N, og_dbg_thresholds th, uint32_t *key,
                                          uint32_t *ids, hipStream_t s)
{
    OG_LAUNCH(og_reorder_dbg_key_kernel, dim3(og_rgrid(N)), dim3(RB), 0, s, deg, N, th, key, ids);
    return og_hip_check(og_launch_status(), "og_reorder_dbg_key_kernel");
}

extern "C" int og_launch_iota(uint32_t *ids, uint32_t N, hipStream_t s)
{
    OG_LAUNCH(og_reorder_iota_kernel, dim3(og_rgrid(N)), dim3(RB), 0, s, ids, N);
    return og_hip_check(og_launch_status(), "og_reorder_iota_kernel");
}

extern "C" int og_launch_reorder_labels(const uint32_t *sorted, uint32_t N, uint32_t reverse, uint32_t *labels,
                                        hipStream_t s)
{
    OG_LAUNCH(og_reorder_label_kernel, dim3(og_rgrid(N)), dim3(RB), 0, s, sorted, N, reverse, labels);
    return og_hip_check(og_launch_status(), "og_reorder_label_kernel");
}

extern "C" int og_launch_relabel(uint32_t *src, uint32_t *dst, uint32_t E, const uint32_t *labels, hipStream_t s)
{
    if (E) OG_LAUNCH(og_relabel_kernel, dim3(og_rgrid(E)), dim3(RB), 0, s, src, dst, E, labels);
    return og_hip_check(og_launch_status(), "og_relabel_kernel");
}

extern "C" int og_launch_compose_labels(uint32_t *label_array, uint32_t *inverse, uint32_t N,
                                        const uint32_t *labels, hipStream_t s)
{
    OG_LAUNCH(og_compose_labels_kernel, dim3(og_rgrid(N)), dim3(RB), 0, s, label_array, inverse, N,
                       labels);
    return og_hip_check(og_launch_status(), "og_compose_labels_kernel");
}
