/*
 * og_spmv.hip -- gfx950 kernels around the reference's benchmark SpMV
 * (SPMV.c:588-1068): the per-vertex terms w * x[src] that the pull gathers
 * (or the push scatters) every iteration, and vector_output in the caller's
 * vertex ids.  The gathers themselves are PageRank's degree-binned pull
 * (og_pr.hip) with its SpMV store.
 *
 * The input vector is fixed (SPMV.c:663-669): x[v] = 1.0f / out-degree(v),
 * 0.001f for a vertex without out-edges; every weight is 0.0001f (the
 * reference's unweighted build).  x and w do not change between iterations,
 * so each term, a function of its source alone, is computed once here with
 * the reference's own arithmetic and gathered as it stands.
 */
#include <hip/hip_runtime.h>

#include "og_pr.h"

#define SB 256

static unsigned og_sgrid(uint32_t n)
{
    uint32_t g = (n + SB - 1) / SB;
    if (g > 4096) g = 4096;
    return g ? g : 1u;
}

/* SPMV.c:663-669 in new ids (a.outdeg is the forward out-degree). */
__device__ __forceinline__ float og_spmv_x(const og_pr_dev &a, uint32_t v)
{
    const uint32_t d = a.outdeg[v];
    return d ? 1.0f / (float)d : 0.001f;
}

/* code 0, 1 (SPMV.c:690, :780): weight * x in fp32. */
__global__ __launch_bounds__(SB) void og_spmv_terms_f32_kernel(og_pr_dev a, float *__restrict__ t)
{
    for (uint32_t v = blockIdx.x * SB + threadIdx.x; v < a.N; v += gridDim.x * SB)
        t[v] = 0.0001f * og_spmv_x(a, v);
}

/* code 2 (SPMV.c:857-893, fixedPoint.h:31 and :67-68): the Q16.16 input
 * FloatToFixed32(x), the weight FloatToFixed32(0.0001f) = 7 held in a
 * float, and MULFixed32V1 = the u64 product >> 16. */
__global__ __launch_bounds__(SB) void og_spmv_terms_fx32_kernel(og_pr_dev a, uint32_t *__restrict__ t)
{
    const float w = (float)(uint32_t)og_x86_f2i64(0.0001f * 65536.0f + 0.5f);
    const uint64_t wu = (uint64_t)og_x86_f2i64(w);
    for (uint32_t v = blockIdx.x * SB + threadIdx.x; v < a.N; v += gridDim.x * SB)
    {
        const float x = og_spmv_x(a, v);
        const uint64_t xq = (uint32_t)og_x86_f2i64(x * 65536.0f + 0.5f);
        t[v] = (uint32_t)((wu * xq) >> 16);
    }
}

/* code 3 (SPMV.c:995-1030, fixedPoint.h:36 and :71-72): Q32.32 input and
 * weight DoubleToFixed64, MULFixed64V1 = the 128-bit product >> 32. */
__global__ __launch_bounds__(SB) void og_spmv_terms_fx64_kernel(og_pr_dev a, uint64_t *__restrict__ t)
{
    const uint64_t wu = og_x86_d2u64((double)0.0001f * 4294967296.0 + 0.5);
    for (uint32_t v = blockIdx.x * SB + threadIdx.x; v < a.N; v += gridDim.x * SB)
    {
        const uint64_t xq = og_x86_d2u64((double)og_spmv_x(a, v) * 4294967296.0 + 0.5);
        const uint64_t lo = wu * xq, hi = __umul64hi(wu, xq);
        t[v] = (hi << 32) | (lo >> 32);
    }
}

/* vector_output in old ids: fp32 y (codes 0, 1), Fixed32ToFloat (code 2,
 * fixedPoint.h:41) or Fixed64ToDouble stored as float (code 3, :39). */
__global__ __launch_bounds__(SB) void og_spmv_out_kernel(og_pr_dev a, uint32_t code, float *__restrict__ out)
{
    for (uint32_t v = blockIdx.x * SB + threadIdx.x; v < a.N; v += gridDim.x * SB)
    {
        const uint32_t n = a.old2new[v];
        float y;
        if (code < 2) y = a.r[n];
        else if (code == 2) y = (float)((const uint64_t *)a.next)[n] / 65536.0f;
        else y = (float)((double)((const uint64_t *)a.next)[n] / 4294967296.0);
        out[v] = y;
    }
}

extern "C" int og_launch_spmv_terms(const og_pr_dev *a, uint32_t code, void *terms, hipStream_t s)
{
    const unsigned g = og_sgrid(a->N);
    if (code < 2) OG_LAUNCH(og_spmv_terms_f32_kernel, dim3(g), dim3(SB), 0, s, *a, (float *)terms);
    else if (code == 2)
        OG_LAUNCH(og_spmv_terms_fx32_kernel, dim3(g), dim3(SB), 0, s, *a, (uint32_t *)terms);
    else OG_LAUNCH(og_spmv_terms_fx64_kernel, dim3(g), dim3(SB), 0, s, *a, (uint64_t *)terms);
    return og_hip_check(og_launch_status(), "og_spmv_terms");
}

extern "C" int og_launch_spmv_out(const og_pr_dev *a, uint32_t code, float *out_old, hipStream_t s)
{
    OG_LAUNCH(og_spmv_out_kernel, dim3(og_sgrid(a->N)), dim3(SB), 0, s, *a, code, out_old);
    return og_hip_check(og_launch_status(), "og_spmv_out_kernel");
}
