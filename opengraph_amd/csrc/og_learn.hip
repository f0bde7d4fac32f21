/*
 * og_learn.hip -- gfx950 kernels of the learned warm start: degree
 * features, the MLP's inference (fp32 VALU, and layer 2 on bf16 MFMA), its
 * training gradient (forward + backward by hand, deterministic block
 * partials), Adam, and the scaling of the prediction into a start vector.
 *
 * The start vector: vertices without in-edges get their fixed-point rank
 * (1 - d) exactly (unnormalized); the others (U) get the predicted shape
 * o / sum_U(o) scaled to the mass the reference's iteration (no dangling
 * redistribution, pageRank.c:989-1028) keeps at its fixed point, from the
 * identity S = 0.15 + 0.85 (S - D), D the mass on vertices without
 * out-edges (og_warm_factor_kernel).
 */
#include <hip/hip_runtime.h>
#include <string.h>

#include "og_internal.h"
#include "og_learn.h"
#include "og_pr.h"
#include "opengraph_amd_ext.h"

#define LB 256

typedef short og_bf16x4 __attribute__((ext_vector_type(4)));
typedef float og_f32x16 __attribute__((ext_vector_type(16)));

__device__ __forceinline__ float og_softplus(float z) { return z > 20.0f ? z : log1pf(expf(z)); }
__device__ __forceinline__ float og_sigmoid(float z) { return 1.0f / (1.0f + expf(-z)); }

/* fp32 -> bf16 bits, round to nearest even (finite inputs). */
__device__ __forceinline__ short og_bf16(float x)
{
    uint32_t u = __float_as_uint(x);
    u += 0x7FFFu + ((u >> 16) & 1u);
    return (short)(u >> 16);
}

/* ------------------------------------------------------------------ */
/* features                                                            */
/* ------------------------------------------------------------------ */

__global__ void og_feat_li_kernel(const uint32_t *__restrict__ row_ptr, uint32_t n, float *__restrict__ li)
{
    for (uint32_t v = blockIdx.x * blockDim.x + threadIdx.x; v < n; v += gridDim.x * blockDim.x)
        li[v] = log1pf((float)(row_ptr[v + 1] - row_ptr[v]));
}

__global__ void og_feat_finish_kernel(const uint32_t *__restrict__ row_ptr, const uint32_t *__restrict__ outdeg,
                                      const float *__restrict__ insum, uint32_t n, float4 *__restrict__ feat)
{
    const float inv_n = 1.0f / (float)n;
    for (uint32_t v = blockIdx.x * blockDim.x + threadIdx.x; v < n; v += gridDim.x * blockDim.x)
    {
        const uint32_t din = row_ptr[v + 1] - row_ptr[v];
        feat[v] = make_float4(log1pf((float)din), log1pf((float)outdeg[v]),
                              din ? insum[v] / (float)din : 0.0f, inv_n);
    }
}

/* ------------------------------------------------------------------ */
/* inference                                                           */
/* ------------------------------------------------------------------ */

__global__ __launch_bounds__(LB) void og_mlp_infer_f32_kernel(const float *__restrict__ P,
                                                             const float4 *__restrict__ feat, uint32_t n,
                                                             float *__restrict__ out)
{
    __shared__ float sW[OG_MLP_NP];
    for (uint32_t i = threadIdx.x; i < OG_MLP_NP; i += LB) sW[i] = P[i];
    __syncthreads();
    for (uint32_t v = blockIdx.x * LB + threadIdx.x; v < n; v += gridDim.x * LB)
    {
        const float4 f = feat[v];
        float h1[OG_MLP_H];
#pragma unroll
        for (int k = 0; k < OG_MLP_H; k++)
        {
            float a = sW[OG_MLP_B1 + k];
            a = fmaf(sW[OG_MLP_W1 + 4 * k + 0], f.x, a);
            a = fmaf(sW[OG_MLP_W1 + 4 * k + 1], f.y, a);
            a = fmaf(sW[OG_MLP_W1 + 4 * k + 2], f.z, a);
            a = fmaf(sW[OG_MLP_W1 + 4 * k + 3], f.w, a);
            h1[k] = fmaxf(a, 0.0f);
        }
        float z3 = sW[OG_MLP_B3];
        for (int m = 0; m < OG_MLP_H; m++)
        {
            float a = sW[OG_MLP_B2 + m];
#pragma unroll
            for (int i = 0; i < OG_MLP_H; i++) a = fmaf(sW[OG_MLP_W2 + OG_MLP_H * m + i], h1[i], a);
            z3 = fmaf(sW[OG_MLP_W3 + m], fmaxf(a, 0.0f), z3);
        }
        out[v] = og_softplus(z3);
    }
}

/* Layer 2 on MFMA: a wavefront takes 32 vertices at a time; H1 (32 x 64,
 * bf16) times W2^T (64 x 64, bf16) accumulates in fp32 through
 * v_mfma_f32_32x32x8_bf16 (two 32-column output tiles, eight k-steps).
 * Operand layout of that instruction: lane l holds A[l%32][4(l/32)+j] and
 * B[4(l/32)+j][l%32] (j < 4); result register r of lane l is
 * D[8(r/4) + 4(l/32) + r%4][l%32].  Layers 1 and 3 stay on the VALU in
 * fp32; layer 3 is a butterfly over the 32 lanes of each half. */
__global__ __launch_bounds__(LB) void og_mlp_infer_mfma_kernel(const float *__restrict__ P,
                                                              const float4 *__restrict__ feat, uint32_t n,
                                                              float *__restrict__ out)
{
    __shared__ float sW1[4 * OG_MLP_H], sB1[OG_MLP_H];
    for (uint32_t i = threadIdx.x; i < 4 * OG_MLP_H; i += LB) sW1[i] = P[OG_MLP_W1 + i];
    for (uint32_t i = threadIdx.x; i < OG_MLP_H; i += LB) sB1[i] = P[OG_MLP_B1 + i];
    __syncthreads();
    const uint32_t lane = threadIdx.x & 63, hs = lane >> 5, col = lane & 31;
    og_bf16x4 bfr[2][8];
#pragma unroll
    for (int tt = 0; tt < 2; tt++)
#pragma unroll
        for (int s = 0; s < 8; s++)
#pragma unroll
            for (int j = 0; j < 4; j++)
                bfr[tt][s][j] = og_bf16(P[OG_MLP_W2 + OG_MLP_H * (32 * tt + col) + 8 * s + 4 * hs + j]);
    const float b2c0 = P[OG_MLP_B2 + col], b2c1 = P[OG_MLP_B2 + 32 + col];
    const float w3c0 = P[OG_MLP_W3 + col], w3c1 = P[OG_MLP_W3 + 32 + col];
    const float b3 = P[OG_MLP_B3];
    const uint32_t nw = gridDim.x * (LB / 64);
    for (uint32_t tile = blockIdx.x * (LB / 64) + (threadIdx.x >> 6); (uint64_t)tile * 32 < n; tile += nw)
    {
        const uint32_t vb = tile * 32;
        const float4 f = feat[min(vb + col, n - 1)];
        og_bf16x4 afr[8];
#pragma unroll
        for (int s = 0; s < 8; s++)
#pragma unroll
            for (int j = 0; j < 4; j++)
            {
                const int k = 8 * s + 4 * hs + j;
                float a = sB1[k];
                a = fmaf(sW1[4 * k + 0], f.x, a);
                a = fmaf(sW1[4 * k + 1], f.y, a);
                a = fmaf(sW1[4 * k + 2], f.z, a);
                a = fmaf(sW1[4 * k + 3], f.w, a);
                afr[s][j] = og_bf16(fmaxf(a, 0.0f));
            }
        og_f32x16 acc0 = {}, acc1 = {};
#pragma unroll
        for (int s = 0; s < 8; s++)
        {
            acc0 = __builtin_amdgcn_mfma_f32_32x32x8bf16_1k(afr[s], bfr[0][s], acc0, 0, 0, 0);
            acc1 = __builtin_amdgcn_mfma_f32_32x32x8bf16_1k(afr[s], bfr[1][s], acc1, 0, 0, 0);
        }
        float p[16];
#pragma unroll
        for (int r = 0; r < 16; r++)
            p[r] = fmaf(w3c0, fmaxf(acc0[r] + b2c0, 0.0f), w3c1 * fmaxf(acc1[r] + b2c1, 0.0f));
#pragma unroll
        for (int o = 1; o < 32; o <<= 1)
#pragma unroll
            for (int r = 0; r < 16; r++) p[r] += __shfl_xor(p[r], o, 64);
        float mine = p[0];
#pragma unroll
        for (int r = 1; r < 16; r++)
            if (col == (uint32_t)r) mine = p[r];
        if (col < 16)
        {
            const uint32_t row = 8 * (col >> 2) + 4 * hs + (col & 3);
            if (vb + row < n) out[vb + row] = og_softplus(mine + b3);
        }
    }
}

/* ------------------------------------------------------------------ */
/* training                                                            */
/* ------------------------------------------------------------------ */

#define TB OG_MLP_TRAIN_BLOCK
#define PART_STRIDE (OG_MLP_NP + 1)   /* gradient partial + loss partial per block */

/* Dataset index of step position x: the Feistel permutation of [0, n)
 * by cycle walking (identity when p.bits == 0). */
__device__ __forceinline__ uint32_t og_data_index(const og_perm_t &p, uint32_t x, uint32_t n)
{
    if (p.bits == 0) return x;
    do x = og_perm_apply(&p, x);
    while (x >= n);
    return x;
}

/* One block: TB vertices.  Each thread runs its vertex forward and
 * backward to dz1, dz2 and g3 = dL/dz3 (L1 loss: sign(o - y) * softplus');
 * the parameter gradients are then block sums over the vertices in index
 * order, formed from LDS (dW2 = sum dz2 (x) h1, dW1 = sum dz1 (x) f). */
__global__ __launch_bounds__(TB) void og_mlp_grad_kernel(const float *__restrict__ P,
                                                        const float4 *__restrict__ feat,
                                                        const float *__restrict__ y, uint32_t n_data,
                                                        og_perm_t perm, uint32_t base, uint32_t count,
                                                        float *__restrict__ part)
{
    __shared__ float sH1[TB][OG_MLP_H + 1];   /* h1, later dz1 */
    __shared__ float sD2[TB][OG_MLP_H + 1];   /* dz2 */
    __shared__ float sG[TB][OG_MLP_H + 1];    /* g3 * h2 */
    __shared__ float sF[TB][4];
    __shared__ float sg3[TB], sloss[TB];
    const uint32_t t = threadIdx.x;
    const uint32_t i = blockIdx.x * TB + t;
    const bool valid = i < count;
    const uint32_t v = valid ? og_data_index(perm, base + i, n_data) : 0;
    const float4 f = valid ? feat[v] : make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    sF[t][0] = f.x; sF[t][1] = f.y; sF[t][2] = f.z; sF[t][3] = f.w;
    float z1[OG_MLP_H];
#pragma unroll
    for (int k = 0; k < OG_MLP_H; k++)
    {
        float a = P[OG_MLP_B1 + k];
        a = fmaf(P[OG_MLP_W1 + 4 * k + 0], f.x, a);
        a = fmaf(P[OG_MLP_W1 + 4 * k + 1], f.y, a);
        a = fmaf(P[OG_MLP_W1 + 4 * k + 2], f.z, a);
        a = fmaf(P[OG_MLP_W1 + 4 * k + 3], f.w, a);
        z1[k] = a;
        sH1[t][k] = fmaxf(a, 0.0f);
    }
    float z2[OG_MLP_H];
    float z3 = P[OG_MLP_B3];
    for (int m = 0; m < OG_MLP_H; m++)
    {
        float a = P[OG_MLP_B2 + m];
#pragma unroll
        for (int k = 0; k < OG_MLP_H; k++) a = fmaf(P[OG_MLP_W2 + OG_MLP_H * m + k], fmaxf(z1[k], 0.0f), a);
        z2[m] = a;
        z3 = fmaf(P[OG_MLP_W3 + m], fmaxf(a, 0.0f), z3);
    }
    const float o = og_softplus(z3);
    const float d = valid ? o - y[v] : 0.0f;
    const float g3 = (float)((d > 0.0f) - (d < 0.0f)) * og_sigmoid(z3);
    sg3[t] = g3;
    sloss[t] = fabsf(d);
    float dz1[OG_MLP_H];
#pragma unroll
    for (int k = 0; k < OG_MLP_H; k++) dz1[k] = 0.0f;
    for (int m = 0; m < OG_MLP_H; m++)
    {
        const float dm = z2[m] > 0.0f ? g3 * P[OG_MLP_W3 + m] : 0.0f;
        sD2[t][m] = dm;
        sG[t][m] = g3 * fmaxf(z2[m], 0.0f);
#pragma unroll
        for (int k = 0; k < OG_MLP_H; k++) dz1[k] = fmaf(dm, P[OG_MLP_W2 + OG_MLP_H * m + k], dz1[k]);
    }
    __syncthreads();
    float *out = part + (size_t)blockIdx.x * PART_STRIDE;
    {
        /* thread t: row t of dW2, db2[t], dW3[t] */
        float acc[OG_MLP_H];
#pragma unroll
        for (int k = 0; k < OG_MLP_H; k++) acc[k] = 0.0f;
        float db2 = 0.0f, dw3 = 0.0f;
        for (uint32_t u = 0; u < TB; u++)
        {
            const float dm = sD2[u][t];
            db2 += dm;
            dw3 += sG[u][t];
#pragma unroll
            for (int k = 0; k < OG_MLP_H; k++) acc[k] = fmaf(dm, sH1[u][k], acc[k]);
        }
#pragma unroll
        for (int k = 0; k < OG_MLP_H; k++) out[OG_MLP_W2 + OG_MLP_H * t + k] = acc[k];
        out[OG_MLP_B2 + t] = db2;
        out[OG_MLP_W3 + t] = dw3;
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < OG_MLP_H; k++) sH1[t][k] = z1[k] > 0.0f ? dz1[k] : 0.0f;
    __syncthreads();
    {
        /* thread t: row t of dW1, db1[t] */
        float a0 = 0.0f, a1 = 0.0f, a2 = 0.0f, a3 = 0.0f, db1 = 0.0f;
        for (uint32_t u = 0; u < TB; u++)
        {
            const float dz = sH1[u][t];
            db1 += dz;
            a0 = fmaf(dz, sF[u][0], a0);
            a1 = fmaf(dz, sF[u][1], a1);
            a2 = fmaf(dz, sF[u][2], a2);
            a3 = fmaf(dz, sF[u][3], a3);
        }
        out[OG_MLP_W1 + 4 * t + 0] = a0;
        out[OG_MLP_W1 + 4 * t + 1] = a1;
        out[OG_MLP_W1 + 4 * t + 2] = a2;
        out[OG_MLP_W1 + 4 * t + 3] = a3;
        out[OG_MLP_B1 + t] = db1;
    }
    if (t == 0)
    {
        float s = 0.0f, l = 0.0f;
        for (uint32_t u = 0; u < TB; u++)
        {
            s += sg3[u];
            l += sloss[u];
        }
        out[OG_MLP_B3] = s;
        out[OG_MLP_NP] = l;
    }
}

/* grad[p] = (sum over blocks in order of part[b][p]) / count; loss sum. */
__global__ __launch_bounds__(LB) void og_mlp_grad_reduce_kernel(const float *__restrict__ part, uint32_t nb,
                                                                uint32_t count, float *__restrict__ grad,
                                                                double *__restrict__ loss_sum)
{
    const uint32_t p = blockIdx.x * LB + threadIdx.x;
    if (p > OG_MLP_NP) return;
    double s = 0.0;
    for (uint32_t b = 0; b < nb; b++) s += part[(size_t)b * PART_STRIDE + p];
    if (p < OG_MLP_NP) grad[p] = (float)(s / (double)count);
    else *loss_sum = s;
}

__global__ void og_mlp_adam_kernel(float *__restrict__ P, const float *__restrict__ g, float *__restrict__ m1,
                                   float *__restrict__ m2, float lr, float b1, float b2, float eps, float c1,
                                   float c2)
{
    const uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= OG_MLP_NP) return;
    const float gp = g[p];
    const float a = b1 * m1[p] + (1.0f - b1) * gp;
    const float b = b2 * m2[p] + (1.0f - b2) * gp * gp;
    m1[p] = a;
    m2[p] = b;
    P[p] -= lr * (a / c1) / (sqrtf(b / c2) + eps);
}

/* ------------------------------------------------------------------ */
/* start vector                                                        */
/* ------------------------------------------------------------------ */

#define WS_BLOCKS 1024
#define WS_SUMS 5
static_assert(WS_SUMS * WS_BLOCKS + 1 == OG_WARM_OK && OG_WARM_OK < OG_WARM_SCRATCH, "warm scratch layout");

/* Block partials of: sum of o over vertices with in-edges (U), the same
 * over those of them without out-edges, the counts of vertices without
 * in-edges (K) and of those without out-edges either, and the count of
 * vertices of U whose prediction is not a finite non-negative number. */
__global__ __launch_bounds__(LB) void og_warm_sums_kernel(const float *__restrict__ o,
                                                         const uint32_t *__restrict__ row_ptr,
                                                         const uint32_t *__restrict__ outdeg, uint32_t n,
                                                         double *__restrict__ part)
{
    __shared__ double sh[WS_SUMS][LB];
    double a[WS_SUMS] = {0.0, 0.0, 0.0, 0.0, 0.0};
    for (uint32_t v = blockIdx.x * LB + threadIdx.x; v < n; v += gridDim.x * LB)
    {
        const bool in = row_ptr[v + 1] != row_ptr[v], dang = outdeg[v] == 0;
        if (in)
        {
            const float x = o[v];
            if (!(x >= 0.0f && x <= 3.402823466e38f)) a[4] += 1.0;   /* NaN, negative or inf */
            else
            {
                a[0] += x;
                if (dang) a[1] += x;
            }
        }
        else
        {
            a[2] += 1.0;
            if (dang) a[3] += 1.0;
        }
    }
    for (int q = 0; q < WS_SUMS; q++) sh[q][threadIdx.x] = a[q];
    __syncthreads();
    for (uint32_t w = LB / 2; w > 0; w >>= 1)
    {
        if (threadIdx.x < w)
            for (int q = 0; q < WS_SUMS; q++) sh[q][threadIdx.x] += sh[q][threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x < WS_SUMS) part[WS_SUMS * blockIdx.x + threadIdx.x] = sh[threadIdx.x][0];
}

/* Mass of the rows with in-edges at the fixed point.  Vertices without
 * in-edges have the fixed-point rank (1-d)/N exactly (normalized), so with
 * their share f_K and the share f_Kd of those also without out-edges, the
 * mass identity S = 0.15 + 0.85 (S - D) gives for the rest
 *   S_U = 0.15 (1 - 0.15 f_K - 0.85 f_Kd) / (0.15 + 0.85 delta_U),
 * delta_U = share of the predicted shape of U on vertices without out-edges.
 * The prediction is accepted only when every value on U is finite and
 * non-negative and their sum is positive (a start with the fixed point's
 * mass exists); otherwise part[OG_WARM_OK] = 0 and the start
 * written is the uniform one. */
__global__ __launch_bounds__(LB) void og_warm_factor_kernel(double *__restrict__ part, uint32_t n)
{
    __shared__ double sh[WS_SUMS][LB];
    double a[WS_SUMS] = {0.0, 0.0, 0.0, 0.0, 0.0};
    for (uint32_t i = threadIdx.x; i < WS_BLOCKS; i += LB)
        for (int q = 0; q < WS_SUMS; q++) a[q] += part[WS_SUMS * i + q];
    for (int q = 0; q < WS_SUMS; q++) sh[q][threadIdx.x] = a[q];
    __syncthreads();
    for (uint32_t w = LB / 2; w > 0; w >>= 1)
    {
        if (threadIdx.x < w)
            for (int q = 0; q < WS_SUMS; q++) sh[q][threadIdx.x] += sh[q][threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0)
    {
        const double sumU = sh[0][0];
        const double nU = (double)n - sh[2][0];
        bool ok = sh[4][0] == 0.0 && (nU == 0.0 || (sumU > 0.0 && sumU <= 1.7976931348623157e308));
        double fac = 0.0;
        if (ok && nU > 0.0)
        {
            const double deltaU = sh[1][0] / sumU;
            const double fK = sh[2][0] / (double)n, fKd = sh[3][0] / (double)n;
            const double SU = 0.15 * (1.0 - 0.15 * fK - 0.85 * fKd) / (0.15 + 0.85 * deltaU);
            fac = (double)n * SU / sumU;                 /* r0 = factor * o on U */
            ok = fac > 0.0 && fac <= 3.402823466e38;
        }
        part[WS_SUMS * WS_BLOCKS] = fac;
        part[OG_WARM_OK] = ok ? 1.0 : 0.0;
    }
}

/* The start vector in old ids: the scaled prediction, or, if it was
 * rejected, the uniform start (1.0 unnormalized = 1/N normalized, as
 * init_mode 1 of the pull). */
__global__ void og_warm_write_kernel(const float *__restrict__ o, const uint32_t *__restrict__ row_ptr,
                                     const uint32_t *__restrict__ new2old, uint32_t n,
                                     const double *__restrict__ part, float *__restrict__ r0_old)
{
    const float fac = (float)part[WS_SUMS * WS_BLOCKS];
    const bool ok = part[OG_WARM_OK] != 0.0;
    for (uint32_t v = blockIdx.x * blockDim.x + threadIdx.x; v < n; v += gridDim.x * blockDim.x)
    {
        /* unnormalized: N * (1 - d) / N for vertices without in-edges */
        const float r = !ok ? 1.0f : row_ptr[v + 1] != row_ptr[v] ? fac * o[v] : OG_BASE_PR;
        r0_old[new2old ? new2old[v] : v] = r;
    }
}

/* ------------------------------------------------------------------ */
/* coarse correction of the start (in-degree classes)                  */
/* ------------------------------------------------------------------ */

__device__ __forceinline__ uint32_t og_wc_class(uint32_t indeg, uint32_t outdeg, uint32_t nin)
{
    const uint32_t o = outdeg ? 1u : 0u;
    if (!indeg) return o;
    const uint32_t l = 31u - (uint32_t)__builtin_clz(indeg);
    return 2u * (1u + min(l, nin - 2u)) + o;
}

__device__ __forceinline__ unsigned long long og_q32(float x)
{
    return (unsigned long long)((double)x * 4294967296.0 + 0.5);
}

/* Per vertex: its class and c0 = r0 / out-degree; per class the mass of r0
 * and the vertex count, in Q32.32 and integers (block sums in LDS, then one
 * global add per class and block). */
__global__ __launch_bounds__(LB) void og_wc_sums_kernel(const float *__restrict__ r0,
                                                       const uint32_t *__restrict__ indeg,
                                                       const uint32_t *__restrict__ outdeg, uint32_t n, uint32_t nin,
                                                       uint8_t *__restrict__ cls, float *__restrict__ c0,
                                                       unsigned long long *__restrict__ mass,
                                                       unsigned long long *__restrict__ count)
{
    __shared__ unsigned long long sm[OG_WC_MAX], sn[OG_WC_MAX];
    if (threadIdx.x < OG_WC_MAX)
    {
        sm[threadIdx.x] = 0ull;
        sn[threadIdx.x] = 0ull;
    }
    __syncthreads();
    for (uint32_t v = blockIdx.x * LB + threadIdx.x; v < n; v += gridDim.x * LB)
    {
        const uint32_t d = outdeg[v];
        const uint32_t c = og_wc_class(indeg[v], d, nin);
        const float x = r0[v];
        cls[v] = (uint8_t)c;
        c0[v] = d ? x / (float)d : 0.0f;
        atomicAdd(&sm[c], og_q32(x));
        atomicAdd(&sn[c], 1ull);
    }
    __syncthreads();
    if (threadIdx.x < 2 * nin)
    {
        if (sm[threadIdx.x]) atomicAdd(&mass[threadIdx.x], sm[threadIdx.x]);
        if (sn[threadIdx.x]) atomicAdd(&count[threadIdx.x], sn[threadIdx.x]);
    }
}

/* mat[class(v) * nc + class(u)] += c0[u] over the forward edges (u, v),
 * in Q32.32: block sums in LDS, then one global add per entry and block.
 * Each thread takes WC_U edges LB apart per step, all their loads (the
 * edge pair, then c0 and both classes) issued before the adds. */
#define WC_U 8
__global__ __launch_bounds__(LB) void og_wc_matrix_kernel(const uint32_t *__restrict__ src,
                                                         const uint32_t *__restrict__ dst, uint32_t E,
                                                         const uint8_t *__restrict__ cls,
                                                         const float *__restrict__ c0, uint32_t nc,
                                                         unsigned long long *__restrict__ mat)
{
    __shared__ unsigned long long sh[OG_WC_MAX * OG_WC_MAX];
    for (uint32_t k = threadIdx.x; k < OG_WC_MAX * OG_WC_MAX; k += LB) sh[k] = 0ull;
    __syncthreads();
    const uint64_t step = (uint64_t)gridDim.x * LB * WC_U;
    for (uint64_t base = (uint64_t)blockIdx.x * LB * WC_U + threadIdx.x; base < E; base += step)
    {
        uint32_t u[WC_U], v[WC_U];
#pragma unroll
        for (uint32_t k = 0; k < WC_U; k++)
        {
            const uint64_t e = min(base + (uint64_t)k * LB, (uint64_t)E - 1);   /* past the end: re-read, not added */
            u[k] = __builtin_nontemporal_load(&src[e]);
            v[k] = __builtin_nontemporal_load(&dst[e]);
        }
        float x[WC_U];
        uint32_t cu[WC_U], cv[WC_U];
#pragma unroll
        for (uint32_t k = 0; k < WC_U; k++)
        {
            x[k] = c0[u[k]];
            cu[k] = cls[u[k]];
            cv[k] = cls[v[k]];
        }
#pragma unroll
        for (uint32_t k = 0; k < WC_U; k++)
            if (base + (uint64_t)k * LB < E && x[k] > 0.0f) atomicAdd(&sh[cv[k] * nc + cu[k]], og_q32(x[k]));
    }
    __syncthreads();
    for (uint32_t k = threadIdx.x; k < nc * nc; k += LB)
        if (sh[k]) atomicAdd(&mat[k], sh[k]);
}

__global__ void og_wc_apply_kernel(float *__restrict__ r0, const uint8_t *__restrict__ cls, uint32_t n,
                                   const float *__restrict__ scale)
{
    for (uint32_t v = blockIdx.x * blockDim.x + threadIdx.x; v < n; v += gridDim.x * blockDim.x)
        r0[v] *= scale[cls[v]];
}

/* ------------------------------------------------------------------ */
/* launchers                                                           */
/* ------------------------------------------------------------------ */

static unsigned og_lgrid(uint64_t n)
{
    uint64_t g = (n + LB - 1) / LB;
    if (g < 1) g = 1;
    if (g > 16384) g = 16384;
    return (unsigned)g;
}

extern "C" int og_launch_wc_sums(const float *r0, const uint32_t *in_degree, const uint32_t *out_degree,
                                 uint32_t n, uint32_t nin, uint8_t *cls, float *c0, double *scratch, hipStream_t s)
{
    if (nin < 2 || nin > OG_WC_IN_MAX)
    {
        og_set_error("coarse start correction: %u in-degree classes (2..%u)", nin, OG_WC_IN_MAX);
        return -1;
    }
    unsigned long long *mass = (unsigned long long *)(scratch + og_wc_mass());
    unsigned long long *count = (unsigned long long *)(scratch + og_wc_count());
    if (og_launch_fill_u32((uint32_t *)mass, 0, 4 * OG_WC_MAX, s)) return -1;   /* mass and count */
    if (n)
        OG_LAUNCH(og_wc_sums_kernel, dim3(og_lgrid(n) < 1024 ? og_lgrid(n) : 1024), dim3(LB), 0, s, r0,
                  in_degree, out_degree, n, nin, cls, c0, mass, count);
    return og_hip_check(og_launch_status(), "og_wc_sums_kernel");
}

extern "C" int og_launch_wc_matrix(const uint32_t *src, const uint32_t *dst, uint32_t E, const uint8_t *cls,
                                   const float *c0, uint32_t nc, unsigned long long *mat, hipStream_t s)
{
    if (og_launch_fill_u32((uint32_t *)mat, 0, 2ull * nc * nc, s)) return -1;
    if (!E) return 0;
    const uint64_t steps = ((uint64_t)E + LB * WC_U - 1) / (LB * WC_U);
    OG_LAUNCH(og_wc_matrix_kernel, dim3(steps < 2048 ? (unsigned)steps : 2048u), dim3(LB), 0, s, src,
                       dst, E, cls, c0, nc, mat);
    return og_hip_check(og_launch_status(), "og_wc_matrix_kernel");
}

extern "C" int og_launch_wc_apply(float *r0, const uint8_t *cls, uint32_t n, const float *scale, hipStream_t s)
{
    OG_LAUNCH(og_wc_apply_kernel, dim3(og_lgrid(n)), dim3(LB), 0, s, r0, cls, n, scale);
    return og_hip_check(og_launch_status(), "og_wc_apply_kernel");
}

extern "C" int og_launch_feat_li(const uint32_t *row_ptr, uint32_t n, float *li, hipStream_t s)
{
    OG_LAUNCH(og_feat_li_kernel, dim3(og_lgrid(n)), dim3(LB), 0, s, row_ptr, n, li);
    return og_hip_check(og_launch_status(), "og_feat_li_kernel");
}

extern "C" int og_launch_feat_finish(const uint32_t *row_ptr, const uint32_t *outdeg, const float *insum,
                                     uint32_t n, float *feat, hipStream_t s)
{
    OG_LAUNCH(og_feat_finish_kernel, dim3(og_lgrid(n)), dim3(LB), 0, s, row_ptr, outdeg, insum, n,
                       (float4 *)feat);
    return og_hip_check(og_launch_status(), "og_feat_finish_kernel");
}

extern "C" int og_launch_mlp_infer(const float *params, const float *feat, uint32_t n, int mfma, float *out,
                                   hipStream_t s)
{
    if (!n) return 0;
    if (mfma)
    {
        uint64_t tiles = ((uint64_t)n + 31) / 32;
        unsigned g = (unsigned)((tiles + 3) / 4 < 8192 ? (tiles + 3) / 4 : 8192);
        OG_LAUNCH(og_mlp_infer_mfma_kernel, dim3(g), dim3(LB), 0, s, params, (const float4 *)feat, n, out);
    }
    else
        OG_LAUNCH(og_mlp_infer_f32_kernel, dim3(og_lgrid(n) < 4096 ? og_lgrid(n) : 4096), dim3(LB), 0,
                           s, params, (const float4 *)feat, n, out);
    return og_hip_check(og_launch_status(), "og_mlp_infer_kernel");
}

extern "C" size_t og_mlp_part_floats(uint32_t count)
{
    return ((size_t)(count + TB - 1) / TB) * PART_STRIDE;
}

extern "C" int og_launch_mlp_grad(const float *params, const float *feat, const float *y, uint32_t n_data,
                                  uint64_t seed, uint32_t base, uint32_t count, float *part, float *grad,
                                  double *loss_sum, hipStream_t s)
{
    if (!count) return 0;
    og_perm_t perm;
    memset(&perm, 0, sizeof(perm));
    if (seed)
    {
        uint32_t bits = 0;
        while (bits < 32 && ((uint64_t)1 << bits) < n_data) bits++;
        perm = og_perm_make(seed, bits);
    }
    const uint32_t nb = (count + TB - 1) / TB;
    OG_LAUNCH(og_mlp_grad_kernel, dim3(nb), dim3(TB), 0, s, params, (const float4 *)feat, y, n_data,
                       perm, base, count, part);
    OG_LAUNCH(og_mlp_grad_reduce_kernel, dim3((OG_MLP_NP + LB) / LB), dim3(LB), 0, s, part, nb, count,
                       grad, loss_sum);
    return og_hip_check(og_launch_status(), "og_mlp_grad_kernel");
}

/* ------------------------------------------------------------------ */
/* gradient partials without an allocation                             */
/* ------------------------------------------------------------------ */

/* The block partials of ogpu_mlp_grad_noalloc: enough for
 * OGPU_MLP_GRAD_NOALLOC_MAX vertices (4.4 MiB), in a static device array
 * that each GPU the library runs on holds until the process ends.  Host
 * code never passes `scratch` itself to a kernel (on the host it names a
 * stand-in, not device memory): og_learn_noalloc_setup looks up its device
 * address once per device, with that device current, and the launches use
 * that address. */
#define OG_NOALLOC_PART_FLOATS \
    ((((size_t)OGPU_MLP_GRAD_NOALLOC_MAX + TB - 1) / TB) * PART_STRIDE)
static __device__ __align__(256) float scratch[OG_NOALLOC_PART_FLOATS];

#define OG_NOALLOC_DEVICES 64
static float *g_scratch_ptr[OG_NOALLOC_DEVICES];

extern "C" int og_learn_noalloc_setup(int device)
{
    if (device < 0 || device >= OG_NOALLOC_DEVICES)
    {
        og_set_error("gradient scratch: device %d out of range", device);
        return -1;
    }
    if (g_scratch_ptr[device]) return 0;
    void *p = NULL;
    if (og_hip_check(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)), "hipGetSymbolAddress(scratch)")) return -1;
    g_scratch_ptr[device] = (float *)p;
    return 0;
}

extern "C" int og_launch_mlp_grad_noalloc(int device, const float *params, const float *feat, const float *y,
                                          uint32_t n, float *grad, double *loss_sum, hipStream_t s)
{
    if (n > OGPU_MLP_GRAD_NOALLOC_MAX)
    {
        og_set_error("ogpu_mlp_grad_noalloc: %u vertices, at most %u per call", n, OGPU_MLP_GRAD_NOALLOC_MAX);
        return -1;
    }
    float *part = device >= 0 && device < OG_NOALLOC_DEVICES ? g_scratch_ptr[device] : NULL;
    if (!part)
    {
        og_set_error("ogpu_mlp_grad_noalloc: no gradient scratch on device %d (library not initialized)", device);
        return -1;
    }
    /* every partial the reduction reads is written by this call's gradient
     * kernel first: nothing needs clearing */
    return og_launch_mlp_grad(params, feat, y, n, 0, 0, n, part, grad, loss_sum, s);
}

extern "C" int og_launch_mlp_adam(float *params, const float *grad, float *m1, float *m2, float lr, float b1,
                                  float b2, float eps, uint32_t step, hipStream_t s)
{
    const float c1 = 1.0f - powf(b1, (float)step), c2 = 1.0f - powf(b2, (float)step);
    OG_LAUNCH(og_mlp_adam_kernel, dim3((OG_MLP_NP + LB - 1) / LB), dim3(LB), 0, s, params, grad, m1, m2,
                       lr, b1, b2, eps, c1, c2);
    return og_hip_check(og_launch_status(), "og_mlp_adam_kernel");
}

extern "C" int og_launch_warm_scale(const float *o, const uint32_t *row_ptr, const uint32_t *outdeg,
                                    const uint32_t *new2old, uint32_t n, double *sums, float *r0_old,
                                    hipStream_t s)
{
    OG_LAUNCH(og_warm_sums_kernel, dim3(WS_BLOCKS), dim3(LB), 0, s, o, row_ptr, outdeg, n, sums);
    OG_LAUNCH(og_warm_factor_kernel, dim3(1), dim3(LB), 0, s, sums, n);
    OG_LAUNCH(og_warm_write_kernel, dim3(og_lgrid(n)), dim3(LB), 0, s, o, row_ptr, new2old, n, sums,
                       r0_old);
    return og_hip_check(og_launch_status(), "og_warm_scale");
}
