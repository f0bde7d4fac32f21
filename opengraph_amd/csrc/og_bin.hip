/*
 * og_bin.hip -- gfx950 kernels of the binned fp32 pull (layout and passes
 * described in og_bin.h) and of its one-time layout build.
 *
 * Exactness.  The reference adds a row's in-edges one after the other into
 * an fp32 running sum (pageRank.c:998-1011).  Here every row is added in
 * that order from LDS: one thread per short row, one wavefront per longer
 * row and one block per row longer than a bin.  Wavefronts and blocks use an
 * exact fast-forward: while the running sum y stays in one binade
 * [2^(e-1), 2^e), every representable value is a multiple of u = 2^(e-24),
 * so for a term a >= 0 the rounded sum is fl(y + a) = y + u * m with m the
 * integer nearest a / u, unless a / u ends in exactly one half (then the
 * tie goes to the even neighbour, which depends on y).  A run of terms is
 * therefore y + u * (prefix sum of their m), an integer scan, up to the
 * first term that is a tie, negative or not finite, or whose sum reaches
 * 2^e; that term is added as an ordinary fp32 add, and the scan restarts
 * after it in the new binade.  The result equals the sequential fp32 sum
 * bit for bit.
 */
#include <hip/hip_runtime.h>

#include "og_bin.h"
#include "og_internal.h"
#include "og_pr.h"

#define BT OG_BIN_THREADS
#define BW (OG_BIN_THREADS / 64)
#define GT OG_BIN_GTHREADS       /* the gather pass's blocks */
#define LT OG_BIN_LTHREADS       /* the long-row pass's blocks */
#define LW (OG_BIN_LTHREADS / 64)
#define GW (OG_BIN_GTHREADS / 64)
#define GATHER_Q 8u   /* runs of 4 consecutive sources per lane in flight in the gather */
#define OG_BIN_HEAD 128u  /* leading terms of a long sum added one by one */

/* ------------------------------------------------------------------ */
/* shared pieces                                                       */
/* ------------------------------------------------------------------ */

__device__ __forceinline__ float og_bin_uniform(float x)
{
    return __int_as_float(__builtin_amdgcn_readfirstlane(__float_as_int(x)));
}

__device__ __forceinline__ float og_bin_lane(float x, uint32_t k)
{
    return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(x), k));
}

/* Inclusive prefix sum over the wavefront with DPP lane moves: shifts by
 * 1, 2, 4 and 8 within each row of 16 lanes (lanes shifted in from outside
 * the row read 0), then row 1 and row 3 add the last lane of the row below
 * (row_bcast:15), then rows 2 and 3 add lane 31 (row_bcast:31).  Every
 * disabled lane adds 0. */
__device__ __forceinline__ uint32_t og_bin_wave_scan(uint32_t v)
{
    v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x111, 0xF, 0xF, true);   /* row_shr:1 */
    v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x112, 0xF, 0xF, true);   /* row_shr:2 */
    v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x114, 0xF, 0xF, true);   /* row_shr:4 */
    v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x118, 0xF, 0xF, true);   /* row_shr:8 */
    v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x142, 0xA, 0xF, false);  /* row_bcast:15 */
    v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x143, 0xC, 0xF, false);  /* row_bcast:31 */
    return v;
}

/* The fast-forward quantities of term a for a running sum in the binade
 * whose unit is 2^-sc: m (valid when !bad) and bad (tie, negative, not
 * finite, or a term of at least the binade's top).  Branch-free: every
 * lane computes, the result is selected. */
__device__ __forceinline__ void og_bin_term(float a, int sc, bool valid, uint32_t &m, bool &bad)
{
    const float t = ldexpf(a, sc);   /* a / u, exact for the ranks' range */
    const float fl = floorf(t);
    const float fr = t - fl;         /* exact */
    const bool in = (t >= 0.0f) & (t < 0x1p24f);   /* false for NaN */
    bad = valid & (!in | (fr == 0.5f));
    const uint32_t mm = (uint32_t)fl + (fr > 0.5f ? 1u : 0u);
    m = (valid & !bad) ? mm : 0u;
}

/* One thread's sequential fp32 sum of slot[j, j1): the 16-byte groups
 * wholly inside the range are read 16 values at a time into two register
 * sets used in turn, each refilled right after it is added, so the loads
 * of one are in flight while the other is added. */
__device__ __forceinline__ void og_bin_add16(float &s, const float4 (&x)[4], uint32_t groups)
{
#pragma unroll
    for (uint32_t u = 0; u < 4; u++)
        if (u < groups)
        {
            s = s + x[u].x;
            s = s + x[u].y;
            s = s + x[u].z;
            s = s + x[u].w;
        }
}

__device__ __forceinline__ void og_bin_load16(const float4 *s4, uint32_t q, float4 (&x)[4])
{
    constexpr uint32_t QMAX = OG_BIN_CAP / 4 - 1;   /* last float4 of slot */
#pragma unroll
    for (uint32_t u = 0; u < 4; u++) x[u] = s4[min(q + u, QMAX)];
}

__device__ __forceinline__ float og_bin_row_sum(const float *slot, uint32_t j, uint32_t j1, float s0)
{
    float s = s0;
    const uint32_t qe = j1 >> 2;                     /* groups [q, qe) lie inside */
    uint32_t q = (j + 3) >> 2;
    if (q >= qe)
    {
        for (; j < j1; j++) s = s + slot[j];
        return s;
    }
    for (; j < 4 * q; j++) s = s + slot[j];          /* head up to a group boundary */
    const float4 *s4 = reinterpret_cast<const float4 *>(slot);
    float4 A[4], B[4];
    og_bin_load16(s4, q, A);
    og_bin_load16(s4, q + 4, B);
    for (; q + 8 <= qe; q += 8)
    {
        /* scheduling barriers keep each refill after the adds of its set
         * and before the adds of the other */
        og_bin_add16(s, A, 4);
        __builtin_amdgcn_sched_barrier(0);
        og_bin_load16(s4, q + 8, A);
        __builtin_amdgcn_sched_barrier(0);
        og_bin_add16(s, B, 4);
        __builtin_amdgcn_sched_barrier(0);
        og_bin_load16(s4, q + 12, B);
        __builtin_amdgcn_sched_barrier(0);
    }
    const uint32_t g = qe - q;                       /* 0..7 groups left, in A then B */
    og_bin_add16(s, A, g < 4 ? g : 4);
    og_bin_add16(s, B, g > 4 ? g - 4 : 0);
    for (j = 4 * qe; j < j1; j++) s = s + slot[j];   /* tail */
    return s;
}

/* x[i + 4 * lane + k] for k < 4 where inside [0, n), else 0 (n > 0;
 * branch-free: the index is clamped, the value selected) */
__device__ __forceinline__ void og_bin_wave_load4(const float *x, uint32_t i, uint32_t n, uint32_t lane,
                                                  float (&a)[4])
{
#pragma unroll
    for (uint32_t k = 0; k < 4; k++)
    {
        const uint32_t p = i + 4 * lane + k;
        const float v = x[min(p, n - 1)];
        a[k] = p < n ? v : 0.0f;
    }
}

/* The sequential fp32 sum of slot[j0, j0 + n) (LDS, n > 0), one
 * wavefront, result uniform: the first OG_BIN_HEAD terms one after the
 * other (where the running sum crosses binades every few terms), then 256
 * terms per step, 4 consecutive ones per lane (a lane's prefix of its own
 * terms, then a wavefront scan of the lanes' totals); the next 256 are
 * read while a step is scanned. */
__device__ float og_bin_wave_sum(const float *slot, uint32_t j0, uint32_t n, float y0)
{
    const uint32_t lane = threadIdx.x & 63;
    const uint32_t h = n < OG_BIN_HEAD ? n : OG_BIN_HEAD;
    float y = og_bin_uniform(og_bin_row_sum(slot, j0, j0 + h, y0));
    const float *x = slot + j0;
    uint32_t i = h;
    float a[4];
    og_bin_wave_load4(x, i, n, lane, a);
    while (i < n)
    {
        const uint32_t left = n - i;
        if (!(y >= 0x1p-100f && y < 0x1p100f))
        {
            /* zero, tiny or not finite: one ordinary add */
            y = og_bin_uniform(y + x[i]);
            i += 1;
            og_bin_wave_load4(x, i, n, lane, a);
            continue;
        }
        float an[4];
        og_bin_wave_load4(x, i + 256, n, lane, an);   /* used when the whole step is taken */
        int ey;
        frexpf(y, &ey);
        const int sc = 24 - ey;
        const uint32_t Y = (uint32_t)ldexpf(y, sc);   /* in [2^23, 2^24) */
        uint32_t L[4];
        bool bad[4], valid[4];
        uint32_t tot = 0;
#pragma unroll
        for (uint32_t k = 0; k < 4; k++)
        {
            uint32_t m;
            valid[k] = 4 * lane + k < left;
            og_bin_term(a[k], sc, valid[k], m, bad[k]);
            tot = min(tot + m, 1u << 25);
            L[k] = tot;
        }
        const uint32_t incl = og_bin_wave_scan(tot);   /* <= 2^31 */
        const uint32_t X = incl - tot;                 /* the earlier lanes' terms */
        uint32_t kfail = 4;
#pragma unroll
        for (int k = 3; k >= 0; k--)
            kfail = (valid[k] & (bad[k] | (Y + X + L[k] >= (1u << 24)))) ? (uint32_t)k : kfail;
        const uint64_t fb = __ballot(kfail < 4);
        if (fb == 0)
        {
            const uint32_t total = (uint32_t)__builtin_amdgcn_readlane(incl, 63);
            y = og_bin_uniform(ldexpf((float)(Y + total), -sc));
            i += left < 256 ? left : 256;
#pragma unroll
            for (uint32_t k = 0; k < 4; k++) a[k] = an[k];
        }
        else
        {
            /* the first failing term: the prefix before it, then an
             * ordinary add of it */
            const uint32_t lf = (uint32_t)__builtin_ctzll(fb);
            uint32_t pre = X;
            float af = a[0];
#pragma unroll
            for (uint32_t k = 1; k < 4; k++)
                if (kfail == k)
                {
                    pre = X + L[k - 1];
                    af = a[k];
                }
            const uint32_t kf = (uint32_t)__builtin_amdgcn_readlane(kfail, lf);
            const uint32_t p = (uint32_t)__builtin_amdgcn_readlane(pre, lf);
            y = ldexpf((float)(Y + p), -sc);
            y = og_bin_uniform(y + og_bin_lane(af, lf));
            i += 4 * lf + kf + 1;
            og_bin_wave_load4(x, i, n, lane, a);
        }
    }
    return y;
}

/* ------------------------------------------------------------------ */
/* gather pass                                                         */
/* ------------------------------------------------------------------ */

typedef uint32_t og_u32x2 __attribute__((ext_vector_type(2)));
typedef float og_f32x4g __attribute__((ext_vector_type(4)));

/* The sources (4 u16 in one 8-byte load) and quad entry of this lane's
 * GATHER_Q runs of 4 sources from f (256 apart; f a multiple of 4, so a run
 * lies in one quad), branch-free (indices clamped to the item's last run)
 * so that all the loads issue before any is waited for. */
__device__ __forceinline__ void og_bin_gather_batch(const og_bin_dev &b, uint32_t f, uint32_t f1,
                                                    og_u32x2 (&sv)[GATHER_Q], uint32_t (&qv)[GATHER_Q])
{
#pragma unroll
    for (uint32_t u = 0; u < GATHER_Q; u++)
    {
        const uint32_t fu = min(f + 256u * u, f1 - 4);
        sv[u] = __builtin_nontemporal_load((const og_u32x2 *)(b.src + fu));
        qv[u] = __builtin_nontemporal_load(&b.qent[fu / OG_BIN_TILE_ALIGN]);
    }
}

__global__ __launch_bounds__(GT) void og_bin_gather_kernel(og_bin_dev b, const float *__restrict__ c,
                                                           const uint32_t *done)
{
    __shared__ __align__(16) float seg[OG_BIN_SEG];
    if (*done) return;
    if (blockIdx.x == 0 && threadIdx.x == 0) *b.ticket = 0u;   /* the long-row pass's row counter */
    const uint32_t lane = threadIdx.x & 63;
    const uint32_t wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    for (uint32_t q = blockIdx.x; q < b.n_items; q += gridDim.x)
    {
        const og_bin_item it = b.items[q];
        const uint32_t v0 = it.seg * OG_BIN_SEG;
        const uint32_t nv = min(OG_BIN_SEG, b.N - v0);
        __syncthreads();   /* the previous item's reads of seg are done */
        {
            const float4 *c4 = (const float4 *)(c + v0);
            float4 *s4 = (float4 *)seg;
            for (uint32_t i = threadIdx.x; i < nv / 4; i += GT) s4[i] = c4[i];
            for (uint32_t i = (nv & ~3u) + threadIdx.x; i < nv; i += GT) seg[i] = c[v0 + i];
        }
        __syncthreads();
        /* wavefront w takes the runs of 4 sources f = f0 + 256 * (w * Q +
         * k * GW * Q + u) + 4 * lane (f0 and f1 are multiples of 16, so a
         * run lies in one quad and is wholly inside or past the item); each
         * run's 4 values go to 4 consecutive entries, one 16-byte store
         * (a quad's entry is a multiple of 16).  Software pipelined: the
         * next batch's loads are issued before this batch's stores (loads
         * complete in order, and a wait for this batch's loads must not
         * also wait for stores issued after them) */
        uint32_t f = it.f0 + 256u * GATHER_Q * wave + 4u * lane;
        og_u32x2 sv[GATHER_Q];
        uint32_t qv[GATHER_Q];
        og_bin_gather_batch(b, f, it.f1, sv, qv);
        for (; f < it.f1; f += 256u * GATHER_Q * GW)
        {
            og_u32x2 sv2[GATHER_Q];
            uint32_t qv2[GATHER_Q];
            og_bin_gather_batch(b, f + 256u * GATHER_Q * GW, it.f1, sv2, qv2);
            og_f32x4g x[GATHER_Q];   /* every lane's sources are valid ones (clamped) */
#pragma unroll
            for (uint32_t u = 0; u < GATHER_Q; u++)
            {
                x[u].x = seg[sv[u].x & 0xFFFFu];
                x[u].y = seg[sv[u].x >> 16];
                x[u].z = seg[sv[u].y & 0xFFFFu];
                x[u].w = seg[sv[u].y >> 16];
            }
            /* lanes past the item's end write the OG_BIN_SPARE spare
             * entries past the last tile instead: unconditional stores keep
             * the count of memory operations in flight static */
#pragma unroll
            for (uint32_t u = 0; u < GATHER_Q; u++)
            {
                const uint32_t fu = f + 256u * u;
                const uint32_t e = fu < it.f1 ? qv[u] + fu % OG_BIN_TILE_ALIGN : b.n_ent + 4u * lane;
                __builtin_nontemporal_store(x[u], (og_f32x4g *)(b.val + e));
            }
#pragma unroll
            for (uint32_t u = 0; u < GATHER_Q; u++)
            {
                sv[u] = sv2[u];
                qv[u] = qv2[u];
            }
        }
    }
}

/* ------------------------------------------------------------------ */
/* sum pass                                                            */
/* ------------------------------------------------------------------ */

/* running statistics of the updates a thread made */
struct og_bin_stat
{
    double err;
    uint32_t active;
};

__device__ __forceinline__ void og_bin_stat_init(og_bin_stat &s)
{
    s.err = 0.0;
    s.active = 0;
}

/* pageRank.c:1016-1028 and the next contribution (pageRank.c:989-996),
 * with the old rank and out-degree already loaded. */
__device__ __forceinline__ void og_bin_epilogue_pf(const og_pr_dev &a, uint32_t v, float sum, float prev,
                                                   uint32_t deg, og_bin_stat &st)
{
    if (a.spmv)
    {
        /* SpMV pass: the running y, started from the old one, is the sum */
        __builtin_nontemporal_store(sum, &a.r[v]);
        return;
    }
    const float nr = OG_BASE_PR + (OG_DAMP * sum);
    __builtin_nontemporal_store(nr, &a.r[v]);
    const double err = fabs((double)(nr - prev));
    st.err += err / (double)a.N;
    st.active += err >= a.epsilon;
    __builtin_nontemporal_store(deg ? nr / (float)deg : 0.0f, &((float *)a.c_out)[v]);
}

__device__ __forceinline__ void og_bin_epilogue(const og_pr_dev &a, uint32_t v, float sum, og_bin_stat &st)
{
    const float prev = __builtin_nontemporal_load(&a.r[v]);
    const uint32_t deg = __builtin_nontemporal_load(&a.outdeg[v]);
    og_bin_epilogue_pf(a, v, sum, prev, deg, st);
}

/* What a bin's rows need besides their values -- slot range (row
 * offsets), old rank, out-degree -- for the OG_BIN_RPT rows a thread holds:
 * row og_bin_meta_row(d, k) for slot k.  Thread rows: thread t holds rows
 * t, t + BT, ...; wavefront rows: lane l of wavefront w holds the
 * wavefront's rows l, l + 64, ... (rows w + BW * l + BT * k).  Loaded with
 * the bin's entries, one bin ahead, so adding a bin's rows issues no
 * global load (a load there would wait for the next bin's entries too:
 * loads complete in order). */
struct og_bin_meta
{
    uint32_t p0[OG_BIN_RPT], p1[OG_BIN_RPT], deg[OG_BIN_RPT];
    float prev[OG_BIN_RPT];
};

__device__ __forceinline__ uint32_t og_bin_meta_row(const og_bin_desc &d, uint32_t k)
{
    const uint32_t t = d.kind == OG_BIN_WAVE_ROWS ? (threadIdx.x >> 6) + BW * (threadIdx.x & 63) : threadIdx.x;
    return d.row0 + t + BT * k;
}

__device__ __forceinline__ void og_bin_meta_load(const og_pr_dev &a, const og_bin_desc &d, og_bin_meta &m)
{
    if (d.kind == OG_BIN_PIECE) return;
#pragma unroll
    for (uint32_t k = 0; k < OG_BIN_RPT; k++)
    {
        const uint32_t v = og_bin_meta_row(d, k);
        if (v < d.row1)
        {
            m.p0[k] = a.row_ptr[v];
            m.p1[k] = a.row_ptr[v + 1];
            m.prev[k] = __builtin_nontemporal_load(&a.r[v]);
            m.deg[k] = __builtin_nontemporal_load(&a.outdeg[v]);
        }
    }
}

/* og_stat over the block of NT threads (fixed order); result in thread 0 */
template <uint32_t NT>
__device__ og_bin_stat og_bin_block_stat(og_bin_stat s)
{
    constexpr uint32_t W = NT / 64;
    __shared__ double sh_err[W];
    __shared__ uint32_t sh_act[W];
    const uint32_t lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    for (int o = 32; o > 0; o >>= 1)
    {
        s.err += __shfl_down(s.err, o, 64);
        s.active += __shfl_down(s.active, o, 64);
    }
    if (lane == 0)
    {
        sh_err[w] = s.err;
        sh_act[w] = s.active;
    }
    __syncthreads();
    og_bin_stat r;
    og_bin_stat_init(r);
    if (threadIdx.x == 0)
        for (uint32_t i = 0; i < W; i++)
        {
            r.err += sh_err[i];
            r.active += sh_act[i];
        }
    __syncthreads();
    return r;
}

/* first-level partial idx of the final reduction (min / max unused: the
 * neutral values) */
__device__ __forceinline__ void og_bin_write_part(const og_pr_dev &a, uint32_t idx, const og_bin_stat &r)
{
    if (threadIdx.x == 0)
    {
        a.part_err[idx] = r.err;
        a.part_act[idx] = r.active;
        a.part_min[idx] = 3.402823466e38f;
        a.part_max[idx] = -3.402823466e38f;
    }
}

/* A bin's entries, 8 per thread and sweep: OG_BIN_SWEEPS sweeps cover a
 * full bin; loaded into registers ahead of time (the next bin's while this
 * one's rows are added). */
#define OG_BIN_SWEEPS ((OG_BIN_CAP + 8 * BT - 1) / (8 * BT))

typedef uint32_t og_u32x4 __attribute__((ext_vector_type(4)));
typedef float og_f32x4 __attribute__((ext_vector_type(4)));

struct og_bin_regs
{
    og_u32x4 p[OG_BIN_SWEEPS];
    og_f32x4 x0[OG_BIN_SWEEPS], x1[OG_BIN_SWEEPS];
};

__device__ __forceinline__ void og_bin_load(const og_bin_dev &b, const og_bin_desc &d, og_bin_regs &r)
{
#pragma unroll
    for (uint32_t k = 0; k < OG_BIN_SWEEPS; k++)
    {
        const uint32_t e = d.ent0 + 8 * (threadIdx.x + k * BT);
        if (e < d.ent1)
        {
            r.p[k] = __builtin_nontemporal_load((const og_u32x4 *)(b.pos + e));
            r.x0[k] = __builtin_nontemporal_load((const og_f32x4 *)(b.val + e));
            r.x1[k] = __builtin_nontemporal_load((const og_f32x4 *)(b.val + e + 4));
        }
    }
}

/* 8 entries' values to their slots (padding entries skipped) */
__device__ __forceinline__ void og_bin_scatter8(float *slot, og_u32x4 p, og_f32x4 x0, og_f32x4 x1)
{
    const uint32_t pp[8] = {p.x & 0xFFFFu, p.x >> 16, p.y & 0xFFFFu, p.y >> 16,
                            p.z & 0xFFFFu, p.z >> 16, p.w & 0xFFFFu, p.w >> 16};
    const float xx[8] = {x0.x, x0.y, x0.z, x0.w, x1.x, x1.y, x1.z, x1.w};
#pragma unroll
    for (uint32_t t = 0; t < 8; t++)
        if (pp[t] != OG_BIN_PAD) slot[pp[t]] = xx[t];
}

/* The bin's entries: the registers' sweeps, then (a bin whose tiles'
 * padding takes it past OG_BIN_SWEEPS sweeps: up to 15 entries per tile)
 * the rest straight from memory. */
__device__ __forceinline__ void og_bin_put(const og_bin_dev &b, float *slot, const og_bin_desc &d,
                                           const og_bin_regs &r)
{
#pragma unroll
    for (uint32_t k = 0; k < OG_BIN_SWEEPS; k++)
    {
        const uint32_t e = d.ent0 + 8 * (threadIdx.x + k * BT);
        if (e < d.ent1) og_bin_scatter8(slot, r.p[k], r.x0[k], r.x1[k]);
    }
    for (uint32_t e = d.ent0 + 8 * (threadIdx.x + OG_BIN_SWEEPS * BT); e < d.ent1; e += 8 * BT)
        og_bin_scatter8(slot, __builtin_nontemporal_load((const og_u32x4 *)(b.pos + e)),
                        __builtin_nontemporal_load((const og_f32x4 *)(b.val + e)),
                        __builtin_nontemporal_load((const og_f32x4 *)(b.val + e + 4)));
}

__global__ __launch_bounds__(BT, 4) void og_bin_sum_kernel(og_bin_dev b, og_pr_dev a)
{
    __shared__ __align__(16) float slot[OG_BIN_CAP];
    if (*a.done) return;
    og_bin_stat st;
    og_bin_stat_init(st);
    const uint32_t lane = threadIdx.x & 63;
    const uint32_t wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    /* the long rows' pieces (bins [0, first_bin)) are the long-row pass's */
    uint32_t q = b.first_bin + blockIdx.x;
    og_bin_desc d;
    og_bin_regs regs;
    og_bin_meta meta;
    if (q < b.n_bins_live)
    {
        d = b.bins[q];
        og_bin_load(b, d, regs);
        og_bin_meta_load(a, d, meta);
    }
    for (; q < b.n_bins_live; q += gridDim.x)
    {
        __syncthreads();   /* the previous bin's reads of slot are done */
        /* this bin's values to their CSR slots */
        og_bin_put(b, slot, d, regs);
        const og_bin_desc cd = d;
        const og_bin_meta cm = meta;
        /* the next bin's entries and row metadata load while this one's
         * rows are added */
        if (q + gridDim.x < b.n_bins_live)
        {
            d = b.bins[q + gridDim.x];
            og_bin_load(b, d, regs);
            og_bin_meta_load(a, d, meta);
        }
        __syncthreads();
        if (cd.kind == OG_BIN_THREAD_ROWS)
        {
#pragma unroll
            for (uint32_t k = 0; k < OG_BIN_RPT; k++)
            {
                const uint32_t v = cd.row0 + threadIdx.x + BT * k;
                if (v >= cd.row1) break;
                /* (an SpMV pass continues the row's running y) */
                const float s = og_bin_row_sum(slot, cm.p0[k] - cd.pos0, cm.p1[k] - cd.pos0,
                                               a.spmv ? cm.prev[k] : 0.0f);
                og_bin_epilogue_pf(a, v, s, cm.prev[k], cm.deg[k], st);
            }
        }
        else
        {
#pragma unroll
            for (uint32_t k = 0; k < OG_BIN_RPT; k++)
            {
                if (cd.row0 + wave + BT * k >= cd.row1) break;
                for (uint32_t l = 0; l < 64; l++)
                {
                    const uint32_t v = cd.row0 + wave + BW * l + BT * k;
                    if (v >= cd.row1) break;
                    const uint32_t j0 = (uint32_t)__builtin_amdgcn_readlane(cm.p0[k], l) - cd.pos0;
                    const uint32_t j1 = (uint32_t)__builtin_amdgcn_readlane(cm.p1[k], l) - cd.pos0;
                    const float prev = og_bin_lane(cm.prev[k], l);
                    const float y = og_bin_wave_sum(slot, j0, j1 - j0, a.spmv ? prev : 0.0f);
                    const uint32_t deg = (uint32_t)__builtin_amdgcn_readlane(cm.deg[k], l);
                    if (lane == 0) og_bin_epilogue_pf(a, v, y, prev, deg, st);
                }
            }
        }
    }
    og_bin_stat r = og_bin_block_stat<BT>(st);
    og_bin_write_part(a, blockIdx.x, r);
}

/* ------------------------------------------------------------------ */
/* long rows                                                           */
/* ------------------------------------------------------------------ */

/* One block per long row (longest first, taken in turn from a counter),
 * running beside the sum pass: the row's pieces in order, each scattered
 * from its entries into LDS in CSR order (as the sum pass does for a bin)
 * and added with the block-wide exact fast-forward, 4 consecutive terms
 * per thread (4096 terms) per step. */
#define OG_BIN_LTT 4u                  /* terms per thread and step */

/* sum (saturated at 2^25) or minimum over the 16 wavefronts' values, read
 * by every lane */
__device__ __forceinline__ void og_bin_waves_read(const uint32_t *v, uint32_t (&w)[LW])
{
    const uint4 *v4 = reinterpret_cast<const uint4 *>(v);   /* 16-byte aligned */
#pragma unroll
    for (uint32_t i = 0; i < LW / 4; i++)
    {
        const uint4 t = v4[i];
        w[4 * i] = t.x;
        w[4 * i + 1] = t.y;
        w[4 * i + 2] = t.z;
        w[4 * i + 3] = t.w;
    }
}

__device__ __forceinline__ uint32_t og_bin_waves_sum(const uint32_t *v)
{
    uint32_t w[LW];
    og_bin_waves_read(v, w);
    uint32_t x = 0;
#pragma unroll
    for (uint32_t i = 0; i < LW; i++) x = min(x + w[i], 1u << 25);
    return x;
}

__device__ __forceinline__ uint32_t og_bin_waves_min(const uint32_t *v)
{
    uint32_t w[LW];
    og_bin_waves_read(v, w);
    uint32_t x = w[0];
#pragma unroll
    for (uint32_t i = 1; i < LW; i++) x = min(x, w[i]);
    return x;
}

/* The block-wide fast-forward of the running sum y (identical in every
 * thread) over xs[p, clen) (LDS); par alternates the partial arrays from
 * step to step. */
__device__ __forceinline__ float og_bin_block_sum(const float *xs, uint32_t p, uint32_t clen, float y,
                                                  uint32_t (&s_wsum)[2][LW], uint32_t (&s_wfail)[2][LW],
                                                  uint32_t &s_pre, uint32_t &par)
{
    const uint32_t lane = threadIdx.x & 63;
    const uint32_t wave = threadIdx.x >> 6;
    while (p < clen)
    {
        const uint32_t left = clen - p;
        if (!(y >= 0x1p-100f && y < 0x1p100f))
        {
            y = y + xs[p];
            p += 1;
            continue;
        }
        int ey;
        frexpf(y, &ey);
        const int sc = 24 - ey;
        const uint32_t Y = (uint32_t)ldexpf(y, sc);   /* in [2^23, 2^24) */
        /* this thread's terms p + 4t .. p + 4t + 3 */
        uint32_t m[OG_BIN_LTT], l[OG_BIN_LTT];
        bool bad[OG_BIN_LTT], valid[OG_BIN_LTT];
        uint32_t tot = 0;
#pragma unroll
        for (uint32_t k = 0; k < OG_BIN_LTT; k++)
        {
            const uint32_t i = OG_BIN_LTT * threadIdx.x + k;
            valid[k] = i < left;
            const float v = xs[min(p + i, clen - 1)];
            og_bin_term(valid[k] ? v : 0.0f, sc, valid[k], m[k], bad[k]);
            tot = min(tot + m[k], 1u << 25);
            l[k] = tot;
        }
        const uint32_t incl_w = og_bin_wave_scan(tot);          /* <= 2^31 */
        if (lane == 63) s_wsum[par][wave] = min(incl_w, 1u << 25);
        __syncthreads();
        uint32_t base = 0;
        {
            /* prefix of the earlier wavefronts' sums */
            uint32_t w[LW];
            og_bin_waves_read(s_wsum[par], w);
#pragma unroll
            for (uint32_t k = 0; k < LW; k++)
                if (k < wave) base = min(base + w[k], 1u << 25);
        }
        const uint32_t X = min(base + incl_w - tot, 1u << 25);   /* before this thread's terms */
        uint32_t kfail = OG_BIN_LTT;
#pragma unroll
        for (int k = (int)OG_BIN_LTT - 1; k >= 0; k--)
            kfail = (valid[k] & (bad[k] | (Y + X + l[k] >= (1u << 24)))) ? (uint32_t)k : kfail;
        const uint64_t fb = __ballot(kfail < OG_BIN_LTT);
        if (lane == 0)
        {
            uint32_t idx = 0xFFFFFFFFu;
            if (fb)
            {
                const uint32_t fl = (uint32_t)__builtin_ctzll(fb);
                idx = OG_BIN_LTT * (64u * wave + fl) + (uint32_t)__builtin_amdgcn_readlane(kfail, fl);
            }
            s_wfail[par][wave] = idx;
        }
        __syncthreads();
        const uint32_t kf = og_bin_waves_min(s_wfail[par]);
        if (kf == 0xFFFFFFFFu)
        {
            const uint32_t total = og_bin_waves_sum(s_wsum[par]);
            y = ldexpf((float)(Y + total), -sc);
            p += left < OG_BIN_LTT * LT ? left : OG_BIN_LTT * LT;
        }
        else
        {
            const uint32_t t = kf / OG_BIN_LTT, kk = kf % OG_BIN_LTT;
            if (threadIdx.x == t)
            {
                uint32_t pre = X;
#pragma unroll
                for (uint32_t k = 0; k < OG_BIN_LTT; k++)
                    if (k + 1 == kk) pre = X + l[k];
                s_pre = pre;
            }
            __syncthreads();
            y = ldexpf((float)(Y + s_pre), -sc);
            y = y + xs[p + kf];
            p += kf + 1;
            __syncthreads();   /* s_pre read by all before its next write */
        }
        par ^= 1;
    }
    return y;
}

__global__ __launch_bounds__(LT) void og_bin_long_kernel(og_bin_dev b, og_pr_dev a, uint32_t part_base)
{
    __shared__ __align__(16) float buf[OG_BIN_CAP];
    __shared__ __align__(16) uint32_t s_wsum[2][LW];
    __shared__ __align__(16) uint32_t s_wfail[2][LW];
    __shared__ uint32_t s_pre, s_row;
    if (*a.done) return;
    og_bin_stat st;
    og_bin_stat_init(st);
    uint32_t par = 0;    /* which s_wsum / s_wfail set a step uses */
    /* rows taken in turn from a counter the gather pass zeroed (longest
     * first: each block takes the next when it is done) */
    for (;;)
    {
        __syncthreads();   /* the previous row's reads of buf and s_row are done */
        if (threadIdx.x == 0) s_row = atomicAdd(b.ticket, 1u);
        __syncthreads();
        const uint32_t q = s_row;
        if (q >= b.n_long) break;
        const uint32_t v = b.long_row[q];
        const uint32_t n = a.row_ptr[v + 1] - a.row_ptr[v];
        float y = a.spmv ? a.r[v] : 0.0f;     /* identical in every thread */
        uint32_t bin = b.long_bin[q];
        for (uint32_t c0 = 0; c0 < n; c0 += OG_BIN_CAP, bin++)
        {
            const og_bin_desc d = b.bins[bin];
            __syncthreads();   /* the previous piece's reads of buf are done */
            /* the piece's entries to their slots, two sweeps of 8 per thread at a time */
            for (uint32_t e = d.ent0 + 8 * threadIdx.x; e < d.ent1; e += 16 * LT)
            {
                const uint32_t e2 = e + 8 * LT;
                const bool two = e2 < d.ent1;
                const uint32_t f = two ? e2 : e;
                const og_u32x4 p0 = __builtin_nontemporal_load((const og_u32x4 *)(b.pos + e));
                const og_f32x4 x0 = __builtin_nontemporal_load((const og_f32x4 *)(b.val + e));
                const og_f32x4 x1 = __builtin_nontemporal_load((const og_f32x4 *)(b.val + e + 4));
                const og_u32x4 p1 = __builtin_nontemporal_load((const og_u32x4 *)(b.pos + f));
                const og_f32x4 y0 = __builtin_nontemporal_load((const og_f32x4 *)(b.val + f));
                const og_f32x4 y1 = __builtin_nontemporal_load((const og_f32x4 *)(b.val + f + 4));
                og_bin_scatter8(buf, p0, x0, x1);
                if (two) og_bin_scatter8(buf, p1, y0, y1);
            }
            __syncthreads();
            uint32_t p = 0;
            if (c0 == 0)
            {
                /* the leading terms one after the other (the running sum
                 * crosses binades every few terms there) */
                p = d.cnt < OG_BIN_HEAD ? d.cnt : OG_BIN_HEAD;
                y = og_bin_uniform(og_bin_row_sum(buf, 0, p, y));
            }
            y = og_bin_block_sum(buf, p, d.cnt, y, s_wsum, s_wfail, s_pre, par);
        }
        if (threadIdx.x == 0) og_bin_epilogue(a, v, y, st);
    }
    og_bin_stat r = og_bin_block_stat<LT>(st);
    og_bin_write_part(a, part_base + blockIdx.x, r);
}

/* ------------------------------------------------------------------ */
/* launch                                                              */
/* ------------------------------------------------------------------ */

extern "C" int og_launch_bin_pull(const og_bin_dev *b, const og_pr_dev *a, uint32_t it, uint32_t sum_grid,
                                  uint32_t long_grid, hipStream_t s, hipEvent_t ev0, hipEvent_t ev1)
{
    (void)it;
    if (ev0 && og_hip_check(hipEventRecord(ev0, s), "hipEventRecord")) return -1;
    if (!b->n_items && og_launch_fill_u32(b->ticket, 0u, 1, s)) return -1;
    if (b->n_items)
        OG_LAUNCH(og_bin_gather_kernel, dim3(og_num_cus() > 0 ? og_num_cus() : 1), dim3(GT), 0, s, *b,
                  (const float *)a->g_in,
                  (const uint32_t *)a->done);
    if (long_grid)
    {
        /* the long rows on the side stream, on CUs of their own, while the
         * sum pass runs on the others */
        hipStream_t side = og_side_stream();
        if (og_hip_check(hipEventRecord(og_side_fork(), s), "hipEventRecord") ||
            og_hip_check(hipStreamWaitEvent(side, og_side_fork(), 0), "hipStreamWaitEvent"))
            return -1;
        OG_LAUNCH(og_bin_long_kernel, dim3(long_grid), dim3(LT), 0, side, *b, *a, sum_grid);
        if (og_hip_check(hipEventRecord(og_side_join(), side), "hipEventRecord")) return -1;
    }
    OG_LAUNCH(og_bin_sum_kernel, dim3(sum_grid), dim3(BT), 0, s, *b, *a);
    if (long_grid && og_hip_check(hipStreamWaitEvent(s, og_side_join(), 0), "hipStreamWaitEvent")) return -1;
    if (ev1 && og_hip_check(hipEventRecord(ev1, s), "hipEventRecord")) return -1;
    return og_hip_check(og_launch_status(), "og_bin_pull");
}

/* ------------------------------------------------------------------ */
/* layout build                                                        */
/* ------------------------------------------------------------------ */

/* key[j] = bin * n_seg + segment of cols[j], val[j] = j for every CSR
 * position j of the bins (one block per bin). */
__global__ void og_bin_keys_kernel(const og_bin_desc *__restrict__ bins, const uint32_t *__restrict__ cols,
                                   uint32_t n_seg, uint32_t *__restrict__ key, uint32_t *__restrict__ val)
{
    const og_bin_desc d = bins[blockIdx.x];
    for (uint32_t k = threadIdx.x; k < d.cnt; k += blockDim.x)
    {
        const uint32_t j = d.pos0 + k;
        key[j] = blockIdx.x * n_seg + cols[j] / OG_BIN_SEG;
        val[j] = j;
    }
}

__global__ void og_bin_flags_kernel(const uint32_t *__restrict__ key, uint32_t n,
                                    const og_bin_desc *__restrict__ bins, uint32_t n_seg,
                                    uint32_t *__restrict__ tile_flag)
{
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
    {
        const uint32_t k = key[i];
        tile_flag[i] = (i == 0 || key[i - 1] != k || i == bins[k / n_seg].pos0) ? 1u : 0u;
    }
}

__global__ void og_bin_tiles_kernel(const uint32_t *__restrict__ key, uint32_t n, uint32_t n_seg,
                                    const uint32_t *__restrict__ scan, uint32_t *__restrict__ tile_seg,
                                    uint32_t *__restrict__ tile_bin, uint32_t *__restrict__ tile_start)
{
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
    {
        if (scan[i + 1] == scan[i]) continue;   /* not a tile start */
        const uint32_t t = scan[i];
        const uint32_t k = key[i];
        const uint32_t bi = k / n_seg;
        tile_seg[t] = k - bi * n_seg;
        tile_bin[t] = bi;
        tile_start[t] = i;
    }
}

__global__ void og_bin_fill_kernel(const uint32_t *__restrict__ key, const uint32_t *__restrict__ val, uint32_t n,
                                   const og_bin_desc *__restrict__ bins, uint32_t n_seg,
                                   const uint32_t *__restrict__ cols, const uint32_t *__restrict__ scan,
                                   const uint32_t *__restrict__ tile_start, const uint32_t *__restrict__ tile_ent,
                                   const uint32_t *__restrict__ tile_src, uint16_t *__restrict__ src,
                                   uint16_t *__restrict__ pos)
{
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
    {
        const uint32_t k = key[i], j = val[i];
        const uint32_t bi = k / n_seg, s = k - bi * n_seg;
        const uint32_t t = scan[i + 1] - 1;          /* the tile holding sorted entry i */
        src[tile_src[t] + (i - tile_start[t])] = (uint16_t)(cols[j] - s * OG_BIN_SEG);
        pos[tile_ent[t] + (i - tile_start[t])] = (uint16_t)(j - bins[bi].pos0);
    }
}

__global__ void og_bin_pads_kernel(const uint32_t *__restrict__ tile_ent, const uint32_t *__restrict__ tile_src,
                                   const uint32_t *__restrict__ tile_cnt, uint32_t n_tiles,
                                   uint16_t *__restrict__ src, uint16_t *__restrict__ pos)
{
    for (uint32_t t = blockIdx.x * blockDim.x + threadIdx.x; t < n_tiles; t += gridDim.x * blockDim.x)
    {
        const uint32_t c = tile_cnt[t];
        const uint32_t pad = (OG_BIN_TILE_ALIGN - c % OG_BIN_TILE_ALIGN) % OG_BIN_TILE_ALIGN;
        for (uint32_t k = 0; k < pad; k++)
        {
            src[tile_src[t] + c + k] = 0;
            pos[tile_ent[t] + c + k] = (uint16_t)OG_BIN_PAD;
        }
    }
}

static unsigned og_bin_grid(uint64_t n)
{
    uint64_t g = (n + 255) / 256;
    if (g > 8192) g = 8192;
    return g ? (unsigned)g : 1u;
}

extern "C" int og_launch_bin_keys(const og_bin_desc *bins, uint32_t n_bins, const uint32_t *cols, uint32_t n_seg,
                                  uint32_t *key, uint32_t *val, hipStream_t s)
{
    if (!n_bins) return 0;
    OG_LAUNCH(og_bin_keys_kernel, dim3(n_bins), dim3(256), 0, s, bins, cols, n_seg, key, val);
    return og_hip_check(og_launch_status(), "og_bin_keys_kernel");
}

extern "C" int og_launch_bin_flags(const uint32_t *key, uint32_t n, const og_bin_desc *bins, uint32_t n_seg,
                                   uint32_t *tile_flag, hipStream_t s)
{
    if (!n) return 0;
    OG_LAUNCH(og_bin_flags_kernel, dim3(og_bin_grid(n)), dim3(256), 0, s, key, n, bins, n_seg, tile_flag);
    return og_hip_check(og_launch_status(), "og_bin_flags_kernel");
}

extern "C" int og_launch_bin_tiles(const uint32_t *key, uint32_t n, uint32_t n_seg, const uint32_t *tile_flag_scan,
                                   uint32_t *tile_seg, uint32_t *tile_bin, uint32_t *tile_start, hipStream_t s)
{
    if (!n) return 0;
    OG_LAUNCH(og_bin_tiles_kernel, dim3(og_bin_grid(n)), dim3(256), 0, s, key, n, n_seg, tile_flag_scan,
                       tile_seg, tile_bin, tile_start);
    return og_hip_check(og_launch_status(), "og_bin_tiles_kernel");
}

extern "C" int og_launch_bin_fill(const uint32_t *key, const uint32_t *val, uint32_t n, const og_bin_desc *bins,
                                  uint32_t n_seg, const uint32_t *cols, const uint32_t *tile_flag_scan,
                                  const uint32_t *tile_start, const uint32_t *tile_ent, const uint32_t *tile_src,
                                  uint16_t *src, uint16_t *pos, hipStream_t s)
{
    if (!n) return 0;
    OG_LAUNCH(og_bin_fill_kernel, dim3(og_bin_grid(n)), dim3(256), 0, s, key, val, n, bins, n_seg, cols,
                       tile_flag_scan, tile_start, tile_ent, tile_src, src, pos);
    return og_hip_check(og_launch_status(), "og_bin_fill_kernel");
}

extern "C" int og_launch_bin_pads(const uint32_t *tile_ent, const uint32_t *tile_src, const uint32_t *tile_cnt,
                                  uint32_t n_tiles, uint16_t *src, uint16_t *pos, hipStream_t s)
{
    if (!n_tiles) return 0;
    OG_LAUNCH(og_bin_pads_kernel, dim3(og_bin_grid(n_tiles)), dim3(256), 0, s, tile_ent, tile_src, tile_cnt,
                       n_tiles, src, pos);
    return og_hip_check(og_launch_status(), "og_bin_pads_kernel");
}
