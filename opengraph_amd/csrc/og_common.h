/*
 * og_common.h -- inline helpers shared by the host C code and the gfx950
 * kernels of libopengraph_amd: the counter-based RNG and vertex permutation
 * of the synthetic graph generators, the x86-64 float->integer conversion
 * results that the reference's fixed-point and quantized casts rely on, and
 * the numeric constants of the PageRank map.
 */
#ifndef OG_COMMON_H
#define OG_COMMON_H

#include <stdint.h>

#if defined(__HIPCC__)
#define OG_HD __host__ __device__ __forceinline__
#else
#define OG_HD static inline
#endif

/* pageRank.h:13 and pageRank.c:53-54. */
#define OG_DAMP 0.85f
/* base_pr = 1.0f - damp = 0.149999976f in fp32 (pageRank.c:54), not 0.15f. */
#define OG_BASE_PR (1.0f - OG_DAMP)
/* reorder.h:19: vertex ids carry two mask bits above bit 30. */
#define OG_ID_MASK 0x3FFFFFFFu

/* ---------------------------------------------------------------------
 * Counter-based random numbers.  og_mix64 is the SplitMix64 output
 * function (Steele, Lea, Flood, "Fast splittable pseudorandom number
 * generators", OOPSLA 2014).  draw(stream, ctr) is the ctr-th output of the
 * SplitMix64 sequence whose state starts at key(seed, stream), so every
 * value depends only on (seed, stream, ctr) and any thread can produce it.
 * ------------------------------------------------------------------- */
#define OG_GAMMA 0x9E3779B97F4A7C15ull

OG_HD uint64_t og_mix64(uint64_t z)
{
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

OG_HD uint64_t og_stream_key(uint64_t seed, uint64_t stream)
{
    return og_mix64(seed * OG_GAMMA + stream);
}

OG_HD uint64_t og_draw(uint64_t key, uint64_t ctr)
{
    return og_mix64(key + (ctr + 1) * OG_GAMMA);
}

/* Streams of the generators. */
#define OG_STREAM_RMAT 0u
#define OG_STREAM_PERM 1u
#define OG_STREAM_ER 2u

/* ---------------------------------------------------------------------
 * Keyed permutation of [0, 2^bits): a 4-round balanced Feistel network on
 * 2*h bits (h = ceil(bits/2)), cycle-walked back into range when bits is
 * odd.  Each round is (L, R) -> (R, L ^ F(R)), F = og_mix64(R ^ k_r) masked
 * to h bits, which is a bijection for any F; cycle walking keeps it one on
 * the sub-range.
 * ------------------------------------------------------------------- */
typedef struct
{
    uint64_t k[4];
    uint32_t bits, half;
} og_perm_t;

OG_HD og_perm_t og_perm_make(uint64_t seed, uint32_t bits)
{
    og_perm_t p;
    uint64_t key = og_stream_key(seed, OG_STREAM_PERM);
    for (int r = 0; r < 4; r++) p.k[r] = og_draw(key, (uint64_t)r);
    p.bits = bits;
    p.half = (bits + 1) / 2;
    return p;
}

OG_HD uint32_t og_perm_apply(const og_perm_t *p, uint32_t x)
{
    if (p->bits == 0) return 0;
    const uint64_t hmask = (1ull << p->half) - 1;
    const uint64_t lim = 1ull << p->bits;
    uint64_t v = x;
    do
    {
        uint64_t L = v >> p->half, R = v & hmask;
        for (int r = 0; r < 4; r++)
        {
            uint64_t nL = R;
            uint64_t nR = L ^ (og_mix64(R ^ p->k[r]) & hmask);
            L = nL;
            R = nR;
        }
        v = (L << p->half) | R;
    } while (v >= lim);
    return (uint32_t)v;
}

/* R-MAT quadrant thresholds on a 32-bit uniform draw (host computes them
 * from a, b, c in double; device and host then agree exactly). */
typedef struct
{
    uint32_t scale;
    uint32_t t_a, t_ab, t_abc;   /* u < t_a: (0,0); < t_ab: (0,1); < t_abc: (1,0); else (1,1) */
    uint32_t permute;
    og_perm_t perm;
    uint64_t key;
} og_rmat_t;

/* Edge e of an R-MAT graph: one 32-bit draw per level (counter e*64 + l),
 * most significant bit first, then the vertex permutation on both ends. */
OG_HD void og_rmat_edge(const og_rmat_t *g, uint64_t e, uint32_t *src, uint32_t *dst)
{
    uint32_t s = 0, d = 0;
    for (uint32_t l = 0; l < g->scale; l++)
    {
        uint32_t u = (uint32_t)(og_draw(g->key, e * 64 + l) >> 32);
        uint32_t sb = u >= g->t_ab;                 /* quadrants (1,0) and (1,1) */
        uint32_t db = (u >= g->t_a && u < g->t_ab) || u >= g->t_abc;
        s = (s << 1) | sb;
        d = (d << 1) | db;
    }
    if (g->permute)
    {
        s = og_perm_apply(&g->perm, s);
        d = og_perm_apply(&g->perm, d);
    }
    *src = s;
    *dst = d;
}

/* Edge e of an Erdos-Renyi G(n, m) multigraph: uniform src and dst in
 * [0, n) from draws 2e and 2e+1 (multiply-shift range reduction). */
OG_HD void og_er_edge(uint64_t key, uint32_t n, uint64_t e, uint32_t *src, uint32_t *dst)
{
    uint64_t a = og_draw(key, 2 * e) >> 32, b = og_draw(key, 2 * e + 1) >> 32;
    *src = (uint32_t)((a * n) >> 32);
    *dst = (uint32_t)((b * n) >> 32);
}

/* ---------------------------------------------------------------------
 * x86-64 results of the reference's float->integer casts (gcc -O3 emits
 * cvttss2si / cvttsd2si; an out-of-range input yields the integer
 * indefinite value, INT_MIN of the instruction's width).  The GPU's own
 * conversions saturate instead, so the kernels call these.
 * ------------------------------------------------------------------- */
OG_HD int32_t og_x86_f2i32(float x)
{
    if (!(x > -2147483904.0f && x < 2147483648.0f)) return INT32_MIN;
    return (int32_t)x;
}
OG_HD int64_t og_x86_f2i64(float x)
{
    if (!(x > -9223373136366403584.0f && x < 9223372036854775808.0f)) return INT64_MIN;
    return (int64_t)x;
}
OG_HD int64_t og_x86_d2i64(double x)
{
    if (!(x > -9223372036854777856.0 && x < 9223372036854775808.0)) return INT64_MIN;
    return (int64_t)x;
}
OG_HD uint64_t og_x86_d2u64(double x)
{
    if (x >= 9223372036854775808.0)
        return (uint64_t)og_x86_d2i64(x - 9223372036854775808.0) ^ 0x8000000000000000ull;
    return (uint64_t)og_x86_d2i64(x);
}

/* fixedPoint.h:36 DoubleToFixed64 of the fp32 quotient r/outdeg. */
OG_HD uint64_t og_fixed64_of(float q)
{
    return og_x86_d2u64((double)q * 4294967296.0 + (q >= 0 ? 0.5 : -0.5));
}

#endif /* OG_COMMON_H */
