/*
 * og_gen.c -- seeded synthetic inputs (R-MAT and Erdos-Renyi) on host and
 * device, and the reference's raw .bin writer.  The reference has no
 * generator (its inputs are files read by edgeList.c:336-594); these produce
 * files in its format and device arrays for benchmarks, from the same
 * counter-based definitions (og_common.h), so host and device agree exactly.
 */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "og_internal.h"
#include "opengraph_amd_ext.h"

static uint32_t og_threshold(double p)
{
    if (p <= 0.0) return 0;
    double x = p * 4294967296.0;
    if (x >= 4294967295.0) return 0xFFFFFFFFu;
    return (uint32_t)x;
}

static int og_rmat_setup(uint32_t scale, double a, double b, double c, uint64_t seed, int permute,
                         og_rmat_t *g)
{
    if (scale == 0 || scale > 30)
    {
        og_set_error("R-MAT scale %u outside [1, 30]", scale);
        return -1;
    }
    if (a < 0 || b < 0 || c < 0 || a + b + c > 1.0 + 1e-12)
    {
        og_set_error("R-MAT probabilities a=%g b=%g c=%g are not a distribution", a, b, c);
        return -1;
    }
    memset(g, 0, sizeof(*g));
    g->scale = scale;
    g->t_a = og_threshold(a);
    g->t_ab = og_threshold(a + b);
    g->t_abc = og_threshold(a + b + c);
    g->permute = permute ? 1u : 0u;
    g->perm = og_perm_make(seed, scale);
    g->key = og_stream_key(seed, OG_STREAM_RMAT);
    return 0;
}

int ogpu_rmat_host(uint32_t scale, uint64_t n, double a, double b, double c, uint64_t seed,
                   int permute, uint32_t *src, uint32_t *dst)
{
    og_rmat_t g;
    if (og_rmat_setup(scale, a, b, c, seed, permute, &g)) return -1;
    #pragma omp parallel for schedule(static)
    for (uint64_t e = 0; e < n; e++) og_rmat_edge(&g, e, &src[e], &dst[e]);
    return 0;
}

int ogpu_rmat_device(uint32_t scale, uint64_t n, double a, double b, double c, uint64_t seed,
                     int permute, uint32_t *d_src, uint32_t *d_dst)
{
    og_rmat_t g;
    if (og_ctx_ensure() || og_rmat_setup(scale, a, b, c, seed, permute, &g)) return -1;
    if (og_launch_rmat(d_src, d_dst, n, &g, og_stream())) return -1;
    return og_hip_check(hipStreamSynchronize(og_stream()), "sync rmat");
}

int ogpu_er_host(uint32_t nv, uint64_t n, uint64_t seed, uint32_t *src, uint32_t *dst)
{
    if (nv == 0) { og_set_error("Erdos-Renyi graph needs at least one vertex"); return -1; }
    uint64_t key = og_stream_key(seed, OG_STREAM_ER);
    #pragma omp parallel for schedule(static)
    for (uint64_t e = 0; e < n; e++) og_er_edge(key, nv, e, &src[e], &dst[e]);
    return 0;
}

int ogpu_er_device(uint32_t nv, uint64_t n, uint64_t seed, uint32_t *d_src, uint32_t *d_dst)
{
    if (nv == 0) { og_set_error("Erdos-Renyi graph needs at least one vertex"); return -1; }
    if (og_ctx_ensure()) return -1;
    if (og_launch_er(d_src, d_dst, n, nv, og_stream_key(seed, OG_STREAM_ER), og_stream())) return -1;
    return og_hip_check(hipStreamSynchronize(og_stream()), "sync er");
}

int ogpu_write_bin(const char *path, const uint32_t *src, const uint32_t *dst, uint64_t n)
{
    FILE *f = fopen(path, "wb");
    if (!f) { og_set_error("cannot open %s for writing", path); return -1; }
    enum { CH = 1 << 16 };
    uint32_t *buf = (uint32_t *)malloc(2 * CH * sizeof(uint32_t));
    if (!buf) { fclose(f); og_set_error("out of host memory"); return -1; }
    for (uint64_t i = 0; i < n; i += CH)
    {
        uint64_t m = n - i < CH ? n - i : CH;
        for (uint64_t k = 0; k < m; k++)
        {
            buf[2 * k] = src[i + k];
            buf[2 * k + 1] = dst[i + k];
        }
        if (fwrite(buf, 8, m, f) != m)
        {
            free(buf);
            fclose(f);
            og_set_error("short write to %s", path);
            return -1;
        }
    }
    free(buf);
    if (fclose(f)) { og_set_error("close of %s failed", path); return -1; }
    return 0;
}
