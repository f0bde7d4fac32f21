/*
 * og_stats.h -- argument block and launchers of the rank-comparison kernels
 * (og_stats.hip), shared with their host side in og_stats.c.
 */
#ifndef OG_STATS_H
#define OG_STATS_H

#include <stdint.h>

#include <hip/hip_runtime_api.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Edit-distance problems up to OG_LEV_SMALL_K run on one wavefront (16
 * rows per lane), up to OG_LEV_LDS_K on eight (24 rows per lane); larger
 * ones walk anti-diagonals kept in global memory. */
#define OG_LEV_SMALL_K 1024u
#define OG_LEV_LDS_K 12288u

typedef struct
{
    uint32_t N, P;                    /* vertices, problems */
    const float *ref_pr, *pr;         /* ranks of the pull run and of the compared run */
    const uint32_t *ref_order, *order;/* realRanks: vertices ascending by rank */
    uint32_t *inv, *ref_inv;          /* position of each vertex */
    const uint32_t *k, *m, *off;      /* per problem: top k, num_vertices, slice offset (P + 1) */
    uint32_t *ids, *ref_ids;          /* slices */
    float *f, *ref_f, *re, *ref_re;
    uint32_t *lev;                    /* results */
    long long *kt;                    /* 6 per problem: s, m1, m2 of the float and the id arrays */
    uint32_t *inter, *mism;
    double *err_f, *err_r;
} og_stats_dev;

int og_launch_stats_inverse(const og_stats_dev *a, hipStream_t s);
int og_launch_stats_slices(const og_stats_dev *a, uint32_t total, hipStream_t s);
int og_launch_lev_small(const og_stats_dev *a, const uint32_t *plist, uint32_t n, hipStream_t s);
int og_launch_lev_large(const og_stats_dev *a, const uint32_t *plist, uint32_t n, hipStream_t s);
int og_launch_lev_global(const og_stats_dev *a, uint32_t p, uint32_t *scratch, hipStream_t s);
/* Kendall counts: problems with k <= OG_KT_SMALL_K one block each (plist),
 * larger ones by groups: n_groups (problem, first row) pairs of u32 */
#define OG_KT_SMALL_K 1024u
int og_launch_kendall_small(const og_stats_dev *a, const uint32_t *plist, uint32_t n, hipStream_t s);
int og_launch_kendall(const og_stats_dev *a, const void *groups, uint32_t n_groups, hipStream_t s);
int og_launch_topk_sums(const og_stats_dev *a, hipStream_t s);

#ifdef __cplusplus
}
#endif

#endif
