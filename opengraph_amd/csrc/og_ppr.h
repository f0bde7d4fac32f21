/*
 * og_ppr.h -- kernels of batched personalized PageRank (og_ppr.hip), shared
 * with the host orchestration in og_ppr.c.
 */
#ifndef OG_PPR_H
#define OG_PPR_H

#include <stdint.h>

#include <hip/hip_runtime_api.h>

#include "og_internal.h"

#ifdef __cplusplus
extern "C" {
#endif

#define OG_PPR_KMAX 64u          /* columns: one per lane of a wavefront */
#define OG_PPR_MAIN_BLOCKS 2048u /* fixed grid of the main kernel (deterministic sums) */
#define OG_PPR_FIN_BLOCKS 64u    /* fixed grid of the finalize kernel */

/* Columns live in "slots" (lanes): slot j holds column colmap[j].  As
 * columns converge the host moves the live ones to the lowest slots, so a
 * gathered row only touches its first kact floats. */
typedef struct og_ppr_dev
{
    uint32_t N, E, k, H, C, n_chunks;
    uint32_t it;                 /* iteration being computed */
    uint32_t max_iter;
    uint32_t kact;               /* slots [0, kact) may hold active columns */
    double epsilon;              /* per-column L1 tolerance */
    const uint32_t *row_ptr, *cols, *outdeg, *chunk_base, *ch_row, *ch_edge;
    const uint32_t *old2new;
    const uint32_t *src;         /* restart vertex (new id) of each slot */
    const uint32_t *colmap;      /* column of each slot */
    float *x;                    /* N x k, row-major */
    const float *c_in;           /* N x k contributions x / outdeg gathered this iteration */
    float *c_out;                /* N x k contributions of the new x */
    float *chunk_part;           /* chunk_base[H] x k: sums of the chunks of rows [0, H) */
    double *part_l1;             /* OG_PPR_MAIN_BLOCKS x k */
    double *fin_l1;              /* OG_PPR_FIN_BLOCKS x k */
    uint64_t *active;            /* bit j: slot j's column still iterating */
    uint32_t *ticket, *done;
    uint32_t *iters;             /* k, by column: iterations applied to it */
    double *l1;                  /* max_iter x k, by column: L1 change of each iteration (0 once frozen) */
} og_ppr_dev;

/* New slot j takes old slot perm[j] (j < k), in x and c. */
typedef struct og_ppr_perm
{
    uint8_t p[OG_PPR_KMAX];
} og_ppr_perm;

int og_launch_ppr_init(const og_ppr_dev *a, hipStream_t s);
int og_launch_ppr_permute(const og_ppr_dev *a, float *c, og_ppr_perm perm, hipStream_t s);
int og_launch_ppr_iter(const og_ppr_dev *a, hipStream_t s, hipEvent_t ev0, hipEvent_t ev1);
int og_launch_ppr_out(const og_ppr_dev *a, float *out_old, hipStream_t s);

#ifdef __cplusplus
}
#endif

#endif
