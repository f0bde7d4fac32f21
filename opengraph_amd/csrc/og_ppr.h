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

typedef struct og_ppr_dev
{
    uint32_t N, k, H, C, n_chunks;
    uint32_t it;                 /* iteration being computed */
    uint32_t max_iter;
    double epsilon;              /* per-column L1 tolerance */
    const uint32_t *row_ptr, *cols, *outdeg, *chunk_base, *ch_row, *ch_edge;
    const uint32_t *old2new;
    const uint32_t *src;         /* k restart vertices, new ids */
    float *x;                    /* N x k, row-major */
    const float *c_in;           /* N x k contributions x / outdeg gathered this iteration */
    float *c_out;                /* N x k contributions of the new x */
    float *chunk_part;           /* chunk_base[H] x k: sums of the chunks of rows [0, H) */
    double *part_l1;             /* OG_PPR_MAIN_BLOCKS x k */
    double *fin_l1;              /* OG_PPR_FIN_BLOCKS x k */
    uint64_t *active;            /* bit j: column j still iterating */
    uint32_t *ticket, *done;
    uint32_t *iters;             /* k: iterations applied to each column */
    double *l1;                  /* max_iter x k: L1 change of each iteration (0 once frozen) */
} og_ppr_dev;

int og_launch_ppr_init(const og_ppr_dev *a, hipStream_t s);
int og_launch_ppr_iter(const og_ppr_dev *a, hipStream_t s, hipEvent_t ev0, hipEvent_t ev1);
int og_launch_ppr_out(const og_ppr_dev *a, float *out_old, hipStream_t s);

#ifdef __cplusplus
}
#endif

#endif
