/*
 * og_plan.h -- launchers of the plan-building kernels (og_plan.hip).
 */
#ifndef OG_PLAN_H
#define OG_PLAN_H

#include <stddef.h>
#include <stdint.h>

#include <hip/hip_runtime_api.h>

#ifdef __cplusplus
extern "C" {
#endif

size_t og_scan_scratch_bytes(uint32_t n);
/* d_out[0..n) = exclusive prefix sums of d_in, d_out[n] = total. */
int og_scan_exclusive(const uint32_t *d_in, uint32_t *d_out, uint32_t n, void *d_scratch,
                      hipStream_t s);
/* Row class of an in-degree (0..3, the order of the plan's row ranges). */
int og_launch_plan_keys(uint32_t N, const uint32_t *outdeg, const uint32_t *indeg, uint32_t seq_t,
                        uint32_t chain_t, uint32_t relabel, uint32_t *key, uint32_t *val, hipStream_t s);
/* old2new, degrees in new order; class_cnt[k] += rows of class k, and
 * class_cnt[4] += rows of more than chain_t in-edges (chain_t > 0). */
int og_launch_plan_perm(uint32_t N, const uint32_t *new2old, const uint32_t *outdeg_old,
                        const uint32_t *indeg_old, uint32_t seq_t, uint32_t chain_t, uint32_t *old2new,
                        uint32_t *outdeg_new, uint32_t *indeg_new, uint32_t *class_cnt,
                        hipStream_t s);
/* cols of rows [0, N): a block per row below H, a wavefront per row above. */
int og_launch_plan_copy(uint32_t H, uint32_t N, const uint32_t *new2old, const uint32_t *old2new,
                        const uint32_t *idx_old, const uint32_t *lists, const uint32_t *row_ptr,
                        uint32_t *cols, hipStream_t s);
/* rowid[e - base] = r for every edge e of the rows [0, R) of row_ptr. */
int og_launch_plan_row_ids_at(uint32_t R, const uint32_t *row_ptr, uint32_t base, uint32_t *rowid,
                              hipStream_t s);
/* cnt[r] = chunks of row r < C (ceil(in-degree / ch)). */
int og_launch_plan_chunks(uint32_t C, const uint32_t *row_ptr, uint32_t ch, uint32_t *cnt,
                          hipStream_t s);
/* Chunk descriptors from the chunk prefix first[0..C]. */
int og_launch_plan_chunk_desc(uint32_t H, uint32_t C, const uint32_t *row_ptr,
                              const uint32_t *first, uint32_t ch, uint32_t *ch_row,
                              uint32_t *ch_edge, hipStream_t s);
int og_launch_plan_map_pairs(uint32_t E, const uint32_t *a, const uint32_t *b,
                             const uint32_t *old2new, uint32_t *oa, uint32_t *ob, hipStream_t s);

#ifdef __cplusplus
}
#endif

#endif
