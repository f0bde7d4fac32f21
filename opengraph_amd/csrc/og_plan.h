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
int og_launch_plan_keys(uint32_t N, const uint32_t *outdeg, const uint32_t *indeg, uint32_t hub_t,
                        uint32_t relabel, uint32_t *key, uint32_t *val, hipStream_t s);
int og_launch_plan_perm(uint32_t N, const uint32_t *new2old, const uint32_t *outdeg_old,
                        const uint32_t *indeg_old, uint32_t hub_t, uint32_t *old2new,
                        uint32_t *outdeg_new, uint32_t *indeg_new, uint32_t *n_hubs, hipStream_t s);
int og_launch_plan_copy(uint32_t H, uint32_t N, const uint32_t *new2old, const uint32_t *old2new,
                        const uint32_t *idx_old, const uint32_t *lists, const uint32_t *row_ptr,
                        uint32_t *cols, hipStream_t s);
int og_launch_plan_chunks(uint32_t H, const uint32_t *row_ptr, uint32_t te, uint32_t *cnt,
                          hipStream_t s);
int og_launch_plan_windows(uint32_t H, uint32_t N, const uint32_t *row_ptr, uint32_t te,
                           uint32_t rw, uint32_t n_windows, uint32_t *win_row, uint32_t *win_edge,
                           hipStream_t s);
int og_launch_plan_map_pairs(uint32_t E, const uint32_t *a, const uint32_t *b,
                             const uint32_t *old2new, uint32_t *oa, uint32_t *ob, hipStream_t s);

#ifdef __cplusplus
}
#endif

#endif
