/*
 * og_learn.h -- kernels of the learned warm start (og_learn.hip), shared
 * with the host side in og_learn.c.
 *
 * The network: per-vertex features f = [log1p in-degree, log1p out-degree,
 * mean log1p in-degree of the in-neighbours, 1/N], then
 *   h1 = relu(W1 f + b1)  (64),  h2 = relu(W2 h1 + b2)  (64),
 *   o  = softplus(W3 . h2 + b3),
 * fp32 parameters packed as W1[64][4] | b1[64] | W2[64][64] (row = output
 * unit) | b2[64] | W3[64] | b3  (4545 floats).
 */
#ifndef OG_LEARN_H
#define OG_LEARN_H

#include <stdint.h>

#include <hip/hip_runtime_api.h>

#ifdef __cplusplus
extern "C" {
#endif

#define OG_MLP_H 64
#define OG_MLP_NP 4545
#define OG_MLP_W1 0
#define OG_MLP_B1 256
#define OG_MLP_W2 320
#define OG_MLP_B2 4416
#define OG_MLP_W3 4480
#define OG_MLP_B3 4544
#define OG_MLP_TRAIN_BLOCK 64    /* vertices per block of the training kernel */

/* feat[v] (float4) from in-degree (row_ptr), out-degree and the in-neighbour
 * sum of log1p in-degree (insum); n vertices. */
int og_launch_feat_li(const uint32_t *row_ptr, uint32_t n, float *li, hipStream_t s);
int og_launch_feat_finish(const uint32_t *row_ptr, const uint32_t *outdeg, const float *insum,
                          uint32_t n, float *feat, hipStream_t s);
/* o[v] = MLP(feat[v]) in fp32 (VALU) or with layer 2 on bf16 MFMA. */
int og_launch_mlp_infer(const float *params, const float *feat, uint32_t n, int mfma, float *out,
                        hipStream_t s);
/* One training step's gradient: vertices idx(base .. base+count) of the
 * dataset, where idx is the Feistel permutation (seed) of [0, n_data)
 * (identity if seed == 0).  Block partials then a fixed-order reduction into
 * grad[4545] and loss_sum[0]. */
int og_launch_mlp_grad(const float *params, const float *feat, const float *y, uint32_t n_data,
                       uint64_t seed, uint32_t base, uint32_t count, float *part, float *grad,
                       double *loss_sum, hipStream_t s);
size_t og_mlp_part_floats(uint32_t count);
int og_launch_mlp_adam(float *params, const float *grad, float *m1, float *m2, float lr, float b1,
                       float b2, float eps, uint32_t step, hipStream_t s);
/* The no-allocation gradient: og_learn_noalloc_setup (once per device,
 * with the device current, at context creation) finds the device address
 * of the static partial array; og_launch_mlp_grad_noalloc runs
 * og_launch_mlp_grad over n <= OGPU_MLP_GRAD_NOALLOC_MAX dataset rows
 * (identity order) with its partials there. */
int og_learn_noalloc_setup(int device);
int og_launch_mlp_grad_noalloc(int device, const float *params, const float *feat, const float *y,
                               uint32_t n, float *grad, double *loss_sum, hipStream_t s);
/* Start ranks in old ids: 1 - d (unnormalized) for vertices without
 * in-edges, N * S_U * o[v] / sum_U(o) for the others, S_U from the
 * fixed-point mass identity (fp64 sums in a fixed order; scratch holds
 * OG_WARM_SCRATCH doubles).  A prediction with a NaN, infinite or negative
 * value on a vertex with in-edges, or with no mass there, is rejected: the
 * start written is then the uniform 1.0, and scratch[OG_WARM_OK] is 0
 * (1 when accepted). */
#define OG_WC_IN_MAX 16                   /* in-degree classes of the coarse correction at most */
#define OG_WC_MAX (2 * OG_WC_IN_MAX)      /* its classes at most (each split by out-edges) */
#define OG_WC_BASE (5 * 1024 + 64)        /* its region of the scratch */
#define OG_WARM_SCRATCH (OG_WC_BASE + OG_WC_MAX * OG_WC_MAX + 3 * OG_WC_MAX)
#define OG_WARM_OK (5 * 1024 + 1)
/* The start's coarse correction over nc = 2 nin classes, on the graph's
 * forward edge list in the caller's ids: vertex v is in class 2 i + o, i
 * its in-degree class among nin (0: no in-edges; i > 0: floor(log2
 * in-degree) = i - 1, the last one open-ended), o = 1 when it has
 * out-edges.  Sums are in Q32.32 (u64, exact in any order).
 * og_launch_wc_sums: cls[v], c0[v] = r0[v] / out-degree, and per class the
 * mass of r0 and the vertex count (u64 at og_wc_mass / og_wc_count).
 * og_launch_wc_matrix: mat[i * nc + j] = sum over edges u -> v with
 * class(v) = i, class(u) = j of c0[u].
 * og_launch_wc_apply: r0[v] *= s[cls[v]]. */
int og_launch_wc_sums(const float *r0, const uint32_t *in_degree, const uint32_t *out_degree, uint32_t n,
                      uint32_t nin, uint8_t *cls, float *c0, double *scratch, hipStream_t s);
int og_launch_wc_matrix(const uint32_t *src, const uint32_t *dst, uint32_t E, const uint8_t *cls,
                        const float *c0, uint32_t nc, unsigned long long *mat, hipStream_t s);
int og_launch_wc_apply(float *r0, const uint8_t *cls, uint32_t n, const float *scale, hipStream_t s);
static inline size_t og_wc_matrix(void) { return OG_WC_BASE; }
static inline size_t og_wc_mass(void) { return OG_WC_BASE + OG_WC_MAX * OG_WC_MAX; }
static inline size_t og_wc_count(void) { return og_wc_mass() + OG_WC_MAX; }
static inline size_t og_wc_scale(void) { return og_wc_count() + OG_WC_MAX; }
int og_launch_warm_scale(const float *o, const uint32_t *row_ptr, const uint32_t *outdeg,
                         const uint32_t *new2old, uint32_t n, double *sums, float *r0_old,
                         hipStream_t s);

#ifdef __cplusplus
}
#endif

#endif
