/*
 * og_oracle_mlp.c -- CPU oracle of the learned warm start (TEST
 * INFRASTRUCTURE ONLY, like og_oracle.c): the per-vertex degree features,
 * the 4 -> 64 -> 64 -> 1 MLP forward pass and the gradient of its L1 loss,
 * all in double precision.  The reference has no learned component; these
 * functions restate the network the library defines (opengraph_amd_ext.h,
 * ogpu_mlp_*), so that the GPU forward, backward and training kernels can
 * be checked against an independent implementation.  The gradient is in
 * turn pinned by central finite differences of orc_mlp_loss in the tests.
 *
 * Parameter vector (4545 floats): W1[64][4] at 0, b1[64] at 256,
 * W2[64][64] at 320 (row = output unit), b2[64] at 4416, W3[64] at 4480,
 * b3 at 4544.
 */
#include "og_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

#define ORC_ID_MASK 0x3FFFFFFFu

enum { H = 64, OW1 = 0, OB1 = 256, OW2 = 320, OB2 = 4416, OW3 = 4480, OB3 = 4544, NP = 4545 };

void orc_mlp_features(const orc_csr *g, float *feat)
{
    const uint32_t N = g->num_vertices;
    for (uint32_t v = 0; v < N; v++)
    {
        const uint32_t din = g->i_out_degree[v];    /* inverse CSR: in-neighbours of v */
        double s = 0.0;
        for (uint32_t j = g->i_edges_idx[v]; j < g->i_edges_idx[v] + din; j++)
            s += log1p((double)g->i_out_degree[g->i_dst[j] & ORC_ID_MASK]);
        feat[4 * (size_t)v + 0] = (float)log1p((double)din);
        feat[4 * (size_t)v + 1] = (float)log1p((double)g->f_out_degree[v]);
        feat[4 * (size_t)v + 2] = din ? (float)(s / din) : 0.0f;
        feat[4 * (size_t)v + 3] = (float)(1.0 / N);
    }
}

static double orc_softplus(double z) { return z > 30.0 ? z : log1p(exp(z)); }
static double orc_sigmoid(double z) { return 1.0 / (1.0 + exp(-z)); }

/* Forward of one vertex; keeps the activations for the backward pass. */
static double orc_mlp_one(const float *P, const float *f, double *z1, double *z2, double *z3)
{
    double h1[H];
    for (int k = 0; k < H; k++)
    {
        double a = P[OB1 + k];
        for (int j = 0; j < 4; j++) a += (double)P[OW1 + 4 * k + j] * f[j];
        z1[k] = a;
        h1[k] = a > 0 ? a : 0.0;
    }
    double o3 = P[OB3];
    for (int n = 0; n < H; n++)
    {
        double a = P[OB2 + n];
        for (int i = 0; i < H; i++) a += (double)P[OW2 + H * n + i] * h1[i];
        z2[n] = a;
        o3 += (double)P[OW3 + n] * (a > 0 ? a : 0.0);
    }
    *z3 = o3;
    return orc_softplus(o3);
}

void orc_mlp_forward(const float *params, const float *feat, uint32_t n, double *out)
{
    for (uint32_t v = 0; v < n; v++)
    {
        double z1[H], z2[H], z3;
        out[v] = orc_mlp_one(params, feat + 4 * (size_t)v, z1, z2, &z3);
    }
}

double orc_mlp_loss(const float *params, const float *feat, const float *y, uint32_t n)
{
    double l = 0.0;
    for (uint32_t v = 0; v < n; v++)
    {
        double z1[H], z2[H], z3;
        l += fabs(orc_mlp_one(params, feat + 4 * (size_t)v, z1, z2, &z3) - y[v]);
    }
    return n ? l / n : 0.0;
}

double orc_mlp_grad(const float *P, const float *feat, const float *y, uint32_t n, double *grad)
{
    memset(grad, 0, NP * sizeof(double));
    double l = 0.0;
    for (uint32_t v = 0; v < n; v++)
    {
        const float *f = feat + 4 * (size_t)v;
        double z1[H], z2[H], z3;
        const double o = orc_mlp_one(P, f, z1, z2, &z3);
        const double d = o - y[v];
        l += fabs(d);
        /* dL/do = sign(o - y) / n, softplus' = sigmoid */
        const double g3 = ((d > 0) - (d < 0)) / (double)n * orc_sigmoid(z3);
        grad[OB3] += g3;
        double dz2[H];
        for (int m = 0; m < H; m++)
        {
            const double h2 = z2[m] > 0 ? z2[m] : 0.0;
            grad[OW3 + m] += g3 * h2;
            dz2[m] = z2[m] > 0 ? g3 * P[OW3 + m] : 0.0;
            grad[OB2 + m] += dz2[m];
        }
        for (int m = 0; m < H; m++)
            for (int i = 0; i < H; i++) grad[OW2 + H * m + i] += dz2[m] * (z1[i] > 0 ? z1[i] : 0.0);
        for (int i = 0; i < H; i++)
        {
            if (z1[i] <= 0) continue;
            double dz1 = 0.0;
            for (int m = 0; m < H; m++) dz1 += dz2[m] * P[OW2 + H * m + i];
            grad[OB1 + i] += dz1;
            for (int j = 0; j < 4; j++) grad[OW1 + 4 * i + j] += dz1 * f[j];
        }
    }
    return n ? l / n : 0.0;
}
