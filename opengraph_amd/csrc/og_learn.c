/*
 * og_learn.c -- host side of the learned warm start (og_learn.hip): the
 * network object, its training set of (features, converged ranks) built on
 * the GPU from graphs, training with Adam, and the start vector handed to a
 * PageRank run.
 */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

#include "og_internal.h"
#include "og_learn.h"
#include "og_pr.h"
#include "opengraph_amd_ext.h"

struct ogpu_mlp
{
    float *params, *grad, *m1, *m2;   /* device, OG_MLP_NP each */
    uint32_t step;                    /* Adam steps taken */
    float *feat, *y;                  /* training set: n_data x 4 features, targets */
    uint32_t n_data, cap;
    float *part;                      /* gradient block partials */
    size_t part_floats;
    double *loss;
};

static double og_learn_now(void)
{
    struct timespec ts;
    clock_gettime(CLOCK_MONOTONIC, &ts);
    return ts.tv_sec + ts.tv_nsec * 1e-9;
}

void ogpu_mlp_free(ogpu_mlp *m)
{
    if (!m) return;
    og_dfree(m->params); og_dfree(m->grad); og_dfree(m->m1); og_dfree(m->m2);
    og_dfree(m->feat); og_dfree(m->y); og_dfree(m->part); og_dfree(m->loss);
    free(m);
}

int ogpu_mlp_create(uint64_t seed, ogpu_mlp **out)
{
    *out = NULL;
    if (og_ctx_ensure()) return -1;
    ogpu_mlp *m = (ogpu_mlp *)calloc(1, sizeof(ogpu_mlp));
    if (!m) { og_set_error("out of host memory"); return -1; }
    const size_t pb = OG_MLP_NP * 4;
    m->params = og_dalloc(pb);
    m->grad = og_dalloc(pb);
    m->m1 = og_dalloc(pb);
    m->m2 = og_dalloc(pb);
    m->loss = og_dalloc(8);
    float *h = (float *)calloc(OG_MLP_NP, 4);
    if (!m->params || !m->grad || !m->m1 || !m->m2 || !m->loss || !h)
    {
        free(h);
        ogpu_mlp_free(m);
        if (!h) og_set_error("out of host memory");
        return -1;
    }
    /* He-uniform weights from SplitMix64 draws, small positive biases */
    uint64_t key = og_stream_key(seed, 7);
    uint64_t ctr = 0;
    for (int i = 0; i < OG_MLP_NP; i++)
    {
        double u = (double)(og_draw(key, ctr++) >> 11) * (1.0 / 9007199254740992.0) * 2.0 - 1.0;
        double lim;
        if (i < OG_MLP_B1) lim = sqrt(6.0 / 4.0);
        else if (i < OG_MLP_W2) lim = 0.0;
        else if (i < OG_MLP_B2) lim = sqrt(6.0 / 64.0);
        else if (i < OG_MLP_W3) lim = 0.0;
        else if (i < OG_MLP_B3) lim = sqrt(3.0 / 64.0);
        else lim = 0.0;
        h[i] = (float)(u * lim);
        if (i >= OG_MLP_B1 && i < OG_MLP_W2) h[i] = 0.01f;
        if (i >= OG_MLP_B2 && i < OG_MLP_W3) h[i] = 0.01f;
    }
    hipStream_t s = og_stream();
    int bad = og_hip_check(hipMemcpyAsync(m->params, h, pb, hipMemcpyHostToDevice, s), "H2D") ||
              og_launch_fill_u32((uint32_t *)m->m1, 0, OG_MLP_NP, s) ||
              og_launch_fill_u32((uint32_t *)m->m2, 0, OG_MLP_NP, s) ||
              og_hip_check(hipStreamSynchronize(s), "sync");
    free(h);
    if (bad) { ogpu_mlp_free(m); return -1; }
    *out = m;
    return 0;
}

int ogpu_mlp_params(ogpu_mlp *m, float *get, const float *set)
{
    hipStream_t s = og_stream();
    if (set && (og_hip_check(hipMemcpyAsync(m->params, set, OG_MLP_NP * 4, hipMemcpyHostToDevice, s), "H2D") ||
                og_launch_fill_u32((uint32_t *)m->m1, 0, OG_MLP_NP, s) ||
                og_launch_fill_u32((uint32_t *)m->m2, 0, OG_MLP_NP, s)))
        return -1;
    if (set) m->step = 0;
    if (get && og_hip_check(hipMemcpyAsync(get, m->params, OG_MLP_NP * 4, hipMemcpyDeviceToHost, s), "D2H"))
        return -1;
    return og_hip_check(hipStreamSynchronize(s), "sync");
}

/* Features of the plan's rows (new ids) into feat (n x 4 floats); uses r as
 * the in-neighbour sum and c_scratch (n floats) for log1p in-degree. */
static int og_plan_features(const og_pr_dev *base, const og_plan *pl, uint32_t n_final, float *c_scratch,
                            float *feat, hipStream_t s)
{
    og_pr_dev a = *base;
    a.g_in = c_scratch;
    a.raw = 1;
    a.dd = 0;
    a.ppr = 0;
    a.allow_stop = 0;
    a.n_parts = og_pull_parts(&a);
    a.n_final_blocks = n_final;
    if (og_launch_fill_u32(a.done, 0, 1, s) || og_launch_feat_li(pl->row_ptr, pl->N, c_scratch, s) ||
        og_launch_pull(&a, 0, OG_NO_STATS, s, NULL, NULL) ||
        og_launch_feat_finish(pl->row_ptr, pl->outdeg, a.r, pl->N, feat, s))
        return -1;
    return 0;
}

/* ---------------- glue with ogpu_pr (og_pagerank.c) ---------------- */

int og_pr_learn_parts(ogpu_pr *p, const og_pr_dev **dev, const og_plan **plan, uint32_t *n_final,
                      float **c1, float **init_d, float **feat, double **scratch);
int og_pr_learn_graph(ogpu_pr *p, og_graph **g, float **tmp_a, float **tmp_b, uint32_t *classes);
void og_pr_use_init(ogpu_pr *p, int accepted);

int ogpu_mlp_add_graph(ogpu_mlp *m, ogpu_graph *g, double tol, uint32_t max_iter, double *l1_out)
{
    /* converged ranks of g: plain pull from the uniform start until the
     * L1 change of the normalized ranks is below tol */
    ogpu_pr_options o;
    ogpu_pr_default_options(&o);
    o.max_iter = max_iter;
    o.epsilon = tol;
    o.stop_rule = 1;
    o.init_mode = 1;
    ogpu_pr *p = NULL;
    if (ogpu_pr_create(g, &o, &p)) return -1;
    ogpu_pr_result res;
    const og_pr_dev *dev;
    const og_plan *pl;
    uint32_t n_final;
    float *c1, *init_d, *feat;
    double *scr;
    int rc = -1;
    if (ogpu_pr_run(p, &res) || og_pr_learn_parts(p, &dev, &pl, &n_final, &c1, &init_d, &feat, &scr)) goto out;
    if (l1_out) *l1_out = res.error_total;
    {
        const uint32_t N = pl->N;
        hipStream_t s = og_stream();
        if (m->n_data + (uint64_t)N > m->cap)
        {
            uint64_t cap = m->cap ? m->cap : 1u << 20;
            while (cap < m->n_data + (uint64_t)N) cap *= 2;
            if (cap > 0xFFFFFFFFu) { og_set_error("training set too large"); goto out; }
            float *nf = og_dalloc(cap * 16), *ny = og_dalloc(cap * 4);
            if (!nf || !ny) { og_dfree(nf); og_dfree(ny); goto out; }
            if (m->n_data &&
                (og_hip_check(hipMemcpyAsync(nf, m->feat, (size_t)m->n_data * 16, hipMemcpyDeviceToDevice, s), "D2D") ||
                 og_hip_check(hipMemcpyAsync(ny, m->y, (size_t)m->n_data * 4, hipMemcpyDeviceToDevice, s), "D2D") ||
                 og_hip_check(hipStreamSynchronize(s), "sync")))
            {
                og_dfree(nf); og_dfree(ny);
                goto out;
            }
            og_dfree(m->feat); og_dfree(m->y);
            m->feat = nf; m->y = ny; m->cap = (uint32_t)cap;
        }
        /* target: relative rank y = N r / sum(r), r the converged
         * unnormalized ranks in the plan's order (what features use) */
        double sum = 0.0;
        float *r_host = (float *)malloc((size_t)N * 4);
        if (!r_host) { og_set_error("out of host memory"); goto out; }
        if (og_hip_check(hipMemcpyAsync(r_host, dev->r, (size_t)N * 4, hipMemcpyDeviceToHost, s), "D2H") ||
            og_hip_check(hipStreamSynchronize(s), "sync"))
        {
            free(r_host);
            goto out;
        }
        for (uint32_t v = 0; v < N; v++) sum += r_host[v];
        const double fac = sum > 0 ? (double)N / sum : 1.0;
        for (uint32_t v = 0; v < N; v++) r_host[v] = (float)(r_host[v] * fac);
        int bad = og_hip_check(hipMemcpyAsync(m->y + m->n_data, r_host, (size_t)N * 4, hipMemcpyHostToDevice, s), "H2D") ||
                  og_plan_features(dev, pl, n_final, c1, m->feat + 4 * (size_t)m->n_data, s) ||
                  og_hip_check(hipStreamSynchronize(s), "sync");
        free(r_host);
        if (bad) goto out;
        m->n_data += N;
    }
    rc = 0;
out:
    ogpu_pr_free(p);
    return rc;
}

uint32_t ogpu_mlp_data_size(const ogpu_mlp *m) { return m->n_data; }

static int og_mlp_ensure_part(ogpu_mlp *m, uint32_t count)
{
    size_t need = og_mlp_part_floats(count);
    if (need <= m->part_floats) return 0;
    og_dfree(m->part);
    m->part = og_dalloc(need * 4);
    if (!m->part) { m->part_floats = 0; return -1; }
    m->part_floats = need;
    return 0;
}

int ogpu_mlp_train(ogpu_mlp *m, uint32_t epochs, uint32_t batch, float lr, uint64_t seed, double *last_loss,
                   double *seconds)
{
    if (!m->n_data || !batch) { og_set_error("ogpu_mlp_train: empty training set or batch"); return -1; }
    hipStream_t s = og_stream();
    if (og_mlp_ensure_part(m, batch)) return -1;
    double t0 = og_learn_now(), loss = 0.0;
    for (uint32_t e = 0; e < epochs; e++)
    {
        const uint64_t es = og_mix64(seed * 0x9E3779B97F4A7C15ull + e + 1) | 1;
        double tot = 0.0;
        for (uint32_t b = 0; b < m->n_data; b += batch)
        {
            const uint32_t cnt = m->n_data - b < batch ? m->n_data - b : batch;
            if (og_launch_mlp_grad(m->params, m->feat, m->y, m->n_data, es, b, cnt, m->part, m->grad, m->loss, s) ||
                og_launch_mlp_adam(m->params, m->grad, m->m1, m->m2, lr, 0.9f, 0.999f, 1e-8f, ++m->step, s))
                return -1;
            double l = 0.0;
            if (og_hip_check(hipMemcpyAsync(&l, m->loss, 8, hipMemcpyDeviceToHost, s), "D2H") ||
                og_hip_check(hipStreamSynchronize(s), "sync"))
                return -1;
            tot += l;
        }
        loss = tot / m->n_data;
    }
    if (last_loss) *last_loss = loss;
    if (seconds) *seconds = og_learn_now() - t0;
    return 0;
}

int ogpu_mlp_grad(ogpu_mlp *m, const float *feat, const float *y, uint32_t n, float *grad, double *loss)
{
    hipStream_t s = og_stream();
    float *df = og_dalloc((size_t)n * 16 + 16), *dy = og_dalloc((size_t)n * 4 + 4);
    int rc = -1;
    if (!df || !dy || og_mlp_ensure_part(m, n)) goto out;
    double l = 0.0;
    if (og_hip_check(hipMemcpyAsync(df, feat, (size_t)n * 16, hipMemcpyHostToDevice, s), "H2D") ||
        og_hip_check(hipMemcpyAsync(dy, y, (size_t)n * 4, hipMemcpyHostToDevice, s), "H2D") ||
        og_launch_mlp_grad(m->params, df, dy, n, 0, 0, n, m->part, m->grad, m->loss, s) ||
        og_hip_check(hipMemcpyAsync(grad, m->grad, OG_MLP_NP * 4, hipMemcpyDeviceToHost, s), "D2H") ||
        og_hip_check(hipMemcpyAsync(&l, m->loss, 8, hipMemcpyDeviceToHost, s), "D2H") ||
        og_hip_check(hipStreamSynchronize(s), "sync"))
        goto out;
    if (loss) *loss = n ? l / n : 0.0;
    rc = 0;
out:
    og_dfree(df); og_dfree(dy);
    return rc;
}

int ogpu_mlp_grad_noalloc(ogpu_mlp *m, const float *d_feat, const float *d_y, uint32_t n, float *d_grad,
                          double *d_loss_sum)
{
    if (og_ctx_ensure()) return -1;
    if (!m || !d_feat || !d_y || !d_grad || !d_loss_sum)
    {
        og_set_error("ogpu_mlp_grad_noalloc: NULL argument");
        return -1;
    }
    return og_launch_mlp_grad_noalloc(og_device(), m->params, d_feat, d_y, n, d_grad, d_loss_sum, og_stream());
}

int ogpu_mlp_forward(ogpu_mlp *m, const float *feat, uint32_t n, int mfma, float *out)
{
    hipStream_t s = og_stream();
    float *df = og_dalloc((size_t)n * 16 + 16), *dout = og_dalloc((size_t)n * 4 + 4);
    int rc = -1;
    if (!df || !dout) goto out;
    if (og_hip_check(hipMemcpyAsync(df, feat, (size_t)n * 16, hipMemcpyHostToDevice, s), "H2D") ||
        og_launch_mlp_infer(m->params, df, n, mfma, dout, s) ||
        og_hip_check(hipMemcpyAsync(out, dout, (size_t)n * 4, hipMemcpyDeviceToHost, s), "D2H") ||
        og_hip_check(hipStreamSynchronize(s), "sync"))
        goto out;
    rc = 0;
out:
    og_dfree(df); og_dfree(dout);
    return rc;
}

int ogpu_pr_features(ogpu_pr *p, float *feat_old)
{
    const og_pr_dev *dev;
    const og_plan *pl;
    uint32_t n_final;
    float *c1, *init_d, *feat;
    double *scr;
    if (og_pr_learn_parts(p, &dev, &pl, &n_final, &c1, &init_d, &feat, &scr)) return -1;
    hipStream_t s = og_stream();
    const uint32_t N = pl->N;
    float *h = (float *)malloc((size_t)N * 16 + 16);
    uint32_t *n2o = (uint32_t *)malloc((size_t)N * 4 + 4);
    int rc = -1;
    if (!h || !n2o) { og_set_error("out of host memory"); goto out; }
    if (og_plan_features(dev, pl, n_final, c1, feat, s) ||
        og_hip_check(hipMemcpyAsync(h, feat, (size_t)N * 16, hipMemcpyDeviceToHost, s), "D2H") ||
        og_hip_check(hipMemcpyAsync(n2o, pl->new2old, (size_t)N * 4, hipMemcpyDeviceToHost, s), "D2H") ||
        og_hip_check(hipStreamSynchronize(s), "sync"))
        goto out;
    for (uint32_t v = 0; v < N; v++) memcpy(feat_old + 4 * (size_t)n2o[v], h + 4 * (size_t)v, 16);
    rc = 0;
out:
    free(h); free(n2o);
    return rc;
}

/* Scale factors of the coarse correction: the classes' mass balance
 * s_i m_i = 0.15 n_i + sum_j A_ij s_j over the classes the start has mass
 * in, A_ij = 0.85 mat_ij / 2^32, by Gaussian elimination with partial
 * pivoting.  diag(m) - A is column diagonally dominant (the iteration keeps
 * at most 0.85 of a class's mass), so the solution is positive; anything
 * else (degenerate input) returns -1 and the start stays uncorrected. */
static int og_wc_solve(uint32_t nc, const unsigned long long *mat, const double *mass, const double *count,
                       float *scale)
{
    double A[OG_WC_MAX][OG_WC_MAX + 1];
    uint32_t idx[OG_WC_MAX], k = 0;
    for (uint32_t i = 0; i < nc; i++)
    {
        scale[i] = 1.0f;
        if (mass[i] > 0.0) idx[k++] = i;
    }
    if (!k) return -1;
    for (uint32_t a = 0; a < k; a++)
    {
        for (uint32_t b = 0; b < k; b++)
            A[a][b] = (a == b ? mass[idx[a]] : 0.0) - 0.85 * (double)mat[idx[a] * nc + idx[b]] / 4294967296.0;
        A[a][k] = 0.15 * count[idx[a]];
    }
    for (uint32_t c = 0; c < k; c++)
    {
        uint32_t piv = c;
        for (uint32_t r = c + 1; r < k; r++)
            if (fabs(A[r][c]) > fabs(A[piv][c])) piv = r;
        if (!(fabs(A[piv][c]) > 0.0)) return -1;
        if (piv != c)
            for (uint32_t q = 0; q <= k; q++)
            {
                double t = A[c][q];
                A[c][q] = A[piv][q];
                A[piv][q] = t;
            }
        for (uint32_t r = c + 1; r < k; r++)
        {
            const double f = A[r][c] / A[c][c];
            for (uint32_t q = c; q <= k; q++) A[r][q] -= f * A[c][q];
        }
    }
    double x[OG_WC_MAX];
    for (int r = (int)k - 1; r >= 0; r--)
    {
        double t = A[r][k];
        for (uint32_t q = (uint32_t)r + 1; q < k; q++) t -= A[r][q] * x[q];
        x[r] = t / A[r][r];
        if (!(x[r] > 0.0 && x[r] < 1e30)) return -1;
    }
    for (uint32_t a = 0; a < k; a++) scale[idx[a]] = (float)x[a];
    return 0;
}

int ogpu_pr_set_init_learned(ogpu_pr *p, ogpu_mlp *m, double *seconds)
{
    const og_pr_dev *dev;
    const og_plan *pl;
    uint32_t n_final;
    float *c1, *init_d, *feat;
    double *scr;
    if (og_pr_learn_parts(p, &dev, &pl, &n_final, &c1, &init_d, &feat, &scr)) return -1;
    hipStream_t s = og_stream();
    double t0 = og_learn_now();
    /* features -> o (bf16 MFMA inference, into c1) -> start ranks (old ids) */
    double ok = 0.0;
    if (og_plan_features(dev, pl, n_final, c1, feat, s) ||
        og_launch_mlp_infer(m->params, feat, pl->N, 1, c1, s) ||
        og_launch_warm_scale(c1, pl->row_ptr, pl->outdeg, pl->new2old, pl->N, scr, init_d, s) ||
        og_hip_check(hipMemcpyAsync(&ok, scr + OG_WARM_OK, 8, hipMemcpyDeviceToHost, s), "D2H") ||
        og_hip_check(hipStreamSynchronize(s), "sync warm start"))
        return -1;
    /* coarse correction of an accepted start over in-degree classes */
    og_graph *g;
    float *tmp_a, *tmp_b;
    uint32_t nin;
    og_pr_learn_graph(p, &g, &tmp_a, &tmp_b, &nin);
    if (ok != 0.0 && nin >= 2)
    {
        const uint32_t nc = 2 * nin;
        unsigned long long mat[OG_WC_MAX * OG_WC_MAX], qmass[OG_WC_MAX], qcount[OG_WC_MAX];
        double mass[OG_WC_MAX], count[OG_WC_MAX];
        float scale[OG_WC_MAX];
        uint8_t *cls = (uint8_t *)tmp_a;
        if (og_launch_wc_sums(init_d, g->f_in_degree, g->f_out_degree, g->N, nin, cls, tmp_b, scr, s) ||
            og_launch_wc_matrix(g->f_src, g->f_dst, g->E, cls, tmp_b, nc,
                                (unsigned long long *)(scr + og_wc_matrix()), s) ||
            og_hip_check(hipMemcpyAsync(mat, scr + og_wc_matrix(), (size_t)nc * nc * 8, hipMemcpyDeviceToHost, s),
                         "D2H") ||
            og_hip_check(hipMemcpyAsync(qmass, scr + og_wc_mass(), (size_t)nc * 8, hipMemcpyDeviceToHost, s), "D2H") ||
            og_hip_check(hipMemcpyAsync(qcount, scr + og_wc_count(), (size_t)nc * 8, hipMemcpyDeviceToHost, s),
                         "D2H") ||
            og_hip_check(hipStreamSynchronize(s), "sync warm correction"))
            return -1;
        for (uint32_t i = 0; i < nc; i++)
        {
            mass[i] = (double)qmass[i] / 4294967296.0;
            count[i] = (double)qcount[i];
        }
        if (!og_wc_solve(nc, mat, mass, count, scale) &&
            (og_hip_check(hipMemcpyAsync(scr + og_wc_scale(), scale, (size_t)nc * 4, hipMemcpyHostToDevice, s),
                          "H2D") ||
             og_launch_wc_apply(init_d, cls, g->N, (const float *)(scr + og_wc_scale()), s) ||
             og_hip_check(hipStreamSynchronize(s), "sync warm correction")))
            return -1;
    }
    /* a rejected prediction (non-finite, negative, no mass) leaves the
     * uniform start in place */
    og_pr_use_init(p, ok != 0.0);
    if (seconds) *seconds = og_learn_now() - t0;
    return 0;
}
