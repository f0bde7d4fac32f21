/*
 * og_bin.c -- one-time build of the binned pull's edge layout (og_bin.h)
 * from a PageRank plan whose rows keep the reference's in-neighbour order.
 *
 * Host: the rows are packed into bins (from the plan's row offsets), the
 * rows of more than OG_BIN_CAP in-edges are cut into pieces, and, once the
 * device has sorted the edges into tiles, the tiles are ordered segment by
 * segment (the gather's source order) and every quad of 16 sources gets
 * its run of entries.
 * Device: the (bin, segment) keys, their stable radix sort, the entries'
 * source and slot indices, and the tile boundaries.
 */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "og_bin.h"
#include "og_internal.h"
#include "og_plan.h"

#define OG_BIN_ZERO_ROWS (OG_BIN_RPT * OG_BIN_THREADS)   /* rows of a bin of rows without in-edges */

void og_bin_plan_destroy(og_bin_plan *b)
{
    if (!b) return;
    og_dfree(b->mem_items); og_dfree(b->mem_qent); og_dfree(b->mem_src);
    og_dfree(b->mem_pos); og_dfree(b->mem_val); og_dfree(b->mem_bins); og_dfree(b->mem_long_row);
    og_dfree(b->mem_long_bin); og_dfree(b->mem_ticket);
    free(b);
}

/* Sum and long-row blocks take half a CU's LDS each: two block slots per
 * CU (a long-row block has twice the threads at half the registers).  The
 * long rows get up to one slot in eight, the sum pass the rest, side by
 * side.  Workgroups are dealt to the XCDs in turn, so the sum pass leaves a
 * whole number of slots per XCD free (the long-row blocks rounded up to a
 * multiple of OG_BIN_XCDS); otherwise a long-row block can find its XCD
 * full and wait for a sum block to end. */
#define OG_BIN_XCDS 8u

uint32_t og_bin_long_grid(const og_bin_plan *b)
{
    /* the long rows' share of the slots: their in-edges' share of E, times
     * 1.6 (an in-edge costs a long-row block about 1.6 times what it costs
     * the sum pass: R-MAT-24 measurements), in whole XCD rounds, between
     * OG_BIN_XCDS and a quarter of the CUs */
    const uint32_t cus = (uint32_t)og_num_cus();
    const uint32_t cap = cus / 4 ? cus / 4 : 1;
    uint64_t l = b->n_edges ? (uint64_t)(2.0 * cus * 1.6 * (double)b->long_terms / (double)b->n_edges) : 0;
    l = (l + OG_BIN_XCDS - 1) / OG_BIN_XCDS * OG_BIN_XCDS;
    if (l < OG_BIN_XCDS) l = OG_BIN_XCDS;
    if (l > cap) l = cap;
    return b->d.n_long < l ? b->d.n_long : (uint32_t)l;
}

uint32_t og_bin_sum_grid(const og_bin_plan *b)
{
    const uint32_t slots = 2u * (uint32_t)og_num_cus();   /* two blocks per CU */
    const uint32_t l = (og_bin_long_grid(b) + OG_BIN_XCDS - 1) / OG_BIN_XCDS * OG_BIN_XCDS;
    return slots > l ? slots - l : 1;
}

/* growable host arrays */
typedef struct
{
    void *p;
    size_t n, cap, elem;
} og_vec;

static void *og_vec_push(og_vec *v)
{
    if (v->n == v->cap)
    {
        size_t c = v->cap ? 2 * v->cap : 1024;
        void *q = realloc(v->p, c * v->elem);
        if (!q) return NULL;
        v->p = q;
        v->cap = c;
    }
    return (char *)v->p + (v->n++) * v->elem;
}

static int og_bits(uint64_t x)
{
    int b = 0;
    while (b < 64 && (1ull << b) <= x) b++;
    return b;
}

/* Device copy of a host array (NULL and error set on failure). */
static void *og_bin_upload(const void *h, size_t bytes)
{
    void *d = og_dalloc(bytes ? bytes : 16);
    if (!d) return NULL;
    if (bytes && og_hip_check(og_copy(d, h, bytes, hipMemcpyHostToDevice), "H2D bin layout"))
    {
        og_dfree(d);
        return NULL;
    }
    return d;
}

/* Host check of every index the binned kernels derive from the layout
 * (enabled by the environment variable OG_BIN_VERIFY): items within the
 * sources, every source's entry within the entries, each source within its
 * segment, each entry gathered once and its source the edge's (the CSR's
 * column at the entry's slot),
 * slots within their bins, bins and pieces within the entries, rows per bin,
 * CSR positions, the pieces first and each long row's pieces covering its
 * in-edges in order.  Runs before any binned kernel, so a bad layout
 * is an error, not a stray access. */
static int og_bin_verify(const og_bin_plan *b, const og_plan *pl, const uint32_t *rp)
{
    const og_bin_dev *d = &b->d;
    const size_t ne = (size_t)b->n_ent, nq = (ne - OG_BIN_SPARE) / OG_BIN_TILE_ALIGN;
    og_bin_item *items = (og_bin_item *)malloc((size_t)d->n_items * sizeof(og_bin_item) + 16);
    uint32_t *qent = (uint32_t *)malloc(nq * 4 + 4);
    uint16_t *src = (uint16_t *)malloc(ne * 2), *pos = (uint16_t *)malloc(ne * 2);
    og_bin_desc *bins = (og_bin_desc *)malloc((size_t)d->n_bins * sizeof(og_bin_desc) + 16);
    uint32_t *lrow = (uint32_t *)malloc((size_t)d->n_long * 4 + 4), *lbin = (uint32_t *)malloc((size_t)d->n_long * 4 + 4);
    uint32_t *gsrc = (uint32_t *)malloc(ne * 4);   /* source id gathered into each entry */
    uint32_t *cols = (uint32_t *)malloc((size_t)pl->E * 4 + 4);
    uint8_t *seen = (uint8_t *)calloc(ne, 1);
    int rc = -1;
#define OG_BIN_FAIL(...) do { og_set_error("binned layout check: " __VA_ARGS__); goto out; } while (0)
    if (!items || !qent || !src || !pos || !bins || !lrow || !lbin || !seen || !gsrc || !cols)
        OG_BIN_FAIL("out of host memory");
    if (og_hip_check(og_copy(items, d->items, (size_t)d->n_items * sizeof(og_bin_item), hipMemcpyDeviceToHost), "D2H") ||
        og_hip_check(og_copy(qent, d->qent, nq * 4, hipMemcpyDeviceToHost), "D2H") ||
        og_hip_check(og_copy(src, d->src, ne * 2, hipMemcpyDeviceToHost), "D2H") ||
        og_hip_check(og_copy(pos, d->pos, ne * 2, hipMemcpyDeviceToHost), "D2H") ||
        og_hip_check(og_copy(bins, d->bins, (size_t)d->n_bins * sizeof(og_bin_desc), hipMemcpyDeviceToHost), "D2H") ||
        og_hip_check(og_copy(lrow, d->long_row, (size_t)d->n_long * 4, hipMemcpyDeviceToHost), "D2H") ||
        og_hip_check(og_copy(lbin, d->long_bin, (size_t)d->n_long * 4, hipMemcpyDeviceToHost), "D2H") ||
        og_hip_check(og_copy(cols, pl->cols, (size_t)pl->E * 4, hipMemcpyDeviceToHost), "D2H"))
        goto out;
    for (uint32_t q = 0; q < d->n_items; q++)
    {
        const og_bin_item it = items[q];
        if (it.seg >= d->n_seg || it.f0 >= it.f1 || it.f1 > nq * OG_BIN_TILE_ALIGN || it.f0 % OG_BIN_TILE_ALIGN ||
            it.f1 % OG_BIN_TILE_ALIGN)
            OG_BIN_FAIL("item %u", q);
        const uint32_t nv = d->N - it.seg * OG_BIN_SEG < OG_BIN_SEG ? d->N - it.seg * OG_BIN_SEG : OG_BIN_SEG;
        for (uint32_t f = it.f0; f < it.f1; f++)
        {
            const uint64_t e = (uint64_t)qent[f / OG_BIN_TILE_ALIGN] + f % OG_BIN_TILE_ALIGN;
            if (e >= ne - OG_BIN_SPARE)
                OG_BIN_FAIL("source %u: entry %llu of %zu", f, (unsigned long long)e, ne - OG_BIN_SPARE);
            if (src[f] >= nv) OG_BIN_FAIL("source %u: %u past segment size %u", f, src[f], nv);
            if (seen[e]) OG_BIN_FAIL("entry %llu gathered twice", (unsigned long long)e);
            seen[e] = 1;
            gsrc[e] = it.seg * OG_BIN_SEG + src[f];
        }
    }
    uint64_t prev_end = 0;
    for (uint32_t q = 0; q < d->n_bins; q++)
    {
        const og_bin_desc bd = bins[q];
        if (bd.ent0 != prev_end || (bd.ent0 % OG_BIN_TILE_ALIGN) || bd.ent1 < bd.ent0 ||
            (bd.ent1 - bd.ent0) % OG_BIN_TILE_ALIGN || bd.ent1 - bd.ent0 < bd.cnt || bd.ent1 > ne - OG_BIN_SPARE ||
            bd.cnt > OG_BIN_CAP)
            OG_BIN_FAIL("bin %u entries [%u, %u) count %u", q, bd.ent0, bd.ent1, bd.cnt);
        prev_end = bd.ent1;
        if ((uint64_t)bd.pos0 + bd.cnt > pl->E) OG_BIN_FAIL("bin %u positions", q);
        if ((bd.kind == OG_BIN_PIECE) != (q < d->first_bin)) OG_BIN_FAIL("bin %u: pieces come first", q);
        if (bd.kind == OG_BIN_PIECE)
        {
            if (bd.row0 >= pl->Z || bd.pos0 < rp[bd.row0] || (uint64_t)bd.pos0 + bd.cnt > rp[bd.row0 + 1])
                OG_BIN_FAIL("piece %u", q);
        }
        else
        {
            if (bd.row0 > bd.row1 || bd.row1 > d->N || bd.row1 - bd.row0 > OG_BIN_RPT * OG_BIN_THREADS)
                OG_BIN_FAIL("bin %u rows", q);
            if (bd.cnt && (rp[bd.row0] != bd.pos0 || rp[bd.row1] - rp[bd.row0] != bd.cnt)) OG_BIN_FAIL("bin %u row offsets", q);
        }
        uint32_t real = 0;
        for (uint32_t e = bd.ent0; e < bd.ent1; e++)
        {
            if (pos[e] != OG_BIN_PAD)
            {
                if (pos[e] >= bd.cnt) OG_BIN_FAIL("entry %u: slot %u in bin %u of %u", e, pos[e], q, bd.cnt);
                if (seen[e] && gsrc[e] != cols[bd.pos0 + pos[e]])
                    OG_BIN_FAIL("entry %u: gathers source %u, the edge's is %u", e, gsrc[e], cols[bd.pos0 + pos[e]]);
                real++;
            }
            if (!seen[e]) OG_BIN_FAIL("entry %u never gathered", e);
        }
        if (real != bd.cnt) OG_BIN_FAIL("bin %u: %u entries for %u in-edges", q, real, bd.cnt);
    }
    uint64_t pieces = 0;
    for (uint32_t q = 0; q < d->n_long; q++)
    {
        const uint32_t r = lrow[q];
        if (r >= pl->Z) OG_BIN_FAIL("long row %u", q);
        const uint32_t n = rp[r + 1] - rp[r];
        for (uint32_t c0 = 0, k = lbin[q]; c0 < n; c0 += OG_BIN_CAP, k++, pieces++)
            if (k >= d->first_bin || bins[k].kind != OG_BIN_PIECE || bins[k].row0 != r || bins[k].pos0 != rp[r] + c0 ||
                bins[k].cnt != (n - c0 < OG_BIN_CAP ? n - c0 : OG_BIN_CAP))
                OG_BIN_FAIL("long row %u: piece %u", q, k);
    }
    if (pieces != d->first_bin) OG_BIN_FAIL("%llu pieces for %u piece bins", (unsigned long long)pieces, d->first_bin);
    rc = 0;
#undef OG_BIN_FAIL
out:
    free(items); free(qent); free(src); free(pos); free(bins); free(lrow); free(lbin); free(seen);
    free(gsrc); free(cols);
    return rc;
}

int og_bin_plan_build(const og_plan *pl, og_bin_plan **out)
{
    *out = NULL;
    const uint32_t N = pl->N, E = pl->E, Z = pl->Z, H = pl->H;
    if (pl->relabel == 1 && H)
    {
        og_set_error("binned pull: the plan sorts long rows by source (relabel 1); it needs CSR order");
        return -1;
    }
    if (!E || !N)
    {
        og_set_error("binned pull: graph without edges");
        return -1;
    }
    og_bin_plan *b = (og_bin_plan *)calloc(1, sizeof(og_bin_plan));
    uint32_t *rp = (uint32_t *)malloc(((size_t)N + 1) * 4);
    og_vec bins = {NULL, 0, 0, sizeof(og_bin_desc)};
    uint32_t *long_rows = NULL, *long_list = NULL, *long_bin_rowordered = NULL, *long_bin = NULL;
    uint32_t *tseg = NULL, *tbin = NULL, *tstart = NULL, *tcnt = NULL, *tent = NULL, *order = NULL;
    uint32_t *seg_cnt = NULL, *tsrc = NULL, *seg_src = NULL, *qent = NULL;
    og_vec items = {NULL, 0, 0, sizeof(og_bin_item)};
    uint32_t *key_a = NULL, *val_a = NULL, *key_b = NULL, *val_b = NULL;
    void *scratch = NULL, *d_tseg = NULL, *d_tbin = NULL, *d_tstart = NULL, *d_tent = NULL, *d_tcnt = NULL;
    void *d_tsrc = NULL;
    hipStream_t s = og_stream();
    int rc = -1;
    if (!b || !rp)
    {
        og_set_error("out of host memory");
        goto done;
    }
    if (og_hip_check(og_copy(rp, pl->row_ptr, ((size_t)N + 1) * 4, hipMemcpyDeviceToHost), "D2H row_ptr"))
        goto done;

    {
        /* ---- long rows, longest first; their pieces first in the bin
         * list (bins of OG_BIN_CAP in-edges each, in CSR order) ---- */
        uint32_t n_long = 0;
        for (uint32_t r = 0; r < Z; r++)
            if (rp[r + 1] - rp[r] > OG_BIN_CAP) n_long++;
        long_rows = (uint32_t *)malloc((size_t)(n_long + 1) * 4);
        long_list = (uint32_t *)malloc((size_t)(n_long + 1) * 4);
        long_bin_rowordered = (uint32_t *)malloc((size_t)(n_long + 1) * 4);
        long_bin = (uint32_t *)malloc((size_t)(n_long + 1) * 4);
        if (!long_rows || !long_list || !long_bin_rowordered || !long_bin)
        {
            og_set_error("out of host memory");
            goto done;
        }
        uint32_t k = 0;
        for (uint32_t r = 0; r < Z; r++)
            if (rp[r + 1] - rp[r] > OG_BIN_CAP)
            {
                const uint32_t deg = rp[r + 1] - rp[r];
                long_bin_rowordered[k] = (uint32_t)bins.n;
                for (uint32_t p0 = 0; p0 < deg; p0 += OG_BIN_CAP)
                {
                    og_bin_desc *d = (og_bin_desc *)og_vec_push(&bins);
                    if (!d) { og_set_error("out of host memory"); goto done; }
                    memset(d, 0, sizeof(*d));
                    d->row0 = r;
                    d->row1 = k;
                    d->pos0 = rp[r] + p0;
                    d->kind = OG_BIN_PIECE;
                    d->cnt = deg - p0 < OG_BIN_CAP ? deg - p0 : OG_BIN_CAP;
                }
                long_rows[k++] = r;
            }
        b->d.first_bin = (uint32_t)bins.n;
        b->n_edges = E;
        for (uint32_t i = 0; i < n_long; i++) b->long_terms += rp[long_rows[i] + 1] - rp[long_rows[i]];
        uint32_t *lord = (uint32_t *)malloc((size_t)(n_long + 1) * 4);
        if (!lord) { og_set_error("out of host memory"); goto done; }
        for (uint32_t i = 0; i < n_long; i++) lord[i] = i;
        for (uint32_t i = 1; i < n_long; i++)   /* insertion sort: few long rows */
        {
            uint32_t x = lord[i], dx = rp[long_rows[x] + 1] - rp[long_rows[x]];
            int j = (int)i - 1;
            while (j >= 0 && rp[long_rows[lord[j]] + 1] - rp[long_rows[lord[j]]] < dx)
            {
                lord[j + 1] = lord[j];
                j--;
            }
            lord[j + 1] = x;
        }
        for (uint32_t i = 0; i < n_long; i++)
        {
            long_list[i] = long_rows[lord[i]];
            long_bin[i] = long_bin_rowordered[lord[i]];
        }
        free(lord);

        /* ---- bins: consecutive rows of one kind (thread rows, wavefront
         * rows of more than OG_BIN_WAVE_ROW in-edges) with at most
         * OG_BIN_CAP in-edges and OG_BIN_RPT * OG_BIN_THREADS rows together
         * (the long rows have their pieces already) ---- */
        uint32_t r = 0;
        while (r < Z)
        {
            const uint32_t deg = rp[r + 1] - rp[r];
            if (deg > OG_BIN_CAP)
            {
                r++;
                continue;
            }
            const uint32_t kind = deg > OG_BIN_WAVE_ROW ? OG_BIN_WAVE_ROWS : OG_BIN_THREAD_ROWS;
            const uint32_t r0 = r;
            uint32_t tot = 0;
            while (r < Z)
            {
                const uint32_t dr = rp[r + 1] - rp[r];
                const uint32_t kr = dr > OG_BIN_WAVE_ROW ? OG_BIN_WAVE_ROWS : OG_BIN_THREAD_ROWS;
                if (kr != kind || dr > OG_BIN_CAP || tot + dr > OG_BIN_CAP || r - r0 == OG_BIN_RPT * OG_BIN_THREADS)
                    break;
                tot += dr;
                r++;
            }
            og_bin_desc *d = (og_bin_desc *)og_vec_push(&bins);
            if (!d) { og_set_error("out of host memory"); goto done; }
            memset(d, 0, sizeof(*d));
            d->row0 = r0;
            d->row1 = r;
            d->pos0 = rp[r0];
            d->kind = kind;
            d->cnt = tot;
        }
        const uint32_t n_edge_bins = (uint32_t)bins.n;
        /* rows without in-edges: epilogue-only bins, last */
        for (uint32_t r0 = Z; r0 < N; r0 += OG_BIN_ZERO_ROWS)
        {
            og_bin_desc *d = (og_bin_desc *)og_vec_push(&bins);
            if (!d) { og_set_error("out of host memory"); goto done; }
            memset(d, 0, sizeof(*d));
            d->row0 = r0;
            d->row1 = N - r0 < OG_BIN_ZERO_ROWS ? N : r0 + OG_BIN_ZERO_ROWS;
            d->pos0 = rp[r0];
            d->kind = OG_BIN_THREAD_ROWS;
        }
        b->n_zero_bins = (uint32_t)bins.n - n_edge_bins;
        const uint32_t n_seg = (N + OG_BIN_SEG - 1) / OG_BIN_SEG;
        if ((uint64_t)n_edge_bins * n_seg > 0xFFFFFFFFull)
        {
            og_set_error("binned pull: graph too large for 32-bit tile keys");
            goto done;
        }
        b->d.N = N;
        b->d.n_seg = n_seg;
        b->d.n_bins = (uint32_t)bins.n;
        b->d.n_bins_live = (uint32_t)bins.n;
        b->d.n_long = n_long;

        /* ---- device: (bin, segment) keys of the CSR positions, stable sort,
         * tile starts ---- */
        og_bin_desc *hb = (og_bin_desc *)bins.p;
        b->mem_bins = og_bin_upload(bins.p, bins.n * sizeof(og_bin_desc));
        key_a = og_dalloc(((size_t)E + 1) * 4);
        val_a = og_dalloc(((size_t)E + 1) * 4);
        key_b = og_dalloc(((size_t)E + 1) * 4);
        val_b = og_dalloc(((size_t)E + 1) * 4);
        size_t sb = og_radix_scratch_bytes(E), sc2 = og_scan_scratch_bytes(E);
        scratch = og_dalloc(sb > sc2 ? sb : sc2);
        if (!b->mem_bins || !key_a || !val_a || !key_b || !val_b || !scratch) goto done;
        og_bin_desc *d_bins = (og_bin_desc *)b->mem_bins;
        if (og_launch_bin_keys(d_bins, n_edge_bins, pl->cols, n_seg, key_b, val_b, s)) goto done;
        int in_b = 0;
        const uint32_t kbits = (uint32_t)og_bits((uint64_t)n_edge_bins * n_seg - 1);
        if (og_radix_sort_pairs(key_b, val_b, key_a, val_a, key_b, val_b, E, kbits ? kbits : 1, scratch, s, &in_b))
            goto done;
        uint32_t *K = in_b ? key_b : key_a, *V = in_b ? val_b : val_a;
        uint32_t *flag = in_b ? key_a : key_b, *fscan = in_b ? val_a : val_b;
        if (og_launch_bin_flags(K, E, d_bins, n_seg, flag, s) || og_scan_exclusive(flag, fscan, E, scratch, s))
            goto done;
        uint32_t n_tiles = 0;
        if (og_hip_check(og_copy(&n_tiles, fscan + E, 4, hipMemcpyDeviceToHost), "D2H tiles")) goto done;
        b->n_tiles = n_tiles;
        d_tseg = og_dalloc((size_t)n_tiles * 4);
        d_tbin = og_dalloc((size_t)n_tiles * 4);
        d_tstart = og_dalloc((size_t)n_tiles * 4);
        if (!d_tseg || !d_tbin || !d_tstart) goto done;
        if (og_launch_bin_tiles(K, E, n_seg, fscan, (uint32_t *)d_tseg, (uint32_t *)d_tbin, (uint32_t *)d_tstart, s))
            goto done;
        tseg = (uint32_t *)malloc((size_t)n_tiles * 4 + 4);
        tbin = (uint32_t *)malloc((size_t)n_tiles * 4 + 4);
        tstart = (uint32_t *)malloc((size_t)n_tiles * 4 + 4);
        tcnt = (uint32_t *)malloc((size_t)n_tiles * 4 + 4);
        tent = (uint32_t *)malloc((size_t)n_tiles * 4 + 4);
        order = (uint32_t *)malloc((size_t)n_tiles * 4 + 4);
        seg_cnt = (uint32_t *)calloc((size_t)n_seg + 1, 4);
        tsrc = (uint32_t *)malloc((size_t)n_tiles * 4 + 4);
        seg_src = (uint32_t *)malloc(((size_t)n_seg + 1) * 4);
        if (!tseg || !tbin || !tstart || !tcnt || !tent || !order || !seg_cnt || !tsrc || !seg_src)
        {
            og_set_error("out of host memory");
            goto done;
        }
        if (og_hip_check(og_copy(tseg, d_tseg, (size_t)n_tiles * 4, hipMemcpyDeviceToHost), "D2H tile seg") ||
            og_hip_check(og_copy(tbin, d_tbin, (size_t)n_tiles * 4, hipMemcpyDeviceToHost), "D2H tile bin") ||
            og_hip_check(og_copy(tstart, d_tstart, (size_t)n_tiles * 4, hipMemcpyDeviceToHost), "D2H tile start"))
            goto done;

        /* ---- host: tile sizes, their padded entry ranges, bins' entries ---- */
        uint64_t ent = 0;
        uint32_t bprev = 0xFFFFFFFFu;
        for (uint32_t t = 0; t < n_tiles; t++)
        {
            tcnt[t] = (t + 1 < n_tiles ? tstart[t + 1] : E) - tstart[t];
            if (tbin[t] != bprev)
            {
                if (bprev != 0xFFFFFFFFu) hb[bprev].ent1 = (uint32_t)ent;
                bprev = tbin[t];
                hb[bprev].ent0 = (uint32_t)ent;
            }
            tent[t] = (uint32_t)ent;
            ent += (tcnt[t] + OG_BIN_TILE_ALIGN - 1) / OG_BIN_TILE_ALIGN * OG_BIN_TILE_ALIGN;
            if (ent + OG_BIN_SPARE > 0xFFFFFFFFull)
            {
                og_set_error("binned pull: graph too large for 32-bit entries");
                goto done;
            }
        }
        if (bprev != 0xFFFFFFFFu) hb[bprev].ent1 = (uint32_t)ent;
        for (uint32_t q = n_edge_bins; q < bins.n; q++) hb[q].ent0 = hb[q].ent1 = (uint32_t)ent;
        b->n_ent = ent + OG_BIN_SPARE;   /* spare entries past the last tile */
        b->d.n_ent = (uint32_t)ent;
        /* the tiles segment by segment (bins ascending), and each tile's
         * first entry in that order: where the gather reads its sources */
        for (uint32_t t = 0; t < n_tiles; t++) seg_cnt[tseg[t] + 1]++;
        for (uint32_t q = 0; q < n_seg; q++) seg_cnt[q + 1] += seg_cnt[q];
        {
            uint32_t *fill = (uint32_t *)malloc(((size_t)n_seg + 1) * 4);
            if (!fill) { og_set_error("out of host memory"); goto done; }
            memcpy(fill, seg_cnt, ((size_t)n_seg + 1) * 4);
            for (uint32_t t = 0; t < n_tiles; t++) order[fill[tseg[t]]++] = t;
            free(fill);
        }
        {
            uint64_t f = 0;
            for (uint32_t q = 0; q < n_seg; q++)
            {
                seg_src[q] = (uint32_t)f;
                for (uint32_t i = seg_cnt[q]; i < seg_cnt[q + 1]; i++)
                {
                    tsrc[order[i]] = (uint32_t)f;
                    f += (tcnt[order[i]] + OG_BIN_TILE_ALIGN - 1) / OG_BIN_TILE_ALIGN * OG_BIN_TILE_ALIGN;
                }
            }
            seg_src[n_seg] = (uint32_t)f;   /* == ent */
        }
        b->mem_src = og_dalloc(b->n_ent * 2);
        b->mem_pos = og_dalloc(b->n_ent * 2);
        b->mem_val = og_dalloc(b->n_ent * 4);
        b->mem_long_row = og_bin_upload(long_list, (size_t)n_long * 4);
        b->mem_long_bin = og_bin_upload(long_bin, (size_t)n_long * 4);
        b->mem_ticket = og_dalloc(256);
        d_tent = og_bin_upload(tent, (size_t)n_tiles * 4);
        d_tcnt = og_bin_upload(tcnt, (size_t)n_tiles * 4);
        d_tsrc = og_bin_upload(tsrc, (size_t)n_tiles * 4);
        if (!b->mem_src || !b->mem_pos || !b->mem_val || !b->mem_long_row || !b->mem_long_bin || !b->mem_ticket ||
            !d_tent || !d_tcnt || !d_tsrc ||
            og_hip_check(og_copy(d_bins, hb, bins.n * sizeof(og_bin_desc), hipMemcpyHostToDevice), "H2D bins"))
            goto done;
        if (og_launch_bin_fill(K, V, E, d_bins, n_seg, pl->cols, fscan, (const uint32_t *)d_tstart,
                               (const uint32_t *)d_tent, (const uint32_t *)d_tsrc, (uint16_t *)b->mem_src,
                               (uint16_t *)b->mem_pos, s) ||
            og_launch_bin_pads((const uint32_t *)d_tent, (const uint32_t *)d_tsrc, (const uint32_t *)d_tcnt, n_tiles,
                               (uint16_t *)b->mem_src, (uint16_t *)b->mem_pos, s) ||
            og_hip_check(hipMemsetAsync((uint16_t *)b->mem_src + ent, 0, OG_BIN_SPARE * 2, s), "memset slack") ||
            og_hip_check(hipStreamSynchronize(s), "sync bin layout"))
            goto done;
        og_dfree(key_a); og_dfree(val_a); og_dfree(key_b); og_dfree(val_b); og_dfree(scratch);
        key_a = val_a = key_b = val_b = NULL;
        scratch = NULL;

        /* ---- host: each quad's first entry (a tile's padded sources are
         * whole quads), and the work items: each segment's sources in runs
         * of at most OG_BIN_ITEM_SRC ---- */
        const uint64_t n_quads = ent / OG_BIN_TILE_ALIGN;
        qent = (uint32_t *)malloc(n_quads * 4 + 4);
        if (!qent) { og_set_error("out of host memory"); goto done; }
        for (uint32_t t = 0; t < n_tiles; t++)
        {
            const uint32_t P = (tcnt[t] + OG_BIN_TILE_ALIGN - 1) / OG_BIN_TILE_ALIGN * OG_BIN_TILE_ALIGN;
            for (uint32_t k = 0; k < P; k += OG_BIN_TILE_ALIGN) qent[(tsrc[t] + k) / OG_BIN_TILE_ALIGN] = tent[t] + k;
        }
        for (uint32_t q = 0; q < n_seg; q++)
            for (uint32_t f0 = seg_src[q]; f0 < seg_src[q + 1]; f0 += OG_BIN_ITEM_SRC)
            {
                og_bin_item *it = (og_bin_item *)og_vec_push(&items);
                if (!it) { og_set_error("out of host memory"); goto done; }
                it->seg = q;
                it->f0 = f0;
                it->f1 = seg_src[q + 1] - f0 < OG_BIN_ITEM_SRC ? seg_src[q + 1] : f0 + OG_BIN_ITEM_SRC;
                it->pad = 0;
            }
        b->mem_qent = og_bin_upload(qent, n_quads * 4);
        b->mem_items = og_bin_upload(items.p, items.n * sizeof(og_bin_item));
        if (!b->mem_qent || !b->mem_items) goto done;
        b->d.n_items = (uint32_t)items.n;
        b->d.items = (const og_bin_item *)b->mem_items;
        b->d.qent = (const uint32_t *)b->mem_qent;
        b->d.src = (const uint16_t *)b->mem_src;
        b->d.pos = (const uint16_t *)b->mem_pos;
        b->d.val = (float *)b->mem_val;
        b->d.bins = d_bins;
        b->d.long_row = (const uint32_t *)b->mem_long_row;
        b->d.long_bin = (const uint32_t *)b->mem_long_bin;
        b->d.ticket = (uint32_t *)b->mem_ticket;
        if (getenv("OG_BIN_VERIFY") && og_bin_verify(b, pl, rp)) goto done;
    }
    rc = 0;
done:
    (void)hipStreamSynchronize(s);
    og_dfree(key_a); og_dfree(val_a); og_dfree(key_b); og_dfree(val_b); og_dfree(scratch);
    og_dfree(d_tseg); og_dfree(d_tbin); og_dfree(d_tstart); og_dfree(d_tent); og_dfree(d_tcnt); og_dfree(d_tsrc);
    free(rp); free(bins.p); free(long_rows); free(long_list); free(long_bin_rowordered); free(long_bin);
    free(tseg); free(tbin); free(tstart); free(tcnt); free(tent); free(order); free(seg_cnt);
    free(items.p); free(tsrc); free(seg_src); free(qent);
    if (rc) og_bin_plan_destroy(b);
    else *out = b;
    return rc;
}
