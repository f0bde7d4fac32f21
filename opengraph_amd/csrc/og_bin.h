/*
 * og_bin.h -- the binned fp32 pull (og_bin.hip, og_bin.c): pageRank.c:998-1011
 * as two streaming passes over an edge layout built once per graph.
 *
 * Rows (new ids, the plan's CSR order kept) are packed into bins whose
 * in-edges fit in one CU's LDS (OG_BIN_CAP values); sources fall in
 * segments of OG_BIN_SEG consecutive ids (their contributions fit in LDS).
 * The edges of bin b whose source lies in segment s form tile (b, s).
 * Entries (one per edge) are stored bin by bin, within a bin tile by tile
 * in segment order, within a tile in CSR order:
 *   pos[e]  u16  the edge's CSR position minus the bin's first position
 *               (OG_BIN_PAD: a padding entry; each tile is padded to
 *               OG_BIN_TILE_ALIGN entries, so every tile's values fill whole
 *               64-byte sectors, written by one block)
 *   val[e]  f32  the gathered contribution, written each iteration
 * The sources are stored in the gather's order instead -- segment by
 * segment, within a segment tile by tile (bins ascending), each tile padded
 * the same way -- so every aligned run of 16 of them (a quad) lies in one
 * tile, and its entries are one run of 16:
 *   src[f]       u16  source id minus the segment's first id
 *   qent[f / 16] u32  the entry of the quad's first source
 * One iteration:
 *   gather pass  one 1024-thread block per CU: per work item (a segment and
 *                a run of its sources), the segment's contributions are
 *                copied into LDS, then val[qent[f / 16] + f % 16] =
 *                lds[src[f]] for the item's sources, 4 consecutive ones
 *                per lane (one 8-byte source load, one quad entry, one
 *                16-byte value store), 256 per wavefront and step.
 *   sum pass     one 1024-thread block per CU: per bin, val[e] is written
 *                to LDS slot pos[e] (the bin's in-edge values in CSR order),
 *                then every row is added from its slots in order -- the
 *                reference's running sum -- by a thread (rows of at most
 *                OG_BIN_WAVE_ROW in-edges) or a wavefront (longer rows, an
 *                exact fast-forward over binades), and the update runs.
 *                The next bin's entries and row metadata are loaded into
 *                registers while a bin's rows are added.
 *   long rows    rows with more than OG_BIN_CAP in-edges are cut into
 *                pieces of OG_BIN_CAP (bins of their own, first in the bin
 *                list); one block per long row, on CUs of its own beside
 *                the sum pass, puts each piece in LDS as the sum pass puts
 *                a bin and adds them in order (exact fast-forward,
 *                block-wide).
 * Every row is therefore summed in the reference's order: the fp32 pull is
 * bit-exact, and every global access is a streaming one.
 */
#ifndef OG_BIN_H
#define OG_BIN_H

#include <stddef.h>
#include <stdint.h>

#include <hip/hip_runtime_api.h>

#include "og_internal.h"
#include "og_pr.h"

#ifdef __cplusplus
extern "C" {
#endif

#define OG_BIN_SEG 39936u      /* sources per segment: 156 KB of fp32 in LDS */
#define OG_BIN_CAP 19968u      /* in-edge values per bin: 78 KB of LDS, two bins per CU */
#define OG_BIN_PAD 0xFFFFu     /* pos of a padding entry */
#define OG_BIN_WAVE_ROW 4096u  /* rows with more in-edges are added by a wavefront */
#define OG_BIN_RPT 1u          /* rows per thread and bin: a bin holds at most RPT * THREADS rows */
#define OG_BIN_TILE_ALIGN 16u  /* tiles padded to 16 entries: 64-byte value runs */
#define OG_BIN_ITEM_SRC (2048u * 64u) /* sources per gather work item at most */
#define OG_BIN_SPARE 256u      /* spare entries past the last tile (the gather's idle stores) */
#define OG_BIN_THREADS 512u    /* sum and long-row blocks (two per CU) */
#define OG_BIN_GTHREADS 1024u  /* gather blocks (one per CU: a segment fills LDS) */
#define OG_BIN_LTHREADS 1024u  /* long-row blocks (a whole CU each: two sum-block slots) */

enum { OG_BIN_THREAD_ROWS = 0, OG_BIN_WAVE_ROWS = 1, OG_BIN_PIECE = 2 };

/* One bin: rows [row0, row1) (a piece: the long row row0 and row1 = the
 * index of that long row in row order), entries [ent0, ent1) (ent1 - ent0
 * a multiple of 16), first CSR position pos0, cnt in-edges. */
typedef struct
{
    uint32_t row0, row1, ent0, ent1;
    uint32_t pos0, kind, pad, cnt;
} og_bin_desc;

/* Gather work item: segment seg, sources [f0, f1) (f0 and f1 multiples
 * of 16, within the segment's sources). */
typedef struct
{
    uint32_t seg, f0, f1, pad;
} og_bin_item;

/* Everything the binned kernels read, by value. */
typedef struct og_bin_dev
{
    uint32_t N, n_seg, n_items, n_bins, n_bins_live, n_long;
    uint32_t first_bin;          /* bins [0, first_bin) are the long rows' pieces */
    uint32_t n_ent;              /* entries in tiles; OG_BIN_SPARE spare ones follow */
    const og_bin_item *items;
    const uint32_t *qent;        /* n_ent / 16 */
    const uint16_t *src, *pos;
    float *val;
    const og_bin_desc *bins;
    const uint32_t *long_row;    /* n_long: the long rows (new ids), longest first */
    const uint32_t *long_bin;    /* n_long: each one's first piece */
    uint32_t *ticket;            /* the long-row pass's row counter (zeroed by the gather) */
} og_bin_dev;

typedef struct og_bin_plan
{
    og_bin_dev d;
    uint64_t n_ent;              /* entries (and sources) including padding and the spares */
    uint32_t n_tiles, n_zero_bins;
    uint64_t n_edges, long_terms;  /* in-edges, and those of the long rows */
    /* device allocations */
    void *mem_items, *mem_qent, *mem_src, *mem_pos, *mem_val, *mem_bins;
    void *mem_long_row, *mem_long_bin, *mem_ticket;
} og_bin_plan;

/* Builds the binned layout of a plan whose rows keep the reference's
 * in-neighbour order (relabel 2 or 0). */
int og_bin_plan_build(const og_plan *pl, og_bin_plan **out);
void og_bin_plan_destroy(og_bin_plan *b);
/* Blocks of the sum pass and of the long-row pass (statistics partials;
 * together one per CU: the passes run side by side). */
uint32_t og_bin_sum_grid(const og_bin_plan *b);
uint32_t og_bin_long_grid(const og_bin_plan *b);

/* One binned pull iteration: gather, then the sum pass and (on the side
 * stream) the long rows; ev0 / ev1 (may be NULL) bracket them.  Statistics
 * partials: the sum pass's blocks, then the long-row blocks; then the final
 * reduction. */
int og_launch_bin_pull(const og_bin_dev *b, const og_pr_dev *a, uint32_t it, uint32_t sum_grid,
                       uint32_t long_grid, hipStream_t s, hipEvent_t ev0, hipEvent_t ev1);

/* plan-build kernels */
int og_launch_bin_keys(const og_bin_desc *bins, uint32_t n_bins, const uint32_t *cols, uint32_t n_seg,
                       uint32_t *key, uint32_t *val, hipStream_t s);
/* tile_flag[i] = 1 where sorted entry i starts a tile (a bin's first entry
 * or a change of key) */
int og_launch_bin_flags(const uint32_t *key, uint32_t n, const og_bin_desc *bins, uint32_t n_seg,
                        uint32_t *tile_flag, hipStream_t s);
/* per tile t (from the exclusive scan of the flags): segment, bin, first
 * sorted entry */
int og_launch_bin_tiles(const uint32_t *key, uint32_t n, uint32_t n_seg, const uint32_t *tile_flag_scan,
                        uint32_t *tile_seg, uint32_t *tile_bin, uint32_t *tile_start, hipStream_t s);
/* sorted entry i of tile t: pos at tile_ent[t] + (i - tile_start[t]), src
 * at tile_src[t] + (i - tile_start[t]); then the padding entries of every
 * tile (src 0, pos PAD) */
int og_launch_bin_fill(const uint32_t *key, const uint32_t *val, uint32_t n, const og_bin_desc *bins,
                       uint32_t n_seg, const uint32_t *cols, const uint32_t *tile_flag_scan,
                       const uint32_t *tile_start, const uint32_t *tile_ent, const uint32_t *tile_src,
                       uint16_t *src, uint16_t *pos, hipStream_t s);
int og_launch_bin_pads(const uint32_t *tile_ent, const uint32_t *tile_src, const uint32_t *tile_cnt,
                       uint32_t n_tiles, uint16_t *src, uint16_t *pos, hipStream_t s);

#ifdef __cplusplus
}
#endif

#endif
