/*
 * og_dd.c -- data-driven PageRank variants (pageRank.c:2405-2896).
 */
#include "og_internal.h"
#include "opengraph_amd_ext.h"

int og_dd_run(ogpu_pr *pr, ogpu_pr_result *res)
{
    (void)pr;
    (void)res;
    og_set_error("data-driven variants (-p 6, 7, 8) are not built yet");
    return -1;
}
