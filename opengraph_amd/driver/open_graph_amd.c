/*
 * open_graph_amd.c -- command-line driver with the reference's flags
 * (open-graph.c:44-327) for the path this library implements: PageRank
 * (-a 1) on CSR (-d 0).  The PageRank trial loop, timing report, rank
 * post-sort and .perf output follow runGraphAlgorithms /
 * runPageRankAlgorithm (graphRun.c:380-535, :1023-1082).
 *
 * Added: -G <spec> writes a seeded synthetic graph to the -f path in the
 * reference's .bin format and exits:
 *     -G rmat,<scale>,<edge factor>,<seed>[,a,b,c]   (a,b,c default .57,.19,.19)
 *     -G er,<vertices>,<edges>,<seed>
 */
#include <argp.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <unistd.h>

#include "opengraph_amd.h"
#include "opengraph_amd_ext.h"

static char doc[] = "OpenGraph PageRank-over-CSR (and SpMV) on MI355X (gfx950).";
static char args_doc[] = "-f <graph file> -a 1|5 -d 0 -p [variant] ...";

static struct argp_option options[] = {
    {"graph-file", 'f', "<FILE>", 0, "Edge list in the binary .bin format."},
    {"graph-file-format", 'z', "[1]", 0, "Only [1] binary edge list is supported."},
    {"algorithm", 'a', "[1]", 0, "[1]-Page-rank [5]-SPMV (the others are not supported)."},
    {"data-structure", 'd', "[0]", 0, "Only [0]-CSR is supported."},
    {"root", 'r', "[0]", 0, "Unused by PageRank."},
    {"direction", 'p', "[0]", 0,
     "[0]-pull [1]-push [2]-pull FP64 [3]-push FP64 [4]-pull Quant32 [5]-push Quant "
     "[6]-DD pull [7]-DD push [8]-DD pull-push [9]-pull FP32 [10]-pull FP16 [11]-pull FP8 "
     "[12]-pull Quant16 [13]-pull Quant8. SPMV (-a 5): [0]-pull [1]-push [2]-pull fixed "
     "[3]-push fixed."},
    {"sort", 'o', "[0]", 0, "[0]-radix-src [1]-radix-src-dest [2]-count-src [3]-count-src-dst."},
    {"pre-num-threads", 'n', "[MAX]", 0, "Host threads for preprocessing."},
    {"algo-num-threads", 'N', "[MAX]", 0, "Host threads for the algorithm."},
    {"Kernel-num-threads", 'K', "[N]", 0, "Host threads for the kernel."},
    {"num-iterations", 'i', "[20]", 0, "Maximum PageRank iterations."},
    {"num-trials", 't', "[1]", 0, "Number of trials."},
    {"tolerance", 'e', "[0.0001]", 0, "Per-vertex tolerance."},
    {"delta", 'b', "[1]", 0, "Unused by PageRank."},
    {"light-reorder-l1", 'l', "[0]", 0,
     "Relabel before the CSR build: 1 out-, 2 in-, 3 in+out-degree sort, 4/5 DBG by out-/in-degree "
     "(other orders are not supported)."},
    {"light-reorder-l2", 'L', "[0]", 0, "A second relabelling after -l (same values)."},
    {"light-reorder-l3", 'O', "[0]", 0, "A third relabelling after -L (same values)."},
    {"convert-format", 'c', "[1]", 0, "Unused."},
    {"generate-weights", 'w', 0, 0, "Records carry a weight (12-byte .wbin); weights are skipped."},
    {"symmetrize", 's', 0, 0, "Add the reverse of every edge."},
    {"serialize", 'x', 0, 0, "Not supported."},
    {"stats", 'S', 0, 0, "PageRank: compare with a pull run (graphStats.c), write <graph>_1_0_<trial>_0.stats."},
    {"bin-size", 'g', "[512]", 0, "Top-k of the -S comparison."},
    {"verbosity", 'j', "[0]", 0, "1: append '<threads> <avg seconds>' to <graph>_1_0_<trials>_<p>.perf."},
    {"remove-duplicate", 'k', 0, 0, "Not supported."},
    {"mask-mode", 'M', "[0]", 0, "Not supported (must be 0)."},
    {"labels-file", 'F', "<FILE>", 0, "Unused."},
    {"cache-size", 'C', "[32768]", 0, "Unused."},
    {"generate", 'G', "<spec>", 0, "Write a synthetic graph to -f and exit (see source header)."},
    {0}};

static const char *g_generate = NULL;

static error_t parse_opt(int key, char *arg, struct argp_state *state)
{
    struct Arguments *a = state->input;
    switch (key)
    {
    case 'f': a->fnameb = arg; break;
    case 'F': a->fnamel = arg; break;
    case 'z': a->fnameb_format = atoi(arg); break;
    case 'd': a->datastructure = atoi(arg); break;
    case 'a': a->algorithm = atoi(arg); break;
    case 'r': a->source = atoi(arg); break;
    case 'n': a->pre_numThreads = atoi(arg); break;
    case 'N': a->algo_numThreads = atoi(arg); break;
    case 'K': a->ker_numThreads = atoi(arg); break;
    case 'i': a->iterations = atoi(arg); break;
    case 't': a->trials = atoi(arg); break;
    case 'e': a->epsilon = atof(arg); break;
    case 'p': a->pushpull = atoi(arg); break;
    case 'o': a->sort = atoi(arg); break;
    case 'l': a->lmode = atoi(arg); break;
    case 'L': a->lmode_l2 = atoi(arg); break;
    case 'O': a->lmode_l3 = atoi(arg); break;
    case 'b': a->delta = atoi(arg); break;
    case 's': a->symmetric = 1; break;
    case 'S': a->Sflag = 1; break;
    case 'w': a->weighted = 1; break;
    case 'x': a->xflag = 1; break;
    case 'c': a->convert_format = atoi(arg); break;
    case 'C': a->cache_size = atoi(arg); break;
    case 'g': a->binSize = atoi(arg); break;
    case 'j': a->verbosity = atoi(arg); break;
    case 'k': a->dflag = 1; break;
    case 'M': a->mmode = atoi(arg); break;
    case 'G': g_generate = arg; break;
    default: return ARGP_ERR_UNKNOWN;
    }
    return 0;
}

static struct argp argp = {options, parse_opt, args_doc, doc};

static int generate(const char *spec, const char *path)
{
    char kind[16] = {0};
    double x[6] = {0, 0, 0, 0.57, 0.19, 0.19};
    int n = sscanf(spec, "%15[^,],%lf,%lf,%lf,%lf,%lf,%lf", kind, &x[0], &x[1], &x[2], &x[3], &x[4], &x[5]);
    if (n < 4 || !path)
    {
        fprintf(stderr, "bad -G spec or missing -f\n");
        return 2;
    }
    uint64_t E;
    uint32_t *src, *dst;
    int rc;
    if (strcmp(kind, "rmat") == 0)
    {
        uint32_t scale = (uint32_t)x[0];
        E = (uint64_t)(x[1] * (double)(1ull << scale));
        src = malloc(E * 4);
        dst = malloc(E * 4);
        if (!src || !dst) { fprintf(stderr, "out of memory\n"); return 1; }
        rc = ogpu_rmat_host(scale, E, x[3], x[4], x[5], (uint64_t)x[2], 1, src, dst);
    }
    else if (strcmp(kind, "er") == 0)
    {
        E = (uint64_t)x[1];
        src = malloc(E * 4);
        dst = malloc(E * 4);
        if (!src || !dst) { fprintf(stderr, "out of memory\n"); return 1; }
        rc = ogpu_er_host((uint32_t)x[0], E, (uint64_t)x[2], src, dst);
    }
    else
    {
        fprintf(stderr, "unknown generator '%s'\n", kind);
        return 2;
    }
    if (rc == 0) rc = ogpu_write_bin(path, src, dst, E);
    if (rc) fprintf(stderr, "%s\n", ogpu_last_error());
    else printf("wrote %llu edges to %s\n", (unsigned long long)E, path);
    free(src);
    free(dst);
    return rc ? 1 : 0;
}

int main(int argc, char **argv)
{
    struct Arguments *a = calloc(1, sizeof(struct Arguments));
    long ncpu = sysconf(_SC_NPROCESSORS_ONLN);
    /* open-graph.c:270-305 */
    a->binSize = 512;
    a->iterations = 20;
    a->trials = 1;
    a->epsilon = 0.0001;
    a->algorithm = 1;
    a->cache_size = 32768;
    a->delta = 1;
    a->pre_numThreads = a->algo_numThreads = a->ker_numThreads = (int)(ncpu > 0 ? ncpu : 1);
    a->fnameb_format = 1;
    a->convert_format = 1;
    argp_parse(&argp, argc, argv, 0, 0, a);
    if (a->dflag) a->sort = 1;   /* open-graph.c:312-313 */

    if (g_generate) return generate(g_generate, a->fnameb);
    if (!a->fnameb) { fprintf(stderr, "missing -f <graph file>\n"); return 2; }
    if ((a->algorithm != 1 && a->algorithm != 5) || (a->datastructure != 0 && a->datastructure != 4) ||
        a->xflag || a->fnameb_format != 1)
    {
        fprintf(stderr, "this driver runs PageRank (-a 1) and SpMV (-a 5) on CSR (-d 0) from binary edge "
                        "lists only\n");
        return 2;
    }

    struct GraphCSR *graph = graphCSRPreProcessingStep(a);
    printf("*-----------------------------------------------------*\n");
    printf("| %-35s %-15d | \n", "Number of Threads Algorithm :", a->algo_numThreads);
    printf(" -----------------------------------------------------\n");
    printf("*-----------------------------------------------------*\n");
    printf("| %-35s %-15d | \n", "Number of Threads Kernel    :", a->ker_numThreads);
    printf(" -----------------------------------------------------\n");
    double time_total = 0.0;
    uint32_t trials = a->trials;
    /* graphRun.c:393-470 for -a 1 and -a 5 */
    while (trials)
    {
        if (a->algorithm == 5)
        {
            struct SPMVStats *stats = SPMVGraphCSR(a, graph);
            if (stats)
            {
                time_total += stats->time_total;
                freeSPMVStats(stats);
            }
            trials--;
            continue;
        }
        struct PageRankStats *stats = pageRankGraphCSR(a, graph);
        if (stats)
        {
            stats->realRanks = radixSortEdgesByPageRank(stats->pageRanks, stats->realRanks, stats->num_vertices);
            time_total += stats->time_total;
            if (a->Sflag)
            {
                /* graphRun.c:414-420: the pull run is the reference ranking, and
                 * -p stays 0 for the trials after it, as in the reference */
                a->pushpull = 0;
                struct PageRankStats *ref_stats = pageRankGraphCSR(a, graph);
                ref_stats->realRanks =
                    radixSortEdgesByPageRank(ref_stats->pageRanks, ref_stats->realRanks, ref_stats->num_vertices);
                collectStatsPageRank(a, ref_stats, stats, trials);
                freePageRankStats(ref_stats);
            }
            freePageRankStats(stats);
        }
        trials--;
    }
    printf("*-----------------------------------------------------*\n");
    printf("| %-21s %-8u %-21f | \n", "Trials Avg Time (S) :", a->trials,
           a->trials ? time_total / a->trials : 0.0);
    printf(" -----------------------------------------------------\n");
    if (a->verbosity > 0)
    {
        char fname[4096];
        snprintf(fname, sizeof(fname), "%s_%u_%u_%u_%u.perf", a->fnameb, a->algorithm, a->datastructure,
                 a->trials, a->pushpull);
        FILE *f = fopen(fname, "a");
        if (f)
        {
            fprintf(f, "%u %lf \n", a->algo_numThreads, a->trials ? time_total / a->trials : 0.0);
            fclose(f);
        }
    }
    graphCSRFree(graph);
    free(a);
    return 0;
}
