"""Write tests/golden/abi_layout.json: sizeof / offsetof of the reference's
boundary structs (Arguments, Vertex, EdgeList, GraphCSR, PageRankStats,
SPMVStats, PageRankCorrelationStats) as
the reference's own headers lay them out on x86-64, by compiling a probe
against /root/reference's include tree.  The ABI test compares
include/opengraph_amd.h and the ctypes mirror against this fixture.
Usage (in a container holding /root/reference): python tests/golden/make_abi_layout.py
"""
import json
import os
import subprocess
import tempfile

REF = "/root/reference/00_graph_bench/include"
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "abi_layout.json")

FIELDS = {
    "Arguments": ["wflag", "xflag", "sflag", "Sflag", "dflag", "binSize", "verbosity", "iterations",
                  "trials", "epsilon", "source", "algorithm", "datastructure", "pushpull", "sort",
                  "lmode", "lmode_l2", "lmode_l3", "mmode", "symmetric", "weighted", "delta",
                  "pre_numThreads", "algo_numThreads", "ker_numThreads", "fnameb", "fnamel",
                  "fnameb_format", "convert_format", "mt19937var", "cache_size"],
    "Vertex": ["out_degree", "in_degree", "edges_idx"],
    "EdgeList": ["num_edges", "num_vertices", "avg_degree", "edges_array_src", "edges_array_dest",
                 "mask_array", "label_array", "inverse_label_array"],
    "GraphCSR": ["num_edges", "num_vertices", "avg_degree", "vertices", "sorted_edges_array",
                 "inverse_vertices", "inverse_sorted_edges_array"],
    "PageRankStats": ["damp", "base_pr", "iterations", "num_vertices", "activeVertices",
                      "realRanks", "pageRanks", "time_total", "error_total"],
    "SPMVStats": ["iterations", "arguments", "num_vertices", "vector_output", "vector_input",
                  "time_total"],
    "PageRankCorrelationStats": ["levenshtein_distance", "float_Kendall", "real_Kendall", "intersection",
                                 "mismatch", "avg_error_float", "avg_error_relative"],
}


def probe_source(includes):
    lines = ["#include <stdio.h>", "#include <stddef.h>"] + [f'#include "{h}"' for h in includes]
    lines.append("int main(void){")
    for s, fs in FIELDS.items():
        lines.append(f'printf("{s} size %zu\\n", sizeof(struct {s}));')
        for f in fs:
            lines.append(f'printf("{s} {f} %zu\\n", offsetof(struct {s}, {f}));')
    lines.append("return 0;}")
    return "\n".join(lines)


def layout(includes, incdirs):
    with tempfile.TemporaryDirectory() as td:
        c = os.path.join(td, "p.c")
        open(c, "w").write(probe_source(includes))
        exe = os.path.join(td, "p")
        subprocess.run(["gcc", "-m64", "-o", exe, c] + [f"-I{d}" for d in incdirs], check=True)
        out = subprocess.run([exe], capture_output=True, text=True, check=True).stdout
    res = {}
    for line in out.splitlines():
        parts = line.split()
        res.setdefault(parts[0], {})[parts[1]] = int(parts[2])
    return res


if __name__ == "__main__":
    incs = [os.path.join(REF, d) for d in ("structures", "algorithms/openmp", "preprocess", "utils", "config")]
    json.dump(layout(["graphConfig.h", "graphCSR.h", "pageRank.h", "SPMV.h", "graphStats.h"], incs), open(OUT, "w"), indent=1)
    print(open(OUT).read()[:300])
