"""Add the PageRank rank-comparison fixtures (graphStats.c:580-1020, -S) to
tests/golden/.

For LAW-enron and R-MAT-16 (sort 0) the rankings are the reference's pull
ranks (-p 0) and those of fixed-point 8-bit (-p 11), 32-bit quantized (-p 4)
and 8-bit quantized (-p 13) pulls (graphbrew, smaller than every top-k of
the table: -p 11 only), epsilon 1e-4, 20 iterations, each
ordered by the reference's radixSortEdgesByPageRank restated in the oracle
(both pinned bit-exact elsewhere).  The reference's own compiled
collectStatsPageRank_topK (oracle/_ref/libogref.so) gives, under
"rank_stats" in golden.json, the seven statistics for top-k 30 .. 10000,
binSize-like 64 and 512, and the 1000-wide chunks at num_vertices
N, N - 1000, N - 2000 and a last partial one.  Its collectStatsPageRank also
writes the .stats file and stdout for verbosity 0, 1 and 2 (binSize 512,
trial 0); those texts are kept as inputs/rank_stats_<graph>_p<code>_v<n>.{stats,stdout}.
Needs oracle/_ref (this container); usage:
python tests/golden/make_golden_rank_stats.py
"""
from __future__ import annotations

import json
import os
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.dirname(HERE))

CODES = (11, 4, 13)


def rankings(po, csr, code):
    r = po.ref_pagerank(csr, code, 1e-4, 20, 4)["ranks"]
    return r, po.sort_labels_by_rank(r, np.arange(csr.num_vertices, dtype=np.uint32))


def problems(N):
    out = [(k, N) for k in (30, 100, 300, 1000, 5000, 10000) if k <= N]
    out += [(64, N), (512, N)]
    if N > 3000:
        out += [(1000, N - 1000 * x) for x in (1, 2)]
        last = (N + 999) // 1000 - 1
        out.append((1000, N - 1000 * last))
    else:
        out += [(N, N), (3, N - 2)]      # a graph smaller than every table size
    return out


def main():
    from oracle import pyoracle as po
    from conftest import golden_input

    if not po.ref_available():
        sys.exit("oracle/_ref missing: make -C oracle ref")
    path = os.path.join(HERE, "golden.json")
    golden = json.load(open(path))
    out = {}
    td = tempfile.mkdtemp()
    for gname in ("graphbrew", "enron", "rmat16"):
        s, d = golden_input(gname, golden)
        csr = po.build_csr(s, d, 0, 4)
        N = csr.num_vertices
        ref_pr, ref_order = rankings(po, csr, 0)
        for code in (CODES if gname != "graphbrew" else (11,)):
            pr, order = rankings(po, csr, code)
            for k, m in problems(N):
                out[f"{gname}/p{code}/k{k}/m{m}"] = po.ref_rank_stats_topk(ref_pr, ref_order, pr, order, k, m)
            for verb in (0, 1, 2):
                base = os.path.join(td, f"{gname}_p{code}_v{verb}")
                so = base + ".stdout"
                po.ref_collect_stats(ref_pr, ref_order, pr, order, base, 512, verb, code, 0, 1.25, 20, 16, so)
                stats_file = f"{base}_1_0_0_{code}.stats"
                name = f"rank_stats_{gname}_p{code}_v{verb}"
                with open(stats_file) as f, open(os.path.join(HERE, "inputs", name + ".stats"), "w") as g:
                    g.write(f.read())
                with open(so) as f, open(os.path.join(HERE, "inputs", name + ".stdout"), "w") as g:
                    g.write(f.read())
            print(gname, code, flush=True)
    golden["rank_stats"] = out
    with open(path, "w") as f:
        json.dump(golden, f, indent=1)


if __name__ == "__main__":
    main()
