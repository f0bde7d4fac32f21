"""The command-line driver (opengraph_amd/open-graph-amd) with -S and -a 5.

-S (graphRun.c:414-420): after the -p 11 run the driver runs the pull, sorts
both rankings and calls collectStatsPageRank; its .stats file must equal
the library's collectStatsPageRank on the same two runs (whose text is
tested against the reference's own elsewhere), apart from the run's own time
and iteration lines.  -a 5 runs the SpMV benchmark; its printed sum (SPMV.c:697-701)
must be the one of the oracle's vector_output (pinned byte-exact against
the reference's).
"""
import os
import shutil
import subprocess

import numpy as np
import pytest

import opengraph_amd as og
from conftest import GOLDEN, golden_input
from oracle import pyoracle as po

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CLI = os.path.join(ROOT, "opengraph_amd", "open-graph-amd")


def _run(args, cwd):
    r = subprocess.run([CLI] + args, cwd=cwd, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    return r.stdout


def _strip(text):
    return [l for l in text.splitlines() if not l.startswith(("Time (S):", "Iterations:"))]


def test_cli_stats_file(golden, tmp_path):
    f = str(tmp_path / "enron.bin")
    shutil.copy(os.path.join(GOLDEN, "inputs", "enron.bin"), f)
    _run(["-f", f, "-a", "1", "-d", "0", "-p", "11", "-o", "0", "-i", "20", "-e", "0.0001", "-S", "-g", "512",
          "-j", "1", "-n", "16", "-N", "16", "-K", "16"], str(tmp_path))
    got = open(f + "_1_0_1_0.stats").read()
    # the same two runs through the library, ordered by the reference's post-sort
    # (pinned in the oracle), and the library's collectStatsPageRank on them
    og.init(0)
    s, d = golden_input("enron", golden)
    g = og.Graph.from_edges(s, d, 0)
    rk = {}
    for code in (0, 11):
        pr = og.PageRank(g, code, 20, 1e-4)
        pr.run()
        r = pr.ranks()
        pr.free()
        rk[code] = (r, po.sort_labels_by_rank(r, np.arange(len(r), dtype=np.uint32)))
    g.free()
    base = str(tmp_path / "lib")
    og.collect_stats_pagerank(*rk[0], *rk[11], base, 512, 1, 0, 1, 0.0, 0, 16)
    want = open(base + "_1_0_1_0.stats").read()
    assert _strip(got) == _strip(want)
    assert "Top K                 30             100            300            1000" in got


def test_cli_spmv_sum(golden, tmp_path):
    f = str(tmp_path / "enron.bin")
    shutil.copy(os.path.join(GOLDEN, "inputs", "enron.bin"), f)
    s, d = golden_input("enron", golden)
    csr = po.build_csr(s, d, 0, 4)
    for code, title in ((0, "Starting SPMV-PULL"), (2, "Starting SPMV-PULL Fixed-Point"),
                        (3, "Starting SPMV-PUSH Fixed-Point")):
        out = _run(["-f", f, "-a", "5", "-d", "0", "-p", str(code), "-o", "0", "-i", "20"], str(tmp_path))
        assert title in out
        y = po.spmv(csr, code, 20)
        # SPMV.c:697-701: (int)(y * 100 + .5) / 100.0, the product in float
        want = sum(int(t) / 100.0 for t in ((y * np.float32(100)).astype(np.float64) + 0.5))
        line = [l for l in out.splitlines() if l.startswith("| 20 ")][-1]
        got = float(line.split("|")[2])
        if code == 0:
            assert abs(got - want) <= 1e-4 * abs(want) + 0.1   # long rows are summed as trees
        else:
            assert f"{got:.6f}" == f"{want:.6f}"
