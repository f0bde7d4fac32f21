"""Rank-comparison statistics (-S, graphStats.c:580-1020) on the GPU.

Bar: every golden case (LAW-enron and R-MAT-16 pull rankings against
fixed-point 8-bit, 32-bit and 8-bit quantized pull rankings; top-k 30 ..
10000, 64, 512, 1000-wide chunks; the 13-vertex graphbrew, smaller than
every table size) equals the reference's
collectStatsPageRank_topK field for field, as one batch per ranking pair and
through the drop-in collectStatsPageRank_topK; the drop-in
collectStatsPageRank writes the reference's .stats file and stdout byte for
byte at verbosity 0, 1 and 2.  Beyond the fixtures: an edit-distance problem
above the register-strip sizes (global-memory diagonals) and at both strip
sizes, all 1000-wide chunks of a
200k-vertex ranking and rankings with many ties, each against the oracle.
"""
import ctypes as C
import os

import numpy as np
import pytest

import opengraph_amd as og
from conftest import GOLDEN, golden_input
from oracle import pyoracle as po

pytestmark = pytest.mark.gpu

CODES = (11, 4, 13)


@pytest.fixture(scope="module", autouse=True)
def device():
    og.init(0)


def _rankings(csr, code):
    r = po.pagerank(csr, code, 1e-4, 20, 4)["ranks"]
    return r, po.sort_labels_by_rank(r, np.arange(csr.num_vertices, dtype=np.uint32))


@pytest.fixture(scope="module")
def rank_cases(golden):
    out = {}
    for gname in ("graphbrew", "enron", "rmat16"):
        s, d = golden_input(gname, golden)
        csr = po.build_csr(s, d, 0, 4)
        out[gname] = {0: _rankings(csr, 0)}
        for code in (CODES if gname != "graphbrew" else (11,)):
            out[gname][code] = _rankings(csr, code)
    return out


def test_golden_cases_batched(golden, rank_cases):
    groups = {}
    for key, want in golden["rank_stats"].items():
        gname, code, k, m = key.split("/")
        groups.setdefault((gname, int(code[1:])), []).append(((int(k[1:]), int(m[1:])), want, key))
    for (gname, code), items in groups.items():
        ref_pr, ref_order = rank_cases[gname][0]
        pr, order = rank_cases[gname][code]
        got = og.rank_stats(ref_pr, ref_order, pr, order, [km for km, _, _ in items])
        for g, (_, want, key) in zip(got, items):
            for f, v in want.items():
                assert g[f] == v, (key, f, g[f], v)
        print(gname, code, len(items), "ok", flush=True)


def test_dropin_topk(golden, rank_cases):
    ref_pr, ref_order = rank_cases["rmat16"][0]
    pr, order = rank_cases["rmat16"][13]
    N = len(pr)
    for k, m in ((30, N), (10000, N), (1000, N - 2000)):
        want = golden["rank_stats"][f"rmat16/p13/k{k}/m{m}"]
        assert og.collect_stats_pagerank_topk(ref_pr, ref_order, pr, order, k, m) == want


def _capture_stdout(fn, path):
    libc = C.CDLL(None)
    libc.fflush(None)
    saved = os.dup(1)
    fd = os.open(path, os.O_WRONLY | os.O_CREAT | os.O_TRUNC, 0o644)
    os.dup2(fd, 1)
    try:
        fn()
        libc.fflush(None)
    finally:
        os.dup2(saved, 1)
        os.close(saved)
        os.close(fd)
    return open(path).read()


@pytest.mark.parametrize("gname,code", [("graphbrew", 11), ("enron", 11), ("rmat16", 4), ("rmat16", 13)])
def test_dropin_collect_stats_file_and_stdout(rank_cases, tmp_path, gname, code):
    ref_pr, ref_order = rank_cases[gname][0]
    pr, order = rank_cases[gname][code]
    for verb in (0, 1, 2):
        base = str(tmp_path / f"{gname}_p{code}_v{verb}")
        text = _capture_stdout(lambda: og.collect_stats_pagerank(ref_pr, ref_order, pr, order, base, 512, verb,
                                                                 code, 0, 1.25, 20, 16), base + ".out")
        name = f"rank_stats_{gname}_p{code}_v{verb}"
        with open(os.path.join(GOLDEN, "inputs", name + ".stats")) as f:
            assert open(f"{base}_1_0_0_{code}.stats").read() == f.read(), (name, "file")
        with open(os.path.join(GOLDEN, "inputs", name + ".stdout")) as f:
            assert text == f.read(), (name, "stdout")


def _random_rankings(seed, N, levels):
    rng = np.random.default_rng(seed)
    ref_pr = (rng.integers(0, levels, N) / levels).astype(np.float32)
    pr = (ref_pr + rng.integers(0, 3, N) / levels).astype(np.float32)
    return (ref_pr, po.sort_labels_by_rank(ref_pr, np.arange(N, dtype=np.uint32)),
            pr, po.sort_labels_by_rank(pr, np.arange(N, dtype=np.uint32)))


def test_large_k_global_diagonals():
    ref_pr, ref_order, pr, order = _random_rankings(7, 40000, 30000)
    probs = [(15000, 40000), (12000, 39000), (11700, 40000), (1025, 40000)]
    got = og.rank_stats(ref_pr, ref_order, pr, order, probs)
    for (k, m), g in zip(probs, got):
        assert g == po.rank_stats_topk(ref_pr, ref_order, pr, order, k, m), (k, m)


def test_all_chunks_and_ties_against_oracle():
    N = 200_000
    ref_pr, ref_order, pr, order = _random_rankings(8, N, 500)    # many ties
    probs = [(1000, N - 1000 * x) for x in range((N + 999) // 1000)] + [(30, N), (512, N)]
    got = og.rank_stats(ref_pr, ref_order, pr, order, probs)
    for (k, m), g in list(zip(probs, got))[::7] + list(zip(probs, got))[-3:]:
        assert g == po.rank_stats_topk(ref_pr, ref_order, pr, order, k, m), (k, m)
    # device-resident inputs give the same
    arrs = [og.DeviceArray.from_numpy(a) for a in (ref_pr, ref_order, pr, order)]
    try:
        assert og.rank_stats_device(*arrs, N, probs) == got
    finally:
        for a in arrs:
            a.free()


def test_rejects_bad_sizes():
    ref_pr, ref_order, pr, order = _random_rankings(9, 100, 50)
    with pytest.raises(og.OgpuError):
        og.rank_stats(ref_pr, ref_order, pr, order, [(10, 101)])
