"""The oracle's rank-comparison statistics (-S, graphStats.c:580-688).

Pins: every golden case (LAW-enron and R-MAT-16 pull rankings against
fixed-point 8-bit, 32-bit and 8-bit quantized pull rankings; top-k 30 ..
10000, 64, 512 and 1000-wide chunks at lower num_vertices; the 13-vertex
graphbrew) equals the
reference's own collectStatsPageRank_topK field for field
(tests/golden/make_golden_rank_stats.py), and so does a live comparison on
random rankings with ties when the reference is built.  Textbook checks:
the Kendall fields against scipy's tau-b on the same top-k arrays; identical
rankings give distance 0, tau 1, full intersection, no mismatch, no error;
a one-position rotation of the top-k list gives edit distance 2.
"""
import numpy as np
import pytest
from scipy import stats as sst

from conftest import golden_input
from oracle import pyoracle as po

CODES = (11, 4, 13)


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


def test_golden_cases_match_reference(golden, rank_cases):
    cases = golden["rank_stats"]
    assert len(cases) == 2 * 3 * 11 + 4
    for key, want in cases.items():
        gname, code, k, m = key.split("/")
        code, k, m = int(code[1:]), int(k[1:]), int(m[1:])
        ref_pr, ref_order = rank_cases[gname][0]
        pr, order = rank_cases[gname][code]
        got = po.rank_stats_topk(ref_pr, ref_order, pr, order, k, m)
        for f, v in want.items():
            assert got[f] == v, (key, f, got[f], v)


def test_kendall_against_scipy(rank_cases):
    ref_pr, ref_order = rank_cases["rmat16"][0]
    pr, order = rank_cases["rmat16"][13]
    N = len(pr)
    for k in (30, 300, 2000):
        got = po.rank_stats_topk(ref_pr, ref_order, pr, order, k)
        top = order[N - k:]
        t = sst.kendalltau(ref_pr[top].astype(np.float64), pr[top].astype(np.float64), variant="b").statistic
        assert abs(got["float_Kendall"] - t) < 1e-5, k
        t = sst.kendalltau(ref_order[N - k:].astype(np.float64), top.astype(np.float64), variant="b").statistic
        assert abs(got["real_Kendall"] - t) < 1e-5, k


def test_identities():
    rng = np.random.default_rng(5)
    N = 3000
    pr = rng.random(N).astype(np.float32)
    order = np.argsort(pr, kind="stable").astype(np.uint32)
    got = po.rank_stats_topk(pr, order, pr, order, 1000)
    assert got["levenshtein_distance"] == 0 and got["intersection"] == 1000 and got["mismatch"] == 0
    assert got["float_Kendall"] == pytest.approx(1.0) and got["real_Kendall"] == pytest.approx(1.0)
    assert got["avg_error_float"] == 0.0 and got["avg_error_relative"] == 0.0
    # rotate the top-k list by one position: one deletion plus one insertion
    rot = order.copy()
    rot[N - 1000:] = np.roll(order[N - 1000:], 1)
    got = po.rank_stats_topk(pr, order, pr, rot, 1000)
    assert got["levenshtein_distance"] == 2 and got["intersection"] == 1000


@pytest.mark.skipif(not po.ref_available(), reason="oracle/_ref not built")
@pytest.mark.parametrize("seed", [1, 2, 3])
def test_live_reference_with_ties(seed):
    rng = np.random.default_rng(seed)
    N = 4000
    ref_pr = rng.integers(0, 50, N).astype(np.float32) / 7     # many ties
    pr = (ref_pr + rng.integers(0, 3, N) / 7).astype(np.float32)
    ref_order = po.sort_labels_by_rank(ref_pr, np.arange(N, dtype=np.uint32))
    order = po.sort_labels_by_rank(pr, np.arange(N, dtype=np.uint32))
    for k, m in ((30, N), (777, N), (1000, N - 1000), (500, 500)):
        assert po.rank_stats_topk(ref_pr, ref_order, pr, order, k, m) == \
            po.ref_rank_stats_topk(ref_pr, ref_order, pr, order, k, m), (k, m)
