"""Pins of the CPU oracle (oracle/og_oracle.c) -- CPU only.

The oracle is the checker of every GPU parity test, so it is pinned first:
* against the golden outputs the compiled reference wrote (tests/golden,
  made by make_golden.py from oracle/_ref/ref_harness): CSR array digests,
  rank-vector digests, iteration counts and per-iteration activeVertices;
* against closed forms and invariants that do not depend on the reference
  (stable-sort / bincount CSR, cycles, stars, a dense linear solve, mass
  balance, fixed-point pull == push);
* with planted mistakes, which these checks must catch.
"""
import hashlib

import numpy as np
import pytest

from conftest import golden_input
from oracle import pyoracle as po


def sha(a):
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


DETERMINISTIC = {0, 2, 3, 4, 6, 9, 10, 11, 12, 13}


def _cases(golden, pred):
    return [c for c in golden["cases"] if pred(c)]


@pytest.fixture(scope="module")
def csr_cache(golden):
    cache = {}

    def get(name, sort):
        if (name, sort) not in cache:
            s, d = golden_input(name, golden)
            cache[(name, sort)] = po.build_csr(s, d, sort, 4)
        return cache[(name, sort)]

    return get


def test_rmat16_input_digest(golden):
    """The generator that writes the R-MAT golden input is pinned by the
    digest of the .bin bytes the reference read."""
    s, d = golden_input("rmat16", golden)
    inter = np.empty(2 * len(s), np.uint32)
    inter[0::2], inter[1::2] = s, d
    assert hashlib.sha256(inter.tobytes()).hexdigest() == golden["inputs"]["rmat16"]["bin_sha256"]


@pytest.mark.parametrize("key", ["graphbrew/sort0", "graphbrew/sort1", "enron/sort0", "enron/sort1",
                                 "rmat16/sort0", "rmat16/sort1"])
def test_oracle_csr_matches_reference(golden, csr_cache, key):
    name, sort = key.split("/")
    g = csr_cache(name, int(sort[-1]))
    ref = golden["csr"][key]
    assert g.num_vertices == ref["num_vertices"] and g.num_edges == ref["num_edges"]
    for k in po.CSR_FIELDS:
        assert sha(g[k]) == ref[k], k


def test_oracle_pagerank_matches_reference(golden, csr_cache):
    """Bit-exact ranks, iteration count and activeVertices rows for every
    deterministic variant; push variants within 1e-5 relative."""
    checked = 0
    for c in golden["cases"]:
        g = csr_cache(c["graph"], c["sort"])
        r = po.pagerank(g, c["code"], c["epsilon"], c["max_iter"], 4)
        exact = c["code"] in DETERMINISTIC or c["graph"] == "graphbrew"
        if exact:
            assert sha(r["ranks"]) == c["ranks_sha256"], c
            assert r["iterations"] == c["iterations"], c
            assert list(r["active"]) == c["active"], c
            assert r["error_total"] == pytest.approx(c["error_total"], rel=1e-9, abs=1e-300)
        else:
            # fp32 atomic push: the order of the adds, hence the last bits of
            # the ranks and when the last vertex drops below epsilon, varies
            # from run to run in the reference as well
            assert c["code"] in (1, 5)
            assert abs(r["iterations"] - c["iterations"]) <= max(3, c["iterations"] // 8)
        checked += 1
    assert checked == len(golden["cases"])


def test_push_within_tolerance_of_reference(golden, golden_ranks, csr_cache):
    for key, ref in golden_ranks.items():
        if not key.startswith("ranks/") or "/p1/" not in key and "/p5/" not in key:
            continue
        _, gname, sort, p, eps, it = key.split("/")
        g = csr_cache(gname, int(sort[-1]))
        r = po.pagerank(g, int(p[1:]), float(eps[3:]), int(it[2:]), 4)
        rel = np.abs(r["ranks"] - ref) / np.maximum(ref, 1e-30)
        assert rel.max() < 1e-5, (key, rel.max())


# ---------------------------------------------------------------------------
# closed forms of the CSR build
# ---------------------------------------------------------------------------

def _numpy_csr(src, dst, mode):
    n = int(max(src.max(), dst.max())) + 1 if len(src) else 1
    if mode in (1, 3):
        order = np.lexsort((dst, src))
    else:
        order = np.argsort(src, kind="stable")
    fs, fd = src[order], dst[order]
    iorder = np.lexsort((fs, fd))      # stable by dst of the src-sorted list
    is_, id_ = fd[iorder], fs[iorder]

    def mapv(keys):
        deg = np.bincount(keys, minlength=n).astype(np.uint32)
        idx = np.zeros(n, np.uint32)
        first = np.ones(len(keys), bool)
        first[1:] = keys[1:] != keys[:-1]
        idx[keys[first]] = np.nonzero(first)[0]
        return deg, idx

    fdeg, fidx = mapv(fs)
    ideg, iidx = mapv(is_)
    return dict(f_src=fs, f_dst=fd, f_out_degree=fdeg, f_in_degree=ideg, f_edges_idx=fidx,
                i_src=is_, i_dst=id_, i_out_degree=ideg, i_in_degree=fdeg, i_edges_idx=iidx), n


@pytest.mark.parametrize("mode", [0, 1, 2, 3])
@pytest.mark.parametrize("seed", [1, 2, 3])
def test_oracle_csr_closed_form(mode, seed):
    rng = np.random.default_rng(seed)
    n = [7, 300, 5000][seed - 1]
    e = [40, 3000, 60000][seed - 1]
    src = rng.integers(0, n, e).astype(np.uint32)
    dst = rng.integers(0, n, e).astype(np.uint32)
    src[:5] = dst[:5]                      # self loops
    src[5:10], dst[5:10] = src[10], dst[10]  # duplicates
    g = po.build_csr(src, dst, mode, 3)
    want, nv = _numpy_csr(src, dst, mode)
    assert g.num_vertices == nv
    for k in po.CSR_FIELDS:
        assert np.array_equal(g[k], want[k]), k


def test_oracle_csr_mode_equivalence():
    rng = np.random.default_rng(9)
    src = rng.integers(0, 1000, 20000).astype(np.uint32)
    dst = rng.integers(0, 1000, 20000).astype(np.uint32)
    a0, a2 = po.build_csr(src, dst, 0, 2), po.build_csr(src, dst, 2, 5)
    a1, a3 = po.build_csr(src, dst, 1, 2), po.build_csr(src, dst, 3, 7)
    for k in po.CSR_FIELDS:
        assert np.array_equal(a0[k], a2[k]) and np.array_equal(a1[k], a3[k])


def test_oracle_empty_graph():
    g = po.build_csr(np.zeros(0, np.uint32), np.zeros(0, np.uint32), 0, 1)
    assert g.num_vertices == 1 and g.num_edges == 0     # edgeList.c:567: max 0 + 1
    r = po.pagerank(g, 0, 1e-4, 20, 1)
    # base_pr = 1.0f - 0.85f = 0.14999998f (pageRank.c:54), divided by N = 1
    assert r["ranks"][0] == np.float32(1.0) - np.float32(0.85)


# ---------------------------------------------------------------------------
# closed forms of the PageRank map
# ---------------------------------------------------------------------------

def test_cycle_converges_to_uniform():
    n = 50
    src = np.arange(n, dtype=np.uint32)
    dst = ((src + 1) % n).astype(np.uint32)
    g = po.build_csr(src, dst, 0, 1)
    r = po.pagerank(g, 0, 1e-7, 400, 1)
    assert np.allclose(r["ranks"], 1.0 / n, rtol=1e-5)
    x, _, _ = po.pagerank_fp64(g, 0.15, 1e-14, 1000)
    assert np.allclose(x, 1.0 / n, rtol=1e-12)


def test_star_closed_form():
    leaves = 20
    src = np.arange(1, leaves + 1, dtype=np.uint32)
    dst = np.zeros(leaves, np.uint32)
    g = po.build_csr(src, dst, 0, 1)
    N = leaves + 1
    r = po.pagerank(g, 0, 1e-7, 50, 1)
    # leaves have no in-edges: exactly base_pr / N with base_pr = 1.0f - 0.85f
    # (pageRank.c:54, :1016, :1040)
    base = np.float32(1.0) - np.float32(0.85)
    assert np.all(r["ranks"][1:] == base / np.float32(N))
    # the hub has no out-edges, so no mass returns: r_hub = 0.15 + 0.85*0.15*L
    assert r["ranks"][0] == pytest.approx((0.15 + 0.85 * 0.15 * leaves) / N, rel=1e-6)


def _dense_solution(g):
    N = g.num_vertices
    P = np.zeros((N, N))
    for u, v in zip(g["f_src"], g["f_dst"]):
        P[u, v] += 1.0 / g["f_out_degree"][u]
    x = np.linalg.solve(np.eye(N) - 0.85 * P.T, 0.15 * np.ones(N))
    return x / N, P


def test_dense_solve_pins_fp64_and_fp32(golden):
    s, d = golden_input("graphbrew", golden)
    g = po.build_csr(s, d, 0, 1)
    x, _ = _dense_solution(g)
    x64, _, _ = po.pagerank_fp64(g, 0.15, 1e-15, 2000)
    assert np.allclose(x64, x, rtol=1e-12, atol=0)
    r32 = po.pagerank(g, 0, 0.0, 200, 1)["ranks"]
    assert np.allclose(r32, x, rtol=2e-6)


def test_planted_transpose_mistake_is_caught(golden):
    """A transposed operand (P instead of P^T) in the closed form must
    disagree with the oracle: the dense-solve pin discriminates."""
    s, d = golden_input("graphbrew", golden)
    g = po.build_csr(s, d, 0, 1)
    _, P = _dense_solution(g)
    N = g.num_vertices
    wrong = np.linalg.solve(np.eye(N) - 0.85 * P, 0.15 * np.ones(N)) / N
    x64, _, _ = po.pagerank_fp64(g, 0.15, 1e-15, 2000)
    assert not np.allclose(x64, wrong, rtol=1e-3)


def test_mass_balance_identity(golden):
    """S_{k+1} = 0.15 N + 0.85 (S_k - D_k), D_k = rank mass on vertices
    without out-edges (no dangling redistribution, pageRank.c:992-995)."""
    s, d = golden_input("enron", golden)
    g = po.build_csr(s, d, 0, 4)
    N = g.num_vertices
    dang = g["f_out_degree"] == 0
    prev = np.full(N, 0.15)
    for k in range(1, 6):
        x, _, _ = po.pagerank_fp64(g, 0.15, 0.0, k)
        cur = x * N
        want = 0.15 * N + 0.85 * (prev.sum() - prev[dang].sum())
        assert cur.sum() == pytest.approx(want, rel=1e-12)
        prev = cur


def test_fixed_point_pull_equals_push(golden):
    s, d = golden_input("rmat16", golden)
    g = po.build_csr(s, d, 0, 8)
    a = po.pagerank(g, 2, 1e-6, 100, 8)
    b = po.pagerank(g, 3, 1e-6, 100, 8)
    assert np.array_equal(a["ranks"].view(np.uint32), b["ranks"].view(np.uint32))
    assert a["iterations"] == b["iterations"] and list(a["active"]) == list(b["active"])


def test_uniform_start_reaches_same_fixed_point(golden):
    """The L1-stop / uniform-start extension runs the same map: both starts
    converge to the FP64 fixed point (L1 contraction, d = 0.85)."""
    s, d = golden_input("enron", golden)
    g = po.build_csr(s, d, 0, 4)
    x, _, _ = po.pagerank_fp64(g, 0.15, 1e-14, 2000, 4)
    u, _, _ = po.pagerank_fp64(g, 1.0, 1e-14, 2000, 4)
    assert np.abs(x - u).sum() < 1e-12
    r = po.pagerank(g, 0, 1e-6, 200, 4, init=np.ones(g.num_vertices, np.float32), stop_l1=True)
    assert r["errors"][-1] < 1e-6
    assert np.abs(r["ranks"] - x).sum() < 1e-5


def test_rank_label_sort_matches_reference_order(golden, golden_ranks):
    """reorder.c:165-221 restated: stable ascending by the fixed-point key."""
    for c in golden["cases"]:
        if c["graph"] != "graphbrew" or c["sort"] != 0:
            continue
        key = f"ranks/graphbrew/sort0/p{c['code']}/eps{c['epsilon']}/it{c['max_iter']}"
        labels = po.sort_labels_by_rank(golden_ranks[key], np.arange(c["num_vertices"], dtype=np.uint32))
        assert sha(labels) == c["labels_sha256"]
