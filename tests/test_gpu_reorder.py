"""Degree and DBG relabelling (-l 1..5) on the GPU, through the C ABI.

Bar: bit-exact against the compiled reference's outputs for every golden
case (graphbrew, LAW-enron, R-MAT-16; sort modes 0 and 1; -l 1..5): the
label of every vertex, every CSR array of the relabelled graph and the
pull ranks and iteration count on it (rows summed in the reference's
order).  Two levels (-l then -L) against the oracle applied twice; the
drop-in graphCSRPreProcessingStep with -l set and the reference-signature
reorderGraphProcess on a host edge list against the same fixtures.
"""
import ctypes as C
import hashlib
import os

import numpy as np
import pytest

import opengraph_amd as og
from conftest import GOLDEN, golden_input
from oracle import pyoracle as po

pytestmark = pytest.mark.gpu

ALL_EXACT = 1 << 31
CSR_KEYS = ["f_src", "f_dst", "f_out_degree", "f_in_degree", "f_edges_idx",
            "i_src", "i_dst", "i_out_degree", "i_in_degree", "i_edges_idx"]


def sha(a):
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


@pytest.fixture(scope="module", autouse=True)
def device():
    og.init(0)


def test_reordered_graphs_match_reference(golden):
    for key, c in golden["reorder"].items():
        gname, sort, lm = key.split("/")
        sort, lmode = int(sort[4:]), int(lm[1:])
        s, d = golden_input(gname, golden)
        g, labels = og.Graph.from_edges_reordered(s, d, sort, (lmode,))
        try:
            assert g.num_vertices == c["num_vertices"] and g.num_edges == c["num_edges"], key
            assert sha(labels) == c["labels_sha256"], key
            arrays = g.arrays()
            for k in CSR_KEYS:
                assert sha(arrays[k]) == c[k], (key, k)
            pr = og.PageRank(g, 0, 100, 1e-4, hub_threshold=ALL_EXACT)
            res = pr.run()
            ranks = pr.ranks()
            pr.free()
            assert res.iterations == c["iterations"] and sha(ranks) == c["ranks_sha256"], key
        finally:
            g.free()


@pytest.mark.parametrize("levels", [(4, 1), (2, 5), (3, 3)])
def test_two_levels_match_oracle(golden, levels):
    s, d = golden_input("rmat16", golden)
    N = int(max(s.max(), d.max())) + 1
    lab = np.arange(N, dtype=np.uint32)
    s1, d1 = s.copy(), d.copy()
    for lm in levels:
        step = po.reorder_labels(s1, d1, N, lm)
        s1, d1, lab = step[s1], step[d1], step[lab]
    o = po.build_csr(s1, d1, 1, 8, N)
    g, labels = og.Graph.from_edges_reordered(s, d, 1, levels)
    try:
        assert np.array_equal(labels, lab)
        arrays = g.arrays()
        for k in CSR_KEYS:
            assert np.array_equal(arrays[k], o[k]), k
    finally:
        g.free()


def test_rejects_other_relabellings(golden):
    s, d = golden_input("graphbrew", golden)
    for lm in (6, 10, 12):
        with pytest.raises(og.OgpuError):
            og.Graph.from_edges_reordered(s, d, 0, (lm,))


def _as(ptr, n):
    return np.ctypeslib.as_array(ptr, shape=(n,)).copy() if n else np.zeros(0, np.uint32)


@pytest.mark.parametrize("lmode", [1, 4])
def test_dropin_preprocessing_with_relabel(golden, lmode):
    """graphCSRPreProcessingStep with -l set: the reference's CSR arrays and
    label_array for LAW-enron."""
    lib = og.load()
    args = og.default_arguments(fnameb=os.path.join(GOLDEN, "inputs", "enron.bin").encode(), sort=0,
                                lmode=lmode)
    gp = lib.graphCSRPreProcessingStep(C.byref(args))
    g = gp.contents
    c = golden["reorder"][f"enron/sort0/l{lmode}"]
    try:
        n, e = g.num_vertices, g.num_edges
        assert n == c["num_vertices"] and e == c["num_edges"]
        fe, ie = g.sorted_edges_array.contents, g.inverse_sorted_edges_array.contents
        fv, iv = g.vertices.contents, g.inverse_vertices.contents
        got = {"f_src": _as(fe.edges_array_src, e), "f_dst": _as(fe.edges_array_dest, e),
               "f_out_degree": _as(fv.out_degree, n), "f_in_degree": _as(fv.in_degree, n),
               "f_edges_idx": _as(fv.edges_idx, n), "i_src": _as(ie.edges_array_src, e),
               "i_dst": _as(ie.edges_array_dest, e), "i_out_degree": _as(iv.out_degree, n),
               "i_in_degree": _as(iv.in_degree, n), "i_edges_idx": _as(iv.edges_idx, n)}
        for k in CSR_KEYS:
            assert sha(got[k]) == c[k], k
        lab = _as(fe.label_array, n)
        assert sha(lab) == c["labels_sha256"]
        inv = _as(fe.inverse_label_array, n)
        assert np.array_equal(inv[lab], np.arange(n, dtype=np.uint32))
    finally:
        lib.graphCSRFree(gp)


@pytest.mark.parametrize("lmode", [2, 3, 5])
def test_dropin_reorder_graph_process(golden, lmode):
    """reorderGraphProcess on the edge list readEdgeListsbin returns: every
    edge relabelled in place (edge order kept) and the label arrays."""
    lib = og.load()
    path = os.path.join(GOLDEN, "inputs", "enron.bin").encode()
    elp = lib.readEdgeListsbin(path, 0, 0, 0)
    el = elp.contents
    s, d = golden_input("enron", golden)
    n, e = el.num_vertices, el.num_edges
    args = og.default_arguments(lmode=lmode)
    try:
        out = lib.reorderGraphProcess(elp, C.byref(args)).contents
        lab = po.reorder_labels(s, d, n, lmode)
        assert np.array_equal(_as(out.edges_array_src, e), lab[s])
        assert np.array_equal(_as(out.edges_array_dest, e), lab[d])
        assert np.array_equal(_as(out.label_array, n), lab)
        assert sha(lab) == golden["reorder"][f"enron/sort0/l{lmode}"]["labels_sha256"]
    finally:
        lib.freeEdgeList(elp)
