"""The .csr file from and into a device graph (ogpu_graph_write_csr /
ogpu_graph_read_csr: the interleaving runs on the GPU).

Bars: a device graph built from the golden edges writes the reference's
bytes (sha256 per graph and sort mode); the reference-written graphbrew
file reads back into the reference's CSR arrays; a written R-MAT-16 file
reads back into the same arrays, and PageRank on the read graph is
bit-identical to PageRank on the built one.
"""
import hashlib
import os

import numpy as np
import pytest

import opengraph_amd as og
from conftest import GOLDEN, golden_input

pytestmark = pytest.mark.gpu

KEYS = ["f_src", "f_dst", "f_out_degree", "f_in_degree", "f_edges_idx",
        "i_src", "i_dst", "i_out_degree", "i_in_degree", "i_edges_idx"]


@pytest.fixture(scope="module", autouse=True)
def device():
    og.init(0)


@pytest.mark.parametrize("name", ["graphbrew", "enron", "rmat16"])
@pytest.mark.parametrize("sort", [0, 1])
def test_device_write_matches_reference(golden, tmp_path, name, sort):
    s, d = golden_input(name, golden)
    g = og.Graph.from_edges(s, d, sort)
    base = str(tmp_path / f"{name}.{sort}")
    try:
        g.write_csr(base)
    finally:
        g.free()
    raw = open(base + ".csr", "rb").read()
    want = golden["csr_file"][f"{name}/sort{sort}"]
    assert len(raw) == want["bytes"] and hashlib.sha256(raw).hexdigest() == want["sha256"]


def test_device_read_reference_file(golden):
    g = og.Graph.read_csr(os.path.join(GOLDEN, "inputs", "graphbrew.sort0.csr"))
    try:
        c = golden["csr"]["graphbrew/sort0"]
        assert g.num_vertices == c["num_vertices"] and g.num_edges == c["num_edges"]
        for k, v in g.arrays().items():
            assert hashlib.sha256(v.tobytes()).hexdigest() == c[k], k
    finally:
        g.free()


def test_device_round_trip_and_pagerank(golden, tmp_path):
    s, d = golden_input("rmat16", golden)
    g = og.Graph.from_edges(s, d, 0)
    base = str(tmp_path / "rmat16")
    g.write_csr(base)
    h = og.Graph.read_csr(base + ".csr")
    try:
        a, b = g.arrays(), h.arrays()
        for k in KEYS:
            assert np.array_equal(a[k], b[k]), k
        out = []
        for gr in (g, h):
            pr = og.PageRank(gr, 0, 100, 1e-6)
            res = pr.run()
            out.append((pr.ranks(), res.iterations))
            pr.free()
        assert out[0][1] == out[1][1] and np.array_equal(out[0][0].view(np.uint32), out[1][0].view(np.uint32))
    finally:
        g.free()
        h.free()


def test_device_read_rejects_truncated_file(tmp_path):
    p = tmp_path / "bad.csr"
    p.write_bytes(np.array([5, 3, 1, 2], np.uint32).tobytes())
    with pytest.raises(og.OgpuError):
        og.Graph.read_csr(str(p))
