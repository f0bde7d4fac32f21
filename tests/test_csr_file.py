"""The .csr file (graphCSR.c:380-558) written and read by the library.

Bars: the bytes the library's writeToBinFileGraphCSR writes equal the
bytes the reference's own writeToBinFileGraphCSR wrote for the same CSR
(sha256 per golden graph and sort mode, tests/golden/make_golden_csrfile.py);
readFromBinFileGraphCSR of the reference-written graphbrew file gives the
reference's CSR arrays; write then read is the identity.  These entry
points are host code (no GPU); the device versions are in
test_gpu_csr_file.py.
"""
import ctypes as C
import hashlib
import os

import numpy as np
import pytest

import opengraph_amd as og
from conftest import GOLDEN, golden_input
from oracle import pyoracle as po

KEYS = ["f_src", "f_dst", "f_out_degree", "f_in_degree", "f_edges_idx",
        "i_src", "i_dst", "i_out_degree", "i_in_degree", "i_edges_idx"]


def _arrays(gp):
    g = gp.contents
    n, e = g.num_vertices, g.num_edges
    fe, ie = g.sorted_edges_array.contents, g.inverse_sorted_edges_array.contents
    fv, iv = g.vertices.contents, g.inverse_vertices.contents

    def a(p, k):
        return np.ctypeslib.as_array(p, shape=(k,)).copy() if k else np.zeros(0, np.uint32)
    return {"f_src": a(fe.edges_array_src, e), "f_dst": a(fe.edges_array_dest, e),
            "f_out_degree": a(fv.out_degree, n), "f_in_degree": a(fv.in_degree, n),
            "f_edges_idx": a(fv.edges_idx, n), "i_src": a(ie.edges_array_src, e),
            "i_dst": a(ie.edges_array_dest, e), "i_out_degree": a(iv.out_degree, n),
            "i_in_degree": a(iv.in_degree, n), "i_edges_idx": a(iv.edges_idx, n)}


@pytest.mark.parametrize("name", ["graphbrew", "enron", "rmat16"])
@pytest.mark.parametrize("sort", [0, 1])
def test_written_bytes_match_reference(golden, tmp_path, name, sort):
    s, d = golden_input(name, golden)
    o = po.build_csr(s, d, sort, 4)
    host = og.HostGraphCSR(o.a)
    base = str(tmp_path / f"{name}.{sort}")
    og.load().writeToBinFileGraphCSR(base.encode(), C.byref(host.g))
    raw = open(base + ".csr", "rb").read()
    want = golden["csr_file"][f"{name}/sort{sort}"]
    assert len(raw) == want["bytes"]
    assert hashlib.sha256(raw).hexdigest() == want["sha256"]


def test_read_reference_file(golden):
    lib = og.load()
    gp = lib.readFromBinFileGraphCSR(os.path.join(GOLDEN, "inputs", "graphbrew.sort0.csr").encode())
    try:
        got = _arrays(gp)
        c = golden["csr"]["graphbrew/sort0"]
        assert gp.contents.num_vertices == c["num_vertices"] and gp.contents.num_edges == c["num_edges"]
        for k in KEYS:
            assert hashlib.sha256(got[k].tobytes()).hexdigest() == c[k], k
    finally:
        lib.graphCSRFree(gp)


def test_write_read_round_trip(golden, tmp_path):
    s, d = golden_input("enron", golden)
    o = po.build_csr(s, d, 1, 4)
    host = og.HostGraphCSR(o.a)
    lib = og.load()
    base = str(tmp_path / "enron")
    lib.writeToBinFileGraphCSR(base.encode(), C.byref(host.g))
    gp = lib.readFromBinFileGraphCSR((base + ".csr").encode())
    try:
        got = _arrays(gp)
        for k in KEYS:
            assert np.array_equal(got[k], o[k]), k
    finally:
        lib.graphCSRFree(gp)
