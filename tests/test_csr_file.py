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


# ---------------------------------------------------------------------------
# ogpu_graph_read_csr / ogpu_graph_from_arrays reject malformed CSRs on the
# host, before any device call (these run without a GPU: the error must be
# the check's, not "no HIP device")
# ---------------------------------------------------------------------------

def _image(o):
    n, e = len(o["f_out_degree"]), len(o["f_src"])
    per_v = np.stack([o["f_out_degree"], o["f_in_degree"], o["f_edges_idx"],
                      o["i_out_degree"], o["i_in_degree"], o["i_edges_idx"]], 1).ravel()
    per_e = np.stack([o["f_src"], o["f_dst"], o["i_src"], o["i_dst"]], 1).ravel()
    return np.concatenate([np.array([e, n], np.uint32), per_v, per_e]).astype(np.uint32)


def _read_error(path):
    lib = og.load()
    h = C.c_void_p()
    rc = lib.ogpu_graph_read_csr(str(path).encode(), C.byref(h))
    return rc, lib.ogpu_last_error().decode()


def _corrupt_cases(golden):
    s, d = golden_input("graphbrew", golden)
    o = po.build_csr(s, d, 0, 1)
    good = _image(o)
    n, e = int(good[1]), int(good[0])
    ve = 2 + 6 * n                                  # first edge word
    cases = {}
    x = good.copy(); x[ve + 4 * 3 + 1] = n          # forward dst id == N
    cases["dst id"] = (x, "not below N")
    x = good.copy(); x[ve + 4 * 0 + 2] = 1 << 30    # inverse src id past the 30-bit space
    cases["inverse id"] = (x, "not below N")
    x = good.copy(); x[2 + 6 * 0 + 0] += 1          # forward out-degree of vertex 0 one too many
    cases["degree sum"] = (x, "vertex")
    x = good.copy(); x[2 + 6 * (n - 1) + 5] = e     # inverse edges_idx past E with a degree
    x[2 + 6 * (n - 1) + 3] = max(1, int(x[2 + 6 * (n - 1) + 3]))
    cases["edges_idx"] = (x, "vertex")
    x = good.copy(); x[0] = e + 5                   # header E larger than the body
    cases["truncated"] = (x, "truncated")
    cases["trailing"] = (np.concatenate([good, np.array([7], np.uint32)]), "bytes after the last edge")
    x = good.copy(); x[1] = (1 << 30) + 1           # N beyond the id space (and the body)
    cases["huge N"] = (x, "truncated")
    return good, cases


def test_read_csr_rejects_malformed_files(golden, tmp_path):
    good, cases = _corrupt_cases(golden)
    p = tmp_path / "good.csr"
    good.tofile(p)
    rc, msg = _read_error(p)
    # a well-formed file passes the checks; without a GPU it then stops at the device
    assert rc == 0 or "device" in msg.lower() or "hip" in msg.lower(), msg
    for name, (img, want) in cases.items():
        p = tmp_path / f"{name.replace(' ', '_')}.csr"
        img.tofile(p)
        rc, msg = _read_error(p)
        assert rc == -1, name
        assert want in msg, (name, msg)


def test_from_arrays_rejects_malformed_csr(golden):
    """ogpu_graph_from_arrays (the upload behind pageRankGraphCSR and
    Graph.from_arrays) applies the same checks."""
    s, d = golden_input("graphbrew", golden)
    o = po.build_csr(s, d, 0, 1)
    lib = og.load()
    n, e = len(o["f_out_degree"]), len(o["f_src"])
    for key, idx, val, want in [("f_dst", 3, n, "not below N"), ("i_src", 0, n + 9, "not below N"),
                                ("f_edges_idx", 2, e, "vertex"), ("i_in_degree", 4, 99, "vertex")]:
        arrs = {k: o[k].astype(np.uint32).copy() for k in KEYS}
        arrs[key][idx] = val
        ptrs = (C.c_void_p * 10)(*[arrs[k].ctypes.data for k in KEYS])
        h = C.c_void_p()
        rc = lib.ogpu_graph_from_arrays(n, e, ptrs, C.byref(h))
        msg = lib.ogpu_last_error().decode()
        assert rc == -1 and want in msg, (key, msg)
