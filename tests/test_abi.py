"""The C-ABI boundary -- CPU only (no compute calls).

* libopengraph_amd.so loads and exports every function include/*.h declares;
* the structs of include/opengraph_amd.h and of the ctypes mirror have the
  layout the reference's own headers give them (tests/golden/abi_layout.json,
  written by make_abi_layout.py from the reference's include tree).
"""
import ctypes as C
import json
import os
import re
import subprocess
import tempfile

import pytest

import opengraph_amd as og
from conftest import GOLDEN, ROOT

HEADERS = [os.path.join(ROOT, "include", h) for h in sorted(os.listdir(os.path.join(ROOT, "include")))
           if h.endswith(".h")]


def declared_functions():
    names = []
    for h in HEADERS:
        text = open(h).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        for m in re.finditer(r"^[A-Za-z_][\w \*]*?\b(\w+)\s*\([^;{]*\)\s*;", text, flags=re.M):
            name = m.group(1)
            if name not in ("if", "while", "for", "return"):
                names.append(name)
    return sorted(set(names))


def test_headers_declare_the_reference_entry_points():
    names = declared_functions()
    for fn in ["pageRankGraphCSR", "pageRankPullGraphCSR", "pageRankPushGraphCSR",
               "pageRankPullFixedPoint64BitGraphCSR", "pageRankPushFixedPointGraphCSR",
               "newPageRankStatsGraphCSR", "freePageRankStats", "graphCSRPreProcessingStep",
               "sortRunAlgorithms", "mapVerticesWithInOutDegree", "readEdgeListsbin", "SPMVGraphCSR",
               "SPMVPullGraphCSR", "SPMVPushGraphCSR", "SPMVPullFixedPointGraphCSR",
               "SPMVPushFixedPointGraphCSR", "newSPMVStatsGraphCSR", "freeSPMVStats",
               "collectStatsPageRank", "collectStatsPageRank_topK"]:
        assert fn in names


def test_library_exports_every_declared_symbol():
    lib = og.load()
    missing = [n for n in declared_functions() if not hasattr(lib, n)]
    assert not missing, missing


def test_ctypes_mirror_covers_every_declared_symbol():
    assert set(declared_functions()) <= set(og.SIGNATURES)


@pytest.fixture(scope="module")
def ref_layout():
    with open(os.path.join(GOLDEN, "abi_layout.json")) as f:
        return json.load(f)


def test_header_struct_layout_matches_reference(ref_layout):
    import importlib.util
    spec = importlib.util.spec_from_file_location("mk", os.path.join(GOLDEN, "make_abi_layout.py"))
    mk = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mk)
    mine = mk.layout(["opengraph_amd.h"], [os.path.join(ROOT, "include")])
    assert mine == ref_layout


def test_ctypes_struct_layout_matches_reference(ref_layout):
    types = {"Arguments": og.Arguments, "Vertex": og.Vertex, "EdgeList": og.EdgeList,
             "GraphCSR": og.GraphCSR, "PageRankStats": og.PageRankStats, "SPMVStats": og.SPMVStats,
             "PageRankCorrelationStats": og.PageRankCorrelationStats}
    for name, t in types.items():
        want = ref_layout[name]
        assert C.sizeof(t) == want["size"], name
        for f, off in want.items():
            if f != "size":
                assert getattr(t, f).offset == off, (name, f)
