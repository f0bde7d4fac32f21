"""The drop-in linked into a program made of the reference's own code.

oracle/_ref/ref_dropin (oracle/Makefile target ``dropin``, built where the
reference's sources are present) is the oracle's reference harness -- the
reference CLI's default Arguments, graphCSRPreProcessingStep
(graphCSR.c:164), pageRankGraphCSR (pageRank.c:871), then
radixSortEdgesByPageRank (reorder.c:141) and a dump -- linked with the
reference's own objects, in which the definitions of the library's entry
points (pageRank.o's seventeen PageRank functions,
graphCSRPreProcessingStep, sortRunAlgorithms) were made local with
objcopy, and with libopengraph_amd.so.  This is the link a maintainer makes
to the reference's driver (graphRun.c:1035-1040 calls pageRankGraphCSR;
graphRun.c itself also needs connectedComponents.c and libJudy, so the
harness stands in for it).  The program's CSR build and PageRank therefore
run in the library, on the GPU, while its edge-list reader,
radixSortEdgesByPageRank and graphCSRFree stay the reference's: their use of
the library's results is part of what is checked.
"""
import hashlib
import os
import subprocess

import numpy as np
import pytest

from conftest import GOLDEN, ROOT

DROPIN = os.path.join(ROOT, "oracle", "_ref", "ref_dropin")
# the entry points the program takes from the library (the reference's
# copies made local), and those of them that the harness calls itself
REPLACED = ("pageRankGraphCSR", "pageRankPullGraphCSR", "pageRankPushGraphCSR",
            "pageRankPullFixedPoint64BitGraphCSR", "pageRankDataDrivenPullPushGraphCSR",
            "newPageRankStatsGraphCSR", "freePageRankStats", "graphCSRPreProcessingStep", "sortRunAlgorithms")
CALLED = ("pageRankGraphCSR", "freePageRankStats", "graphCSRPreProcessingStep")
CSR_NAMES = ("f_src", "f_dst", "f_out_degree", "f_in_degree", "f_edges_idx",
             "i_src", "i_dst", "i_out_degree", "i_in_degree", "i_edges_idx")
# (graph, sort) -> variants run: those whose reference result does not
# depend on thread timing, with the drop-in's default options (the fp32 pull
# is binned: every row in CSR order); the data-driven pull (6) sums
# LAW-enron's rows of more than 16 in-edges as fixed trees, so it is held to
# the bits on graphbrew only
CODES = {("graphbrew", 0): (0, 2, 3, 4, 6, 9, 10, 11, 12, 13), ("enron", 0): (0, 2, 9, 13),
         ("graphbrew", 1): (0, 2), ("enron", 1): (0, 2)}

needs_dropin = pytest.mark.skipif(not os.path.exists(DROPIN),
                                  reason="oracle/_ref/ref_dropin not built (make -C oracle ref dropin)")


def _sha(a):
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


def _symbols(path):
    """(undefined, local text, global text) symbol names of a program"""
    out = subprocess.run(["nm", path], capture_output=True, text=True, check=True).stdout
    undef, local, glob = set(), set(), set()
    for line in out.splitlines():
        f = line.split()
        if len(f) == 2 and f[0] == "U":
            undef.add(f[1])
        elif len(f) == 3 and f[1] in "tT":
            (local if f[1] == "t" else glob).add(f[2])
    return undef, local, glob


@needs_dropin
def test_dropin_program_binds_to_library():
    """The program calls the library's definitions: the reference's copy of
    each replaced entry point is local to its object (no global definition
    is left in the program), the ones the harness calls are undefined
    symbols of the program, and libopengraph_amd.so is a needed library.
    What the library does not replace keeps the reference's definitions."""
    undef, local, glob = _symbols(DROPIN)
    for s in REPLACED:
        assert s in local and s not in glob, s
    for s in CALLED:
        assert s in undef, s
    assert {"graphCSRFree", "radixSortEdgesByPageRank"} <= glob
    dyn = subprocess.run(["readelf", "-d", DROPIN], capture_output=True, text=True, check=True).stdout
    assert "[libopengraph_amd.so]" in dyn


def _read_out(path, dump_csr):
    raw = open(path, "rb").read()
    hdr = np.frombuffer(raw, dtype=np.uint32, count=4)
    assert hdr[0] == 0x4652474F
    n, e, iters = int(hdr[1]), int(hdr[2]), int(hdr[3])
    err, _ = np.frombuffer(raw, dtype=np.float64, count=2, offset=16)
    o = 32
    ranks = np.frombuffer(raw, dtype=np.float32, count=n, offset=o)
    o += 4 * n
    labels = np.frombuffer(raw, dtype=np.uint32, count=n, offset=o)
    o += 4 * n
    csr = {}
    if dump_csr:
        for side in "fi":
            for name, cnt in (("src", e), ("dst", e), ("out_degree", n), ("in_degree", n), ("edges_idx", n)):
                csr[f"{side}_{name}"] = np.frombuffer(raw, dtype=np.uint32, count=cnt, offset=o)
                o += 4 * cnt
    assert o == len(raw)
    return iters, float(err), ranks, labels, csr


@pytest.mark.gpu
@needs_dropin
def test_dropin_program_matches_reference(golden, tmp_path):
    """The program's ranks, iteration count, error_total, PageRank order and
    CSR arrays equal the reference's own run of the same case (the golden
    cases, made by the reference harness) bit for bit, and both entry points
    report from the library (OGPU_TRACE)."""
    env = dict(os.environ, OGPU_TRACE="1", OMP_NUM_THREADS="4")
    n = 0
    for c in golden["cases"]:
        name, code = c["graph"], c["code"]
        if code not in CODES.get((name, c["sort"]), ()) or c["epsilon"] != 1e-4:
            continue
        out = str(tmp_path / f"{name}_{c['sort']}_{code}.bin")
        dump = code == 0
        cmd = [DROPIN, os.path.join(GOLDEN, "inputs", f"{name}.bin"), str(c["sort"]), str(code),
               repr(c["epsilon"]), str(c["max_iter"]), "4", out, "1" if dump else "0"]
        print("case", name, "sort", c["sort"], "code", code, flush=True)
        p = subprocess.run(cmd, capture_output=True, text=True, env=env, timeout=60)
        assert p.returncode == 0, p.stderr[-2000:]
        assert "opengraph_amd: graphCSRPreProcessingStep on HIP device" in p.stderr, p.stderr[-2000:]
        assert "opengraph_amd: pageRankGraphCSR on HIP device" in p.stderr, p.stderr[-2000:]
        iters, err, ranks, labels, csr = _read_out(out, dump)
        assert iters == c["iterations"], c
        assert _sha(ranks) == c["ranks_sha256"], c
        assert _sha(labels) == c["labels_sha256"], c
        assert err == pytest.approx(c["error_total"], rel=1e-9, abs=1e-300), c
        gc = golden["csr"][f"{name}/sort{c['sort']}"]   # sha256 of each of the reference's arrays
        for k in csr:
            assert _sha(csr[k]) == gc[k], (name, c["sort"], k)
        n += 1
    assert n == 18
