"""Add the degree-relabelling fixtures to tests/golden/golden.json.

Runs oracle/_ref/ref_harness (the reference's own graphCSRPreProcessingStep,
compiled from its sources by ``make -C oracle ref``) with the reference's
-l option 1..5 (out-, in-, in+out-degree sort; DBG by out- and in-degree,
reorder.c:275-574) on the golden graphs, sort modes 0 and 1, and records
under "reorder": sha256 of every CSR array of the relabelled graph, of the
forward label_array, and of the pull ranks (-p 0, epsilon 1e-4, 100
iterations) with their iteration count.  The existing entries of
golden.json are kept as they are.

Needs /root/reference (this container); usage:
python tests/golden/make_golden_reorder.py
"""
from __future__ import annotations

import json
import os
import subprocess
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, HERE)

from make_golden import HARNESS, sha  # noqa: E402

CSR_KEYS = [("f_src", "e"), ("f_dst", "e"), ("f_out_degree", "n"), ("f_in_degree", "n"),
            ("f_edges_idx", "n"), ("i_src", "e"), ("i_dst", "e"), ("i_out_degree", "n"),
            ("i_in_degree", "n"), ("i_edges_idx", "n")]


def run(path, sort, lmode, threads=4):
    with tempfile.TemporaryDirectory() as td:
        out = os.path.join(td, "o.bin")
        subprocess.run([HARNESS, path, str(sort), "0", "0.0001", "100", str(threads), out, "1", str(lmode)],
                       capture_output=True, text=True, check=True)
        raw = open(out, "rb").read()
    hdr = np.frombuffer(raw[:16], np.uint32)
    n, e, it = int(hdr[1]), int(hdr[2]), int(hdr[3])
    off = 32
    ranks = np.frombuffer(raw[off:off + 4 * n], np.float32)
    off += 8 * n                                   # ranks, then the rank-sorted labels
    arrays = {}
    for k, kind in CSR_KEYS:
        cnt = e if kind == "e" else n
        arrays[k] = np.frombuffer(raw[off:off + 4 * cnt], np.uint32)
        off += 4 * cnt
    labels = np.frombuffer(raw[off:off + 4 * n], np.uint32)
    assert off + 4 * n == len(raw)
    return n, e, it, ranks, arrays, labels


def main():
    if not os.path.exists(HARNESS):
        sys.exit("oracle/_ref/ref_harness missing: make -C oracle ref")
    import opengraph_amd as og
    from conftest import golden_input

    path = os.path.join(HERE, "golden.json")
    golden = json.load(open(path))
    tmp = tempfile.mkdtemp()
    files = {"graphbrew": os.path.join(HERE, "inputs", "graphbrew.bin"),
             "enron": os.path.join(HERE, "inputs", "enron.bin")}
    s, d = golden_input("rmat16", golden)
    files["rmat16"] = os.path.join(tmp, "rmat16.bin")
    og.write_bin(files["rmat16"], s, d)
    out = {}
    for gname, p in files.items():
        for sort in (0, 1):
            for lmode in (1, 2, 3, 4, 5):
                n, e, it, ranks, arrays, labels = run(p, sort, lmode)
                out[f"{gname}/sort{sort}/l{lmode}"] = dict(
                    num_vertices=n, num_edges=e, iterations=it, ranks_sha256=sha(ranks),
                    labels_sha256=sha(labels), **{k: sha(v) for k, v in arrays.items()})
                print(gname, sort, lmode, "->", it, flush=True)
    os.remove(files["rmat16"])
    golden["reorder"] = out
    with open(path, "w") as f:
        json.dump(golden, f, indent=1)


if __name__ == "__main__":
    main()
