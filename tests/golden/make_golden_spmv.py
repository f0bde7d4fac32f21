"""Add the SpMV fixtures (SPMV.c:588-1068) to tests/golden/golden.json.

Builds the golden graphs' CSR with the oracle (itself pinned bit-exact
against the reference's CSR digests) and runs the reference's own compiled
SPMVGraphCSR (oracle/_ref/libogref.so) with -p 0 (pull), 1 (push),
2 (fixed-point pull) and 3 (fixed-point push), 20 iterations, on one thread
(so the push's atomic additions happen in source order).  Records under
"spmv" the sha256 of vector_output, its first values, and whether 4
threads gave the same bytes (they must for 0, 2 and 3: ordered float sums
per row and integer sums).  Needs oracle/_ref (this container); usage:
python tests/golden/make_golden_spmv.py
"""
from __future__ import annotations

import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, HERE)
sys.path.insert(0, os.path.dirname(HERE))

from make_golden import sha  # noqa: E402

ITERATIONS = 20


def main():
    from oracle import pyoracle as po
    from conftest import golden_input

    if not po.ref_available():
        sys.exit("oracle/_ref missing: make -C oracle ref")
    path = os.path.join(HERE, "golden.json")
    golden = json.load(open(path))
    out = {}
    for gname in ("graphbrew", "enron", "rmat16"):
        s, d = golden_input(gname, golden)
        for sort in (0, 1):
            csr = po.build_csr(s, d, sort, 4)
            for code in (0, 1, 2, 3):
                y1, _ = po.ref_spmv(csr, code, ITERATIONS, 1)
                y4, _ = po.ref_spmv(csr, code, ITERATIONS, 4)
                out[f"{gname}/sort{sort}/p{code}"] = dict(
                    iterations=ITERATIONS, num_vertices=csr.num_vertices,
                    output_sha256=sha(y1), head=[float(v) for v in y1[:8]],
                    sum=float(np.sum(y1, dtype=np.float64)),
                    threads4_same=bool(np.array_equal(y1.view(np.uint32), y4.view(np.uint32))))
                print(gname, sort, code, out[f"{gname}/sort{sort}/p{code}"]["threads4_same"], flush=True)
    golden["spmv"] = out
    with open(path, "w") as f:
        json.dump(golden, f, indent=1)


if __name__ == "__main__":
    main()
