"""Add the drop-in output-contract fixtures (pageRankGraphCSR's printed
table and PageRankStats fields, pageRank.c:47-75, :968-974, :1031,
:1049-1053 and the same lines of every variant) to tests/golden/.

For graphbrew (every -p code 0..13) and LAW-enron (sort 0, the codes whose
reference result does not depend on thread timing) the reference's own
compiled pageRankGraphCSR (oracle/_ref/libogref.so, one thread) runs with
epsilon 1e-4 and 20 iterations; its stdout, with the time column blanked
(conftest.strip_times), is kept as inputs/dropin_<graph>_p<code>.stdout and
the stats fields under "dropin" in golden.json.  Needs oracle/_ref (this
container); usage: python tests/golden/make_golden_dropin.py
"""
from __future__ import annotations

import json
import os
import sys
import tempfile

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.dirname(HERE))

CASES = [("graphbrew", c) for c in range(14)] + [("enron", c) for c in (0, 2, 3, 4, 6, 9, 10, 11, 12, 13)]
EPS, ITERS = 1e-4, 20


def main():
    from oracle import pyoracle as po
    from conftest import golden_input, strip_times

    if not po.ref_available():
        sys.exit("oracle/_ref missing: make -C oracle ref")
    path = os.path.join(HERE, "golden.json")
    golden = json.load(open(path))
    out = {}
    csrs = {}
    for name, code in CASES:
        if name not in csrs:
            s, d = golden_input(name, golden)
            csrs[name] = po.build_csr(s, d, 0, 4)
        with tempfile.TemporaryDirectory() as td:
            tmp = os.path.join(td, "out.txt")
            _, fields = po.ref_pagerank_stdout(csrs[name], code, EPS, ITERS, 1, tmp)
            text = open(tmp).read()
        fn = f"dropin_{name}_p{code}.stdout"
        with open(os.path.join(HERE, "inputs", fn), "w") as f:
            f.write(strip_times(text) + "\n")
        out[f"{name}/p{code}"] = dict(fields, graph=name, sort=0, code=code, epsilon=EPS, max_iter=ITERS,
                                      stdout=fn)
        print(name, code, fields["iterations"], flush=True)
    golden["dropin"] = out
    with open(path, "w") as f:
        json.dump(golden, f, indent=1)


if __name__ == "__main__":
    main()
