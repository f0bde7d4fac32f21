"""Write the golden fixtures of tests/golden/ from the compiled reference.

Runs oracle/_ref/ref_harness -- the reference's own graphCSRPreProcessingStep
and pageRankGraphCSR, compiled from its sources by ``make -C oracle ref`` --
on the reference's test graphs and on a seeded R-MAT scale-16 graph written
by this library's generator, and records for each (graph, sort, -p, epsilon,
iterations) case:

* sha256 of the rank vector bytes, the iteration count, error_total, the
  per-iteration activeVertices rows and the PR sum printed by the reference;
* sha256 of each CSR array per (graph, sort);
* full rank vectors for the pull and fixed-point pull cases (ranks.npz).

Needs /root/reference (this container); the fixtures it writes are data and
travel with the repository.  Usage: python tests/golden/make_golden.py
"""
from __future__ import annotations

import hashlib
import json
import os
import re
import subprocess
import sys
import tempfile

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

HARNESS = os.path.join(ROOT, "oracle", "_ref", "ref_harness")
OUT = os.path.dirname(os.path.abspath(__file__))

ROW = re.compile(r"^\|\s*(\d+)\s*\|\s*(\d+)\s*\|\s*([0-9.]+)\s*\|\s*([0-9.]+)\s*\|")
FINAL = re.compile(r"^\|\s*(\d+)\s*\|\s*([0-9.]+)\s*\|\s*([0-9.]+)\s*\|\s*([0-9.]+)\s*\|")


def sha(a: np.ndarray) -> str:
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


def run_case(path, sort, code, eps, iters, threads=4, dump_csr=False):
    with tempfile.TemporaryDirectory() as td:
        out = os.path.join(td, "o.bin")
        p = subprocess.run([HARNESS, path, str(sort), str(code), repr(eps), str(iters), str(threads),
                            out, "1" if dump_csr else "0"], capture_output=True, text=True, check=True)
        raw = open(out, "rb").read()
    hdr = np.frombuffer(raw[:16], np.uint32)
    n, e, it = int(hdr[1]), int(hdr[2]), int(hdr[3])
    err, tt = np.frombuffer(raw[16:32], np.float64)
    off = 32
    ranks = np.frombuffer(raw[off:off + 4 * n], np.float32).copy()
    off += 4 * n
    labels = np.frombuffer(raw[off:off + 4 * n], np.uint32).copy()
    off += 4 * n
    csr = None
    if dump_csr:
        csr = {}
        for k, cnt in [("f_src", e), ("f_dst", e), ("f_out_degree", n), ("f_in_degree", n),
                       ("f_edges_idx", n), ("i_src", e), ("i_dst", e), ("i_out_degree", n),
                       ("i_in_degree", n), ("i_edges_idx", n)]:
            csr[k] = np.frombuffer(raw[off:off + 4 * cnt], np.uint32).copy()
            off += 4 * cnt
    # per-iteration rows: between the "Iteration" header and the final table
    lines = p.stdout.splitlines()
    start = max(i for i, l in enumerate(lines) if "Iteration " in l and "Active" in l)
    active, errs = [], []
    final = None
    for l in lines[start + 1:]:
        if "Iterations" in l and "PR Sum" in l:
            break
        m = ROW.match(l)
        if m:
            active.append(int(m.group(2)))
            errs.append(float(m.group(3)))
    for l in lines:
        m = FINAL.match(l)
        if m and "." in m.group(2):
            final = float(m.group(2))
    return dict(n=n, e=e, iterations=it, error_total=float(err), ranks=ranks, labels=labels,
                active=active, errors=errs, pr_sum=final, csr=csr)


def main():
    if not os.path.exists(HARNESS):
        sys.exit("oracle/_ref/ref_harness missing: make -C oracle ref")
    import opengraph_amd as og

    tmp = tempfile.mkdtemp()
    rmat = os.path.join(tmp, "rmat16.bin")
    src, dst = og.rmat_host(16, 16 << 16, 0.57, 0.19, 0.19, 27, True)
    og.write_bin(rmat, src, dst)
    graphs = {
        "graphbrew": os.path.join(OUT, "inputs", "graphbrew.bin"),
        "enron": os.path.join(OUT, "inputs", "enron.bin"),
        "rmat16": rmat,
    }
    cases = []
    for sort in (0, 1):
        for code in range(14):
            for eps in (1e-4, 1e-6):
                cases.append(("graphbrew", sort, code, eps, 100))
        for code in (0, 1, 2, 3, 4, 5, 6, 9, 10, 11, 12, 13):
            cases.append(("enron", sort, code, 1e-4, 100))
        cases.append(("enron", sort, 0, 1e-6, 100))
        cases.append(("enron", sort, 0, 0.0, 10))
        for code in (0, 2, 4, 6, 9):
            cases.append(("rmat16", sort, code, 1e-6, 100))
        cases.append(("rmat16", sort, 0, 1e-4, 100))
        cases.append(("rmat16", sort, 0, 0.0, 12))

    golden = {"inputs": {"rmat16": {"generator": "rmat", "scale": 16, "edges": 16 << 16, "a": 0.57,
                                     "b": 0.19, "c": 0.19, "seed": 27, "permute": True,
                                     "bin_sha256": hashlib.sha256(open(rmat, "rb").read()).hexdigest()}},
              "csr": {}, "cases": []}
    full = {}
    for gname, sort, code, eps, iters in cases:
        key_csr = f"{gname}/sort{sort}"
        need_csr = key_csr not in golden["csr"]
        r = run_case(graphs[gname], sort, code, eps, iters, dump_csr=need_csr)
        if need_csr:
            golden["csr"][key_csr] = {"num_vertices": r["n"], "num_edges": r["e"],
                                      **{k: sha(v) for k, v in r["csr"].items()}}
            if gname == "graphbrew":
                for k, v in r["csr"].items():
                    full[f"csr/{key_csr}/{k}"] = v
        case = dict(graph=gname, sort=sort, code=code, epsilon=eps, max_iter=iters,
                    iterations=r["iterations"], error_total=r["error_total"],
                    ranks_sha256=sha(r["ranks"]), labels_sha256=sha(r["labels"]),
                    active=r["active"], errors=r["errors"], pr_sum=r["pr_sum"],
                    num_vertices=r["n"], num_edges=r["e"])
        golden["cases"].append(case)
        if sort == 0 and (code in (0, 1, 2, 5, 7, 8) or gname == "graphbrew"):
            full[f"ranks/{gname}/sort{sort}/p{code}/eps{eps}/it{iters}"] = r["ranks"]
        print(gname, sort, code, eps, iters, "->", r["iterations"], len(r["active"]), flush=True)
    with open(os.path.join(OUT, "golden.json"), "w") as f:
        json.dump(golden, f, indent=1)
    np.savez_compressed(os.path.join(OUT, "ranks.npz"), **full)
    os.remove(rmat)


if __name__ == "__main__":
    main()
