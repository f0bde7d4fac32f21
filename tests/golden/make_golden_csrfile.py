"""Add the .csr file fixtures (graphCSR.c:380-558) to tests/golden/.

Builds the golden graphs' CSR with the oracle (itself pinned bit-exact
against the reference's CSR digests) and has the reference's own compiled
writeToBinFileGraphCSR (oracle/_ref/libogref.so) write it; records under
"csr_file" in golden.json the size and sha256 of each file, and keeps
graphbrew's sort-0 file (a few hundred bytes) as inputs/graphbrew.sort0.csr
for the reader tests.  Needs oracle/_ref (this container); usage:
python tests/golden/make_golden_csrfile.py
"""
from __future__ import annotations

import hashlib
import json
import os
import shutil
import sys
import tempfile

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.dirname(HERE))


def main():
    from oracle import pyoracle as po
    from conftest import golden_input

    if not po.ref_available():
        sys.exit("oracle/_ref missing: make -C oracle ref")
    path = os.path.join(HERE, "golden.json")
    golden = json.load(open(path))
    out = {}
    tmp = tempfile.mkdtemp()
    for gname in ("graphbrew", "enron", "rmat16"):
        s, d = golden_input(gname, golden)
        for sort in (0, 1):
            o = po.build_csr(s, d, sort, 4)
            base = os.path.join(tmp, f"{gname}.sort{sort}")
            po.ref_write_csr(o, base)
            raw = open(base + ".csr", "rb").read()
            out[f"{gname}/sort{sort}"] = dict(bytes=len(raw), sha256=hashlib.sha256(raw).hexdigest())
            if gname == "graphbrew" and sort == 0:
                shutil.copy(base + ".csr", os.path.join(HERE, "inputs", "graphbrew.sort0.csr"))
            print(gname, sort, len(raw), flush=True)
    shutil.rmtree(tmp)
    golden["csr_file"] = out
    with open(path, "w") as f:
        json.dump(golden, f, indent=1)


if __name__ == "__main__":
    main()
