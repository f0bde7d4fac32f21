"""The oracle's SpMV (-p 0..3 of SPMVGraphCSR) against the compiled reference.

Pins: for every golden case (graphbrew, LAW-enron, R-MAT-16; sort modes 0
and 1; pull, push, fixed-point pull, fixed-point push; 20 iterations) the
oracle's vector_output has the same bytes as the reference's, run on one
thread (tests/golden/make_golden_spmv.py).  Closed forms in double: every
output is iterations * weight * (sum of the inputs over the in-edges), with
weight 0.0001 for pull, push and fixed-point push and 7/65536 (the Q16.16
rounding of 0.0001, SPMV.c:870-876) for the fixed-point pull; a dropped
iteration, a wrong direction or a transposed CSR fails these.
"""
import hashlib

import numpy as np
import pytest

from conftest import golden_input
from oracle import pyoracle as po


def sha(a):
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


def test_spmv_cases_match_reference(golden):
    cases = golden["spmv"]
    assert len(cases) == 24
    built = {}
    for key, c in cases.items():
        gname, sort, code = key.split("/")
        sort, code = int(sort[4:]), int(code[1:])
        if (gname, sort) not in built:
            s, d = golden_input(gname, golden)
            built[(gname, sort)] = po.build_csr(s, d, sort, 4)
        y = po.spmv(built[(gname, sort)], code, c["iterations"])
        assert sha(y) == c["output_sha256"], key
        assert [float(v) for v in y[:8]] == c["head"], key


def _closed_form(csr, code, iterations):
    N = csr.num_vertices
    od = csr["f_out_degree"].astype(np.float64)
    x = np.where(od > 0, (1.0 / np.maximum(od, 1)).astype(np.float32), np.float32(0.001))
    x = x.astype(np.float64)
    src, dst = csr["f_src"].astype(np.int64), csr["f_dst"].astype(np.int64)
    s = np.bincount(dst, weights=x[src], minlength=N)
    w = 7.0 / 65536.0 if code == 2 else 0.0001
    return iterations * w * s, iterations * w * np.bincount(dst, minlength=N)


@pytest.mark.parametrize("code", [0, 1, 2, 3])
def test_spmv_closed_form(golden, code):
    s, d = golden_input("rmat16", golden)
    csr = po.build_csr(s, d, 0, 4)
    y = po.spmv(csr, code, 20).astype(np.float64)
    want, wdeg = _closed_form(csr, code, 20)
    terms = 20.0 * np.bincount(csr["f_dst"], minlength=csr.num_vertices)
    if code == 2:
        # Q16.16 inputs round by at most 2^-17 each, every product is floored to a
        # multiple of 2^-16: one unit per term at most, plus the final float rounding
        bound = terms * (7.0 / 65536.0 / 131072.0 + 1.0 / 65536.0) + 2.0 ** -23 * want + 1e-12
    elif code == 3:
        # Q32.32: one unit in 2^-32 per term, the weight's rounding (below 1e-6
        # relative) and the final float rounding
        bound = terms * 2.0 ** -32 + 2e-6 * want + 1e-12
    else:
        # a sequential float sum of n terms is within (n + 1) * 2^-24 of the exact one
        bound = (terms + 2.0) * 2.0 ** -24 * want + 1e-12
    assert np.all(np.abs(y - want) <= bound)
    assert np.all(y[wdeg == 0] == 0.0)
    if code != 2:
        # a dropped iteration (5 %) would leave the bound on the vertices with many terms
        assert np.mean(np.abs(y - want) <= 0.01 * want) > 0.99


def test_spmv_rejects_unknown_code(golden):
    s, d = golden_input("graphbrew", golden)
    with pytest.raises(ValueError):
        po.spmv(po.build_csr(s, d, 0, 1), 4, 1)


@pytest.mark.skipif(not po.ref_available(), reason="oracle/_ref not built")
@pytest.mark.parametrize("seed", [3, 11])
def test_spmv_live_reference(seed):
    rng = np.random.default_rng(seed)
    n, e = 300, 2500
    s = rng.integers(0, n, e).astype(np.uint32)
    d = (rng.zipf(1.6, e) % n).astype(np.uint32)
    csr = po.build_csr(s, d, 1, 2)
    for code in (0, 1, 2, 3):
        want, _ = po.ref_spmv(csr, code, 7, 1)
        assert np.array_equal(po.spmv(csr, code, 7).view(np.uint32), want.view(np.uint32)), code
