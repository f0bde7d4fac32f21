"""The oracle's degree relabelling (-l 1..5) against the compiled reference.

Pins: for every golden case (graphbrew, LAW-enron, R-MAT-16; sort modes 0
and 1; -l 1 out-, 2 in-, 3 in+out-degree sort, 4/5 DBG by out-/in-degree)
the labels, every CSR array of the relabelled graph and the pull ranks and
iteration count equal the reference's (tests/golden/make_golden_reorder.py).
Invariants on random graphs: the labels are a permutation; along the
labels, degrees never increase (degree sort) and DBG buckets never
increase, with ties in the order the reference gives them (decreasing id
for the degree sort, increasing id within a DBG bucket).
"""
import hashlib

import numpy as np
import pytest

from conftest import golden_input
from oracle import pyoracle as po

CSR_KEYS = ["f_src", "f_dst", "f_out_degree", "f_in_degree", "f_edges_idx",
            "i_src", "i_dst", "i_out_degree", "i_in_degree", "i_edges_idx"]


def sha(a):
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


def test_reorder_cases_match_reference(golden):
    assert len(golden["reorder"]) == 30
    for key, c in golden["reorder"].items():
        gname, sort, lm = key.split("/")
        sort, lmode = int(sort[4:]), int(lm[1:])
        s, d = golden_input(gname, golden)
        N = c["num_vertices"]
        lab = po.reorder_labels(s, d, N, lmode)
        assert sha(lab) == c["labels_sha256"], key
        o = po.build_csr(lab[s], lab[d], sort, 4, N)
        for k in CSR_KEYS:
            assert sha(o[k]) == c[k], (key, k)
        r = po.pagerank(o, 0, 1e-4, 100, 4)
        assert r["iterations"] == c["iterations"] and sha(r["ranks"]) == c["ranks_sha256"], key


def _degrees(s, d, N, lmode):
    out = np.bincount(s, minlength=N)
    inn = np.bincount(d, minlength=N)
    return {1: out, 2: inn, 3: out + inn, 4: out, 5: inn}[lmode]


@pytest.mark.parametrize("lmode", [1, 2, 3, 4, 5])
def test_reorder_invariants(lmode):
    rng = np.random.default_rng(lmode)
    N, E = 3000, 20000
    s = (rng.pareto(1.2, E) * 10).astype(np.uint32) % N
    d = rng.integers(0, N, E, dtype=np.uint32)
    lab = po.reorder_labels(s, d, N, lmode)
    assert np.array_equal(np.sort(lab), np.arange(N, dtype=np.uint32))
    deg = _degrees(s, d, N, lmode)
    by_label = np.empty(N, np.int64)
    by_label[lab] = np.arange(N)            # vertex holding each label
    if lmode <= 3:
        dl = deg[by_label]
        assert np.all(dl[:-1] >= dl[1:])
        ties = dl[:-1] == dl[1:]
        assert np.all(by_label[:-1][ties] > by_label[1:][ties])
    else:
        avg = E // N
        th = [1 if avg <= 1 else avg // 2]
        for _ in range(9):
            th.append(th[-1] * 2)
        th.append(2**32 - 1)
        bucket = np.searchsorted(np.array(th, np.int64), deg, side="left")
        bl = bucket[by_label]
        assert np.all(bl[:-1] >= bl[1:])
        same = bl[:-1] == bl[1:]
        assert np.all(by_label[:-1][same] < by_label[1:][same])


def test_reorder_rejects_other_modes():
    s = np.array([0, 1], np.uint32)
    d = np.array([1, 0], np.uint32)
    for lmode in (0, 6, 10, 11):
        with pytest.raises(ValueError):
            po.reorder_labels(s, d, 2, lmode)
