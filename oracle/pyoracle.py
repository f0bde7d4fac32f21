"""ctypes wrapper of the CPU oracle (TEST INFRASTRUCTURE ONLY).

Loads oracle/libogoracle.so (the C/OpenMP restatement of the reference,
og_oracle.c) and, when present, oracle/_ref/libogref.so (the reference's own
PageRank objects compiled from its sources, ref_lib.c).  Only tests/,
__graft_entry__.smoke() and bench.py's cpu_baseline leg use this module, and
only as the checker / the timed CPU baseline; the product never imports it.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
# OG_ORACLE_LIB names another build of the same restatement (the sanitized one)
ORACLE_LIB = os.environ.get("OG_ORACLE_LIB") or os.path.join(_HERE, "libogoracle.so")
REF_LIB = os.path.join(_HERE, "_ref", "libogref.so")
REF_HARNESS = os.path.join(_HERE, "_ref", "ref_harness")

u32p = C.POINTER(C.c_uint32)
f64p = C.POINTER(C.c_double)
f32p = C.POINTER(C.c_float)

CSR_FIELDS = ["f_src", "f_dst", "f_out_degree", "f_in_degree", "f_edges_idx",
              "i_src", "i_dst", "i_out_degree", "i_in_degree", "i_edges_idx"]


class OrcCSR(C.Structure):
    _fields_ = [("num_vertices", C.c_uint32), ("num_edges", C.c_uint32),
                ("f_src", u32p), ("f_dst", u32p),
                ("f_out_degree", u32p), ("f_in_degree", u32p), ("f_edges_idx", u32p),
                ("i_src", u32p), ("i_dst", u32p),
                ("i_out_degree", u32p), ("i_in_degree", u32p), ("i_edges_idx", u32p)]


class OrcResult(C.Structure):
    _fields_ = [("iterations", C.c_uint32), ("rows", C.c_uint32), ("error_total", C.c_double),
                ("pr_sum", C.c_double), ("seconds_total", C.c_double)]


class OrcCorrStats(C.Structure):
    """orc_corr_stats = graphStats.h:28-37 PageRankCorrelationStats."""
    _fields_ = [("levenshtein_distance", C.c_uint32), ("float_Kendall", C.c_float),
                ("real_Kendall", C.c_float), ("intersection", C.c_uint32), ("mismatch", C.c_uint32),
                ("avg_error_float", C.c_double), ("avg_error_relative", C.c_double)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


_lib = None
_ref = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(ORACLE_LIB):
            raise RuntimeError(f"{ORACLE_LIB} missing: make -C oracle")
        L = C.CDLL(ORACLE_LIB)
        L.orc_build_csr.argtypes = [u32p, u32p, C.c_uint32, C.c_uint32, C.c_int, C.POINTER(OrcCSR)]
        L.orc_build_csr.restype = C.c_int
        L.orc_build_csr_n.argtypes = [u32p, u32p, C.c_uint32, C.c_uint32, C.c_uint32, C.c_int,
                                      C.POINTER(OrcCSR)]
        L.orc_build_csr_n.restype = C.c_int
        L.orc_free_csr.argtypes = [C.POINTER(OrcCSR)]
        L.orc_pagerank.argtypes = [C.POINTER(OrcCSR), C.c_uint32, C.c_double, C.c_uint32, C.c_int,
                                   f32p, C.c_int, f32p, u32p, f64p, C.POINTER(OrcResult)]
        L.orc_pagerank.restype = C.c_int
        L.orc_pagerank_fp64.argtypes = [C.POINTER(OrcCSR), C.c_double, C.c_double, C.c_uint32,
                                        C.c_int, f64p, f64p]
        L.orc_pagerank_fp64.restype = C.c_uint32
        L.orc_ppr_fp64.argtypes = [C.POINTER(OrcCSR), C.c_uint32, C.c_double, C.c_uint32, f64p, f64p]
        L.orc_ppr_fp64.restype = C.c_uint32
        L.orc_ppr_f32.argtypes = [C.POINTER(OrcCSR), C.c_uint32, C.c_double, C.c_uint32, f32p, f64p]
        L.orc_ppr_f32.restype = C.c_uint32
        L.orc_mlp_features.argtypes = [C.POINTER(OrcCSR), f32p]
        L.orc_mlp_forward.argtypes = [f32p, f32p, C.c_uint32, f64p]
        L.orc_mlp_loss.argtypes = [f32p, f32p, f32p, C.c_uint32]
        L.orc_mlp_loss.restype = C.c_double
        L.orc_mlp_grad.argtypes = [f32p, f32p, f32p, C.c_uint32, f64p]
        L.orc_mlp_grad.restype = C.c_double
        L.orc_sort_labels_by_rank.argtypes = [f32p, u32p, C.c_uint32]
        L.orc_stable_sort_pairs.argtypes = [u32p, u32p, C.c_uint32, C.c_int]
        L.orc_stable_sort_pairs.restype = C.c_int
        L.orc_reorder_labels.argtypes = [u32p, u32p, C.c_uint32, C.c_uint32, C.c_uint32, u32p]
        L.orc_reorder_labels.restype = C.c_int
        L.orc_spmv.argtypes = [C.POINTER(OrcCSR), C.c_uint32, C.c_uint32, f32p]
        L.orc_rank_stats_topk.argtypes = [f32p, u32p, f32p, u32p, C.c_uint32, C.c_uint32, C.c_uint32,
                                          C.POINTER(OrcCorrStats)]
        L.orc_rank_stats_topk.restype = C.c_int
        L.orc_spmv.restype = C.c_int
        _lib = L
    return _lib


def ref_available() -> bool:
    return os.path.exists(REF_LIB)


def ref_lib():
    global _ref
    if _ref is None:
        R = C.CDLL(REF_LIB)
        R.ogref_pagerank.argtypes = [C.c_uint32, C.c_uint32] + [u32p] * 10 + [
            C.c_uint32, C.c_double, C.c_uint32, C.c_int, f32p, u32p, f64p, f64p]
        R.ogref_pagerank.restype = C.c_int
        R.ogref_pagerank_stdout.argtypes = [C.c_uint32, C.c_uint32] + [u32p] * 10 + [
            C.c_uint32, C.c_double, C.c_uint32, C.c_int, C.c_char_p, f32p, f64p]
        R.ogref_pagerank_stdout.restype = C.c_int
        R.ogref_reorder.argtypes = [C.c_uint32, C.c_uint32, u32p, u32p, C.c_uint32, C.c_int, u32p, f64p]
        R.ogref_reorder.restype = C.c_int
        R.ogref_csr_build.argtypes = [C.c_uint32, C.c_uint32, u32p, u32p, C.c_uint32, C.c_int,
                                      C.POINTER(u32p), f64p]
        R.ogref_csr_build.restype = C.c_int
        R.ogref_write_csr.argtypes = [C.c_uint32, C.c_uint32] + [u32p] * 10 + [C.c_char_p]
        R.ogref_write_csr.restype = C.c_int
        R.ogref_spmv.argtypes = [C.c_uint32, C.c_uint32] + [u32p] * 10 + [
            C.c_uint32, C.c_uint32, C.c_int, f32p, f64p]
        R.ogref_spmv.restype = C.c_int
        R.ogref_rank_stats_batch.argtypes = [C.c_uint32, f32p, u32p, f32p, u32p, C.c_uint32, u32p, u32p,
                                             u32p, f32p, f64p, f64p]
        R.ogref_rank_stats_batch.restype = C.c_int
        R.ogref_collect_stats.argtypes = [C.c_uint32, f32p, u32p, f32p, u32p, C.c_char_p, C.c_uint32,
                                          C.c_uint32, C.c_uint32, C.c_uint32, C.c_double, C.c_uint32,
                                          C.c_int, C.c_char_p, f64p]
        R.ogref_collect_stats.restype = C.c_int
        _ref = R
    return _ref


def _p(a, t=u32p):
    return a.ctypes.data_as(t)


class CSR:
    """Oracle CSR: numpy copies of the ten reference arrays."""

    def __init__(self, arrays: dict, num_vertices: int, num_edges: int):
        self.a = {k: np.ascontiguousarray(arrays[k], dtype=np.uint32) for k in CSR_FIELDS}
        self.num_vertices = num_vertices
        self.num_edges = num_edges
        self.s = OrcCSR(num_vertices, num_edges, *[_p(self.a[k]) for k in CSR_FIELDS])

    def __getitem__(self, k):
        return self.a[k]


def build_csr(src, dst, sort_mode=0, nthreads=0, min_vertices=0) -> CSR:
    L = lib()
    src = np.ascontiguousarray(src, dtype=np.uint32)
    dst = np.ascontiguousarray(dst, dtype=np.uint32)
    g = OrcCSR()
    rc = L.orc_build_csr_n(_p(src), _p(dst), len(src), min_vertices, sort_mode, nthreads, C.byref(g))
    if rc:
        raise RuntimeError("orc_build_csr failed")
    N, E = g.num_vertices, g.num_edges
    out = {}
    for k in CSR_FIELDS:
        n = E if k in ("f_src", "f_dst", "i_src", "i_dst") else N
        ptr = getattr(g, k)
        out[k] = np.ctypeslib.as_array(ptr, shape=(n,)).copy() if n else np.zeros(0, np.uint32)
    L.orc_free_csr(C.byref(g))
    return CSR(out, N, E)


def pagerank(csr: CSR, code=0, epsilon=1e-4, max_iter=20, nthreads=0, init=None, stop_l1=False):
    """Returns dict(ranks, iterations, rows, error_total, pr_sum, seconds, active, errors)."""
    L = lib()
    N = csr.num_vertices
    ranks = np.zeros(N, np.float32)
    act = np.zeros(max(max_iter, 1), np.uint32)
    err = np.zeros(max(max_iter, 1), np.float64)
    res = OrcResult()
    ip = None
    if init is not None:
        init = np.ascontiguousarray(init, dtype=np.float32)
        ip = _p(init, f32p)
    rc = L.orc_pagerank(C.byref(csr.s), code, epsilon, max_iter, nthreads, ip, int(stop_l1),
                        _p(ranks, f32p), _p(act), _p(err, f64p), C.byref(res))
    if rc:
        raise RuntimeError(f"orc_pagerank failed ({rc})")
    return dict(ranks=ranks, iterations=res.iterations, rows=res.rows,
                error_total=res.error_total, pr_sum=res.pr_sum, seconds=res.seconds_total,
                active=act[:res.rows].copy(), errors=err[:res.rows].copy())


def pagerank_fp64(csr: CSR, init=0.15, tol=1e-13, max_iter=1000, nthreads=0):
    L = lib()
    out = np.zeros(csr.num_vertices, np.float64)
    l1 = C.c_double()
    it = L.orc_pagerank_fp64(C.byref(csr.s), init, tol, max_iter, nthreads, _p(out, f64p), C.byref(l1))
    return out, it, l1.value


def ppr_fp64(csr: CSR, source, tol=1e-12, max_iter=1000):
    L = lib()
    out = np.zeros(csr.num_vertices, np.float64)
    l1 = C.c_double()
    it = L.orc_ppr_fp64(C.byref(csr.s), source, tol, max_iter, _p(out, f64p), C.byref(l1))
    return out, it, l1.value


def ppr_f32(csr: CSR, source, tol=1e-6, max_iter=100):
    """fp32 personalized PageRank in the pull kernel's arithmetic: (x, iterations, l1 per iteration)."""
    L = lib()
    out = np.zeros(csr.num_vertices, np.float32)
    hist = np.zeros(max(max_iter, 1), np.float64)
    it = L.orc_ppr_f32(C.byref(csr.s), source, tol, max_iter, _p(out, f32p), _p(hist, f64p))
    return out, it, hist[:it]


MLP_PARAMS = 4545


def mlp_features(csr: CSR):
    out = np.zeros((csr.num_vertices, 4), np.float32)
    lib().orc_mlp_features(C.byref(csr.s), _p(out, f32p))
    return out


def mlp_forward(params, feat):
    p = np.ascontiguousarray(params, dtype=np.float32)
    f = np.ascontiguousarray(feat, dtype=np.float32)
    out = np.zeros(len(f), np.float64)
    lib().orc_mlp_forward(_p(p, f32p), _p(f, f32p), len(f), _p(out, f64p))
    return out


def mlp_loss(params, feat, y):
    p = np.ascontiguousarray(params, dtype=np.float32)
    f = np.ascontiguousarray(feat, dtype=np.float32)
    yy = np.ascontiguousarray(y, dtype=np.float32)
    return lib().orc_mlp_loss(_p(p, f32p), _p(f, f32p), _p(yy, f32p), len(f))


def mlp_grad(params, feat, y):
    """(loss, gradient[4545]) in double precision."""
    p = np.ascontiguousarray(params, dtype=np.float32)
    f = np.ascontiguousarray(feat, dtype=np.float32)
    yy = np.ascontiguousarray(y, dtype=np.float32)
    g = np.zeros(MLP_PARAMS, np.float64)
    loss = lib().orc_mlp_grad(_p(p, f32p), _p(f, f32p), _p(yy, f32p), len(f), _p(g, f64p))
    return loss, g


def sort_labels_by_rank(ranks, labels):
    r = np.ascontiguousarray(ranks, dtype=np.float32)
    lab = np.ascontiguousarray(labels, dtype=np.uint32).copy()
    lib().orc_sort_labels_by_rank(_p(r, f32p), _p(lab), len(lab))
    return lab


def reorder_labels(src, dst, num_vertices, lmode):
    """New label of every vertex for the reference's -l lmode (1..5)."""
    s = np.ascontiguousarray(src, dtype=np.uint32)
    d = np.ascontiguousarray(dst, dtype=np.uint32)
    lab = np.zeros(num_vertices, np.uint32)
    if lib().orc_reorder_labels(_p(s), _p(d), len(s), num_vertices, lmode, _p(lab)):
        raise ValueError(f"lmode {lmode} is not a degree or DBG relabelling")
    return lab


def spmv(csr: CSR, code=0, iterations=20):
    """vector_output of the reference's SPMVGraphCSR (-p code 0..3) for
    `iterations` iterations, run sequentially."""
    out = np.zeros(csr.num_vertices, np.float32)
    if lib().orc_spmv(C.byref(csr.s), code, iterations, _p(out, f32p)):
        raise ValueError(f"SpMV code {code} is not 0..3")
    return out


def ref_pagerank(csr: CSR, code=0, epsilon=1e-4, iterations=20, threads=1):
    """Run the reference's own compiled pageRankGraphCSR on these arrays."""
    R = ref_lib()
    N, E = csr.num_vertices, csr.num_edges
    ranks = np.zeros(N, np.float32)
    it = C.c_uint32()
    err = C.c_double()
    sec = C.c_double()
    arrays = [csr.a[k] for k in ["f_out_degree", "f_in_degree", "f_edges_idx", "f_src", "f_dst",
                                  "i_out_degree", "i_in_degree", "i_edges_idx", "i_src", "i_dst"]]
    rc = R.ogref_pagerank(N, E, *[_p(x) for x in arrays], code, epsilon, iterations, threads,
                          _p(ranks, f32p), C.byref(it), C.byref(err), C.byref(sec))
    if rc:
        raise RuntimeError("ogref_pagerank failed")
    return dict(ranks=ranks, iterations=it.value, error_total=err.value, seconds=sec.value)


def ref_pagerank_stdout(csr: CSR, code, epsilon, iterations, threads, stdout_path):
    """The reference's own pageRankGraphCSR with its printed table written to
    stdout_path; returns the ranks and the PageRankStats fields."""
    R = ref_lib()
    N, E = csr.num_vertices, csr.num_edges
    arrays = [csr.a[k] for k in ["f_out_degree", "f_in_degree", "f_edges_idx", "f_src", "f_dst",
                                  "i_out_degree", "i_in_degree", "i_edges_idx", "i_src", "i_dst"]]
    ranks = np.zeros(N, np.float32)
    f = np.zeros(7, np.float64)
    rc = R.ogref_pagerank_stdout(N, E, *[_p(x) for x in arrays], code, epsilon, iterations, threads,
                                 stdout_path.encode(), _p(ranks, f32p), _p(f, f64p))
    if rc:
        raise RuntimeError("ogref_pagerank_stdout failed")
    return ranks, dict(damp=float(f[0]), base_pr=float(f[1]), iterations=int(f[2]),
                       num_vertices=int(f[3]), error_total=float(f[4]), realRanks_identity=bool(f[6]))


def ref_write_csr(csr: CSR, fname: str):
    """The reference's own writeToBinFileGraphCSR: writes fname + '.csr'."""
    arrays = [csr.a[k] for k in ["f_out_degree", "f_in_degree", "f_edges_idx", "f_src", "f_dst",
                                  "i_out_degree", "i_in_degree", "i_edges_idx", "i_src", "i_dst"]]
    if ref_lib().ogref_write_csr(csr.num_vertices, csr.num_edges, *[_p(x) for x in arrays], fname.encode()):
        raise RuntimeError("ogref_write_csr failed")


def ref_reorder(src, dst, num_vertices, lmode, threads=1):
    """Run the reference's own compiled reorderGraphProcess (-l lmode) on a
    copy of the edge list; returns (labels, seconds)."""
    R = ref_lib()
    s = np.ascontiguousarray(src, dtype=np.uint32)
    d = np.ascontiguousarray(dst, dtype=np.uint32)
    lab = np.zeros(num_vertices, np.uint32)
    sec = C.c_double()
    if R.ogref_reorder(num_vertices, len(s), _p(s), _p(d), lmode, threads, _p(lab), C.byref(sec)):
        raise RuntimeError("ogref_reorder failed")
    return lab, sec.value


def ref_csr_build(src, dst, num_vertices, sort_mode=0, threads=1):
    """Run the reference's own compiled CSR preprocessing (sortRunAlgorithms,
    graphCSRAssignEdgeList with mapVerticesWithInOutDegree, the inverse list)
    on a copy of the edge list; returns (dict of the ten arrays, seconds)."""
    R = ref_lib()
    s = np.ascontiguousarray(src, dtype=np.uint32)
    d = np.ascontiguousarray(dst, dtype=np.uint32)
    E = len(s)
    out = {k: np.zeros(E if k in ("f_src", "f_dst", "i_src", "i_dst") else num_vertices, np.uint32)
           for k in CSR_FIELDS}
    ptrs = (u32p * 10)(*[_p(out[k]) for k in CSR_FIELDS])
    sec = C.c_double()
    if R.ogref_csr_build(num_vertices, E, _p(s), _p(d), sort_mode, threads, ptrs, C.byref(sec)):
        raise RuntimeError("ogref_csr_build failed")
    return out, sec.value


def rank_stats_topk(ref_pr, ref_order, pr, order, topk, num_vertices=None):
    """graphStats.c collectStatsPageRank_topK restated: dict of the seven fields."""
    rp, ro = np.ascontiguousarray(ref_pr, np.float32), np.ascontiguousarray(ref_order, np.uint32)
    p, o = np.ascontiguousarray(pr, np.float32), np.ascontiguousarray(order, np.uint32)
    n = len(rp)
    out = OrcCorrStats()
    if lib().orc_rank_stats_topk(_p(rp, f32p), _p(ro), _p(p, f32p), _p(o), n, topk,
                                 n if num_vertices is None else num_vertices, C.byref(out)):
        raise ValueError("bad sizes")
    return out.as_dict()


def ref_rank_stats_batch(ref_pr, ref_order, pr, order, problems):
    """The reference's own collectStatsPageRank_topK for each (topk,
    num_vertices) in problems, inverse rankings formed once; returns (list of
    dicts, seconds spent in the _topK calls)."""
    rp, ro = np.ascontiguousarray(ref_pr, np.float32), np.ascontiguousarray(ref_order, np.uint32)
    p, o = np.ascontiguousarray(pr, np.float32), np.ascontiguousarray(order, np.uint32)
    P = len(problems)
    ks = np.ascontiguousarray([k for k, _ in problems], np.uint32)
    ms = np.ascontiguousarray([m for _, m in problems], np.uint32)
    u = np.zeros(3 * P, np.uint32)
    f = np.zeros(2 * P, np.float32)
    d = np.zeros(2 * P, np.float64)
    sec = C.c_double()
    if ref_lib().ogref_rank_stats_batch(len(rp), _p(rp, f32p), _p(ro), _p(p, f32p), _p(o), P, _p(ks), _p(ms),
                                        _p(u), _p(f, f32p), _p(d, f64p), C.byref(sec)):
        raise RuntimeError("ogref_rank_stats_batch failed")
    out = [dict(levenshtein_distance=int(u[3 * i]), float_Kendall=float(f[2 * i]),
                real_Kendall=float(f[2 * i + 1]), intersection=int(u[3 * i + 1]), mismatch=int(u[3 * i + 2]),
                avg_error_float=float(d[2 * i]), avg_error_relative=float(d[2 * i + 1])) for i in range(P)]
    return out, sec.value


def ref_rank_stats_topk(ref_pr, ref_order, pr, order, topk, num_vertices=None):
    """The reference's own collectStatsPageRank_topK: dict of the seven fields."""
    n = len(ref_pr)
    return ref_rank_stats_batch(ref_pr, ref_order, pr, order, [(topk, n if num_vertices is None else num_vertices)])[0][0]


def ref_collect_stats(ref_pr, ref_order, pr, order, fnameb, bin_size=512, verbosity=0, pushpull=0,
                      trial=0, time_total=0.0, iterations=0, threads=1, stdout_path=None):
    """The reference's own collectStatsPageRank: appends to
    <fnameb>_1_0_<trial>_<pushpull>.stats; returns the call's seconds."""
    rp, ro = np.ascontiguousarray(ref_pr, np.float32), np.ascontiguousarray(ref_order, np.uint32)
    p, o = np.ascontiguousarray(pr, np.float32), np.ascontiguousarray(order, np.uint32)
    sec = C.c_double()
    ref_lib().ogref_collect_stats(len(rp), _p(rp, f32p), _p(ro), _p(p, f32p), _p(o), fnameb.encode(),
                                  bin_size, verbosity, pushpull, trial, time_total, iterations, threads,
                                  stdout_path.encode() if stdout_path else None, C.byref(sec))
    return sec.value


def ref_spmv(csr: CSR, code=0, iterations=20, threads=1):
    """Run the reference's own compiled SPMVGraphCSR (-p code) on these
    arrays; returns (vector_output, time_total seconds)."""
    R = ref_lib()
    out = np.zeros(csr.num_vertices, np.float32)
    sec = C.c_double()
    arrays = [csr.a[k] for k in ["f_out_degree", "f_in_degree", "f_edges_idx", "f_src", "f_dst",
                                  "i_out_degree", "i_in_degree", "i_edges_idx", "i_src", "i_dst"]]
    if R.ogref_spmv(csr.num_vertices, csr.num_edges, *[_p(x) for x in arrays], code, iterations,
                    threads, _p(out, f32p), C.byref(sec)):
        raise RuntimeError("ogref_spmv failed")
    return out, sec.value


def read_bin(path):
    raw = np.fromfile(path, dtype=np.uint32)
    raw = raw[: (len(raw) // 2) * 2]
    return raw[0::2].copy(), raw[1::2].copy()
