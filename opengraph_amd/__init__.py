"""Python host mirror of libopengraph_amd's C ABI (ctypes, no torch).

The library is the product: PageRank over CSR and the edge-list -> CSR
builder of OpenGraph, on MI355X (gfx950).  This module only binds its
``extern "C"`` entry points (include/opengraph_amd.h, opengraph_amd_ext.h)
so tests, the benchmark and Python callers can reach them:

* ``Arguments``, ``Vertex``, ``EdgeList``, ``GraphCSR``, ``PageRankStats`` are
  the reference's structs (graphConfig.h:11-44, vertex.h:13-18,
  edgeList.h:8-23, graphCSR.h:12-29, pageRank.h:19-30) as ctypes types, so
  ``pagerank_graph_csr`` calls the drop-in ``pageRankGraphCSR`` exactly as
  the reference's driver does (graphRun.c:1035-1040).
* ``DeviceArray``, ``Graph`` and ``PageRank`` wrap the device-resident
  ``ogpu_*`` API.

Loading fails loudly when the shared library has not been built
(``python -c "import __graft_entry__ as g; g.build()"`` or
``make -C opengraph_amd``); there is no CPU fallback.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
# OGPU_LIB names another build of the same library (the sanitized host build)
LIB_PATH = os.environ.get("OGPU_LIB") or os.path.join(_HERE, "libopengraph_amd.so")

u32 = C.c_uint32
u32p = C.POINTER(C.c_uint32)


class MT19937State(C.Structure):
    """mt19937.h:59-63."""
    _fields_ = [("mt", C.c_ulong * 624), ("mti", C.c_int)]


class Arguments(C.Structure):
    """graphConfig.h:11-44."""
    _fields_ = [
        ("wflag", C.c_int), ("xflag", C.c_int), ("sflag", C.c_int), ("Sflag", C.c_int),
        ("dflag", C.c_int), ("binSize", u32), ("verbosity", u32), ("iterations", u32),
        ("trials", u32), ("epsilon", C.c_double), ("source", C.c_int), ("algorithm", u32),
        ("datastructure", u32), ("pushpull", u32), ("sort", u32), ("lmode", u32),
        ("lmode_l2", u32), ("lmode_l3", u32), ("mmode", u32), ("symmetric", u32),
        ("weighted", u32), ("delta", u32), ("pre_numThreads", C.c_int),
        ("algo_numThreads", C.c_int), ("ker_numThreads", C.c_int), ("fnameb", C.c_char_p),
        ("fnamel", C.c_char_p), ("fnameb_format", u32), ("convert_format", u32),
        ("mt19937var", MT19937State), ("cache_size", u32),
    ]


class Vertex(C.Structure):
    """vertex.h:13-18."""
    _fields_ = [("out_degree", u32p), ("in_degree", u32p), ("edges_idx", u32p)]


class EdgeList(C.Structure):
    """edgeList.h:8-23."""
    _fields_ = [
        ("num_edges", u32), ("num_vertices", u32), ("avg_degree", u32),
        ("edges_array_src", u32p), ("edges_array_dest", u32p), ("mask_array", u32p),
        ("label_array", u32p), ("inverse_label_array", u32p),
    ]


class GraphCSR(C.Structure):
    """graphCSR.h:12-29 (WEIGHTED 0, DIRECTED 1)."""
    _fields_ = [
        ("num_edges", u32), ("num_vertices", u32), ("avg_degree", u32),
        ("vertices", C.POINTER(Vertex)), ("sorted_edges_array", C.POINTER(EdgeList)),
        ("inverse_vertices", C.POINTER(Vertex)),
        ("inverse_sorted_edges_array", C.POINTER(EdgeList)),
    ]


class PageRankStats(C.Structure):
    """pageRank.h:19-30."""
    _fields_ = [
        ("damp", C.c_float), ("base_pr", C.c_float), ("iterations", u32), ("num_vertices", u32),
        ("activeVertices", u32), ("realRanks", u32p), ("pageRanks", C.POINTER(C.c_float)),
        ("time_total", C.c_double), ("error_total", C.c_double),
    ]


class SPMVStats(C.Structure):
    """SPMV.h:19-28."""
    _fields_ = [
        ("iterations", u32), ("arguments", C.c_void_p), ("num_vertices", u32),
        ("vector_output", C.POINTER(C.c_float)), ("vector_input", C.POINTER(C.c_float)),
        ("time_total", C.c_double),
    ]


class PageRankCorrelationStats(C.Structure):
    """graphStats.h:28-37."""
    _fields_ = [("levenshtein_distance", u32), ("float_Kendall", C.c_float), ("real_Kendall", C.c_float),
                ("intersection", u32), ("mismatch", u32), ("avg_error_float", C.c_double),
                ("avg_error_relative", C.c_double)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class SpmvOptions(C.Structure):
    """ogpu_spmv_options (opengraph_amd_ext.h)."""
    _fields_ = [("code", u32), ("hub_threshold", u32), ("hot_cache", u32), ("scatter", u32)]


class MultiTile(C.Structure):
    """ogpu_multi_tile (opengraph_amd_ext.h): a tile shape of the multi-graph pull."""
    _fields_ = [("block", u32), ("chunks_per_wave", u32), ("row_batch", u32), ("pad", u32)]

    def shape(self):
        return (self.block, self.chunks_per_wave, self.row_batch)


class PrOptions(C.Structure):
    """ogpu_pr_options (opengraph_amd_ext.h)."""
    _fields_ = [
        ("variant", u32), ("max_iter", u32), ("epsilon", C.c_double), ("stop_rule", u32),
        ("init_mode", u32), ("relabel", u32), ("hub_threshold", u32), ("ppr", u32),
        ("ppr_source", u32), ("hot_cache", u32), ("exact_above", u32), ("binned", u32),
        ("warm_classes", u32),
    ]


class PrResult(C.Structure):
    """ogpu_pr_result (opengraph_amd_ext.h)."""
    _fields_ = [
        ("iterations", u32), ("rows", u32), ("active", u32), ("pad", u32),
        ("error_total", C.c_double), ("pr_sum", C.c_double), ("seconds_total", C.c_double),
        ("seconds_plan", C.c_double),
    ]


class PprResult(C.Structure):
    """ogpu_ppr_result (opengraph_amd_ext.h)."""
    _fields_ = [
        ("iterations", u32), ("converged", u32), ("seconds_total", C.c_double),
        ("seconds_plan", C.c_double), ("kernel_ms_per_iter", C.c_double),
    ]


# Every entry point of include/opengraph_amd.h and include/opengraph_amd_ext.h:
# name -> (restype, argtypes).
_VP = C.c_void_p
_PR_FN = (C.POINTER(PageRankStats), [C.POINTER(Arguments), C.POINTER(GraphCSR)])
_SPMV_FN = (C.POINTER(SPMVStats), [C.POINTER(Arguments), C.POINTER(GraphCSR)])
SIGNATURES = {
    # opengraph_amd.h
    "pageRankGraphCSR": _PR_FN, "pageRankPullGraphCSR": _PR_FN, "pageRankPushGraphCSR": _PR_FN,
    "pageRankPullFixedPoint64BitGraphCSR": _PR_FN, "pageRankPullFixedPoint32BitGraphCSR": _PR_FN,
    "pageRankPullFixedPoint16BitGraphCSR": _PR_FN, "pageRankPullFixedPoint8BitGraphCSR": _PR_FN,
    "pageRankPushFixedPointGraphCSR": _PR_FN, "pageRankPullQuant32BitGraphCSR": _PR_FN,
    "pageRankPullQuant16BitGraphCSR": _PR_FN, "pageRankPullQuant8BitGraphCSR": _PR_FN,
    "pageRankPushQuantGraphCSR": _PR_FN, "pageRankDataDrivenPullGraphCSR": _PR_FN,
    "pageRankDataDrivenPushGraphCSR": _PR_FN, "pageRankDataDrivenPullPushGraphCSR": _PR_FN,
    "newPageRankStatsGraphCSR": (C.POINTER(PageRankStats), [C.POINTER(GraphCSR)]),
    "freePageRankStats": (None, [C.POINTER(PageRankStats)]),
    "SPMVGraphCSR": _SPMV_FN, "SPMVPullGraphCSR": _SPMV_FN, "SPMVPushGraphCSR": _SPMV_FN,
    "SPMVPullFixedPointGraphCSR": _SPMV_FN, "SPMVPushFixedPointGraphCSR": _SPMV_FN,
    "newSPMVStatsGraphCSR": (C.POINTER(SPMVStats), [C.POINTER(GraphCSR)]),
    "freeSPMVStats": (None, [C.POINTER(SPMVStats)]),
    "collectStatsPageRank": (None, [C.POINTER(Arguments), C.POINTER(PageRankStats), C.POINTER(PageRankStats), u32]),
    "collectStatsPageRank_topK": (PageRankCorrelationStats, [C.POINTER(PageRankStats), C.POINTER(PageRankStats),
                                                             u32p, u32p, u32p, u32, u32, C.c_void_p, u32]),
    "graphCSRPreProcessingStep": (C.POINTER(GraphCSR), [C.POINTER(Arguments)]),
    "graphCSRNew": (C.POINTER(GraphCSR), [u32, u32, C.c_uint8]),
    "graphCSRAssignEdgeList": (C.POINTER(GraphCSR), [C.POINTER(GraphCSR), C.POINTER(EdgeList), C.c_uint8]),
    "graphCSRFree": (None, [C.POINTER(GraphCSR)]),
    "graphCSRPrint": (None, [C.POINTER(GraphCSR)]),
    "mapVerticesWithInOutDegree": (C.POINTER(GraphCSR), [C.POINTER(GraphCSR), C.c_uint8]),
    "newVertexArray": (C.POINTER(Vertex), [u32]),
    "freeVertexArray": (None, [C.POINTER(Vertex)]),
    "readEdgeListsbin": (C.POINTER(EdgeList), [C.c_char_p, C.c_uint8, u32, u32]),
    "readEdgeListsMem": (C.POINTER(EdgeList), [C.POINTER(EdgeList), C.c_uint8, u32, u32]),
    "newEdgeList": (C.POINTER(EdgeList), [u32]),
    "freeEdgeList": (None, [C.POINTER(EdgeList)]),
    "sortRunAlgorithms": (C.POINTER(EdgeList), [C.POINTER(EdgeList), u32]),
    "radixSortEdgesBySource": (C.POINTER(EdgeList), [C.POINTER(EdgeList)]),
    "reorderGraphGenerateInOutDegrees": (u32p, [u32p, C.POINTER(EdgeList), u32]),
    "radixSortEdgesByDegree": (u32p, [u32p, u32p, u32]),
    "relabelEdgeList": (C.POINTER(EdgeList), [C.POINTER(EdgeList), u32p]),
    "reorderGraphListDegree": (C.POINTER(EdgeList), [C.POINTER(EdgeList), u32p, u32]),
    "reorderGraphProcessDegree": (C.POINTER(EdgeList), [u32, C.POINTER(EdgeList), u32]),
    "reorderGraphListDBG": (C.POINTER(EdgeList), [C.POINTER(EdgeList), u32p, u32p, u32, u32]),
    "reorderGraphProcessDBG": (C.POINTER(EdgeList), [u32, C.POINTER(EdgeList), u32]),
    "reorderGraphProcess": (C.POINTER(EdgeList), [C.POINTER(EdgeList), C.POINTER(Arguments)]),
    "radixSortEdgesBySourceAndDestination": (C.POINTER(EdgeList), [C.POINTER(EdgeList)]),
    "countSortEdgesBySource": (C.POINTER(EdgeList), [C.POINTER(EdgeList)]),
    "countSortEdgesBySourceAndDestination": (C.POINTER(EdgeList), [C.POINTER(EdgeList)]),
    "radixSortEdgesByPageRank": (u32p, [C.POINTER(C.c_float), u32p, u32]),
    # opengraph_amd_ext.h
    "ogpu_init": (C.c_int, [C.c_int]),
    "ogpu_last_error": (C.c_char_p, []),
    "ogpu_device_info": (C.c_int, [C.c_char_p, C.c_size_t, C.POINTER(C.c_int), C.POINTER(C.c_size_t)]),
    "ogpu_malloc": (C.c_int, [C.POINTER(_VP), C.c_size_t]),
    "ogpu_free": (C.c_int, [_VP]),
    "ogpu_memcpy_h2d": (C.c_int, [_VP, _VP, C.c_size_t]),
    "ogpu_memcpy_d2h": (C.c_int, [_VP, _VP, C.c_size_t]),
    "ogpu_synchronize": (C.c_int, []),
    "ogpu_device_copy": (C.c_int, [_VP, _VP, C.c_size_t]),
    "ogpu_rmat_device": (C.c_int, [u32, C.c_uint64, C.c_double, C.c_double, C.c_double, C.c_uint64, C.c_int, _VP, _VP]),
    "ogpu_rmat_host": (C.c_int, [u32, C.c_uint64, C.c_double, C.c_double, C.c_double, C.c_uint64, C.c_int, _VP, _VP]),
    "ogpu_er_device": (C.c_int, [u32, C.c_uint64, C.c_uint64, _VP, _VP]),
    "ogpu_er_host": (C.c_int, [u32, C.c_uint64, C.c_uint64, _VP, _VP]),
    "ogpu_write_bin": (C.c_int, [C.c_char_p, _VP, _VP, C.c_uint64]),
    "ogpu_graph_build": (C.c_int, [_VP, _VP, u32, u32, C.POINTER(_VP)]),
    "ogpu_graph_build_reordered": (C.c_int, [_VP, _VP, u32, u32, u32p, u32, _VP, C.POINTER(_VP)]),
    "ogpu_graph_write_csr": (C.c_int, [_VP, C.c_char_p]),
    "ogpu_graph_read_csr": (C.c_int, [C.c_char_p, C.POINTER(_VP)]),
    "writeToBinFileGraphCSR": (None, [C.c_char_p, C.POINTER(GraphCSR)]),
    "readFromBinFileGraphCSR": (C.POINTER(GraphCSR), [C.c_char_p]),
    "ogpu_graph_from_arrays": (C.c_int, [u32, u32, C.POINTER(_VP), C.POINTER(_VP)]),
    "ogpu_graph_info": (C.c_int, [_VP, u32p, u32p]),
    "ogpu_graph_array": (C.c_int, [_VP, u32, _VP]),
    "ogpu_graph_free": (None, [_VP]),
    "ogpu_pr_default_options": (None, [C.POINTER(PrOptions)]),
    "ogpu_pr_create": (C.c_int, [_VP, C.POINTER(PrOptions), C.POINTER(_VP)]),
    "ogpu_pr_set_init": (C.c_int, [_VP, _VP]),
    "ogpu_pr_init_fallback": (C.c_int, [_VP]),
    "ogpu_pr_get_init": (C.c_int, [_VP, _VP]),
    "ogpu_pr_run": (C.c_int, [_VP, C.POINTER(PrResult)]),
    "ogpu_pr_reset": (C.c_int, [_VP]),
    "ogpu_pr_enqueue": (C.c_int, [_VP, u32, C.c_int]),
    "ogpu_pr_kernel_ms": (C.c_int, [_VP, C.POINTER(C.c_double), u32p]),
    "ogpu_pr_ranks": (C.c_int, [_VP, _VP]),
    "ogpu_pr_rows": (C.c_int, [_VP, u32p, C.POINTER(C.c_double), u32]),
    "ogpu_pr_bytes_per_iteration": (C.c_double, [_VP]),
    "ogpu_pr_free": (None, [_VP]),
    "ogpu_mlp_create": (C.c_int, [C.c_uint64, C.POINTER(_VP)]),
    "ogpu_mlp_params": (C.c_int, [_VP, _VP, _VP]),
    "ogpu_mlp_add_graph": (C.c_int, [_VP, _VP, C.c_double, u32, C.POINTER(C.c_double)]),
    "ogpu_mlp_data_size": (u32, [_VP]),
    "ogpu_mlp_train": (C.c_int, [_VP, u32, u32, C.c_float, C.c_uint64, C.POINTER(C.c_double),
                                 C.POINTER(C.c_double)]),
    "ogpu_mlp_grad": (C.c_int, [_VP, _VP, _VP, u32, _VP, C.POINTER(C.c_double)]),
    "ogpu_mlp_grad_noalloc": (C.c_int, [_VP, _VP, _VP, u32, _VP, _VP]),
    "ogpu_mlp_forward": (C.c_int, [_VP, _VP, u32, C.c_int, _VP]),
    "ogpu_mlp_free": (None, [_VP]),
    "ogpu_pr_features": (C.c_int, [_VP, _VP]),
    "ogpu_pr_set_init_learned": (C.c_int, [_VP, _VP, C.POINTER(C.c_double)]),
    "ogpu_tune_hub_threshold": (C.c_int, [_VP, u32p, u32, u32, u32p, C.POINTER(C.c_double)]),
    "ogpu_multi_create": (C.c_int, [C.POINTER(_VP), u32, C.POINTER(PrOptions), u32p, C.POINTER(_VP)]),
    "ogpu_multi_create_tiled": (C.c_int, [C.POINTER(_VP), u32, C.POINTER(PrOptions), u32p,
                                          C.POINTER(MultiTile), C.POINTER(_VP)]),
    "ogpu_multi_tile_shapes": (u32, [C.POINTER(MultiTile), u32]),
    "ogpu_multi_tile_of": (C.c_int, [_VP, u32, C.POINTER(MultiTile)]),
    "ogpu_multi_run": (C.c_int, [_VP, C.c_int, C.POINTER(C.c_double)]),
    "ogpu_multi_result": (C.c_int, [_VP, u32, C.POINTER(PrResult), _VP]),
    "ogpu_multi_hub_threshold": (u32, [_VP, u32]),
    "ogpu_multi_free": (None, [_VP]),
    "ogpu_ppr_create": (C.c_int, [_VP, u32p, u32, u32, C.c_double, C.POINTER(_VP)]),
    "ogpu_ppr_run": (C.c_int, [_VP, C.POINTER(PprResult)]),
    "ogpu_ppr_ranks": (C.c_int, [_VP, _VP]),
    "ogpu_ppr_columns": (C.c_int, [_VP, u32p, C.POINTER(C.c_double), u32]),
    "ogpu_ppr_bytes_per_iteration": (C.c_double, [_VP]),
    "ogpu_ppr_free": (None, [_VP]),
    "ogpu_spmv_default_options": (None, [C.POINTER(SpmvOptions)]),
    "ogpu_spmv_create": (C.c_int, [_VP, C.POINTER(SpmvOptions), C.POINTER(_VP)]),
    "ogpu_spmv_run": (C.c_int, [_VP, u32, C.POINTER(C.c_double)]),
    "ogpu_spmv_output": (C.c_int, [_VP, _VP]),
    "ogpu_spmv_iteration_seconds": (C.c_int, [_VP, u32, C.POINTER(C.c_double)]),
    "ogpu_spmv_reset": (C.c_int, [_VP]),
    "ogpu_spmv_enqueue": (C.c_int, [_VP, u32, C.c_int]),
    "ogpu_spmv_kernel_ms": (C.c_int, [_VP, C.POINTER(C.c_double), u32p]),
    "ogpu_spmv_bytes_per_iteration": (C.c_double, [_VP]),
    "ogpu_spmv_free": (None, [_VP]),
    "ogpu_rank_stats": (C.c_int, [_VP, _VP, _VP, _VP, u32, u32p, u32p, u32, C.POINTER(PageRankCorrelationStats)]),
    "ogpu_rank_stats_device": (C.c_int, [_VP, _VP, _VP, _VP, u32, u32p, u32p, u32,
                                         C.POINTER(PageRankCorrelationStats)]),
}

GRAPH_ARRAYS = ["f_src", "f_dst", "f_out_degree", "f_in_degree", "f_edges_idx",
                "i_src", "i_dst", "i_out_degree", "i_in_degree", "i_edges_idx"]

_lib = None


class OgpuError(RuntimeError):
    pass


def load():
    """Load libopengraph_amd.so and declare every entry point (raises if absent)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise OgpuError(f"{LIB_PATH} is not built; run __graft_entry__.build() or make -C opengraph_amd")
    lib = C.CDLL(LIB_PATH)
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def check(rc):
    if rc != 0:
        raise OgpuError(load().ogpu_last_error().decode())


def init(device: int = 0):
    check(load().ogpu_init(device))


def device_info():
    lib = load()
    name = C.create_string_buffer(256)
    cus = C.c_int()
    mem = C.c_size_t()
    check(lib.ogpu_device_info(name, 256, C.byref(cus), C.byref(mem)))
    return name.value.decode(), cus.value, mem.value


def _ptr(a: np.ndarray):
    return a.ctypes.data_as(C.c_void_p)


class DeviceArray:
    """A device allocation owned by Python (ogpu_malloc / ogpu_free)."""

    def __init__(self, nbytes: int):
        self.nbytes = int(nbytes)
        self.ptr = C.c_void_p()
        check(load().ogpu_malloc(C.byref(self.ptr), max(self.nbytes, 1)))

    @classmethod
    def from_numpy(cls, a: np.ndarray):
        a = np.ascontiguousarray(a)
        d = cls(a.nbytes)
        if a.nbytes:
            check(load().ogpu_memcpy_h2d(d.ptr, _ptr(a), a.nbytes))
        return d

    def to_numpy(self, dtype, count):
        out = np.empty(count, dtype=dtype)
        if out.nbytes:
            check(load().ogpu_memcpy_d2h(_ptr(out), self.ptr, out.nbytes))
        return out

    def free(self):
        if self.ptr:
            load().ogpu_free(self.ptr)
            self.ptr = C.c_void_p()

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def rmat_host(scale, num_edges, a=0.57, b=0.19, c=0.19, seed=1, permute=True):
    src = np.empty(num_edges, np.uint32)
    dst = np.empty(num_edges, np.uint32)
    check(load().ogpu_rmat_host(scale, num_edges, a, b, c, seed, int(permute), _ptr(src), _ptr(dst)))
    return src, dst


def er_host(num_vertices, num_edges, seed):
    src = np.empty(num_edges, np.uint32)
    dst = np.empty(num_edges, np.uint32)
    check(load().ogpu_er_host(num_vertices, num_edges, seed, _ptr(src), _ptr(dst)))
    return src, dst


def rmat_device(scale, num_edges, a=0.57, b=0.19, c=0.19, seed=1, permute=True):
    src = DeviceArray(4 * num_edges)
    dst = DeviceArray(4 * num_edges)
    check(load().ogpu_rmat_device(scale, num_edges, a, b, c, seed, int(permute), src.ptr, dst.ptr))
    return src, dst


def er_device(num_vertices, num_edges, seed):
    src = DeviceArray(4 * num_edges)
    dst = DeviceArray(4 * num_edges)
    check(load().ogpu_er_device(num_vertices, num_edges, seed, src.ptr, dst.ptr))
    return src, dst


def write_bin(path, src, dst):
    check(load().ogpu_write_bin(path.encode(), _ptr(src), _ptr(dst), len(src)))


class Graph:
    """Device forward + inverse CSR (ogpu_graph)."""

    def __init__(self, handle):
        self.h = handle
        n, e = u32(), u32()
        check(load().ogpu_graph_info(self.h, C.byref(n), C.byref(e)))
        self.num_vertices, self.num_edges = n.value, e.value

    @classmethod
    def build(cls, d_src: DeviceArray, d_dst: DeviceArray, num_edges: int, sort_mode: int = 0):
        h = C.c_void_p()
        check(load().ogpu_graph_build(d_src.ptr, d_dst.ptr, num_edges, sort_mode, C.byref(h)))
        return cls(h)

    @classmethod
    def build_reordered(cls, d_src: DeviceArray, d_dst: DeviceArray, num_edges: int, num_vertices: int,
                        sort_mode: int = 0, lmodes=(1,)):
        """CSR after the reference's -l relabellings (ogpu_graph_build_reordered)
        of device edge arrays whose largest id is num_vertices - 1; returns
        (graph, labels of the input vertices)."""
        dl = DeviceArray(4 * max(num_vertices, 1))
        lm = np.ascontiguousarray(lmodes, dtype=np.uint32)
        h = C.c_void_p()
        try:
            check(load().ogpu_graph_build_reordered(d_src.ptr, d_dst.ptr, num_edges, sort_mode,
                                                    lm.ctypes.data_as(u32p), len(lm), dl.ptr, C.byref(h)))
            labels = dl.to_numpy(np.uint32, num_vertices)
        finally:
            dl.free()
        return cls(h), labels

    @classmethod
    def from_edges_reordered(cls, src: np.ndarray, dst: np.ndarray, sort_mode: int = 0, lmodes=(1,)):
        """The same from host edge arrays."""
        ds, dd = DeviceArray.from_numpy(src.astype(np.uint32)), DeviceArray.from_numpy(dst.astype(np.uint32))
        n = int(max(src.max(), dst.max())) + 1 if len(src) else 1
        try:
            return cls.build_reordered(ds, dd, len(src), n, sort_mode, lmodes)
        finally:
            ds.free(); dd.free()

    @classmethod
    def read_csr(cls, path: str):
        """A device graph from a reference .csr file (ogpu_graph_read_csr)."""
        h = C.c_void_p()
        check(load().ogpu_graph_read_csr(path.encode(), C.byref(h)))
        return cls(h)

    def write_csr(self, fname: str):
        """Write fname + '.csr' in the reference's format (ogpu_graph_write_csr)."""
        check(load().ogpu_graph_write_csr(self.h, fname.encode()))

    @classmethod
    def from_edges(cls, src: np.ndarray, dst: np.ndarray, sort_mode: int = 0):
        ds, dd = DeviceArray.from_numpy(src.astype(np.uint32)), DeviceArray.from_numpy(dst.astype(np.uint32))
        g = cls.build(ds, dd, len(src), sort_mode)
        ds.free()
        dd.free()
        return g

    @classmethod
    def from_arrays(cls, arrays: dict):
        """arrays: the 10 CSR arrays named as GRAPH_ARRAYS (numpy uint32)."""
        keep = [np.ascontiguousarray(arrays[k], dtype=np.uint32) for k in GRAPH_ARRAYS]
        ptrs = (C.c_void_p * 10)(*[k.ctypes.data for k in keep])
        h = C.c_void_p()
        n = len(keep[2])
        e = len(keep[0])
        check(load().ogpu_graph_from_arrays(n, e, ptrs, C.byref(h)))
        return cls(h)

    def array(self, name: str) -> np.ndarray:
        i = GRAPH_ARRAYS.index(name)
        n = self.num_edges if name in ("f_src", "f_dst", "i_src", "i_dst") else self.num_vertices
        out = np.empty(n, np.uint32)
        check(load().ogpu_graph_array(self.h, i, _ptr(out)))
        return out

    def arrays(self) -> dict:
        return {k: self.array(k) for k in GRAPH_ARRAYS}

    def free(self):
        if self.h:
            load().ogpu_graph_free(self.h)
            self.h = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class PageRank:
    """A prepared PageRank run on a device graph (ogpu_pr)."""

    def __init__(self, graph: Graph, variant=0, max_iter=20, epsilon=1e-4, stop_rule=0,
                 init_mode=0, relabel=1, hub_threshold=0, init=None, ppr_source=None, hot_cache=0,
                 exact_above=None, binned=None, warm_classes=None):
        lib = load()
        o = PrOptions()
        lib.ogpu_pr_default_options(C.byref(o))
        if exact_above is not None:
            o.exact_above = exact_above
        if binned is not None:
            o.binned = binned
        if warm_classes is not None:
            o.warm_classes = warm_classes
        o.variant, o.max_iter, o.epsilon = variant, max_iter, epsilon
        o.stop_rule, o.init_mode, o.relabel, o.hub_threshold = stop_rule, init_mode, relabel, hub_threshold
        if ppr_source is not None:
            o.ppr, o.ppr_source = 1, ppr_source
        o.hot_cache = hot_cache
        self.graph = graph
        self.h = C.c_void_p()
        check(lib.ogpu_pr_create(graph.h, C.byref(o), C.byref(self.h)))
        if init is not None:
            self.set_init(init)

    def set_init(self, init: np.ndarray):
        """Start vector of the following runs (unnormalized, caller's ids); a
        NaN, infinite or negative entry or no mass falls back to uniform."""
        a = np.ascontiguousarray(init, dtype=np.float32)
        check(load().ogpu_pr_set_init(self.h, _ptr(a)))

    def init_fallback(self) -> bool:
        """True when the last start offered was rejected for the uniform one."""
        return load().ogpu_pr_init_fallback(self.h) == 1

    def get_init(self) -> np.ndarray:
        """The start vector of the next runs (old ids, unnormalized)."""
        out = np.empty(self.graph.num_vertices, np.float32)
        check(load().ogpu_pr_get_init(self.h, _ptr(out)))
        return out

    def run(self) -> PrResult:
        r = PrResult()
        check(load().ogpu_pr_run(self.h, C.byref(r)))
        return r

    def ranks(self) -> np.ndarray:
        out = np.empty(self.graph.num_vertices, np.float32)
        check(load().ogpu_pr_ranks(self.h, _ptr(out)))
        return out

    def rows(self, max_rows=100000):
        act = np.zeros(max_rows, np.uint32)
        err = np.zeros(max_rows, np.float64)
        n = load().ogpu_pr_rows(self.h, act.ctypes.data_as(u32p),
                                err.ctypes.data_as(C.POINTER(C.c_double)), max_rows)
        return act[:n], err[:n]

    def reset(self):
        check(load().ogpu_pr_reset(self.h))

    def enqueue(self, n, time_kernels=False):
        check(load().ogpu_pr_enqueue(self.h, n, int(time_kernels)))

    def kernel_ms(self):
        t, c = C.c_double(), u32()
        check(load().ogpu_pr_kernel_ms(self.h, C.byref(t), C.byref(c)))
        return t.value, c.value

    def bytes_per_iteration(self):
        return load().ogpu_pr_bytes_per_iteration(self.h)

    def features(self) -> np.ndarray:
        """N x 4 learned-start features, caller's vertex order (variant 0)."""
        out = np.empty((self.graph.num_vertices, 4), np.float32)
        check(load().ogpu_pr_features(self.h, _ptr(out)))
        return out

    def set_init_learned(self, mlp: "Mlp") -> float:
        """Learned start for the following runs; returns the host seconds it took."""
        s = C.c_double()
        check(load().ogpu_pr_set_init_learned(self.h, mlp.h, C.byref(s)))
        return s.value

    def free(self):
        if self.h:
            load().ogpu_pr_free(self.h)
            self.h = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


MLP_PARAMS = 4545


class Mlp:
    """The learned warm start's network (ogpu_mlp)."""

    def __init__(self, seed=1):
        self.h = C.c_void_p()
        check(load().ogpu_mlp_create(seed, C.byref(self.h)))

    def params(self) -> np.ndarray:
        out = np.empty(MLP_PARAMS, np.float32)
        check(load().ogpu_mlp_params(self.h, _ptr(out), None))
        return out

    def set_params(self, p):
        a = np.ascontiguousarray(p, dtype=np.float32)
        assert a.size == MLP_PARAMS
        check(load().ogpu_mlp_params(self.h, None, _ptr(a)))

    def add_graph(self, graph: Graph, tol=1e-9, max_iter=500) -> float:
        l1 = C.c_double()
        check(load().ogpu_mlp_add_graph(self.h, graph.h, tol, max_iter, C.byref(l1)))
        return l1.value

    def data_size(self) -> int:
        return load().ogpu_mlp_data_size(self.h)

    def train(self, epochs=10, batch=65536, lr=3e-3, seed=1):
        """(last epoch's mean loss, seconds)."""
        loss, sec = C.c_double(), C.c_double()
        check(load().ogpu_mlp_train(self.h, epochs, batch, lr, seed, C.byref(loss), C.byref(sec)))
        return loss.value, sec.value

    def grad(self, feat, y):
        f = np.ascontiguousarray(feat, dtype=np.float32)
        yy = np.ascontiguousarray(y, dtype=np.float32)
        g = np.empty(MLP_PARAMS, np.float32)
        loss = C.c_double()
        check(load().ogpu_mlp_grad(self.h, _ptr(f), _ptr(yy), len(yy), _ptr(g), C.byref(loss)))
        return loss.value, g

    def forward(self, feat, mfma=False):
        f = np.ascontiguousarray(feat, dtype=np.float32)
        out = np.empty(len(f), np.float32)
        check(load().ogpu_mlp_forward(self.h, _ptr(f), len(f), int(mfma), _ptr(out)))
        return out

    def free(self):
        if self.h:
            load().ogpu_mlp_free(self.h)
            self.h = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def tune_hub_threshold(graph: Graph, cands=(8, 16, 32, 64, 128), iters=5):
    """(best threshold, ms per iteration of each candidate) (ogpu_tune_hub_threshold)."""
    c = np.ascontiguousarray(cands, dtype=np.uint32)
    ms = np.zeros(len(c), np.float64)
    best = u32()
    check(load().ogpu_tune_hub_threshold(graph.h, c.ctypes.data_as(u32p), len(c), iters, C.byref(best),
                                         ms.ctypes.data_as(C.POINTER(C.c_double))))
    return best.value, ms


class Multi:
    """Plain PageRank on many graphs at once (ogpu_multi): mode 0 one graph
    after another, 1 all graphs per launch, 2 the same in one hipGraph."""

    def __init__(self, graphs, max_iter=100, epsilon=1e-6, hub_thresholds=None, hub_threshold=0,
                 tiles=None):
        """tiles: None (every graph's tile shape autotuned), or one
        (block, chunks_per_wave, row_batch) per graph, (0, 0, 0) to tune it."""
        lib = load()
        o = PrOptions()
        lib.ogpu_pr_default_options(C.byref(o))
        o.max_iter, o.epsilon, o.hub_threshold = max_iter, epsilon, hub_threshold
        self.graphs = list(graphs)
        hs = (_VP * len(self.graphs))(*[g.h.value if isinstance(g.h, C.c_void_p) else g.h for g in self.graphs])
        ht = None
        if hub_thresholds is not None:
            self._ht = np.ascontiguousarray(hub_thresholds, dtype=np.uint32)
            ht = self._ht.ctypes.data_as(u32p)
        tl = None
        if tiles is not None:
            tl = (MultiTile * len(self.graphs))(*[MultiTile(b, k, r, 0) for (b, k, r) in tiles])
        self.h = C.c_void_p()
        check(lib.ogpu_multi_create_tiled(hs, len(self.graphs), C.byref(o), ht, tl, C.byref(self.h)))

    def run(self, mode=2) -> float:
        s = C.c_double()
        check(load().ogpu_multi_run(self.h, mode, C.byref(s)))
        return s.value

    def result(self, g, ranks=True):
        r = PrResult()
        out = np.empty(self.graphs[g].num_vertices, np.float32) if ranks else None
        check(load().ogpu_multi_result(self.h, g, C.byref(r), _ptr(out) if ranks else None))
        return r, out

    def hub_threshold(self, g):
        return load().ogpu_multi_hub_threshold(self.h, g)

    def tile(self, g):
        """Graph g's tile shape (block, chunks_per_wave, row_batch)."""
        t = MultiTile()
        check(load().ogpu_multi_tile_of(self.h, g, C.byref(t)))
        return t.shape()

    @staticmethod
    def tile_shapes():
        """The compiled tile shapes of the multi-graph pull."""
        lib = load()
        n = lib.ogpu_multi_tile_shapes(None, 0)
        arr = (MultiTile * n)()
        lib.ogpu_multi_tile_shapes(arr, n)
        return [t.shape() for t in arr]

    def free(self):
        if self.h:
            load().ogpu_multi_free(self.h)
            self.h = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class Ppr:
    """Batched personalized PageRank: k <= 64 sources in one N x k block (ogpu_ppr)."""

    def __init__(self, graph: Graph, sources, max_iter=100, epsilon=1e-6):
        src = np.ascontiguousarray(sources, dtype=np.uint32)
        self.graph, self.k = graph, len(src)
        self.h = C.c_void_p()
        check(load().ogpu_ppr_create(graph.h, src.ctypes.data_as(u32p), self.k, max_iter, epsilon,
                                     C.byref(self.h)))

    def run(self) -> PprResult:
        r = PprResult()
        check(load().ogpu_ppr_run(self.h, C.byref(r)))
        return r

    def ranks(self) -> np.ndarray:
        """N x k, caller's vertex order."""
        out = np.empty((self.graph.num_vertices, self.k), np.float32)
        check(load().ogpu_ppr_ranks(self.h, _ptr(out)))
        return out

    def columns(self, max_rows=100000):
        """(iterations per column, L1 history rows x k)."""
        it = np.zeros(self.k, np.uint32)
        hist = np.zeros((max_rows, self.k), np.float64)
        n = load().ogpu_ppr_columns(self.h, it.ctypes.data_as(u32p),
                                    hist.ctypes.data_as(C.POINTER(C.c_double)), max_rows)
        if n < 0:
            check(n)
        return it, hist[:n]

    def bytes_per_iteration(self):
        return load().ogpu_ppr_bytes_per_iteration(self.h)

    def free(self):
        if self.h:
            load().ogpu_ppr_free(self.h)
            self.h = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class Spmv:
    """The reference's SpMV benchmark (-p code 0..3) on a device graph (ogpu_spmv)."""

    def __init__(self, graph: Graph, code=0, hub_threshold=0, hot_cache=0, scatter=0):
        o = SpmvOptions()
        load().ogpu_spmv_default_options(C.byref(o))
        o.code, o.hub_threshold, o.hot_cache, o.scatter = code, hub_threshold, hot_cache, scatter
        self.graph = graph
        self.h = C.c_void_p()
        check(load().ogpu_spmv_create(graph.h, C.byref(o), C.byref(self.h)))

    def run(self, iterations=20) -> float:
        """Device seconds of the whole run."""
        s = C.c_double()
        check(load().ogpu_spmv_run(self.h, iterations, C.byref(s)))
        return s.value

    def output(self) -> np.ndarray:
        out = np.empty(self.graph.num_vertices, np.float32)
        check(load().ogpu_spmv_output(self.h, _ptr(out)))
        return out

    def reset(self):
        check(load().ogpu_spmv_reset(self.h))

    def enqueue(self, n, time_kernels=False):
        check(load().ogpu_spmv_enqueue(self.h, n, int(time_kernels)))

    def kernel_ms(self):
        t, c = C.c_double(), u32()
        check(load().ogpu_spmv_kernel_ms(self.h, C.byref(t), C.byref(c)))
        return t.value, c.value

    def bytes_per_iteration(self):
        return load().ogpu_spmv_bytes_per_iteration(self.h)

    def free(self):
        if self.h:
            load().ogpu_spmv_free(self.h)
            self.h = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def synchronize():
    check(load().ogpu_synchronize())


def device_copy(dst: DeviceArray, src: DeviceArray, nbytes: int):
    check(load().ogpu_device_copy(dst.ptr, src.ptr, nbytes))


# --------------------------------------------------------------------------
# Drop-in calls through the reference's structs.
# --------------------------------------------------------------------------

class HostGraphCSR:
    """A reference-layout GraphCSR built from numpy arrays (kept alive here)."""

    def __init__(self, arrays: dict):
        self._keep = {k: np.ascontiguousarray(arrays[k], dtype=np.uint32) for k in GRAPH_ARRAYS}
        k = self._keep
        n = len(k["f_out_degree"])
        e = len(k["f_src"])

        def p(a):
            return a.ctypes.data_as(u32p)

        self.fv = Vertex(p(k["f_out_degree"]), p(k["f_in_degree"]), p(k["f_edges_idx"]))
        self.iv = Vertex(p(k["i_out_degree"]), p(k["i_in_degree"]), p(k["i_edges_idx"]))
        self.fe = EdgeList(e, n, e // max(n, 1), p(k["f_src"]), p(k["f_dst"]), None, None, None)
        self.ie = EdgeList(e, n, e // max(n, 1), p(k["i_src"]), p(k["i_dst"]), None, None, None)
        self.g = GraphCSR(e, n, e // max(n, 1), C.pointer(self.fv), C.pointer(self.fe),
                          C.pointer(self.iv), C.pointer(self.ie))


def default_arguments(**kw) -> Arguments:
    """open-graph.c:270-305 defaults, overridden by keyword."""
    a = Arguments()
    a.binSize, a.iterations, a.trials, a.epsilon = 512, 20, 1, 0.0001
    a.algorithm, a.cache_size, a.delta = 1, 32768, 1
    a.pre_numThreads = a.algo_numThreads = a.ker_numThreads = os.cpu_count() or 1
    a.fnameb_format = a.convert_format = 1
    for k, v in kw.items():
        setattr(a, k, v)
    return a


def pagerank_graph_csr(graph: HostGraphCSR, pushpull=0, iterations=20, epsilon=1e-4):
    """Call the drop-in pageRankGraphCSR; returns (ranks, iterations, error_total, time)."""
    lib = load()
    args = default_arguments(pushpull=pushpull, iterations=iterations, epsilon=epsilon)
    st = lib.pageRankGraphCSR(C.byref(args), C.byref(graph.g))
    if not st:
        raise OgpuError("pageRankGraphCSR returned NULL")
    s = st.contents
    n = s.num_vertices
    ranks = np.ctypeslib.as_array(s.pageRanks, shape=(n,)).copy() if n else np.zeros(0, np.float32)
    out = (ranks, s.iterations, s.error_total, s.time_total)
    lib.freePageRankStats(st)
    return out


def pagerank_graph_csr_stats(graph: HostGraphCSR, pushpull=0, iterations=20, epsilon=1e-4):
    """Call the drop-in pageRankGraphCSR; returns (ranks, fields) with every
    PageRankStats field (pageRank.h:19-30) and whether realRanks is the
    identity."""
    lib = load()
    args = default_arguments(pushpull=pushpull, iterations=iterations, epsilon=epsilon)
    st = lib.pageRankGraphCSR(C.byref(args), C.byref(graph.g))
    if not st:
        raise OgpuError("pageRankGraphCSR returned NULL")
    s = st.contents
    n = s.num_vertices
    ranks = np.ctypeslib.as_array(s.pageRanks, shape=(n,)).copy() if n else np.zeros(0, np.float32)
    real = np.ctypeslib.as_array(s.realRanks, shape=(n,)) if n else np.zeros(0, np.uint32)
    fields = dict(damp=s.damp, base_pr=s.base_pr, iterations=s.iterations, num_vertices=n,
                  activeVertices=s.activeVertices, error_total=s.error_total, time_total=s.time_total,
                  realRanks_identity=bool(np.array_equal(real, np.arange(n, dtype=np.uint32))))
    lib.freePageRankStats(st)
    return ranks, fields


def spmv_graph_csr(graph: HostGraphCSR, pushpull=0, iterations=20):
    """Call the drop-in SPMVGraphCSR; returns (vector_output, vector_input, iterations, time)."""
    lib = load()
    args = default_arguments(pushpull=pushpull, iterations=iterations)
    st = lib.SPMVGraphCSR(C.byref(args), C.byref(graph.g))
    if not st:
        raise OgpuError("SPMVGraphCSR returned NULL")
    s = st.contents
    n = s.num_vertices
    y = np.ctypeslib.as_array(s.vector_output, shape=(n,)).copy() if n else np.zeros(0, np.float32)
    x = np.ctypeslib.as_array(s.vector_input, shape=(n,)).copy() if n else np.zeros(0, np.float32)
    out = (y, x, s.iterations, s.time_total)
    lib.freeSPMVStats(st)
    return out


def rank_stats(ref_pr, ref_order, pr, order, problems):
    """ogpu_rank_stats: problems = [(topk, num_vertices), ...]; list of dicts
    with the seven PageRankCorrelationStats fields."""
    rp, ro = np.ascontiguousarray(ref_pr, np.float32), np.ascontiguousarray(ref_order, np.uint32)
    p, o = np.ascontiguousarray(pr, np.float32), np.ascontiguousarray(order, np.uint32)
    ks = np.ascontiguousarray([k for k, _ in problems], np.uint32)
    ms = np.ascontiguousarray([m for _, m in problems], np.uint32)
    out = (PageRankCorrelationStats * len(problems))()
    check(load().ogpu_rank_stats(_ptr(rp), _ptr(ro), _ptr(p), _ptr(o), len(rp), ks.ctypes.data_as(u32p),
                                 ms.ctypes.data_as(u32p), len(problems), out))
    return [c.as_dict() for c in out]


def rank_stats_device(d_ref_pr, d_ref_order, d_pr, d_order, n, problems):
    """ogpu_rank_stats_device on DeviceArrays of n entries."""
    ks = np.ascontiguousarray([k for k, _ in problems], np.uint32)
    ms = np.ascontiguousarray([m for _, m in problems], np.uint32)
    out = (PageRankCorrelationStats * len(problems))()
    check(load().ogpu_rank_stats_device(d_ref_pr.ptr, d_ref_order.ptr, d_pr.ptr, d_order.ptr, n,
                                        ks.ctypes.data_as(u32p), ms.ctypes.data_as(u32p), len(problems), out))
    return [c.as_dict() for c in out]


def _pr_stats(pr, order):
    pr = np.ascontiguousarray(pr, np.float32)
    order = np.ascontiguousarray(order, np.uint32)
    st = PageRankStats()
    st.damp, st.base_pr, st.num_vertices = 0.85, 0.15, len(pr)
    st.pageRanks = pr.ctypes.data_as(C.POINTER(C.c_float))
    st.realRanks = order.ctypes.data_as(u32p)
    return st, (pr, order)


def collect_stats_pagerank(ref_pr, ref_order, pr, order, fnameb, bin_size=512, verbosity=0, pushpull=0,
                           trial=0, time_total=0.0, iterations=0, threads=1):
    """The drop-in collectStatsPageRank: prints its tables and appends to
    <fnameb>_1_0_<trial>_<pushpull>.stats."""
    rs, keep1 = _pr_stats(ref_pr, ref_order)
    st, keep2 = _pr_stats(pr, order)
    st.time_total, st.iterations = time_total, iterations
    fb = fnameb.encode()
    args = default_arguments(binSize=bin_size, verbosity=verbosity, pushpull=pushpull, algorithm=1,
                             datastructure=0, fnameb=fb)
    args.pre_numThreads = args.algo_numThreads = args.ker_numThreads = threads
    load().collectStatsPageRank(C.byref(args), C.byref(rs), C.byref(st), trial)
    del keep1, keep2, fb


def collect_stats_pagerank_topk(ref_pr, ref_order, pr, order, topk, num_vertices=None):
    """The drop-in collectStatsPageRank_topK (verbose 0, no file)."""
    rs, keep1 = _pr_stats(ref_pr, ref_order)
    st, keep2 = _pr_stats(pr, order)
    n = len(keep1[0])
    inv = np.empty(n, np.uint32)
    inv[keep2[1]] = np.arange(n, dtype=np.uint32)
    rinv = np.empty(n, np.uint32)
    rinv[keep1[1]] = np.arange(n, dtype=np.uint32)
    c = load().collectStatsPageRank_topK(C.byref(rs), C.byref(st), keep1[1].ctypes.data_as(u32p),
                                         rinv.ctypes.data_as(u32p), inv.ctypes.data_as(u32p), topk,
                                         n if num_vertices is None else num_vertices, None, 0)
    return c.as_dict()
