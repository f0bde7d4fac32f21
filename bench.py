"""Benchmark: PageRank pull iterations on R-MAT scale 24 (one MI355X).

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu] [--no-extensions]
                    [--extensions spmv,rank_stats,...]

Workload: directed R-MAT, scale 24 (2^24 vertices), edge factor 16 (2^28
edges), a=0.57 b=0.19 c=0.19, vertex ids permuted, seed 1, duplicates and
self-loops kept, generated on the GPU; CSR built on the GPU (radix sort by
source, -o 0); fp32 pull PageRank (-p 0, damping 0.85).  A step is one full
pull iteration (every edge gathered once); the stop rule is disabled while
timing so every step does the full work.

Printed JSON line: value = GTEPS (edges traversed per second, 1e9) over the
K timed steps, with inputs resident in HBM; roofline = algorithmic bytes per
iteration (8E + 20N) over the iteration's average kernel time, measured with
HIP events on the library's stream around its three passes (the gather, then
the sum pass beside the long-row pass on the side stream); traffic = the HBM
bytes per iteration that rocprofv3 FETCH_SIZE + WRITE_SIZE counted for those
three kernels on this same command (profiles/pull_rmat24_counters.json,
scaled by a plain copy);
cpu_baseline = the reference's own pageRankPullGraphCSR (compiled from its
sources into oracle/_ref) on the same CSR on this host, over a bounded number
of iterations.

"extensions" holds the other measured forms, each also on one GPU: push and
fixed-point variants with their host runs, the reference's SpMV benchmark
(-p 0..3) on the same graph against its host run, the -S rank comparison of
two rankings of that graph against the reference's statistics code, the DBG
relabelling and the .csr file write / load, the Erdos-Renyi graph of the same
size, batched personalized PageRank (R-MAT 22, k = 64, against 64
single-source runs), 256 uneven graphs in one hipGraph launch (against
per-graph launches) and the learned warm start (time to L1 < 1e-6 against the
uniform start on R-MAT 24 graphs).  A failure there is reported in the line
and does not affect the headline number.
"""
import argparse
import json
import os
import sys
import time
import traceback

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "GTEPS + HBM GB/s per PR iter; time-to-L1<1e-6, learned vs plain, RMAT-24"
HBM_PEAK_GBS = 8000.0
COUNTERS = os.path.join(ROOT, "profiles", "pull_rmat24_counters.json")


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def _ref_csr(g):
    from oracle import pyoracle as po
    return po, po.CSR(g.arrays(), g.num_vertices, g.num_edges)


def cpu_time(g, code, iterations, threads, csr=None, ranks=False):
    """Seconds per iteration of the reference's compiled PageRank variant
    (or of the oracle port when the reference could not be built); with
    ranks, also its normalized ranks after those iterations."""
    po, c = _ref_csr(g) if csr is None else csr
    kind = "reference" if po.ref_available() else "port"
    if kind == "reference":
        r = po.ref_pagerank(c, code, 0.0, iterations, threads)
    else:
        r = po.pagerank(c, code, 0.0, iterations, threads)
    if ranks:
        return r["seconds"] / iterations, kind, r["ranks"]
    return r["seconds"] / iterations, kind


def _cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return None


def gpu_iter_ms(g, variant, steps=10, warmup=2, **kw):
    import opengraph_amd as og
    pr = og.PageRank(g, variant=variant, max_iter=100, epsilon=1e-6, **kw)
    pr.reset()
    pr.enqueue(warmup)
    og.synchronize()
    pr.kernel_ms()
    t0 = time.perf_counter()
    pr.enqueue(steps, time_kernels=True)
    og.synchronize()
    t = (time.perf_counter() - t0) / steps
    kms, kc = pr.kernel_ms()
    bpi = pr.bytes_per_iteration()
    pr.free()
    return t * 1e3, kms / max(kc, 1), bpi


def ext_variants(g, threads, csr, do_cpu):
    """Pull/push/fixed-point on the headline graph, each against the host run."""
    out, cpu = {}, {}
    for code, name, kw in ((0, "pull_fp32_gather_default", {"binned": 2}),
                           (0, "pull_fp32_gather_every_row_exact", {"hub_threshold": 1 << 31, "binned": 2}),
                           (0, "pull_fp32_gather_chunk_split_only", {"exact_above": 0, "binned": 2}),
                           (1, "push_fp32", {}), (1, "push_fp32_lds_accumulators", {"hot_cache": 1}),
                           (2, "pull_fixed64", {}), (3, "push_fixed64", {}), (11, "pull_fixed8", {})):
        ms, kms, bpi = gpu_iter_ms(g, code, **kw)
        d = dict(ms_per_iteration=round(ms, 4), gteps=round(g.num_edges / ms / 1e6, 2),
                 algorithmic_gb_per_s=round(bpi / (kms / 1e3) / 1e9, 1))
        if do_cpu and code in (1, 2, 3):
            if code not in cpu:
                cpu[code] = cpu_time(g, code, 2, threads, csr)
            s, kind = cpu[code]
            d.update(cpu_seconds_per_iteration=round(s, 4), cpu_kind=kind, speedup_vs_cpu=round(s / (ms / 1e3), 1))
        out[name] = d
        log(f"  {name}: {d}")
    return out


def spmv_iter_ms(g, code, steps=10, warmup=2, **kw):
    import opengraph_amd as og
    s = og.Spmv(g, code, **kw)
    s.reset()
    s.enqueue(warmup)
    og.synchronize()
    s.kernel_ms()
    t0 = time.perf_counter()
    s.enqueue(steps, time_kernels=True)
    og.synchronize()
    t = (time.perf_counter() - t0) / steps
    kms, kc = s.kernel_ms()
    bpi = s.bytes_per_iteration()
    s.free()
    return t * 1e3, kms / max(kc, 1), bpi


def ext_spmv(g, threads, csr):
    """The reference's SpMV benchmark (SPMVGraphCSR -p 0..3) on the headline
    graph: ms per pass and algorithmic GB/s of each form, the reference's own
    compiled SpMV on the host cores over 2 passes, and the GPU's 2-pass
    vector_output compared with the reference's (fixed point and the binned
    fp32 pull: bit-exact, recorded with the largest relative difference)."""
    import opengraph_amd as og
    out = {}
    forms = ((0, "pull_fp32", {}, 10), (0, "pull_fp32_gather", {"hub_threshold": 16}, 10),
             (0, "pull_fp32_gather_every_row_exact", {"hub_threshold": 1 << 31}, 3),
             (1, "push_fp32", {}, 5), (1, "push_fp32_plain_atomics", {"hot_cache": 2}, 5),
             (2, "pull_fixed32", {}, 10), (3, "push_fixed64_gathered", {}, 10),
             (3, "push_fixed64_scattered", {"scatter": 1}, 5))
    for code, name, kw, steps in forms:
        ms, kms, bpi = spmv_iter_ms(g, code, steps=steps, **kw)
        out[name] = dict(ms_per_pass=round(ms, 4), kernel_ms=round(kms, 4), gteps=round(g.num_edges / ms / 1e6, 2),
                         algorithmic_gb_per_s=round(bpi / (kms / 1e3) / 1e9, 1))
        log(f"  spmv {name}: {out[name]}")
    if csr is not None:
        po, c = csr
        if po.ref_available():
            indeg = g.array("i_out_degree")
            short = indeg <= 16
            for code, name in ((0, "pull_fp32"), (2, "pull_fixed32")):
                ref, sec = po.ref_spmv(c, code, 2, threads)
                s = og.Spmv(g, code)
                s.run(2)
                y = s.output()
                s.free()
                d = out[name]
                d.update(cpu_reference_seconds_per_pass=round(sec / 2, 4), cpu_threads=threads,
                         speedup_vs_cpu=round(sec / 2 / (d["ms_per_pass"] / 1e3), 1))
                if code == 2:
                    d["two_pass_output_bit_exact_vs_reference"] = bool(np.array_equal(y.view(np.uint32),
                                                                                      ref.view(np.uint32)))
                else:
                    rel = np.abs(y.astype(np.float64) - ref) / np.maximum(np.abs(ref.astype(np.float64)), 1e-30)
                    d["two_pass_output_bit_exact_vs_reference"] = bool(np.array_equal(y.view(np.uint32),
                                                                                      ref.view(np.uint32)))
                    d["two_pass_max_rel_diff_long_rows"] = float(rel[~short].max()) if (~short).any() else 0.0
                    # both against the exact (fp64) two-pass value: the difference is
                    # the reference's own fp32 running-sum drift on the long rows
                    od = c["f_out_degree"].astype(np.float64)
                    x = np.where(od > 0, (1.0 / np.maximum(od, 1)).astype(np.float32), np.float32(0.001))
                    exact = 2 * 0.0001 * np.bincount(c["f_dst"], weights=x.astype(np.float64)[c["f_src"]],
                                                     minlength=g.num_vertices)
                    den = np.maximum(exact, 1e-30)
                    d["two_pass_max_rel_err_vs_fp64_gpu"] = float((np.abs(y - exact) / den).max())
                    d["two_pass_max_rel_err_vs_fp64_reference"] = float((np.abs(ref - exact) / den).max())
                log(f"  spmv {name} vs reference: {d}")
    return out


def ext_rank_stats(g):
    """The -S comparison of the fixed-point 8-bit pull ranking (-p 11) with
    the pull ranking on the headline graph, both ordered by the library's
    radixSortEdgesByPageRank: every problem collectStatsPageRank computes at
    verbosity 2 (top-k 30 .. 10000, all 1000-wide chunks, top 512) on the
    GPU with the rankings resident, against the reference's own compiled
    collectStatsPageRank_topK on the host for the top-k table and a sample
    of 40 chunks (field-for-field equal to the GPU's), scaled to all chunks."""
    import ctypes as C
    import opengraph_amd as og
    from oracle import pyoracle as po
    N = g.num_vertices
    lib = og.load()
    rk = {}
    for code in (0, 11):
        pr = og.PageRank(g, code, 20, 1e-4)
        pr.run()
        r = pr.ranks()
        pr.free()
        lab = np.arange(N, dtype=np.uint32)
        lib.radixSortEdgesByPageRank(r.ctypes.data_as(C.POINTER(C.c_float)), lab.ctypes.data_as(C.POINTER(C.c_uint32)), N)
        rk[code] = (r, lab)
    (ref_pr, ref_order), (pr, order) = rk[0], rk[11]
    tops = [k for k in (30, 100, 300, 1000, 5000, 10000) if k <= N]
    chunks = [(1000, N - 1000 * x) for x in range((N + 999) // 1000)]
    probs = [(k, N) for k in tops] + chunks + [(512, N)]
    arrs = [og.DeviceArray.from_numpy(a) for a in (ref_pr, ref_order, pr, order)]
    try:
        times = []
        for _ in range(3):
            og.synchronize()
            t0 = time.perf_counter()
            got = og.rank_stats_device(*arrs, N, probs)
            times.append(time.perf_counter() - t0)
        t0 = time.perf_counter()
        og.rank_stats(ref_pr, ref_order, pr, order, [(k, N) for k in tops])
        t_table_h2d = time.perf_counter() - t0
        t0 = time.perf_counter()
        og.rank_stats_device(*arrs, N, [(k, N) for k in tops])
        t_table = time.perf_counter() - t0
    finally:
        for a in arrs:
            a.free()
    out = dict(workload=f"R-MAT-24 -p 11 vs -p 0 rankings; {len(probs)} problems (top-k {tops}, "
                        f"{len(chunks)} chunks of 1000, top 512)",
               gpu_all_problems_seconds=round(min(times), 4), gpu_topk_table_seconds=round(t_table, 4),
               gpu_topk_table_with_upload_seconds=round(t_table_h2d, 4),
               top10000=got[len(tops) - 1] if tops else None)
    if po.ref_available():
        ref_tab, t_ref_table = po.ref_rank_stats_batch(ref_pr, ref_order, pr, order, [(k, N) for k in tops])
        sample = list(range(0, len(chunks), max(1, len(chunks) // 40)))[:40]
        ref_ch, t_ref_ch = po.ref_rank_stats_batch(ref_pr, ref_order, pr, order, [chunks[x] for x in sample])
        t_ref_chunk = t_ref_ch / len(sample)
        same = ref_tab == got[:len(tops)] and ref_ch == [got[len(tops) + x] for x in sample]
        out.update(cpu_reference_topk_table_seconds=round(t_ref_table, 3),
                   cpu_reference_seconds_per_chunk=round(t_ref_chunk, 5),
                   cpu_reference_all_problems_seconds_scaled=round(t_ref_table + t_ref_chunk * len(chunks), 1),
                   cpu_sample=f"top-k table + {len(sample)} of {len(chunks)} chunks, one thread (the reference's "
                              f"collectStatsPageRank is serial), collectStatsPageRank_topK calls only",
                   results_equal_to_reference=bool(same),
                   speedup_topk_table=round(t_ref_table / t_table, 1),
                   speedup_all_problems=round((t_ref_table + t_ref_chunk * len(chunks)) / min(times), 1))
    return out


def ext_relabel(g, threads, csr):
    """DBG relabelling (-l 4) of the headline edge list and the CSR build of
    the relabelled graph on the GPU, against the plain CSR build, and the
    reference's own compiled reorderGraphProcess on the host cores (the
    relabelling alone) with a check that both give the same labels."""
    import opengraph_amd as og
    f_src, f_dst = g.array("f_src"), g.array("f_dst")
    N, E = g.num_vertices, g.num_edges
    ds, dd = og.DeviceArray.from_numpy(f_src), og.DeviceArray.from_numpy(f_dst)
    t_plain, t_re, labels = [], [], None
    for _ in range(3):
        og.synchronize()
        t0 = time.perf_counter()
        h = og.Graph.build(ds, dd, E, 0)
        og.synchronize()
        t_plain.append(time.perf_counter() - t0)
        h.free()
        t0 = time.perf_counter()
        h, labels = og.Graph.build_reordered(ds, dd, E, N, 0, (4,))
        og.synchronize()
        t_re.append(time.perf_counter() - t0)
        h.free()
    ds.free(); dd.free()
    out = dict(workload=f"R-MAT-24 forward edge list, -l 4 (DBG by out-degree), sort 0",
               gpu_relabel_and_csr_seconds=round(min(t_re), 4), gpu_csr_only_seconds=round(min(t_plain), 4),
               gpu_relabel_seconds=round(min(t_re) - min(t_plain), 4))
    if csr is not None:
        po, _ = csr
        if po.ref_available():
            ref_labels, sec = po.ref_reorder(f_src, f_dst, N, 4, threads)
            out.update(cpu_reference_relabel_seconds=round(sec, 4), cpu_threads=threads,
                       labels_equal_to_reference=bool(np.array_equal(ref_labels, labels)),
                       speedup_relabel_vs_cpu=round(sec / max(min(t_re) - min(t_plain), 1e-9), 1))
    return out


def ext_csr_file():
    """The reference's .csr file of an R-MAT-20 graph (seed 9): write it from
    the device graph and load it back (GPU interleaving, one file read or
    write), against building the CSR from the edge list on the GPU."""
    import tempfile
    import opengraph_amd as og
    scale, E = 20, 16 << 20
    ds, dd = og.rmat_device(scale, E, 0.57, 0.19, 0.19, 9, True)
    og.synchronize()
    t0 = time.perf_counter()
    g = og.Graph.build(ds, dd, E, 0)
    og.synchronize()
    t_build = time.perf_counter() - t0
    ds.free(); dd.free()
    with tempfile.TemporaryDirectory() as td:
        base = os.path.join(td, "rmat20")
        t0 = time.perf_counter()
        g.write_csr(base)
        t_write = time.perf_counter() - t0
        size = os.path.getsize(base + ".csr")
        t0 = time.perf_counter()
        h = og.Graph.read_csr(base + ".csr")
        og.synchronize()
        t_read = time.perf_counter() - t0
    same = all(np.array_equal(a, b) for a, b in zip(g.arrays().values(), h.arrays().values()))
    g.free(); h.free()
    return dict(workload="R-MAT-20 (seed 9), E = 16 Mi, sort 0", file_bytes=size,
                csr_build_from_edges_seconds=round(t_build, 4), write_seconds=round(t_write, 4),
                read_to_device_seconds=round(t_read, 4), read_back_identical=same)


def ext_er(args, threads, do_cpu):
    """The Erdos-Renyi graph of the headline's size (seed 2): binned pull
    (the default) and the gather pull, against the reference's own pull on
    the host cores, with the binned ranks compared to the reference's."""
    import opengraph_amd as og
    E = args.edge_factor << args.scale
    ds, dd = og.er_device(1 << args.scale, E, 2)
    g = og.Graph.build(ds, dd, E, 0)
    ds.free(); dd.free()
    out = dict(workload=f"Erdos-Renyi N=2^{args.scale} E={E} seed 2, pull fp32")
    for name, kw in (("binned", {}), ("gather", {"binned": 2})):
        ms, kms, bpi = gpu_iter_ms(g, 0, **kw)
        out[name] = dict(ms_per_iteration=round(ms, 4), gteps=round(E / ms / 1e6, 2),
                         algorithmic_gb_per_s=round(bpi / (kms / 1e3) / 1e9, 1),
                         roofline_frac=round(bpi / (kms / 1e3) / 1e9 / HBM_PEAK_GBS, 4))
    if do_cpu:
        csr = _ref_csr(g)
        s, kind, ref_ranks = cpu_time(g, 0, args.cpu_iterations, threads, csr, ranks=True)
        pr = og.PageRank(g, 0, args.cpu_iterations, 0.0)
        pr.run()
        x = pr.ranks()
        pr.free()
        out.update(cpu_seconds_per_iteration=round(s, 4), cpu_kind=kind, cpu_threads=threads,
                   speedup_vs_cpu=round(s / (out["binned"]["ms_per_iteration"] / 1e3), 1),
                   binned_ranks_bit_identical_to_reference=bool(np.array_equal(x.view(np.uint32),
                                                                              ref_ranks.view(np.uint32))),
                   parity_iterations=args.cpu_iterations)
    g.free()
    return out


def ext_csr_build(args, threads, do_cpu):
    """The CSR builder at the headline size: R-MAT-24 (seed 1) edges on the
    device, forward + inverse CSR by the GPU radix sort (-o 0: by source;
    -o 1: by destination then source), best of 3, inputs and outputs in HBM;
    algorithmic bytes 20 B per pair per 8-bit LSD pass (4 B key read for the
    histogram, 8 B pair read, 8 B pair write) over both lists, plus the
    degree count and offsets (4E + 8N).  Against the reference's own
    sortRunAlgorithms + graphCSRAssignEdgeList (forward and inverse) on the
    host cores, with all ten arrays compared."""
    import opengraph_amd as og
    from oracle import pyoracle as po
    E = args.edge_factor << args.scale
    ds, dd = og.rmat_device(args.scale, E, 0.57, 0.19, 0.19, args.seed, True)
    out = dict(workload=f"R-MAT-{args.scale} (seed {args.seed}) edge list in HBM -> forward and inverse CSR")
    src = dst = None
    for sort in (0, 1):
        times = []
        for _ in range(3):
            og.synchronize()
            t0 = time.perf_counter()
            g = og.Graph.build(ds, dd, E, sort)
            og.synchronize()
            times.append(time.perf_counter() - t0)
            if _ < 2:
                g.free()
        N = g.num_vertices
        key_bits = max(1, int(N - 1).bit_length())
        passes = 2 * ((key_bits + 7) // 8) * (2 if sort == 1 else 1)
        bpb = 20 * E * passes + 4 * E + 8 * N
        t = min(times)
        d = dict(seconds=round(t, 4), lsd_passes=passes, algorithmic_bytes=bpb,
                 algorithmic_gb_per_s=round(bpb / t / 1e9, 1), roofline_frac=round(bpb / t / 1e9 / HBM_PEAK_GBS, 4),
                 medges_per_second=round(E / t / 1e6, 1))
        if do_cpu and po.ref_available():
            if src is None:
                src, dst = ds.to_numpy(np.uint32, E), dd.to_numpy(np.uint32, E)
            ref, sec = po.ref_csr_build(src, dst, N, sort, threads)
            same = all(np.array_equal(g.array(k), ref[k]) for k in po.CSR_FIELDS)
            del ref
            d.update(cpu_reference_seconds=round(sec, 3), cpu_threads=threads, cpu_kind="reference",
                     cpu_sample="the reference's sortRunAlgorithms (radix) and graphCSRAssignEdgeList "
                                "(mapVerticesWithInOutDegree), forward and inverse, full size",
                     speedup_vs_cpu=round(sec / t, 1), arrays_bit_identical_to_reference=bool(same))
        g.free()
        out[f"sort{sort}"] = d
        log(f"  csr build -o {sort}: {d}")
    ds.free(); dd.free()
    return out


def ext_ppr():
    """Batched personalized PageRank, R-MAT 22 (seed 3), k = 64 one-hot
    sources (seed 4), tol 1e-6 per column, against 64 single-source runs."""
    import opengraph_amd as og
    scale, E = 22, 16 << 22
    ds, dd = og.rmat_device(scale, E, 0.57, 0.19, 0.19, 3, True)
    g = og.Graph.build(ds, dd, E, 0)
    ds.free(); dd.free()
    src = np.random.default_rng(4).integers(0, g.num_vertices, size=64, dtype=np.uint32)
    p = og.Ppr(g, src, 100, 1e-6)
    p.run()                                   # first run pays one-time costs
    r = p.run()
    bpi = p.bytes_per_iteration()
    X = p.ranks()
    iters, _ = p.columns()
    p.free()
    t_single, it_single, maxdiff = 0.0, [], 0.0
    for j in range(64):
        pr = og.PageRank(g, 0, 100, 1e-6, ppr_source=int(src[j]))
        t0 = time.perf_counter()
        res = pr.run()
        t_single += time.perf_counter() - t0
        it_single.append(res.rows)
        if j < 8:
            x = pr.ranks()
            maxdiff = max(maxdiff, float(np.max(np.abs(X[:, j] - x) / np.maximum(np.abs(x), 1e-30))))
        pr.free()
    g.free()
    return dict(workload="R-MAT 22 (seed 3), k=64 one-hot sources (seed 4), tol 1e-6 per column",
                bytes_note="algorithmic bytes per iteration count only the column slots still gathered "
                           "(converged columns leave them), averaged over the executed iterations",
                iterations=int(r.iterations), converged=int(r.converged),
                batched_seconds=round(r.seconds_total, 4), spmm_kernel_ms_per_iteration=round(r.kernel_ms_per_iter, 4),
                algorithmic_gb_per_s=round(bpi / (r.kernel_ms_per_iter / 1e3) / 1e9, 1),
                roofline_frac=round(bpi / (r.kernel_ms_per_iter / 1e3) / 1e9 / HBM_PEAK_GBS, 4),
                single_source_seconds_total=round(t_single, 4),
                single_source_iterations_mean=round(float(np.mean(it_single)), 2),
                speedup_vs_single_source=round(t_single / r.seconds_total, 2),
                max_rel_diff_vs_single_first8=maxdiff)


def ext_multi():
    """256 uneven R-MAT graphs (scale U{12..20}, edge factor U{8..32},
    a U[0.45, 0.65], b = c = (1 - a - 0.05)/2, master seed 5): plain PageRank
    (tol 1e-6) in one hipGraph launch against per-graph launches."""
    import opengraph_amd as og
    rng = np.random.default_rng(5)
    gs, sumN, sumE = [], 0, 0
    t0 = time.perf_counter()
    for i in range(256):
        sc = int(rng.integers(12, 21)); ef = int(rng.integers(8, 33)); a = float(rng.uniform(0.45, 0.65))
        b = (1 - a - 0.05) / 2
        E = ef << sc
        ds, dd = og.rmat_device(sc, E, a, b, b, 5000 + i, True)
        gs.append(og.Graph.build(ds, dd, E, 0))
        ds.free(); dd.free()
        sumN += gs[-1].num_vertices; sumE += E
    t_build = time.perf_counter() - t0
    t0 = time.perf_counter()
    m = og.Multi(gs, 100, 1e-6)               # autotuned row-class thresholds and tile shapes
    t_setup = time.perf_counter() - t0
    m.run(2)                                  # capture + first launch
    t_graph = min(m.run(2) for _ in range(3))
    t_each = min(m.run(0) for _ in range(2))
    t_launch = m.run(1)
    iters = [m.result(i, ranks=False)[0].iterations for i in range(0, 256, 16)]
    hubs = [m.hub_threshold(i) for i in range(256)]
    by_class = {}
    for i, g in enumerate(gs):
        cls = max(int(g.num_vertices - 1).bit_length(), 0)   # ceil(log2 N)
        by_class.setdefault(cls, (hubs[i], m.tile(i)))
    m.free()
    # the same batch with the one tile shape all graphs had before tuning
    m = og.Multi(gs, 100, 1e-6, hub_thresholds=hubs, tiles=[(256, 4, 8)] * 256)
    m.run(2)
    t_fixed = min(m.run(2) for _ in range(2))
    m.free()
    for g in gs:
        g.free()
    tuned = {f"2^{c}": dict(hub_threshold=h, block=t[0], chunks_per_wave=t[1], row_batch=t[2])
             for c, (h, t) in sorted(by_class.items())}
    return dict(workload="256 R-MAT graphs, scale U{12..20}, ef U{8..32}, a U[0.45,0.65], seed 5; tol 1e-6",
                sum_vertices=sumN, sum_edges=sumE, build_seconds=round(t_build, 2),
                setup_and_autotune_seconds=round(t_setup, 2), tuned_thresholds=sorted(set(hubs)),
                tuned_per_size_class=tuned,
                one_hipgraph_seconds=round(t_graph, 4), per_graph_launch_seconds=round(t_each, 4),
                batched_launch_seconds=round(t_launch, 4), speedup_hipgraph_vs_per_graph=round(t_each / t_graph, 2),
                fixed_tile_256_4_8_hipgraph_seconds=round(t_fixed, 4),
                speedup_tuned_tiles_vs_fixed=round(t_fixed / t_graph, 3),
                sample_iterations=iters)


def ext_learned(n_train=200, n_eval=8, epochs=8):
    """Learned warm start: MLP trained on converged ranks of 200 R-MAT-18
    graphs (a U[0.45, 0.65], ef U{8..32}, seed 6) by the library's own
    kernels; evaluated on R-MAT-24 graphs (seeds 100..) by wall time to
    L1 < 1e-6 against the uniform start (warm-start time included)."""
    import opengraph_amd as og

    def graph(scale, seed, rng):
        a = float(rng.uniform(0.45, 0.65)); ef = int(rng.integers(8, 33)); b = (1 - a - 0.05) / 2
        E = ef << scale
        ds, dd = og.rmat_device(scale, E, a, b, b, seed, True)
        g = og.Graph.build(ds, dd, E, 0)
        ds.free(); dd.free()
        return g, a, ef

    rng = np.random.default_rng(6)
    m = og.Mlp(6)
    t0 = time.perf_counter()
    worst = 0.0
    for i in range(n_train):
        g, _, _ = graph(18, 6000 + i, rng)
        worst = max(worst, m.add_graph(g, 1e-9, 2000))
        g.free()
    t_data = time.perf_counter() - t0
    loss, t_train = m.train(epochs, 65536, 3e-3, 1)
    n_vertices = m.data_size()
    # the method without the trained network: a constant prediction (zero
    # weights, softplus(1) everywhere) through the same scaling and class
    # correction
    const = og.Mlp(6)
    pc = np.zeros(const.params().size, np.float32)
    pc[-1] = 1.0
    const.set_params(pc)
    evals = []
    for k in range(n_eval):
        g, a, ef = graph(24, 100 + k, np.random.default_rng(100 + k))
        row = dict(seed=100 + k, a=round(a, 3), edge_factor=ef)
        for name, net in (("uniform", None), ("constant_corrected", const), ("learned", m)):
            pr = og.PageRank(g, 0, 1000, 1e-6, stop_rule=1, init_mode=1)
            og.synchronize()
            t = time.perf_counter()
            ws = pr.set_init_learned(net) if net is not None else 0.0
            r = pr.run()
            t = time.perf_counter() - t
            row[name] = dict(iterations=int(r.rows), seconds=round(t, 4), warm_start_seconds=round(ws, 4),
                             last_l1=r.error_total)
            pr.free()
        row["time_saved_frac"] = round(1 - row["learned"]["seconds"] / row["uniform"]["seconds"], 4)
        evals.append(row)
        log(f"  learned eval {row}")
        g.free()
    m.free()
    const.free()
    tu = sum(e["uniform"]["seconds"] for e in evals)
    tl = sum(e["learned"]["seconds"] for e in evals)
    tc = sum(e["constant_corrected"]["seconds"] for e in evals)
    return dict(workload=f"train {n_train} R-MAT-18 (seed 6), eval {n_eval} R-MAT-24 (seeds 100..), L1<1e-6",
                training_vertices=int(n_vertices), data_seconds=round(t_data, 2),
                train_epochs=epochs, train_seconds=round(t_train, 2), final_train_loss=round(loss, 5),
                worst_training_target_l1=worst, eval=evals,
                total_seconds_uniform=round(tu, 4), total_seconds_learned=round(tl, 4),
                time_saved_frac=round(1 - tl / tu, 4),
                iterations_uniform_mean=float(np.mean([e["uniform"]["iterations"] for e in evals])),
                iterations_learned_mean=float(np.mean([e["learned"]["iterations"] for e in evals])),
                total_seconds_constant_corrected=round(tc, 4),
                iterations_constant_corrected_mean=float(np.mean([e["constant_corrected"]["iterations"]
                                                                   for e in evals])),
                time_saved_frac_vs_constant_corrected=round(1 - tl / tc, 4))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--scale", type=int, default=24)
    ap.add_argument("--edge-factor", type=int, default=16)
    ap.add_argument("--seed", type=int, default=1)
    ap.add_argument("--cpu-iterations", type=int, default=3)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-extensions", action="store_true")
    ap.add_argument("--extensions", default="",
                    help="comma-separated subset of the extension legs to run (default: all)")
    ap.add_argument("--copy-reference", action="store_true",
                    help="after timing, run one 1 GiB device copy (og_copy_kernel): the reference "
                         "point that scales rocprofv3 FETCH_SIZE / WRITE_SIZE to bytes")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported: the library runs "
              f"each graph on a single GPU (sharding by destination rows is not built).", file=sys.stderr)
        sys.exit(2)

    import opengraph_amd as og
    og.init(0)
    dev, cus, mem = og.device_info()
    log(f"device {dev}, {cus} CUs")
    scale, E = args.scale, args.edge_factor << args.scale

    t0 = time.perf_counter()
    d_src, d_dst = og.rmat_device(scale, E, 0.57, 0.19, 0.19, args.seed, True)
    og.synchronize()
    t_gen = time.perf_counter() - t0
    t0 = time.perf_counter()
    g = og.Graph.build(d_src, d_dst, E, 0)
    og.synchronize()
    t_csr = time.perf_counter() - t0
    if not args.copy_reference:
        d_src.free()
        d_dst.free()
    N = g.num_vertices
    log(f"R-MAT-{scale}: N={N} E={E}; generate {t_gen:.3f}s, CSR build {t_csr:.3f}s")

    t0 = time.perf_counter()
    pr = og.PageRank(g, variant=0, max_iter=100, epsilon=1e-6)
    t_plan = time.perf_counter() - t0
    bytes_iter = pr.bytes_per_iteration()

    # warmup, then exactly K timed steps bracketed by synchronization
    pr.reset()
    pr.enqueue(args.warmup)
    og.synchronize()
    pr.kernel_ms()           # drop warmup events
    og.synchronize()
    t0 = time.perf_counter()
    pr.enqueue(args.steps, time_kernels=True)
    og.synchronize()
    t1 = time.perf_counter()
    kms, kcount = pr.kernel_ms()
    step_s = (t1 - t0) / args.steps
    kern_s = kms / 1e3 / max(kcount, 1)
    gteps = E * args.steps / (t1 - t0) / 1e9
    achieved = bytes_iter / kern_s / 1e9
    log(f"pull: {step_s * 1e3:.3f} ms/iter wall, dominant kernel {kern_s * 1e3:.3f} ms, "
        f"{gteps:.1f} GTEPS, {achieved:.0f} GB/s algorithmic")

    if args.copy_reference:
        og.device_copy(d_dst, d_src, 1 << 30)
        d_src.free()
        d_dst.free()

    # convergence runs for the record: the reference stop rule, and L1 < 1e-6
    # from the uniform start (the plain side of the learned comparison)
    res = pr.run()
    pr.free()
    pr1 = og.PageRank(g, 0, 1000, 1e-6, stop_rule=1, init_mode=1)
    t0 = time.perf_counter()
    r1 = pr1.run()
    t_l1 = time.perf_counter() - t0
    pr1.free()
    extra = dict(converge_iterations=int(res.iterations), converge_seconds=res.seconds_total,
                 uniform_to_l1_1e6_iterations=int(r1.rows), uniform_to_l1_1e6_seconds=round(t_l1, 4),
                 plan_seconds=t_plan, csr_build_seconds=t_csr, generate_seconds=t_gen,
                 ms_per_iteration_wall=step_s * 1e3, dominant_kernel_ms=kern_s * 1e3,
                 algorithmic_gb_per_s=achieved)

    cpu = None
    csr = None
    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    if not args.no_cpu:
        try:
            csr = _ref_csr(g)
            s, kind, ref_ranks = cpu_time(g, 0, args.cpu_iterations, threads, csr, ranks=True)
            cpu = dict(value=round(E / s / 1e9, 4), unit="GTEPS", cores=threads, kind=kind,
                       cpu_model=_cpu_model(), nproc=os.cpu_count(),
                       sample=f"R-MAT-{scale} pull, {args.cpu_iterations} full iterations (epsilon 0) of "
                              + ("the reference pageRankPullGraphCSR compiled from its sources"
                                 if kind == "reference" else "the oracle C/OpenMP restatement")
                              + f", OpenMP on {threads} threads, timed by its own timer: {s:.3f} s/iteration")
            extra["speedup_vs_cpu_per_iteration"] = round(s / step_s, 1)
            log(f"cpu baseline ({kind}, {threads} threads): {s:.3f} s/iter")
            # full-size parity on the same iterations: the GPU's ranks against the
            # reference's, and both against the FP64 oracle of the same map
            po, c = csr
            x64, _, _ = po.pagerank_fp64(c, 0.15, 0.0, args.cpu_iterations, threads)
            par = dict(iterations=args.cpu_iterations,
                       l1_vs_fp64_reference=float(np.abs(ref_ranks.astype(np.float64) - x64).sum()))
            for form, kw in (("binned_default", {}),
                             ("gather_exact_above_4096", {"binned": 2}),
                             ("gather_chunk_split_only", {"exact_above": 0, "binned": 2}),
                             ("gather_every_row_exact", {"hub_threshold": 1 << 31, "binned": 2})):
                pr3 = og.PageRank(g, 0, args.cpu_iterations, 0.0, **kw)
                pr3.run()
                gpu_ranks = pr3.ranks()
                pr3.free()
                rel = np.abs(gpu_ranks.astype(np.float64) - ref_ranks) / np.maximum(ref_ranks.astype(np.float64), 1e-300)
                par[form] = dict(
                    max_rel_err_vs_reference=float(rel.max()),
                    ranks_bit_identical_fraction=float(np.mean(gpu_ranks.view(np.uint32) == ref_ranks.view(np.uint32))),
                    l1_vs_fp64=float(np.abs(gpu_ranks.astype(np.float64) - x64).sum()))
            extra["parity_full_size"] = par
            log(f"full-size parity: {extra['parity_full_size']}")
        except Exception as e:   # the baseline is reported, never required
            log("cpu baseline failed:", e)

    traffic = None
    try:
        with open(COUNTERS) as f:
            cnt = json.load(f)
        traffic = cnt.get("hbm_bytes_per_launch")
    except OSError:
        pass

    extensions = {}
    if not args.no_extensions:
        for name, fn in (("variants", lambda: ext_variants(g, threads, csr, csr is not None)),
                         ("spmv", lambda: ext_spmv(g, threads, csr)),
                         ("rank_stats", lambda: ext_rank_stats(g)),
                         ("relabel_dbg", lambda: ext_relabel(g, threads, csr)),
                         ("csr_file", ext_csr_file),
                         ("csr_build", lambda: ext_csr_build(args, threads, csr is not None)),
                         ("erdos_renyi", lambda: ext_er(args, threads, csr is not None)), ("batched_ppr", ext_ppr),
                         ("uneven_batch", ext_multi), ("learned_warm_start", ext_learned)):
            if args.extensions and name not in args.extensions.split(","):
                continue
            t0 = time.perf_counter()
            try:
                if name == "batched_ppr" and g is not None:
                    g.free()          # free the headline graph before the big extensions
                    g = None
                extensions[name] = fn()
            except Exception as e:
                extensions[name] = {"error": f"{type(e).__name__}: {e}"}
                log(traceback.format_exc())
            extensions[name]["seconds"] = round(time.perf_counter() - t0, 2)
            log(f"{name}: {extensions[name]}")
    if g is not None:
        g.free()

    line = {
        "metric": METRIC, "value": round(gteps, 3), "unit": "GTEPS", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(step_s * 1e3, 4),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
        "data": f"synthetic R-MAT scale {scale}, edge factor {args.edge_factor}, seed {args.seed}, "
                f"generated on the GPU",
        "config": {"workload": f"rmat{scale}_pull_pagerank", "scale": scale, "edge_factor": args.edge_factor,
                   "a": 0.57, "b": 0.19, "c": 0.19, "seed": args.seed, "num_vertices": N, "num_edges": E,
                   "variant": "pull fp32 (-p 0)", "sort": 0},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic,
                     "bytes_per_iteration": bytes_iter,
                     "kernel": "og_bin_gather_kernel, then og_bin_sum_kernel beside og_bin_long_kernel on the side "
                               "stream (one iteration's passes; HIP events on the library stream from before the "
                               "gather to after the join)",
                     "traffic_source": "profiles/pull_rmat24_counters.json" if traffic else None},
        "cpu_baseline": cpu,
        "extra": extra,
        "extensions": extensions,
        "device": dev,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
