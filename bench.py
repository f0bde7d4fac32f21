"""Benchmark: PageRank pull iterations on R-MAT scale 24 (one MI355X).

    python bench.py [--gpus 1] [--steps K] [--warmup W]

Workload: directed R-MAT, scale 24 (2^24 vertices),
edge factor 16 (2^28 edges), a=0.57 b=0.19 c=0.19, vertex ids permuted,
seed 1, duplicates and self-loops kept, generated on the GPU; CSR built on
the GPU (radix sort by source, -o 0); fp32 pull PageRank (-p 0, damping
0.85).  A step is one full pull iteration (every edge gathered once); the
stop rule is disabled while timing so every step does the full work.

Printed JSON line: value = GTEPS (edges traversed per second, 1e9) over the
K timed steps, with inputs resident in HBM; roofline = algorithmic bytes per
iteration (8E + 20N) over the dominant kernel's average duration, measured
with HIP events on the library's stream; cpu_baseline = the reference's own
pageRankPullGraphCSR (compiled from its sources into oracle/_ref) on the same
CSR on this host, over a bounded number of iterations.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "GTEPS + HBM GB/s per PR iter; time-to-L1<1e-6, learned vs plain, RMAT-24"
HBM_PEAK_GBS = 8000.0


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def cpu_baseline(g, iterations, threads):
    """Time the reference's compiled pull (or the oracle port) on this host."""
    from oracle import pyoracle as po
    arrays = g.arrays()
    csr = po.CSR(arrays, g.num_vertices, g.num_edges)
    kind = "reference" if po.ref_available() else "port"
    t0 = time.perf_counter()
    if kind == "reference":
        r = po.ref_pagerank(csr, 0, 0.0, iterations, threads)
        secs = r["seconds"]
    else:
        r = po.pagerank(csr, 0, 0.0, iterations, threads)
        secs = r["seconds"]
    wall = time.perf_counter() - t0
    per_iter = secs / iterations
    return dict(value=g.num_edges / per_iter / 1e9, unit="GTEPS", cores=threads, kind=kind,
                sample=f"R-MAT-24 pull, {iterations} full iterations (epsilon 0) of "
                       f"{'the reference pageRankPullGraphCSR compiled from its sources' if kind == 'reference' else 'the oracle C/OpenMP restatement'}"
                       f", timed by its own timer; {secs:.2f} s ({wall:.2f} s wall incl. setup)",
                seconds_per_iteration=per_iter), r["ranks"]


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
    ap.add_argument("--copy-reference", action="store_true",
                    help="after timing, run one 1 GiB device copy (og_copy_kernel): the reference "
                         "point that scales rocprofv3 FETCH_SIZE / WRITE_SIZE to bytes")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported (the path does not "
              f"span GPUs; see DESIGN.md for how it would shard).", file=sys.stderr)
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

    # convergence run (reference stop rule) for the record
    res = pr.run()
    ranks = pr.ranks()
    extra = dict(converge_iterations=int(res.iterations), converge_seconds=res.seconds_total,
                 plan_seconds=t_plan, csr_build_seconds=t_csr, generate_seconds=t_gen,
                 ms_per_iteration_wall=step_s * 1e3, dominant_kernel_ms=kern_s * 1e3,
                 algorithmic_gb_per_s=achieved)

    cpu = None
    if not args.no_cpu:
        threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
        try:
            cpu, cpu_ranks = cpu_baseline(g, args.cpu_iterations, threads)
            extra["speedup_vs_cpu_per_iteration"] = cpu["seconds_per_iteration"] / step_s
            log(f"cpu baseline ({cpu['kind']}, {threads} threads): {cpu['seconds_per_iteration']:.3f} s/iter")
        except Exception as e:   # the baseline is reported, never required
            log("cpu baseline failed:", e)

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
                     "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None,
                     "bytes_per_iteration": bytes_iter, "kernel": "og_pull_kernel<0>"},
        "cpu_baseline": cpu and {k: v for k, v in cpu.items() if k != "seconds_per_iteration"},
        "extra": extra,
        "device": dev,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
