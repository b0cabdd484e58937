#!/usr/bin/env python3
"""Benchmark of the FRW hot path on one MI355X.

Workload (default, the plain GPU FRW configuration): 10^4 non-overlapping
rects in a 10^5 box, sides U{10..500} (Philox seed 1), device quadtree build,
master rect 0; one step = 10^7 walks from the master.  The metric is walk
steps/s (hops per second of walk-phase time, tree build excluded), with
time-to-1% capacitance (3 sigma) beside it.

Prints ONE JSON line (rank 0) with a roofline object for the walk kernel and
a cpu_baseline object: the reference's own CPU code (oracle/_ref, built from
the reference sources) timed on the host on a bounded sample of the same
workload.
"""
import argparse
import json
import math
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "engineer-thesis_amd"))
sys.path.insert(0, os.path.join(ROOT, "tests"))

HBM_PEAK_GBS = 8000.0

WORKLOADS = {
    # name: (n rects, box, side_min, side_max, cluster_frac, sigma, n_clusters, walks per step)
    "c2": (10_000, 100_000, 10, 500, 0.0, 0.0, 0, 10_000_000),
    "c3": (1_000_000, 10_000_000, 5, 300, 0.2, 1e5, 100, 10_000_000),
    "c1": (100, 10_000, 10, 200, 0.0, 0.0, 0, 1_000_000),
}
DESCR = {
    "c2": "plain GPU FRW: 1e4 rects, box 1e5, sides U[10,500], master 0, 1e7 walks/step",
    "c3": "FRW at scale: 1e6 rects, box 1e7, sides U[5,300], 20% in 100 Gaussian blobs "
          "(sigma 1e5), master 0, 1e7 walks/step",
    "c1": "CPU-reference layout: 100 rects, box 1e4, sides U[10,200], master 0, 1e6 walks/step",
}


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def parse_args():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", choices=sorted(WORKLOADS), default="c2")
    ap.add_argument("--mode", choices=["exact", "bisect"], default="exact")
    ap.add_argument("--walks", type=int, default=0, help="walks per step (0 = workload default)")
    ap.add_argument("--leaf-cap", type=int, default=16)
    ap.add_argument("--block", type=int, default=0)
    ap.add_argument("--cpu-seconds", type=float, default=12.0)
    ap.add_argument("--no-cpu", action="store_true")
    return ap.parse_args()


def cpu_baseline(rects, bounds, master, seconds):
    """Reference CPU walk (the reference's Tree/RectHost code) on one host
    thread over walk ids disjoint from the GPU's, sized to ~`seconds`."""
    import numpy as np
    import oracle_ffi as O
    r = rects.astype(np.float64)
    b = tuple(float(v) for v in bounds)
    kind = "reference" if O.ref_available() else "port"
    first = 1 << 31
    if kind == "reference":
        t = O.RefTree(r, b)

        def run(n, first_walk):
            t0 = time.perf_counter()
            _, _, hops, _ = t.walk(r[master], n, seed=1, first_walk=first_walk)
            return hops, time.perf_counter() - t0
    else:
        ix = O.Index(r, b, O.REFTREE_DFS)

        def run(n, first_walk):
            t0 = time.perf_counter()
            _, hops, _, _ = ix.walk(master, n, seed=1, first_walk=first_walk, nthreads=1)
            return hops, time.perf_counter() - t0
    n = 200
    hops, dt = run(n, first)
    while dt < 0.5 and n < 10 ** 7:
        n *= 4
        hops, dt = run(n, first)
    n_final = max(n, int(n * seconds / max(dt, 1e-6)))
    hops, dt = run(n_final, first + n)
    return {
        "value": hops / dt,
        "unit": "walk steps/s",
        "cores": 1,
        "kind": kind,
        "sample": f"{n_final} walks of the same layout and master (walk ids from 2^31), "
                  f"{hops} hops in {dt:.1f} s, single host thread (the reference CPU path is "
                  f"serial), per-walk /dev/urandom re-seed of the reference excluded",
        "walks_per_s": n_final / dt,
    }


def main():
    args = parse_args()
    world = int(os.environ.get("WORLD_SIZE", "1"))
    if args.gpus != 1 or world != 1:
        print("bench.py measures one GPU only (walks would shard by walk id across GPUs, "
              f"but no multi-GPU path is built); --gpus {args.gpus} is not supported.",
              file=sys.stderr)
        sys.exit(2)
    import numpy as np
    import torch
    import frw

    if not torch.cuda.is_available():
        print("bench.py needs a HIP device", file=sys.stderr)
        sys.exit(1)
    torch.cuda.init()
    n, box, smin, smax, cfrac, sigma, ncl, wps = WORKLOADS[args.workload]
    walks = args.walks or wps
    mode = frw.MODE_EXACT if args.mode == "exact" else frw.MODE_BISECT
    master = 0

    lay = frw.Layout.generate(n, box, smin, smax, seed=1, cluster_frac=cfrac, sigma=sigma,
                              n_clusters=ncl)
    tree = frw.Tree(lay, leaf_cap=args.leaf_cap)
    info = tree.info()
    nr = info.n_rects
    log(f"layout {n} rects, tree {info.n_nodes} nodes depth {info.depth}, build {info.build_ms:.2f} ms")

    stream = torch.cuda.Stream()
    sh = stream.cuda_stream
    counts = torch.zeros(nr + 2, dtype=torch.int64, device="cuda")
    hops = torch.zeros(1, dtype=torch.int64, device="cuda")

    def launch(step):
        tree.walk_async(master, walks, counts, hops, seed=1, mode=mode,
                        first_walk=step * walks, block=args.block, stream=sh)

    for w in range(args.warmup):
        launch(w)
    torch.cuda.synchronize()
    log("warmup done")
    counts.zero_()
    hops.zero_()
    torch.cuda.synchronize()
    ev = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True))
          for _ in range(args.steps)]
    t0 = time.perf_counter()
    with torch.cuda.stream(stream):
        for k in range(args.steps):
            ev[k][0].record(stream)
            launch(args.warmup + k)
            ev[k][1].record(stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    launch_ms = [a.elapsed_time(b) for a, b in ev]
    log(f"timed {args.steps} steps in {wall:.3f} s")
    total_hops = int(hops.item())
    c = counts.cpu().numpy()
    n_walks = walks * args.steps
    steps_per_s = total_hops / wall
    walks_per_s = n_walks / wall
    f = 1.0 - c[master] / n_walks
    n1 = math.ceil(9e4 * (1.0 - f) / f) if f > 0 else float("inf")
    t1pct = info.build_ms / 1e3 + n1 / walks_per_s

    # algorithmic bytes: node (32 B) and rect (16 B) records the traversals read,
    # counted by an instrumented run of the first 10^6 walks of the workload
    sub = min(walks, 1_000_000)
    tr_nodes, tr_rects, tr_hops = tree.walk_traffic(master, sub, seed=1, mode=mode,
                                                    first_walk=args.warmup * walks)
    bytes_per_hop = (32.0 * tr_nodes + 16.0 * tr_rects) / max(tr_hops, 1)
    log(f"traffic run: {tr_hops} hops, {bytes_per_hop:.1f} B/hop")
    avg_ms = sum(launch_ms) / len(launch_ms)
    bytes_per_launch = bytes_per_hop * total_hops / args.steps
    achieved = bytes_per_launch / (avg_ms * 1e-3) / 1e9

    out = {
        "metric": "walk steps/s",
        "value": steps_per_s,
        "unit": "walk steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": wall * 1e3 / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {
            "workload": DESCR[args.workload],
            "rects": nr,
            "walks_per_step": walks,
            "mode": args.mode,
            "leaf_cap": info.leaf_cap,
            "tree_nodes": info.n_nodes,
            "tree_depth": info.depth,
        },
        "walks_per_s": walks_per_s,
        "hops_per_walk": total_hops / n_walks,
        "fraction_other": f,
        "escaped": int(c[nr]) / n_walks,
        "tree_build_ms": info.build_ms,
        "time_to_1pct_s": t1pct,
        "walks_for_1pct": n1,
        "roofline": {
            "bound": "hbm",
            "achieved": achieved,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS,
            "traffic": None,
            "kernel": "walk_kernel",
            "bytes_per_hop": bytes_per_hop,
            "node_records_per_hop": tr_nodes / max(tr_hops, 1),
            "rect_records_per_hop": tr_rects / max(tr_hops, 1),
            "avg_launch_ms": avg_ms,
        },
    }
    if not args.no_cpu:
        log("cpu baseline ...")
        cb = cpu_baseline(lay.rects(), lay.bounds(), master, args.cpu_seconds)
        cb["time_to_1pct_s"] = n1 / cb.pop("walks_per_s")
        out["cpu_baseline"] = cb
    else:
        out["cpu_baseline"] = None
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
