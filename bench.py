#!/usr/bin/env python3
"""Benchmark of the FRW hot path on one MI355X.

Default workload (learned + batched at 10^6 rects): the 10^6-rect layout
(box 10^7, sides U{5..300}, 20 % of the rects in 100 Gaussian blobs of sigma
10^5, Philox seed 1), device quadtree build, 16 masters (rect 0 and 15 drawn
uniformly) with walks per master U[5e5, 2e6], all in one persistent launch per
step replayed from a captured HIP graph, launch shape from the autotuner.  The
learned proposal's network is trained first on 10^7 (point, exact square)
pairs harvested from exact walks of the same masters.  One step = one launch
over every job.  The metric is walk steps/s (hops per second of walk time)
with walks/s and time-to-1 % capacitance (3 sigma) beside it; the same batch
runs in the same process with the plain FRW's own square search (the
reference's bisection, one collision query per step) and with the exact
search (the bisection replayed from one nearest-rect query: the plain FRW's
very walks, computed faster).

Prints ONE JSON line with a roofline object for the dominant kernel (the
learned walk kernel: MFMA FLOP/s live from HIP events, HBM traffic and
instruction counts from the committed rocprofv3 counter summary) and a
cpu_baseline object: the reference's own CPU code (oracle/_ref, built from the
reference sources) timed on the host on a bounded sample of the same layout.
"""
import argparse
import json
import math
import os
import re
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "engineer-thesis_amd"))
sys.path.insert(0, os.path.join(ROOT, "tests"))

HBM_PEAK_GBS = 8000.0
BF16_PEAK_TFLOPS = 2500.0
# one VALU (or SALU) wave instruction per CU per cycle: 256 CUs x 4 SIMDs x
# 2.4 GHz / 4 cycles per wave64 instruction
ISSUE_PEAK = 256 * 4 * 2.4e9 / 4
# network FLOP per proposal: 2 (67*64 + 64*64 + 64*64 + 64)
FLOP_PER_PROPOSAL = 2 * (67 * 64 + 64 * 64 + 64 * 64 + 64)
# training FLOP per sample: the forward pass, the weight gradients (as many
# multiply-adds again) and the data gradients of layers 3, 2 and the output
# (64 + 2 * 64*64)
TRAIN_FLOP_PER_SAMPLE = FLOP_PER_PROPOSAL * 2 + 2 * (64 + 2 * 64 * 64)
COUNTERS_MD = os.path.join(ROOT, "profiles", "walk_kernel_batched_1e6rects_16masters.md")

WORKLOADS = {
    # name: (n rects, box, side_min, side_max, cluster_frac, sigma, n_clusters, walks per step)
    "c2": (10_000, 100_000, 10, 500, 0.0, 0.0, 0, 10_000_000),
    "c3": (1_000_000, 10_000_000, 5, 300, 0.2, 1e5, 100, 100_000_000),
    "c1": (100, 10_000, 10, 200, 0.0, 0.0, 0, 1_000_000),
    "c5": (1_000_000, 10_000_000, 5, 300, 0.2, 1e5, 100, 100_000_000),
}
DESCR = {
    "c2": "plain GPU FRW: 1e4 rects, box 1e5, sides U[10,500], master 0, 1e7 walks/step",
    "c3": "FRW at scale: 1e6 rects, box 1e7, sides U[5,300], 20% in 100 Gaussian blobs "
          "(sigma 1e5), master 0, 1e8 walks/step",
    "c1": "CPU-reference layout: 100 rects, box 1e4, sides U[10,200], master 0, 1e6 walks/step",
    "c4": "batched: 64 layouts, rects log-uniform in [1e2,1e5] (box 1000 sqrt(n), sides "
          "U[10,500]), master 0 each, walks per layout U[1e5,1e6], one persistent launch per "
          "step, graph replay, autotuned launch shape",
    "c5": "learned proposal at scale: the 1e6-rect layout, network trained on pairs harvested "
          "from exact walks, master 0, 1e8 walks/step, exact fallback",
    "c6": "learned+batched at 1e6 rects: the 1e6-rect layout (box 1e7, sides U[5,300], 20% in 100 "
          "Gaussian blobs), 16 masters (rect 0 and 15 drawn uniformly), walks per master U[5e5,2e6], "
          "one persistent launch per step, graph replay, autotuned launch shape; network trained on "
          "1e7 pairs harvested from exact walks of the 16 masters; plain FRW (bisection) and exact "
          "search timed on the same batch",
}


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def parse_args():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", choices=sorted(WORKLOADS) + ["c4", "c6"], default="c6")
    ap.add_argument("--mode", choices=["exact", "bisect", "learned"], default=None)
    ap.add_argument("--alpha", type=float, default=0.9, help="learned: safety factor")
    ap.add_argument("--pairs", type=int, default=10_000_000, help="learned: training pairs")
    ap.add_argument("--epochs", type=int, default=3, help="learned: training epochs")
    ap.add_argument("--train-batch", type=int, default=4096, help="learned: rows per optimiser step")
    ap.add_argument("--lr", type=float, default=1.2e-2, help="learned: Adam step size")
    ap.add_argument("--save-net", default="", help="learned: save the trained parameters (.npy)")
    ap.add_argument("--load-net", default="",
                    help="learned: load the parameters (.npy) instead of harvesting and training")
    ap.add_argument("--train-seed", type=int, default=1, help="learned: network init and shuffle seed")
    ap.add_argument("--max-level", type=int, default=15, help="learned: finest occupancy-pyramid level")
    ap.add_argument("--compare", action="store_true", default=True,
                    help="also time bisect / exact / learned on the same layout (default)")
    ap.add_argument("--no-compare", dest="compare", action="store_false")
    ap.add_argument("--no-autotune", action="store_true", help="batched: default launch shape")
    ap.add_argument("--compare-walks", type=int, default=10_000_000)
    ap.add_argument("--walks", type=int, default=0, help="walks per step (0 = workload default)")
    ap.add_argument("--leaf-cap", type=int, default=4)
    ap.add_argument("--block", type=int, default=0)
    ap.add_argument("--cpu-seconds", type=float, default=12.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--replay-walks", type=int, default=10_000,
                    help="walks of job 0 replayed by the CPU oracle")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only copy 256 MiB three times (scales the FETCH_SIZE / WRITE_SIZE counters)")
    return ap.parse_args()


# ---------------------------------------------------------------- CPU baseline

def cpu_baseline(rects, bounds, master, seconds):
    """Reference CPU walk (the reference's Tree/RectHost code) on one host
    thread over walk ids disjoint from the GPU's, sized to ~`seconds`; beside
    it the oracle's restatement (reference tree, single-path query fixed) over
    all host cores with OpenMP, sized to ~seconds/2."""
    import numpy as np
    import oracle_ffi as O
    r = rects.astype(np.float64)
    b = tuple(float(v) for v in bounds)
    kind = "reference" if O.ref_available() else "port"
    first = 1 << 31
    t0 = time.perf_counter()
    ix = O.Index(r, b, O.REFTREE_DFS)
    port_build = time.perf_counter() - t0
    ref_build = None
    if kind == "reference":
        t0 = time.perf_counter()
        t = O.RefTree(r, b)
        ref_build = time.perf_counter() - t0

        def run(n, first_walk):
            t0 = time.perf_counter()
            _, _, hops, _ = t.walk(r[master], n, seed=1, first_walk=first_walk)
            return hops, time.perf_counter() - t0
    else:
        def run(n, first_walk):
            t0 = time.perf_counter()
            _, hops, _, _ = ix.walk(master, n, seed=1, first_walk=first_walk, nthreads=1)
            return hops, time.perf_counter() - t0

    def sized(fn, secs):
        n = 200
        hops, dt = fn(n, first)
        while dt < 0.5 and n < 10 ** 8:
            n *= 4
            hops, dt = fn(n, first)
        n_final = max(n, int(n * secs / max(dt, 1e-6)))
        hops, dt = fn(n_final, first + n)
        return n_final, hops, dt

    n_final, hops, dt = sized(run, seconds)
    cores = min(os.cpu_count() or 1, 16)

    def run_all(n, first_walk):
        t0 = time.perf_counter()
        _, h, _, _ = ix.walk(master, n, seed=1, first_walk=first_walk, nthreads=cores)
        return h, time.perf_counter() - t0

    n_all, hops_all, dt_all = sized(run_all, seconds / 2)
    return {
        "value": hops / dt,
        "unit": "walk steps/s",
        "cores": 1,
        "kind": kind,
        "sample": f"{n_final} walks of the same layout from rect {master} (walk ids from 2^31), "
                  f"{hops} hops in {dt:.1f} s, single host thread (the reference CPU path is "
                  f"serial), per-walk /dev/urandom re-seed of the reference excluded",
        "walks_per_s": n_final / dt,
        "tree_build_s": ref_build if ref_build is not None else port_build,
        "all_cores": {
            "value": hops_all / dt_all,
            "unit": "walk steps/s",
            "cores": cores,
            "kind": "port",
            "walks_per_s": n_all / dt_all,
            "sample": f"{n_all} walks, {hops_all} hops in {dt_all:.1f} s, OpenMP over walk ids "
                      f"(the oracle's restatement on the reference tree with its single-path "
                      f"query fixed)",
        },
    }


# ---------------------------------------------------------------- helpers

def make_layout(frw, workload, seed=1):
    n, box, smin, smax, cfrac, sigma, ncl, _ = WORKLOADS[workload]
    return frw.Layout.generate(n, box, smin, smax, seed=seed, cluster_frac=cfrac, sigma=sigma,
                               n_clusters=ncl)


def walks_for_1pct(f):
    """Walks for a 3-sigma half-width of 1% of the fraction f of walks that end
    anywhere but the master: 3 sqrt(f (1 - f) / N) <= 0.01 f."""
    return math.ceil(9e4 * (1.0 - f) / f) if f > 0 else float("inf")


def read_counters():
    """The counter figures of the committed rocprofv3 summary of this
    workload (one 'counters-json:' line), or None."""
    try:
        with open(COUNTERS_MD) as f:
            for line in f:
                if line.startswith("counters-json:"):
                    return json.loads(line[len("counters-json:"):])
    except (OSError, ValueError):
        pass
    return None


def time_walks(torch, launch, steps, warmup, sync_stream):
    """Runs warmup + timed launches on sync_stream; returns (wall seconds of
    the timed launches, per-launch ms from HIP events on that stream)."""
    with torch.cuda.stream(sync_stream):
        for w in range(warmup):
            launch(w)
    torch.cuda.synchronize()
    ev = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True))
          for _ in range(steps)]
    t0 = time.perf_counter()
    with torch.cuda.stream(sync_stream):
        for k in range(steps):
            ev[k][0].record(sync_stream)
            launch(warmup + k)
            ev[k][1].record(sync_stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    return wall, [a.elapsed_time(b) for a, b in ev]


# ---------------------------------------------------------------- batched (default)

def train_batched(frw, torch, trees, harvest, pyr_of, n_pairs, epochs, seed, net, batch=4096, lr=1.2e-2):
    """Trains `net` (one of the trees' proposal objects, whose weights every
    job uses) on pairs harvested (in walk order) from exact walks of every
    (tree, master) of `harvest`, walk ids from 3 * 2^30 (disjoint from the
    measured ones), features from each tree's own pyramid.  One batched
    harvest (two launches) serves every (tree, master); a job short of its
    share of pairs is harvested again with proportionally more walks."""
    import numpy as np
    t0 = time.perf_counter()
    n_h = len(harvest)
    per = max(1024, n_pairs // n_h)
    walks = [max(1000, per // 8)] * n_h
    cap = 3 * per * n_h
    pts = torch.empty((cap, 2), dtype=torch.float64, device="cuda")
    hh = torch.empty(cap, dtype=torch.float64, device="cuda")
    ds = torch.empty(cap, dtype=torch.float64, device="cuda")

    def harvest_all(walks):
        hb = frw.Batch(trees, [{"tree": i, "master": m, "n_walks": w, "first_walk": 3 << 30, "job": 0}
                               for (i, m), w in zip(harvest, walks)])
        n = hb.harvest(pts, hh, ds, seed=1)
        return n, np.concatenate([[0], np.cumsum(n)])

    n, offs = harvest_all(walks)
    if (n < per).any():
        walks = [w if k >= per else int(w * per / max(k, 1) * 1.2) + 1000 for w, k in zip(walks, n)]
        n, offs = harvest_all(walks)
    got_total = 0
    for k, (i, m) in enumerate(harvest):
        got = int(min(n[k], per))
        lo = int(offs[k])
        net.add_samples(pts[lo:lo + got], hh[lo:lo + got], ds[lo:lo + got], features_of=pyr_of[i])
        got_total += got
    del pts, hh, ds
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    loss = net.train(epochs=epochs, batch=batch, lr=lr, seed=seed)
    t2 = time.perf_counter()
    return net, {"pairs": got_total, "harvest_s": t1 - t0, "train_s": t2 - t1, "epochs": epochs,
                 "batch": batch, "lr": lr, "final_loss": loss, "seed": seed}


def bench_batched(args, frw, torch, np):
    """One persistent launch per step over many jobs, captured in a graph,
    launch shape from the autotuner.  c6 (default): the 1e6-rect layout with
    16 masters, walks per master U[5e5, 2e6].  c4: 64 layouts, rect counts
    log-uniform in [1e2, 1e5] (density of the 1e4 / 1e5-box layout, sides
    U[10,500]), master 0 each, walks per layout U[1e5, 1e6].  Learned mode:
    one network trained on pairs harvested from every layout (every master),
    each job on its own layout's pyramid."""
    rng = np.random.default_rng(1)
    trees, jobs, lays = [], [], []
    t0 = time.perf_counter()
    if args.workload == "c4":
        n_lay = 64
        sizes = np.exp(rng.uniform(np.log(1e2), np.log(1e5), n_lay)).astype(int)
        wl = rng.integers(100_000, 1_000_001, n_lay)
        for i, (n, w) in enumerate(zip(sizes, wl)):
            box = int(1000 * math.sqrt(n))
            lay = frw.Layout.generate(int(n), box, 10, 500, seed=100 + i)
            lays.append(lay)
            trees.append(frw.Tree(lay, leaf_cap=args.leaf_cap))
            jobs.append({"tree": i, "master": 0, "n_walks": int(w), "job": i})
        harvest = [(i, 0) for i in range(n_lay)]
        max_level = min(args.max_level, 13)
    else:
        lay = make_layout(frw, "c3")
        lays.append(lay)
        frw.Tree(lay, leaf_cap=args.leaf_cap)  # first build of the process: code-object loading
        trees.append(frw.Tree(lay, leaf_cap=args.leaf_cap))
        sizes = np.array([WORKLOADS["c3"][0]])
        masters = [0] + [int(m) for m in rng.choice(np.arange(1, WORKLOADS["c3"][0]), 15, replace=False)]
        wl = rng.integers(500_000, 2_000_001, len(masters))
        for i, (m, w) in enumerate(zip(masters, wl)):
            jobs.append({"tree": 0, "master": m, "n_walks": int(w), "job": i})
        harvest = [(0, m) for m in masters]
        max_level = args.max_level
    n_jobs = len(jobs)
    torch.cuda.synchronize()
    build_s = time.perf_counter() - t0
    build_ms = max(t.info().build_ms for t in trees)
    walks_per_step = int(wl.sum())
    log(f"{len(trees)} layout(s) in {build_s:.2f} s (largest device build {build_ms:.2f} ms); {n_jobs} jobs, "
        f"{walks_per_step} walks per step")
    batch = frw.Batch(trees, jobs)
    modes = ["bisect", "exact", "learned"] if args.compare else [args.mode or "learned"]
    net, train_info = None, None
    if "learned" in modes:
        t0 = time.perf_counter()
        pyrs = [frw.Mlp(t, seed=args.train_seed, max_level=max_level) for t in trees]
        torch.cuda.synchronize()
        pyr_s = time.perf_counter() - t0
        net = pyrs[int(np.argmax(sizes))]  # its weights serve every job, each job its own pyramid
        if args.load_net:
            net.set_params(np.load(args.load_net))
            train_info = loaded_net_info(args)
        else:
            net, train_info = train_batched(frw, torch, trees, harvest, pyrs, args.pairs, args.epochs,
                                            args.train_seed, net, args.train_batch, args.lr)
            if args.save_net:
                np.save(args.save_net, net.params())
        train_info["pyramid_s"] = pyr_s
        train_info["pyramid_max_level"] = max_level
        batch.set_proposal(pyrs)
        log(f"proposal: pyramids in {pyr_s:.2f} s, {train_info['pairs']} pairs harvested in "
            f"{train_info['harvest_s']:.2f} s, trained {train_info['epochs']} epochs in {train_info['train_s']:.2f} s, "
            f"loss {train_info['final_loss']}")
    counts = torch.zeros(batch.counts_size(), dtype=torch.int64, device="cuda")
    hops = torch.zeros(n_jobs, dtype=torch.int64, device="cuda")
    stats = torch.zeros(2, dtype=torch.int64, device="cuda")
    stream = torch.cuda.Stream()
    sh = stream.cuda_stream
    results = {}
    runs = list(modes)
    if "learned" in modes:
        runs.append("network_only")
    for m in runs:
        mode_id = {"exact": frw.MODE_EXACT, "bisect": frw.MODE_BISECT}.get(m, frw.MODE_LEARNED)
        alpha = 1e-9 if m == "network_only" else args.alpha
        if m == "network_only":
            shape = results["learned"]["autotuned"]
        elif args.no_autotune:
            shape = {"block": 0, "grid": 0}
        elif mode_id == frw.MODE_LEARNED:
            shape, rate = batch.autotune_learned(net, alpha=alpha, seed=1)
            log(f"autotune {m}: {shape} at {rate:.3e} steps/s")
        else:
            shape, rate = batch.autotune(seed=1, mode=mode_id)
            log(f"autotune {m}: {shape} at {rate:.3e} steps/s")

        def launch(step, mode_id=mode_id, shape=shape, alpha=alpha):
            if mode_id == frw.MODE_LEARNED:
                batch.walk_learned(net, counts, hops, stats, alpha=alpha, seed=1 + step,
                                   block=shape["block"], grid=shape["grid"], use_graph=True, stream=sh)
            else:
                batch.walk(counts, hops, seed=1 + step, mode=mode_id, block=shape["block"], grid=shape["grid"],
                           use_graph=True, stream=sh)

        steps = args.steps
        with torch.cuda.stream(stream):
            for w in range(args.warmup):
                launch(w)
        torch.cuda.synchronize()
        counts.zero_()
        hops.zero_()
        stats.zero_()
        torch.cuda.synchronize()
        wall, launch_ms = time_walks(torch, lambda k: launch(args.warmup + k), steps, 0, stream)
        batch.check()
        total_hops = int(hops.sum().item())
        n_walks = walks_per_step * steps
        c = counts.cpu().numpy()
        st = stats.cpu().numpy()
        per_job = []
        worst = 0.0  # steps until every job's fraction is known to 1% (3 sigma)
        for j in range(n_jobs):
            off = batch.counts_offset(j)
            nj = int(wl[j]) * steps
            fj = 1.0 - c[off + jobs[j]["master"]] / nj
            worst = max(worst, walks_for_1pct(fj) / wl[j])
            per_job.append(fj)
        results[m] = {"walk_steps_per_s": total_hops / wall, "walks_per_s": n_walks / wall,
                      "hops_per_walk": total_hops / n_walks, "wall": wall, "steps": steps,
                      "avg_launch_ms": sum(launch_ms) / len(launch_ms), "autotuned": shape,
                      "steps_for_all_jobs_1pct": worst, "time_to_all_1pct_s": worst * wall / steps,
                      "fractions": per_job, "counts": c, "hops_total": total_hops}
        if mode_id == frw.MODE_LEARNED:
            # stats = (accepted, rejected) proposals: hops that evaluated the network
            results[m]["accept_rate"] = int(st[0]) / max(1, total_hops)
            results[m]["proposals"] = int(st[0] + st[1])
            results[m]["proposal_accept_rate"] = int(st[0]) / max(1, int(st[0] + st[1]))
        log(f"{m}: {results[m]['walk_steps_per_s']:.3e} steps/s, {results[m]['walks_per_s']:.3e} walks/s, "
            f"{results[m]['hops_per_walk']:.2f} hops/walk, launch {results[m]['avg_launch_ms']:.2f} ms")
    head = "learned" if "learned" in results else modes[0]
    r = results[head]
    setup_s = build_ms / 1e3 + ((train_info["pyramid_s"] + train_info["harvest_s"] + train_info["train_s"])
                                if head == "learned" and train_info else 0.0)
    out = {
        "metric": "walk steps/s",
        "value": r["walk_steps_per_s"],
        "unit": "walk steps/s",
        "n_gpus": 1,
        "steps": r["steps"],
        "warmup": args.warmup,
        "ms_per_step": r["wall"] * 1e3 / r["steps"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": DESCR[args.workload], "layouts": len(trees), "jobs": n_jobs,
                   "rects": int(sizes.max()), "walks_per_step": walks_per_step, "mode": head,
                   "autotuned": r["autotuned"], "graph": True, "leaf_cap": args.leaf_cap,
                   "alpha": args.alpha if head == "learned" else None},
        "walks_per_s": r["walks_per_s"],
        "hops_per_walk": r["hops_per_walk"],
        "time_to_all_jobs_1pct_s": setup_s + r["time_to_all_1pct_s"],
        "time_to_all_jobs_1pct_note": "setup (warm tree build" + (", pyramid, harvest and training" if head ==
                                      "learned" else "") + ") + steps until every job's f is known to 1 % (3 "
                                      "sigma) x the step's wall time",
        "tree_build_ms": build_ms,
    }
    if train_info:
        out["learned"] = dict(train_info, alpha=args.alpha)
        tb = train_info["batch"]
        out["learned"]["train_tflops"] = (train_info["pairs"] // tb * tb * train_info["epochs"] *
                                          TRAIN_FLOP_PER_SAMPLE / max(train_info["train_s"], 1e-9) / 1e12)
        out["learned"]["train_note"] = ("algorithmic training FLOP/s (forward, weight and data gradients; "
                                        f"batches of {tb}, one optimiser step per batch: latency-bound)")
        if "learned" in results:
            out["learned"]["accept_rate"] = results["learned"]["accept_rate"]
            out["learned"]["proposal_accept_rate"] = results["learned"]["proposal_accept_rate"]
            out["learned"]["accept_note"] = ("accept_rate: accepted proposals per hop; proposal_accept_rate: "
                                             "accepted per network evaluation (hops that asked)")
    if args.compare:
        comp = {}
        for m in runs:
            q = results[m]
            comp[m] = {k: q[k] for k in ("walk_steps_per_s", "walks_per_s", "hops_per_walk", "avg_launch_ms",
                                         "autotuned", "steps_for_all_jobs_1pct", "time_to_all_1pct_s")}
            if "accept_rate" in q:
                comp[m]["accept_rate"] = q["accept_rate"]
                comp[m]["proposal_accept_rate"] = q["proposal_accept_rate"]
        comp["network_only"]["note"] = ("learned walks whose every proposal is rejected (alpha 1e-9): the "
                                        "network and its features run on every hop that asks (previous square at "
                                        "least the finest verification cell), then the exact step")
        comp["plain"] = "bisect: the reference's bisection square search, same batch, same launch form"
        for m in ("exact", "learned"):
            comp[f"{m}_vs_plain_steps_per_s"] = results[m]["walk_steps_per_s"] / results["bisect"]["walk_steps_per_s"]
            comp[f"{m}_vs_plain_walks_per_s"] = results[m]["walks_per_s"] / results["bisect"]["walks_per_s"]
        le = results["learned"]["walks_per_s"] / results["exact"]["walks_per_s"]
        comp["learned_vs_exact_walks_per_s"] = le
        comp["learned_vs_exact_time_to_1pct"] = (results["exact"]["time_to_all_1pct_s"] + build_ms / 1e3) / \
            out["time_to_all_jobs_1pct_s"]
        # per-hop cost bound: a learned walk costs at least its hops times the
        # network's own per-hop cost, and takes no fewer hops than the exact walk
        # per-hop cost bound: a learned walk pays at least the network's share
        # per hop and the exact step for every hop without an accepted
        # proposal, and takes no fewer hops than the exact walk
        t_exact = 1.0 / results["exact"]["walk_steps_per_s"]
        t_net = 1.0 / results["network_only"]["walk_steps_per_s"] - t_exact
        acc = results["learned"]["accept_rate"]
        per_hop = t_net + (1.0 - acc) * t_exact
        bound = 1.0 / (results["learned"]["hops_per_walk"] * per_hop) if per_hop > 0 else float("inf")
        need = t_exact * (results["exact"]["hops_per_walk"] / results["learned"]["hops_per_walk"] - (1.0 - acc))
        comp["bound"] = {
            "exact_hop_ps": t_exact * 1e12, "network_per_hop_ps": t_net * 1e12,
            "network_over_exact_hop": t_net / t_exact, "accept_rate": acc,
            "learned_walks_per_s_upper_bound": bound,
            "network_per_hop_ps_needed_to_beat_exact": need * 1e12,
            "statement": (f"learned loses to the exact search: with {acc:.1%} of its hops accepted it still pays the "
                          f"exact step on the others, and its network costs {t_net * 1e12:.0f} ps per hop against "
                          f"the {max(need, 0) * 1e12:.0f} ps that would break even, so the learned walk is capped at "
                          f"{bound / results['exact']['walks_per_s']:.2f}x the exact walks/s" if le < 1
                          else "learned beats the exact search"),
        }
        comp["verdict"] = (f"learned+batched reaches {comp['learned_vs_plain_steps_per_s']:.2f}x the plain FRW's "
                           f"walk steps/s ({comp['learned_vs_plain_walks_per_s']:.2f}x walks/s) and "
                           f"{le:.2f}x the exact search's walks/s")
        out["comparison"] = comp
        log(comp["verdict"])
        log(comp["bound"]["statement"])
    # per-conductor row of job 0 (master rect 0): the largest couplings
    # P_0j = hits on j / N with their 3-sigma half-widths, exact and learned
    if "exact" in results:
        ce = results["exact"]["counts"][:int(trees[jobs[0]["tree"]].info().n_rects) + 2]
        nr0 = len(ce) - 2
        top = np.argsort(-ce[:nr0])[:8]
        row = {"master": jobs[0]["master"], "walks_exact": int(wl[0]) * results["exact"]["steps"], "entries": []}
        n0 = int(wl[0]) * results["exact"]["steps"]
        cl = results["learned"]["counts"][:nr0 + 2] if "learned" in results else None
        n0l = int(wl[0]) * results["learned"]["steps"] if cl is not None else 0
        for j in list(top) + ["escaped"]:
            k = nr0 if j == "escaped" else int(j)
            p = ce[k] / n0
            e = {"rect": "escaped" if j == "escaped" else k, "p_exact": p,
                 "p_exact_3sigma": 3 * math.sqrt(max(p * (1 - p), 0) / n0)}
            if cl is not None:
                pl = cl[k] / n0l
                e["p_learned"] = pl
                e["p_learned_3sigma"] = 3 * math.sqrt(max(pl * (1 - pl), 0) / n0l)
            row["entries"].append(e)
        fe = results["exact"]["fractions"][0]
        row["f_exact"] = fe
        row["f_exact_rel_3sigma"] = 3 * math.sqrt(fe * (1 - fe) / n0) / fe if fe > 0 else None
        if cl is not None:
            fl = results["learned"]["fractions"][0]
            row["f_learned"] = fl
            row["f_learned_rel_err_vs_exact"] = abs(fl - fe) / fe
            row["f_rel_3sigma_combined"] = 3 * math.sqrt(fe * (1 - fe) / n0 + fl * (1 - fl) / n0l) / fe
        out["per_conductor_row_job0"] = row
    # the CPU oracle replays the first walks of job 0 (exact search, same
    # Philox streams as the first timed step): hit rect and hop count per walk
    if args.replay_walks > 0 and args.workload == "c6":
        import oracle_ffi as O
        t0 = time.perf_counter()
        lay0 = lays[0]
        tr0 = trees[0]
        nrp = min(args.replay_walks, int(wl[0]))
        nr = tr0.info().n_rects
        pc = torch.zeros(nr + 2, dtype=torch.int64, device="cuda")
        ph = torch.zeros(1, dtype=torch.int64, device="cuda")
        hit = torch.zeros(nrp, dtype=torch.int32, device="cuda")
        hp = torch.zeros(nrp, dtype=torch.int32, device="cuda")
        seed = 1 + args.warmup
        tr0.walk_async(jobs[0]["master"], nrp, pc, ph, hit, hp, seed=seed, job=0)
        torch.cuda.synchronize()
        ix = O.Index(lay0.rects().astype(np.float64), tuple(float(v) for v in lay0.bounds()), O.REFTREE_DFS)
        oc, _, ohit, ohp = ix.walk(jobs[0]["master"], nrp, seed=seed, job=0, per_walk=True, nthreads=16,
                                   max_hops=10 ** 7)
        same = int(np.sum((hit.cpu().numpy() == ohit) & (hp.cpu().numpy() == ohp)))
        fo = 1 - oc[jobs[0]["master"]] / nrp
        fg = 1 - int(pc[jobs[0]["master"]].item()) / nrp
        out["oracle_replay"] = {"walks": nrp, "identical_walks": same, "f_gpu": fg, "f_oracle": fo,
                                "rel_err_vs_oracle": abs(fg - fo) / fo if fo > 0 else 0.0,
                                "seconds": time.perf_counter() - t0,
                                "note": "job 0's first walks of the first timed step, exact search, against the "
                                        "FP64 CPU oracle (reference tree, single-path query fixed)"}
        log(f"oracle replay: {same}/{nrp} walks identical")
    # roofline of the dominant kernel (the learned walk kernel): MFMA FLOP/s
    # of the network's algorithmic work, live; HBM traffic and instruction
    # counts from the committed counter summary of the same workload
    cnt = read_counters() if args.workload == "c6" else None
    if "learned" in results:
        q = results["learned"]
        props_per_launch = q["proposals"] / q["steps"]
        achieved = props_per_launch * FLOP_PER_PROPOSAL / (q["avg_launch_ms"] * 1e-3) / 1e12
        roof = {"bound": "mfma", "achieved": achieved, "peak": BF16_PEAK_TFLOPS, "unit": "TFLOP/s",
                "frac": achieved / BF16_PEAK_TFLOPS, "traffic": None, "kernel": "walk_kernel<learned> (batched)",
                "flop_per_proposal": FLOP_PER_PROPOSAL, "proposals_per_launch": props_per_launch,
                "avg_launch_ms": q["avg_launch_ms"]}
        add_counter_roofs(roof, cnt, "learned", q, walks_per_step)
        out["roofline"] = roof
    for m in ("exact", "bisect"):
        if m not in results:
            continue
        q = results[m]
        # the search and replay run from the caches: instruction issue binds
        roof = {"bound": "issue", "kernel": f"walk_kernel<{m}> (batched)", "avg_launch_ms": q["avg_launch_ms"],
                "traffic": None}
        add_counter_roofs(roof, cnt, m, q, walks_per_step)
        out.setdefault("rooflines_other", {})[m] = roof
    if "roofline" not in out:
        out["roofline"] = out.get("rooflines_other", {}).get(head)
    if not args.no_cpu and args.workload == "c6":
        log("cpu baseline ...")
        cb = cpu_baseline(lays[0].rects(), lays[0].bounds(), jobs[0]["master"], args.cpu_seconds)
        f0 = results.get("exact", results[head])["fractions"][0]
        n1 = walks_for_1pct(f0)
        cb["time_to_1pct_s"] = cb["tree_build_s"] + n1 / cb["walks_per_s"]
        cb["all_cores"]["time_to_1pct_s"] = n1 / cb["all_cores"]["walks_per_s"]
        cb["time_to_1pct_note"] = f"job 0 (rect {jobs[0]['master']}) alone: host tree build + {n1} walks"
        out["cpu_baseline"] = cb
        log(f"cpu baseline ({cb['kind']}): {cb['value']:.3e} steps/s on 1 thread")
    else:
        out["cpu_baseline"] = None
    print(json.dumps(out), flush=True)


def add_counter_roofs(roof, cnt, mode, q, walks_per_step):
    """HBM traffic per launch (FETCH_SIZE and WRITE_SIZE of the mode's
    dispatch, each scaled by the factor a plain copy calibrates), the issue
    roof (VALU and SALU wave instructions per hop against one per CU per
    cycle) and the VALU lane utilisation, from the counter summary."""
    hops_per_s = q["walk_steps_per_s"]
    hops_per_launch = q["hops_total"] / q["steps"]
    if not cnt or mode not in cnt.get("kernels", {}):
        return
    k = cnt["kernels"][mode]
    f = cnt["copy_factor"]
    if "FETCH_SIZE_KiB" in k and "WRITE_SIZE_KiB" in k:
        traffic = (k["FETCH_SIZE_KiB"] * f["FETCH_SIZE"] + k["WRITE_SIZE_KiB"] * f["WRITE_SIZE"]) * 1024.0
        roof["traffic"] = traffic
        roof["traffic_unit"] = "bytes per launch (HBM, counters x copy factor)"
        roof["hbm_gbs"] = traffic / (q["avg_launch_ms"] * 1e-3) / 1e9
        roof["hbm_frac"] = roof["hbm_gbs"] / HBM_PEAK_GBS
        if roof.get("bound") == "hbm":
            roof.update(achieved=roof["hbm_gbs"], peak=HBM_PEAK_GBS, unit="GB/s", frac=roof["hbm_frac"])
    if "SQ_INSTS_VALU" in k and k.get("hops"):
        valu = k["SQ_INSTS_VALU"] / k["hops"]
        salu = k.get("SQ_INSTS_SALU", 0.0) / k["hops"]
        roof["issue"] = {"valu_per_hop": valu, "salu_per_hop": salu,
                         "valu_per_s": valu * hops_per_s, "salu_per_s": salu * hops_per_s,
                         "peak_per_s": ISSUE_PEAK, "frac": max(valu, salu) * hops_per_s / ISSUE_PEAK}
        if k.get("SQ_ACTIVE_INST_VALU"):
            roof["issue"]["valu_lane_utilisation"] = k["SQ_THREAD_CYCLES_VALU"] / (k["SQ_ACTIVE_INST_VALU"] * 64)
        if roof.get("bound") == "issue":
            roof.update(achieved=valu * hops_per_s, peak=ISSUE_PEAK, unit="VALU wave instructions/s",
                        frac=roof["issue"]["frac"])
    if "bytes_per_hop" in k:
        roof["algorithmic_bytes_per_hop"] = k["bytes_per_hop"]
        ab = k["bytes_per_hop"] * hops_per_launch / (q["avg_launch_ms"] * 1e-3) / 1e9
        roof["algorithmic_gbs"] = ab
        if roof.get("bound") == "hbm" and roof.get("traffic") is None:
            roof.update(achieved=ab, peak=HBM_PEAK_GBS, unit="GB/s", frac=ab / HBM_PEAK_GBS)
        if cnt.get("l2_read_gbs"):
            roof["algorithmic_frac_of_measured_l2_read"] = ab / cnt["l2_read_gbs"]
    if "SQ_INSTS_VALU_MFMA_MOPS_BF16" in k and k.get("hops"):
        executed = k["SQ_INSTS_VALU_MFMA_MOPS_BF16"] * 512 / k["hops"] * hops_per_s / 1e12
        roof["mfma_executed_tflops"] = executed
        roof["mfma_executed_frac"] = executed / BF16_PEAK_TFLOPS


# ---------------------------------------------------------------- single layout

def loaded_net_info(args):
    """Training record of a network loaded with --load-net (no harvest, no
    training in this run: a counter run profiles the walks alone)."""
    return {"pairs": 0, "harvest_s": 0.0, "train_s": 0.0, "epochs": 0, "batch": args.train_batch, "lr": args.lr,
            "final_loss": None, "seed": args.train_seed, "loaded_from": args.load_net}


def train_proposal(frw, torch, tree, master, n_pairs, epochs, seed=1, batch=4096, lr=1.2e-2):
    """Harvests (point, previous half-size, d*) pairs from exact walks (walk ids
    from 3 * 2^30, disjoint from the measured ones) and trains the network."""
    t0 = time.perf_counter()
    mlp = frw.Mlp(tree, seed=seed)
    pts = torch.empty((n_pairs, 2), dtype=torch.float64, device="cuda")
    hh = torch.empty(n_pairs, dtype=torch.float64, device="cuda")
    ds = torch.empty(n_pairs, dtype=torch.float64, device="cuda")
    walks = max(1000, n_pairs // 8)
    got = frw.walk_harvest(tree, master, walks, pts, hh, ds, seed=seed, first_walk=3 << 30)
    if got < n_pairs:  # more walks to fill the set
        more = int(walks * (n_pairs / max(got, 1) - 1) * 1.1) + 1000
        got = frw.walk_harvest(tree, master, walks + more, pts, hh, ds, seed=seed, first_walk=3 << 30)
    got = min(got, n_pairs)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    mlp.set_samples(pts[:got].contiguous(), hh[:got].contiguous(), ds[:got].contiguous())
    del pts, hh, ds
    loss = mlp.train(epochs=epochs, batch=batch, lr=lr, seed=seed)
    t2 = time.perf_counter()
    log(f"proposal: {got} pairs harvested in {t1 - t0:.2f} s, {epochs} epochs in {t2 - t1:.2f} s, "
        f"loss {loss:.4f}")
    return mlp, {"pairs": got, "harvest_s": t1 - t0, "train_s": t2 - t1, "epochs": epochs,
                 "batch": batch, "lr": lr, "final_loss": loss}


def walk_mode(frw, torch, tree, mlp, mode, master, walks, steps, warmup, stream, alpha, block=0,
              first_base=0):
    nr = tree.info().n_rects
    counts = torch.zeros(nr + 2, dtype=torch.int64, device="cuda")
    hops = torch.zeros(1, dtype=torch.int64, device="cuda")
    stats = torch.zeros(2, dtype=torch.int64, device="cuda")
    sh = stream.cuda_stream

    def launch(step):
        if mode == "learned":
            frw.walk_learned_async(tree, mlp, master, walks, counts, hops, stats, seed=1,
                                   first_walk=first_base + step * walks, alpha=alpha, block=block,
                                   stream=sh)
        else:
            m = frw.MODE_EXACT if mode == "exact" else frw.MODE_BISECT
            tree.walk_async(master, walks, counts, hops, seed=1, mode=m,
                            first_walk=first_base + step * walks, block=block, stream=sh)

    # counts/hops/stats cover the timed launches only; warmup launches take the
    # first walk ranges, timed ones the next
    with torch.cuda.stream(stream):
        for w in range(warmup):
            launch(w)
    torch.cuda.synchronize()
    counts.zero_()
    hops.zero_()
    stats.zero_()
    torch.cuda.synchronize()
    wall, launch_ms = time_walks(torch, lambda k: launch(warmup + k), steps, 0, stream)
    c = counts.cpu().numpy()
    st = stats.cpu().numpy()
    total_hops = int(hops.item())
    n_walks = walks * steps
    return {
        "wall": wall, "launch_ms": launch_ms, "hops": total_hops, "walks": n_walks,
        "counts": c, "accepted": int(st[0]), "rejected": int(st[1]),
        "steps_per_s": total_hops / wall, "walks_per_s": n_walks / wall,
    }


def fraction_other(c, master, n_walks):
    return 1.0 - c[master] / n_walks


def traffic_bytes(frw, tree, mlp, mode, master, walks, alpha, first):
    """Algorithmic bytes per hop from an instrumented run: 32 B per node record,
    16 B per rect record and 4 B per start-table entry the searches read;
    learned mode adds 16 B per pyramid probe and 64 B per feature window
    (8 rows x two words)."""
    if mode == "learned":
        nodes, rects, h, acc, fb, tabs = frw.walk_learned_traffic(tree, mlp, master, walks, seed=1,
                                                                  alpha=alpha, first_walk=first)
        b = 32.0 * nodes + 16.0 * rects + 4.0 * tabs + 16.0 * (acc + fb) + 64.0 * h
    else:
        m = frw.MODE_EXACT if mode == "exact" else frw.MODE_BISECT
        nodes, rects, h, tabs = tree.walk_traffic(master, walks, seed=1, mode=m, first_walk=first)
        b = 32.0 * nodes + 16.0 * rects + 4.0 * tabs
    return b / max(h, 1), nodes / max(h, 1), rects / max(h, 1)


def weighted_run(frw, torch, tree, master, nw, stream, flags, build_ms):
    """Weighted estimator over nw walks (one warm-up launch on disjoint walk
    ids first): the master's self term, its 3-sigma error and time to 1 %."""
    nr = tree.info().n_rects
    wc = torch.zeros(nr + 2, dtype=torch.int64, device="cuda")
    wh = torch.zeros(1, dtype=torch.int64, device="cuda")
    wsum = torch.zeros(nr + 2, dtype=torch.float64, device="cuda")
    wsq = torch.zeros(nr + 2, dtype=torch.float64, device="cuda")
    tree.walk_weighted_async(master, nw, wc, wh, wsum, wsq, seed=1, first_walk=1 << 31,
                             stream=stream.cuda_stream, flags=flags)  # warm-up launch
    torch.cuda.synchronize()
    for t_ in (wc, wh, wsum, wsq):
        t_.zero_()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    tree.walk_weighted_async(master, nw, wc, wh, wsum, wsq, seed=1, first_walk=0, stream=stream.cuda_stream,
                             flags=flags)
    torch.cuda.synchronize()
    wt = time.perf_counter() - t0
    ws, wq = wsum.cpu().numpy(), wsq.cpu().numpy()
    c_self = ws[master] / nw
    var = max(wq[master] / nw - c_self * c_self, 0.0)
    n1w = math.ceil(var * (300.0 / c_self) ** 2) if c_self > 0 else float("inf")
    return {
        "walks": nw,
        "walks_per_s": nw / wt,
        "walk_steps_per_s": int(wh.item()) / wt,
        "c_self": c_self,
        "c_self_rel_3sigma": 3.0 * math.sqrt(var / nw) / c_self if c_self > 0 else None,
        "var_per_walk": var,
        "c_escape": ws[nr] / nw,
        "sum_over_bins": float(ws.sum() / nw),
        "walks_for_1pct": n1w,
        "time_to_1pct_s": build_ms / 1e3 + n1w / (nw / wt),
        "units": "eps per unit length",
    }


def adaptive_runs(frw, tree, master, build_ms, mode):
    """Time to 1 % (3 sigma) measured by adaptive rounds: the hit fraction f
    in the bench's walk mode (exact unless learned was asked for) and the
    master's self-capacitance with the importance-sampled weighted walks."""
    res = {"goal_rel_3sigma": 0.01, "build_ms": build_ms}
    hit_mode = frw.MODE_BISECT if mode == "bisect" else frw.MODE_EXACT
    runs = [("fraction_other", dict(mode=hit_mode)),
            ("c_self", dict(mode=frw.MODE_WEIGHTED, flags=frw.WEIGHT_IMPORTANCE))]
    for name, kw in runs:
        tree.walk_adaptive(master, rel_err=0.01, first_round=65536, seed=9, first_walk=3 << 30,
                           max_walks=1 << 24, **kw)  # warm-up (code paths, allocations)
        r, _, _, _ = tree.walk_adaptive(master, rel_err=0.01, first_round=65536, seed=1,
                                        max_walks=(1 << 31) - 1, **kw)
        res[name] = {"value": r.value, "rel_err_3sigma": r.rel_err_3sigma, "converged": bool(r.converged),
                     "n_walks": r.n_walks, "rounds": r.rounds, "hops": r.hops, "wall_ms": r.wall_ms,
                     "time_to_1pct_s": build_ms / 1e3 + r.wall_ms / 1e3}
    return res


def refcompat_run(frw, torch, lay, master, stream):
    """The reference-compatible mode (the reference CPU tree and queries, one
    lane per walk): the same walks as the reference's own CPU code."""
    t0 = time.perf_counter()
    rt = frw.Tree(lay, ref_compat=True)
    build_s = time.perf_counter() - t0
    nr = rt.info().n_rects
    counts = torch.zeros(nr + 2, dtype=torch.int64, device="cuda")
    hops = torch.zeros(1, dtype=torch.int64, device="cuda")
    n = 1 << 14
    while True:
        counts.zero_()
        hops.zero_()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        rt.walk_async(master, n, counts, hops, seed=1, mode=frw.MODE_REFCOMPAT, first_walk=0,
                      stream=stream.cuda_stream)
        torch.cuda.synchronize()
        dt = time.perf_counter() - t0
        if dt > 0.5 or n >= 1 << 24:
            break
        n *= 4
    c = counts.cpu().numpy()
    return {"walks": n, "walks_per_s": n / dt, "walk_steps_per_s": int(hops.item()) / dt,
            "fraction_other": (n - int(c[master])) / n, "host_tree_build_s": build_s,
            "tree_nodes": rt.refcompat_info()[0],
            "note": "the reference CPU tree (insertion behaviour, single-path queries, literal bisection) on the "
                    "device; walk for walk the reference CPU FRW's walks for the same random stream"}


def bench_single(args, frw, torch, np):
    walks = args.walks or WORKLOADS[args.workload][7]
    master = 0
    mode = args.mode or ("learned" if args.workload == "c5" else "exact")
    lay = make_layout(frw, args.workload)
    # the first build of a process also pays one-time code-object loading;
    # time-to-1% uses the build time of a second (warm) build of the layout
    first_build_ms = frw.Tree(lay, leaf_cap=args.leaf_cap).info().build_ms
    tree = frw.Tree(lay, leaf_cap=args.leaf_cap)
    info = tree.info()
    nr = info.n_rects
    log(f"layout {nr} rects, tree {info.n_nodes} nodes depth {info.depth}, build {info.build_ms:.2f} ms")
    stream = torch.cuda.Stream()
    mlp, train_info = None, None
    if (mode == "learned" or args.compare) and args.load_net:
        mlp = frw.Mlp(tree, seed=args.train_seed)
        mlp.set_params(np.load(args.load_net))
        train_info = loaded_net_info(args)
    elif mode == "learned" or args.compare:
        mlp, train_info = train_proposal(frw, torch, tree, master, args.pairs, args.epochs,
                                         batch=args.train_batch, lr=args.lr)
        if args.save_net:
            np.save(args.save_net, mlp.params())
    r = walk_mode(frw, torch, tree, mlp, mode, master, walks, args.steps, args.warmup, stream,
                  args.alpha, block=args.block)
    log(f"timed {args.steps} steps of {walks} walks ({mode}) in {r['wall']:.3f} s")
    f = fraction_other(r["counts"], master, r["walks"])
    n1 = walks_for_1pct(f)
    setup_s = info.build_ms / 1e3 + (train_info["harvest_s"] + train_info["train_s"]
                                     if mode == "learned" else 0.0)
    t1pct = setup_s + n1 / r["walks_per_s"]
    sub = min(walks, 1_000_000)
    bph, nph, rph = traffic_bytes(frw, tree, mlp, mode, master, sub, args.alpha, first=args.warmup * walks)
    avg_ms = sum(r["launch_ms"]) / len(r["launch_ms"])
    bytes_per_launch = bph * r["hops"] / args.steps
    achieved = bytes_per_launch / (avg_ms * 1e-3) / 1e9
    out = {
        "metric": "walk steps/s", "value": r["steps_per_s"], "unit": "walk steps/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": r["wall"] * 1e3 / args.steps,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64", "data": "synthetic",
        "config": {"workload": DESCR[args.workload], "rects": nr, "walks_per_step": walks, "mode": mode,
                   "leaf_cap": info.leaf_cap, "tree_nodes": info.n_nodes, "tree_depth": info.depth},
        "walks_per_s": r["walks_per_s"], "hops_per_walk": r["hops"] / r["walks"], "fraction_other": f,
        "escaped": int(r["counts"][nr]) / r["walks"], "tree_build_ms": info.build_ms,
        "tree_build_first_ms": first_build_ms, "time_to_1pct_s": t1pct, "walks_for_1pct": n1,
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": None, "kernel": "walk_kernel",
                     "bytes_per_hop": bph, "node_records_per_hop": nph, "rect_records_per_hop": rph,
                     "avg_launch_ms": avg_ms},
    }
    l2_gbs = frw.probe_read_bandwidth(2 << 20, 20000)
    hbm_gbs = frw.probe_read_bandwidth(4 << 30, 4)
    out["roofline"]["measured_read_gbs"] = {"l2_resident_2MiB": l2_gbs, "hbm_4GiB": hbm_gbs}
    out["roofline"]["frac_of_measured_l2_read"] = achieved / l2_gbs
    nw = min(walks, 10_000_000)
    out["weighted"] = weighted_run(frw, torch, tree, master, nw, stream, 0, info.build_ms)
    out["weighted"]["importance"] = weighted_run(frw, torch, tree, master, nw, stream, frw.WEIGHT_IMPORTANCE,
                                                 info.build_ms)
    out["adaptive"] = adaptive_runs(frw, tree, master, info.build_ms, mode)
    if nr <= 200_000:
        out["refcompat"] = refcompat_run(frw, torch, lay, master, stream)
    if mode == "learned":
        out["learned"] = dict(train_info, alpha=args.alpha,
                              accept_rate=r["accepted"] / max(1, r["hops"]),
                              proposal_accept_rate=r["accepted"] / max(1, r["accepted"] + r["rejected"]))
    if args.compare:
        cw = min(walks, args.compare_walks)
        comp = {}
        for m in ("bisect", "exact", "learned"):
            q = walk_mode(frw, torch, tree, mlp, m, master, cw, 1, 1, stream, args.alpha,
                          block=args.block, first_base=2 << 30)
            fm = fraction_other(q["counts"], master, q["walks"])
            setup = info.build_ms / 1e3 + (train_info["harvest_s"] + train_info["train_s"]
                                           if m == "learned" else 0.0)
            comp[m] = {"walk_steps_per_s": q["steps_per_s"], "walks_per_s": q["walks_per_s"],
                       "hops_per_walk": q["hops"] / q["walks"], "fraction_other": fm,
                       "time_to_1pct_s": setup + walks_for_1pct(fm) / q["walks_per_s"]}
            if m == "learned":
                comp[m]["accept_rate"] = q["accepted"] / max(1, q["hops"])
            log(f"compare {m}: {q['steps_per_s']:.3e} steps/s, {q['walks_per_s']:.3e} walks/s")
        comp["walks"] = cw
        comp["plain"] = "bisect: the reference's bisection square search on the same device tree"
        for m in ("exact", "learned"):
            comp[f"{m}_vs_plain_walks_per_s"] = comp[m]["walks_per_s"] / comp["bisect"]["walks_per_s"]
            comp[f"{m}_vs_plain_steps_per_s"] = comp[m]["walk_steps_per_s"] / comp["bisect"]["walk_steps_per_s"]
        comp["learned_vs_exact_walks_per_s"] = comp["learned"]["walks_per_s"] / comp["exact"]["walks_per_s"]
        out["comparison"] = comp
    if not args.no_cpu:
        cb = cpu_baseline(lay.rects(), lay.bounds(), master, args.cpu_seconds)
        cb["time_to_1pct_s"] = cb["tree_build_s"] + n1 / cb.pop("walks_per_s")
        out["cpu_baseline"] = cb
        if "refcompat" in out and cb["kind"] == "reference":
            out["refcompat"]["vs_reference_cpu_steps_per_s"] = out["refcompat"]["walk_steps_per_s"] / cb["value"]
    else:
        out["cpu_baseline"] = None
    print(json.dumps(out), flush=True)


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
    if args.copy_calibration:
        a = torch.empty(64 * 1024 * 1024, dtype=torch.float32, device="cuda")
        b = torch.empty_like(a)
        a.fill_(1.0)
        for _ in range(3):
            b.copy_(a)
        torch.cuda.synchronize()
        log("copied 3 x 256 MiB")
        return
    if args.workload in ("c4", "c6"):
        return bench_batched(args, frw, torch, np)
    return bench_single(args, frw, torch, np)


if __name__ == "__main__":
    main()
