"""Batched walks: one persistent launch over many (layout, master, walks)
jobs gives, job by job, exactly the tallies of the single-layout walks with the
same Philox streams; the captured graph replays identically; the autotuned
launch shape changes nothing but speed."""
import numpy as np
import pytest

import frw
from conftest import FIXTURES
from gpu_common import load_fixture, torch_dev

pytestmark = pytest.mark.gpu


def _single(tree, master, n, job, first, mode=frw.MODE_EXACT):
    torch = torch_dev()
    nr = tree.info().n_rects
    c = torch.zeros(nr + 2, dtype=torch.int64, device="cuda")
    h = torch.zeros(1, dtype=torch.int64, device="cuda")
    tree.walk_async(master, n, c, h, seed=1, mode=mode, job=job, first_walk=first)
    torch.cuda.synchronize()
    return c.cpu().numpy(), int(h.item())


def _fixture_batch():
    trees, jobs = [], []
    for ti, (name, master) in enumerate(FIXTURES):
        lay, rects, _ = load_fixture(name)
        trees.append(frw.Tree(lay, leaf_cap=8 if ti % 2 else 16))
        for k, m in enumerate(sorted({0, master, len(rects) // 2})):
            jobs.append({"tree": ti, "master": m, "n_walks": 3000 + 1000 * k,
                         "first_walk": 500 * k, "job": 10 * ti + k})
    return trees, jobs


def _run(batch, mode=frw.MODE_EXACT, **kw):
    torch = torch_dev()
    counts = torch.zeros(batch.counts_size(), dtype=torch.int64, device="cuda")
    hops = torch.zeros(batch.n_jobs, dtype=torch.int64, device="cuda")
    batch.walk(counts, hops, seed=1, mode=mode, **kw)
    torch.cuda.synchronize()
    batch.check()
    return counts.cpu().numpy(), hops.cpu().numpy()


@pytest.mark.parametrize("mode", [frw.MODE_EXACT, frw.MODE_BISECT])
def test_batch_equals_single_walks(mode):
    trees, jobs = _fixture_batch()
    batch = frw.Batch(trees, jobs)
    counts, hops = _run(batch, mode)
    for j, q in enumerate(jobs):
        off = batch.counts_offset(j)
        nr = trees[q["tree"]].info().n_rects
        c1, h1 = _single(trees[q["tree"]], q["master"], q["n_walks"], q["job"], q["first_walk"], mode)
        assert np.array_equal(counts[off:off + nr + 2], c1), j
        assert hops[j] == h1, j


def test_batch_graph_replay_and_autotune():
    torch = torch_dev()
    trees, jobs = _fixture_batch()
    batch = frw.Batch(trees, jobs)
    ref_counts, ref_hops = _run(batch)
    counts = torch.zeros(batch.counts_size(), dtype=torch.int64, device="cuda")
    hops = torch.zeros(batch.n_jobs, dtype=torch.int64, device="cuda")
    for rep in range(3):
        counts.zero_()
        hops.zero_()
        torch.cuda.synchronize()
        batch.walk(counts, hops, seed=1, use_graph=True)
        torch.cuda.synchronize()
        assert np.array_equal(counts.cpu().numpy(), ref_counts), rep
        assert np.array_equal(hops.cpu().numpy(), ref_hops), rep
    best, rate = batch.autotune()
    assert best["block"] in (64, 128, 192, 256) and best["grid"] > 0 and rate > 0
    c2, h2 = _run(batch, block=best["block"], grid=best["grid"])
    assert np.array_equal(c2, ref_counts) and np.array_equal(h2, ref_hops)


def test_batch_many_layouts():
    """64 generated layouts (rect counts log-uniform in [10^2, 10^4]) with
    uneven walk counts; per-job conservation and agreement with single walks
    on a sample of jobs."""
    rng = np.random.default_rng(4)
    trees, jobs = [], []
    for k in range(64):
        n = int(10 ** rng.uniform(2, 4))
        box = int(1e5 * np.sqrt(n / 1e4))
        lay = frw.Layout.generate(n, max(box, 3000), 10, 500 if box > 20000 else 100, seed=100 + k)
        trees.append(frw.Tree(lay))
        jobs.append({"tree": k, "master": 0, "n_walks": int(rng.integers(2000, 20000)),
                     "first_walk": 0, "job": k})
    batch = frw.Batch(trees, jobs)
    counts, hops = _run(batch)
    for j, q in enumerate(jobs):
        off = batch.counts_offset(j)
        nr = trees[j].info().n_rects
        assert counts[off:off + nr + 2].sum() == q["n_walks"]
    for j in (0, 17, 63):
        q = jobs[j]
        off = batch.counts_offset(j)
        nr = trees[j].info().n_rects
        c1, h1 = _single(trees[j], 0, q["n_walks"], q["job"], 0)
        assert np.array_equal(counts[off:off + nr + 2], c1) and hops[j] == h1


def _shared_params(n_params, seed=3):
    rng = np.random.default_rng(seed)
    P = rng.normal(0.0, 0.3, n_params).astype(np.float32)
    P[: 64 * 72].reshape(64, 72)[:, 67:] = 0.0
    return P


def test_batch_learned_equals_single_learned():
    """A batched learned launch (one network, each job on its own layout's
    occupancy pyramid) gives, job by job, the tallies of the single-layout
    learned walks with the same network and Philox streams; graph replays
    repeat them.  The network is random, so walks are capped at 2000 hops
    (a poor proposal makes walks long, never wrong)."""
    torch = torch_dev()
    trees, jobs = _fixture_batch()
    batch = frw.Batch(trees, jobs)
    pyrs = [frw.Mlp(t, seed=1) for t in trees]
    P = _shared_params(pyrs[0].n_params)
    for m in pyrs:
        m.set_params(P)
    net = pyrs[0]
    batch.set_proposal(pyrs)
    counts = torch.zeros(batch.counts_size(), dtype=torch.int64, device="cuda")
    hops = torch.zeros(batch.n_jobs, dtype=torch.int64, device="cuda")
    stats = torch.zeros(2, dtype=torch.int64, device="cuda")
    batch.walk_learned(net, counts, hops, stats, alpha=0.9, seed=1, max_hops=2000)
    torch.cuda.synchronize()
    batch.check()
    ref_c, ref_h, ref_s = counts.cpu().numpy(), hops.cpu().numpy(), stats.cpu().numpy()
    # (accepted, rejected) proposals: at most one per hop
    assert ref_s[0] > 0 and ref_s[0] + ref_s[1] <= ref_h.sum()
    acc_sum = 0
    for j, q in enumerate(jobs):
        t = trees[q["tree"]]
        nr = t.info().n_rects
        c1 = torch.zeros(nr + 2, dtype=torch.int64, device="cuda")
        h1 = torch.zeros(1, dtype=torch.int64, device="cuda")
        acc, fb = frw.walk_learned(t, pyrs[q["tree"]], q["master"], q["n_walks"], c1, h1, seed=1,
                                   first_walk=q["first_walk"], job=q["job"], alpha=0.9, max_hops=2000)
        off = batch.counts_offset(j)
        assert np.array_equal(ref_c[off:off + nr + 2], c1.cpu().numpy()), j
        assert ref_h[j] == int(h1.item()), j
        acc_sum += acc
    assert acc_sum == ref_s[0]
    for rep in range(3):
        counts.zero_()
        hops.zero_()
        stats.zero_()
        torch.cuda.synchronize()
        batch.walk_learned(net, counts, hops, stats, alpha=0.9, seed=1, use_graph=True, max_hops=2000)
        torch.cuda.synchronize()
        assert np.array_equal(counts.cpu().numpy(), ref_c), rep
        assert np.array_equal(stats.cpu().numpy(), ref_s), rep
    print(f"\nbatched learned (random network): {ref_h.sum()} hops, accepted {ref_s[0]}, "
          f"capped walks {sum(ref_c[batch.counts_offset(j) + trees[q['tree']].info().n_rects + 1] for j, q in enumerate(jobs))}")
    best, rate = batch.autotune_learned(net, alpha=0.9, max_hops=500)
    # learned blocks of 8 and 12 waves hold one LDS copy of the weights
    assert best["block"] in (64, 128, 192, 256, 512, 768) and best["grid"] > 0 and rate > 0
    # any launch shape, the tuned one and a 12-wave block included, gives the same walks
    for shape in (best, {"block": 768, "grid": 64}, {"block": 64, "grid": 300}):
        counts.zero_()
        hops.zero_()
        stats.zero_()
        batch.walk_learned(net, counts, hops, stats, alpha=0.9, seed=1, block=shape["block"], grid=shape["grid"],
                           max_hops=2000)
        torch.cuda.synchronize()
        assert np.array_equal(counts.cpu().numpy(), ref_c), shape
        assert np.array_equal(hops.cpu().numpy(), ref_h), shape
        assert np.array_equal(stats.cpu().numpy(), ref_s), shape


def test_batch_learned_needs_pyramids():
    torch = torch_dev()
    trees, jobs = _fixture_batch()
    batch = frw.Batch(trees, jobs)
    net = frw.Mlp(trees[0], seed=1)
    counts = torch.zeros(batch.counts_size(), dtype=torch.int64, device="cuda")
    hops = torch.zeros(batch.n_jobs, dtype=torch.int64, device="cuda")
    with pytest.raises(frw.FrwError):
        batch.walk_learned(net, counts, hops)
    with pytest.raises(frw.FrwError):  # a network built for another tree
        batch.set_proposal([net] * len(trees))


@pytest.mark.parametrize("mode", [frw.MODE_EXACT, frw.MODE_BISECT])
def test_batch_beyond_lds_job_table(mode):
    """200 jobs: the walk kernel keeps the walk offsets and the tallies of
    the first 128 jobs in LDS and reads / adds those of the others in global
    memory.  Every job, on both sides of that line, equals its single
    launch."""
    trees = []
    for name, master in FIXTURES[:4]:
        lay, rects, _ = load_fixture(name)
        trees.append(frw.Tree(lay))
    rng = np.random.default_rng(7)
    jobs = []
    for k in range(200):
        ti = k % len(trees)
        nr = trees[ti].info().n_rects
        jobs.append({"tree": ti, "master": int(rng.integers(0, nr)), "n_walks": int(rng.integers(200, 1500)),
                     "first_walk": int(rng.integers(0, 1 << 20)), "job": k})
    batch = frw.Batch(trees, jobs)
    counts, hops = _run(batch, mode)
    for j, q in enumerate(jobs):
        tree = trees[q["tree"]]
        nr = tree.info().n_rects
        off = batch.counts_offset(j)
        c1, h1 = _single(tree, q["master"], q["n_walks"], q["job"], q["first_walk"], mode)
        assert np.array_equal(counts[off:off + nr + 2], c1), j
        assert hops[j] == h1, j


def test_batch_harvest_equals_single_harvests():
    """One batched harvest (one launch per pass for every job) gives, job
    after job, the records of each job's own walk_harvest; a capacity below
    the total keeps a prefix of them; the batch's later walks are unchanged
    (its job table is uploaded again)."""
    torch = torch_dev()
    trees, jobs = _fixture_batch()
    batch = frw.Batch(trees, jobs)
    ref_counts, ref_hops = _run(batch)
    cap = 2_000_000

    def bufs(n):
        return (torch.zeros((n, 2), dtype=torch.float64, device="cuda"),
                torch.zeros(n, dtype=torch.float64, device="cuda"),
                torch.zeros(n, dtype=torch.float64, device="cuda"))

    pts, hh, ds = bufs(cap)
    n = batch.harvest(pts, hh, ds, seed=5)
    total = int(n.sum())
    assert 0 < total < cap and (n > 0).all()
    p1, h1, d1 = bufs(cap)
    off = 0
    for j, q in enumerate(jobs):
        got = frw.walk_harvest(trees[q["tree"]], q["master"], q["n_walks"], p1, h1, d1, seed=5,
                               first_walk=q["first_walk"], job=q["job"])
        assert got == n[j], j
        assert torch.equal(pts[off:off + got], p1[:got]), j
        assert torch.equal(hh[off:off + got], h1[:got]) and torch.equal(ds[off:off + got], d1[:got]), j
        off += got
    small = total // 3
    p2, h2, d2 = bufs(small)
    n2 = batch.harvest(p2, h2, d2, seed=5)
    assert int(n2.sum()) == small
    assert np.array_equal(np.cumsum(n2)[-1], small) and (n2 <= n).all()
    assert torch.equal(p2, pts[:small]) and torch.equal(h2, hh[:small]) and torch.equal(d2, ds[:small])
    counts, hops = _run(batch)
    assert np.array_equal(counts, ref_counts) and np.array_equal(hops, ref_hops)
