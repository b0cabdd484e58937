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
