"""Walks on deep trees, where the walk kernel keeps its traversal stacks as
packed runs of siblings instead of single node indices.

The kernel takes the packed layout when the single-node stack (3 depth + 2
entries per lane) would hold fewer lanes per CU in LDS than its registers
allow: exact and bisection walks from 13 tree levels on, learned walks
from 14 (their activation tile already takes 11 KB per wave).  A layout of
dense clusters with one rect per leaf gives a tree of 18 levels.

* Exact and bisection walks on that tree replay the oracle walk for walk,
  and equal the walks on a shallow tree of the same layout (64-rect leaves,
  single-node stacks).
* Learned walks whose proposals can never pass (a safety factor of 1e-30
  makes every proposed square smaller than the 2-unit minimum) take the
  exact fallback on every hop, so they equal the exact walks walk for walk.
* A batched launch of several masters on the deep tree equals the single
  launches."""
import numpy as np
import pytest

import frw
import oracle_ffi as O
from gpu_common import oracle_index, torch_dev

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def deep():
    torch_dev()
    lay = frw.Layout.generate(20_000, 1_000_000, 5, 40, seed=9, cluster_frac=0.9, sigma=3e3, n_clusters=3)
    rects = lay.rects()
    bounds = tuple(float(v) for v in lay.bounds())
    tree = frw.Tree(lay, leaf_cap=1)
    assert tree.info().depth >= 14, tree.info().depth
    return lay, rects, bounds, tree


def _walks(tree, master, n, mode, seed=1, job=0):
    torch = torch_dev()
    nr = tree.info().n_rects
    counts = torch.zeros(nr + 2, dtype=torch.int64, device="cuda")
    hops = torch.zeros(1, dtype=torch.int64, device="cuda")
    hit = torch.full((n,), -7, dtype=torch.int32, device="cuda")
    hp = torch.full((n,), -7, dtype=torch.int32, device="cuda")
    tree.walk_async(master, n, counts, hops, hit, hp, seed=seed, mode=mode, job=job)
    torch.cuda.synchronize()
    return counts.cpu().numpy(), int(hops.item()), hit.cpu().numpy(), hp.cpu().numpy()


@pytest.mark.parametrize("mode", [frw.MODE_EXACT, frw.MODE_BISECT])
def test_deep_tree_walks_replay_oracle(deep, mode):
    lay, rects, bounds, tree = deep
    n = 4000 if mode == frw.MODE_EXACT else 1000
    counts, hops, hit, hp = _walks(tree, 0, n, mode)
    ix = oracle_index(rects, bounds, O.REFTREE_DFS)
    oc, ohops, ohit, ohp = ix.walk(0, n, seed=1, per_walk=True)
    assert np.array_equal(hit, ohit), np.nonzero(hit != ohit)[0][:10]
    assert np.array_equal(hp, ohp)
    assert hops == ohops
    assert np.array_equal(counts, oc)
    shallow = frw.Tree(lay, leaf_cap=64)
    assert shallow.info().depth < tree.info().depth
    c2, h2, hit2, hp2 = _walks(shallow, 0, n, mode)
    assert np.array_equal(hit2, hit) and np.array_equal(hp2, hp) and h2 == hops


def test_deep_tree_learned_fallback_equals_exact(deep):
    torch = torch_dev()
    lay, rects, bounds, tree = deep
    n = 4000
    net = frw.Mlp(tree, seed=1, max_level=13)
    nr = tree.info().n_rects
    counts = torch.zeros(nr + 2, dtype=torch.int64, device="cuda")
    hops = torch.zeros(1, dtype=torch.int64, device="cuda")
    hit = torch.full((n,), -7, dtype=torch.int32, device="cuda")
    hp = torch.full((n,), -7, dtype=torch.int32, device="cuda")
    acc, fb = frw.walk_learned(tree, net, 0, n, counts, hops, hit, hp, seed=1, alpha=1e-30)
    torch.cuda.synchronize()
    assert acc == 0 and 0 <= fb <= int(hops.item())  # every proposal rejected
    ec, eh, ehit, ehp = _walks(tree, 0, n, frw.MODE_EXACT)
    assert np.array_equal(hit.cpu().numpy(), ehit)
    assert np.array_equal(hp.cpu().numpy(), ehp)
    assert np.array_equal(counts.cpu().numpy(), ec)


def test_deep_tree_batch_equals_single(deep):
    torch = torch_dev()
    lay, rects, bounds, tree = deep
    masters = [0, 7, 1234, 19_999]
    n = 3000
    jobs = [{"tree": 0, "master": m, "n_walks": n, "job": i} for i, m in enumerate(masters)]
    batch = frw.Batch([tree], jobs)
    counts = torch.zeros(batch.counts_size(), dtype=torch.int64, device="cuda")
    hops = torch.zeros(len(jobs), dtype=torch.int64, device="cuda")
    batch.walk(counts, hops, seed=1, mode=frw.MODE_EXACT)
    torch.cuda.synchronize()
    batch.check()
    c = counts.cpu().numpy()
    nr = tree.info().n_rects
    for i, m in enumerate(masters):
        sc, sh, _, _ = _walks(tree, m, n, frw.MODE_EXACT, job=i)
        off = batch.counts_offset(i)
        assert np.array_equal(c[off:off + nr + 2], sc), i
        assert int(hops[i].item()) == sh, i
