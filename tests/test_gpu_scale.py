"""Parity at the 10^6-rect scale (box 10^7, sides U[5,300], 20% of the rects
in 100 Gaussian blobs of sigma 10^5; Morton leaves of at most 64 rects).

* The first 10^4 walks replay the oracle's walks on the reference tree with
  its single-path query fixed: same hit rect and hop count for every walk,
  so every query those walks made was answered identically.
* 10^6 Q1 and Q2 queries and 2x10^5 Q3 queries (random points plus rect
  corners, edges and the p+-1 / p+-2 shell) against the same oracle tree.
* Full size (10^8 walks): conservation, determinism and launch-shape
  independence of the tallies."""
import os

import numpy as np
import pytest

import frw
import oracle_ffi as O
from gpu_common import query_points, to_dev, torch_dev

pytestmark = pytest.mark.gpu

THREADS = min(16, os.cpu_count() or 1)


@pytest.fixture(scope="module")
def c3():
    torch_dev()
    lay = frw.Layout.generate(10 ** 6, 10 ** 7, 5, 300, seed=1, cluster_frac=0.2, sigma=1e5, n_clusters=100)
    rects = lay.rects()
    bounds = tuple(float(v) for v in lay.bounds())
    tree = frw.Tree(lay, leaf_cap=64)
    ix = O.Index(rects.astype(np.float64), bounds, O.REFTREE_DFS)
    return lay, rects, bounds, tree, ix


def test_first_10k_walks_replay_oracle(c3):
    torch = torch_dev()
    lay, rects, bounds, tree, ix = c3
    n = 10_000
    nr = len(rects)
    counts = torch.zeros(nr + 2, dtype=torch.int64, device="cuda")
    hops = torch.zeros(1, dtype=torch.int64, device="cuda")
    hit = torch.full((n,), -7, dtype=torch.int32, device="cuda")
    hp = torch.full((n,), -7, dtype=torch.int32, device="cuda")
    tree.walk_async(0, n, counts, hops, hit, hp, seed=1)
    torch.cuda.synchronize()
    oc, ohops, ohit, ohp = ix.walk(0, n, seed=1, per_walk=True, nthreads=THREADS)
    g_hit, g_hp = hit.cpu().numpy(), hp.cpu().numpy()
    bad = np.nonzero((g_hit != ohit) | (g_hp != ohp))[0]
    assert len(bad) == 0, (bad[:10], g_hit[bad[:5]], ohit[bad[:5]])
    assert int(hops.item()) == ohops
    assert np.array_equal(counts.cpu().numpy(), oc)


def test_million_queries(c3):
    torch = torch_dev()
    lay, rects, bounds, tree, ix = c3
    allp = query_points(rects, bounds, 400_000, seed=21)  # random + corners, edges, shell
    rng = np.random.default_rng(22)
    pts = np.ascontiguousarray(allp[rng.permutation(len(allp))[:1_000_000]])
    n = len(pts)
    assert n == 10 ** 6
    h = rng.exponential(2e4, n)
    boxes = np.column_stack([pts[:, 0] - h, pts[:, 1] - h, pts[:, 0] + h, pts[:, 1] + h])
    n_q3 = 200_000
    dp = to_dev(torch, pts)
    q1 = torch.zeros(n, dtype=torch.uint8, device="cuda")
    q2 = torch.zeros(n, dtype=torch.int32, device="cuda")
    q3 = torch.zeros((n_q3, 4), dtype=torch.float64, device="cuda")
    tree.collide(to_dev(torch, boxes), q1)
    tree.point(dp, q2)
    tree.maxsquare(to_dev(torch, pts[:n_q3]), q3, mode=frw.MODE_EXACT)
    torch.cuda.synchronize()
    o1, o2, o3 = ix.queries_many(pts, boxes, n_q3=n_q3, nthreads=THREADS)
    assert np.array_equal(q1.cpu().numpy(), o1), np.nonzero(q1.cpu().numpy() != o1)[0][:10]
    assert np.array_equal(q2.cpu().numpy(), o2), np.nonzero(q2.cpu().numpy() != o2)[0][:10]
    g3 = q3.cpu().numpy()
    bad = np.nonzero((g3 != o3).any(1))[0]
    assert len(bad) == 0, (bad[:10], pts[bad[:3]])
    assert o1.mean() > 0.05 and (o2 >= 0).mean() > 0.05  # both answers occur


def test_full_size_tallies(c3):
    torch = torch_dev()
    lay, rects, bounds, tree, ix = c3
    nr = len(rects)

    def run(n, first, block=0, grid=0):
        c = torch.zeros(nr + 2, dtype=torch.int64, device="cuda")
        h = torch.zeros(1, dtype=torch.int64, device="cuda")
        tree.walk_async(0, n, c, h, seed=1, first_walk=first, block=block, grid=grid)
        torch.cuda.synchronize()
        return c.cpu().numpy(), int(h.item())

    c, h = run(10 ** 8, 0)
    assert c.sum() == 10 ** 8 and h >= 10 ** 8
    f = 1.0 - c[0] / 1e8
    assert 0.0 < f < 1.0
    c2, h2 = run(10 ** 7, 10 ** 8)
    c3_, h3 = run(10 ** 7, 10 ** 8, block=128, grid=517)
    assert np.array_equal(c2, c3_) and h2 == h3
    # two disjoint walk ranges agree within 4 sigma on the headline fraction
    f2 = 1.0 - c2[0] / 1e7
    sigma = np.sqrt(f * (1 - f) * (1 / 1e8 + 1 / 1e7))
    assert abs(f - f2) < 4 * sigma
