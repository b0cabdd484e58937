"""Device queries bit-exact against the oracle (brute force / full-DFS tree):
Q1 box collision, Q2 point location, nearest-rect distance, Q3 biggest square
in both modes (replayed bisection and literal bisection), Gaussian surface."""
import numpy as np
import pytest

import frw
import oracle_ffi as O
from conftest import FIXTURES
from gpu_common import load_fixture, oracle_index, query_points, to_dev, torch_dev

pytestmark = pytest.mark.gpu


def _gpu_queries(tree, pts, boxes):
    torch = torch_dev()
    dp = to_dev(torch, pts)
    db = to_dev(torch, boxes)
    n = len(pts)
    q1 = torch.zeros(len(boxes), dtype=torch.uint8, device="cuda")
    q2 = torch.zeros(n, dtype=torch.int32, device="cuda")
    dn = torch.zeros(n, dtype=torch.float64, device="cuda")
    q3e = torch.zeros((n, 4), dtype=torch.float64, device="cuda")
    q3b = torch.zeros((n, 4), dtype=torch.float64, device="cuda")
    tree.collide(db, q1)
    tree.point(dp, q2)
    tree.nearest(dp, dn)
    tree.maxsquare(dp, q3e, mode=frw.MODE_EXACT)
    tree.maxsquare(dp, q3b, mode=frw.MODE_BISECT)
    torch.cuda.synchronize()
    return (q1.cpu().numpy(), q2.cpu().numpy(), dn.cpu().numpy(), q3e.cpu().numpy(),
            q3b.cpu().numpy())


def _check_layout(lay, rects, bounds, n_random, seed, cap=16, n_q3=None):
    tree = frw.Tree(lay, leaf_cap=cap)
    pts = query_points(rects, bounds, n_random, seed)
    rng = np.random.default_rng(seed + 1)
    W = bounds[2] - bounds[0]
    h = rng.exponential(W * 0.01, len(pts))
    boxes = np.column_stack([pts[:, 0] - h, pts[:, 1] - h, pts[:, 0] + h, pts[:, 1] + h])
    q1, q2, dn, q3e, q3b = _gpu_queries(tree, pts, boxes)
    ex = oracle_index(rects, bounds, O.EXACT)
    df = oracle_index(rects, bounds, O.REFTREE_DFS)
    n_q3 = len(pts) if n_q3 is None else n_q3
    for i in range(len(pts)):
        px, py = pts[i]
        assert q1[i] == df.q1(tuple(boxes[i])), ("Q1", i, boxes[i])
        assert q2[i] == ex.q2(px, py), ("Q2", i, px, py)
        d, _ = ex.nearest(px, py)
        assert dn[i] == d, ("nearest", i, px, py, dn[i], d)
        if i < n_q3:
            want, _ = df.q3(px, py)
            assert tuple(q3b[i]) == want, ("Q3 bisect", i, px, py)
            assert tuple(q3e[i]) == want, ("Q3 exact", i, px, py)
    return tree


@pytest.mark.parametrize("name,master", FIXTURES)
def test_queries_fixtures(name, master):
    lay, rects, bounds = load_fixture(name)
    tree = _check_layout(lay, rects, bounds, 3000, seed=len(name), n_q3=1500)
    df = oracle_index(rects, bounds, O.REFTREE_DFS)
    for m in sorted({0, master, len(rects) - 1}):
        assert tree.gauss_surface(m) == df.gauss_from(m)


@pytest.mark.parametrize("cap", [1, 4, 64])
def test_queries_leaf_caps(cap):
    lay, rects, bounds = load_fixture("test1000.txt")
    _check_layout(lay, rects, bounds, 1500, seed=cap, cap=cap, n_q3=400)


def test_queries_generated_dense():
    """10^4 rects (box 10^5, sides U[10,500]) and a clustered layout."""
    for lay in (frw.Layout.generate(10_000, 100_000, 10, 500, seed=1),
                frw.Layout.generate(5_000, 400_000, 5, 300, seed=2, cluster_frac=0.5,
                                    sigma=8000, n_clusters=6)):
        rects = lay.rects()
        bounds = tuple(float(v) for v in lay.bounds())
        _check_layout(lay, rects, bounds, 4000, seed=7, n_q3=800)


def test_queries_touching_and_single():
    """Touching rects (shared edges and corners) and a one-rect layout."""
    rects = np.array([[0, 0, 10, 10], [10, 0, 20, 10], [0, 10, 10, 20], [20, 10, 30, 30],
                      [100, 100, 140, 120]], dtype=np.int32)
    lay = frw.Layout.from_rects(rects)
    _check_layout(lay, rects, tuple(float(v) for v in lay.bounds()), 500, seed=3)
    one = np.array([[5, 5, 17, 9]], dtype=np.int32)
    lay1 = frw.Layout.from_rects(one)
    _check_layout(lay1, one, tuple(float(v) for v in lay1.bounds()), 300, seed=4)


def test_queries_large_layout():
    """10^6 rects (box 10^7): sampled queries against the full-DFS tree."""
    lay = frw.Layout.generate(10 ** 6, 10 ** 7, 5, 300, seed=1, cluster_frac=0.2, sigma=1e5,
                              n_clusters=100)
    rects = lay.rects()
    bounds = tuple(float(v) for v in lay.bounds())
    tree = frw.Tree(lay, leaf_cap=16)
    pts = query_points(rects, bounds, 4000, seed=11)
    rng = np.random.default_rng(12)
    h = rng.exponential(2e4, len(pts))
    boxes = np.column_stack([pts[:, 0] - h, pts[:, 1] - h, pts[:, 0] + h, pts[:, 1] + h])
    q1, q2, dn, q3e, q3b = _gpu_queries(tree, pts, boxes)
    df = oracle_index(rects, bounds, O.REFTREE_DFS)
    for i in range(len(pts)):
        px, py = pts[i]
        assert q1[i] == df.q1(tuple(boxes[i]))
        k = df.q2(px, py)
        assert q2[i] == k
        if i < 600:
            want, _ = df.q3(px, py)
            assert tuple(q3b[i]) == want and tuple(q3e[i]) == want, (i, px, py)
