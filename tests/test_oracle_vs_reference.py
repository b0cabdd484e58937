"""Pins the C restatement (oracle/liboracle.so) against the reference's own CPU
sources compiled into oracle/_ref (randomWalk/CPU/{quadTree,rectHost,Parser}.cpp).

Bit-for-bit: the layout reader and bounds, getPointFromNindex, the
bug-compatible Tree's three queries and Gaussian surface, and whole walks.
Also: the full-DFS tree agrees with brute force on every query, and the
reference's single-path rect query (D1) really does miss rects.
"""
import numpy as np
import pytest

import oracle_ffi as O
from conftest import FIXTURES, layout_path

pytestmark = pytest.mark.skipif(
    not (O.ref_available() or __import__("os").path.isdir(O.REFERENCE_SRC)),
    reason="oracle/_ref not built and the reference tree is absent")


def _load(name):
    n, r, nl = O.parse(layout_path(name))
    assert n > 0
    return r, O.bounds(r)


@pytest.mark.parametrize("name", [f for f, _ in FIXTURES])
def test_parser_and_bounds_match_reference(oracle_lib, name):
    n, r, nl = O.parse(layout_path(name))
    rn, rr, rb, rnl = O.ref_parse(layout_path(name))
    assert (n, nl) == (rn, rnl)
    assert np.array_equal(r, rr)
    assert O.bounds(r) == tuple(rb)


def test_point_from_index_bit_exact(oracle_lib):
    rng = np.random.default_rng(11)
    squares = []
    for _ in range(300):
        x0 = rng.uniform(-1e6, 1e6)
        y0 = rng.uniform(-1e6, 1e6)
        w = rng.choice([rng.uniform(0.01, 5), rng.uniform(1, 1e4), float(rng.integers(1, 500))])
        h = rng.choice([w, rng.uniform(0.01, 1e4)])
        squares.append((x0, y0, x0 + w, y0 + h))
    squares += [(0, 0, 2, 2), (10, 10, 14, 14), (-3.5, 7.25, 96.5, 107.25), (0, 0, 0.1, 0.1)]
    for sq in squares:
        for k in range(200):
            assert O.point_from_index(sq, k) == O.ref_point_from_index(sq, k), (sq, k)


def _random_boxes(b, n, rng):
    W = b[2] - b[0]
    out = []
    for _ in range(n):
        px = rng.uniform(b[0], b[2])
        py = rng.uniform(b[1], b[3])
        h = rng.exponential(W * 0.02)
        out.append((px, py, h))
    return out


@pytest.mark.parametrize("name", ["test", "test1", "test2", "test100.txt", "test1000.txt"])
def test_reftree_matches_reference_tree(oracle_lib, name):
    r, b = _load(name)
    ref = O.RefTree(r, b)
    bt = O.Index(r, b, O.REFTREE)
    rng = np.random.default_rng(5)
    for i, (px, py, h) in enumerate(_random_boxes(b, 1500, rng)):
        box = (px - h, py - h, px + h, py + h)
        assert bt.q1(box) == ref.q1(box)
        hit, rect = ref.q2(px, py)
        k = bt.q2(px, py)
        assert hit == (k >= 0) and (not hit or tuple(r[k]) == rect)
        if i < 200:
            assert bt.q3(px, py)[0] == ref.q3(px, py)
    for m in range(0, len(r), max(1, len(r) // 25)):
        assert bt.gauss_from(m) == ref.gauss_from(r[m])
        # ignore-by-value: the master never collides with its own surface
        assert bt.q1(tuple(r[m]), ignore=tuple(r[m])) == ref.q1(tuple(r[m]), tuple(r[m]))


@pytest.mark.parametrize("name", ["test", "test100.txt", "test1000.txt", "test10000.txt"])
def test_dfs_tree_is_exact(oracle_lib, name):
    r, b = _load(name)
    ex = O.Index(r, b, O.EXACT)
    df = O.Index(r, b, O.REFTREE_DFS)
    bt = O.Index(r, b, O.REFTREE)
    rng = np.random.default_rng(9)
    d1 = 0
    for i, (px, py, h) in enumerate(_random_boxes(b, 1000, rng)):
        box = (px - h, py - h, px + h, py + h)
        e = ex.q1(box)
        assert df.q1(box) == e
        d1 += bt.q1(box) != e
        assert df.q2(px, py) == ex.q2(px, py)
        if i < 100:
            assert df.q3(px, py)[0] == ex.q3(px, py)[0]
    if name != "test1":
        assert d1 > 0   # the single-path query (quadTree.cpp:148-217) misses rects


@pytest.mark.parametrize("name,master,nw", [("test", 10, 3000), ("test1", 0, 2000),
                                            ("test2", 5, 2000), ("test100.txt", 0, 1500),
                                            ("test1000.txt", 0, 400), ("test10000.txt", 0, 150)])
def test_walks_replay_reference(oracle_lib, name, master, nw):
    """Same Philox stream: the restated walk equals the reference code walk, hop for hop."""
    r, b = _load(name)
    ref = O.RefTree(r, b)
    bt = O.Index(r, b, O.REFTREE)
    hit, hops, total, other = ref.walk(r[master], nw, seed=1)
    c, tot2, h2, hp2 = bt.walk(master, nw, seed=1, per_walk=True, nthreads=1)
    hv = np.array([tuple(r[k]) if k >= 0 else (-1, -1, -1, -1) for k in h2])
    assert np.array_equal(hv, hit)
    assert np.array_equal(hops, hp2)
    assert total == tot2
    assert other == nw - c[master]
