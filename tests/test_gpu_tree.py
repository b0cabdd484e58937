"""Device Morton build: structural invariants (the checks of the reference's
checkQuadTree, randomWalk/Cuda/createquadtree.cu:373-425, plus partition and
tightness) on the reference fixtures and on generated layouts up to 10^6 rects."""
import numpy as np
import pytest

import frw
from conftest import FIXTURES
from gpu_common import load_fixture, torch_dev

pytestmark = pytest.mark.gpu

INT_MAX, INT_MIN = 2 ** 31 - 1, -2 ** 31


def check_tree(lay, tree, leaf_cap):
    torch_dev()
    rects = lay.rects()
    info = tree.info()
    nodes, trects, ids = tree.export()
    n = len(rects)
    assert info.n_rects == n and info.n_nodes == len(nodes) and info.n_nodes % 4 == 0
    # rect array is a permutation of the layout
    assert np.array_equal(np.sort(ids), np.arange(n))
    assert np.array_equal(trects, rects[ids])
    box, fc, rb, re, lvl = nodes[:, 0:4], nodes[:, 4], nodes[:, 5], nodes[:, 6], nodes[:, 7]
    # every rect is owned by exactly one node
    owned = np.zeros(n, dtype=np.int64)
    for a, b in zip(rb, re):
        owned[a:b] += 1
    assert np.all(owned == 1)
    # subtree walk: node boxes are tight, children inside parents, leaves small
    K = info.morton_bits
    stack = [(0, None)]
    seen = 0
    while stack:
        ni, parent_box = stack.pop()
        seen += 1
        bx = box[ni]
        # subtree rects
        sub = [ni]
        rs = []
        while sub:
            q = sub.pop()
            rs.append(trects[rb[q]:re[q]])
            if fc[q] >= 0:
                sub.extend(range(fc[q], fc[q] + 4))
        rs = np.concatenate(rs) if rs else np.zeros((0, 4), np.int32)
        if len(rs) == 0:
            assert tuple(bx) == (INT_MAX, INT_MAX, INT_MIN, INT_MIN) and fc[ni] < 0
            continue
        assert tuple(bx) == (rs[:, 0].min(), rs[:, 1].min(), rs[:, 2].max(), rs[:, 3].max())
        if parent_box is not None:
            assert bx[0] >= parent_box[0] and bx[1] >= parent_box[1]
            assert bx[2] <= parent_box[2] and bx[3] <= parent_box[3]
        if fc[ni] >= 0:
            assert fc[ni] % 4 == 0
            for c in range(4):
                assert lvl[fc[ni] + c] == lvl[ni] + 1
                stack.append((fc[ni] + c, bx))
    assert seen == len(nodes) - 3   # the root's three empty sibling slots
    # Morton cell containment (loose quadtree): every rect's doubled centre
    # (x0 + x1, y0 + y1) lies in its node's cell; internal nodes own no rects.
    B = lay.bounds().astype(np.int64)
    size = 1 << K
    code_x = np.zeros(len(nodes), dtype=np.int64)
    code_y = np.zeros(len(nodes), dtype=np.int64)
    for ni in range(len(nodes)):
        if fc[ni] >= 0:
            half = size >> (lvl[ni] + 1)
            for c in range(4):
                code_x[fc[ni] + c] = code_x[ni] + (half if c & 1 else 0)
                code_y[fc[ni] + c] = code_y[ni] + (half if c & 2 else 0)
    for ni in range(len(nodes)):
        if fc[ni] >= 0:
            assert rb[ni] == re[ni], "internal node owns rects"
            continue
        if rb[ni] == re[ni]:
            continue
        cs = size >> lvl[ni]
        r = trects[rb[ni]:re[ni]].astype(np.int64)
        cx = r[:, 0] + r[:, 2] - 2 * B[0]
        cy = r[:, 1] + r[:, 3] - 2 * B[1]
        assert np.all(cx >= code_x[ni]) and np.all(cx < code_x[ni] + cs)
        assert np.all(cy >= code_y[ni]) and np.all(cy < code_y[ni] + cs)
        assert re[ni] - rb[ni] <= leaf_cap or lvl[ni] >= K
    return info


@pytest.mark.parametrize("name", [f for f, _ in FIXTURES])
@pytest.mark.parametrize("cap", [1, 16])
def test_tree_equals_host_build(name, cap):
    """The device build equals the sequential host restatement node for node."""
    from test_hostsim import HostTree
    lay, _, _ = load_fixture(name)
    torch_dev()
    g_nodes, g_rects, g_ids = frw.Tree(lay, leaf_cap=cap).export()
    h_nodes, h_rects, h_ids, info = HostTree(lay, cap).export()
    assert np.array_equal(g_nodes, h_nodes)
    assert np.array_equal(g_rects, h_rects) and np.array_equal(g_ids, h_ids)


@pytest.mark.parametrize("name", [f for f, _ in FIXTURES])
@pytest.mark.parametrize("cap", [4, 16, 64])
def test_tree_invariants_fixtures(name, cap):
    lay, _, _ = load_fixture(name)
    t = frw.Tree(lay, leaf_cap=cap)
    check_tree(lay, t, cap)


def test_tree_invariants_generated():
    lay = frw.Layout.generate(20000, 200_000, 10, 500, seed=3, cluster_frac=0.3, sigma=5000,
                              n_clusters=8)
    for cap in (8, 32):
        check_tree(lay, frw.Tree(lay, leaf_cap=cap), cap)


def test_tree_large_deterministic():
    """10^6 rects (box 10^7, sides U[5,300], 20% in blobs): two builds are
    identical and every rect is owned once."""
    torch_dev()
    lay = frw.Layout.generate(10 ** 6, 10 ** 7, 5, 300, seed=1, cluster_frac=0.2, sigma=1e5,
                              n_clusters=100)
    a = frw.Tree(lay, leaf_cap=16)
    b = frw.Tree(lay, leaf_cap=16)
    na, ra, ia = a.export()
    nb, rb_, ib = b.export()
    assert np.array_equal(na, nb) and np.array_equal(ia, ib)
    from test_hostsim import HostTree
    hn, hr, hi, _ = HostTree(lay, 16).export()
    assert np.array_equal(na, hn) and np.array_equal(ia, hi)
    owned = np.zeros(len(ia), dtype=np.int64)
    np.add.at(owned, np.concatenate([np.arange(s, e) for s, e in zip(na[:, 5], na[:, 6]) if e > s]), 1)
    assert np.all(owned == 1)
    assert a.info().depth <= 25
