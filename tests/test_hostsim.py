"""CPU build of the library's device traversal code (tests/hostsim, compiled
from engineer-thesis_amd/csrc/frw_device.h) against the oracle: the same
__host__ __device__ functions the walk kernel runs, driven sequentially.
Catches logic errors in the replayed bisection, the nearest-rect search, the
point query and the Morton build without a GPU."""
import ctypes as C
import os
import subprocess

import numpy as np
import pytest

import frw
import oracle_ffi as O
from conftest import FIXTURES, TESTS, layout_path
from gpu_common import query_points

HS_DIR = os.path.join(TESTS, "hostsim")
HS_SO = os.path.join(HS_DIR, "libhostsim.so")


def hostsim():
    src = os.path.join(HS_DIR, "frw_hostsim.cpp")
    dev = os.path.join(os.path.dirname(TESTS), "engineer-thesis_amd", "csrc", "frw_device.h")
    if (not os.path.exists(HS_SO) or os.path.getmtime(HS_SO) < os.path.getmtime(src)
            or os.path.getmtime(HS_SO) < os.path.getmtime(dev)):
        subprocess.run(["make", "-s", "-B", "-C", HS_DIR], check=True)
    L = C.CDLL(HS_SO)
    vp = C.c_void_p
    L.hs_build.restype = vp
    L.hs_build.argtypes = [vp, C.c_int, vp, C.c_int]
    L.hs_free.argtypes = [vp]
    L.hs_info.argtypes = [vp, vp]
    L.hs_export.argtypes = [vp, vp, vp, vp]
    L.hs_queries.argtypes = [vp, vp, C.c_int, vp, vp, vp, vp, vp]
    L.hs_gauss.argtypes = [vp, C.c_int, vp]
    L.hs_walk.restype = C.c_int64
    L.hs_walk.argtypes = [vp, C.c_int, C.c_int, C.c_uint64, C.c_uint32, C.c_uint32, C.c_int64,
                          C.c_uint32, vp, vp, vp, vp]
    L.hs_set_stack.argtypes = [C.c_int]
    L.hs_set_lane_pad.argtypes = [C.c_int]
    L.hs_closed_form_count.restype = C.c_longlong
    L.hs_closed_form_direct_count.restype = C.c_longlong
    L.hs_walk_hinted.restype = C.c_int64
    L.hs_walk_hinted.argtypes = [vp, C.c_int, C.c_int, C.c_uint64, C.c_uint32, C.c_uint32, C.c_int64,
                                 C.c_uint32, vp, vp, vp, vp, C.c_int]
    return L


class HostTree:
    """runs: traverse with the packed sibling-run stack (RunStack) instead of
    the single-node one (PlainStack)."""

    def __init__(self, lay, cap=16, runs=0):
        self.L = hostsim()
        self.runs = int(runs)
        self.r = np.ascontiguousarray(lay.rects())
        self.b = np.ascontiguousarray(lay.bounds())
        self.h = self.L.hs_build(self.r.ctypes.data, len(self.r), self.b.ctypes.data, cap)

    def __del__(self):
        try:
            self.L.hs_free(self.h)
        except Exception:
            pass

    def export(self):
        info = np.zeros(3, np.int64)
        self.L.hs_info(self.h, info.ctypes.data)
        nodes = np.zeros((info[0], 8), np.int32)
        rects = np.zeros((len(self.r), 4), np.int32)
        ids = np.zeros(len(self.r), np.int32)
        self.L.hs_export(self.h, nodes.ctypes.data, rects.ctypes.data, ids.ctypes.data)
        return nodes, rects, ids, info

    def queries(self, pts):
        n = len(pts)
        pts = np.ascontiguousarray(pts, dtype=np.float64)
        out = (np.zeros(n), np.zeros(n, np.int32), np.zeros((n, 4)), np.zeros((n, 4)),
               np.zeros(n, np.int32))
        self.L.hs_set_stack(self.runs)
        self.L.hs_queries(self.h, pts.ctypes.data, n, *[a.ctypes.data for a in out])
        return out

    def walk(self, master, n, mode=0, seed=1, job=0, first=0, max_hops=10 ** 7, use_hint=False):
        """Per-walk hits and hops; use_hint: the exact mode seeds each search
        with the previous hop's nearest rect and hands that rect to the
        bisection replay, as the walk kernel does."""
        T = frw.green_thresholds()
        hit = np.zeros(n, np.int32)
        hp = np.zeros(n, np.int32)
        self.L.hs_set_stack(self.runs)
        tot = self.L.hs_walk_hinted(self.h, master, mode, seed, job, first, n, max_hops, T.ctypes.data,
                                    hit.ctypes.data, hp.ctypes.data, None, int(use_hint))
        assert tot >= 0, "bisection error flag"
        return hit, hp, tot

    def walk_counted(self, master, n, mode=0, seed=1, job=0, first=0, use_hint=True, max_hops=10 ** 7):
        """(node records read, rect records read, hops, start-table entries
        read) of the same walks, the exact mode seeding each search with the
        previous hop's nearest rect as the walk kernel does."""
        T = frw.green_thresholds()
        hit = np.zeros(n, np.int32)
        hp = np.zeros(n, np.int32)
        tr = np.zeros(3, np.int64)
        self.L.hs_set_stack(self.runs)
        tot = self.L.hs_walk_hinted(self.h, master, mode, seed, job, first, n, max_hops, T.ctypes.data,
                                    hit.ctypes.data, hp.ctypes.data, tr.ctypes.data, int(use_hint))
        assert tot >= 0, "bisection error flag"
        return int(tr[0]), int(tr[1]), int(tot), int(tr[2])


@pytest.mark.parametrize("name,master", FIXTURES)
@pytest.mark.parametrize("cap", [1, 16, 64])
@pytest.mark.parametrize("runs", [0, 1])
def test_hostsim_walks_replay_oracle(oracle_lib, name, master, cap, runs):
    lay = frw.Layout.load(layout_path(name))
    ht = HostTree(lay, cap, runs)
    ix = O.Index(lay.rects().astype(np.float64), tuple(float(v) for v in lay.bounds()),
                 O.REFTREE_DFS)
    _, tot_o, ohit, ohp = ix.walk(master, 1500, per_walk=True)
    for mode, hint in ((frw.MODE_EXACT, False), (frw.MODE_EXACT, True), (frw.MODE_BISECT, False)):
        hit, hp, tot = ht.walk(master, 1500, mode=mode, use_hint=hint)
        assert np.array_equal(hit, ohit) and np.array_equal(hp, ohp) and tot == tot_o


@pytest.mark.parametrize("name,master", FIXTURES)
@pytest.mark.parametrize("pad", [1, 3, 20])
def test_hostsim_walks_lockstep_replay(oracle_lib, name, master, pad):
    """The square replay's decided-steps loop runs in step with the other
    lanes of its wave: here every lane that stops goes on for `pad` more
    steps (adding 0), as beside lanes that still go; walks still replay the
    oracle (a lane's late steps change nothing)."""
    lay = frw.Layout.load(layout_path(name))
    ht = HostTree(lay, 8, 1)
    ix = O.Index(lay.rects().astype(np.float64), tuple(float(v) for v in lay.bounds()),
                 O.REFTREE_DFS)
    _, tot_o, ohit, ohp = ix.walk(master, 1500, per_walk=True)
    ht.L.hs_set_lane_pad(pad)
    try:
        hit, hp, tot = ht.walk(master, 1500, mode=frw.MODE_EXACT, use_hint=True)
    finally:
        ht.L.hs_set_lane_pad(0)
    assert np.array_equal(hit, ohit) and np.array_equal(hp, ohp) and tot == tot_o


@pytest.mark.parametrize("name,master", FIXTURES[:4])
@pytest.mark.parametrize("runs", [0, 1])
def test_hostsim_queries_and_gauss(oracle_lib, name, master, runs):
    lay = frw.Layout.load(layout_path(name))
    r = lay.rects()
    b = tuple(float(v) for v in lay.bounds())
    ht = HostTree(lay, 4, runs)
    pts = query_points(r, b, 1500, seed=2)
    dist, q2, q3e, q3b, nfb = ht.queries(pts)
    ex = O.Index(r.astype(np.float64), b, O.EXACT)
    assert np.array_equal(q3e, q3b)
    for i, (px, py) in enumerate(pts):
        assert dist[i] == ex.nearest(px, py)[0]
        assert q2[i] == ex.q2(px, py)
        if i % 5 == 0:
            assert tuple(q3e[i]) == ex.q3(px, py)[0]
    g = np.zeros(4)
    ht.L.hs_set_stack(runs)
    ht.L.hs_gauss(ht.h, master, g.ctypes.data)
    assert tuple(g) == ex.gauss_from(master)


def test_hostsim_large_layout(oracle_lib):
    """10^5 rects, clustered: replayed bisection == literal bisection on every
    query; walks replay the oracle."""
    lay = frw.Layout.generate(100_000, 2_000_000, 5, 300, seed=5, cluster_frac=0.2, sigma=4e4,
                              n_clusters=20)
    r = lay.rects()
    b = tuple(float(v) for v in lay.bounds())
    ht = HostTree(lay, 16)
    pts = query_points(r, b, 20000, seed=3)
    dist, q2, q3e, q3b, nfb = ht.queries(pts)
    assert np.array_equal(q3e, q3b)
    ix = O.Index(r.astype(np.float64), b, O.REFTREE_DFS)
    for i in range(0, len(pts), 11):
        assert q2[i] == ix.q2(*pts[i])
    hit, hp, tot = ht.walk(0, 3000)
    _, tot_o, ohit, ohp = ix.walk(0, 3000, per_walk=True)
    assert np.array_equal(hit, ohit) and np.array_equal(hp, ohp)


@pytest.mark.parametrize("box", [2_000_000, 1_999_999, 123_457, 10_000_000])
def test_hostsim_square_closed_form(box):
    """The replayed square search takes its decisions from the closed form
    of its decided steps (growth by dist0, then a binary search down to
    dist0 / 2^m) on most queries, and equals the literal bisection on every
    query, for boxes whose first step 0.05 W is a short binary fraction (h
    exact: the box is p -/+ L at once) and ones whose is not (h rounds in the
    loops: the decisions are the lattice's away from the band, and the
    steps are replayed on the box)."""
    lay = frw.Layout.generate(20_000, box, 5, max(10, box // 4000), seed=7, cluster_frac=0.3, sigma=box / 50,
                              n_clusters=10)
    r = lay.rects()
    b = tuple(float(v) for v in lay.bounds())
    ht = HostTree(lay, 8)
    rng = np.random.default_rng(11)
    pts = query_points(r, b, 20000, seed=4)
    # and points with long fractions, as walks produce (panel points of squares)
    f = rng.random((5000, 2))
    extra = np.column_stack([b[0] + f[:, 0] * (b[2] - b[0]), b[1] + f[:, 1] * (b[3] - b[1])])
    # and points just below the top of a binade, where p + h crosses it and
    # a running sum rounds (in a growth step, replayed as the loops add it)
    tops = [float(2 ** e) for e in range(10, 24) if b[0] < 2 ** e < b[2]]
    near = np.array([[t - rng.random() * 0.08 * (b[2] - b[0]), b[1] + rng.random() * (b[3] - b[1])]
                     for t in tops for _ in range(400)])
    pts = np.vstack([pts, extra, near, near[:, ::-1]])
    pts = pts[(pts[:, 1] > b[1]) & (pts[:, 1] < b[3]) & (pts[:, 0] > b[0]) & (pts[:, 0] < b[2])]
    c0, d0 = ht.L.hs_closed_form_count(), ht.L.hs_closed_form_direct_count()
    dist, q2, q3e, q3b, nfb = ht.queries(pts)
    used = ht.L.hs_closed_form_count() - c0
    direct = ht.L.hs_closed_form_direct_count() - d0
    assert np.array_equal(q3e, q3b)
    free = int(np.count_nonzero(dist > 2.0))
    print(f"\nbox {box}: closed form on {used} of {free} searches past the first two tests, {direct} without "
          f"replaying the halving steps")
    assert used > 0.9 * free
    if box != 1_999_999:  # h exact (10^7: the default layout's geometry, dist0 = 510000)
        assert direct > 0.95 * used


def test_hostsim_tree_invariants():
    lay = frw.Layout.load(layout_path("test10000.txt"))
    nodes, rects, ids, info = HostTree(lay, 16).export()
    assert np.array_equal(np.sort(ids), np.arange(len(ids)))
    owned = np.zeros(len(ids), np.int64)
    for s, e in zip(nodes[:, 5], nodes[:, 6]):
        owned[s:e] += 1
    assert np.all(owned == 1)
    assert len(nodes) % 4 == 0 and info[1] >= 2


@pytest.mark.parametrize("runs", [0, 1])
def test_hostsim_deep_tree_stack_bound(runs):
    """Deep trees (dense clusters, one rect per leaf): every traversal of the
    queries and of exact, hinted and bisection walks stays within the stack
    entries the kernels allocate for either stack layout (the host build
    checks each stack access and aborts on overflow), and both layouts visit
    the same nodes in the same order (equal traffic counts)."""
    lay = frw.Layout.generate(20_000, 1_000_000, 5, 40, seed=9, cluster_frac=0.9, sigma=3e3,
                              n_clusters=3)
    r = lay.rects()
    b = tuple(float(v) for v in lay.bounds())
    ht = HostTree(lay, 1, runs)
    _, _, _, info = ht.export()
    assert info[1] >= 14, info
    pts = query_points(r, b, 4000, seed=4)
    dist, q2, q3e, q3b, nfb = ht.queries(pts)
    assert np.array_equal(q3e, q3b)
    ex = O.Index(r.astype(np.float64), b, O.EXACT)
    for i in range(0, len(pts), 7):
        assert dist[i] == ex.nearest(*pts[i])[0]
        assert q2[i] == ex.q2(*pts[i])
    hinted = ht.walk_counted(0, 400)
    plain = ht.walk_counted(0, 400, use_hint=False)
    assert hinted[2] == plain[2]
    bis = ht.walk_counted(0, 400, mode=frw.MODE_BISECT)
    hit_b, hp_b, tot_b = ht.walk(0, 400, mode=frw.MODE_BISECT)
    hit_e, hp_e, tot_e = ht.walk(0, 400, mode=frw.MODE_EXACT)
    assert np.array_equal(hit_b, hit_e) and tot_b == tot_e == hinted[2]
    other = HostTree(lay, 1, 1 - runs)
    assert other.walk_counted(0, 400) == hinted
    assert other.walk_counted(0, 400, mode=frw.MODE_BISECT) == bis
