"""The reference's own call sequence through the shim (Parser -> Tree ->
createTree -> RandomWalk and the Tree queries, include/randomwalk.h) against
the oracle's restatement of the reference CPU tree: the same hit rect and hop
count for every walk, the same Q1 / Q2 / Q3 answers and Gaussian surface."""
import os
import subprocess

import numpy as np
import pytest

import oracle_ffi as O
from conftest import layout_path
from gpu_common import load_fixture, oracle_index, query_points

pytestmark = pytest.mark.gpu

DRIVER = os.path.join(os.path.dirname(__file__), "shim", "shim_driver")


def run(*args):
    if not os.path.exists(DRIVER):
        pytest.fail("tests/shim/shim_driver is not built (make -C tests/shim)")
    out = subprocess.run([DRIVER, *map(str, args)], capture_output=True, text=True, timeout=120,
                         env=dict(os.environ, FRW_SEED="1"))
    assert out.returncode == 0, out.stderr
    return [list(map(float, line.split())) for line in out.stdout.splitlines()]


@pytest.mark.parametrize("name,master", [("test", 10), ("test100.txt", 0)])
def test_random_walk_calls_replay_oracle(name, master):
    lay, rects, bounds = load_fixture(name)
    n = 200
    rows = run("walk", layout_path(name), master, n)
    _, _, ohit, ohp = oracle_index(rects, bounds, O.REFTREE).walk(master, n, seed=1, per_walk=True)
    r = np.asarray(rects, dtype=np.float64)
    for i, row in enumerate(rows):
        assert int(row[0]) == i and int(row[1]) == ohp[i], (i, row, ohp[i])
        want = r[ohit[i]] if ohit[i] >= 0 else np.full(4, -1.0)
        assert np.array_equal(np.asarray(row[2:6]), want), (i, row, ohit[i])


def test_tree_queries_replay_oracle(tmp_path):
    lay, rects, bounds = load_fixture("test1000.txt")
    pts = query_points(rects, bounds, 150, seed=9)
    h = np.random.default_rng(2).choice([0.5, 1.0, 2.0, 10.0, 80.0], len(pts))
    f = tmp_path / "pts"
    f.write_text("\n".join(f"{x!r} {y!r} {hh!r}" for (x, y), hh in zip(pts.tolist(), h.tolist())))
    rows = run("query", layout_path("test1000.txt"), f)
    ix = oracle_index(rects, bounds, O.REFTREE)
    r = np.asarray(rects, dtype=np.float64)
    assert len(rows) == len(pts)
    for (x, y), hh, row in zip(pts, h, rows):
        assert int(row[0]) == ix.q1((x - hh, y - hh, x + hh, y + hh))
        hit = ix.q2(x, y)
        assert int(row[1]) == (hit >= 0)
        if hit >= 0:
            assert np.array_equal(np.asarray(row[2:6]), r[hit])
        sq, _ = ix.q3(x, y)
        assert tuple(row[6:10]) == tuple(sq)


def test_gaussian_surface_call():
    lay, rects, bounds = load_fixture("test")
    ix = oracle_index(rects, bounds, O.REFTREE)
    for m in (0, 10, 16):
        assert tuple(run("gauss", layout_path("test"), m, 1.5)[0]) == ix.gauss_from(m)
    assert tuple(run("gauss", layout_path("test"), 3, 0.5)[0]) == tuple(float(v) for v in rects[3])
