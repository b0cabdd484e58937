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


# ---------------------------------------------------------------- the GPU half
# randomWalkCudaInit -> randomWalkCudaWrapper -> countAvg -> freeQuadTreeManager,
# createQuadTree with the caller's spaceSize and doCheck, getAvgPathLenCUDA,
# runRandomWalk(GPU=true) and frw_cli -G (randomWalk/Cuda/mainkernels.h:9-15,
# randomWalk/Cuda/randomwalk.cu:64-134, randomWalk/Cuda/getAvgPathLenCUDA.cu:50-80),
# against the oracle's walks on the reference tree with its single-path query
# fixed (the walks of the exact-search engine).

CLI = os.path.join(os.path.dirname(os.path.dirname(__file__)), "engineer-thesis_amd", "frw", "frw_cli")


def run_raw(*args, cwd=None):
    if not os.path.exists(DRIVER):
        pytest.fail("tests/shim/shim_driver is not built (make -C tests/shim)")
    print("shim driver:", *args, flush=True)
    out = subprocess.run([DRIVER, *map(str, args)], capture_output=True, text=True, timeout=120,
                         env=dict(os.environ, FRW_SEED="1"), cwd=cwd)
    assert out.returncode == 0, out.stderr
    return out


def expected_flags(rects, ohit, master):
    """Per-walk flag (1: the hit rect differs by value from the master,
    escapes and stuck walks included) and rect ((-1,-1,-1,-1) if none)."""
    r = np.asarray(rects, dtype=np.float64)
    want_rect = np.where((ohit >= 0)[:, None], r[np.maximum(ohit, 0)], -1.0)
    flag = np.array([0 if h >= 0 and np.array_equal(r[h], r[master]) else 1 for h in ohit])
    return flag, want_rect


def parse_walks(stdout):
    lines = stdout.strip().splitlines()
    assert lines[-1].startswith("avg ")
    rows = np.array([list(map(float, ln.split())) for ln in lines[:-1]])
    return rows[:, 0].astype(int), rows[:, 1:5], float(lines[-1].split()[1])


@pytest.mark.parametrize("name,master,seed", [("test", 10, 1), ("test1000.txt", 0, 7)])
def test_cuda_wrapper_flags_and_rects_replay_oracle(name, master, seed):
    lay, rects, bounds = load_fixture(name)
    n = 1500
    out = run_raw("cudawalk", layout_path(name), master, n, seed)
    flag, rect, avg = parse_walks(out.stdout)
    _, _, ohit, _ = oracle_index(rects, bounds, O.REFTREE_DFS).walk(master, n, seed=seed, per_walk=True)
    wflag, wrect = expected_flags(rects, ohit, master)
    assert np.array_equal(flag, wflag)
    assert np.array_equal(rect, wrect)
    assert avg == flag.sum() / n  # countAvg
    assert "Stworzono pomyslnie" in out.stderr  # randomWalkCudaInit checks its tree


def test_create_quad_tree_takes_space_size_and_checks():
    name, master, seed, pad, n = "test100.txt", 0, 3, 700, 1000
    lay, rects, bounds = load_fixture(name)
    out = run_raw("cudatree", layout_path(name), master, n, seed, pad, 1)
    assert "Sprawdzanie drzewa." in out.stderr and "Stworzono pomyslnie" in out.stderr
    flag, rect, avg = parse_walks(out.stdout)
    grown = (bounds[0] - pad, bounds[1] - pad, bounds[2] + pad, bounds[3] + pad)
    _, _, ohit, _ = oracle_index(rects, grown, O.REFTREE_DFS).walk(master, n, seed=seed, per_walk=True)
    wflag, wrect = expected_flags(rects, ohit, master)
    assert np.array_equal(flag, wflag) and np.array_equal(rect, wrect)
    # the caller's bounds change the walks (fewer escapes than with the padded layer size)
    _, _, dhit, _ = oracle_index(rects, bounds, O.REFTREE_DFS).walk(master, n, seed=seed, per_walk=True)
    assert (ohit == -1).sum() < (dhit == -1).sum()
    # no check requested: nothing logged
    quiet = run_raw("cudatree", layout_path(name), master, 10, seed, pad, 0)
    assert "Sprawdzanie" not in quiet.stderr


def oracle_fraction(name, master, n, kind):
    lay, rects, bounds = load_fixture(name)
    _, _, ohit, _ = oracle_index(rects, bounds, kind).walk(master, n, seed=1, per_walk=True)
    flag, _ = expected_flags(rects, ohit, master)
    return flag.sum() / n


@pytest.mark.parametrize("gpu,kind", [(1, O.REFTREE_DFS), (0, O.REFTREE)])
def test_get_avg_path_len_matches_oracle(tmp_path, gpu, kind):
    name, master, n = "test", 10, 3000
    out = run_raw("avg", layout_path(name), master, n, gpu, cwd=tmp_path)
    got = float(out.stdout.strip().splitlines()[-1])
    assert got == oracle_fraction(name, master, n, kind)
    assert len(list(tmp_path.glob("*wyniki.txt"))) == 1  # the results file is written


def test_run_random_walk_gpu_and_cli(tmp_path):
    name, master, n = "test", 10, 2000
    f = oracle_fraction(name, master, n, O.REFTREE_DFS)
    want = f"[GPU] Ile sciezek trafiło do innego elementu: {f * 100:f}%"
    out = run_raw("run", layout_path(name), master, n, 1, cwd=tmp_path)
    assert want in out.stdout, out.stdout
    if not os.path.exists(CLI):
        pytest.fail("engineer-thesis_amd/frw/frw_cli is not built")
    print("frw_cli -G", flush=True)
    cli = subprocess.run([CLI, "-S", layout_path(name), "-I", str(n), "-O", str(master), "-G"],
                         capture_output=True, text=True, timeout=120, env=dict(os.environ, FRW_SEED="1"),
                         cwd=tmp_path)
    assert cli.returncode == 0, cli.stderr
    assert want in cli.stdout, cli.stdout
