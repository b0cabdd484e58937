"""Reference-compatible GPU mode against the oracle's restatement of the
reference CPU tree (single-path rectangle query, insert split quirk,
first-containing-child point query, literal bisection).

The oracle's reference tree is itself checked bit for bit against the
reference's own sources compiled into oracle/_ref (test_oracle_vs_reference),
so equality here carries through to the reference CPU FRW:

* the rebuilt tree has the oracle's node count and object entries;
* Q1 / Q2 / Q3 on adversarial points and boxes are identical (Q3 to the bit);
* the Gaussian surface (factor search on the single-path query) is identical;
* every walk ends on the same rect after the same number of hops, and the
  command-line driver's CPU flag (getAvgPathLen) prints the oracle's fraction.
"""
import os
import re
import subprocess

import numpy as np
import pytest

import frw
import oracle_ffi as O
from conftest import FIXTURES, layout_path
from gpu_common import load_fixture, oracle_index, query_points, to_dev, torch_dev

pytestmark = pytest.mark.gpu

CLI = os.path.join(os.path.dirname(frw.__file__), "frw_cli")


def _boxes(pts, rng):
    """Boxes around the points: tiny, touching-size and large half-widths."""
    h = rng.choice([0.25, 1.0, 2.0, 7.5, 40.0, 300.0], len(pts))
    return np.ascontiguousarray(np.column_stack([pts[:, 0] - h, pts[:, 1] - h, pts[:, 0] + h, pts[:, 1] + h]))


def gpu_ref_walks(tree, master, n, seed=1, first=0, job=0, max_hops=0):
    torch = torch_dev()
    nr = tree.info().n_rects
    counts = torch.zeros(nr + 2, dtype=torch.int64, device="cuda")
    hops = torch.zeros(1, dtype=torch.int64, device="cuda")
    hit = torch.full((n,), -7, dtype=torch.int32, device="cuda")
    hp = torch.full((n,), -7, dtype=torch.int32, device="cuda")
    tree.walk_async(master, n, counts, hops, hit, hp, seed=seed, mode=frw.MODE_REFCOMPAT, first_walk=first,
                    job=job, max_hops=max_hops)
    torch.cuda.synchronize()
    return counts.cpu().numpy(), int(hops.item()), hit.cpu().numpy(), hp.cpu().numpy()


@pytest.mark.parametrize("name", [f for f, _ in FIXTURES])
def test_refcompat_tree_shape(name):
    lay, rects, bounds = load_fixture(name)
    tree = frw.Tree(lay, ref_compat=True)
    nodes, objs = tree.refcompat_info()
    st = oracle_index(rects, bounds, O.REFTREE).stats()
    assert nodes == st[0] and objs == st[2], (nodes, objs, st)


@pytest.mark.parametrize("name", [f for f, _ in FIXTURES])
def test_refcompat_queries(name):
    torch = torch_dev()
    lay, rects, bounds = load_fixture(name)
    tree = frw.Tree(lay, ref_compat=True)
    rng = np.random.default_rng(17)
    pts = query_points(rects, bounds, 3000, seed=3)
    boxes = _boxes(pts, rng)
    n = len(pts)
    q1 = torch.zeros(n, dtype=torch.uint8, device="cuda")
    q2 = torch.zeros(n, dtype=torch.int32, device="cuda")
    q3 = torch.zeros((n, 4), dtype=torch.float64, device="cuda")
    tree.refcompat_queries(to_dev(torch, pts), to_dev(torch, boxes), q1, q2, q3)
    ix = oracle_index(rects, bounds, O.REFTREE)
    oq1, oq2, oq3 = ix.queries_many(pts, boxes, n_q3=n)
    assert np.array_equal(q1.cpu().numpy(), oq1)
    assert np.array_equal(q2.cpu().numpy(), oq2)
    assert np.array_equal(q3.cpu().numpy(), oq3)
    # Q1 with a rect ignored by value (the Gaussian-surface search's form)
    ig = rects[len(rects) // 2]
    sel = slice(0, 400)
    tree.refcompat_queries(boxes=to_dev(torch, boxes[sel]), q1=q1[sel], ignore=ig)
    want = [ix.q1(tuple(b), ignore=tuple(float(v) for v in ig)) for b in boxes[sel]]
    assert np.array_equal(q1[sel].cpu().numpy(), np.array(want, dtype=np.uint8))
    # the single-path queries differ from the full search somewhere on the
    # larger fixtures; a mode that answered the full search would fail here
    if len(rects) >= 1000:
        dq1, dq2, _ = oracle_index(rects, bounds, O.REFTREE_DFS).queries_many(pts, boxes)
        assert (dq1 != oq1).any() or (dq2 != oq2).any()


@pytest.mark.parametrize("name,master", FIXTURES)
def test_refcompat_gauss_surface(name, master):
    lay, rects, bounds = load_fixture(name)
    tree = frw.Tree(lay, ref_compat=True)
    ix = oracle_index(rects, bounds, O.REFTREE)
    for m in sorted({master, 0, len(rects) // 2, len(rects) - 1}):
        assert tree.refcompat_gauss_surface(m) == ix.gauss_from(m), m
    # factor below one: the reference returns the rect itself
    assert tree.refcompat_gauss_surface(master, factor=0.5) == tuple(float(v) for v in rects[master])


@pytest.mark.parametrize("name,master", FIXTURES)
def test_refcompat_walks_replay_oracle(name, master):
    lay, rects, bounds = load_fixture(name)
    tree = frw.Tree(lay, ref_compat=True)
    n = 3000 if len(rects) < 5000 else 1000
    counts, hops, hit, hp = gpu_ref_walks(tree, master, n, seed=1)
    oc, ohops, ohit, ohp = oracle_index(rects, bounds, O.REFTREE).walk(master, n, seed=1, per_walk=True)
    assert np.array_equal(hit, ohit), np.nonzero(hit != ohit)[0][:10]
    assert np.array_equal(hp, ohp)
    assert hops == ohops
    assert np.array_equal(counts, oc)


def test_refcompat_walk_from_explicit_square():
    """Walks from the caller's start square equal the master's walks when the
    square is the master's Gaussian surface, and the oracle's walks from the
    same square otherwise."""
    torch = torch_dev()
    lay, rects, bounds = load_fixture("test100.txt")
    tree = frw.Tree(lay, ref_compat=True)
    ix = oracle_index(rects, bounds, O.REFTREE)
    nr, n = len(rects), 1500
    c0, h0, hit0, hp0 = gpu_ref_walks(tree, 3, n, seed=2)

    def from_square(sq):
        counts = torch.zeros(nr + 2, dtype=torch.int64, device="cuda")
        hops = torch.zeros(1, dtype=torch.int64, device="cuda")
        hit = torch.full((n,), -7, dtype=torch.int32, device="cuda")
        hp = torch.full((n,), -7, dtype=torch.int32, device="cuda")
        tree.refcompat_walk_from(sq, n, counts, hops, hit, hp, seed=2)
        torch.cuda.synchronize()
        return counts.cpu().numpy(), hit.cpu().numpy(), hp.cpu().numpy()

    c1, hit1, hp1 = from_square(tree.refcompat_gauss_surface(3))
    assert np.array_equal(hit1, hit0) and np.array_equal(hp1, hp0) and np.array_equal(c1, c0)
    # a square of the caller's: the oracle walk from it (its 'master' only
    # names the start surface, taken here from a rect with that surface)
    r = rects[7]
    sq = tuple(float(v) for v in (r[0] - 3, r[1] - 3, r[2] + 3, r[3] + 3))
    c2, hit2, hp2 = from_square(sq)
    T = O.green_thresholds()
    for i in range(0, n, 97):
        oh, ohops = ix.walk_one_from(sq, T, 2, 0, i)
        assert hit2[i] == oh and hp2[i] == ohops, i


def test_refcompat_seed_job_offset_and_max_hops():
    lay, rects, bounds = load_fixture("test100.txt")
    tree = frw.Tree(lay, ref_compat=True)
    ix = oracle_index(rects, bounds, O.REFTREE)
    for seed, job, first, mh in [(7, 0, 0, 0), (1, 5, 123456, 0), ((3 << 32) | 9, 2, 10, 3)]:
        c, h, hit, hp = gpu_ref_walks(tree, 3, 1500, seed=seed, first=first, job=job, max_hops=mh)
        oc, oh, ohit, ohp = ix.walk(3, 1500, seed=seed, job=job, first_walk=first,
                                    max_hops=mh if mh else 10 ** 6, per_walk=True)
        assert np.array_equal(hit, ohit) and np.array_equal(hp, ohp) and np.array_equal(c, oc)


def test_refcompat_sync_walk_and_errors():
    lay, rects, bounds = load_fixture("test")
    tree = frw.Tree(lay, ref_compat=True)
    counts, s = tree.walk(10, 2000, seed=4, mode=frw.MODE_REFCOMPAT)
    oc, oh, _, _ = oracle_index(rects, bounds, O.REFTREE).walk(10, 2000, seed=4)
    assert np.array_equal(counts, oc) and s.hops == oh
    # a tree built without the reference-compatible index refuses the mode
    plain = frw.Tree(lay)
    with pytest.raises(frw.FrwError):
        plain.walk(10, 100, mode=frw.MODE_REFCOMPAT)
    with pytest.raises(frw.FrwError):
        plain.refcompat_info()


def test_cli_cpu_flag_runs_reference_compatible_mode(tmp_path):
    """The driver without -G prints the oracle's reference-tree fraction."""
    if not os.path.exists(CLI):
        pytest.fail("frw_cli is not built")
    lay, rects, bounds = load_fixture("test")
    n, master = 4000, 10
    env = dict(os.environ, FRW_SEED="3")
    out = subprocess.run([CLI, "-S", layout_path("test"), "-O", str(master), "-I", str(n)], cwd=tmp_path,
                         env=env, capture_output=True, text=True, timeout=120)
    assert out.returncode == 0, out.stderr
    m = re.search(r"\[CPU\][^:]*: ([0-9.]+)%", out.stdout)
    assert m, out.stdout
    oc, _, _, _ = oracle_index(rects, bounds, O.REFTREE).walk(master, n, seed=3)
    assert abs(float(m.group(1)) - 100.0 * (n - oc[master]) / n) < 1e-5
