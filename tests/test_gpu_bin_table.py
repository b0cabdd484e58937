"""Tallies of walks that end on a rect other than their job's master go
through a per-block table in LDS (open addressing on (job, bin), 8 probes,
flushed once per block; a key that finds no slot goes straight to its bin):
64 slots for single-job and bisection launches, 512 for batched exact and
learned launches, whose blocks walk many jobs in turn.  These tests give the
table more distinct bins per block than it has slots, so both of its paths
run, and check that every bin still holds
exactly the walks that ended there (per-walk outputs and the oracle's replay
are the reference), for several launch shapes, for the weighted estimator's
exact fixed-point sums, and for batched jobs whose keys differ only in the
job."""
import numpy as np
import pytest

import frw
import oracle_ffi as O
from gpu_common import oracle_index, torch_dev
from test_gpu_weighted import fixed_point_sums

pytestmark = pytest.mark.gpu

N_WALKS = 200_000
SHAPES = [(0, 0), (256, 8), (64, 3)]  # the default shape, then few blocks: many walks (and bins) per block


def _ringed_master():
    """Rect 0, a 2000-unit square master, ringed by three rings of 8-unit
    squares 24 units apart (200, 232 and 264 units from its edges): walks
    from its Gaussian surface end on the master or on one of hundreds of
    small rects."""
    r = [[10_000, 10_000, 12_000, 12_000]]
    for k, d in enumerate((200, 232, 264)):
        lo, hi = 10_000 - d, 12_000 + d
        for x in range(lo + 12 * (k % 2), hi - 8, 24):
            r.append([x, lo - 8, x + 8, lo])
            r.append([x, hi, x + 8, hi + 8])
        for y in range(lo + 16 + 12 * (k % 2), hi - 24, 24):
            r.append([lo - 8, y, lo, y + 8])
            r.append([hi, y, hi + 8, y + 8])
    return np.array(r, dtype=np.int32)


@pytest.fixture(scope="module")
def dense():
    lay = frw.Layout.from_rects(_ringed_master())
    tree = frw.Tree(lay)
    rects = lay.rects()
    bounds = tuple(float(v) for v in lay.bounds())
    return lay, tree, rects, bounds


def _bins(hit, nr):
    return np.where(hit >= 0, hit, np.where(hit == -1, nr, nr + 1))


def test_counts_equal_per_walk_outcomes_for_any_shape(dense):
    torch = torch_dev()
    lay, tree, rects, bounds = dense
    nr = len(rects)
    master = 0
    ref = None
    for block, grid in SHAPES:
        counts = torch.zeros(nr + 2, dtype=torch.int64, device="cuda")
        hops = torch.zeros(1, dtype=torch.int64, device="cuda")
        hit = torch.full((N_WALKS,), -7, dtype=torch.int32, device="cuda")
        tree.walk_async(master, N_WALKS, counts, hops, hit, None, seed=5, block=block, grid=grid)
        torch.cuda.synchronize()
        c, h = counts.cpu().numpy(), hit.cpu().numpy()
        assert (h != -7).all()
        assert np.array_equal(c, np.bincount(_bins(h, nr), minlength=nr + 2))
        if ref is None:
            ref = (c, h)
        else:
            assert np.array_equal(c, ref[0]) and np.array_equal(h, ref[1]), (block, grid)
    c, h = ref
    other = np.count_nonzero(c[:nr] > 0) - (1 if c[master] > 0 else 0)
    # far more distinct bins than the 64 slots one block holds
    assert other > 4 * 64, other
    # the walks themselves are the oracle's
    n_rep = 20_000
    oc, _, ohit, _ = oracle_index(rects, bounds, O.REFTREE_DFS).walk(master, n_rep, seed=5, per_walk=True)
    assert np.array_equal(h[:n_rep], ohit)


@pytest.mark.parametrize("block,grid", SHAPES[1:])
def test_weighted_sums_exact_with_a_full_table(dense, block, grid):
    torch = torch_dev()
    lay, tree, rects, bounds = dense
    nr = len(rects)
    counts = torch.zeros(nr + 2, dtype=torch.int64, device="cuda")
    hops = torch.zeros(1, dtype=torch.int64, device="cuda")
    wsum = torch.zeros(nr + 2, dtype=torch.float64, device="cuda")
    wsq = torch.zeros(nr + 2, dtype=torch.float64, device="cuda")
    w_out = torch.zeros(N_WALKS, dtype=torch.float64, device="cuda")
    hit = torch.zeros(N_WALKS, dtype=torch.int32, device="cuda")
    tree.walk_weighted_async(0, N_WALKS, counts, hops, wsum, wsq, w_out=w_out, hit_out=hit, seed=11,
                             block=block, grid=grid)
    torch.cuda.synchronize()
    h = hit.cpu().numpy()
    bins = _bins(h, nr)
    assert np.array_equal(counts.cpu().numpy(), np.bincount(bins, minlength=nr + 2))
    S, Q = fixed_point_sums(w_out.cpu().numpy(), bins, nr + 2)
    assert np.array_equal(wsum.cpu().numpy(), S) and np.array_equal(wsq.cpu().numpy(), Q)


def test_batched_jobs_on_one_layout_keep_their_bins_apart(dense):
    """Jobs on the same tree hit the same rects: their table keys differ
    only in the job, and each job's bins must equal its own single launch.
    The jobs' distinct (job, bin) keys outnumber the 512 slots of the
    batched table."""
    torch = torch_dev()
    lay, tree, rects, bounds = dense
    nr = len(rects)
    jobs = [{"tree": 0, "master": 0, "n_walks": 150_000, "job": 0},
            {"tree": 0, "master": 200, "n_walks": 120_000, "job": 1},
            {"tree": 0, "master": 0, "n_walks": 90_000, "job": 2, "first_walk": 150_000}]
    batch = frw.Batch([tree], jobs)
    counts = torch.zeros(batch.counts_size(), dtype=torch.int64, device="cuda")
    hops = torch.zeros(len(jobs), dtype=torch.int64, device="cuda")
    batch.walk(counts, hops, seed=3, block=128, grid=5)
    torch.cuda.synchronize()
    batch.check()
    c = counts.cpu().numpy()
    keys = sum(np.count_nonzero(c[batch.counts_offset(j):batch.counts_offset(j) + nr] > 0) for j in range(len(jobs)))
    assert keys > 512, keys
    for j, q in enumerate(jobs):
        c1 = torch.zeros(nr + 2, dtype=torch.int64, device="cuda")
        h1 = torch.zeros(1, dtype=torch.int64, device="cuda")
        tree.walk_async(q["master"], q["n_walks"], c1, h1, seed=3, job=q["job"], first_walk=q.get("first_walk", 0))
        torch.cuda.synchronize()
        off = batch.counts_offset(j)
        assert np.array_equal(c[off:off + nr + 2], c1.cpu().numpy()), j
        assert int(hops[j].item()) == int(h1.item()), j
