"""The batched and learned configurations at their stated sizes.

C4 (batched): 64 layouts with rect counts log-uniform in [10^2, 10^5] (box
1000 sqrt(n), sides U[10,500]), master 0 each, walks per layout U[10^5, 10^6]
(about 3.5 x 10^7 in one launch): per-job conservation, batch = single launch
for 8 sampled jobs (the largest included), the first 10^3 walks of 3 jobs
replayed by the oracle (reference tree, single-path query fixed), graph
replay of the autotuned shape, and learned batched walks conserving per job.

C5 (learned): the 10^6-rect layout, 10^7 training pairs harvested from exact
walks, 10^8 learned walks: conservation, every accepted-type proposal of a
sample certified free by the exact collision query, a chi-square homogeneity
test of the hit distribution against 10^8 exact walks, and walks whose every
proposal is rejected equal the exact walks walk for walk (first 10^4)."""
import math
import os

import numpy as np
import pytest

import frw
import oracle_ffi as O
from gpu_common import torch_dev

pytestmark = pytest.mark.gpu

THREADS = min(16, os.cpu_count() or 1)


def _homogeneity_p(c1, c2, min_expected=20):
    from scipy.stats import chi2_contingency
    keep = (c1 + c2) >= 2 * min_expected
    table = np.vstack([np.append(c1[keep], c1[~keep].sum()), np.append(c2[keep], c2[~keep].sum())])
    table = table[:, table.sum(0) > 0]
    return chi2_contingency(table)[1]


# ---------------------------------------------------------------- C4

@pytest.fixture(scope="module")
def c4():
    torch_dev()
    rng = np.random.default_rng(1)
    sizes = np.exp(rng.uniform(np.log(1e2), np.log(1e5), 64)).astype(int)
    walks = rng.integers(100_000, 1_000_001, 64)
    lays, trees, jobs = [], [], []
    for i, (n, w) in enumerate(zip(sizes, walks)):
        lay = frw.Layout.generate(int(n), int(1000 * math.sqrt(n)), 10, 500, seed=100 + i)
        lays.append(lay)
        trees.append(frw.Tree(lay))
        jobs.append({"tree": i, "master": 0, "n_walks": int(w), "job": i})
    print(f"C4: {int(sizes.sum())} rects in 64 layouts ({sizes.min()}..{sizes.max()}), {int(walks.sum())} walks",
          flush=True)
    return sizes, walks, lays, trees, jobs, frw.Batch(trees, jobs)


def _batch_run(torch, batch, n_jobs, **kw):
    counts = torch.zeros(batch.counts_size(), dtype=torch.int64, device="cuda")
    hops = torch.zeros(n_jobs, dtype=torch.int64, device="cuda")
    batch.walk(counts, hops, seed=1, **kw)
    torch.cuda.synchronize()
    batch.check()
    return counts.cpu().numpy(), hops.cpu().numpy()


def test_c4_conservation_and_single_launch_equality(c4):
    torch = torch_dev()
    sizes, walks, lays, trees, jobs, batch = c4
    counts, hops = _batch_run(torch, batch, len(jobs))
    for j, q in enumerate(jobs):
        off = batch.counts_offset(j)
        nr = trees[j].info().n_rects
        assert counts[off:off + nr + 2].sum() == q["n_walks"], j
        assert hops[j] >= q["n_walks"] - counts[off + nr]  # every walk that did not escape hopped
    sample = sorted({int(np.argmax(sizes)), int(np.argmin(sizes)), 0, 9, 21, 33, 47, 63})
    assert len(sample) == 8
    for j in sample:
        q = jobs[j]
        nr = trees[j].info().n_rects
        c = torch.zeros(nr + 2, dtype=torch.int64, device="cuda")
        h = torch.zeros(1, dtype=torch.int64, device="cuda")
        trees[j].walk_async(0, q["n_walks"], c, h, seed=1, job=q["job"])
        torch.cuda.synchronize()
        off = batch.counts_offset(j)
        assert np.array_equal(counts[off:off + nr + 2], c.cpu().numpy()), j
        assert hops[j] == int(h.item()), j


def test_c4_first_walks_replay_oracle(c4):
    torch = torch_dev()
    sizes, walks, lays, trees, jobs, batch = c4
    n = 1000
    for j in (int(np.argmax(sizes)), 5, 40):
        lay, tree, q = lays[j], trees[j], jobs[j]
        nr = tree.info().n_rects
        c = torch.zeros(nr + 2, dtype=torch.int64, device="cuda")
        h = torch.zeros(1, dtype=torch.int64, device="cuda")
        hit = torch.zeros(n, dtype=torch.int32, device="cuda")
        hp = torch.zeros(n, dtype=torch.int32, device="cuda")
        tree.walk_async(0, n, c, h, hit, hp, seed=1, job=q["job"])
        torch.cuda.synchronize()
        ix = O.Index(lay.rects().astype(np.float64), tuple(float(v) for v in lay.bounds()), O.REFTREE_DFS)
        _, _, ohit, ohp = ix.walk(0, n, seed=1, job=q["job"], per_walk=True, nthreads=THREADS, max_hops=10 ** 7)
        assert np.array_equal(hit.cpu().numpy(), ohit), j
        assert np.array_equal(hp.cpu().numpy(), ohp), j


def test_c4_autotuned_graph_replays(c4):
    torch = torch_dev()
    sizes, walks, lays, trees, jobs, batch = c4
    ref_c, ref_h = _batch_run(torch, batch, len(jobs))
    shape, rate = batch.autotune(seed=1)
    assert rate > 0
    counts = torch.zeros(batch.counts_size(), dtype=torch.int64, device="cuda")
    hops = torch.zeros(len(jobs), dtype=torch.int64, device="cuda")
    for rep in range(2):
        counts.zero_()
        hops.zero_()
        torch.cuda.synchronize()
        batch.walk(counts, hops, seed=1, block=shape["block"], grid=shape["grid"], use_graph=True)
        torch.cuda.synchronize()
        assert np.array_equal(counts.cpu().numpy(), ref_c) and np.array_equal(hops.cpu().numpy(), ref_h), rep


def test_c4_learned_batched_conserves(c4):
    torch = torch_dev()
    sizes, walks, lays, trees, jobs, batch = c4
    pyrs = [frw.Mlp(t, seed=1, max_level=13) for t in trees]
    net = pyrs[int(np.argmax(sizes))]
    per = 20_000
    pts = torch.empty((per, 2), dtype=torch.float64, device="cuda")
    hh = torch.empty(per, dtype=torch.float64, device="cuda")
    ds = torch.empty(per, dtype=torch.float64, device="cuda")
    for i in range(0, 64, 4):
        got = min(per, frw.walk_harvest(trees[i], 0, 4000, pts, hh, ds, seed=1, first_walk=3 << 30))
        net.add_samples(pts[:got], hh[:got], ds[:got], features_of=pyrs[i])
    net.train(epochs=2, batch=1024, lr=2e-3, seed=1)
    batch.set_proposal(pyrs)
    counts = torch.zeros(batch.counts_size(), dtype=torch.int64, device="cuda")
    hops = torch.zeros(len(jobs), dtype=torch.int64, device="cuda")
    stats = torch.zeros(2, dtype=torch.int64, device="cuda")
    batch.walk_learned(net, counts, hops, stats, alpha=0.9, seed=1)
    torch.cuda.synchronize()
    batch.check()
    c, hp, st = counts.cpu().numpy(), hops.cpu().numpy(), stats.cpu().numpy()
    for j, q in enumerate(jobs):
        off = batch.counts_offset(j)
        nr = trees[j].info().n_rects
        assert c[off:off + nr + 2].sum() == q["n_walks"], j
    assert st[0] > 0 and st[0] + st[1] <= hp.sum()  # (accepted, rejected): at most one proposal a hop


# ---------------------------------------------------------------- C5

@pytest.fixture(scope="module")
def c5():
    torch = torch_dev()
    lay = frw.Layout.generate(10 ** 6, 10 ** 7, 5, 300, seed=1, cluster_frac=0.2, sigma=1e5, n_clusters=100)
    tree = frw.Tree(lay)
    mlp = frw.Mlp(tree, seed=1)
    n_pairs = 10_000_000
    pts = torch.empty((n_pairs, 2), dtype=torch.float64, device="cuda")
    hh = torch.empty(n_pairs, dtype=torch.float64, device="cuda")
    ds = torch.empty(n_pairs, dtype=torch.float64, device="cuda")
    got = frw.walk_harvest(tree, 0, 1_000_000, pts, hh, ds, seed=1, first_walk=3 << 30)
    assert got == n_pairs, got
    mlp.set_samples(pts, hh, ds)
    loss = mlp.train(epochs=3, batch=1024, lr=2e-3, seed=1)
    print(f"C5: {got} pairs, final loss {loss:.4f}", flush=True)
    return lay, tree, mlp, pts, hh, ds


def test_c5_learned_walks_conserve_and_match_exact(c5):
    torch = torch_dev()
    lay, tree, mlp, pts, hh, ds = c5
    n = 100_000_000
    nr = tree.info().n_rects
    cl = torch.zeros(nr + 2, dtype=torch.int64, device="cuda")
    hl = torch.zeros(1, dtype=torch.int64, device="cuda")
    st = torch.zeros(2, dtype=torch.int64, device="cuda")
    frw.walk_learned_async(tree, mlp, 0, n, cl, hl, st, seed=1, alpha=0.9)
    ce = torch.zeros(nr + 2, dtype=torch.int64, device="cuda")
    he = torch.zeros(1, dtype=torch.int64, device="cuda")
    tree.walk_async(0, n, ce, he, seed=2)
    torch.cuda.synchronize()
    c_l, c_e = cl.cpu().numpy(), ce.cpu().numpy()
    acc, fb = (int(v) for v in st.cpu().numpy())
    print(f"C5 learned: {int(hl.item()) / n:.2f} hops/walk, {acc / (acc + fb):.3f} of proposals and "
          f"{acc / int(hl.item()):.3f} of hops accepted; "
          f"exact {int(he.item()) / n:.2f} hops/walk", flush=True)
    assert c_l.sum() == n and c_e.sum() == n
    assert acc > 0 and acc + fb <= int(hl.item())
    assert c_l[nr + 1] == 0  # no walk reached the hop cap
    p = _homogeneity_p(c_l, c_e)
    assert p > 1e-4, p


def test_c5_proposals_certified_free_are_free(c5):
    """The proposal pipeline of the learned walk on harvested points: features,
    network, d = alpha cell 2^clamp(y, -4, 4), the pyramid probe; every square
    the probe certifies is free under the exact collision query."""
    torch = torch_dev()
    lay, tree, mlp, pts, hh, ds = c5
    m = 1_000_000
    P, H = pts[:m].contiguous(), hh[:m].contiguous()
    feats = torch.zeros((m, 72), dtype=torch.float32, device="cuda")
    mlp.features(P, H, feats)
    y = torch.zeros(m, dtype=torch.float32, device="cuda")
    mlp.forward(feats, y)
    K0 = mlp.levels()[0]
    # the feature level and cell of each point, as the walk computes them
    e = torch.where(H >= 1.0, (torch.frexp(H)[1] - 1).double(), torch.zeros_like(H))  # ilogb
    lvl = torch.clamp(K0 - e + 1, 0, mlp.levels()[1])
    cell = torch.pow(2.0, K0 - lvl)
    d = cell * torch.pow(2.0, torch.clamp(y.double(), -4.0, 4.0)) * 0.9
    ok = torch.zeros(m, dtype=torch.uint8, device="cuda")
    mlp.certify(P, d, ok)
    cert = ok.bool() & (d > 2.0)
    boxes = torch.stack([P[:, 0] - d, P[:, 1] - d, P[:, 0] + d, P[:, 1] + d], dim=1)[cert].contiguous()
    hit = torch.zeros(boxes.shape[0], dtype=torch.uint8, device="cuda")
    tree.collide(boxes, hit)
    torch.cuda.synchronize()
    n_cert = boxes.shape[0]
    print(f"C5 probe: {n_cert} of {m} proposals certified", flush=True)
    assert n_cert > 1000
    assert int(hit.sum().item()) == 0
    # and every certified square lies within the exact free distance
    assert bool((d[cert] < ds[:m][cert]).all())


def test_c5_rejected_proposals_walk_as_exact(c5):
    torch = torch_dev()
    lay, tree, mlp, pts, hh, ds = c5
    n = 10_000
    nr = tree.info().n_rects
    out = []
    for learned in (True, False):
        c = torch.zeros(nr + 2, dtype=torch.int64, device="cuda")
        h = torch.zeros(1, dtype=torch.int64, device="cuda")
        hit = torch.zeros(n, dtype=torch.int32, device="cuda")
        hp = torch.zeros(n, dtype=torch.int32, device="cuda")
        if learned:
            frw.walk_learned_async(tree, mlp, 0, n, c, h, None, hit, hp, seed=5, alpha=1e-9)
        else:
            tree.walk_async(0, n, c, h, hit, hp, seed=5)
        torch.cuda.synchronize()
        out.append((hit.cpu().numpy(), hp.cpu().numpy()))
    assert np.array_equal(out[0][0], out[1][0]) and np.array_equal(out[0][1], out[1][1])
