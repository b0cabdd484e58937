"""Weighted (physical) FRW estimator on the GPU against the CPU oracle.

* walk for walk: the same start point, first square, panel and weight, then
  the same exact walk (hit rect and hop count) as the oracle's restatement;
* the per-conductor weight sums are exact fixed-point sums of the per-walk
  weights (each rounded once to a multiple of 2^-32): bit-identical for any
  launch shape, equal to the same sums taken on the host, and equal to the
  oracle's sums within the oracle's own FP64 rounding;
* physics: a parallel-plate pair couples with about -length/gap (plus
  fringe), the self term is positive, and all weights (escapes included)
  average to zero within the Monte Carlo error, at 10^6 rects too."""
import numpy as np
import pytest

import frw
import oracle_ffi as O
from conftest import FIXTURES
from gpu_common import load_fixture, oracle_index, torch_dev

pytestmark = pytest.mark.gpu


def gpu_weighted(tree, master, n, seed=1, first=0, job=0, max_hops=0, eps=1.0, flags=0):
    torch = torch_dev()
    nr = tree.info().n_rects
    counts = torch.zeros(nr + 2, dtype=torch.int64, device="cuda")
    hops = torch.zeros(1, dtype=torch.int64, device="cuda")
    wsum = torch.zeros(nr + 2, dtype=torch.float64, device="cuda")
    wsq = torch.zeros(nr + 2, dtype=torch.float64, device="cuda")
    w_out = torch.zeros(n, dtype=torch.float64, device="cuda")
    hit = torch.full((n,), -7, dtype=torch.int32, device="cuda")
    hp = torch.full((n,), -7, dtype=torch.int32, device="cuda")
    tree.walk_weighted_async(master, n, counts, hops, wsum, wsq, w_out=w_out, hit_out=hit, hops_out=hp,
                             seed=seed, eps=eps, first_walk=first, job=job, max_hops=max_hops, flags=flags)
    torch.cuda.synchronize()
    return (counts.cpu().numpy(), int(hops.item()), wsum.cpu().numpy(), wsq.cpu().numpy(),
            w_out.cpu().numpy(), hit.cpu().numpy(), hp.cpu().numpy())


@pytest.mark.parametrize("name,master", FIXTURES[:4])
def test_weighted_walks_replay_oracle(name, master):
    lay, rects, bounds = load_fixture(name)
    tree = frw.Tree(lay)
    n = 3000
    counts, hops, wsum, wsq, w, hit, hp = gpu_weighted(tree, master, n, seed=5)
    ix = oracle_index(rects, bounds, O.REFTREE_DFS)
    ohit, ohp, ow = ix.walk_weighted(master, n, seed=5)
    assert np.array_equal(hit, ohit), np.nonzero(hit != ohit)[0][:10]
    assert np.array_equal(hp, ohp)
    np.testing.assert_allclose(w, ow, rtol=1e-13, atol=0)
    assert hops == ohp.sum()
    nr = len(rects)
    bins = np.where(ohit >= 0, ohit, np.where(ohit == -1, nr, nr + 1))
    assert np.array_equal(counts, np.bincount(bins, minlength=nr + 2))
    want = np.bincount(bins, weights=ow, minlength=nr + 2)
    want_q = np.bincount(bins, weights=ow * ow, minlength=nr + 2)
    scale = np.abs(ow).sum() + 1e-300
    assert np.abs(wsum - want).max() <= 1e-12 * scale
    assert np.abs(wsq - want_q).max() <= 1e-12 * (ow * ow).sum()


@pytest.mark.parametrize("name,master", FIXTURES[:4])
def test_importance_walks_replay_oracle(name, master):
    """Importance-sampled first hop: same panels, weights and walks as the oracle."""
    lay, rects, bounds = load_fixture(name)
    tree = frw.Tree(lay)
    n = 3000
    counts, hops, wsum, wsq, w, hit, hp = gpu_weighted(tree, master, n, seed=8, flags=frw.WEIGHT_IMPORTANCE)
    ohit, ohp, ow = oracle_index(rects, bounds, O.REFTREE_DFS).walk_weighted(master, n, seed=8, importance=True)
    assert np.array_equal(hit, ohit), np.nonzero(hit != ohit)[0][:10]
    assert np.array_equal(hp, ohp)
    np.testing.assert_allclose(w, ow, rtol=1e-13, atol=0)
    assert hops == ohp.sum()


def test_importance_parallel_plates():
    """Same coupling as the plain estimator within the Monte Carlo error,
    zero-sum weights, smaller variance of the self term."""
    lay = frw.Layout.from_rects(_plates())
    tree = frw.Tree(lay)
    n = 2_000_000
    _, ws_p, wq_p, _ = tree.walk_weighted(0, n, seed=11)
    _, ws_i, wq_i, _ = tree.walk_weighted(0, n, seed=12, flags=frw.WEIGHT_IMPORTANCE)
    cp, ci = ws_p / n, ws_i / n
    vp, vi = wq_p / n - cp * cp, wq_i / n - ci * ci
    for j in (0, 1):
        assert abs(cp[j] - ci[j]) < 4 * np.sqrt((vp[j] + vi[j]) / n), (j, cp[j], ci[j])
    assert -25.0 < ci[1] < -20.0, ci
    tot = ws_i.sum() / n
    assert abs(tot) < 4 * np.sqrt(max(wq_i.sum() / n - tot * tot, 0) / n)
    assert vi[0] < 0.9 * vp[0], (vi[0], vp[0])


def test_weighted_seed_job_offset_and_max_hops():
    lay, rects, bounds = load_fixture("test100.txt")
    tree = frw.Tree(lay, leaf_cap=4)
    ix = oracle_index(rects, bounds, O.REFTREE_DFS)
    # (the third range crosses walk id 2^32: the ids' high word enters both
    # the 'GAUS' start and the 'WALK' hop streams)
    for seed, job, first, mh in [(9, 3, 12345, 0), (1, 0, 0, 2), (7, 1, (1 << 32) - 700, 0)]:
        c, h, ws, wq, w, hit, hp = gpu_weighted(tree, 3, 1500, seed=seed, first=first, job=job, max_hops=mh)
        ohit, ohp, ow = ix.walk_weighted(3, 1500, seed=seed, job=job, first_walk=first,
                                         max_hops=mh if mh else 10 ** 6)
        assert np.array_equal(hit, ohit) and np.array_equal(hp, ohp)
        np.testing.assert_allclose(w, ow, rtol=1e-13, atol=0)


def fixed_point_sums(w, bins, nb):
    """The sums the kernel forms: q = rint(w 2^32) per walk, sum of q and of
    q^2 >> 32 per bin as exact integers, converted as sign and magnitude
    (hi 2^64 + lo) and scaled by 2^-32."""
    q = np.rint(np.asarray(w) * 2.0 ** 32).astype(np.int64).tolist()
    S, Q = [0] * nb, [0] * nb
    for b, v in zip(np.asarray(bins).tolist(), q):
        S[b] += v
        Q[b] += (v * v) >> 32

    def conv(v):
        m = -v if v < 0 else v
        d = float(m >> 64) * 18446744073709551616.0 + float(m & ((1 << 64) - 1))
        return (-d if v < 0 else d) * (1.0 / 4294967296.0)

    return np.array([conv(v) for v in S]), np.array([conv(v) for v in Q])


@pytest.mark.parametrize("flags", [0, frw.WEIGHT_IMPORTANCE])
def test_weighted_sums_exact_for_any_launch_shape(flags):
    torch = torch_dev()
    lay, rects, bounds = load_fixture("test1000.txt")
    tree = frw.Tree(lay)
    nr = len(rects)
    n = 200_000
    outs = []
    for block, grid in ((0, 0), (64, 37), (256, 1000)):
        counts = torch.zeros(nr + 2, dtype=torch.int64, device="cuda")
        hops = torch.zeros(1, dtype=torch.int64, device="cuda")
        wsum = torch.zeros(nr + 2, dtype=torch.float64, device="cuda")
        wsq = torch.zeros(nr + 2, dtype=torch.float64, device="cuda")
        w_out = torch.zeros(n, dtype=torch.float64, device="cuda")
        hit = torch.zeros(n, dtype=torch.int32, device="cuda")
        tree.walk_weighted_async(0, n, counts, hops, wsum, wsq, w_out=w_out, hit_out=hit, seed=3, block=block,
                                 grid=grid, flags=flags)
        torch.cuda.synchronize()
        outs.append([t.cpu().numpy() for t in (counts, wsum, wsq, w_out, hit)])
    for o in outs[1:]:
        for a, b in zip(outs[0], o):
            assert np.array_equal(a, b)  # bit for bit, whatever the shape
    counts, wsum, wsq, w, hit = outs[0]
    bins = np.where(hit >= 0, hit, np.where(hit == -1, nr, nr + 1))
    S, Q = fixed_point_sums(w, bins, nr + 2)
    assert np.array_equal(wsum, S) and np.array_equal(wsq, Q)
    # a second launch into the same bins adds its own converted sums
    wsum2 = torch.from_numpy(wsum.copy()).cuda()
    wsq2 = torch.from_numpy(wsq.copy()).cuda()
    c2 = torch.zeros(nr + 2, dtype=torch.int64, device="cuda")
    h2 = torch.zeros(1, dtype=torch.int64, device="cuda")
    tree.walk_weighted_async(0, n, c2, h2, wsum2, wsq2, seed=3, flags=flags)
    torch.cuda.synchronize()
    assert np.array_equal(wsum2.cpu().numpy(), wsum + S) and np.array_equal(wsq2.cpu().numpy(), wsq + Q)


def _plates(gap=100, length=2000, thick=10):
    return np.array([[0, 0, length, thick], [0, thick + gap, length, 2 * thick + gap]], dtype=np.int32)


def test_parallel_plates():
    lay = frw.Layout.from_rects(_plates())
    tree = frw.Tree(lay)
    n = 2_000_000
    counts, wsum, wsq, hops = tree.walk_weighted(0, n, seed=11)
    c = wsum / n
    sig = np.sqrt(np.maximum(wsq / n - c * c, 0) / n)
    assert counts.sum() == n
    assert -25.0 < c[1] < -20.0, (c, sig)  # -length/gap = -20, plus fringe
    assert c[0] > 0 and c[2] < 0           # self term; the grounded boundary
    tot = wsum.sum() / n
    sig_tot = np.sqrt(max(wsq.sum() / n - tot * tot, 0) / n)
    assert abs(tot) < 4 * sig_tot, (tot, sig_tot)
    # the same walks again: identical tallies and sums, bit for bit
    counts2, wsum2, wsq2, hops2 = tree.walk_weighted(0, n, seed=11)
    assert np.array_equal(counts, counts2) and hops == hops2
    assert np.array_equal(wsum2, wsum) and np.array_equal(wsq2, wsq)


def test_weighted_at_scale():
    lay = frw.Layout.generate(10 ** 6, 10 ** 7, 5, 300, seed=1, cluster_frac=0.2, sigma=1e5, n_clusters=100)
    tree = frw.Tree(lay)
    n = 10 ** 7
    counts, wsum, wsq, hops = tree.walk_weighted(0, n, seed=2)
    assert counts.sum() == n and np.isfinite(wsum).all() and np.isfinite(wsq).all()
    c = wsum / n
    assert c[0] > 0
    tot = wsum.sum() / n
    sig_tot = np.sqrt(max(wsq.sum() / n - tot * tot, 0) / n)
    assert abs(tot) < 4 * sig_tot, (tot, sig_tot)


def test_weighted_rejects_bad_arguments():
    lay, rects, bounds = load_fixture("test")
    tree = frw.Tree(lay)
    with pytest.raises(frw.FrwError):
        tree.walk_weighted(0, 100, eps=0.0)
    with pytest.raises(frw.FrwError):
        tree.walk_weighted(len(rects), 100)
    with pytest.raises(frw.FrwError):
        tree.walk_weighted(0, 100, flags=8)
