"""Learned square-size proposal on the GPU against the CPU oracle.

* occupancy bitmaps and feature windows: bit-exact against a numpy
  rasterisation and the oracle's brute force over all rects;
* MFMA forward pass and the training gradients: within bf16 tolerance of the
  oracle's restatement (weights and activations rounded to bf16, FP32 sums);
* harvest: every recorded d* equals the exact nearest-rect distance;
* training on harvested pairs lowers the loss;
* learned walks: every accepted square is collision-free by construction, so
  the hit distribution must equal the exact method's (chi-square homogeneity
  over independent walk ranges);
* what the network buys: walks/s against the reference's bisection search,
  and the measured per-hop bound that keeps it below the exact search."""
import numpy as np
import pytest

import frw
import oracle_ffi as O
from gpu_common import oracle_index, to_dev, torch_dev

pytestmark = pytest.mark.gpu

INP = O.MLP_IN_PAD


def _layout(n=3000, box=60000, seed=5):
    lay = frw.Layout.generate(n, box, 10, 500, seed=seed)
    return lay, lay.rects(), tuple(float(v) for v in lay.bounds())


def _k0(bnds):
    m = max(bnds[2] - bnds[0], bnds[3] - bnds[1])
    k = 0
    while 2 ** k <= m:
        k += 1
    return k


@pytest.fixture(scope="module")
def setup():
    torch_dev()
    lay, rects, bnds = _layout()
    tree = frw.Tree(lay)
    mlp = frw.Mlp(tree, seed=1)
    return lay, rects, bnds, tree, mlp


def test_pyramid_levels(setup):
    lay, rects, bnds, tree, mlp = setup
    K0, Lmax, Vmax, Pmax = mlp.levels()
    assert K0 == _k0(bnds) and Lmax == min(K0, 15)
    assert Vmax == O.mlp_vmax(bnds, len(rects)) and Pmax == max(Lmax, Vmax)


def test_occupancy_bitmaps(setup):
    lay, rects, bnds, tree, mlp = setup
    K0 = _k0(bnds)
    R = rects.astype(np.int64)
    ox, oy = int(bnds[0]), int(bnds[1])
    Pmax = mlp.levels()[3]
    for lvl in range(0, min(Pmax, 12) + 1):
        n = 1 << lvl
        s = 2 ** (K0 - lvl)
        g = np.zeros((n, n), dtype=np.uint8)
        for x0, y0, x1, y1 in R:
            i0, i1 = (x0 - ox) // s, min((x1 - ox) // s, n - 1)
            j0, j1 = (y0 - oy) // s, min((y1 - oy) // s, n - 1)
            g[j0:j1 + 1, i0:i1 + 1] = 1
        assert np.array_equal(mlp.occupancy(lvl), g), lvl
    # the finest level on sampled cells, against every rect
    lvl = Pmax
    s = 2 ** (K0 - lvl)
    g = mlp.occupancy(lvl)
    rng = np.random.default_rng(1)
    cells = rng.integers(0, 1 << lvl, (4000, 2))
    sel = R[rng.integers(0, len(R), 1000)]  # plus cells under rect corners
    cells = np.vstack([cells, np.column_stack([(sel[:, 0] - ox) // s, (sel[:, 1] - oy) // s])])
    for i, j in cells:
        a, b = i * s, (i + 1) * s
        c, d = j * s, (j + 1) * s
        occ = np.any((R[:, 0] - ox < b) & (R[:, 2] - ox >= a) & (R[:, 1] - oy < d) & (R[:, 3] - oy >= c))
        assert g[j, i] == occ, (i, j)


def test_features_bit_exact(setup):
    torch = torch_dev()
    lay, rects, bnds, tree, mlp = setup
    rng = np.random.default_rng(2)
    n = 700
    pts = np.column_stack([rng.uniform(bnds[0], bnds[2], n), rng.uniform(bnds[1], bnds[3], n)])
    # points next to the bounds (windows that leave the grid) and at rect corners
    pts[:20, 0] = bnds[0] + rng.uniform(0, 3, 20)
    pts[20:40, 1] = bnds[3] - rng.uniform(0, 3, 20)
    pts[40:80] = rects[:40, :2] - 0.5
    hs = np.exp(rng.uniform(np.log(0.2), np.log(2e5), n))
    out = torch.zeros((n, INP), dtype=torch.float32, device="cuda")
    mlp.features(to_dev(torch, pts), to_dev(torch, hs), out)
    got = out.cpu().numpy()
    want = O.mlp_features(rects.astype(np.float64), bnds, pts, hs)
    bad = np.nonzero((got != want).any(1))[0]
    assert len(bad) == 0, (bad[:5], pts[bad[:5]], hs[bad[:5]])


def test_certify_bit_exact(setup):
    torch = torch_dev()
    lay, rects, bnds, tree, mlp = setup
    rng = np.random.default_rng(6)
    n = 3000
    pts = np.column_stack([rng.uniform(bnds[0], bnds[2], n), rng.uniform(bnds[1], bnds[3], n)])
    dpts = to_dev(torch, pts)
    dstar = torch.zeros(n, dtype=torch.float64, device="cuda")
    tree.nearest(dpts, dstar)
    ds = dstar.cpu().numpy() * rng.uniform(0.2, 1.1, n)
    ds[:10] = [0.0, -1.0, 1e-3, 0.5, 2.0, 3.0, 1e5, 1e9, np.inf, np.nan]
    out = torch.zeros(n, dtype=torch.uint8, device="cuda")
    mlp.certify(dpts, to_dev(torch, ds), out)
    got = out.cpu().numpy()
    want = O.mlp_certify(rects.astype(np.float64), bnds, pts, ds)
    assert np.array_equal(got, want), np.nonzero(got != want)[0][:10]
    assert 0.2 < got.mean() < 1.0
    # certified squares are free (GPU exact Q1 on the same squares)
    sel = np.nonzero(got)[0]
    boxes = np.column_stack([pts[sel, 0] - ds[sel], pts[sel, 1] - ds[sel],
                             pts[sel, 0] + ds[sel], pts[sel, 1] + ds[sel]])
    hit = torch.zeros(len(sel), dtype=torch.uint8, device="cuda")
    tree.collide(to_dev(torch, boxes), hit)
    assert not hit.any()


def _random_params(rng, n_params, scale=0.3):
    P = rng.normal(0.0, scale, n_params).astype(np.float32)
    P[: 64 * INP].reshape(64, INP)[:, 67:] = 0.0
    return P


def _random_features(rng, n):
    X = np.zeros((n, INP), dtype=np.float32)
    X[:, :64] = rng.integers(0, 2, (n, 64))
    X[:, 64] = rng.integers(0, 13, n) / 16.0
    X[:, 65:67] = rng.random((n, 2))
    return X


def test_forward_matches_oracle(setup):
    torch = torch_dev()
    mlp = setup[4]
    rng = np.random.default_rng(3)
    P = _random_params(rng, mlp.n_params)
    mlp.set_params(P)
    assert np.array_equal(mlp.params(), P)
    for n in (1000, 64, 5):  # ragged last wave, one wave, a partial wave
        X = _random_features(rng, n)
        y = torch.zeros(n, dtype=torch.float32, device="cuda")
        mlp.forward(to_dev(torch, X), y)
        got = y.cpu().numpy()
        want = O.mlp_forward(P, X)
        d = np.abs(got - want)
        # FP32 sums in another order: a hidden activation may round to the
        # neighbouring bf16 value, so the bound is on the aggregate
        assert d.mean() <= 2e-3 * (np.abs(want).mean() + 1), (n, d.mean())
        assert d.max() <= 3e-2 * (np.abs(want).max() + 1), (n, d.max())


def test_gradient_matches_oracle(setup):
    torch = torch_dev()
    mlp = setup[4]
    rng = np.random.default_rng(4)
    P = _random_params(rng, mlp.n_params)
    mlp.set_params(P)
    n = 256
    X = _random_features(rng, n)
    t = rng.normal(0, 2, n).astype(np.float32)
    g, loss = mlp.gradient(to_dev(torch, X), to_dev(torch, t))
    oloss, og = O.mlp_backward(P, X, t)
    assert loss == pytest.approx(oloss, rel=2e-3)
    offs = np.cumsum([0, 64 * INP, 64, 64 * 64, 64, 64 * 64, 64, 64, 1])
    for name, a, b in zip(("W1", "b1", "W2", "b2", "W3", "b3", "w4", "b4"), offs[:-1], offs[1:]):
        ga, oa = g[a:b], og[a:b]
        err = np.linalg.norm(ga - oa) / (np.linalg.norm(oa) + 1e-12)
        assert err < 2e-2, (name, err)
    # with a wrong target the gradient moves (the check is not vacuous)
    g2, _ = mlp.gradient(to_dev(torch, X), to_dev(torch, t + 1.0))
    assert np.linalg.norm(g2 - g) > 1e-3 * np.linalg.norm(g)


@pytest.fixture(scope="module")
def harvested(setup):
    torch = torch_dev()
    lay, rects, bnds, tree, mlp = setup
    cap = 400_000
    pts = torch.zeros((cap, 2), dtype=torch.float64, device="cuda")
    h = torch.zeros(cap, dtype=torch.float64, device="cuda")
    ds = torch.zeros(cap, dtype=torch.float64, device="cuda")
    n = frw.walk_harvest(tree, 0, 40_000, pts, h, ds, seed=11)
    assert 0 < n <= cap
    return pts[:n].contiguous(), h[:n].contiguous(), ds[:n].contiguous()


def test_harvest_is_deterministic_and_ordered(setup, harvested):
    """The records come in walk and hop order: the same walks give the same
    arrays, and a smaller capacity gives a prefix of them."""
    torch = torch_dev()
    lay, rects, bnds, tree, mlp = setup
    pts, h, ds = harvested
    n = pts.shape[0]
    cap = 400_000
    p2 = torch.zeros((cap, 2), dtype=torch.float64, device="cuda")
    h2 = torch.zeros(cap, dtype=torch.float64, device="cuda")
    d2 = torch.zeros(cap, dtype=torch.float64, device="cuda")
    assert frw.walk_harvest(tree, 0, 40_000, p2, h2, d2, seed=11) == n
    assert torch.equal(p2[:n], pts) and torch.equal(h2[:n], h) and torch.equal(d2[:n], ds)
    small = 12_345
    p3 = torch.zeros((small, 2), dtype=torch.float64, device="cuda")
    h3 = torch.zeros(small, dtype=torch.float64, device="cuda")
    d3 = torch.zeros(small, dtype=torch.float64, device="cuda")
    assert frw.walk_harvest(tree, 0, 40_000, p3, h3, d3, seed=11) == small
    assert torch.equal(p3, pts[:small]) and torch.equal(d3, ds[:small])
    # the first walk's records are its hops in order: the first point lies on
    # the master's Gaussian surface
    g = tree.gauss_surface(0)
    x, y = pts[0].tolist()
    assert g[0] <= x <= g[2] and g[1] <= y <= g[3] and (x in (g[0], g[2]) or y in (g[1], g[3]))


def test_harvest_records_exact_distances(setup, harvested):
    torch = torch_dev()
    lay, rects, bnds, tree, mlp = setup
    pts, h, ds = harvested
    out = torch.zeros(pts.shape[0], dtype=torch.float64, device="cuda")
    tree.nearest(pts, out)
    assert torch.equal(out, ds)
    assert bool((ds > 0).all()) and bool((h > 0).all())
    idx = oracle_index(rects, bnds, O.EXACT)
    P = pts[:3000].cpu().numpy()
    D = ds[:3000].cpu().numpy()
    for (px, py), d in zip(P, D):
        assert idx.nearest(px, py)[0] == d


@pytest.fixture(scope="module")
def trained(setup, harvested):
    lay, rects, bnds, tree, mlp = setup
    pts, h, ds = harvested
    mlp2 = frw.Mlp(tree, seed=7)
    mlp2.set_samples(pts, h, ds)
    # loss of the initial network: one pass with a negligible step
    first = mlp2.train(epochs=1, batch=1024, lr=1e-12, seed=1)
    last = mlp2.train(epochs=6, batch=1024, lr=2e-3, seed=2)
    return mlp2, first, last


def test_samples_in_pieces_train_alike(setup, harvested):
    """The training set grows geometrically: samples appended in three calls
    (the second and third beyond the capacity the first left) train to the
    same parameters as the same samples set at once, and a smaller set
    written over a larger one (the capacity kept) trains as on a fresh
    network."""
    lay, rects, bnds, tree, mlp = setup
    pts, h, ds = harvested
    n = pts.shape[0]
    assert n >= 4096
    cuts = [0, n // 7, n // 7 + 1000, n]

    def trained_params(feed, seed=3):
        m = frw.Mlp(tree, seed=7)
        feed(m)
        loss = m.train(epochs=1, batch=1024, lr=2e-3, seed=seed)
        return m.params(), loss

    def pieces(m):
        for lo, hi in zip(cuts[:-1], cuts[1:]):
            m.add_samples(pts[lo:hi], h[lo:hi], ds[lo:hi])

    whole, lw = trained_params(lambda m: m.set_samples(pts, h, ds))
    parts, lp = trained_params(pieces)
    assert np.array_equal(whole, parts) and lw == lp
    half = n // 2

    def shrink(m):
        m.set_samples(pts, h, ds)
        m.set_samples(pts[:half], h[:half], ds[:half])

    small, ls = trained_params(shrink)
    fresh, lf = trained_params(lambda m: m.set_samples(pts[:half], h[:half], ds[:half]))
    assert np.array_equal(small, fresh) and ls == lf
    assert not np.array_equal(small, whole)


def test_training_lowers_loss(trained):
    mlp2, first, last = trained
    assert np.isfinite(first) and np.isfinite(last)
    assert last < 0.5 * first, (first, last)


def test_trained_forward_still_matches_oracle(setup, harvested, trained):
    torch = torch_dev()
    mlp2 = trained[0]
    pts, h, ds = harvested
    n = 512
    feats = torch.zeros((n, INP), dtype=torch.float32, device="cuda")
    mlp2.features(pts[:n].contiguous(), h[:n].contiguous(), feats)
    y = torch.zeros(n, dtype=torch.float32, device="cuda")
    mlp2.forward(feats, y)
    want = O.mlp_forward(mlp2.params(), feats.cpu().numpy())
    d = np.abs(y.cpu().numpy() - want)
    assert d.mean() <= 2e-3 * (np.abs(want).mean() + 1)


def _learned(tree, mlp, master, n, seed, alpha=0.8, first=0):
    torch = torch_dev()
    nr = tree.info().n_rects
    counts = torch.zeros(nr + 2, dtype=torch.int64, device="cuda")
    hops = torch.zeros(1, dtype=torch.int64, device="cuda")
    acc, fb = frw.walk_learned(tree, mlp, master, n, counts, hops, seed=seed, alpha=alpha,
                               first_walk=first)
    return counts.cpu().numpy(), int(hops.item()), acc, fb


def _exact(tree, master, n, seed, first=0):
    torch = torch_dev()
    nr = tree.info().n_rects
    counts = torch.zeros(nr + 2, dtype=torch.int64, device="cuda")
    hops = torch.zeros(1, dtype=torch.int64, device="cuda")
    tree.walk_async(master, n, counts, hops, seed=seed, first_walk=first)
    torch.cuda.synchronize()
    return counts.cpu().numpy(), int(hops.item())


def _homogeneous(c1, c2, min_expected=20):
    from scipy.stats import chi2_contingency
    keep = (c1 + c2) >= 2 * min_expected
    table = np.vstack([np.append(c1[keep], c1[~keep].sum()), np.append(c2[keep], c2[~keep].sum())])
    table = table[:, table.sum(0) > 0]
    return chi2_contingency(table)[1]


def test_learned_walk_distribution(setup, trained):
    lay, rects, bnds, tree, mlp = setup
    mlp2 = trained[0]
    n = 200_000
    c_l, hops_l, acc, fb = _learned(tree, mlp2, 0, n, seed=21)
    c_e, hops_e = _exact(tree, 0, n, seed=22)
    assert c_l.sum() == n and c_e.sum() == n
    assert acc + fb <= hops_l  # (accepted, rejected) proposals: at most one a hop
    assert acc > 0.3 * hops_l, (acc, fb)
    p = _homogeneous(c_l, c_e)
    assert p > 1e-4, p
    # determinism: the same walks again give the same tallies
    c_l2, hops_l2, acc2, fb2 = _learned(tree, mlp2, 0, n, seed=21)
    assert np.array_equal(c_l, c_l2) and hops_l2 == hops_l and acc2 == acc


def test_training_is_deterministic_and_its_seed_spread(setup, harvested):
    """Training on the same (ordered) harvest with the same seed gives the
    same parameters bit for bit; over five seeds the acceptance rate of the
    learned walks is reported and stays within a factor of two."""
    torch = torch_dev()
    lay, rects, bnds, tree, mlp = setup
    pts, h, ds = harvested
    acc_rates, params = [], []
    for seed in (1, 2, 3, 4, 5, 1):
        m = frw.Mlp(tree, seed=seed)
        m.set_samples(pts, h, ds)
        m.train(epochs=3, batch=1024, lr=2e-3, seed=seed)
        params.append(m.params())
        if len(acc_rates) < 5:
            c, hp, acc, fb = _learned(tree, m, 0, 50_000, seed=61)
            acc_rates.append(acc / (acc + fb))
    print("acceptance over seeds 1-5:", " ".join(f"{a:.3f}" for a in acc_rates))
    assert np.array_equal(params[0], params[5])
    assert not np.array_equal(params[0], params[1])
    assert min(acc_rates) > 0.5 * max(acc_rates), acc_rates


def _timed_walks(launch, reps=3):
    """Walks per second of launch() (HIP events on the current stream, best
    of `reps` after a warm-up)."""
    torch = torch_dev()
    launch()
    torch.cuda.synchronize()
    best = None
    for _ in range(reps):
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        launch()
        e1.record()
        torch.cuda.synchronize()
        ms = e0.elapsed_time(e1)
        best = ms if best is None else min(best, ms)
    return best


def _walk_rates(tree, mlp2, n, alpha=0.8):
    """(walks/s, hops/walk) of bisection, exact and learned walks over the
    same walk ids, and of learned walks whose every proposal is rejected
    (alpha so small that no proposal reaches the 2-unit minimum: the
    network and its features run on every hop that asks, then the exact
    step)."""
    torch = torch_dev()
    nr = tree.info().n_rects
    counts = torch.zeros(nr + 2, dtype=torch.int64, device="cuda")
    hops = torch.zeros(1, dtype=torch.int64, device="cuda")
    stream = torch.cuda.current_stream().cuda_stream
    out = {}

    total = {}

    def run(name, launch):
        hops.zero_()
        ms = _timed_walks(launch)
        total[name] = int(hops.item())
        per = total[name] / 4 / n  # warm-up + 3 timed launches
        out[name] = (n / (ms * 1e-3), per)

    run("bisect", lambda: tree.walk_async(0, n, counts, hops, seed=41, mode=frw.MODE_BISECT, stream=stream))
    run("exact", lambda: tree.walk_async(0, n, counts, hops, seed=41, stream=stream))
    stats = torch.zeros(2, dtype=torch.int64, device="cuda")
    run("learned", lambda: frw.walk_learned_async(tree, mlp2, 0, n, counts, hops, stats, seed=41, alpha=alpha,
                                                  stream=stream))
    st = stats.cpu().numpy()
    # accepted proposals per hop (the stats count the hops that asked the
    # network: accepted + rejected), over the same four launches as the hops
    out["accept_rate"] = int(st[0]) / max(1, total["learned"])
    run("network_only", lambda: frw.walk_learned_async(tree, mlp2, 0, n, counts, hops, seed=41, alpha=1e-9,
                                                       stream=stream))
    for k, v in out.items():
        if k != "accept_rate":
            print(f"{k}: {v[0]:.3e} walks/s, {v[1]:.2f} hops/walk")
    return out


def test_learned_gain_over_the_plain_bisection(setup, trained):
    """What the trained network gains over the plain FRW it was built for:
    the reference's bisection square search (20-45 collision queries per
    hop, randomWalk/CPU/quadTree.cpp:264-312) on the same walk ids."""
    lay, rects, bnds, tree, mlp = setup
    r = _walk_rates(tree, trained[0], 200_000)
    assert r["learned"][0] > 1.5 * r["bisect"][0], r


def test_network_cost_bounds_learned_below_exact(setup, trained):
    """Why the learned walk does not beat the exact search.  Per hop a learned
    walk pays at least the network's share t_net (the proposal features and
    the network for the hops that ask, measured with every proposal rejected)
    and, for the fraction 1 - a of hops without an accepted proposal, the exact
    step t_exact; and it takes no fewer hops than the exact walk (an accepted
    square is certified free, so it is no larger than the exact one).  So its
    walks/s are at most 1 / (H_learned (t_net + (1 - a) t_exact)), which the
    measured rates respect and which lies below the exact search's walks/s."""
    lay, rects, bnds, tree, mlp = setup
    r = _walk_rates(tree, trained[0], 200_000)
    t_exact = 1.0 / (r["exact"][0] * r["exact"][1])           # s per exact hop
    t_net = 1.0 / (r["network_only"][0] * r["network_only"][1]) - t_exact
    a = r["accept_rate"]
    bound = 1.0 / (r["learned"][1] * (t_net + (1.0 - a) * t_exact))
    print(f"per hop: exact {t_exact * 1e12:.1f} ps, network {t_net * 1e12:.1f} ps, accepted {a:.3f}; "
          f"learned walks/s bound {bound:.3e}")
    assert r["learned"][1] >= 0.98 * r["exact"][1]
    assert t_net > 0
    assert r["learned"][0] <= 1.05 * bound
    assert bound < r["exact"][0]


@pytest.mark.xfail(strict=True, reason="measured: the learned walk runs below the exact search's walks/s "
                   "(the network's per-hop cost alone exceeds an exact hop; see the bound test above)")
def test_learned_walks_per_second_reach_the_exact_search(setup, trained):
    lay, rects, bnds, tree, mlp = setup
    r = _walk_rates(tree, trained[0], 200_000)
    assert r["learned"][0] >= r["exact"][0], r


def test_learned_alpha_sweep(setup, trained):
    """Reports acceptance and hops per walk against the safety factor."""
    lay, rects, bnds, tree, mlp = setup
    mlp2, first, last = trained
    print(f"\ntraining loss: first epoch {first:.4f}, last {last:.4f}")
    te = tree.walk_traffic(0, 20_000, seed=23)
    print(f"exact: hops/walk {te[2] / 20_000:.2f} bytes/walk {(32 * te[0] + 16 * te[1] + 4 * te[3]) / 20_000:.0f}")
    for alpha in (0.5, 0.6, 0.7, 0.8, 0.9, 1.0):
        tl = frw.walk_learned_traffic(tree, mlp2, 0, 20_000, seed=23, alpha=alpha)
        b = 32 * tl[0] + 16 * tl[1] + 16 * (tl[3] + tl[4]) + 4 * tl[5]
        print(f"alpha {alpha}: hops/walk {tl[2] / 20_000:.2f} accept {tl[3] / max(1, tl[2]):.3f} "
              f"bytes/walk {b / 20_000:.0f}")
        assert tl[3] + tl[4] <= tl[2]  # proposals: at most one a hop


def test_untrained_network_falls_back(setup):
    """Whatever the network says, the walk distribution stays the exact one:
    a random network mostly falls back, and the tallies still agree."""
    lay, rects, bnds, tree, mlp = setup
    rng = np.random.default_rng(9)
    mlp.set_params(_random_params(rng, mlp.n_params, scale=0.5))
    n = 100_000
    c_l, hops_l, acc, fb = _learned(tree, mlp, 3, n, seed=31, alpha=1.0)
    c_e, hops_e = _exact(tree, 3, n, seed=32)
    assert acc + fb <= hops_l
    assert _homogeneous(c_l, c_e) > 1e-4


def test_learned_rejects_bad_arguments(setup, trained):
    torch = torch_dev()
    lay, rects, bnds, tree, mlp = setup
    counts = torch.zeros(len(rects) + 2, dtype=torch.int64, device="cuda")
    hops = torch.zeros(1, dtype=torch.int64, device="cuda")
    with pytest.raises(frw.FrwError):
        frw.walk_learned(tree, trained[0], 0, 10, counts, hops, alpha=0.0)
    with pytest.raises(frw.FrwError):
        frw.walk_learned(tree, trained[0], len(rects), 10, counts, hops)
    other = frw.Tree(frw.Layout.generate(100, 5000, 10, 100, seed=2))
    with pytest.raises(frw.FrwError):
        frw.walk_learned(other, trained[0], 0, 10, counts, hops)
