"""The oracle's restatement of the learned proposal (oracle/orc_mlp.c), pinned
against independent computations: torch's bfloat16 conversion, a float64
autograd of the same network, and a numpy rasterisation of the occupancy
windows.  CPU only."""
import numpy as np
import pytest
import torch

import oracle_ffi as O

HID, INP = 64, O.MLP_IN_PAD


def _random_params(rng, scale=0.3):
    P = rng.normal(0.0, scale, O.mlp_param_count()).astype(np.float32)
    W1 = P[: HID * INP].reshape(HID, INP)
    W1[:, 67:] = 0.0  # padding columns of the first layer are zero
    return P


def _random_features(rng, n):
    X = np.zeros((n, INP), dtype=np.float32)
    X[:, :64] = rng.integers(0, 2, (n, 64))
    X[:, 64] = rng.integers(0, 13, n) / 16.0
    return X


def _split(P):
    o = 0
    out = {}
    for name, shape in (("W1", (HID, INP)), ("b1", (HID,)), ("W2", (HID, HID)), ("b2", (HID,)),
                        ("W3", (HID, HID)), ("b3", (HID,)), ("w4", (HID,)), ("b4", (1,))):
        n = int(np.prod(shape))
        out[name] = P[o:o + n].reshape(shape)
        o += n
    assert o == len(P)
    return out


def _torch_loss(P, X, t, rounding):
    """Same network in torch (float64 unless rounding emulates bf16)."""
    p = _split(P)
    dt = torch.float32 if rounding else torch.float64
    T = {k: torch.tensor(v, dtype=dt, requires_grad=True) for k, v in p.items()}

    def r(x):
        return x.to(torch.bfloat16).to(dt) if rounding else x

    a = r(torch.tensor(X, dtype=dt))
    for wi, bi in (("W1", "b1"), ("W2", "b2"), ("W3", "b3")):
        a = r(torch.nn.functional.gelu(a @ r(T[wi]).T + T[bi]))
    y = a @ r(T["w4"]) + T["b4"]
    loss = 0.5 * ((y - torch.tensor(t, dtype=dt)) ** 2).mean()
    return loss, y, T


def test_bf16_matches_torch_rounding():
    rng = np.random.default_rng(3)
    xs = np.concatenate([
        rng.normal(0, 1, 2000), rng.normal(0, 1e-3, 500), rng.normal(0, 1e30, 200),
        # exact ties: mantissa low half 0x8000, with even and odd upper halves
        np.array([1.0 + 2.0 ** -8, 1.0 + 3 * 2.0 ** -8, -(1.0 + 2.0 ** -8), 1e-40, -0.0, 0.0]),
    ]).astype(np.float32)
    want = torch.tensor(xs).to(torch.bfloat16).to(torch.float32).numpy()
    got = np.array([O.bf16(float(x)) for x in xs], dtype=np.float32)
    assert np.array_equal(got.view(np.uint32), want.view(np.uint32))


def test_level_and_cell():
    # K0 = 11 for a 1024-wide box (2^10 is not > 1024), Lmax = 11
    assert O.mlp_level((0, 0, 1024, 1024), 1.0) == (11, 1.0)       # e = 0 -> l = 12 -> clamp 11
    assert O.mlp_level((0, 0, 1024, 1024), 0.25) == (11, 1.0)      # h < 1 counts as 1
    assert O.mlp_level((0, 0, 1024, 1024), 37.0) == (7, 16.0)      # e = 5 -> l = 7, s = 2^4
    assert O.mlp_level((0, 0, 1024, 1024), 5000.0) == (0, 2048.0)  # clamp at 0, s = 2^K0
    assert O.mlp_level((0, 0, 10 ** 7, 10 ** 7), 3.0) == (15, 2.0 ** 9)  # Lmax = 15 cap (K0 = 24)


def _raster(rects, bnds, level, K0):
    """Occupancy of the 2^l x 2^l grid by integer cell ranges (numpy)."""
    n = 1 << level
    s = 2 ** (K0 - level)
    g = np.zeros((n, n), dtype=np.uint8)
    for x0, y0, x1, y1 in rects:
        i0, i1 = int((x0 - bnds[0]) // s), min(int((x1 - bnds[0]) // s), n - 1)
        j0, j1 = int((y0 - bnds[1]) // s), min(int((y1 - bnds[1]) // s), n - 1)
        g[j0:j1 + 1, i0:i1 + 1] = 1
    return g


def _window(grid, ci, cj):
    n = grid.shape[0]
    f = np.ones(64, dtype=np.float32)
    for v in range(8):
        for u in range(8):
            i, j = ci - 4 + u, cj - 4 + v
            if 0 <= i < n and 0 <= j < n:
                f[v * 8 + u] = grid[j, i]
    return f


def test_features_hand_case():
    # one rect [100,100]-[150,120] in a 1024 box; h = 37 -> level 7, cell 16
    rects = np.array([[100, 100, 150, 120]], dtype=np.float64)
    b = (0.0, 0.0, 1024.0, 1024.0)
    X = O.mlp_features(rects, b, [(130.0, 160.0)], [37.0])
    f = X[0, :64].reshape(8, 8)  # [v][u]
    # p's cell (8, 10); window columns i = 4..11, rows j = 6..13
    # rect cells: i = 6..9 (100//16=6, 150//16=9), j = 6..7 (100//16=6, 120//16=7)
    want = np.zeros((8, 8), dtype=np.float32)
    want[0:2, 2:6] = 1.0
    assert np.array_equal(f, want)
    assert X[0, 64] == pytest.approx(7 / 16)
    assert X[0, 65] == 0.125 and X[0, 66] == 0.0  # 130/16 = 8.125, 160/16 = 10
    assert not X[0, 67:].any()
    # a point near the corner: cells outside the grid read as occupied
    X = O.mlp_features(rects, b, [(5.0, 5.0)], [37.0])
    f = X[0, :64].reshape(8, 8)
    assert f[:4, :].all() and f[:, :4].all() and not f[4:, 4:].any()


def test_features_match_rasterisation():
    rng = np.random.default_rng(7)
    n = 60
    x0 = rng.integers(0, 3800, n)
    y0 = rng.integers(0, 3800, n)
    rects = np.stack([x0, y0, x0 + rng.integers(1, 200, n), y0 + rng.integers(1, 200, n)], 1)
    rects = rects.astype(np.float64)
    b = (-40.0, -40.0, 4100.0, 4100.0)
    K0 = int(np.floor(np.log2(4140))) + 1
    pts = rng.uniform(-39, 4099, (300, 2))
    hs = np.exp(rng.uniform(np.log(0.3), np.log(5000), 300))
    X = O.mlp_features(rects, b, pts, hs)
    for (px, py), h, row in zip(pts, hs, X):
        lvl, s = O.mlp_level(b, h)
        g = _raster(rects, b, lvl, K0)
        ci, cj = int(np.floor((px - b[0]) / s)), int(np.floor((py - b[1]) / s))
        assert np.array_equal(row[:64], _window(g, ci, cj))
        assert row[64] == np.float32(lvl / 16)
        fx = torch.tensor((px - b[0]) / s - ci, dtype=torch.float32).to(torch.bfloat16).item()
        fy = torch.tensor((py - b[1]) / s - cj, dtype=torch.float32).to(torch.bfloat16).item()
        assert row[65] == fx and row[66] == fy


def test_forward_and_gradient_match_autograd_without_rounding():
    rng = np.random.default_rng(11)
    P = _random_params(rng)
    X = _random_features(rng, 48)
    t = rng.normal(0, 2, 48).astype(np.float32)
    y = O.mlp_forward(P, X, rounding=False)
    loss, g = O.mlp_backward(P, X, t, rounding=False)
    L, Y, T = _torch_loss(P, X, t, rounding=False)
    L.backward()
    np.testing.assert_allclose(y, Y.detach().numpy(), rtol=1e-4, atol=1e-4)
    assert loss == pytest.approx(L.item(), rel=1e-4)
    o = 0
    for name in ("W1", "b1", "W2", "b2", "W3", "b3", "w4", "b4"):
        ref = T[name].grad.numpy().ravel()
        got = g[o:o + ref.size]
        o += ref.size
        scale = np.abs(ref).max() + 1e-12
        assert np.abs(got - ref).max() <= 2e-4 * scale, name


def test_rounded_forward_matches_torch_bf16_emulation():
    rng = np.random.default_rng(12)
    P = _random_params(rng)
    X = _random_features(rng, 64)
    t = rng.normal(0, 2, 64).astype(np.float32)
    y = O.mlp_forward(P, X, rounding=True)
    _, Y, _ = _torch_loss(P, X, t, rounding=True)
    # summation order differs, so an activation can round to the other bf16
    # neighbour now and then: compare in aggregate
    d = np.abs(y - Y.detach().numpy())
    assert d.mean() < 2e-3 * (np.abs(y).mean() + 1) and d.max() < 3e-2 * (np.abs(y).max() + 1)
    # and rounding changes the answer (the switch is live)
    y0 = O.mlp_forward(P, X, rounding=False)
    assert np.abs(y0 - y).max() > 1e-4



def test_vmax_rule():
    # spacing = sqrt(W H / n); finest verification cell about spacing / 128
    assert O.mlp_vmax((0, 0, 100000, 100000), 10000) == 17 - 2   # spacing 1000 -> 7.8 -> 2^2
    assert O.mlp_vmax((0, 0, 60000, 60000), 3000) == 16 - 3      # spacing 1095 -> 8.6 -> 2^3
    assert O.mlp_vmax((0, 0, 10 ** 7, 10 ** 7), 10 ** 6) == 15   # capped at 15
    assert O.mlp_vmax((0, 0, 100, 100), 10 ** 6) == 7            # cell >= 1 -> level K0


def test_certify_against_rasterisation_and_exact_q1():
    rng = np.random.default_rng(17)
    n = 80
    x0 = rng.integers(0, 3800, n)
    y0 = rng.integers(0, 3800, n)
    rects = np.stack([x0, y0, x0 + rng.integers(1, 200, n), y0 + rng.integers(1, 200, n)], 1)
    rects = rects.astype(np.float64)
    b = (-40.0, -40.0, 4100.0, 4100.0)
    K0 = int(np.floor(np.log2(4140))) + 1
    Vmax = O.mlp_vmax(b, n)
    idx = O.Index(rects, b, O.EXACT)
    m = 1500
    pts = rng.uniform(-39, 4099, (m, 2))
    dstar = np.array([idx.nearest(px, py)[0] for px, py in pts])
    ds = dstar * rng.uniform(0.3, 1.2, m)
    got = O.mlp_certify(rects, b, pts, ds)
    for (px, py), d, c in zip(pts, ds, got):
        if not d > 0:
            assert c == 0
            continue
        sq = (px - d, py - d, px + d, py + d)
        inside = sq[0] >= b[0] and sq[1] >= b[1] and sq[2] <= b[2] and sq[3] <= b[3]
        v = min(max(K0 - int(np.floor(np.log2(d))) + 4, 0), Vmax)
        s = 2.0 ** (K0 - v)
        g = _raster(rects, b, v, K0)
        i0, i1 = int(np.floor((sq[0] - b[0]) / s)), int(np.floor((sq[2] - b[0]) / s))
        j0, j1 = int(np.floor((sq[1] - b[1]) / s)), int(np.floor((sq[3] - b[1]) / s))
        in_grid = i0 >= 0 and j0 >= 0 and i1 < g.shape[0] and j1 < g.shape[0]
        want = inside and in_grid and not g[j0:j1 + 1, i0:i1 + 1].any()
        assert bool(c) == want
        if c:  # a certified square is free: exact Q1 finds no rect
            assert not idx.q1(sq)
    # the probe is not vacuous: well inside d*, most squares certify
    assert got[ds < 0.5 * dstar].mean() > 0.5
