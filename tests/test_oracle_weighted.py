"""The oracle's weighted (physical) FRW estimator, pinned by the mathematics:
the exit density series against the Green CDF it extends, the gradient
tables against finite differences of the full series for every side, the
zero-sum of the weights, and a parallel-plate capacitor against its closed
form.  CPU only."""
import numpy as np
import pytest

import oracle_ffi as O

PER = O.NSAMPLE // 4


def _side_coords(side, x0, y0):
    """(u, v) of a source at (x0, y0) for the exit density of a side (y down,
    panels clockwise from the top-left corner)."""
    return [(x0, y0), (y0, 1 - x0), (1 - x0, 1 - y0), (1 - y0, x0)][side]


def test_density_at_centre_is_the_green_density():
    cdf = O.green_cdf()
    w = np.diff(cdf)  # normalised panel masses
    raw = np.array([O.square_density((j + 0.5) / PER, 0.5, 0.5) for j in range(PER)]) / PER
    # the cdf is built from the same density (19-term series): same shape
    np.testing.assert_allclose(w[:PER] / w[:PER].sum(), raw / raw.sum(), rtol=2e-6, atol=1e-9)
    # and the four sides are equal
    for s in range(4):
        np.testing.assert_allclose(w[s * PER:(s + 1) * PER], w[:PER], rtol=1e-12)


@pytest.mark.parametrize("side", [0, 1, 2, 3])
def test_gradient_matches_finite_differences(side):
    gx, gy = O.green_grad()
    cdf = O.green_cdf()
    T = sum(O.square_density((j + 0.5) / PER, 0.5, 0.5) for j in range(PER)) / PER * 4
    d = 1e-5
    for j in range(0, PER, 7):
        s = (j + 0.5) / PER
        k = side * PER + j

        def mass(x0, y0):
            u, v = _side_coords(side, x0, y0)
            return O.square_density(s, u, v) / (PER * T)

        fx = (mass(0.5 + d, 0.5) - mass(0.5 - d, 0.5)) / (2 * d)
        fy = (mass(0.5, 0.5 + d) - mass(0.5, 0.5 - d)) / (2 * d)
        assert gx[k] == pytest.approx(fx, rel=1e-5, abs=1e-9), (side, j)
        assert gy[k] == pytest.approx(fy, rel=1e-5, abs=1e-9), (side, j)


def test_gradient_direction_and_zero_sum():
    gx, gy = O.green_grad()
    assert abs(gx.sum()) < 1e-12 and abs(gy.sum()) < 1e-12
    # moving the source up (y decreasing, y down) raises the top side's mass
    assert (gy[:PER] < 0).all()
    # moving it right raises the right side's mass
    assert (gx[PER:2 * PER] > 0).all()


def _plates(gap=100, length=2000, thick=10):
    a = [0, 0, length, thick]
    b = [0, thick + gap, length, 2 * thick + gap]
    return np.array([a, b], dtype=np.float64)


def test_parallel_plates_capacitance():
    """Two plates of length 2000 and gap 100 (eps = 1): the coupling is
    about -length / gap = -20 (plus fringe), and the weights of all walks,
    escapes included, average to zero."""
    rects = _plates()
    ix = O.Index(rects, O.bounds(rects), O.EXACT)
    n = 40_000
    hit, hops, w = ix.walk_weighted(0, n, seed=3)
    assert np.isfinite(w).all() and (hops >= 0).all()
    c01 = w[hit == 1].sum() / n
    c00 = w[hit == 0].sum() / n
    sig = w.std() / np.sqrt(n)
    assert -26.0 < c01 < -19.0, c01
    assert c00 > 0
    assert abs(w.mean()) < 4 * sig, (w.mean(), sig)


def test_weighted_walks_are_deterministic():
    rects = _plates()
    ix = O.Index(rects, O.bounds(rects), O.REFTREE_DFS)
    a = ix.walk_weighted(0, 500, seed=7)
    b = ix.walk_weighted(0, 500, seed=7)
    for x, y in zip(a, b):
        assert np.array_equal(x, y)


def test_importance_tables():
    """Thresholds drawing panel k with probability proportional to |g_k|:
    every panel reachable, probabilities summing to one exactly and within
    two threshold units of |g_k| / sum |g|."""
    gx, gy = O.green_grad()
    for g in (gx, gy):
        T, q = O.importance_table(g)
        assert T[0] == 0 and (np.diff(T.astype(np.int64)) > 0).all()
        assert (q > 0).all() and q.sum() == 1.0
        a = np.abs(g) / np.abs(g).sum()
        assert np.abs(q - a).max() <= 2.0 ** -31
        # the panels a side's normal sees most are drawn most: |gx| peaks on
        # the left/right sides, |gy| on the top/bottom sides
    qx, qy = O.importance_table(gx)[1], O.importance_table(gy)[1]
    side = np.repeat(np.arange(4), PER)
    assert qx[side % 2 == 1].sum() > qx[side % 2 == 0].sum()
    assert qy[side % 2 == 0].sum() > qy[side % 2 == 1].sum()


def test_importance_sampling_same_capacitance_lower_variance():
    """The importance-sampled first hop changes which panel is drawn and the
    weight's divisor, not the expectation: on the parallel plates both
    estimators agree within their Monte Carlo error, the weights still sum
    to zero, and the per-walk variance of the self term drops."""
    rects = _plates()
    ix = O.Index(rects, O.bounds(rects), O.EXACT)
    n = 40_000
    hp, _, wp = ix.walk_weighted(0, n, seed=3)
    hi, hopi, wi = ix.walk_weighted(0, n, seed=4, importance=True)
    assert np.isfinite(wi).all() and (hopi >= 0).all()
    for j in (0, 1):
        xp = np.where(hp == j, wp, 0.0)
        xi = np.where(hi == j, wi, 0.0)
        sig = np.hypot(xp.std(), xi.std()) / np.sqrt(n)
        assert abs(xp.mean() - xi.mean()) < 4 * sig, (j, xp.mean(), xi.mean(), sig)
    assert abs(wi.mean()) < 4 * wi.std() / np.sqrt(n)
    c01 = wi[hi == 1].sum() / n
    assert -26.0 < c01 < -19.0, c01
    v_plain = np.where(hp == 0, wp, 0.0).var()
    v_imp = np.where(hi == 0, wi, 0.0).var()
    assert v_imp < 0.9 * v_plain, (v_imp, v_plain)  # about 0.8 on these and the fixture layouts
