"""Adaptive stopping (rounds sized by n_rem = n (err^2 / goal^2 - 1)):
the rounds take consecutive walk ids, so the tallies equal one launch of the
same total; the reported estimate and its 3-sigma error are recomputed from
those tallies; the goal is met, or max_walks ends the loop."""
import numpy as np
import pytest

import frw
from gpu_common import load_fixture

pytestmark = pytest.mark.gpu


def _f_err(counts, master, n):
    f = (n - counts[master]) / n
    return f, 3 * np.sqrt(f * (1 - f) / n) / f


@pytest.mark.parametrize("mode", [frw.MODE_EXACT, frw.MODE_BISECT, frw.MODE_REFCOMPAT])
def test_adaptive_hit_fraction_equals_one_launch(mode):
    lay, rects, bounds = load_fixture("test")
    tree = frw.Tree(lay, ref_compat=mode == frw.MODE_REFCOMPAT)
    r, counts, _, _ = tree.walk_adaptive(10, rel_err=0.01, mode=mode, first_round=4096, seed=5)
    assert r.converged == 1 and r.rounds >= 2
    f, err = _f_err(counts, 10, r.n_walks)
    assert err <= 0.01 and abs(r.value - f) < 1e-15 and abs(r.rel_err_3sigma - err) < 1e-12
    one, s = tree.walk(10, r.n_walks, seed=5, mode=mode)
    assert np.array_equal(counts, one) and s.hops == r.hops
    # the last round was sized by eq. 32: the total stays near the walks the
    # goal needs (N_1% = 9e4 (1 - f) / f), not a multiple of it
    assert r.n_walks < 1.6 * 9e4 * (1 - f) / f + 4096


def test_adaptive_weighted_self_term():
    lay, rects, bounds = load_fixture("test100.txt")
    tree = frw.Tree(lay)
    r, counts, wsum, wsq = tree.walk_adaptive(0, rel_err=0.05, mode=frw.MODE_WEIGHTED, first_round=8192,
                                              seed=3, flags=frw.WEIGHT_IMPORTANCE)
    n = r.n_walks
    assert r.converged == 1 and counts.sum() == n
    c = wsum[0] / n
    err = 3 * np.sqrt(max(wsq[0] / n - c * c, 0) / n) / abs(c)
    assert abs(r.value - c) <= 1e-12 * abs(c) and err <= 0.05
    oc, ows, owq, oh = tree.walk_weighted(0, n, seed=3, flags=frw.WEIGHT_IMPORTANCE)
    assert np.array_equal(counts, oc) and oh == r.hops
    np.testing.assert_allclose(wsum, ows, rtol=1e-9, atol=1e-9 * np.abs(ows).max())


def test_adaptive_stops_at_max_walks_and_rejects_bad_arguments():
    lay, rects, bounds = load_fixture("test")
    tree = frw.Tree(lay)
    r, counts, _, _ = tree.walk_adaptive(10, rel_err=1e-4, first_round=2000, max_walks=10_000)
    assert r.converged == 0 and r.n_walks == 10_000 and counts.sum() == 10_000
    with pytest.raises(frw.FrwError):
        tree.walk_adaptive(10, rel_err=0.0)
    with pytest.raises(frw.FrwError):
        tree.walk_adaptive(10, mode=frw.MODE_LEARNED)
