"""Edge layouts walked on the device against the oracle, walk for walk: one
rect, degenerate (zero-width and point) rects, overlapping and nested rects
(the point query's deepest-containment and lowest-id rules decide the hit),
touching rects, a layout narrower than 100 (no bounds padding, the master on
the bounds), coordinates near 2^27, and ragged walk counts.  Exact,
bisection, weighted and reference-compatible modes."""
import numpy as np
import pytest

import frw
import oracle_ffi as O
from gpu_common import torch_dev

pytestmark = pytest.mark.gpu


def _layouts():
    big = 1 << 27
    return {
        "single": [[0, 0, 100, 40]],
        "degenerate": [[0, 0, 1000, 10], [500, 200, 500, 400], [800, 600, 800, 600], [0, 900, 1000, 910]],
        "nested": [[0, 0, 1000, 1000], [100, 100, 400, 400], [200, 200, 300, 300], [600, 600, 900, 900],
                   [600, 600, 900, 900], [2000, 0, 2100, 1000]],
        "touching": [[0, 0, 100, 100], [100, 0, 200, 100], [0, 100, 100, 200], [200, 200, 300, 300],
                     [300, 300, 400, 400]],
        "narrow": [[0, 0, 20, 5], [0, 10, 20, 15], [25, 0, 30, 15]],
        # coordinates near 2^27, bounds about 2^26 wide (the build takes widths below 2^28)
        "far": [[big - 5000, big - 5000, big - 4000, big - 4900], [big - 3000, big - 4000, big - 2000, big - 3900],
                [big - (1 << 26), big - (1 << 26), big - (1 << 26) + 100, big - (1 << 26) + 100],
                [big - 10, big - 10, big, big]],
    }


def _walks(tree, master, n, mode, seed=1, ref=False):
    torch = torch_dev()
    nr = tree.info().n_rects
    c = torch.zeros(nr + 2, dtype=torch.int64, device="cuda")
    h = torch.zeros(1, dtype=torch.int64, device="cuda")
    hit = torch.full((n,), -7, dtype=torch.int32, device="cuda")
    hp = torch.full((n,), -7, dtype=torch.int32, device="cuda")
    tree.walk_async(master, n, c, h, hit, hp, seed=seed, mode=mode)
    torch.cuda.synchronize()
    return c.cpu().numpy(), int(h.item()), hit.cpu().numpy(), hp.cpu().numpy()


@pytest.mark.parametrize("name", sorted(_layouts()))
@pytest.mark.parametrize("mode", [frw.MODE_EXACT, frw.MODE_BISECT])
def test_edge_layout_walks_replay_oracle(name, mode):
    rects = np.array(_layouts()[name], dtype=np.int32)
    lay = frw.Layout.from_rects(rects)
    b = tuple(float(v) for v in lay.bounds())
    tree = frw.Tree(lay)
    ix = O.Index(rects.astype(np.float64), b, O.EXACT)
    for master in range(len(rects)):
        for n in (1, 63, 1001):
            c, h, hit, hp = _walks(tree, master, n, mode, seed=master + 7)
            oc, oh, ohit, ohp = ix.walk(master, n, seed=master + 7, per_walk=True)
            assert np.array_equal(hit, ohit), (name, master, n, np.nonzero(hit != ohit)[0][:5])
            assert np.array_equal(hp, ohp) and h == oh and np.array_equal(c, oc)


@pytest.mark.parametrize("name", sorted(_layouts()))
def test_edge_layout_weighted_and_refcompat(name):
    torch = torch_dev()
    rects = np.array(_layouts()[name], dtype=np.int32)
    lay = frw.Layout.from_rects(rects)
    b = tuple(float(v) for v in lay.bounds())
    tree = frw.Tree(lay, ref_compat=True)
    ex = O.Index(rects.astype(np.float64), b, O.EXACT)
    rt = O.Index(rects.astype(np.float64), b, O.REFTREE)
    nr, n = len(rects), 777
    for master in range(nr):
        c = torch.zeros(nr + 2, dtype=torch.int64, device="cuda")
        h = torch.zeros(1, dtype=torch.int64, device="cuda")
        ws = torch.zeros(nr + 2, dtype=torch.float64, device="cuda")
        wq = torch.zeros(nr + 2, dtype=torch.float64, device="cuda")
        w = torch.zeros(n, dtype=torch.float64, device="cuda")
        hit = torch.full((n,), -7, dtype=torch.int32, device="cuda")
        tree.walk_weighted_async(master, n, c, h, ws, wq, w_out=w, hit_out=hit, seed=3, flags=frw.WEIGHT_IMPORTANCE)
        torch.cuda.synchronize()
        ohit, ohp, ow = ex.walk_weighted(master, n, seed=3, importance=True)
        assert np.array_equal(hit.cpu().numpy(), ohit), (name, master)
        np.testing.assert_allclose(w.cpu().numpy(), ow, rtol=1e-13, atol=0)
        rc, rh, rhit, rhp = _walks(tree, master, n, frw.MODE_REFCOMPAT, seed=5)
        oc, oh, ohit2, ohp2 = rt.walk(master, n, seed=5, per_walk=True)
        assert np.array_equal(rhit, ohit2) and np.array_equal(rhp, ohp2), (name, master)
