"""Device walks against the oracle.

Replay: with the same Philox stream, every GPU walk (both modes) ends on the
same rect after the same number of hops as the oracle's restatement of the
reference walk over an exact index (reference CPU tree with the single-path
query fixed).  Statistics: independent walk ranges agree within 3 sigma, and
the reference's own (bug-compatible) CPU tree is compared where its single-path
query does not change the answer.  Sizes up to 10^6 rects check
size-independent properties (conservation, determinism, launch-shape
independence)."""
import numpy as np
import pytest

import frw
import oracle_ffi as O
from conftest import FIXTURES
from gpu_common import load_fixture, oracle_index, torch_dev

pytestmark = pytest.mark.gpu


def gpu_walks(tree, master, n, mode, seed=1, first=0, block=0, grid=0, job=0, max_hops=0):
    torch = torch_dev()
    nr = tree.info().n_rects
    counts = torch.zeros(nr + 2, dtype=torch.int64, device="cuda")
    hops = torch.zeros(1, dtype=torch.int64, device="cuda")
    hit = torch.full((n,), -7, dtype=torch.int32, device="cuda")
    hp = torch.full((n,), -7, dtype=torch.int32, device="cuda")
    tree.walk_async(master, n, counts, hops, hit, hp, seed=seed, mode=mode, first_walk=first,
                    block=block, grid=grid, job=job, max_hops=max_hops)
    torch.cuda.synchronize()
    return counts.cpu().numpy(), int(hops.item()), hit.cpu().numpy(), hp.cpu().numpy()


@pytest.mark.parametrize("name,master", FIXTURES)
@pytest.mark.parametrize("mode", [frw.MODE_EXACT, frw.MODE_BISECT])
def test_walk_replays_oracle(name, master, mode):
    lay, rects, bounds = load_fixture(name)
    tree = frw.Tree(lay)
    n = 3000 if len(rects) < 5000 else 1000
    counts, hops, hit, hp = gpu_walks(tree, master, n, mode, seed=1)
    df = oracle_index(rects, bounds, O.REFTREE_DFS)
    oc, ohops, ohit, ohp = df.walk(master, n, seed=1, per_walk=True)
    assert np.array_equal(hit, ohit), np.nonzero(hit != ohit)[0][:10]
    assert np.array_equal(hp, ohp)
    assert hops == ohops
    assert np.array_equal(counts, oc)


def test_walk_replay_seed_job_offset():
    lay, rects, bounds = load_fixture("test100.txt")
    tree = frw.Tree(lay, leaf_cap=4)
    df = oracle_index(rects, bounds, O.REFTREE_DFS)
    for seed, job, first in [(7, 0, 0), (1, 5, 123456), ((3 << 32) | 9, 2, 10)]:
        c, h, hit, hp = gpu_walks(tree, 3, 1500, frw.MODE_EXACT, seed=seed, first=first, job=job)
        oc, oh, ohit, ohp = df.walk(3, 1500, seed=seed, job=job, first_walk=first, per_walk=True)
        assert np.array_equal(hit, ohit) and np.array_equal(hp, ohp)


def test_walk_ids_beyond_32_bits():
    """Walk ids run to 2^48: a range across 2^32 replays the oracle walk for
    walk (the id's high word enters the Philox domain word), walks 2^32 + k
    are not walks k again, the top of the id range replays too, and a range
    past 2^48 is refused before anything runs."""
    lay, rects, bounds = load_fixture("test100.txt")
    tree = frw.Tree(lay, leaf_cap=4)
    df = oracle_index(rects, bounds, O.REFTREE_DFS)
    first = (1 << 32) - 700
    c, h, hit, hp = gpu_walks(tree, 3, 1400, frw.MODE_EXACT, seed=5, first=first)
    oc, oh, ohit, ohp = df.walk(3, 1400, seed=5, first_walk=first, per_walk=True)
    assert np.array_equal(hit, ohit) and np.array_equal(hp, ohp) and h == oh
    _, _, hit0, hp0 = gpu_walks(tree, 3, 700, frw.MODE_EXACT, seed=5, first=0)
    assert not (np.array_equal(hit[700:], hit0) and np.array_equal(hp[700:], hp0))
    top = (1 << 48) - 300
    c, h, hit, hp = gpu_walks(tree, 3, 300, frw.MODE_BISECT, seed=5, first=top)
    oc, oh, ohit, ohp = df.walk(3, 300, seed=5, first_walk=top, per_walk=True)
    assert np.array_equal(hit, ohit) and np.array_equal(hp, ohp)
    with pytest.raises(frw.FrwError) as e:
        gpu_walks(tree, 3, 301, frw.MODE_EXACT, seed=5, first=top)
    assert e.value.code == frw.E_RANGE


def test_walk_launch_shape_invariance():
    lay, rects, bounds = load_fixture("test1000.txt")
    tree = frw.Tree(lay)
    ref = gpu_walks(tree, 0, 20000, frw.MODE_EXACT)
    for block, grid in [(64, 1), (128, 7), (256, 0), (192, 33)]:
        got = gpu_walks(tree, 0, 20000, frw.MODE_EXACT, block=block, grid=grid)
        assert np.array_equal(got[0], ref[0]) and got[1] == ref[1]
        assert np.array_equal(got[2], ref[2]) and np.array_equal(got[3], ref[3])


def test_walk_max_hops_stuck_bin():
    lay, rects, bounds = load_fixture("test")
    tree = frw.Tree(lay)
    c, h, hit, hp = gpu_walks(tree, 10, 2000, frw.MODE_EXACT, max_hops=3)
    df = oracle_index(rects, bounds, O.REFTREE_DFS)
    oc, oh, ohit, ohp = df.walk(10, 2000, seed=1, max_hops=3, per_walk=True)
    assert np.array_equal(hit, ohit) and np.array_equal(hp, ohp)
    assert c[len(rects) + 1] == np.sum(hit == O.HIT_STUCK) > 0
    assert hp.max() <= 3


def _frac_other(counts, master, n):
    return 1.0 - counts[master] / n


@pytest.mark.parametrize("name,master", [("test1", 0), ("test100.txt", 0), ("test1000.txt", 0)])
def test_walk_statistics_3sigma(name, master):
    """Disjoint walk ranges: GPU (10^6 walks) vs oracle (10^5 walks) hit
    fractions within 3 sigma, per conductor and for f = 1 - P(master)."""
    lay, rects, bounds = load_fixture(name)
    tree = frw.Tree(lay)
    N = 10 ** 6
    counts, _, _, _ = gpu_walks(tree, master, N, frw.MODE_EXACT, seed=1)
    df = oracle_index(rects, bounds, O.REFTREE_DFS)
    M = 10 ** 5
    oc, _, _, _ = df.walk(master, M, seed=1, first_walk=N)
    p1, p2 = counts / N, oc / M
    sig = np.sqrt(p1 * (1 - p1) / N + p2 * (1 - p2) / M)
    big = (p1 > 1e-3) | (p2 > 1e-3)
    z = np.abs(p1 - p2)[big] / np.maximum(sig[big], 1e-12)
    assert z.max() < 4.5, z.max()     # max over several bins
    f1, f2 = _frac_other(counts, master, N), _frac_other(oc, master, M)
    assert abs(f1 - f2) < 3 * np.sqrt(f1 * (1 - f1) / N + f2 * (1 - f2) / M)


def test_walk_vs_reference_code_3sigma():
    """On tests/test1 the reference's single-path CPU query never changes an
    answer, so the reference code's own walks (oracle/_ref, 5*10^4 walks) and
    the GPU (10^6 walks) must agree within 3 sigma."""
    if not O.ref_available():
        pytest.skip("oracle/_ref not built")
    lay, rects, bounds = load_fixture("test1")
    tree = frw.Tree(lay)
    N = 10 ** 6
    counts, _, _, _ = gpu_walks(tree, 0, N, frw.MODE_EXACT, seed=1)
    ref = O.RefTree(rects.astype(np.float64), bounds)
    M = 50_000
    hit, hops, total, other = ref.walk(rects[0].astype(np.float64), M, seed=3)
    f1, f2 = _frac_other(counts, 0, N), other / M
    assert abs(f1 - f2) < 3 * np.sqrt(f1 * (1 - f1) / N + f2 * (1 - f2) / M)


def test_walk_c1_config_replay():
    """100 rects in a 10^4 box (sides U[10,200]), rect 0 master, 10^5 walks."""
    lay = frw.Layout.generate(100, 10_000, 10, 200, seed=1)
    rects = lay.rects()
    bounds = tuple(float(v) for v in lay.bounds())
    tree = frw.Tree(lay)
    c, h, hit, hp = gpu_walks(tree, 0, 10 ** 5, frw.MODE_EXACT)
    df = oracle_index(rects, bounds, O.REFTREE_DFS)
    oc, oh, ohit, ohp = df.walk(0, 10 ** 5, seed=1, per_walk=True)
    assert np.array_equal(hit, ohit) and np.array_equal(hp, ohp) and h == oh


def test_walk_large_layout_properties():
    """10^6 rects: conservation, determinism, mode equivalence, replay of a
    walk subset."""
    lay = frw.Layout.generate(10 ** 6, 10 ** 7, 5, 300, seed=1, cluster_frac=0.2, sigma=1e5,
                              n_clusters=100)
    rects = lay.rects()
    bounds = tuple(float(v) for v in lay.bounds())
    tree = frw.Tree(lay, leaf_cap=16)
    n = 200_000
    a = gpu_walks(tree, 0, n, frw.MODE_EXACT)
    b = gpu_walks(tree, 0, n, frw.MODE_EXACT, block=128)
    assert a[0].sum() == n and a[1] == int(a[3].sum())
    assert np.array_equal(a[0], b[0]) and np.array_equal(a[2], b[2])
    c = gpu_walks(tree, 0, 3000, frw.MODE_BISECT)
    assert np.array_equal(c[2], a[2][:3000]) and np.array_equal(c[3], a[3][:3000])
    df = oracle_index(rects, bounds, O.REFTREE_DFS)
    oc, oh, ohit, ohp = df.walk(0, 400, seed=1, per_walk=True)
    assert np.array_equal(a[2][:400], ohit) and np.array_equal(a[3][:400], ohp)
