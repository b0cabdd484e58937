"""The records per hop that price the walk kernel's roofline (node records
32 B, rect records 16 B) are counted by the instrumented device walk; here
they equal, exactly, the counts of a sequential host replay of the same
Philox walks on the same flat tree (tests/hostsim, the device traversal code
built for the CPU), on the default bench layout and at 10^6 rects."""
import pytest

import frw
import test_hostsim as HS

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("mode,n", [(frw.MODE_EXACT, 20_000), (frw.MODE_BISECT, 3_000)])
def test_traffic_counts_equal_host_replay(mode, n):
    lay = frw.Layout.generate(10_000, 100_000, 10, 500, seed=1)
    tree = frw.Tree(lay)
    got = tree.walk_traffic(0, n, seed=3, mode=mode)
    ht = HS.HostTree(lay, cap=tree.info().leaf_cap)
    assert got == ht.walk_counted(0, n, mode=mode, seed=3, use_hint=mode == frw.MODE_EXACT)


def test_traffic_counts_equal_host_replay_at_scale():
    lay = frw.Layout.generate(10 ** 6, 10 ** 7, 5, 300, seed=1, cluster_frac=0.2, sigma=1e5, n_clusters=100)
    tree = frw.Tree(lay)
    got = tree.walk_traffic(0, 5_000, seed=4)
    assert got == HS.HostTree(lay, cap=tree.info().leaf_cap).walk_counted(0, 5_000, seed=4)
