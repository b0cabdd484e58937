"""frw_walk_noalloc (the entry point that allocates nothing: every scratch
buffer in the library's fixed device array) against frw_walk, the entry point
it varies, through the C ABI: identical counts and summaries on consecutive
calls, the second on a smaller layout and fewer walks (the shared array keeps
the previous call's contents, so its bins must be cleared), in exact and
bisection modes; and the size limit."""
import numpy as np
import pytest

import frw
from conftest import layout_path
from gpu_common import torch_dev

pytestmark = pytest.mark.gpu


def same_summary(a, b):
    return (a.n_walks, a.hops, a.n_master, a.n_escaped, a.n_stuck, a.fraction_other) == \
        (b.n_walks, b.hops, b.n_master, b.n_escaped, b.n_stuck, b.fraction_other)


@pytest.mark.parametrize("mode", [frw.MODE_EXACT, frw.MODE_BISECT])
def test_noalloc_equals_walk_on_consecutive_calls(mode):
    torch = torch_dev()
    torch.cuda.set_device(0)
    big = frw.Tree(frw.Layout.generate(20_000, 140_000, 10, 500, seed=4))
    small = frw.Tree(frw.Layout.load(layout_path("test")))
    runs = [(big, 0, 200_000 if mode == frw.MODE_EXACT else 20_000, 11),
            (small, 10, 3_000, 12),
            (big, 17, 5_000, 13)]
    for tree, master, n, seed in runs:
        print(f"noalloc mode {mode}: {tree.info().n_rects} rects, master {master}, {n} walks", flush=True)
        want_c, want_s = tree.walk(master, n, seed=seed, mode=mode, first_walk=100)
        got_c, got_s = tree.walk_noalloc(master, n, seed=seed, mode=mode, first_walk=100)
        assert np.array_equal(got_c, want_c)
        assert same_summary(got_s, want_s)
        assert got_c.sum() == n


def test_noalloc_on_a_stream_and_size_limit():
    torch = torch_dev()
    torch.cuda.set_device(0)
    tree = frw.Tree(frw.Layout.load(layout_path("test1000.txt")))
    st = torch.cuda.Stream()
    want_c, want_s = tree.walk(0, 50_000, seed=3)
    got_c, got_s = tree.walk_noalloc(0, 50_000, seed=3, stream=st.cuda_stream)
    assert np.array_equal(got_c, want_c) and same_summary(got_s, want_s)
    with pytest.raises(frw.FrwError) as e:
        tree.walk_noalloc(0, 10, mode=frw.MODE_WEIGHTED)
    assert e.value.code == frw.E_INVALID
    huge = frw.Layout.generate((1 << 20) + 1, 200_000_000, 5, 20, seed=2)
    big = frw.Tree(huge)
    with pytest.raises(frw.FrwError) as e:
        big.walk_noalloc(0, 10)
    assert e.value.code == frw.E_RANGE and "at most" in str(e.value)
