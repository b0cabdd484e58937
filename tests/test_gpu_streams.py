"""Launches that share one handle's device scratch, queued on different
streams without a host wait between them: each walk on one tree, and each
launch of one batch, must equal its result when run alone (the handle orders
its launches).  Also the argument checks of the batched entry points."""
import numpy as np
import pytest

import frw
from gpu_common import torch_dev

pytestmark = pytest.mark.gpu


def test_two_streams_one_tree_equal_serial():
    torch = torch_dev()
    lay = frw.Layout.generate(5_000, 70_000, 10, 500, seed=21)
    tree = frw.Tree(lay)
    nr = tree.info().n_rects
    jobs = [(0, 300_000, 5), (7, 120_000, 6), (3, 200_000, 8)]

    def outputs(n):
        return (torch.zeros(nr + 2, dtype=torch.int64, device="cuda"), torch.zeros(1, dtype=torch.int64, device="cuda"),
                torch.zeros(n, dtype=torch.int32, device="cuda"), torch.zeros(n, dtype=torch.int32, device="cuda"))

    serial = []
    for m, n, seed in jobs:
        o = outputs(n)
        tree.walk_async(m, n, *o, seed=seed)
        torch.cuda.synchronize()
        serial.append([t.cpu().numpy() for t in o])
    streams = [torch.cuda.Stream() for _ in jobs]
    outs = [outputs(n) for _, n, _ in jobs]
    torch.cuda.synchronize()
    print("three walks on one tree, three streams", flush=True)
    for (m, n, seed), st, o in zip(jobs, streams, outs):
        tree.walk_async(m, n, *o, seed=seed, stream=st.cuda_stream)
    torch.cuda.synchronize()
    for want, o in zip(serial, outs):
        for w, g in zip(want, o):
            assert np.array_equal(g.cpu().numpy(), w)


def test_two_streams_one_batch_equal_serial():
    torch = torch_dev()
    trees = [frw.Tree(frw.Layout.generate(n, int(1000 * np.sqrt(n)), 10, 500, seed=30 + i))
             for i, n in enumerate((300, 2_000))]
    jobs = [{"tree": 0, "master": 0, "n_walks": 60_000}, {"tree": 1, "master": 4, "n_walks": 90_000}]
    b = frw.Batch(trees, jobs)
    size = b.counts_size()
    seeds = (3, 4)
    serial = []
    for s in seeds:
        c = torch.zeros(size, dtype=torch.int64, device="cuda")
        h = torch.zeros(2, dtype=torch.int64, device="cuda")
        b.walk(c, h, seed=s)
        torch.cuda.synchronize()
        serial.append((c.cpu().numpy(), h.cpu().numpy()))
    streams = [torch.cuda.Stream() for _ in seeds]
    outs = [(torch.zeros(size, dtype=torch.int64, device="cuda"), torch.zeros(2, dtype=torch.int64, device="cuda"))
            for _ in seeds]
    torch.cuda.synchronize()
    print("two launches of one batch, two streams", flush=True)
    for s, st, (c, h) in zip(seeds, streams, outs):
        b.walk(c, h, seed=s, stream=st.cuda_stream)
    torch.cuda.synchronize()
    for (wc, wh), (c, h) in zip(serial, outs):
        assert np.array_equal(c.cpu().numpy(), wc) and np.array_equal(h.cpu().numpy(), wh)


def test_batched_entry_points_reject_bad_max_hops():
    torch = torch_dev()
    tree = frw.Tree(frw.Layout.generate(200, 20_000, 10, 200, seed=5))
    b = frw.Batch([tree], [{"tree": 0, "master": 0, "n_walks": 100}])
    c = torch.zeros(b.counts_size(), dtype=torch.int64, device="cuda")
    h = torch.zeros(1, dtype=torch.int64, device="cuda")
    for bad in (1 << 32, -1):
        with pytest.raises(frw.FrwError) as e:
            b.walk(c, h, max_hops=bad)
        assert e.value.code == frw.E_RANGE
    mlp = frw.Mlp(tree, seed=1, max_level=10)
    b.set_proposal([mlp])
    with pytest.raises(frw.FrwError) as e:
        b.walk_learned(mlp, c, h, max_hops=1 << 32)
    assert e.value.code == frw.E_RANGE
    with pytest.raises(frw.FrwError) as e:
        b.autotune_learned(mlp, max_hops=1 << 33)
    assert e.value.code == frw.E_RANGE
    with pytest.raises(frw.FrwError) as e:
        frw.walk_learned(tree, mlp, 0, 10, c[:202], h, max_hops=1 << 32)
    assert e.value.code == frw.E_RANGE
    torch.cuda.synchronize()
    assert int(c.sum().item()) == 0  # nothing ran
