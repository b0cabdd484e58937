"""Helpers shared by the GPU parity tests."""
import numpy as np

import frw
import oracle_ffi as O
from conftest import layout_path


def torch_dev():
    import torch
    if not torch.cuda.is_available():
        raise RuntimeError("GPU test run without a HIP device")
    return torch


def load_fixture(name):
    lay = frw.Layout.load(layout_path(name))
    return lay, lay.rects(), tuple(float(v) for v in lay.bounds())


def oracle_index(rects, bounds, kind=O.REFTREE_DFS):
    return O.Index(np.asarray(rects, dtype=np.float64), bounds, kind)


def query_points(rects, bounds, n_random, seed):
    """Random points plus adversarial ones: rect corners and edges, points at
    L-inf distance 0.5..2.5 from rects (the p+-1 / p+-2 shell), bounds edges."""
    rng = np.random.default_rng(seed)
    b = bounds
    pts = [np.column_stack([rng.uniform(b[0], b[2], n_random), rng.uniform(b[1], b[3], n_random)])]
    r = np.asarray(rects, dtype=np.float64)
    k = min(len(r), max(1, n_random // 8))
    sel = r[rng.choice(len(r), k, replace=False)]
    pts.append(sel[:, [0, 1]])
    pts.append(sel[:, [2, 3]])
    pts.append(np.column_stack([sel[:, 0], (sel[:, 1] + sel[:, 3]) / 2]))
    pts.append(np.column_stack([(sel[:, 0] + sel[:, 2]) / 2, sel[:, 3]]))
    for d in (0.5, 1.0, 1.5, 2.0, 2.5):
        pts.append(np.column_stack([sel[:, 2] + d, (sel[:, 1] + sel[:, 3]) / 2]))
        pts.append(np.column_stack([(sel[:, 0] + sel[:, 2]) / 2, sel[:, 1] - d]))
    pts.append(np.column_stack([rng.uniform(b[0], b[2], 16), np.full(16, b[1])]))
    pts.append(np.column_stack([np.full(16, b[2]), rng.uniform(b[1], b[3], 16)]))
    return np.ascontiguousarray(np.concatenate(pts), dtype=np.float64)


def to_dev(torch, a, dtype=None):
    t = torch.from_numpy(np.ascontiguousarray(a))
    if dtype is not None:
        t = t.to(dtype)
    return t.cuda()
