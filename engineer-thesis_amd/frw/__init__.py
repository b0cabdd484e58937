"""Python binding of the gfx950 FRW library (libfrw.so, C ABI in include/frw.h).

A thin ctypes layer: device buffers are torch tensors on ``cuda`` (HIP) passed
by ``data_ptr()``; nothing here computes on the host.  If libfrw.so is not
built, importing this module raises.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libfrw.so")

MODE_EXACT, MODE_BISECT, MODE_LEARNED, MODE_WEIGHTED, MODE_REFCOMPAT = 0, 1, 2, 3, 4
TREE_MORTON, TREE_REFCOMPAT = 0, 1
WEIGHT_IMPORTANCE = 1
GEN_UNIFORM, GEN_CLUSTERED = 0, 1
NSAMPLE = 200

OK, E_INVALID, E_IO, E_PARSE, E_LAYER, E_RANGE, E_DEVICE, E_NOMEM, E_KERNEL = 0, -1, -2, -3, -4, -5, -6, -7, -8


class FrwError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"frw error {code}: {msg}")
        self.code = code


class GenParams(C.Structure):
    _fields_ = [("kind", C.c_int32), ("box", C.c_int32), ("n", C.c_int64),
                ("side_min", C.c_int32), ("side_max", C.c_int32),
                ("cluster_frac", C.c_double), ("sigma", C.c_double),
                ("n_clusters", C.c_int32), ("reserved", C.c_int32), ("seed", C.c_uint64)]


class TreeParams(C.Structure):
    _fields_ = [("leaf_cap", C.c_int32), ("kind", C.c_int32), ("reserved", C.c_int32 * 6)]


class TreeInfo(C.Structure):
    _fields_ = [("n_rects", C.c_int64), ("n_nodes", C.c_int64), ("depth", C.c_int32),
                ("leaf_cap", C.c_int32), ("bounds", C.c_int32 * 4), ("morton_bits", C.c_int32),
                ("reserved", C.c_int32), ("build_ms", C.c_double)]


class WalkParams(C.Structure):
    _fields_ = [("master", C.c_int32), ("mode", C.c_int32), ("n_walks", C.c_int64),
                ("first_walk", C.c_int64), ("seed", C.c_uint64), ("job", C.c_uint32),
                ("block", C.c_int32), ("grid", C.c_int32), ("reserved", C.c_int32),
                ("max_hops", C.c_int64)]


class WalkOutputs(C.Structure):
    _fields_ = [("counts", C.c_void_p), ("hops", C.c_void_p), ("hit_out", C.c_void_p),
                ("hops_out", C.c_void_p)]


class BatchJob(C.Structure):
    _fields_ = [("tree", C.c_int32), ("master", C.c_int32), ("n_walks", C.c_int64),
                ("first_walk", C.c_int64), ("job", C.c_uint32), ("reserved", C.c_int32)]


class BatchParams(C.Structure):
    _fields_ = [("mode", C.c_int32), ("block", C.c_int32), ("grid", C.c_int32),
                ("use_graph", C.c_int32), ("seed", C.c_uint64), ("max_hops", C.c_int64)]


class AdaptiveResult(C.Structure):
    _fields_ = [("n_walks", C.c_int64), ("hops", C.c_int64), ("rounds", C.c_int32), ("converged", C.c_int32),
                ("value", C.c_double), ("rel_err_3sigma", C.c_double), ("wall_ms", C.c_double)]


class WalkSummary(C.Structure):
    _fields_ = [("n_walks", C.c_int64), ("hops", C.c_int64), ("n_master", C.c_int64),
                ("n_escaped", C.c_int64), ("n_stuck", C.c_int64), ("fraction_other", C.c_double),
                ("walk_ms", C.c_double)]


#: Every symbol include/frw.h declares (checked by the test-suite).
EXPORTS = [
    "frw_last_error_string", "frw_version", "frw_layout_load", "frw_layout_layer_count",
    "frw_layout_from_rects", "frw_layout_generate", "frw_layout_save", "frw_layout_size",
    "frw_layout_rects", "frw_layout_bounds", "frw_layout_free", "frw_green_thresholds",
    "frw_tree_build", "frw_tree_info_get", "frw_tree_export", "frw_tree_free",
    "frw_query_collide", "frw_query_point", "frw_query_nearest", "frw_query_maxsquare",
    "frw_gauss_surface", "frw_walk_async", "frw_walk", "frw_walk_traffic", "frw_write_results",
    "frw_batch_create", "frw_batch_counts_size", "frw_batch_counts_offset", "frw_batch_walk",
    "frw_batch_autotune", "frw_batch_check", "frw_batch_free",
    "frw_mlp_create", "frw_mlp_free", "frw_mlp_param_count", "frw_mlp_get_params",
    "frw_mlp_set_params", "frw_mlp_occupancy", "frw_mlp_features", "frw_mlp_forward",
    "frw_mlp_gradient", "frw_mlp_set_samples", "frw_mlp_train", "frw_walk_harvest",
    "frw_walk_learned", "frw_walk_learned_async", "frw_walk_learned_traffic",
    "frw_mlp_levels", "frw_mlp_certify", "frw_mlp_add_samples", "frw_batch_set_proposal",
    "frw_batch_walk_learned", "frw_batch_autotune_learned", "frw_walk_weighted_async", "frw_walk_weighted", "frw_walk_adaptive", "frw_weight_tables", "frw_probe_read_bandwidth",
    "frw_query_refcompat", "frw_refcompat_gauss_surface", "frw_refcompat_info", "frw_refcompat_walk_from",
    "frw_layout_set_bounds", "frw_walk_noalloc", "frw_batch_harvest",
]

_lib = None


def lib():
    """Load libfrw.so (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError(f"{LIB_PATH} is not built; run `make -C engineer-thesis_amd`")
    # One HIP runtime per process: when torch is present, let it load its
    # libamdhip64 first so libfrw.so binds to the same instance (same SONAME)
    # instead of pulling in a second copy from /opt/rocm.
    try:
        import torch  # noqa: F401
    except ImportError:
        pass
    L = C.CDLL(LIB_PATH)
    vp, i32, i64, u64 = C.c_void_p, C.c_int32, C.c_int64, C.c_uint64
    L.frw_last_error_string.restype = C.c_char_p
    L.frw_version.restype = C.c_char_p
    L.frw_layout_load.argtypes = [C.c_char_p, i32, C.POINTER(vp)]
    L.frw_layout_layer_count.argtypes = [C.c_char_p, C.POINTER(i32)]
    L.frw_layout_from_rects.argtypes = [vp, i64, C.POINTER(vp)]
    L.frw_layout_generate.argtypes = [C.POINTER(GenParams), C.POINTER(vp)]
    L.frw_layout_save.argtypes = [vp, C.c_char_p]
    L.frw_layout_size.restype = i64
    L.frw_layout_size.argtypes = [vp]
    L.frw_layout_rects.argtypes = [vp, vp]
    L.frw_layout_bounds.argtypes = [vp, vp]
    L.frw_layout_free.argtypes = [vp]
    L.frw_layout_free.restype = None
    L.frw_green_thresholds.argtypes = [vp]
    L.frw_tree_build.argtypes = [vp, C.POINTER(TreeParams), vp, C.POINTER(vp)]
    L.frw_tree_info_get.argtypes = [vp, C.POINTER(TreeInfo)]
    L.frw_tree_export.argtypes = [vp, vp, vp, vp]
    L.frw_tree_free.argtypes = [vp]
    L.frw_tree_free.restype = None
    L.frw_query_collide.argtypes = [vp, vp, i64, vp, vp]
    L.frw_query_refcompat.argtypes = [vp, vp, vp, i64, vp, vp, vp, vp, vp]
    L.frw_refcompat_gauss_surface.argtypes = [vp, vp, C.c_double, vp]
    L.frw_refcompat_walk_from.argtypes = [vp, vp, C.POINTER(WalkParams), C.POINTER(WalkOutputs), vp]
    L.frw_layout_set_bounds.argtypes = [vp, vp]
    L.frw_refcompat_info.argtypes = [vp, vp]
    L.frw_query_point.argtypes = [vp, vp, i64, vp, vp]
    L.frw_query_nearest.argtypes = [vp, vp, i64, vp, vp]
    L.frw_query_maxsquare.argtypes = [vp, vp, i64, i32, vp, vp]
    L.frw_gauss_surface.argtypes = [vp, i32, vp]
    L.frw_walk_async.argtypes = [vp, C.POINTER(WalkParams), C.POINTER(WalkOutputs), vp]
    L.frw_walk.argtypes = [vp, C.POINTER(WalkParams), vp, C.POINTER(WalkSummary)]
    L.frw_walk_noalloc.argtypes = [vp, C.POINTER(WalkParams), vp, vp, C.POINTER(WalkSummary)]
    L.frw_walk_traffic.argtypes = [vp, C.POINTER(WalkParams), vp]
    L.frw_walk_weighted_async.argtypes = [vp, C.POINTER(WalkParams), C.POINTER(WalkOutputs), C.c_double,
                                          C.c_uint32, vp, vp, vp, vp]
    L.frw_walk_weighted.argtypes = [vp, C.POINTER(WalkParams), C.c_double, C.c_uint32, vp, vp, vp, vp]
    L.frw_probe_read_bandwidth.argtypes = [i64, i32, C.POINTER(C.c_double)]
    L.frw_weight_tables.argtypes = [vp, vp]
    L.frw_weight_tables.restype = None
    L.frw_walk_adaptive.argtypes = [vp, C.POINTER(WalkParams), C.c_double, i64, C.c_double, C.c_uint32, vp, vp,
                                    vp, C.POINTER(AdaptiveResult)]
    L.frw_write_results.argtypes = [vp, vp, i64, C.c_char_p]
    L.frw_batch_create.argtypes = [vp, i32, vp, i32, C.POINTER(vp)]
    L.frw_batch_counts_size.restype = i64
    L.frw_batch_counts_size.argtypes = [vp]
    L.frw_batch_counts_offset.argtypes = [vp, i32, C.POINTER(i64)]
    L.frw_batch_walk.argtypes = [vp, C.POINTER(BatchParams), vp, vp, vp]
    L.frw_batch_autotune.argtypes = [vp, C.POINTER(BatchParams), C.POINTER(BatchParams),
                                     C.POINTER(C.c_double)]
    L.frw_batch_check.argtypes = [vp]
    L.frw_batch_free.argtypes = [vp]
    L.frw_batch_free.restype = None
    L.frw_mlp_create.argtypes = [vp, u64, i32, C.POINTER(vp)]
    L.frw_mlp_add_samples.argtypes = [vp, vp, vp, vp, vp, i64]
    L.frw_batch_set_proposal.argtypes = [vp, vp, i32]
    L.frw_batch_walk_learned.argtypes = [vp, C.POINTER(BatchParams), vp, C.c_float, vp, vp, vp, vp]
    L.frw_batch_autotune_learned.argtypes = [vp, C.POINTER(BatchParams), vp, C.c_float,
                                             C.POINTER(BatchParams), C.POINTER(C.c_double)]
    L.frw_batch_harvest.argtypes = [vp, C.POINTER(BatchParams), vp, vp, vp, i64, vp]
    L.frw_mlp_free.argtypes = [vp]
    L.frw_mlp_free.restype = None
    L.frw_mlp_get_params.argtypes = [vp, vp, i32]
    L.frw_mlp_set_params.argtypes = [vp, vp, i32]
    L.frw_mlp_occupancy.argtypes = [vp, i32, vp, i64]
    L.frw_mlp_levels.argtypes = [vp, vp]
    L.frw_mlp_certify.argtypes = [vp, vp, vp, i64, vp, vp]
    L.frw_mlp_features.argtypes = [vp, vp, vp, i64, vp, vp]
    L.frw_mlp_forward.argtypes = [vp, vp, i64, vp, vp]
    L.frw_mlp_gradient.argtypes = [vp, vp, vp, i64, vp, C.POINTER(C.c_float)]
    L.frw_mlp_set_samples.argtypes = [vp, vp, vp, vp, i64]
    L.frw_mlp_train.argtypes = [vp, i32, i32, C.c_float, u64, C.POINTER(C.c_float)]
    L.frw_walk_harvest.argtypes = [vp, C.POINTER(WalkParams), vp, vp, vp, i64, C.POINTER(i64)]
    L.frw_walk_learned.argtypes = [vp, vp, C.POINTER(WalkParams), C.POINTER(WalkOutputs),
                                   C.c_float, vp]
    L.frw_walk_learned_async.argtypes = [vp, vp, C.POINTER(WalkParams), C.POINTER(WalkOutputs),
                                         C.c_float, vp, vp]
    L.frw_walk_learned_traffic.argtypes = [vp, vp, C.POINTER(WalkParams), C.c_float, vp]
    _lib = L
    return L


def _check(rc):
    if rc != OK:
        raise FrwError(rc, lib().frw_last_error_string().decode(errors="replace"))
    return rc


def green_thresholds():
    t = np.zeros(NSAMPLE, dtype=np.uint32)
    _check(lib().frw_green_thresholds(t.ctypes.data))
    return t


def probe_read_bandwidth(nbytes, repeats):
    """GB/s of repeated read sweeps over a device buffer of nbytes."""
    g = C.c_double()
    _check(lib().frw_probe_read_bandwidth(nbytes, repeats, C.byref(g)))
    return g.value


def weight_tables():
    """(gx, gy, pw, qx, qy, tx, ty): the weighted estimator's first-hop tables."""
    wt = np.zeros(5 * NSAMPLE, dtype=np.float64)
    it = np.zeros(2 * NSAMPLE, dtype=np.uint32)
    lib().frw_weight_tables(wt.ctypes.data, it.ctypes.data)
    return tuple(wt.reshape(5, NSAMPLE)) + tuple(it.reshape(2, NSAMPLE))


class Layout:
    """A layer of rects (int32 x0 y0 x1 y1) with its padded absorbing bounds."""

    def __init__(self, handle):
        self.h = C.c_void_p(handle)

    @classmethod
    def load(cls, path, layer=0):
        h = C.c_void_p()
        _check(lib().frw_layout_load(str(path).encode(), layer, C.byref(h)))
        return cls(h.value)

    @classmethod
    def from_rects(cls, rects):
        r = np.ascontiguousarray(rects, dtype=np.int32).reshape(-1, 4)
        h = C.c_void_p()
        _check(lib().frw_layout_from_rects(r.ctypes.data, len(r), C.byref(h)))
        return cls(h.value)

    @classmethod
    def generate(cls, n, box, side_min, side_max, seed=1, cluster_frac=0.0, sigma=0.0,
                 n_clusters=0):
        p = GenParams(kind=GEN_CLUSTERED if cluster_frac > 0 else GEN_UNIFORM, box=box, n=n,
                      side_min=side_min, side_max=side_max, cluster_frac=cluster_frac,
                      sigma=sigma, n_clusters=n_clusters, reserved=0, seed=seed)
        h = C.c_void_p()
        _check(lib().frw_layout_generate(C.byref(p), C.byref(h)))
        return cls(h.value)

    @staticmethod
    def layer_count(path):
        n = C.c_int32()
        _check(lib().frw_layout_layer_count(str(path).encode(), C.byref(n)))
        return n.value

    def __len__(self):
        return lib().frw_layout_size(self.h)

    def set_bounds(self, b):
        """Replace the padded bounds with the caller's (every rect inside them)."""
        arr = np.asarray(b, dtype=np.int32)
        _check(lib().frw_layout_set_bounds(self.h, arr.ctypes.data))

    def rects(self):
        out = np.zeros((len(self), 4), dtype=np.int32)
        if len(out):
            _check(lib().frw_layout_rects(self.h, out.ctypes.data))
        return out

    def bounds(self):
        out = np.zeros(4, dtype=np.int32)
        _check(lib().frw_layout_bounds(self.h, out.ctypes.data))
        return out

    def save(self, path):
        _check(lib().frw_layout_save(self.h, str(path).encode()))

    def write_results(self, counts, n_walks, path):
        c = np.ascontiguousarray(counts, dtype=np.int64)
        _check(lib().frw_write_results(self.h, c.ctypes.data, n_walks, str(path).encode()))

    def __del__(self):
        try:
            if self.h:
                lib().frw_layout_free(self.h)
                self.h = None
        except Exception:
            pass


class Tree:
    """Device quadtree of a layout (Morton sort-and-split build)."""

    def __init__(self, layout, leaf_cap=0, stream=None, ref_compat=False):
        self.layout = layout
        p = TreeParams(leaf_cap=leaf_cap, kind=TREE_REFCOMPAT if ref_compat else TREE_MORTON)
        h = C.c_void_p()
        _check(lib().frw_tree_build(layout.h, C.byref(p), stream, C.byref(h)))
        self.h = h

    def info(self):
        i = TreeInfo()
        _check(lib().frw_tree_info_get(self.h, C.byref(i)))
        return i

    def layout_rect(self, index):
        return self.layout.rects()[index]

    def export(self):
        i = self.info()
        nodes = np.zeros((i.n_nodes, 8), dtype=np.int32)
        rects = np.zeros((i.n_rects, 4), dtype=np.int32)
        ids = np.zeros(i.n_rects, dtype=np.int32)
        _check(lib().frw_tree_export(self.h, nodes.ctypes.data, rects.ctypes.data, ids.ctypes.data))
        return nodes, rects, ids

    # device-buffer queries (torch tensors on the GPU)
    def collide(self, boxes, out, stream=None):
        _check(lib().frw_query_collide(self.h, boxes.data_ptr(), boxes.shape[0], out.data_ptr(), stream))

    def point(self, pts, out, stream=None):
        _check(lib().frw_query_point(self.h, pts.data_ptr(), pts.shape[0], out.data_ptr(), stream))

    def nearest(self, pts, out, stream=None):
        _check(lib().frw_query_nearest(self.h, pts.data_ptr(), pts.shape[0], out.data_ptr(), stream))

    def maxsquare(self, pts, out, mode=MODE_EXACT, stream=None):
        _check(lib().frw_query_maxsquare(self.h, pts.data_ptr(), pts.shape[0], mode, out.data_ptr(),
                                         stream))

    def refcompat_queries(self, pts=None, boxes=None, q1=None, q2=None, q3=None, stream=None, ignore=None):
        """Reference-compatible Q1 (boxes, skipping a rect equal to `ignore`),
        Q2 and Q3 (points) into device tensors."""
        n = pts.shape[0] if pts is not None else boxes.shape[0]
        ptr = lambda x: x.data_ptr() if x is not None else None
        ig = np.asarray(ignore, dtype=np.int32) if ignore is not None else None
        _check(lib().frw_query_refcompat(self.h, ptr(pts), ptr(boxes), n,
                                         ig.ctypes.data if ig is not None else None,
                                         ptr(q1), ptr(q2), ptr(q3), stream))

    def refcompat_gauss_surface(self, master=None, rect=None, factor=1.5):
        """Tree::creatGaussianSurfFrom(rect, factor) on the reference tree (rect = the
        master's by default)."""
        r = np.asarray(self.layout_rect(master) if rect is None else rect, dtype=np.int32)
        out = np.zeros(4, dtype=np.float64)
        _check(lib().frw_refcompat_gauss_surface(self.h, r.ctypes.data, factor, out.ctypes.data))
        return tuple(out)

    def refcompat_walk_from(self, surface, n_walks, counts, hops, hit_out=None, hops_out=None, seed=1,
                            job=0, first_walk=0, max_hops=0, stream=None):
        """Reference-compatible walks from an explicit start square (device outputs)."""
        sq = np.asarray(surface, dtype=np.float64)
        p = WalkParams(master=0, mode=MODE_REFCOMPAT, n_walks=n_walks, first_walk=first_walk, seed=seed,
                       job=job, block=0, grid=0, reserved=0, max_hops=max_hops)
        o = WalkOutputs(counts.data_ptr(), hops.data_ptr(),
                        hit_out.data_ptr() if hit_out is not None else None,
                        hops_out.data_ptr() if hops_out is not None else None)
        _check(lib().frw_refcompat_walk_from(self.h, sq.ctypes.data, C.byref(p), C.byref(o), stream))

    def refcompat_info(self):
        out = np.zeros(2, dtype=np.int64)
        _check(lib().frw_refcompat_info(self.h, out.ctypes.data))
        return tuple(int(v) for v in out)

    def gauss_surface(self, master):
        out = np.zeros(4, dtype=np.float64)
        _check(lib().frw_gauss_surface(self.h, master, out.ctypes.data))
        return tuple(out)

    def walk(self, master, n_walks, seed=1, mode=MODE_EXACT, job=0, first_walk=0, max_hops=0,
             block=0, grid=0):
        """Synchronous walk; returns (counts[n+2], summary)."""
        p = WalkParams(master=master, mode=mode, n_walks=n_walks, first_walk=first_walk, seed=seed,
                       job=job, block=block, grid=grid, reserved=0, max_hops=max_hops)
        counts = np.zeros(self.info().n_rects + 2, dtype=np.int64)
        s = WalkSummary()
        _check(lib().frw_walk(self.h, C.byref(p), counts.ctypes.data, C.byref(s)))
        return counts, s

    def walk_noalloc(self, master, n_walks, seed=1, mode=MODE_EXACT, job=0, first_walk=0, max_hops=0,
                     block=0, grid=0, stream=None):
        """frw_walk_noalloc: the walks of walk() with the library's fixed
        scratch array instead of per-call device allocations; returns
        (counts[n+2], summary)."""
        p = WalkParams(master=master, mode=mode, n_walks=n_walks, first_walk=first_walk, seed=seed,
                       job=job, block=block, grid=grid, reserved=0, max_hops=max_hops)
        counts = np.zeros(self.info().n_rects + 2, dtype=np.int64)
        s = WalkSummary()
        _check(lib().frw_walk_noalloc(self.h, C.byref(p), stream, counts.ctypes.data, C.byref(s)))
        return counts, s

    def walk_traffic(self, master, n_walks, seed=1, mode=MODE_EXACT, job=0, first_walk=0,
                     max_hops=0, block=0, grid=0):
        """Instrumented walk: (node records read, rect records read, hops,
        start-table entries read)."""
        p = WalkParams(master=master, mode=mode, n_walks=n_walks, first_walk=first_walk, seed=seed,
                       job=job, block=block, grid=grid, reserved=0, max_hops=max_hops)
        out = np.zeros(4, dtype=np.int64)
        _check(lib().frw_walk_traffic(self.h, C.byref(p), out.ctypes.data))
        return tuple(int(v) for v in out)

    def walk_async(self, master, n_walks, counts, hops, hit_out=None, hops_out=None, seed=1,
                   mode=MODE_EXACT, job=0, first_walk=0, max_hops=0, block=0, grid=0, stream=None):
        """Asynchronous walk into device tensors (counts int64[n+2], hops int64[1])."""
        p = WalkParams(master=master, mode=mode, n_walks=n_walks, first_walk=first_walk, seed=seed,
                       job=job, block=block, grid=grid, reserved=0, max_hops=max_hops)
        o = WalkOutputs(counts.data_ptr(), hops.data_ptr(),
                        hit_out.data_ptr() if hit_out is not None else None,
                        hops_out.data_ptr() if hops_out is not None else None)
        _check(lib().frw_walk_async(self.h, C.byref(p), C.byref(o), stream))

    def walk_weighted_async(self, master, n_walks, counts, hops, wsum, wsq, w_out=None, hit_out=None,
                            hops_out=None, seed=1, eps=1.0, job=0, first_walk=0, max_hops=0, block=0,
                            grid=0, stream=None, flags=0):
        """Weighted (physical) estimator into device tensors: wsum[j] / n_walks
        estimates the capacitance C_master,j (eps per unit length).  flags:
        WEIGHT_IMPORTANCE draws the first panel proportional to |n . grad G|."""
        p = WalkParams(master=master, mode=MODE_WEIGHTED, n_walks=n_walks, first_walk=first_walk, seed=seed,
                       job=job, block=block, grid=grid, reserved=0, max_hops=max_hops)
        o = WalkOutputs(counts.data_ptr(), hops.data_ptr(),
                        hit_out.data_ptr() if hit_out is not None else None,
                        hops_out.data_ptr() if hops_out is not None else None)
        _check(lib().frw_walk_weighted_async(self.h, C.byref(p), C.byref(o), eps, flags, wsum.data_ptr(),
                                             wsq.data_ptr(), w_out.data_ptr() if w_out is not None else None,
                                             stream))

    def walk_weighted(self, master, n_walks, seed=1, eps=1.0, job=0, first_walk=0, max_hops=0, flags=0):
        """Synchronous weighted walks: (counts, wsum, wsq, hops) on the host."""
        p = WalkParams(master=master, mode=MODE_WEIGHTED, n_walks=n_walks, first_walk=first_walk, seed=seed,
                       job=job, block=0, grid=0, reserved=0, max_hops=max_hops)
        nb = self.info().n_rects + 2
        counts = np.zeros(nb, dtype=np.int64)
        wsum = np.zeros(nb, dtype=np.float64)
        wsq = np.zeros(nb, dtype=np.float64)
        hops = C.c_int64()
        _check(lib().frw_walk_weighted(self.h, C.byref(p), eps, flags, counts.ctypes.data, wsum.ctypes.data,
                                       wsq.ctypes.data, C.byref(hops)))
        return counts, wsum, wsq, hops.value

    def walk_adaptive(self, master, rel_err=0.01, mode=MODE_EXACT, first_round=0, max_walks=0, seed=1,
                      eps=1.0, flags=0, first_walk=0, job=0, max_hops=0):
        """Walks in rounds until the 3-sigma relative error of f (hit modes)
        or of C_mm (MODE_WEIGHTED) reaches rel_err: (result, counts, wsum, wsq)."""
        p = WalkParams(master=master, mode=mode, n_walks=first_round, first_walk=first_walk, seed=seed,
                       job=job, block=0, grid=0, reserved=0, max_hops=max_hops)
        nb = self.info().n_rects + 2
        counts = np.zeros(nb, dtype=np.int64)
        wsum = np.zeros(nb, dtype=np.float64)
        wsq = np.zeros(nb, dtype=np.float64)
        r = AdaptiveResult()
        _check(lib().frw_walk_adaptive(self.h, C.byref(p), rel_err, max_walks, eps, flags, counts.ctypes.data,
                                       wsum.ctypes.data, wsq.ctypes.data, C.byref(r)))
        return r, counts, wsum, wsq

    def __del__(self):
        try:
            if self.h:
                lib().frw_tree_free(self.h)
                self.h = None
        except Exception:
            pass


class Batch:
    """Many (tree, master, walk count) jobs in one persistent walk launch."""

    def __init__(self, trees, jobs):
        """trees: list of Tree; jobs: list of dicts with keys tree, master,
        n_walks and optionally first_walk, job."""
        self.trees = list(trees)
        arr = (C.c_void_p * len(self.trees))(*[t.h.value for t in self.trees])
        js = (BatchJob * len(jobs))()
        for i, j in enumerate(jobs):
            js[i] = BatchJob(tree=j["tree"], master=j["master"], n_walks=j["n_walks"],
                             first_walk=j.get("first_walk", 0), job=j.get("job", i), reserved=0)
        self.n_jobs = len(jobs)
        h = C.c_void_p()
        _check(lib().frw_batch_create(arr, len(self.trees), js, len(jobs), C.byref(h)))
        self.h = h

    def counts_size(self):
        return lib().frw_batch_counts_size(self.h)

    def counts_offset(self, job):
        o = C.c_int64()
        _check(lib().frw_batch_counts_offset(self.h, job, C.byref(o)))
        return o.value

    def walk(self, counts, hops, seed=1, mode=MODE_EXACT, block=0, grid=0, use_graph=False,
             max_hops=0, stream=None):
        p = BatchParams(mode=mode, block=block, grid=grid, use_graph=int(use_graph), seed=seed,
                        max_hops=max_hops)
        _check(lib().frw_batch_walk(self.h, C.byref(p), counts.data_ptr(), hops.data_ptr(), stream))

    def autotune(self, seed=1, mode=MODE_EXACT, use_graph=False):
        base = BatchParams(mode=mode, block=0, grid=0, use_graph=int(use_graph), seed=seed,
                           max_hops=0)
        best = BatchParams()
        rate = C.c_double()
        _check(lib().frw_batch_autotune(self.h, C.byref(base), C.byref(best), C.byref(rate)))
        return {"block": best.block, "grid": best.grid}, rate.value

    def set_proposal(self, per_tree):
        """per_tree: one Mlp per tree of the batch (their occupancy pyramids)."""
        arr = (C.c_void_p * len(per_tree))(*[m.h.value for m in per_tree])
        self._pyr = list(per_tree)
        _check(lib().frw_batch_set_proposal(self.h, arr, len(per_tree)))

    def walk_learned(self, net, counts, hops, stats=None, alpha=0.9, seed=1, block=0, grid=0,
                     use_graph=False, max_hops=0, stream=None):
        p = BatchParams(mode=MODE_LEARNED, block=block, grid=grid, use_graph=int(use_graph), seed=seed,
                        max_hops=max_hops)
        _check(lib().frw_batch_walk_learned(self.h, C.byref(p), net.h, alpha, counts.data_ptr(),
                                            hops.data_ptr(), stats.data_ptr() if stats is not None else None,
                                            stream))

    def autotune_learned(self, net, alpha=0.9, seed=1, use_graph=False, max_hops=0):
        base = BatchParams(mode=MODE_LEARNED, block=0, grid=0, use_graph=int(use_graph), seed=seed,
                           max_hops=max_hops)
        best = BatchParams()
        rate = C.c_double()
        _check(lib().frw_batch_autotune_learned(self.h, C.byref(base), net.h, alpha, C.byref(best),
                                                C.byref(rate)))
        return {"block": best.block, "grid": best.grid}, rate.value

    def harvest(self, pts, h, dstar, seed=1, block=0, grid=0, max_hops=0):
        """walk_harvest for every job in one launch per pass: the records of
        job 0's walks, then job 1's, ... (walk and hop order), up to
        pts.shape[0] in all; returns each job's record count (numpy int64)."""
        p = BatchParams(mode=MODE_EXACT, block=block, grid=grid, use_graph=0, seed=seed, max_hops=max_hops)
        n = np.zeros(self.n_jobs, dtype=np.int64)
        _check(lib().frw_batch_harvest(self.h, C.byref(p), pts.data_ptr(), h.data_ptr(), dstar.data_ptr(),
                                       pts.shape[0], n.ctypes.data))
        return n

    def check(self):
        _check(lib().frw_batch_check(self.h))

    def __del__(self):
        try:
            if self.h:
                lib().frw_batch_free(self.h)
                self.h = None
        except Exception:
            pass


class Mlp:
    """Learned square-size proposal bound to one tree (occupancy pyramid +
    67-64-64-64-1 network)."""

    def __init__(self, tree, seed=1, max_level=0):
        self.tree = tree
        h = C.c_void_p()
        _check(lib().frw_mlp_create(tree.h, seed, max_level, C.byref(h)))
        self.h = h
        self.n_params = lib().frw_mlp_param_count()

    def params(self):
        p = np.zeros(self.n_params, dtype=np.float32)
        _check(lib().frw_mlp_get_params(self.h, p.ctypes.data, self.n_params))
        return p

    def set_params(self, p):
        p = np.ascontiguousarray(p, dtype=np.float32)
        _check(lib().frw_mlp_set_params(self.h, p.ctypes.data, self.n_params))

    def occupancy(self, level):
        n = (1 << level) * (1 << level)
        words = np.zeros((n + 31) // 32, dtype=np.uint32)
        _check(lib().frw_mlp_occupancy(self.h, level, words.ctypes.data, len(words)))
        bits = np.unpackbits(words.view(np.uint8), bitorder="little")[:n]
        return bits.reshape(1 << level, 1 << level)   # [j][i]

    def levels(self):
        """(K0, finest window level, finest verification level, finest bitmap level)."""
        out = np.zeros(4, dtype=np.int32)
        _check(lib().frw_mlp_levels(self.h, out.ctypes.data))
        return tuple(int(v) for v in out)

    def certify(self, pts, d, out, stream=None):
        """Verification probe on device tensors: out uint8[n]."""
        _check(lib().frw_mlp_certify(self.h, pts.data_ptr(), d.data_ptr(), pts.shape[0], out.data_ptr(),
                                     stream))

    def features(self, pts, h, out, stream=None):
        _check(lib().frw_mlp_features(self.h, pts.data_ptr(), h.data_ptr(), pts.shape[0],
                                      out.data_ptr(), stream))

    def forward(self, X, y, stream=None):
        _check(lib().frw_mlp_forward(self.h, X.data_ptr(), X.shape[0], y.data_ptr(), stream))

    def gradient(self, X, t):
        g = np.zeros(self.n_params, dtype=np.float32)
        loss = C.c_float()
        _check(lib().frw_mlp_gradient(self.h, X.data_ptr(), t.data_ptr(), X.shape[0], g.ctypes.data,
                                      C.byref(loss)))
        return g, loss.value

    def set_samples(self, pts, h, dstar):
        _check(lib().frw_mlp_set_samples(self.h, pts.data_ptr(), h.data_ptr(), dstar.data_ptr(),
                                         pts.shape[0]))

    def add_samples(self, pts, h, dstar, features_of=None):
        """Appends samples with features from `features_of`'s pyramid (default: own)."""
        src = features_of.h if features_of is not None else None
        _check(lib().frw_mlp_add_samples(self.h, src, pts.data_ptr(), h.data_ptr(), dstar.data_ptr(),
                                         pts.shape[0]))

    def train(self, epochs=2, batch=1024, lr=1e-3, seed=1):
        loss = C.c_float()
        _check(lib().frw_mlp_train(self.h, epochs, batch, lr, seed, C.byref(loss)))
        return loss.value

    def __del__(self):
        try:
            if self.h:
                lib().frw_mlp_free(self.h)
                self.h = None
        except Exception:
            pass


def walk_harvest(tree, master, n_walks, pts, h, dstar, seed=1, first_walk=0, job=0):
    """Exact walks that record (point, previous half-size, d*) of non-final hops
    into device tensors; returns the number of records."""
    p = WalkParams(master=master, mode=MODE_EXACT, n_walks=n_walks, first_walk=first_walk, seed=seed,
                   job=job, block=0, grid=0, reserved=0, max_hops=0)
    n = C.c_int64()
    _check(lib().frw_walk_harvest(tree.h, C.byref(p), pts.data_ptr(), h.data_ptr(), dstar.data_ptr(),
                                  pts.shape[0], C.byref(n)))
    return n.value


def walk_learned(tree, mlp, master, n_walks, counts, hops, hit_out=None, hops_out=None, seed=1,
                 first_walk=0, job=0, alpha=0.8, block=0, grid=0, max_hops=0):
    """Walks with the learned proposal; returns (accepted, rejected) proposal
    counts (their sum: network evaluations; hops - accepted: exact searches)."""
    p = WalkParams(master=master, mode=MODE_LEARNED, n_walks=n_walks, first_walk=first_walk,
                   seed=seed, job=job, block=block, grid=grid, reserved=0, max_hops=max_hops)
    o = WalkOutputs(counts.data_ptr(), hops.data_ptr(),
                    hit_out.data_ptr() if hit_out is not None else None,
                    hops_out.data_ptr() if hops_out is not None else None)
    st = np.zeros(2, dtype=np.int64)
    _check(lib().frw_walk_learned(tree.h, mlp.h, C.byref(p), C.byref(o), alpha, st.ctypes.data))
    return int(st[0]), int(st[1])


def walk_learned_async(tree, mlp, master, n_walks, counts, hops, stats=None, hit_out=None,
                       hops_out=None, seed=1, first_walk=0, job=0, alpha=0.8, block=0, grid=0,
                       stream=None):
    """Asynchronous learned walks into device tensors; `stats` (device uint64[2]
    or int64[2]) accumulates (accepted, rejected) proposals when given."""
    p = WalkParams(master=master, mode=MODE_LEARNED, n_walks=n_walks, first_walk=first_walk,
                   seed=seed, job=job, block=block, grid=grid, reserved=0, max_hops=0)
    o = WalkOutputs(counts.data_ptr(), hops.data_ptr(),
                    hit_out.data_ptr() if hit_out is not None else None,
                    hops_out.data_ptr() if hops_out is not None else None)
    _check(lib().frw_walk_learned_async(tree.h, mlp.h, C.byref(p), C.byref(o), alpha,
                                        stats.data_ptr() if stats is not None else None, stream))


def walk_learned_traffic(tree, mlp, master, n_walks, seed=1, alpha=0.8, first_walk=0, job=0):
    """Instrumented learned walks: (node records, rect records, hops, accepted, rejected,
    start-table entries)."""
    p = WalkParams(master=master, mode=MODE_LEARNED, n_walks=n_walks, first_walk=first_walk,
                   seed=seed, job=job, block=0, grid=0, reserved=0, max_hops=0)
    out = np.zeros(6, dtype=np.int64)
    _check(lib().frw_walk_learned_traffic(tree.h, mlp.h, C.byref(p), alpha, out.ctypes.data))
    return tuple(int(v) for v in out)
