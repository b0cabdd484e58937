"""ctypes bindings for the CPU oracle (oracle/liboracle.so) and for the
reference's own CPU sources compiled into oracle/_ref/libfrwref.so.

Test infrastructure only: tests/, ``__graft_entry__.smoke()`` and bench.py's
``cpu_baseline`` leg are the only users.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_DIR = os.path.join(ROOT, "oracle")
ORACLE_SO = os.path.join(ORACLE_DIR, "liboracle.so")
REF_SO = os.path.join(ORACLE_DIR, "_ref", "libfrwref.so")
REFERENCE_SRC = "/root/reference/randomWalk"

NSAMPLE = 200
HIT_ESCAPED = -1
HIT_STUCK = -2
EXACT, REFTREE, REFTREE_DFS = 0, 1, 2


class Box(C.Structure):
    _fields_ = [("x0", C.c_double), ("y0", C.c_double), ("x1", C.c_double), ("y1", C.c_double)]

    def t(self):
        return (self.x0, self.y0, self.x1, self.y1)


def build_oracle(ref: bool = True) -> None:
    """Compile the oracle (and, where the reference tree exists, _ref)."""
    subprocess.run(["make", "-s", "-C", ORACLE_DIR], check=True)
    if ref and os.path.isdir(REFERENCE_SRC):
        subprocess.run(["make", "-s", "-C", ORACLE_DIR, "ref"], check=True)


_orc = None
_ref = None

_D = C.POINTER(C.c_double)
_I32 = C.POINTER(C.c_int32)
_I64 = C.POINTER(C.c_int64)
_U32 = C.POINTER(C.c_uint32)
_F32 = C.POINTER(C.c_float)


def _ptr(a, t):
    return a.ctypes.data_as(t) if a is not None else None


def oracle():
    global _orc
    if _orc is not None:
        return _orc
    if not os.path.exists(ORACLE_SO):
        build_oracle(ref=False)
    L = C.CDLL(ORACLE_SO)
    L.orc_philox.argtypes = [_U32, _U32, _U32]
    L.orc_walk_u32.restype = C.c_uint32
    L.orc_walk_u32.argtypes = [C.c_uint64, C.c_uint64, C.c_uint32, C.c_uint32]
    L.orc_green_cdf.argtypes = [_D]
    L.orc_green_thresholds.argtypes = [_U32]
    L.orc_sample_index.argtypes = [_U32, C.c_uint32]
    L.orc_get_index_fp.argtypes = [_D, C.c_double]
    L.orc_point_from_index.argtypes = [C.POINTER(Box), C.c_int, C.c_int, _D, _D]
    L.orc_gauss_surface.restype = Box
    L.orc_gauss_surface.argtypes = [C.POINTER(Box), C.c_double, C.c_double]
    L.orc_parse.argtypes = [C.c_char_p, C.c_int, _D, C.c_int, C.POINTER(C.c_int)]
    L.orc_bounds.argtypes = [_D, C.c_int, C.POINTER(Box)]
    L.orc_index_new.restype = C.c_void_p
    L.orc_index_new.argtypes = [_D, C.c_int, C.POINTER(Box), C.c_int]
    L.orc_index_free.argtypes = [C.c_void_p]
    L.orc_index_stats.argtypes = [C.c_void_p, _I64]
    L.orc_q1.argtypes = [C.c_void_p, C.POINTER(Box), C.POINTER(Box)]
    L.orc_q2.argtypes = [C.c_void_p, C.c_double, C.c_double]
    L.orc_nearest.restype = C.c_double
    L.orc_nearest.argtypes = [C.c_void_p, C.c_double, C.c_double, C.POINTER(C.c_int)]
    L.orc_q3.restype = Box
    L.orc_q3.argtypes = [C.c_void_p, C.c_double, C.c_double, C.POINTER(C.c_int)]
    L.orc_gauss_from.restype = Box
    L.orc_gauss_from.argtypes = [C.c_void_p, C.c_int]
    L.orc_queries_many.argtypes = [C.c_void_p, _D, _D, C.c_int64, C.c_int64, C.POINTER(C.c_uint8), _I32,
                                   _D, C.c_int]
    L.orc_square_density.restype = C.c_double
    L.orc_square_density.argtypes = [C.c_double, C.c_double, C.c_double]
    L.orc_green_grad.argtypes = [_D, _D]
    L.orc_walk_weighted_many.restype = C.c_int64
    L.orc_walk_weighted_many.argtypes = [C.c_void_p, C.c_int, C.c_uint64, C.c_uint32, C.c_int64, C.c_int64,
                                         C.c_int64, C.c_double, C.c_int, _I32, _I32, _D]
    L.orc_importance_table.argtypes = [_D, _U32, _D]
    L.orc_bf16.restype = C.c_float
    L.orc_bf16.argtypes = [C.c_float]
    L.orc_mlp_set_rounding.argtypes = [C.c_int]
    L.orc_mlp_level.argtypes = [C.POINTER(Box), C.c_double, _D]
    L.orc_mlp_features.argtypes = [_D, C.c_int, C.POINTER(Box), C.c_double, C.c_double,
                                   C.c_double, _F32]
    L.orc_mlp_forward.argtypes = [_F32, _F32, C.c_int, _F32]
    L.orc_mlp_vmax.argtypes = [C.POINTER(Box), C.c_int]
    L.orc_mlp_certify.argtypes = [_D, C.c_int, C.POINTER(Box), C.c_double, C.c_double, C.c_double]
    L.orc_mlp_backward.restype = C.c_float
    L.orc_mlp_backward.argtypes = [_F32, _F32, _F32, C.c_int, _F32]
    L.orc_walk_one.restype = C.c_int
    L.orc_walk_one.argtypes = [C.c_void_p, C.POINTER(Box), _U32, C.c_uint64, C.c_uint32, C.c_uint64, C.c_int64,
                               _I64]
    L.orc_walk_many.restype = C.c_int64
    L.orc_walk_many.argtypes = [C.c_void_p, C.c_int, C.c_uint64, C.c_uint32, C.c_int64,
                                C.c_int64, C.c_int64, _I64, _I32, _I32, C.c_int]
    _orc = L
    return L


def ref_available() -> bool:
    return os.path.exists(REF_SO)


def reference():
    global _ref
    if _ref is not None:
        return _ref
    if not os.path.exists(REF_SO):
        if os.path.isdir(REFERENCE_SRC):
            build_oracle(ref=True)
        else:
            raise RuntimeError("oracle/_ref/libfrwref.so is not built")
    L = C.CDLL(REF_SO)
    L.ref_tree_new.restype = C.c_void_p
    L.ref_tree_new.argtypes = [_D, C.c_int, _D]
    L.ref_q1.argtypes = [C.c_void_p, _D, _D]
    L.ref_q2.argtypes = [C.c_void_p, C.c_double, C.c_double, _D]
    L.ref_q3.argtypes = [C.c_void_p, C.c_double, C.c_double, _D]
    L.ref_gauss_from.argtypes = [C.c_void_p, _D, _D]
    L.ref_in_bounds_point.argtypes = [C.c_void_p, C.c_double, C.c_double]
    L.ref_point_from_index.argtypes = [_D, C.c_int, C.c_int, _D]
    L.ref_parse.argtypes = [C.c_char_p, C.c_int, _D, C.c_int, _D, C.POINTER(C.c_int)]
    L.ref_walk_many.restype = C.c_int64
    L.ref_walk_many.argtypes = [C.c_void_p, _D, C.c_uint64, C.c_uint32, C.c_int64, C.c_int64,
                                C.c_int64, _D, _I32, _I64]
    _ref = L
    return L


# ---------------------------------------------------------------- helpers

def philox(ctr, key):
    L = oracle()
    c = np.asarray(ctr, dtype=np.uint32)
    k = np.asarray(key, dtype=np.uint32)
    o = np.zeros(4, dtype=np.uint32)
    L.orc_philox(_ptr(c, _U32), _ptr(k, _U32), _ptr(o, _U32))
    return o


def green_cdf():
    o = np.zeros(NSAMPLE + 1, dtype=np.float64)
    oracle().orc_green_cdf(_ptr(o, _D))
    return o


def green_thresholds():
    o = np.zeros(NSAMPLE, dtype=np.uint32)
    oracle().orc_green_thresholds(_ptr(o, _U32))
    return o


def point_from_index(sq, index, nsample=NSAMPLE):
    x, y = C.c_double(), C.c_double()
    oracle().orc_point_from_index(C.byref(Box(*sq)), index, nsample, C.byref(x), C.byref(y))
    return (x.value, y.value)


def parse(path, layer=0, cap=10_000_000):
    L = oracle()
    nl = C.c_int()
    n = L.orc_parse(path.encode(), layer, None, 0, C.byref(nl))
    if n < 0:
        return n, None, nl.value
    buf = np.zeros((max(n, 1), 4), dtype=np.float64)
    L.orc_parse(path.encode(), layer, _ptr(buf, _D), n, C.byref(nl))
    return n, buf[:n], nl.value


def bounds(rects):
    r = np.ascontiguousarray(rects, dtype=np.float64)
    b = Box()
    oracle().orc_bounds(_ptr(r, _D), len(r), C.byref(b))
    return b.t()


class Index:
    """Oracle spatial index (EXACT, REFTREE or REFTREE_DFS)."""

    def __init__(self, rects, bnds=None, kind=EXACT):
        self.L = oracle()
        self.r = np.ascontiguousarray(rects, dtype=np.float64).reshape(-1, 4)
        self.n = len(self.r)
        self.bounds = tuple(bnds) if bnds is not None else bounds(self.r)
        self.kind = kind
        self.h = self.L.orc_index_new(_ptr(self.r, _D), self.n, C.byref(Box(*self.bounds)), kind)

    def __del__(self):
        try:
            self.L.orc_index_free(self.h)
        except Exception:
            pass

    def stats(self):
        o = np.zeros(3, dtype=np.int64)
        self.L.orc_index_stats(self.h, _ptr(o, _I64))
        return o

    def q1(self, box, ignore=None):
        ig = C.byref(Box(*ignore)) if ignore is not None else None
        return self.L.orc_q1(self.h, C.byref(Box(*box)), ig)

    def q2(self, px, py):
        return self.L.orc_q2(self.h, px, py)

    def queries_many(self, pts, boxes, n_q3=0, nthreads=0):
        """(q1[n], q2[n], q3[n_q3, 4]) for points and boxes, OpenMP over queries."""
        P = np.ascontiguousarray(pts, dtype=np.float64)
        B = np.ascontiguousarray(boxes, dtype=np.float64)
        n = len(P)
        q1 = np.zeros(n, dtype=np.uint8)
        q2 = np.zeros(n, dtype=np.int32)
        q3 = np.zeros((max(n_q3, 1), 4), dtype=np.float64)
        self.L.orc_queries_many(self.h, _ptr(P, _D), _ptr(B, _D), n, n_q3,
                                q1.ctypes.data_as(C.POINTER(C.c_uint8)), _ptr(q2, _I32), _ptr(q3, _D),
                                nthreads)
        return q1, q2, q3[:n_q3]

    def walk_weighted(self, master, n_walks, seed=1, job=0, first_walk=0, max_hops=10**6, eps=1.0,
                      importance=False):
        """(hit[n], hops[n], weight[n]) of weighted walks (Gaussian-surface start;
        importance: first panel proportional to |n . grad G|)."""
        hit = np.zeros(n_walks, dtype=np.int32)
        hops = np.zeros(n_walks, dtype=np.int32)
        w = np.zeros(n_walks, dtype=np.float64)
        self.L.orc_walk_weighted_many(self.h, master, seed, job, first_walk, n_walks, max_hops, eps,
                                      int(importance), _ptr(hit, _I32), _ptr(hops, _I32), _ptr(w, _D))
        return hit, hops, w

    def nearest(self, px, py):
        a = C.c_int()
        d = self.L.orc_nearest(self.h, px, py, C.byref(a))
        return d, a.value

    def q3(self, px, py):
        nq = C.c_int()
        b = self.L.orc_q3(self.h, px, py, C.byref(nq))
        return b.t(), nq.value

    def gauss_from(self, master):
        return self.L.orc_gauss_from(self.h, master).t()

    def walk(self, master, n_walks, seed=1, job=0, first_walk=0, max_hops=10**6,
             per_walk=False, nthreads=0):
        counts = np.zeros(self.n + 2, dtype=np.int64)
        hit = np.zeros(n_walks, dtype=np.int32) if per_walk else None
        hops = np.zeros(n_walks, dtype=np.int32) if per_walk else None
        total = self.L.orc_walk_many(self.h, master, seed, job, first_walk, n_walks, max_hops,
                                     _ptr(counts, _I64), _ptr(hit, _I32), _ptr(hops, _I32),
                                     nthreads)
        return counts, total, hit, hops

    def walk_one_from(self, surface, T, seed, job, walk, max_hops=10**6):
        """(hit, hops) of one walk from an explicit start square."""
        T = np.ascontiguousarray(T, dtype=np.uint32)
        h = C.c_int64()
        hit = self.L.orc_walk_one(self.h, C.byref(Box(*surface)), _ptr(T, _U32), seed, job, walk, max_hops,
                                  C.byref(h))
        return hit, h.value


class RefTree:
    """The reference's own Tree (randomWalk/CPU/quadTree.cpp), built from source."""

    def __init__(self, rects, bnds=None):
        self.L = reference()
        self.r = np.ascontiguousarray(rects, dtype=np.float64).reshape(-1, 4)
        self.bounds = np.asarray(bnds if bnds is not None else bounds(self.r), dtype=np.float64)
        self.h = self.L.ref_tree_new(_ptr(self.r, _D), len(self.r), _ptr(self.bounds, _D))

    def q1(self, box, ignore=None):
        b = np.asarray(box, dtype=np.float64)
        ig = np.asarray(ignore, dtype=np.float64) if ignore is not None else None
        return self.L.ref_q1(self.h, _ptr(b, _D), _ptr(ig, _D))

    def q2(self, px, py):
        o = np.zeros(4)
        hit = self.L.ref_q2(self.h, px, py, _ptr(o, _D))
        return hit, tuple(o)

    def q3(self, px, py):
        o = np.zeros(4)
        self.L.ref_q3(self.h, px, py, _ptr(o, _D))
        return tuple(o)

    def gauss_from(self, rect):
        r = np.asarray(rect, dtype=np.float64)
        o = np.zeros(4)
        self.L.ref_gauss_from(self.h, _ptr(r, _D), _ptr(o, _D))
        return tuple(o)

    def walk(self, master_rect, n_walks, seed=1, job=0, first_walk=0, max_hops=10**6):
        m = np.asarray(master_rect, dtype=np.float64)
        hit = np.zeros((n_walks, 4), dtype=np.float64)
        hops = np.zeros(n_walks, dtype=np.int32)
        other = C.c_int64()
        total = self.L.ref_walk_many(self.h, _ptr(m, _D), seed, job, first_walk, n_walks,
                                     max_hops, _ptr(hit, _D), _ptr(hops, _I32), C.byref(other))
        return hit, hops, total, other.value


def ref_point_from_index(sq, index, nsample=NSAMPLE):
    s = np.asarray(sq, dtype=np.float64)
    o = np.zeros(2)
    reference().ref_point_from_index(_ptr(s, _D), index, nsample, _ptr(o, _D))
    return (o[0], o[1])


def ref_parse(path, layer=0, cap=10_000_000):
    L = reference()
    nl = C.c_int()
    b = np.zeros(4)
    n = L.ref_parse(path.encode(), layer, None, 0, None, C.byref(nl))
    if n < 0:
        return n, None, None, nl.value
    buf = np.zeros((max(n, 1), 4), dtype=np.float64)
    L.ref_parse(path.encode(), layer, _ptr(buf, _D), n, _ptr(b, _D), C.byref(nl))
    return n, buf[:n], tuple(b), nl.value


# ---------------------------------------------------------------- learned proposal

MLP_IN_PAD = 72


def bf16(x: float) -> float:
    return oracle().orc_bf16(x)


def mlp_param_count() -> int:
    return oracle().orc_mlp_param_count()


def mlp_level(bnds, h):
    cell = C.c_double()
    lvl = oracle().orc_mlp_level(C.byref(Box(*bnds)), h, C.byref(cell))
    return lvl, cell.value


def mlp_features(rects, bnds, pts, hs):
    """[n, 72] float32 feature rows (67 features + zero padding), brute force."""
    r = np.ascontiguousarray(rects, dtype=np.float64)
    b = Box(*bnds)
    out = np.zeros((len(pts), MLP_IN_PAD), dtype=np.float32)
    row = np.zeros(67, dtype=np.float32)
    L = oracle()
    for i, ((px, py), h) in enumerate(zip(pts, hs)):
        L.orc_mlp_features(_ptr(r, _D), len(r), C.byref(b), float(px), float(py), float(h), _ptr(row, _F32))
        out[i, :67] = row
    return out


def mlp_vmax(bnds, n_rects):
    return oracle().orc_mlp_vmax(C.byref(Box(*bnds)), n_rects)


def mlp_certify(rects, bnds, pts, ds):
    """1 where the square of half-size d at p is certified free (brute force)."""
    r = np.ascontiguousarray(rects, dtype=np.float64)
    b = Box(*bnds)
    L = oracle()
    return np.array([L.orc_mlp_certify(_ptr(r, _D), len(r), C.byref(b), float(px), float(py), float(d))
                     for (px, py), d in zip(pts, ds)], dtype=np.uint8)


class _Rounding:
    def __init__(self, on):
        self.on = on

    def __enter__(self):
        oracle().orc_mlp_set_rounding(1 if self.on else 0)

    def __exit__(self, *a):
        oracle().orc_mlp_set_rounding(1)


def mlp_forward(params, X, rounding=True):
    P = np.ascontiguousarray(params, dtype=np.float32)
    Xc = np.ascontiguousarray(X, dtype=np.float32)
    y = np.zeros(len(Xc), dtype=np.float32)
    with _Rounding(rounding):
        oracle().orc_mlp_forward(_ptr(P, _F32), _ptr(Xc, _F32), len(Xc), _ptr(y, _F32))
    return y


def mlp_backward(params, X, t, rounding=True):
    P = np.ascontiguousarray(params, dtype=np.float32)
    Xc = np.ascontiguousarray(X, dtype=np.float32)
    tc = np.ascontiguousarray(t, dtype=np.float32)
    g = np.zeros(len(P), dtype=np.float32)
    with _Rounding(rounding):
        loss = oracle().orc_mlp_backward(_ptr(P, _F32), _ptr(Xc, _F32), _ptr(tc, _F32), len(Xc), _ptr(g, _F32))
    return float(loss), g


# ---------------------------------------------------------------- weighted estimator

def square_density(s, u, v):
    return oracle().orc_square_density(s, u, v)


def importance_table(g):
    """(thresholds u32[200], probabilities[200]) drawing panel k proportional to |g[k]|."""
    g = np.ascontiguousarray(g, dtype=np.float64)
    T = np.zeros(NSAMPLE, dtype=np.uint32)
    q = np.zeros(NSAMPLE, dtype=np.float64)
    oracle().orc_importance_table(_ptr(g, _D), _ptr(T, _U32), _ptr(q, _D))
    return T, q


def green_grad():
    gx = np.zeros(NSAMPLE, dtype=np.float64)
    gy = np.zeros(NSAMPLE, dtype=np.float64)
    oracle().orc_green_grad(_ptr(gx, _D), _ptr(gy, _D))
    return gx, gy
