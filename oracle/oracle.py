"""ctypes front end of the CPU oracle (oracle/emst_oracle.c).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg, always as the checker or the timed CPU baseline,
never as part of the product path.
"""
import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle_emst.so")
_lib = None

MAX_ROUNDS = 128


class Grid(ctypes.Structure):
    _fields_ = [
        ("dim", ctypes.c_int), ("R", ctypes.c_int),
        ("ext", ctypes.c_int32 * 3), ("pc", ctypes.c_int32 * 3), ("PC", ctypes.c_int32 * 3),
        ("ext_base", ctypes.c_int32 * 3), ("PCD", ctypes.c_int32 * 3),
        ("ext_dual", ctypes.c_int32 * 3), ("p1", ctypes.c_float * 3),
        ("d_max", ctypes.c_int32), ("ncell", ctypes.c_int64), ("pdim", ctypes.c_int),
    ]


class Trace(ctypes.Structure):
    _fields_ = [
        ("rounds", ctypes.c_int32),
        ("cum_edges", ctypes.c_int32 * MAX_ROUNDS),
        ("searched", ctypes.c_int64 * MAX_ROUNDS),
        ("examined", ctypes.c_int64 * MAX_ROUNDS),
        ("overfull_cells", ctypes.c_int64),
        ("cycles_cut", ctypes.c_int64),
        ("ms_after_cells", ctypes.c_double),
    ]


def build():
    """Compile the oracle with its Makefile (gcc)."""
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def _bind(L):
    P = ctypes.c_void_p
    i64, i32, dbl = ctypes.c_int64, ctypes.c_int, ctypes.c_double
    L.or_bbox.argtypes = [P, i64, i32, P, P]
    L.or_rescale.argtypes = [P, i64, i32, P, P, P, P, P]
    L.or_rescale.restype = dbl
    L.or_grid_init.argtypes = [P, i32, i32, P, P]
    L.or_grid_init_pd.argtypes = [P, i32, i32, i32, P, P]
    L.or_find_cell.argtypes = [P, P, P]
    L.or_d2.argtypes = [P, P, i32]
    L.or_d2.restype = ctypes.c_float
    L.or_d2_pairs.argtypes = [P, i32, P, P, i64, P]
    L.or_forest_components.argtypes = [i64, P, P, i64, P]
    L.or_forest_components.restype = i64
    L.or_cell_build.argtypes = [P, P, i64, P, P, P]
    L.or_cell_build.restype = i64
    L.or_octants_all.argtypes = [i32, P, P, i64, i32, i32, P, P, P, P]
    L.or_boruvka.argtypes = [i32, P, P, i64, i32, i32, i32, P, P, P, P, P, P]
    L.or_boruvka.restype = i64
    L.or_prim_float.argtypes = [P, i64, i32, P, P]
    L.or_prim_float.restype = i64
    L.or_prim_fp64.argtypes = [P, i64, i32, P, P, P]
    L.or_prim_fp64.restype = i64
    L.or_weight.argtypes = [P, i32, P, P, i64]
    L.or_weight.restype = dbl
    L.or_kruskal.argtypes = [P, i64, i32, P, P, i64, P, P]
    L.or_kruskal.restype = i64
    L.or_threads.argtypes = [i32]
    L.or_threads.restype = i32
    return L


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        _lib = _bind(ctypes.CDLL(_LIB_PATH))
    return _lib


def build_native(out_dir):
    """The timed CPU baseline's build: the same source compiled with -O3
    -march=native for the host that runs it (into out_dir).  Returns the
    loaded library, for select()."""
    out = os.path.join(out_dir, "liboracle_emst_native.so")
    subprocess.run(["make", "-s", "-C", _HERE, "native", f"NATIVE_OUT={out}"], check=True)
    return _bind(ctypes.CDLL(out))


def select(L):
    """Make L (a library from lib() or build_native()) the one every function
    here calls; returns the previous one."""
    global _lib
    old = lib()
    _lib = L
    return old


def _p(a):
    return a.ctypes.data_as(ctypes.c_void_p)


def ncones(dim):
    return 2 * dim * (1 << (dim - 1))


def bbox(pts):
    pts = np.ascontiguousarray(pts, dtype=np.float32)
    n, dim = pts.shape
    lo = np.zeros(dim, np.float32)
    hi = np.zeros(dim, np.float32)
    lib().or_bbox(_p(pts), n, dim, _p(lo), _p(hi))
    return lo, hi


def rescale(pts):
    """A1+A2: returns (rescaled points, unitX, newmin[3], newmax[3])."""
    pts = np.ascontiguousarray(pts, dtype=np.float32)
    n, dim = pts.shape
    lo, hi = bbox(pts)
    out = np.empty_like(pts)
    nmin = np.zeros(3)
    nmax = np.zeros(3)
    u = lib().or_rescale(_p(pts), n, dim, _p(lo), _p(hi), _p(out), _p(nmin), _p(nmax))
    return out, u, nmin, nmax


def grid(dim, R, nmin, nmax, pdim=None):
    """Reference cell grid of dimension dim over points of dimension pdim
    (default dim; dim 2 over 3D points is the 2.5D mode's grid)."""
    g = Grid()
    lib().or_grid_init_pd(ctypes.byref(g), dim, pdim or dim, R,
                          _p(np.ascontiguousarray(nmin, np.float64)),
                          _p(np.ascontiguousarray(nmax, np.float64)))
    return g


def cell_build(g, rescaled):
    rescaled = np.ascontiguousarray(rescaled, dtype=np.float32)
    n = rescaled.shape[0]
    cell_of = np.zeros((n, 3), np.int32)
    start = np.zeros(g.ncell + 1, np.int64)
    cpts = np.zeros(max(n, 1), np.int32)
    outside = lib().or_cell_build(ctypes.byref(g), _p(rescaled), n, _p(cell_of), _p(start), _p(cpts))
    return cell_of, start, cpts[:n], outside


def d2(a, b):
    a = np.ascontiguousarray(a, dtype=np.float32)
    b = np.ascontiguousarray(b, dtype=np.float32)
    return float(lib().or_d2(_p(a), _p(b), a.shape[0]))


def octants(rescaled, labels=None, mode=0, R=1, cdim=None):
    """One cone-search stage over all points from the initial state.
    mode 0 = exact cone minimum, 1 = the reference ring search on an R grid;
    cdim = cone dimension (default the point dimension; 2 over 3D points is
    the 2.5D mode).  Returns (cand_id [n,K], cand_d2 [n,K], examined)."""
    rescaled = np.ascontiguousarray(rescaled, dtype=np.float32)
    n, dim = rescaled.shape
    cdim = cdim or dim
    K = ncones(cdim)
    if labels is None:
        labels = np.arange(n, dtype=np.int32)
    labels = np.ascontiguousarray(labels, dtype=np.int32)
    cid = np.full((n, K), -1, np.int32)
    cd2 = np.full((n, K), np.inf, np.float32)
    ex = ctypes.c_int64(0)
    g = None
    if mode == 1:
        _, _, nmin, nmax = _extremes(rescaled)
        g = grid(cdim, R, nmin, nmax, pdim=dim)
    lib().or_octants_all(mode, ctypes.byref(g) if g is not None else None, _p(rescaled), n, dim, cdim,
                         _p(labels), _p(cid), _p(cd2), ctypes.byref(ex))
    return cid, cd2, ex.value


def _extremes(rescaled):
    n, dim = rescaled.shape
    nmin = np.zeros(3)
    nmax = np.zeros(3)
    if n:
        nmin[:dim] = rescaled.min(axis=0).astype(np.float64)
        nmax[:dim] = rescaled.max(axis=0).astype(np.float64)
    return None, None, nmin, nmax


def boruvka(rescaled, mode=0, R=1, max_rounds=100, want_cands=False, cdim=None):
    """A8-A14 on already rescaled points.  Returns a dict with edges (u<v,
    in insertion order), per-edge round, final labels, trace.  cdim: cone
    dimension (default the point dimension; 2 over 3D points = 2.5D)."""
    rescaled = np.ascontiguousarray(rescaled, dtype=np.float32)
    n, dim = rescaled.shape
    cdim = cdim or dim
    K = ncones(cdim)
    eu = np.zeros(max(n - 1, 1), np.int32)
    ev = np.zeros(max(n - 1, 1), np.int32)
    er = np.zeros(max(n - 1, 1), np.int32)
    lab = np.zeros(max(n, 1), np.int32)
    cand = np.zeros((max(n, 1), K), np.int32) if want_cands else None
    tr = Trace()
    g = None
    if mode == 1:
        _, _, nmin, nmax = _extremes(rescaled)
        g = grid(cdim, R, nmin, nmax, pdim=dim)
    m = lib().or_boruvka(mode, ctypes.byref(g) if g is not None else None, _p(rescaled), n, dim, cdim,
                         max_rounds, _p(eu), _p(ev), _p(er), _p(lab),
                         _p(cand) if cand is not None else None, ctypes.byref(tr))
    rounds = tr.rounds
    return {
        "u": eu[:m].copy(), "v": ev[:m].copy(), "round": er[:m].copy(),
        "labels": lab[:n].copy(), "rounds": rounds,
        "cum_edges": list(tr.cum_edges[:rounds]),
        "searched": list(tr.searched[:rounds]),
        "examined": list(tr.examined[:rounds]),
        "overfull_cells": tr.overfull_cells,
        "cycles_cut": tr.cycles_cut,
        "ms_after_cells": tr.ms_after_cells,
        "cands": cand[:n] if cand is not None else None,
    }


def prim_float(rescaled):
    rescaled = np.ascontiguousarray(rescaled, dtype=np.float32)
    n, dim = rescaled.shape
    eu = np.zeros(max(n - 1, 1), np.int32)
    ev = np.zeros(max(n - 1, 1), np.int32)
    m = lib().or_prim_float(_p(rescaled), n, dim, _p(eu), _p(ev))
    return eu[:m].copy(), ev[:m].copy()


def prim_fp64(pts):
    pts = np.ascontiguousarray(pts, dtype=np.float32)
    n, dim = pts.shape
    eu = np.zeros(max(n - 1, 1), np.int32)
    ev = np.zeros(max(n - 1, 1), np.int32)
    w = ctypes.c_double(0)
    m = lib().or_prim_fp64(_p(pts), n, dim, _p(eu), _p(ev), ctypes.byref(w))
    return eu[:m].copy(), ev[:m].copy(), w.value


def threads(n=0):
    """Threads of the oracle's per-point searches; n > 0 sets them."""
    return lib().or_threads(n)


def weight(orig, u, v):
    orig = np.ascontiguousarray(orig, dtype=np.float32)
    u = np.ascontiguousarray(u, dtype=np.int32)
    v = np.ascontiguousarray(v, dtype=np.int32)
    return lib().or_weight(_p(orig), orig.shape[1], _p(u), _p(v), len(u))


def d2_pairs(rescaled, u, v):
    """A4 for every edge (u[i], v[i]): float32 d2 with the reference's fused
    operation order."""
    rescaled = np.ascontiguousarray(rescaled, dtype=np.float32)
    u = np.ascontiguousarray(u, dtype=np.int32)
    v = np.ascontiguousarray(v, dtype=np.int32)
    out = np.empty(len(u), np.float32)
    lib().or_d2_pairs(_p(rescaled), rescaled.shape[1], _p(u), _p(v), len(u), _p(out))
    return out


def forest_components(n, u, v):
    """(components of the graph on n vertices, edges that closed a cycle)."""
    u = np.ascontiguousarray(u, dtype=np.int32)
    v = np.ascontiguousarray(v, dtype=np.int32)
    cyc = ctypes.c_int64(0)
    comps = lib().or_forest_components(n, _p(u), _p(v), len(u), ctypes.byref(cyc))
    return comps, cyc.value


def kruskal(rescaled, cu, cv):
    """Kruskal under the float total order over the candidate pairs (cu[i],
    cv[i]); returns the chosen edges (u < v) in key order."""
    rescaled = np.ascontiguousarray(rescaled, dtype=np.float32)
    n, dim = rescaled.shape
    cu = np.ascontiguousarray(cu, dtype=np.int32)
    cv = np.ascontiguousarray(cv, dtype=np.int32)
    eu = np.zeros(max(n - 1, 1), np.int32)
    ev = np.zeros(max(n - 1, 1), np.int32)
    m = lib().or_kruskal(_p(rescaled), n, dim, _p(cu), _p(cv), len(cu), _p(eu), _p(ev))
    return eu[:m].copy(), ev[:m].copy()


def canon(u, v):
    """Edge set as a sorted (m,2) int array with u < v."""
    u = np.asarray(u, np.int64)
    v = np.asarray(v, np.int64)
    a = np.minimum(u, v)
    b = np.maximum(u, v)
    key = np.lexsort((b, a))
    return np.stack([a[key], b[key]], axis=1)


def emst(pts, mode=0, R=1, max_rounds=100, cdim=None):
    """Whole path on input coordinates: rescale, Boruvka, FP64 weight on the
    original coordinates.  cdim=2 on 3D points is the 2.5D mode (the rescale
    is the 3D one, SolutionEMST.cu:305-336 with DimP = 3)."""
    rs, unit, _, _ = rescale(pts)
    res = boruvka(rs, mode=mode, R=R, max_rounds=max_rounds, cdim=cdim)
    res["rescaled"] = rs
    res["unitX"] = unit
    res["weight"] = weight(pts, res["u"], res["v"])
    return res
