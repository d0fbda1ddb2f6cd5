"""ctypes binding of libemst_amd.so (the C ABI in include/emst.h).

The class `Calculateur` mirrors the reference's entry surface
(emst_coalition/include/CalculateurEMST.h:43-58: initialize, run, activate,
getLinks, getAdaptiveMap) over an in-memory point array.  The library must be
built (see build()); nothing here falls back to a CPU path: a missing or
unloadable library raises.
"""
import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libemst_amd.so")

EMST_OK = 0
STATUS = {0: "EMST_OK", 1: "EMST_E_ARG", 2: "EMST_E_HIP", 3: "EMST_E_OOM", 4: "EMST_E_NONFINITE",
          5: "EMST_E_STATE", 6: "EMST_E_NODEVICE", 7: "EMST_E_LIMIT", 8: "EMST_E_IO"}

# every symbol include/emst.h declares
EXPORTS = [
    "emst_default_options", "emst_create", "emst_destroy", "emst_last_error", "emst_set_points",
    "emst_run", "emst_step", "emst_get_edges", "emst_get_weight", "emst_get_labels",
    "emst_get_rescaled", "emst_get_trace", "emst_get_candidates", "emst_get_cell_grid",
    "emst_get_times", "emst_get_search_stats", "emst_synth", "emst_version",
    "emst_set_radius_model", "emst_get_radius_samples", "emst_mlp_train", "emst_mlp_eval", "emst_mlp_grad",
    "emst_set_clouds", "emst_get_cloud_results", "emst_autotune", "emst_get_shape", "emst_get_options",
    "emst_read_points",
    "emst_write_solution", "emst_weight_noalloc", "emst_get_round_ends",
]
CLI_PATH = os.path.join(_HERE, "emst_cli")

MLP_PARAMS = 4818


class EmstError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"{STATUS.get(code, code)}: {msg}")
        self.code = code


class Options(ctypes.Structure):
    _fields_ = [("dim", ctypes.c_int32), ("level_radius", ctypes.c_int32),
                ("max_rounds", ctypes.c_int32), ("bucket_size", ctypes.c_int32),
                ("trace", ctypes.c_int32), ("device", ctypes.c_int32),
                ("radius_mode", ctypes.c_int32), ("seed_scale", ctypes.c_float),
                ("cone_dim", ctypes.c_int32), ("search_split", ctypes.c_int32), ("round_search", ctypes.c_int32),
                ("log_samples", ctypes.c_int32), ("one_launch", ctypes.c_int32)]


class RoundTrace(ctypes.Structure):
    _fields_ = [("iteration", ctypes.c_int32), ("size", ctypes.c_int32), ("length", ctypes.c_double),
                ("time_flatten", ctypes.c_float), ("time_termination", ctypes.c_float),
                ("time_find_next_closest", ctypes.c_float), ("time_find_min_pair", ctypes.c_float),
                ("time_connect_union", ctypes.c_float), ("searched", ctypes.c_int64)]


def build(jobs=4):
    """Compile libemst_amd.so for gfx950 with hipcc (package Makefile)."""
    subprocess.run(["make", "-s", f"-j{jobs}", "-C", _HERE], check=True)


_lib = None


def lib():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"{LIB_PATH} is not built: run build() (hipcc, gfx950) first")
    L = ctypes.CDLL(LIB_PATH)
    P, i32, i64 = ctypes.c_void_p, ctypes.c_int32, ctypes.c_int64
    L.emst_default_options.argtypes = [P]
    L.emst_create.argtypes = [P]
    L.emst_create.restype = P
    L.emst_destroy.argtypes = [P]
    L.emst_last_error.argtypes = [P]
    L.emst_last_error.restype = ctypes.c_char_p
    L.emst_set_points.argtypes = [P, P, i64, i32, i32]
    L.emst_run.argtypes = [P]
    L.emst_step.argtypes = [P, P]
    L.emst_get_edges.argtypes = [P, P, P, P, P]
    L.emst_get_weight.argtypes = [P, P, P]
    L.emst_get_labels.argtypes = [P, P]
    L.emst_get_rescaled.argtypes = [P, P, P]
    L.emst_get_trace.argtypes = [P, P, P]
    L.emst_get_candidates.argtypes = [P, P, P, P]
    L.emst_get_cell_grid.argtypes = [P, i32, P, P, P, P]
    L.emst_get_times.argtypes = [P, P, P]
    L.emst_get_search_stats.argtypes = [P, P, P]
    L.emst_synth.argtypes = [i32, i64, i32, ctypes.c_uint64, P]
    L.emst_set_radius_model.argtypes = [P, P]
    L.emst_get_radius_samples.argtypes = [P, P, P, P]
    L.emst_mlp_train.argtypes = [i32, P, P, i64, i32, i32, ctypes.c_float, ctypes.c_uint64, P, P]
    L.emst_mlp_eval.argtypes = [i32, P, P, i64, P]
    L.emst_mlp_grad.argtypes = [i32, P, P, P, i64, P, P]
    L.emst_set_clouds.argtypes = [P, i32, P, P, i32, i32]
    L.emst_get_cloud_results.argtypes = [P, P, P, P]
    L.emst_autotune.argtypes = [P, P, P, P]
    L.emst_get_shape.argtypes = [P, P, P]
    L.emst_get_options.argtypes = [P, P]
    L.emst_read_points.argtypes = [ctypes.c_char_p, i32, P, P]
    L.emst_write_solution.argtypes = [P, ctypes.c_char_p]
    L.emst_weight_noalloc.argtypes = [P, i32, P, P, i64, i32, P, P]
    L.emst_get_round_ends.argtypes = [P, P, P]
    L.emst_version.restype = ctypes.c_char_p
    _lib = L
    return L


def _p(a):
    return None if a is None else a.ctypes.data_as(ctypes.c_void_p)


def synth(kind, n, dim, seed):
    """Synthetic cloud (0 uniform, 1 Gaussian mixture, 2 power-law), n x dim float32."""
    out = np.empty((n, dim), np.float32)
    rc = lib().emst_synth(kind, n, dim, seed, _p(out))
    if rc:
        raise EmstError(rc, "emst_synth")
    return out


def batch_workload(clouds=4096, seed=5):
    """The batched workload: `clouds` clouds with sizes round(exp(U[ln 1e2,
    ln 1e5])), half 2D and half 3D (shuffled), each uniform (kind 0) or eight
    Gaussians (kind 3) by a fair coin.  Returns (sizes, dims, kinds, make):
    make(i) is cloud i's points, synth(kinds[i], sizes[i], dims[i],
    seed * 1000003 + i)."""
    rng = np.random.default_rng(seed)
    sizes = np.rint(np.exp(rng.uniform(np.log(1e2), np.log(1e5), clouds))).astype(np.int64)
    dims = np.array([2] * (clouds // 2) + [3] * (clouds - clouds // 2))
    rng.shuffle(dims)
    kinds = np.where(rng.random(clouds) < 0.5, 0, 3)

    def make(i):
        return synth(int(kinds[i]), int(sizes[i]), int(dims[i]), seed * 1_000_003 + int(i))

    return sizes, dims, kinds, make


def read_points(path, dim):
    """InputRW::readFile: points of a .2d / .3d / .tsp / .points file as an
    n x dim float32 array (host only)."""
    L = lib()
    n = ctypes.c_int64(0)
    rc = L.emst_read_points(os.fsencode(path), dim, None, ctypes.byref(n))
    if rc:
        raise EmstError(rc, f"emst_read_points({path})")
    out = np.empty((n.value, dim), np.float32)
    if n.value:
        rc = L.emst_read_points(os.fsencode(path), dim, _p(out), ctypes.byref(n))
        if rc:
            raise EmstError(rc, f"emst_read_points({path})")
    return out


class Calculateur:
    """One EMST context on one device (CalculateurEMST / SolutionEMST)."""

    def __init__(self, dim=3, level_radius=1, max_rounds=100, bucket_size=0, trace=False, device=0,
                 seed_scale=0.0, cone_dim=0, search_split=0, round_search=0, log_samples=0, one_launch=0):
        """cone_dim=2 with dim=3: the 2.5D mode (SolutionEMST<3,2>, 3D points,
        wedges over the xy plane)."""
        L = lib()
        o = Options()
        L.emst_default_options(ctypes.byref(o))
        o.dim, o.level_radius, o.max_rounds = dim, level_radius, max_rounds
        o.bucket_size, o.trace, o.device = bucket_size, int(trace), device
        o.seed_scale = seed_scale
        o.cone_dim = cone_dim
        o.search_split = search_split
        o.round_search = round_search
        o.log_samples = log_samples
        o.one_launch = one_launch
        self._L = L
        self._c = L.emst_create(ctypes.byref(o))
        if not self._c:
            raise EmstError(6, L.emst_last_error(None).decode())
        self.dim = dim
        self.n = 0

    def close(self):
        if getattr(self, "_c", None):
            self._L.emst_destroy(self._c)
            self._c = None

    def __del__(self):
        self.close()

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def _chk(self, rc, what):
        if rc != EMST_OK:
            raise EmstError(rc, f"{what}: {self._L.emst_last_error(self._c).decode()}")

    # -- reference entry surface -----------------------------------------
    def initialize(self, points):
        """SolutionEMST::initialize(NetLink&, pMin, pMax) over a host array."""
        pts = np.ascontiguousarray(points, dtype=np.float32)
        self._keep = pts
        n, dim = pts.shape
        self._chk(self._L.emst_set_points(self._c, _p(pts), n, dim, 0), "set_points")
        self.n, self.dim = n, dim
        self.offsets = np.array([0, n], np.int64)

    def initialize_device(self, dev_ptr, n, dim):
        """Same, for points already resident on the device (AoS float32)."""
        self._chk(self._L.emst_set_points(self._c, ctypes.c_void_p(dev_ptr), n, dim, 1), "set_points")
        self.n, self.dim = n, dim
        self.offsets = np.array([0, n], np.int64)

    def initialize_clouds(self, clouds):
        """A batch of independent clouds of this context's dimension (list of
        n_i x dim arrays), packed back to back; edges then carry global ids
        (cloud offset + local id)."""
        pts = [np.ascontiguousarray(c, dtype=np.float32).reshape(-1, self.dim) for c in clouds]
        sizes = np.array([len(c) for c in pts], np.int64)
        packed = np.ascontiguousarray(np.concatenate(pts) if pts else np.zeros((0, self.dim), np.float32))
        self._keep = packed
        self._chk(self._L.emst_set_clouds(self._c, len(pts), _p(sizes), _p(packed), self.dim, 0), "set_clouds")
        self.n = int(sizes.sum())
        self.offsets = np.concatenate([[0], np.cumsum(sizes)]).astype(np.int64)

    def initialize_clouds_device(self, dev_ptr, sizes):
        """Same, for a packed batch already resident on the device."""
        sizes = np.ascontiguousarray(sizes, dtype=np.int64)
        self._sizes = sizes
        self._chk(self._L.emst_set_clouds(self._c, len(sizes), _p(sizes), ctypes.c_void_p(dev_ptr), self.dim, 1),
                  "set_clouds")
        self.n = int(sizes.sum())
        self.offsets = np.concatenate([[0], np.cumsum(sizes)]).astype(np.int64)

    def writeSolution(self, path):
        """SolutionEMST::writeSolution: .links, .correspondenceMap,
        .disjointSetMap and .points next to `path` (cut at its first '.' or '_')."""
        self._chk(self._L.emst_write_solution(self._c, os.fsencode(path)), "write_solution")

    def autotune(self):
        """Leaf size and seed scale picked by timing the current points; kept
        for later runs.  Returns (bucket, scale, build + search ms)."""
        b = ctypes.c_int32(0)
        sc = ctypes.c_float(0)
        ms = ctypes.c_float(0)
        self._chk(self._L.emst_autotune(self._c, ctypes.byref(b), ctypes.byref(sc), ctypes.byref(ms)), "autotune")
        return b.value, sc.value, ms.value

    def options(self):
        """The context's current options (after autotune: the chosen leaf
        size, seed scale and search form) as a dict of Calculateur arguments."""
        o = Options()
        self._chk(self._L.emst_get_options(self._c, ctypes.byref(o)), "options")
        return {"level_radius": o.level_radius, "max_rounds": o.max_rounds, "bucket_size": o.bucket_size,
                "trace": bool(o.trace), "device": o.device, "seed_scale": o.seed_scale, "cone_dim": o.cone_dim,
                "search_split": o.search_split, "round_search": o.round_search, "log_samples": o.log_samples,
                "one_launch": o.one_launch}

    def cloud_results(self):
        """Per cloud: (FP64 weight, edge count, completion round)."""
        k = len(self.offsets) - 1
        w = np.empty(k, np.float64)
        e = np.empty(k, np.int32)
        r = np.empty(k, np.int32)
        self._chk(self._L.emst_get_cloud_results(self._c, _p(w), _p(e), _p(r)), "cloud_results")
        return w, e, r

    def run(self):
        self._chk(self._L.emst_run(self._c), "run")

    def activate(self):
        """One Boruvka round; returns False once the termination test holds."""
        fin = ctypes.c_int32(0)
        self._chk(self._L.emst_step(self._c, ctypes.byref(fin)), "step")
        return fin.value == 0

    def getLinks(self):
        """Forest edges (u < v, sorted) and their float32 d2 (rescaled)."""
        m = ctypes.c_int64(0)
        self._chk(self._L.emst_get_edges(self._c, None, None, None, ctypes.byref(m)), "edges")
        u = np.empty(m.value, np.int32)
        v = np.empty(m.value, np.int32)
        d = np.empty(m.value, np.float32)
        self._chk(self._L.emst_get_edges(self._c, _p(u), _p(v), _p(d), ctypes.byref(m)), "edges")
        return u, v, d

    def getAdaptiveMap(self):
        out = np.empty((self.n, self.dim), np.float32)
        ux = ctypes.c_double(0)
        self._chk(self._L.emst_get_rescaled(self._c, _p(out), ctypes.byref(ux)), "rescaled")
        return out, ux.value

    def evaluate(self):
        """(FP64 weight on input coordinates, edge count)."""
        w = ctypes.c_double(0)
        m = ctypes.c_int64(0)
        self._chk(self._L.emst_get_weight(self._c, ctypes.byref(w), ctypes.byref(m)), "weight")
        return w.value, m.value

    def labels(self):
        out = np.empty(self.n, np.int32)
        if self.n:
            self._chk(self._L.emst_get_labels(self._c, _p(out)), "labels")
        return out

    def trace(self):
        k = ctypes.c_int32(0)
        self._chk(self._L.emst_get_trace(self._c, None, ctypes.byref(k)), "trace")
        rows = (RoundTrace * max(k.value, 1))()
        self._chk(self._L.emst_get_trace(self._c, rows, ctypes.byref(k)), "trace")
        return [dict((f, getattr(rows[i], f)) for f, _ in RoundTrace._fields_) for i in range(k.value)]

    def candidates(self):
        k = ctypes.c_int32(0)
        self._chk(self._L.emst_get_candidates(self._c, None, None, ctypes.byref(k)), "candidates")
        cand = np.empty((self.n, k.value), np.int32)
        cd2 = np.empty((self.n, k.value), np.float32)
        self._chk(self._L.emst_get_candidates(self._c, _p(cand), _p(cd2), ctypes.byref(k)), "candidates")
        return cand, cd2

    def cell_grid(self, level_radius=1):
        ext = np.zeros(3, np.int32)
        self._chk(self._L.emst_get_cell_grid(self._c, level_radius, _p(ext), None, None, None), "cell grid")
        ncell = int(ext[0]) * int(ext[1]) * int(ext[2])
        cell_of = np.empty((self.n, 3), np.int32)
        start = np.empty(ncell + 1, np.int64)
        cpts = np.empty(max(self.n, 1), np.int32)
        self._chk(self._L.emst_get_cell_grid(self._c, level_radius, _p(ext), _p(cell_of), _p(start),
                                             _p(cpts)), "cell grid")
        return ext, cell_of, start, cpts[: self.n]

    def set_radius_model(self, model):
        """Learned seed radius for the next runs (None = plain seed)."""
        if model is None:
            self._chk(self._L.emst_set_radius_model(self._c, None), "radius model")
            return
        m = np.ascontiguousarray(model, dtype=np.float32)
        assert m.size == MLP_PARAMS
        self._chk(self._L.emst_set_radius_model(self._c, _p(m)), "radius model")

    def radius_samples(self):
        """(features m x 8, targets m): every point's, from the last cone search
        (round_search 0), or those logged by the last component-search run
        (round_search 1 with log_samples=1)."""
        m = ctypes.c_int64(0)
        self._chk(self._L.emst_get_radius_samples(self._c, None, None, ctypes.byref(m)), "samples")
        X = np.empty((m.value, 8), np.float32)
        T = np.empty(m.value, np.float32)
        self._chk(self._L.emst_get_radius_samples(self._c, _p(X), _p(T), ctypes.byref(m)), "samples")
        return X, T

    def round_ends(self):
        """One-launch runs: ms from the run's start to the end of each round."""
        k = ctypes.c_int32(0)
        self._chk(self._L.emst_get_round_ends(self._c, None, ctypes.byref(k)), "round_ends")
        out = np.empty(k.value, np.float32)
        if k.value:
            self._chk(self._L.emst_get_round_ends(self._c, _p(out), ctypes.byref(k)), "round_ends")
        return out

    def times(self):
        t = (ctypes.c_float * 6)()
        k = ctypes.c_int32(6)
        self._chk(self._L.emst_get_times(self._c, t, ctypes.byref(k)), "times")
        return list(t)[: k.value]


def train_radius_model(X, T, epochs=8, batch=1024, lr=2e-3, seed=1, device=0):
    """Train the seed-radius MLP on the GPU; returns (model, loss per epoch)."""
    X = np.ascontiguousarray(X, dtype=np.float32)
    T = np.ascontiguousarray(T, dtype=np.float32)
    model = np.empty(MLP_PARAMS, np.float32)
    losses = np.empty(epochs, np.float32)
    rc = lib().emst_mlp_train(device, _p(X), _p(T), len(T), epochs, batch, ctypes.c_float(lr), seed,
                              _p(model), _p(losses))
    if rc:
        raise EmstError(rc, "emst_mlp_train")
    return model, losses


def mlp_eval(model, X, device=0):
    X = np.ascontiguousarray(X, dtype=np.float32)
    out = np.empty(len(X), np.float32)
    rc = lib().emst_mlp_eval(device, _p(np.ascontiguousarray(model, np.float32)), _p(X), len(X), _p(out))
    if rc:
        raise EmstError(rc, "emst_mlp_eval")
    return out


def _stats_names():
    return ["packets", "nodes", "leaves", "leaf_points", "pairs", "group_points", "grouped", "cycles_cut", "resumed",
            "comp_queries", "comp_nodes", "comp_leaves", "comp_pairs", "knn_nodes", "knn_leaves", "knn_pairs",
            "comp_walks", "knn_walks"]


def search_stats(calc):
    t = (ctypes.c_int64 * 18)()
    k = ctypes.c_int32(18)
    calc._chk(calc._L.emst_get_search_stats(calc._c, t, ctypes.byref(k)), "stats")
    return dict(zip(_stats_names(), list(t)[: k.value]))


def emst(points, **kw):
    """Whole path on a host array: returns (u, v, d2, weight)."""
    pts = np.ascontiguousarray(points, dtype=np.float32)
    with Calculateur(dim=pts.shape[1], **kw) as c:
        c.initialize(pts)
        c.run()
        u, v, d = c.getLinks()
        w, _ = c.evaluate()
    return u, v, d, w
