"""TSPLIB sweep: every TSPLIB instance of the reference's 2D tests (79 files,
up to 238 025 cities), read with the library's TSPLIB reader.  Every
instance: the forest is a spanning tree whose FP64 weight equals the exact
EMST weight from a Delaunay triangulation (scipy, float64 input coordinates)
to 1e-6 relative (the tree is chosen in rescaled float32, so near-ties may
pick another tree of the same weight to that precision).  Instances up to
6000 cities: the edge set and the per-round forest sizes equal the oracle's
exact Boruvka."""
import gzip
import os

import numpy as np
import pytest

from conftest import GOLDEN

pytestmark = pytest.mark.gpu
TSP_DIR = os.path.join(GOLDEN, "tsplib")
NAMES = sorted(os.listdir(TSP_DIR))


def _load(pkg, name, tmp_path):
    p = tmp_path / name[:-3]
    p.write_bytes(gzip.open(os.path.join(TSP_DIR, name)).read())
    return pkg.read_points(str(p), 2)


def delaunay_emst_weight(pts):
    from scipy.sparse import coo_matrix
    from scipy.sparse.csgraph import minimum_spanning_tree
    from scipy.spatial import Delaunay
    x = np.unique(pts.astype(np.float64), axis=0)      # coincident cities add zero-length edges
    tri = Delaunay(x).simplices
    e = np.concatenate([tri[:, [0, 1]], tri[:, [1, 2]], tri[:, [0, 2]]])
    e = np.unique(np.sort(e, axis=1), axis=0)
    w = np.sqrt(((x[e[:, 0]] - x[e[:, 1]]) ** 2).sum(1))
    g = coo_matrix((w, (e[:, 0], e[:, 1])), shape=(len(x), len(x)))
    return float(minimum_spanning_tree(g).sum())


def spanning(n, u, v):
    from scipy.sparse import coo_matrix
    from scipy.sparse.csgraph import connected_components
    if len(u) != n - 1:
        return False
    g = coo_matrix((np.ones(len(u)), (u, v)), shape=(n, n))
    return connected_components(g, directed=False)[0] == 1


@pytest.mark.parametrize("name", NAMES)
def test_tsplib_instance(pkg, oracle_mod, name, tmp_path):
    pts = _load(pkg, name, tmp_path)
    n = len(pts)
    print("tsplib", name, n, flush=True)
    with pkg.Calculateur(dim=2, trace=n <= 6000) as c:
        c.initialize(pts)
        c.run()
        u, v, _ = c.getLinks()
        w, m = c.evaluate()
        tr = c.trace() if n <= 6000 else None
    assert m == n - 1 and spanning(n, u, v)
    wd = delaunay_emst_weight(pts)
    assert abs(w - wd) <= 1e-6 * wd, (w, wd)
    if n <= 6000:
        ref = oracle_mod.emst(pts, mode=0)
        assert np.array_equal(np.stack([u, v], 1).astype(np.int64), oracle_mod.canon(ref["u"], ref["v"]))
        assert [r["size"] for r in tr] == ref["cum_edges"]
