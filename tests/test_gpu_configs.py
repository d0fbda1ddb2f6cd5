"""Parity at the sizes of the benchmarked workloads, through the C ABI.

C1  10^4 uniform 2D points, seed 0: edges equal the complete-graph MST under
    the float order (oracle Prim), the FP64 weight equals the FP64 Prim's.
C2  10^6 uniform points, 2D seed 1 and 3D seed 2: edges equal the oracle's
    restatement of the reference ring search (mode 1), in both search forms.
C3  10^7 3D points, 64 Gaussians + 10% background, seed 3, and
C4  10^7 2D power-law points, seed 4: every form (cone search plain and with
    the shipped seed model, component search plain and with its model) builds
    the same tree; that tree is a spanning tree, every d2 equals the oracle's
    fused d2 bit for bit, the FP64 weight equals the oracle's sum, it holds
    every point's nearest neighbour (scipy k-d tree, independent of both),
    and Kruskal over the tree plus the 3 nearest neighbours of every point
    picks exactly the tree (no neighbour edge can replace a tree edge).
    C4 also equals the oracle's restatement of the reference ring search
    (mode 1) through committed digests of its edges and d2 bits
    (tests/golden/oracle_digests.json, made by make_oracle_digests.py: the
    oracle takes an hour on it).
C5  the 4096-cloud batch (seed 5): every cloud gets n-1 edges, every 64th
    cloud equals its own single-cloud run, every cloud of at most 5000
    points equals the oracle's complete-graph MST.
"""
import hashlib
import json
import os
from concurrent.futures import ThreadPoolExecutor

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

WTOL = 1e-9
DIGESTS = json.load(open(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "oracle_digests.json")))


def gpu_tree(pkg, pts, model=None, **kw):
    with pkg.Calculateur(dim=pts.shape[1], **kw) as c:
        if model is not None:
            c.set_radius_model(model)
        c.initialize(pts)
        c.run()
        u, v, d = c.getLinks()
        w, m = c.evaluate()
        resc, _ = c.getAdaptiveMap()
    return {"u": u, "v": v, "d": d, "w": w, "m": m, "resc": resc}


def edges(t):
    return np.stack([t["u"], t["v"]], axis=1).astype(np.int64)


def load_model(name):
    import __graft_entry__ as ge
    path = os.path.join(ge.PKG_DIR, "models", name)
    return np.load(path) if os.path.exists(path) else None


def check_tree(oracle_mod, pts, t):
    """Spanning tree in canonical order, exact fused d2 per edge, FP64
    weight on the input coordinates."""
    n = len(pts)
    u, v, d = t["u"], t["v"], t["d"]
    assert t["m"] == n - 1 and len(u) == n - 1
    assert np.all(u < v)
    key = u.astype(np.int64) * n + v
    assert np.all(np.diff(key) > 0)
    comps, cycles = oracle_mod.forest_components(n, u, v)
    assert comps == 1 and cycles == 0
    rescaled, _, _, _ = oracle_mod.rescale(pts)
    assert np.array_equal(t["resc"].view(np.uint32), rescaled.view(np.uint32))
    want = oracle_mod.d2_pairs(rescaled, u, v)
    assert np.array_equal(want.view(np.uint32), d.view(np.uint32)), \
        f"{int(np.sum(want.view(np.uint32) != d.view(np.uint32)))} edges with another d2"
    w = oracle_mod.weight(pts, u, v)
    assert abs(t["w"] - w) <= WTOL * w, (t["w"], w)
    return rescaled


def check_neighbours(oracle_mod, rescaled, t, k=4):
    """Independent of the library and of the reference's search: the tree
    holds every point's nearest neighbour, and Kruskal over the tree plus
    each point's k-1 nearest neighbours (scipy k-d tree, float64 distances)
    chooses exactly the tree."""
    from scipy.spatial import cKDTree

    n = len(rescaled)
    u, v = t["u"], t["v"]
    _, nb = cKDTree(rescaled).query(rescaled, k=k, workers=16)
    nb = nb.astype(np.int32)
    src = np.repeat(np.arange(n, dtype=np.int32), k - 1)
    dst = nb[:, 1:].reshape(-1)
    # smallest incident tree key per point must not exceed the key to the
    # k-d tree's nearest other point (a float64 ordering; a near-tie may put
    # an equal-d2 point of larger id first, hence <=, with the d2 equal)
    d2t = t["d"].view(np.uint32).astype(np.uint64)
    tk = np.full(n, np.iinfo(np.uint64).max, np.uint64)
    for a, b in ((u, v), (v, u)):
        np.minimum.at(tk, a, (d2t << np.uint64(32)) | b.astype(np.uint64))
    first = nb[:, 1]
    self_hit = first == np.arange(n)        # coincident points: the k-d tree may list the point itself second
    first = np.where(self_hit, nb[:, 0], first)
    dn = oracle_mod.d2_pairs(rescaled, np.arange(n, dtype=np.int32), first).view(np.uint32).astype(np.uint64)
    nk = (dn << np.uint64(32)) | first.astype(np.uint64)
    assert np.all((tk >> np.uint64(32)) <= (nk >> np.uint64(32)))
    cu = np.concatenate([u, src])
    cv = np.concatenate([v, dst])
    ku, kv = oracle_mod.kruskal(rescaled, cu, cv)
    assert np.array_equal(oracle_mod.canon(ku, kv), edges(t))


def test_c1_uniform_2d_1e4(pkg, oracle_mod):
    """C1: 10^4 uniform 2D points, seed 0, both search forms: the edge set is
    the complete-graph MST under (d2 float, min id, max id), the weight the
    FP64 Prim's on the input coordinates (1e-6 relative: the FP64 tree may
    order near-equal edges differently), the reference ring search agrees."""
    pts = pkg.synth(0, 10_000, 2, 0)
    print("case C1 10^4 2D seed 0", flush=True)
    rescaled, _, _, _ = oracle_mod.rescale(pts)
    pu, pv = oracle_mod.prim_float(rescaled)
    want = oracle_mod.canon(pu, pv)
    _, _, w64 = oracle_mod.prim_fp64(pts)
    ref = oracle_mod.emst(pts, mode=1, R=1)
    assert np.array_equal(oracle_mod.canon(ref["u"], ref["v"]), want)
    for rs in (0, 1):
        t = gpu_tree(pkg, pts, round_search=rs)
        check_tree(oracle_mod, pts, t)
        assert np.array_equal(edges(t), want), f"round_search={rs}"
        assert abs(t["w"] - w64) <= 1e-6 * w64, (t["w"], w64)


@pytest.mark.parametrize("dim,seed", [(2, 1), (3, 2)])
def test_c2_uniform_1e6(pkg, oracle_mod, dim, seed):
    """C2: 10^6 uniform points: both search forms equal the oracle's
    restatement of the reference ring search (R = 1) edge for edge, with the
    same d2 bits and weight."""
    n = 1_000_000
    pts = pkg.synth(0, n, dim, seed)
    print(f"case C2 10^6 {dim}D seed {seed}", flush=True)
    ref = oracle_mod.emst(pts, mode=1, R=1)
    want = oracle_mod.canon(ref["u"], ref["v"])
    for rs in (0, 1):
        t = gpu_tree(pkg, pts, round_search=rs)
        check_tree(oracle_mod, pts, t)
        assert np.array_equal(edges(t), want), f"round_search={rs}"
        assert abs(t["w"] - ref["weight"]) <= WTOL * ref["weight"]


@pytest.mark.parametrize("name,kind,dim,seed", [("C3", 1, 3, 3), ("C4", 2, 2, 4)])
def test_c3_c4_full_size(pkg, oracle_mod, name, kind, dim, seed):
    """C3 / C4 at 10^7 points: every form builds the same tree, and that tree
    passes the full-size checks (spanning, exact d2, weight, neighbours)."""
    n = 10_000_000
    pts = pkg.synth(kind, n, dim, seed)
    print(f"case {name} 10^7 {dim}D kind {kind} seed {seed}", flush=True)
    forms = {"cone_plain": gpu_tree(pkg, pts, round_search=0),
             "comp_plain": gpu_tree(pkg, pts, round_search=1)}
    cone_model = load_model(f"radius_{dim}d.npy")
    if cone_model is not None:
        forms["cone_learned"] = gpu_tree(pkg, pts, model=cone_model, round_search=0)
    comp_model = load_model(f"radius_comp_{dim}d.npy")
    if comp_model is not None:
        forms["comp_learned"] = gpu_tree(pkg, pts, model=comp_model, round_search=1)
    ref = forms["comp_plain"]
    for k, t in forms.items():
        assert np.array_equal(t["u"], ref["u"]) and np.array_equal(t["v"], ref["v"]), k
        assert np.array_equal(t["d"].view(np.uint32), ref["d"].view(np.uint32)), k
        assert t["w"] == ref["w"], k
    rescaled = check_tree(oracle_mod, pts, ref)
    check_neighbours(oracle_mod, rescaled, ref)
    dg = [v for v in DIGESTS.values() if (v["kind"], v["n"], v["dim"], v["seed"]) == (kind, n, dim, seed)]
    if name == "C4":
        assert len(dg) == 1
    for want in dg:
        e = edges(ref).astype("<i4")
        assert len(e) == want["edges"]
        assert hashlib.sha256(e.tobytes()).hexdigest() == want["edges_sha256"]
        assert hashlib.sha256(ref["d"].astype("<f4").tobytes()).hexdigest() == want["d2_sha256"]
        assert abs(ref["w"] - want["weight"]) <= WTOL * want["weight"]


def test_c5_batch_4096(pkg, oracle_mod):
    """C5: the 4096-cloud batch (seed 5), in both search forms: every cloud
    gets n-1 edges; every 64th cloud equals its single-cloud run (edges, d2
    bits, weight); every cloud of at most 5000 points equals the oracle's
    complete-graph MST under the float order."""
    sizes, dims, kinds, make = pkg.batch_workload(4096, 5)
    print("case C5 4096 clouds seed 5", flush=True)
    clouds = [make(i) for i in range(len(sizes))]

    def oracle_tree(i):
        rs, _, _, _ = oracle_mod.rescale(clouds[i])
        pu, pv = oracle_mod.prim_float(rs)
        return oracle_mod.canon(pu, pv)

    small = [i for i in range(len(sizes)) if sizes[i] <= 5000]
    with ThreadPoolExecutor(16) as ex:
        want_small = dict(zip(small, ex.map(oracle_tree, small)))
    for rs in (0, 1):
        for d in (2, 3):
            idx = np.nonzero(dims == d)[0]
            with pkg.Calculateur(dim=d, round_search=rs) as c:
                c.initialize_clouds([clouds[i] for i in idx])
                c.run()
                u, v, dd = c.getLinks()
                w, e, _ = c.cloud_results()
                off = c.offsets
            assert np.array_equal(e, np.maximum(sizes[idx] - 1, 0)), f"{d}D rs={rs}: a cloud is not a tree"
            cloud_of = np.searchsorted(off, u, side="right") - 1
            assert np.array_equal(cloud_of, np.searchsorted(off, v, side="right") - 1)
            starts = np.searchsorted(cloud_of, np.arange(len(idx) + 1))
            for j, i in enumerate(idx):
                a, b = starts[j], starts[j + 1]
                got = np.stack([u[a:b] - off[j], v[a:b] - off[j]], axis=1).astype(np.int64)
                if i in want_small:
                    assert np.array_equal(got, want_small[i]), f"cloud {i} ({sizes[i]} pts {d}D) rs={rs}"
                if i % 64 == 0:
                    t = gpu_tree(pkg, clouds[i], round_search=rs)
                    assert np.array_equal(got, edges(t)), f"cloud {i} rs={rs}"
                    assert np.array_equal(dd[a:b].view(np.uint32), t["d"].view(np.uint32))
                    assert abs(w[j] - t["w"]) <= 1e-12 * max(t["w"], 1.0)
