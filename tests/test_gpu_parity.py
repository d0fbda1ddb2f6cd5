"""GPU parity: the HIP path through the C ABI against the oracle and the
reference's own golden data.  Integer outputs (edges, labels, candidate ids,
cells) must be identical; float32 outputs (rescaled coordinates, d2) must be
bit-identical; the FP64 weight may differ from the oracle's sequential sum
only by summation order (relative 1e-12)."""
import json
import os

import numpy as np
import pytest

from conftest import GOLDEN, g6, golden_path, read_points

pytestmark = pytest.mark.gpu

FIX = json.load(open(os.path.join(GOLDEN, "fixtures.json")))
WTOL = 1e-12


def run_gpu(pkg, pts, **kw):
    with pkg.Calculateur(dim=pts.shape[1], **kw) as c:
        c.initialize(pts)
        c.run()
        u, v, d = c.getLinks()
        w, m = c.evaluate()
        lab = c.labels()
        tr = c.trace() if kw.get("trace") else None
    return dict(e=np.stack([u, v], axis=1).astype(np.int64), d2=d, w=w, m=m, labels=lab, trace=tr)


@pytest.mark.parametrize("name", sorted(FIX["traces"]))
def test_reference_traces(pkg, oracle_mod, name):
    """Per-round forest size and length equal the reference's logged runs
    (emst_coalition/test_emst/test_*_2/output), and the final tree equals the
    oracle's."""
    pts = read_points(golden_path(name))
    print("case", name, pts.shape, flush=True)
    g = run_gpu(pkg, pts, trace=True)
    rows = FIX["traces"][name]
    assert [r["size"] for r in g["trace"]] == [r["size"] for r in rows]
    assert [g6(np.float32(r["length"])) for r in g["trace"]] == [r["length"] for r in rows]
    ref = oracle_mod.emst(pts, mode=0 if len(pts) <= 4000 else 1)
    assert np.array_equal(g["e"], oracle_mod.canon(ref["u"], ref["v"]))
    assert np.array_equal(g["labels"], ref["labels"])
    assert abs(g["w"] - ref["weight"]) <= WTOL * ref["weight"]


def test_outypu_full_output(pkg, oracle_mod):
    """3000randomCity.2d: edge set of outypu.links, the single root offset 840
    of outypu.disjointSetMap (point 210, 4-byte labels) and the rescaled
    coordinates of outypu.points."""
    pts = read_points(golden_path("3000randomCity.2d"))
    with pkg.Calculateur(dim=2) as c:
        c.initialize(pts)
        c.run()
        u, v, _ = c.getLinks()
        lab = c.labels()
        resc, unit = c.getAdaptiveMap()
    assert sorted(zip(u.tolist(), v.tolist())) == [tuple(e) for e in FIX["outypu"]["edges"]]
    assert sorted(set((lab * 4).tolist())) == FIX["outypu"]["disjoint_set_values"]
    ores, ounit, _, _ = oracle_mod.rescale(pts)
    assert unit == ounit
    assert np.array_equal(resc.view(np.uint32), ores.view(np.uint32))
    printed = open(golden_path("outypu.points")).read().split()[6:]
    printed = [x for i, x in enumerate(printed) if i % 3 != 2]
    assert ["%.6g" % x for x in resc.astype(np.float64).ravel()] == printed


@pytest.mark.parametrize("dim,kind,n,seed", [(2, 0, 3000, 1), (3, 0, 3000, 2), (2, 1, 2500, 3),
                                             (3, 1, 2500, 4), (2, 2, 2000, 5), (3, 2, 2000, 6),
                                             (3, 0, 70, 7), (2, 0, 9, 8)])
def test_cone_candidates_bit_exact(pkg, oracle_mod, dim, kind, n, seed):
    """Round-1 cone minima (id and float32 d2 per cone) equal the oracle's
    exhaustive restatement of NodeComputeOctantMD's contract."""
    pts = pkg.synth(kind, n, dim, seed)
    print("case", dim, kind, n, seed, flush=True)
    with pkg.Calculateur(dim=dim) as c:
        c.initialize(pts)
        cand, cd2 = c.candidates()
        resc, _ = c.getAdaptiveMap()
    ocid, ocd2, _ = oracle_mod.octants(resc, mode=0)
    assert np.array_equal(cand, ocid)
    assert np.array_equal(cd2.view(np.uint32), ocd2.view(np.uint32))


@pytest.mark.parametrize("dim,split", [(3, 1), (3, 2), (3, 3), (3, 0), (2, 2), (2, 3)])
def test_split_search_forms(pkg, oracle_mod, dim, split):
    """The round-1 cone search in one launch, or a first pass followed by the
    queries left with open cones, resumed in place or compacted into full
    waves: every form gives the exact cone minima."""
    pts = pkg.synth(1, 12000, dim, 81)
    print("case", dim, split, flush=True)
    with pkg.Calculateur(dim=dim, search_split=split) as c:
        c.initialize(pts)
        cand, cd2 = c.candidates()
        st = pkg.search_stats(c)
        resc, _ = c.getAdaptiveMap()
    assert (st["resumed"] > 0) == (split in (2, 3))      # auto: one launch below 2^21 points
    ocid, ocd2, _ = oracle_mod.octants(resc, mode=0)
    assert np.array_equal(cand, ocid)
    assert np.array_equal(cd2.view(np.uint32), ocd2.view(np.uint32))


@pytest.mark.parametrize("dim,kind,n,seed", [(2, 0, 6000, 21), (3, 0, 6000, 22), (2, 1, 6000, 23),
                                             (3, 1, 6000, 24), (2, 2, 5000, 25), (3, 2, 5000, 26)])
def test_tree_equals_complete_graph_prim(pkg, oracle_mod, dim, kind, n, seed):
    """The tree is the unique MST of the complete graph under
    (float32 d2, min id, max id) in rescaled coordinates."""
    pts = pkg.synth(kind, n, dim, seed)
    print("case", dim, kind, n, seed, flush=True)
    g = run_gpu(pkg, pts)
    rs, _, _, _ = oracle_mod.rescale(pts)
    pu, pv = oracle_mod.prim_float(rs)
    assert np.array_equal(g["e"], oracle_mod.canon(pu, pv))
    w = oracle_mod.weight(pts, pu, pv)
    assert abs(g["w"] - w) <= WTOL * w
    # FP64 Kruskal/Prim on exact distances in input coordinates
    _, _, w64 = oracle_mod.prim_fp64(pts)
    assert abs(g["w"] - w64) <= 1e-6 * w64


def test_forest_after_each_round(pkg, oracle_mod):
    """activate(): the forest after round r is the oracle's edges of rounds <= r."""
    pts = read_points(golden_path("4000randomCity.2d"))
    ref = oracle_mod.emst(pts, mode=0)
    with pkg.Calculateur(dim=2) as c:
        c.initialize(pts)
        r = 0
        while c.activate():
            r += 1
            u, v, _ = c.getLinks()
            sel = ref["round"] <= r
            assert np.array_equal(np.stack([u, v], 1).astype(np.int64),
                                  oracle_mod.canon(ref["u"][sel], ref["v"][sel]))
        assert r == ref["rounds"]


@pytest.mark.parametrize("name,dim", [("1000randomCity.2d", 2), ("2000randomCity.3d", 3)])
@pytest.mark.parametrize("R", [1, 2, 3])
def test_reference_cell_grid(pkg, oracle_mod, name, dim, R):
    """findCell per point and the per-cell point lists equal the oracle's
    restatement of ViewGridQuadMD (ViewGrid.h:251-359)."""
    pts = read_points(golden_path(name))
    with pkg.Calculateur(dim=dim) as c:
        c.initialize(pts)
        ext, cell_of, start, cpts = c.cell_grid(R)
        resc, _ = c.getAdaptiveMap()
    _, _, nmin, nmax = oracle_mod.rescale(pts)
    g = oracle_mod.grid(dim, R, nmin, nmax)
    ocell, ostart, ocpts, outside = oracle_mod.cell_build(g, resc)
    assert outside == 0
    assert list(ext) == list(g.ext_dual)
    assert np.array_equal(cell_of, ocell)
    assert np.array_equal(start, ostart)
    assert np.array_equal(cpts, ocpts)


def test_duplicates_tsp(pkg, oracle_mod):
    """test20.tsp holds a duplicated city (lines 7-8): zero-length edge and the
    repeated cone search of NodeEMST.h:1051-1077."""
    pts = read_points(golden_path("test20.tsp"))
    g = run_gpu(pkg, pts, trace=True)
    ref = oracle_mod.emst(pts, mode=0)
    assert np.array_equal(g["e"], oracle_mod.canon(ref["u"], ref["v"]))
    assert [r["size"] for r in g["trace"]] == ref["cum_edges"]
    assert np.array_equal(g["labels"], ref["labels"])


@pytest.mark.parametrize("case", ["identical", "pairs", "lattice2", "lattice3", "line", "two"])
def test_degenerate_inputs(pkg, oracle_mod, case):
    """Duplicates, exact distance ties and collinear points: the tie order
    (d2, min id, max id) decides every choice."""
    rng = np.random.default_rng(5)
    if case == "identical":
        pts = np.full((40, 3), 0.25, np.float32)
    elif case == "pairs":
        base = rng.random((60, 2), dtype=np.float32)
        pts = np.concatenate([base, base[::-1], base[:7]], axis=0)
    elif case == "lattice2":
        g = np.stack(np.meshgrid(np.arange(12), np.arange(9)), -1).reshape(-1, 2)
        pts = rng.permutation(g).astype(np.float32)
    elif case == "lattice3":
        g = np.stack(np.meshgrid(np.arange(6), np.arange(5), np.arange(4)), -1).reshape(-1, 3)
        pts = rng.permutation(g).astype(np.float32)
    elif case == "line":
        pts = np.stack([rng.random(300).astype(np.float32), np.zeros(300, np.float32)], 1)
    else:
        pts = np.array([[0.0, 0.0], [1.0, 1.0]], np.float32)
    print("case", case, pts.shape, flush=True)
    g = run_gpu(pkg, pts, trace=True)
    ref = oracle_mod.emst(pts, mode=0)
    assert np.array_equal(g["e"], oracle_mod.canon(ref["u"], ref["v"]))
    assert [r["size"] for r in g["trace"]] == ref["cum_edges"]
    assert np.array_equal(g["labels"], ref["labels"])


def _morton_keys(resc):
    """Host restatement of the Morton quantisation (per-axis cell index)."""
    resc = np.asarray(resc, np.float32)
    dim = resc.shape[1]
    maxq = (1 << (16 if dim == 3 else 24)) - 1
    mx = np.float32(resc.max())
    qs = np.float32(maxq / np.float64(mx))
    q = np.minimum((resc * qs).astype(np.float64), maxq).astype(np.uint64)
    return [tuple(r) for r in q]


def _coincident_case(case, oracle_mod):
    rng = np.random.default_rng(17)
    if case == "groups2":
        base = rng.random((1500, 2), dtype=np.float32)
        pts = np.repeat(base, rng.integers(1, 7, len(base)), axis=0)
    elif case == "groups3":
        base = rng.random((1200, 3), dtype=np.float32)
        pts = np.repeat(base, rng.integers(1, 5, len(base)), axis=0)
    elif case == "big_group":
        pts = np.concatenate([rng.random((2000, 2), dtype=np.float32), np.full((3000, 2), 0.5, np.float32)])
    elif case == "huge_groups":
        # two locations repeated beyond what one wave walks (k_dup_big)
        pts = np.concatenate([rng.random((1500, 2), dtype=np.float32), np.full((5000, 2), 0.25, np.float32),
                              np.full((4500, 2), 0.75, np.float32)])
    elif case == "near_twins":
        base = rng.random((800, 3), dtype=np.float32)
        step = np.nextafter(base, np.float32(2))
        pts = np.concatenate([base, step, base[:300], step[:200]])
    elif case in ("wide_run", "huge_mixed"):
        # points on two distinct locations of one Morton cell: a run too long
        # to split (1400 points; 5000, one the grid fills), so grouping turns off
        copies = 700 if case == "wide_run" else 2500
        other = rng.random((500, 3), dtype=np.float32)
        for a in np.linspace(0.3, 0.7, 41, dtype=np.float32):
            b = np.float32(a + np.spacing(a))
            pts = np.concatenate([other, np.full((copies, 3), a, np.float32), np.full((copies, 3), b, np.float32)])
            rs = oracle_mod.rescale(pts)[0]
            keys = _morton_keys(rs)
            if not np.array_equal(rs[500], rs[500 + copies]) and keys[500] == keys[500 + copies]:
                break
        else:
            pytest.skip("no two locations found in one Morton cell")
    else:   # underflow: distinct points at float distance 0, plus copies
        pts = np.concatenate([rng.random((200, 2), dtype=np.float32),
                              np.zeros((3, 2), np.float32), np.array([[1e-30, 0.0]], np.float32)])
    return rng.permutation(pts)


@pytest.mark.parametrize("case,grouped", [("groups2", 1), ("groups3", 1), ("big_group", 1), ("huge_groups", 1),
                                          ("near_twins", 1), ("wide_run", 0), ("huge_mixed", 0), ("underflow", 0)])
def test_coincident_groups(pkg, oracle_mod, case, grouped):
    """Points sharing a location search once per location; their cones,
    the tree, the per-round sizes and the labels equal the per-point oracle.
    Long mixed runs and distinct points at float distance 0 search per point."""
    pts = _coincident_case(case, oracle_mod)
    print("case", case, pts.shape, flush=True)
    with pkg.Calculateur(dim=pts.shape[1], trace=True) as c:
        c.initialize(pts)
        cand, cd2 = c.candidates()
        st = pkg.search_stats(c)
        resc, _ = c.getAdaptiveMap()
        c.run()
        u, v, _ = c.getLinks()
        tr = c.trace()
        lab = c.labels()
    _, counts = np.unique(resc, axis=0, return_counts=True)
    assert st["grouped"] == grouped
    if grouped:
        assert st["group_points"] == int(counts[counts > 1].sum())
    ocid, ocd2, _ = oracle_mod.octants(resc, mode=0)
    assert np.array_equal(cand, ocid)
    assert np.array_equal(cd2.view(np.uint32), ocd2.view(np.uint32))
    ref = oracle_mod.emst(pts, mode=0)
    assert np.array_equal(np.stack([u, v], 1).astype(np.int64), oracle_mod.canon(ref["u"], ref["v"]))
    assert [r["size"] for r in tr] == ref["cum_edges"]
    assert np.array_equal(lab, ref["labels"])


@pytest.mark.parametrize("case,split", [("groups3", 2), ("groups3", 3), ("groups2", 3), ("near_twins", 3)])
def test_coincident_groups_split_search(pkg, oracle_mod, case, split):
    """Coincident groups with the split cone search (first pass, then the
    open queries resumed in place or compacted): the representatives' saved
    '+' cones and every point's cones stay exact."""
    pts = _coincident_case(case, oracle_mod)
    print("case", case, split, pts.shape, flush=True)
    with pkg.Calculateur(dim=pts.shape[1], search_split=split) as c:
        c.initialize(pts)
        cand, cd2 = c.candidates()
        st = pkg.search_stats(c)
        resc, _ = c.getAdaptiveMap()
        c.run()
        u, v, _ = c.getLinks()
    assert st["grouped"] == 1 and st["resumed"] > 0
    ocid, ocd2, _ = oracle_mod.octants(resc, mode=0)
    assert np.array_equal(cand, ocid)
    assert np.array_equal(cd2.view(np.uint32), ocd2.view(np.uint32))
    ref = oracle_mod.emst(pts, mode=0)
    assert np.array_equal(np.stack([u, v], 1).astype(np.int64), oracle_mod.canon(ref["u"], ref["v"]))


@pytest.mark.parametrize("n", [0, 1])
def test_tiny(pkg, n):
    pts = np.zeros((n, 3), np.float32)
    g = run_gpu(pkg, pts)
    assert g["m"] == 0 and len(g["e"]) == 0 and g["w"] == 0


def test_nonfinite_rejected(pkg):
    pts = np.zeros((10, 2), np.float32)
    pts[3, 1] = np.nan
    with pkg.Calculateur(dim=2) as c:
        c.initialize(pts)
        with pytest.raises(pkg.EmstError) as ei:
            c.run()
        assert ei.value.code == 4


def test_round_cap(pkg, oracle_mod):
    pts = pkg.synth(0, 5000, 2, 31)
    ref = oracle_mod.emst(pts, mode=0, max_rounds=2)
    with pkg.Calculateur(dim=2, max_rounds=2) as c:
        c.initialize(pts)
        c.run()
        u, v, _ = c.getLinks()
    assert np.array_equal(np.stack([u, v], 1).astype(np.int64), oracle_mod.canon(ref["u"], ref["v"]))


def _is_spanning_tree(n, u, v):
    parent = np.arange(n)

    def find(x):
        while parent[x] != x:
            parent[x] = parent[parent[x]]
            x = parent[x]
        return x

    for a, b in zip(u.tolist(), v.tolist()):
        ra, rb = find(a), find(b)
        if ra == rb:
            return False
        parent[ra] = rb
    return len(u) == n - 1


@pytest.mark.parametrize("dim,kind", [(3, 1), (2, 2), (3, 0)])
def test_large_ring_search_parity(pkg, oracle_mod, dim, kind):
    """2e5 points: the tree equals the oracle's restatement of the reference's
    ring search (R = 1), and is a spanning tree; the tree does not depend on
    the leaf size of the search hierarchy."""
    n = 200_000
    pts = pkg.synth(kind, n, dim, 41 + kind)
    print("case", dim, kind, n, flush=True)
    g = run_gpu(pkg, pts)
    g2 = run_gpu(pkg, pts, bucket_size=16)
    assert np.array_equal(g["e"], g2["e"])
    assert _is_spanning_tree(n, g["e"][:, 0], g["e"][:, 1])
    if kind == 0:
        ref = oracle_mod.emst(pts, mode=1, R=1)
        assert np.array_equal(g["e"], oracle_mod.canon(ref["u"], ref["v"]))
        assert abs(g["w"] - ref["weight"]) <= WTOL * ref["weight"]


@pytest.mark.parametrize("dim,kind", [(3, 1), (2, 0)])
def test_full_size_properties(pkg, oracle_mod, dim, kind):
    """2e6 points: spanning tree, canonical order, every d2 equal bit for bit
    to the oracle's fused d2 of the rescaled coordinates, weight equals the
    sum of edge lengths, two runs identical."""
    n = 2_000_000
    pts = pkg.synth(kind, n, dim, 77)
    print("case", dim, kind, n, flush=True)
    with pkg.Calculateur(dim=dim) as c:
        c.initialize(pts)
        c.run()
        u, v, d = c.getLinks()
        w, m = c.evaluate()
        resc, _ = c.getAdaptiveMap()
        c.initialize(pts)
        c.run()
        u2, v2, d2 = c.getLinks()
    assert m == n - 1
    assert np.array_equal(u, u2) and np.array_equal(v, v2) and np.array_equal(d.view(np.uint32), d2.view(np.uint32))
    assert np.all(u < v)
    key = u.astype(np.int64) * n + v
    assert np.all(np.diff(key) > 0)
    assert _is_spanning_tree(n, u, v)
    rescaled, _, _, _ = oracle_mod.rescale(pts)
    assert np.array_equal(resc.view(np.uint32), rescaled.view(np.uint32))
    rd2 = oracle_mod.d2_pairs(rescaled, u, v)
    assert np.array_equal(rd2.view(np.uint32), d.view(np.uint32))
    a, b = pts[u].astype(np.float64), pts[v].astype(np.float64)
    w64 = np.sqrt(((a - b) ** 2).sum(1)).sum()
    assert abs(w - w64) <= 1e-6 * w64
