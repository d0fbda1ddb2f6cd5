"""GPU parity of the 2.5D mode (SolutionEMST<3,2>: 3D points, cones over the
xy plane, 3D distances; emst_coalition/src/CalculateurEMST.cu:19,154-160)
through the C ABI, against the oracle: wedge minima (id and float32 d2)
bit-exact, edge sets, per-round forest sizes, labels and the cycle rule's
withheld hooks identical; the FP64 weight within summation order."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

WTOL = 1e-12


def cloud(seed, n, zscale, ties=False):
    rng = np.random.default_rng(seed)
    pts = rng.random((n, 3), dtype=np.float32)
    pts[:, 2] *= np.float32(zscale)
    if ties:
        # an xy lattice (wedge-boundary ties), a column of points sharing one xy
        # (zero xy offset: every '+' wedge) and exact duplicates
        k = n // 4
        pts[:k, :2] = np.floor(pts[:k, :2] * 8) / 8
        pts[k:k + 6, :2] = pts[k, :2]
        pts[k + 6:k + 9] = pts[k + 9]
    return pts


def run25(pkg, pts, trace=False, step=False, **kw):
    with pkg.Calculateur(dim=3, cone_dim=2, trace=trace, **kw) as c:
        c.initialize(pts)
        if step:
            while c.activate():
                pass
        else:
            c.run()
        u, v, d = c.getLinks()
        w, m = c.evaluate()
        out = dict(e=np.stack([u, v], 1).astype(np.int64), d2=d, w=w, m=m, labels=c.labels(),
                   cut=pkg.search_stats(c)["cycles_cut"], trace=c.trace() if trace else None)
    return out


@pytest.mark.parametrize("seed,n,zscale,ties,split", [(1, 3000, 1.0, False, 0), (2, 2500, 0.05, True, 0),
                                                      (3, 2000, 5.0, True, 0), (4, 70, 1.0, False, 0),
                                                      (5, 3000, 1.0, True, 2), (6, 3000, 0.3, False, 3)])
def test_wedge_candidates_bit_exact(pkg, oracle_mod, seed, n, zscale, ties, split):
    pts = cloud(seed, n, zscale, ties)
    print("case", seed, n, zscale, ties, split, flush=True)
    with pkg.Calculateur(dim=3, cone_dim=2, search_split=split) as c:
        c.initialize(pts)
        cand, cd2 = c.candidates()
        resc, _ = c.getAdaptiveMap()
    ores, _, _, _ = oracle_mod.rescale(pts)
    assert np.array_equal(resc.view(np.uint32), ores.view(np.uint32))
    ocid, ocd2, _ = oracle_mod.octants(resc, mode=0, cdim=2)
    assert cand.shape == (n, 8)
    assert np.array_equal(cand, ocid)
    assert np.array_equal(cd2.view(np.uint32), ocd2.view(np.uint32))


@pytest.mark.parametrize("seed,n,zscale,ties", [(11, 4000, 0.01, False), (12, 4000, 0.3, False),
                                                (13, 4000, 1.0, True), (14, 3000, 10.0, True),
                                                (15, 5000, 2.0, False)])
def test_forest_and_cycle_rule_equal_oracle(pkg, oracle_mod, seed, n, zscale, ties):
    """Edges, per-round forest sizes, labels and the number of withheld hooks
    equal the oracle's (exact wedge minima), from the graph-replayed rounds."""
    pts = cloud(seed, n, zscale, ties)
    print("case", seed, n, zscale, ties, flush=True)
    g = run25(pkg, pts)
    ref = oracle_mod.emst(pts, mode=0, cdim=2)
    assert np.array_equal(g["e"], oracle_mod.canon(ref["u"], ref["v"]))
    assert np.array_equal(g["labels"], ref["labels"])
    assert g["cut"] == ref["cycles_cut"]
    assert abs(g["w"] - ref["weight"]) <= WTOL * ref["weight"]


@pytest.mark.parametrize("seed,zscale", [(21, 1.0), (22, 0.2)])
def test_rounds_one_at_a_time(pkg, oracle_mod, seed, zscale):
    """activate() rounds with the trace: the forest size after every round
    equals the oracle's, and the result equals the replayed-graph run."""
    pts = cloud(seed, 3000, zscale, True)
    print("case", seed, zscale, flush=True)
    ref = oracle_mod.emst(pts, mode=0, cdim=2)
    s = run25(pkg, pts, trace=True, step=True)
    assert [r["size"] for r in s["trace"]] == ref["cum_edges"]
    g = run25(pkg, pts)
    assert np.array_equal(s["e"], g["e"])
    assert s["cut"] == g["cut"] == ref["cycles_cut"]


@pytest.mark.parametrize("seed,n,zscale", [(31, 60000, 1.0), (32, 100000, 0.1)])
def test_large_against_ring_search(pkg, oracle_mod, seed, n, zscale):
    """Sizes where the exhaustive oracle is too slow: the oracle's restated
    ring search (R = 1, a 2D grid over x and y) builds the same tree."""
    pts = cloud(seed, n, zscale)
    print("case", seed, n, zscale, flush=True)
    g = run25(pkg, pts)
    ref = oracle_mod.emst(pts, mode=1, R=1, cdim=2)
    assert np.array_equal(g["e"], oracle_mod.canon(ref["u"], ref["v"]))
    assert g["cut"] == ref["cycles_cut"]


def test_flat_input_equals_the_2d_mode(pkg):
    pts = cloud(41, 20000, 0.0)
    pts[:, 2] = np.float32(0.25)
    g = run25(pkg, pts)
    with pkg.Calculateur(dim=2) as c:
        c.initialize(np.ascontiguousarray(pts[:, :2]))
        c.run()
        u, v, _ = c.getLinks()
    assert g["cut"] == 0
    assert np.array_equal(g["e"], np.stack([u, v], 1).astype(np.int64))


def test_batch_equals_single_clouds(pkg):
    rng = np.random.default_rng(51)
    clouds = [cloud(60 + i, int(rng.integers(50, 3000)), float(rng.choice([0.1, 1.0, 4.0])), i % 2 == 0)
              for i in range(9)]
    with pkg.Calculateur(dim=3, cone_dim=2) as c:
        c.initialize_clouds(clouds)
        c.run()
        u, v, d = c.getLinks()
        w, e, r = c.cloud_results()
        off = c.offsets
    for i, pts in enumerate(clouds):
        s = run25(pkg, pts)
        sel = (u >= off[i]) & (u < off[i + 1])
        eb = np.stack([u[sel] - off[i], v[sel] - off[i]], 1).astype(np.int64)
        assert np.array_equal(eb, s["e"])
        assert np.array_equal(d[sel].view(np.uint32), s["d2"].view(np.uint32))
        assert e[i] == len(pts) - 1
        assert abs(w[i] - s["w"]) <= WTOL * s["w"]


def test_cell_grid_over_xy(pkg, oracle_mod):
    """The reference's 2D cell grid of the 2.5D mode: extents and cells from
    x and y only (SolutionEMST.cu:422-430 with DimCM = 2)."""
    pts = cloud(71, 5000, 3.0)
    with pkg.Calculateur(dim=3, cone_dim=2) as c:
        c.initialize(pts)
        ext, cell_of, start, cpts = c.cell_grid(2)
        resc, _ = c.getAdaptiveMap()
    _, _, nmin, nmax = oracle_mod.rescale(pts)
    g = oracle_mod.grid(2, 2, nmin, nmax, pdim=3)
    ocell, ostart, ocpts, outside = oracle_mod.cell_build(g, resc)
    assert outside == 0
    assert list(ext) == list(g.ext_dual) and ext[2] == 1
    assert np.array_equal(cell_of, ocell)
    assert np.array_equal(start, ostart)
    assert np.array_equal(cpts, ocpts)


def test_learned_radius_is_refused(pkg):
    with pkg.Calculateur(dim=3, cone_dim=2) as c:
        with pytest.raises(pkg.EmstError):
            c.set_radius_model(np.zeros(pkg.MLP_PARAMS, np.float32))
