"""The oracle pinned against the reference's own golden data, closed forms and
the complete-graph MST (CPU only)."""
import json
import os
from fractions import Fraction

import numpy as np
import pytest

from conftest import GOLDEN, g6, golden_path, read_points

FIX = json.load(open(os.path.join(GOLDEN, "fixtures.json")))


@pytest.mark.parametrize("name", sorted(FIX["traces"]))
@pytest.mark.parametrize("mode", [0, 1])
def test_traces_match_reference_logs(oracle_mod, name, mode):
    """Per-round forest size and float32 length of both oracle forms (exact
    cone minima, and the reference's ring search with R = 1) equal the logged
    reference runs."""
    pts = read_points(golden_path(name))
    if mode == 0 and len(pts) > 4000:
        pytest.skip("exhaustive form is quadratic; ring form covers this size")
    r = oracle_mod.emst(pts, mode=mode, R=1)
    rows = FIX["traces"][name]
    assert r["cum_edges"] == [x["size"] for x in rows]
    # per-round length: weight of the edges added up to round k
    lengths = []
    for k in range(1, r["rounds"] + 1):
        sel = r["round"] <= k
        lengths.append(g6(np.float32(oracle_mod.weight(pts, r["u"][sel], r["v"][sel]))))
    assert lengths == [x["length"] for x in rows]


def test_outypu_links_points_labels(oracle_mod):
    pts = read_points(golden_path("3000randomCity.2d"))
    r = oracle_mod.emst(pts, mode=1, R=1)
    assert oracle_mod.canon(r["u"], r["v"]).tolist() == FIX["outypu"]["edges"]
    assert sorted(set((r["labels"] * 4).tolist())) == FIX["outypu"]["disjoint_set_values"]
    printed = open(golden_path("outypu.points")).read().split()[6:]
    printed = [x for i, x in enumerate(printed) if i % 3 != 2]
    assert ["%.6g" % x for x in r["rescaled"].astype(np.float64).ravel()] == printed


def _f32(fr):
    """Round a Fraction to the nearest float32 (ties to even), exactly."""
    if fr == 0:
        return np.float32(0)
    sign = -1 if fr < 0 else 1
    a = abs(fr)
    e = a.numerator.bit_length() - a.denominator.bit_length()
    while Fraction(2) ** e > a:
        e -= 1
    while Fraction(2) ** (e + 1) <= a:
        e += 1
    ulp = Fraction(2) ** (e - 23)
    q, rem = divmod(a, ulp)
    q = int(q)
    if rem * 2 > ulp or (rem * 2 == ulp and q % 2 == 1):
        q += 1
    return np.float32(sign * float(q * ulp))


def test_d2_is_fused_and_ordered(oracle_mod):
    """d2 = fl(fl(dx*dx) + dy*dy) [+ dz*dz] with each '+ a*a' fused
    (Node.h:241-250 under nvcc --use_fast_math, lib.pro:138), checked with
    exact rational arithmetic on random and on constructed operands."""
    a = np.array([0.0, 0.0, 0.0], np.float32)
    b = np.array([1.0, 4097.0, 0.0], np.float32)
    # fused: round(4097^2 + 1) = 16785410; unfused would give 16785408
    assert oracle_mod.d2(a[:2], b[:2]) == np.float32(16785410.0)
    rng = np.random.default_rng(0)
    for dim in (2, 3):
        for _ in range(2000):
            p = rng.random(dim, dtype=np.float32) * np.float32(rng.choice([1, 100, 1e4]))
            q = rng.random(dim, dtype=np.float32) * np.float32(rng.choice([1, 100, 1e4]))
            d = (p - q).astype(np.float32)
            f = _f32(Fraction(float(d[0])) ** 2)
            f = _f32(Fraction(float(d[1])) ** 2 + Fraction(float(f)))
            if dim == 3:
                f = _f32(Fraction(float(d[2])) ** 2 + Fraction(float(f)))
            assert oracle_mod.d2(p, q) == f


def test_rescale_closed_form(oracle_mod):
    """unitX = (ceil(_wnd) + 1) / wg (SolutionEMST.cu:305-342) on a box whose
    _wnd is known exactly."""
    pts = np.array([[0, 0], [100, 25], [50, 10], [3, 7]], np.float32)
    out, unit, nmin, nmax = oracle_mod.rescale(pts)
    wnd = np.sqrt(4 * 100 / 25) * 2          # = 8
    assert unit == (np.ceil(wnd) + 1) / 100.0
    assert np.array_equal(out, (pts - pts.min(0)) * np.float32(unit))
    assert nmin[0] == 0 and nmax[0] == np.float32(100 * np.float32(unit))


def test_grid_geometry_closed_form(oracle_mod):
    """ViewGridQuadMD recurrences and findCell on a hand-computed case:
    ext = 11 -> pc = 5; R = 2: PC = (5+1)/2 = 3, ext_base = 4 + (6+1)/2 = 7,
    PCD = 1, ext_dual = 2 + (4+1)/2 = 4; FDual(0) = (0 + 1)*2 - 6 + 5 = 1,
    cells 4 wide."""
    g = oracle_mod.grid(2, 2, np.zeros(3), np.array([10.0, 10.0, 0.0]))
    assert list(g.ext)[:2] == [11, 11] and list(g.pc)[:2] == [5, 5]
    assert list(g.PC)[:2] == [3, 3] and list(g.ext_base)[:2] == [7, 7]
    assert list(g.ext_dual)[:2] == [4, 4] and g.p1[0] == 1.0
    pts = np.array([[0.0, 0.0], [2.99, 3.0], [3.0, 6.99], [10.0, 7.0]], np.float32)
    cell, _, _, outside = oracle_mod.cell_build(g, pts)
    # floor((p - 1 + 2) / 4)
    assert cell[:, :2].tolist() == [[0, 0], [0, 1], [1, 1], [2, 2]]
    assert outside == 0


@pytest.mark.parametrize("dim,n,seed", [(2, 1500, 1), (3, 1500, 2), (2, 600, 3), (3, 600, 4)])
def test_boruvka_equals_prim(oracle_mod, dim, n, seed):
    """Both oracle forms give the unique complete-graph MST under the total
    order (Prim), on uniform and on clustered clouds."""
    rng = np.random.default_rng(seed)
    if seed % 2:
        pts = rng.random((n, dim), dtype=np.float32)
    else:
        c = rng.random((5, dim))
        pts = (c[rng.integers(0, 5, n)] + 0.01 * rng.standard_normal((n, dim))).astype(np.float32)
    rs, _, _, _ = oracle_mod.rescale(pts)
    want = oracle_mod.canon(*oracle_mod.prim_float(rs))
    for mode in (0, 1):
        r = oracle_mod.boruvka(rs, mode=mode, R=1)
        assert np.array_equal(oracle_mod.canon(r["u"], r["v"]), want)


def test_exact_cones_are_true_minima(oracle_mod):
    """Brute force per cone on a tiny cloud with a pure-Python classifier
    built from the cone definition (dominant axis and signs)."""
    rng = np.random.default_rng(9)
    pts = rng.random((40, 3), dtype=np.float32)
    cid, cd2, _ = oracle_mod.octants(pts, mode=0)
    for p in range(len(pts)):
        for k in range(24):
            f, s = divmod(k, 4)
            ax, neg = divmod(f, 2)
            others = [i for i in range(3) if i != ax]
            best = None
            for q in range(len(pts)):
                if q == p:
                    continue
                d = (pts[q] - pts[p]).astype(np.float32)
                a = np.abs(d)
                if a[ax] < a.max() or (d[ax] < 0) != bool(neg):
                    continue
                ok = all((d[o] <= 0) if (s >> j) & 1 else (d[o] >= 0) for j, o in enumerate(others))
                if not ok:
                    continue
                v = oracle_mod.d2(pts[p], pts[q])
                if best is None or (v, q) < best:
                    best = (v, q)
            if best is None:
                assert cid[p, k] == -1
            else:
                assert (cd2[p, k], cid[p, k]) == best


def test_fp64_prim_weight(oracle_mod):
    rng = np.random.default_rng(3)
    pts = rng.random((800, 2), dtype=np.float32)
    r = oracle_mod.emst(pts, mode=0)
    _, _, w64 = oracle_mod.prim_fp64(pts)
    # float32 d2 per edge: relative error below 2^-23 for every length
    assert abs(r["weight"] - w64) <= 2 ** -22 * w64
