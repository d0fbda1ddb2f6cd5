"""Batched clouds (emst_set_clouds): every cloud of a batch gets exactly the
forest, weight and per-round completion it gets when run alone, and small
clouds match the oracle's exact tree."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _single(pkg, pts):
    with pkg.Calculateur(dim=pts.shape[1]) as c:
        c.initialize(pts)
        c.run()
        u, v, d = c.getLinks()
        w, m = c.evaluate()
    return u, v, d, w


def _clouds(pkg, dim, seed):
    rng = np.random.default_rng(seed)
    sizes = [0, 1, 2, 3, 7, 64, 65, 500, 3000, 12000, 40000]
    sizes += list(rng.integers(2, 2000, 20))
    clouds = []
    for i, n in enumerate(sizes):
        kind = [0, 1, 2, 3][i % 4]
        pts = pkg.synth(kind, int(n), dim, 1000 + i)
        if i % 5 == 4 and n > 4:     # coincident points inside a cloud
            pts[: n // 4] = pts[n // 4: 2 * (n // 4)]
        clouds.append(pts)
    return clouds


@pytest.mark.parametrize("dim", [2, 3])
def test_batch_equals_single_clouds(pkg, dim):
    clouds = _clouds(pkg, dim, dim)
    print("batch", dim, len(clouds), sum(len(c) for c in clouds), flush=True)
    with pkg.Calculateur(dim=dim) as c:
        c.initialize_clouds(clouds)
        c.run()
        u, v, d = c.getLinks()
        w, e, r = c.cloud_results()
        off = c.offsets
    # global ids never cross clouds; edges come sorted by (u, v), so each
    # cloud's edges are one contiguous run
    cu = np.searchsorted(off, u, side="right") - 1
    cv = np.searchsorted(off, v, side="right") - 1
    assert np.array_equal(cu, cv)
    for k, pts in enumerate(clouds):
        sel = cu == k
        su, sv, sd, sw = _single(pkg, pts)
        assert e[k] == len(su) == sel.sum()
        assert np.array_equal(u[sel] - off[k], su) and np.array_equal(v[sel] - off[k], sv)
        assert np.array_equal(d[sel].view(np.uint32), sd.view(np.uint32))
        assert abs(w[k] - sw) <= 1e-12 * max(sw, 1.0), (k, w[k], sw)


@pytest.mark.parametrize("dim", [2, 3])
def test_batch_small_clouds_match_oracle(pkg, oracle_mod, dim):
    rng = np.random.default_rng(40 + dim)
    clouds = [pkg.synth(int(rng.integers(0, 4)), int(rng.integers(2, 400)), dim, 77 + i) for i in range(300)]
    print("small batch", dim, len(clouds), flush=True)
    with pkg.Calculateur(dim=dim) as c:
        c.initialize_clouds(clouds)
        c.run()
        u, v, _ = c.getLinks()
        w, e, r = c.cloud_results()
        off = c.offsets
    cu = np.searchsorted(off, u, side="right") - 1
    for k in range(0, len(clouds), 7):
        ref = oracle_mod.emst(clouds[k], mode=0)
        sel = cu == k
        got = np.stack([u[sel] - off[k], v[sel] - off[k]], 1).astype(np.int64)
        assert np.array_equal(got, oracle_mod.canon(ref["u"], ref["v"])), k
        assert r[k] == ref["rounds"], (k, r[k], ref["rounds"])


def test_batch_of_singletons_and_pairs(pkg):
    clouds = [np.zeros((1, 3), np.float32)] * 5 + [np.array([[0, 0, 0], [1, 1, 1]], np.float32)] * 3
    with pkg.Calculateur(dim=3) as c:
        c.initialize_clouds(clouds)
        c.run()
        u, v, _ = c.getLinks()
        w, e, r = c.cloud_results()
    assert list(e) == [0] * 5 + [1] * 3
    assert np.allclose(w[5:], np.sqrt(3.0), rtol=1e-7) and np.all(w[:5] == 0)
    assert list(u) == [5, 7, 9] and list(v) == [6, 8, 10]
