"""Run-time machinery around the kernels: rounds replayed from a captured
graph give exactly the rounds issued one by one from the host, on every
replay; the autotuner changes only the work, never the tree."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _edges(c):
    u, v, d = c.getLinks()
    return np.stack([u, v], 1), d


@pytest.mark.parametrize("dim,kind,n", [(3, 1, 60_000), (2, 2, 80_000), (2, 0, 1000)])
def test_graph_rounds_equal_host_rounds(pkg, dim, kind, n):
    pts = pkg.synth(kind, n, dim, 21)
    if kind == 2:
        pts[: n // 10] = pts[n // 10: 2 * (n // 10)]      # coincident groups (grouped rounds)
    with pkg.Calculateur(dim=dim, trace=True) as c:      # traced: one round per host call
        c.initialize(pts)
        c.run()
        ref, refd = _edges(c)
        ref_rounds = len(c.trace())
    with pkg.Calculateur(dim=dim) as c:                  # untraced: captured rounds
        for rep in range(3):
            c.initialize(pts)
            c.run()
            e, d = _edges(c)
            assert np.array_equal(e, ref) and np.array_equal(d.view(np.uint32), refd.view(np.uint32)), rep
            w, m = c.evaluate()
            assert m == n - 1
            # eager work on the same stream between replays
            _ = c.candidates()


def test_round_cap_with_captured_rounds(pkg, oracle_mod):
    pts = pkg.synth(0, 5000, 2, 31)
    ref = oracle_mod.emst(pts, mode=0, max_rounds=3)
    with pkg.Calculateur(dim=2, max_rounds=3) as c:
        c.initialize(pts)
        c.run()
        u, v, _ = c.getLinks()
    assert np.array_equal(np.stack([u, v], 1).astype(np.int64), oracle_mod.canon(ref["u"], ref["v"]))


@pytest.mark.parametrize("dim,kind", [(3, 1), (2, 0)])
def test_autotune_keeps_the_tree(pkg, dim, kind):
    pts = pkg.synth(kind, 100_000, dim, 8)
    with pkg.Calculateur(dim=dim) as c:
        c.initialize(pts)
        c.run()
        ref, _ = _edges(c)
        b, sc, ms = c.autotune()
        # 1.0: the 3D component search's seed, when that search wins
        assert b in (16, 32, 64) and sc in (1.0, 2.0, 3.0, 4.0) and ms > 0
        c.run()
        e, _ = _edges(c)
        assert np.array_equal(e, ref)
        c.initialize(pts)          # the tuned leaf size stays for later point sets
        c.run()
        e, _ = _edges(c)
        assert np.array_equal(e, ref)
        opt = c.options()
        assert opt["bucket_size"] == b and opt["seed_scale"] == sc and opt["search_split"] in (0, 1, 2, 3)
        assert opt["round_search"] in (0, 1)
    # a new context built from the tuned options runs the same way
    keys = ("bucket_size", "seed_scale", "search_split", "round_search")
    with pkg.Calculateur(dim=dim, **{k: opt[k] for k in keys}) as c2:
        c2.initialize(pts)
        c2.run()
        e, _ = _edges(c2)
        assert np.array_equal(e, ref)
