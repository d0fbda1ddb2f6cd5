"""options.one_launch: the whole run (build, neighbour lists, rounds) as one
captured graph launch gives exactly the default run's forest, for single
clouds, batches, repeated runs on new points of the same shape, a cloud whose
leaves outnumber the graph's slots (redone eagerly), and NaN input (refused)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _run(pkg, pts, **kw):
    with pkg.Calculateur(dim=pts.shape[1], round_search=1, **kw) as c:
        c.initialize(pts)
        c.run()
        u, v, d = c.getLinks()
        w, m = c.evaluate()
    return u, v, d, w


def _same(a, b):
    assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1])
    assert np.array_equal(a[2].view(np.uint32), b[2].view(np.uint32))
    assert a[3] == b[3]


@pytest.mark.parametrize("kind,n,dim", [(1, 200_000, 3), (0, 100_000, 2), (2, 50_000, 2), (3, 3000, 3)])
def test_one_launch_equals_default(pkg, kind, n, dim):
    pts = pkg.synth(kind, n, dim, 60 + kind)
    print("case one_launch", kind, n, dim, flush=True)
    _same(_run(pkg, pts, one_launch=1), _run(pkg, pts))


def test_one_launch_round_ends(pkg):
    """The device clock stamps inside the launch: one end time per round run,
    increasing, after the build and within the run's wall time."""
    pts = pkg.synth(1, 200_000, 3, 61)
    print("case one_launch round ends", flush=True)
    with pkg.Calculateur(dim=3, round_search=1, one_launch=1) as c:
        c.initialize(pts)
        c.run()
        ends = c.round_ends()
        t = c.times()
        _, _, last = c.cloud_results()
    assert len(ends) == last[0] > 0
    assert np.all(np.diff(ends) > 0)
    assert 0 < t[0] < ends[0] <= t[3] * 1.05 + 0.05
    assert ends[-1] <= t[3] * 1.05 + 0.05


def test_one_launch_replays_on_new_points(pkg, oracle_mod):
    """The captured graph is replayed for new points of the same shape; each
    replay builds that cloud's tree."""
    print("case one_launch replay", flush=True)
    with pkg.Calculateur(dim=3, round_search=1, one_launch=1) as c:
        for seed in (1, 2, 3):
            pts = pkg.synth(1, 3000, 3, seed)
            c.initialize(pts)
            c.run()
            u, v, _ = c.getLinks()
            ref = oracle_mod.emst(pts, mode=0)
            assert np.array_equal(np.stack([u, v], 1).astype(np.int64), oracle_mod.canon(ref["u"], ref["v"]))


def test_one_launch_batch_equals_single_clouds(pkg):
    rng = np.random.default_rng(9)
    clouds = [pkg.synth(int(rng.integers(0, 4)), int(rng.integers(1, 5000)), 3, 300 + i) for i in range(60)]
    print("case one_launch batch", len(clouds), flush=True)
    with pkg.Calculateur(dim=3, round_search=1, one_launch=1) as c:
        c.initialize_clouds(clouds)
        c.run()
        u, v, d = c.getLinks()
        w, e, _ = c.cloud_results()
        off = c.offsets
    cu = np.searchsorted(off, u, side="right") - 1
    for k, pts in enumerate(clouds):
        sel = cu == k
        su, sv, sd, sw = _run(pkg, pts)
        assert e[k] == len(su) == sel.sum()
        assert np.array_equal(u[sel] - off[k], su) and np.array_equal(v[sel] - off[k], sv)
        assert np.array_equal(d[sel].view(np.uint32), sd.view(np.uint32))


def test_one_launch_more_leaves_than_slots(pkg):
    """Points halving their distance to the origin each time: the prefix tree
    peels one point per level, so the leaves (at most 8 points) outnumber the
    graph's slots; the run is redone eagerly with the same result."""
    x = np.float32(2.0) ** -np.arange(40, dtype=np.float32)
    pts = np.zeros((len(x) + 20, 3), np.float32)
    pts[: len(x), 0] = x
    pts[len(x):, 1] = np.linspace(0, 1, 20, dtype=np.float32)
    print("case one_launch overflow", flush=True)
    _same(_run(pkg, pts, one_launch=1, bucket_size=8), _run(pkg, pts, bucket_size=8))


def test_one_launch_refuses_nan(pkg):
    pts = pkg.synth(0, 1000, 3, 5)
    pts[17, 1] = np.nan
    with pkg.Calculateur(dim=3, round_search=1, one_launch=1) as c:
        c.initialize(pts)
        with pytest.raises(pkg.EmstError) as e:
            c.run()
        assert e.value.code == 4
        # and the context still works on good points afterwards
        good = pkg.synth(0, 1000, 3, 6)
        c.initialize(good)
        c.run()
        assert c.evaluate()[1] == 999
