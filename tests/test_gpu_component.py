"""The per-round component search (round_search = 1; NodeSpiralSearchMD,
emst_coalition/include/NodeEMST.h:67-509, the reference's alternative to the
round-1 cone search) against the oracle and the cone search: the same edge
sets, per-round forest sizes and labels.  Each round's component minimum is
the component's shortest outgoing edge under (d2, min id, max id) in both
searches, so the forests must agree round by round."""
import numpy as np
import pytest

from conftest import golden_path, read_points

pytestmark = pytest.mark.gpu
WTOL = 1e-12


def run(pkg, pts, step=False, **kw):
    with pkg.Calculateur(dim=pts.shape[1], trace=True, **kw) as c:
        c.initialize(pts)
        if step:
            while c.activate():
                pass
        else:
            c.run()
        u, v, d = c.getLinks()
        w, m = c.evaluate()
        out = dict(e=np.stack([u, v], 1).astype(np.int64), d2=d, w=w, labels=c.labels(),
                   sizes=[r["size"] for r in c.trace()])
    return out


@pytest.mark.parametrize("dim,kind,n,seed", [(2, 0, 5000, 31), (3, 0, 5000, 32), (3, 1, 5000, 33),
                                             (2, 2, 5000, 34), (3, 2, 4000, 35), (2, 0, 9, 36)])
def test_matches_oracle_round_by_round(pkg, oracle_mod, dim, kind, n, seed):
    pts = pkg.synth(kind, n, dim, seed)
    print("case", dim, kind, n, seed, flush=True)
    g = run(pkg, pts, round_search=1)
    ref = oracle_mod.emst(pts, mode=0)
    assert np.array_equal(g["e"], oracle_mod.canon(ref["u"], ref["v"]))
    assert g["sizes"] == ref["cum_edges"]
    assert np.array_equal(g["labels"], ref["labels"])
    assert abs(g["w"] - ref["weight"]) <= WTOL * ref["weight"]


@pytest.mark.parametrize("name", ["3000randomCity.2d", "1000randomCity.2d", "2000randomCity.3d"])
def test_reference_data(pkg, oracle_mod, name):
    """The reference's own point files, stepped round by round."""
    pts = read_points(golden_path(name))
    g = run(pkg, pts, round_search=1, step=True)
    ref = oracle_mod.emst(pts, mode=0)
    assert np.array_equal(g["e"], oracle_mod.canon(ref["u"], ref["v"]))
    assert g["sizes"] == ref["cum_edges"]


@pytest.mark.parametrize("dim,kind,n,seed", [(3, 0, 200_000, 41), (2, 2, 200_000, 42), (3, 1, 150_000, 43)])
def test_equals_cone_search(pkg, dim, kind, n, seed):
    """Larger clouds: identical edges (and d2 bits) and per-round sizes as the cone search."""
    pts = pkg.synth(kind, n, dim, seed)
    print("case", dim, kind, n, seed, flush=True)
    a = run(pkg, pts, round_search=1)
    b = run(pkg, pts, round_search=0)
    assert np.array_equal(a["e"], b["e"])
    assert np.array_equal(a["d2"].view(np.uint32), b["d2"].view(np.uint32))
    assert a["sizes"] == b["sizes"]
    assert np.array_equal(a["labels"], b["labels"])


def test_coincident_points(pkg, oracle_mod):
    rng = np.random.default_rng(44)
    base = rng.random((1200, 3), dtype=np.float32)
    pts = rng.permutation(np.repeat(base, rng.integers(1, 6, len(base)), axis=0))
    g = run(pkg, pts, round_search=1)
    ref = oracle_mod.emst(pts, mode=0)
    assert np.array_equal(g["e"], oracle_mod.canon(ref["u"], ref["v"]))
    assert g["sizes"] == ref["cum_edges"]


def test_batch_equals_single_clouds(pkg):
    rng = np.random.default_rng(45)
    clouds = [pkg.synth(int(rng.integers(0, 3)), int(rng.integers(2, 4000)), 3, 500 + i) for i in range(7)]
    with pkg.Calculateur(dim=3, round_search=1) as c:
        c.initialize_clouds(clouds)
        c.run()
        u, v, _ = c.getLinks()
        off = c.offsets
    for i, pts in enumerate(clouds):
        s = run(pkg, pts, round_search=1)
        sel = (u >= off[i]) & (u < off[i + 1])
        eb = np.stack([u[sel] - off[i], v[sel] - off[i]], 1).astype(np.int64)
        assert np.array_equal(eb, s["e"])


def test_refusals(pkg):
    pts = pkg.synth(0, 100, 3, 46)
    with pkg.Calculateur(dim=3, round_search=1) as c:
        c.initialize(pts)
        with pytest.raises(pkg.EmstError):
            c.candidates()
        with pytest.raises(pkg.EmstError):
            c.radius_samples()          # no log without log_samples=1
    with pytest.raises(pkg.EmstError):
        pkg.Calculateur(dim=3, cone_dim=2, round_search=1)


@pytest.mark.parametrize("dim", [2, 3])
def test_large_coincident_groups_equal_cone_search(pkg, dim):
    """Runs of thousands of identical points (ties at d2 = 0 broken on ids)."""
    rng = np.random.default_rng(47 + dim)
    base = rng.random((3000, dim), dtype=np.float32)
    big = np.repeat(base[:3], [6000, 2500, 40], axis=0)
    pts = rng.permutation(np.concatenate([base, big, base[10:400]]))
    a = run(pkg, pts, round_search=1)
    b = run(pkg, pts, round_search=0)
    assert np.array_equal(a["e"], b["e"])
    assert a["sizes"] == b["sizes"]


def _learned_run(pkg, pts, model):
    with pkg.Calculateur(dim=pts.shape[1], trace=True, round_search=1) as c:
        c.set_radius_model(model)
        c.initialize(pts)
        c.run()
        u, v, d = c.getLinks()
        return np.stack([u, v], 1).astype(np.int64), [r["size"] for r in c.trace()], c.labels()


@pytest.mark.parametrize("dim,kind", [(3, 1), (2, 2)])
def test_learned_seed_logs_trains_and_keeps_the_forest(pkg, dim, kind):
    """Pairs logged from a plain run train a model on the GPU; any model and
    offset (too small, too large) only changes the work: the forests of every
    round equal the plain search's."""
    pts = pkg.synth(kind, 60_000, dim, 49 + dim)
    with pkg.Calculateur(dim=dim, round_search=1, log_samples=1) as c:
        c.initialize(pts)
        c.run()
        X, T = c.radius_samples()
    assert len(T) > 1000
    assert np.isfinite(X).all() and np.isfinite(T).all()
    # round and size features: round 1 (log 1 = 0) with single points (log 1 = 0)
    assert X[:, 0].min() == 0.0 and X[:, 0].max() > 0.0
    assert X[X[:, 0] == 0.0][:, 1].max() == 0.0
    model, losses = pkg.train_radius_model(X, T, epochs=3, batch=1024, seed=dim)
    assert losses[-1] < float(np.var(T))
    plain = run(pkg, pts, round_search=1)
    for off in (-3.0, 0.0, 3.0):
        m = model.copy()
        m[-1] = off
        e, sizes, labels = _learned_run(pkg, pts, m)
        assert np.array_equal(e, plain["e"])
        assert sizes == plain["sizes"]
        assert np.array_equal(labels, plain["labels"])


def test_learned_seed_batch_and_oracle(pkg, oracle_mod):
    """The learned seed on a batch of clouds, and against the oracle."""
    rng = np.random.default_rng(51)
    model = np.zeros(pkg.MLP_PARAMS, np.float32)
    model[-1] = 0.5                      # a constant prediction: e^0.5 leaf spacings
    clouds = [pkg.synth(int(rng.integers(0, 3)), int(rng.integers(2, 3000)), 3, 600 + i) for i in range(5)]
    with pkg.Calculateur(dim=3, round_search=1) as c:
        c.set_radius_model(model)
        c.initialize_clouds(clouds)
        c.run()
        u, v, _ = c.getLinks()
        off = c.offsets
    for i, pts in enumerate(clouds):
        ref = oracle_mod.emst(pts, mode=0)
        sel = (u >= off[i]) & (u < off[i + 1])
        eb = np.stack([u[sel] - off[i], v[sel] - off[i]], 1).astype(np.int64)
        assert np.array_equal(eb, oracle_mod.canon(ref["u"], ref["v"]))


@pytest.mark.parametrize("n", [0, 1, 2, 3])
def test_tiny_inputs(pkg, oracle_mod, n):
    pts = pkg.synth(0, max(n, 1), 3, 53)[:n]
    with pkg.Calculateur(dim=3, round_search=1) as c:
        c.initialize(pts)
        c.run()
        u, v, _ = c.getLinks()
        w, m = c.evaluate()
    assert m == max(n - 1, 0) and len(u) == m
    if n >= 2:
        ref = oracle_mod.emst(pts, mode=0)
        assert np.array_equal(np.stack([u, v], 1).astype(np.int64), oracle_mod.canon(ref["u"], ref["v"]))


@pytest.mark.parametrize("dim", [2, 3])
def test_all_points_coincident(pkg, dim):
    """One location repeated 5000 times: every point's nearest is the smallest
    other id, so round 1 builds the star around id 0."""
    pts = np.repeat(np.full((1, dim), 0.25, np.float32), 5000, axis=0)
    g = run(pkg, pts, round_search=1)
    assert g["sizes"] == [4999]
    e = g["e"]
    assert np.array_equal(e[:, 0], np.zeros(4999)) and np.array_equal(e[:, 1], np.arange(1, 5000))


def test_batch_with_tiny_clouds_and_round_cap(pkg):
    rng = np.random.default_rng(54)
    sizes = [1, 2, 0, 3, 500, 1, 2000]
    clouds = [pkg.synth(int(rng.integers(0, 3)), max(k, 1), 2, 700 + i)[:k] for i, k in enumerate(sizes)]
    out = {}
    for rs in (0, 1):
        with pkg.Calculateur(dim=2, round_search=rs, max_rounds=3) as c:
            c.initialize_clouds(clouds)
            c.run()
            u, v, d = c.getLinks()
            out[rs] = (np.stack([u, v], 1), d.view(np.uint32), c.cloud_results())
    assert np.array_equal(out[0][0], out[1][0]) and np.array_equal(out[0][1], out[1][1])
    for a, b in zip(out[0][2], out[1][2]):
        assert np.array_equal(a, b)


def test_context_reuse_across_point_sets(pkg):
    """One context, several point sets of different sizes and dimensions'
    distributions in a row (buffers regrown, the captured rounds re-captured):
    each run equals a fresh context's."""
    sets = [pkg.synth(1, 30_000, 3, 55), pkg.synth(0, 5_000, 3, 56), pkg.synth(2, 80_000, 3, 57),
            pkg.synth(0, 5_000, 3, 56)]
    with pkg.Calculateur(dim=3, round_search=1) as c:
        for pts in sets:
            c.initialize(pts)
            c.run()
            u, v, _ = c.getLinks()
            fresh = run(pkg, pts, round_search=1)
            assert np.array_equal(np.stack([u, v], 1).astype(np.int64), fresh["e"])
            c.run()                     # a second run on the same points
            u2, v2, _ = c.getLinks()
            assert np.array_equal(u2, u) and np.array_equal(v2, v)


def test_many_random_clouds_against_oracle(pkg, oracle_mod):
    """60 seeded clouds of every generator kind, 2D and 3D, 2-2500 points:
    edges, per-round sizes and labels equal the oracle's."""
    rng = np.random.default_rng(58)
    for i in range(60):
        dim = 2 + (i % 2)
        kind = int(rng.integers(0, 3))
        n = int(rng.integers(2, 2500))
        pts = pkg.synth(kind, n, dim, 900 + i)
        g = run(pkg, pts, round_search=1)
        ref = oracle_mod.emst(pts, mode=0)
        assert np.array_equal(g["e"], oracle_mod.canon(ref["u"], ref["v"])), (i, dim, kind, n)
        assert g["sizes"] == ref["cum_edges"], (i, dim, kind, n)
        assert np.array_equal(g["labels"], ref["labels"]), (i, dim, kind, n)


@pytest.mark.parametrize("dim", [2, 3])
def test_neighbour_lists_short_and_used_up(pkg, oracle_mod, dim):
    """The round-1 lists hold each point's K nearest other points (K = 6 in
    2D, 10 in 3D).  Clouds of 2 to K + 6 points leave them short (nothing
    outside once the listed points are all inside), exactly full, or used up;
    500 tight clusters of K + 4 points use every list up inside its own
    cluster by round 2, so the searches start from the lists' lower bounds.
    Edges, per-round sizes and labels equal the oracle's."""
    K = 10 if dim == 3 else 6
    cases = [pkg.synth(int(k % 3), k, dim, 1200 + k) for k in range(2, K + 7)]
    rng = np.random.default_rng(59 + dim)
    centres = rng.random((500, dim)).astype(np.float32)
    jitter = (rng.random((500, K + 4, dim)) * 1e-4).astype(np.float32)
    cases.append((centres[:, None, :] + jitter).reshape(-1, dim))
    for pts in cases:
        g = run(pkg, pts, round_search=1)
        ref = oracle_mod.emst(pts, mode=0)
        assert np.array_equal(g["e"], oracle_mod.canon(ref["u"], ref["v"])), len(pts)
        assert g["sizes"] == ref["cum_edges"], len(pts)
        assert np.array_equal(g["labels"], ref["labels"]), len(pts)
