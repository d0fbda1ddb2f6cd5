"""The oracle's full-size checkers (no GPU): the batched fused d2, the
union-find component count and Kruskal under the float total order, each
pinned against an independent computation."""
import numpy as np
import pytest


def _pts(pkg, kind, n, dim, seed):
    import oracle
    rs, _, _, _ = oracle.rescale(pkg.synth(kind, n, dim, seed))
    return rs


@pytest.mark.parametrize("dim", [2, 3])
def test_d2_pairs_equals_single_d2(pkg, oracle_mod, dim):
    rs = _pts(pkg, 1, 500, dim, 3)
    rng = np.random.default_rng(dim)
    u = rng.integers(0, 500, 300).astype(np.int32)
    v = rng.integers(0, 500, 300).astype(np.int32)
    got = oracle_mod.d2_pairs(rs, u, v)
    want = np.array([oracle_mod.d2(rs[a], rs[b]) for a, b in zip(u, v)], np.float32)
    assert np.array_equal(got.view(np.uint32), want.view(np.uint32))


def test_forest_components(oracle_mod):
    # a path 0-1-2, a pair 3-4, a lone 5, and one edge closing a cycle
    u = np.array([0, 1, 3, 2], np.int32)
    v = np.array([1, 2, 4, 0], np.int32)
    assert oracle_mod.forest_components(6, u, v) == (3, 1)
    assert oracle_mod.forest_components(3, u[:2], v[:2]) == (1, 0)


@pytest.mark.parametrize("dim,kind", [(2, 0), (3, 1), (2, 2)])
def test_kruskal_complete_graph_equals_prim(pkg, oracle_mod, dim, kind):
    """With every pair as a candidate, Kruskal under (d2, min id, max id) is
    the complete-graph MST of Prim under the same order."""
    rs = _pts(pkg, kind, 600, dim, 10 + kind)
    iu, iv = np.triu_indices(len(rs), 1)
    ku, kv = oracle_mod.kruskal(rs, iv, iu)
    pu, pv = oracle_mod.prim_float(rs)
    assert np.array_equal(oracle_mod.canon(ku, kv), oracle_mod.canon(pu, pv))


@pytest.mark.parametrize("dim", [2, 3])
def test_kruskal_lattice_ties(oracle_mod, dim):
    """Exact ties everywhere: the order's id tie-break decides, as in Prim."""
    g = np.stack(np.meshgrid(*[np.arange(5)] * dim), -1).reshape(-1, dim).astype(np.float32)
    g = np.concatenate([g, g[:7]])          # coincident points too
    iu, iv = np.triu_indices(len(g), 1)
    ku, kv = oracle_mod.kruskal(g, iu, iv)
    pu, pv = oracle_mod.prim_float(g)
    assert np.array_equal(oracle_mod.canon(ku, kv), oracle_mod.canon(pu, pv))


def test_kruskal_rejects_a_replaced_edge(pkg, oracle_mod):
    """The full-size tree check: Kruskal over a tree plus near neighbours
    returns the tree when it is the MST, and not when one of its edges was
    swapped for a longer one."""
    rs = _pts(pkg, 0, 800, 2, 5)
    pu, pv = oracle_mod.prim_float(rs)
    iu, iv = np.triu_indices(len(rs), 1)
    d = oracle_mod.d2_pairs(rs, iu.astype(np.int32), iv.astype(np.int32))
    near = d < np.quantile(d, 0.02)
    cu = np.concatenate([pu, iu[near]])
    cv = np.concatenate([pv, iv[near]])
    ku, kv = oracle_mod.kruskal(rs, cu, cv)
    assert np.array_equal(oracle_mod.canon(ku, kv), oracle_mod.canon(pu, pv))
    # drop the tree's shortest edge and reconnect its two sides by a longer pair
    k = int(np.argmin(oracle_mod.d2_pairs(rs, pu, pv)))
    bu, bv = np.delete(pu, k), np.delete(pv, k)
    parent = list(range(len(rs)))

    def find(x):
        while parent[x] != x:
            parent[x] = parent[parent[x]]
            x = parent[x]
        return x

    for a, b in zip(bu.tolist(), bv.tolist()):
        parent[find(a)] = find(b)
    side = np.array([find(i) for i in range(len(rs))])
    cross = np.nonzero((side[iu] != side[iv]) & near)[0]
    cross = cross[np.argsort(d[cross])]
    j = cross[1]                            # the second-shortest crossing pair
    tu, tv = np.append(bu, iu[j]).astype(np.int32), np.append(bv, iv[j]).astype(np.int32)
    assert oracle_mod.forest_components(len(rs), tu, tv) == (1, 0)
    ku, kv = oracle_mod.kruskal(rs, np.concatenate([tu, iu[near]]), np.concatenate([tv, iv[near]]))
    assert not np.array_equal(oracle_mod.canon(ku, kv), oracle_mod.canon(tu, tv))
