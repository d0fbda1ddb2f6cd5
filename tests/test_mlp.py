"""Seed-radius MLP: the numpy reference pinned by finite differences (CPU),
then the GPU forward/backward MFMA kernels checked against it, training
reduces the loss, and learned seeds change the work but not the tree."""
import numpy as np
import pytest

import mlp_ref as R


def test_reference_backward_matches_finite_differences():
    """float64 form: analytic gradient of sum (y - t)^2 equals central
    differences for every parameter block (a dropped term or a transposed
    operand fails here)."""
    rng = np.random.default_rng(1)
    X = rng.normal(size=(40, R.F))
    T = rng.normal(size=40)
    P = R.random_model(2, X).astype(np.float64)
    G, _ = R.grad(P, X, T, bf16_on=False)
    idx = np.concatenate([rng.choice(np.arange(a, b), 12, replace=False)
                          for a, b in ((R.oW1, R.oB1), (R.oB1, R.oW2), (R.oW2, R.oB2),
                                       (R.oB2, R.oW3), (R.oW3, R.oB3))] + [[R.oB3]])
    h = 1e-6
    for i in idx:
        Pp, Pm = P.copy(), P.copy()
        Pp[i] += h
        Pm[i] -= h
        lp = ((R.forward(Pp, X, False)[0] - T) ** 2).sum()
        lm = ((R.forward(Pm, X, False)[0] - T) ** 2).sum()
        num = (lp - lm) / (2 * h)
        assert abs(num - G[i]) <= 1e-5 * max(1.0, abs(num)), (i, num, G[i])


def test_bf16_rounding_is_nearest_even():
    x = np.array([1.0, 1.00390625, 1.005859375, 1.0078125, -2.5, 3.14159], np.float32)
    # 1 + 2^-8 is a tie between 1 and 1 + 2^-7 -> even (1.0); 1 + 1.5*2^-8 rounds up
    assert R.bf16(x).tolist() == [1.0, 1.0, 1.0078125, 1.0078125, -2.5, 3.140625]


def _data(n, seed):
    rng = np.random.default_rng(seed)
    X = rng.normal(size=(n, R.F)).astype(np.float32)
    X[:, 0] += 1.5
    T = (0.7 * np.tanh(X[:, 0]) - 0.3 * X[:, 5] + 0.1 * rng.normal(size=n)).astype(np.float32)
    return X, T


@pytest.mark.gpu
def test_gpu_forward_matches_reference(pkg):
    import ctypes
    X, _ = _data(1000, 3)
    P = R.random_model(4, X)
    out = np.empty(len(X), np.float32)
    L = pkg.lib()
    rc = L.emst_mlp_eval(0, P.ctypes.data_as(ctypes.c_void_p), X.ctypes.data_as(ctypes.c_void_p), len(X),
                         out.ctypes.data_as(ctypes.c_void_p))
    assert rc == 0
    ref, _ = R.forward(P, X, True)
    # bf16 activations: the MFMA accumulation order may move an activation by
    # one bf16 step (2^-8 relative); outputs agree to ~1e-2 relative of scale
    assert np.max(np.abs(out - ref)) <= 2e-2 * max(1.0, np.max(np.abs(ref)))
    assert np.corrcoef(out, ref)[0, 1] > 0.9999


@pytest.mark.gpu
def test_gpu_gradient_matches_reference(pkg):
    import ctypes
    X, T = _data(512, 5)
    P = R.random_model(6, X)
    G = np.empty(R.NPARAM, np.float32)
    loss = ctypes.c_double(0)
    L = pkg.lib()
    rc = L.emst_mlp_grad(0, P.ctypes.data_as(ctypes.c_void_p), X.ctypes.data_as(ctypes.c_void_p),
                         T.ctypes.data_as(ctypes.c_void_p), len(X), G.ctypes.data_as(ctypes.c_void_p),
                         ctypes.byref(loss))
    assert rc == 0
    Gr, lr = R.grad(P, X, T, True)
    assert abs(loss.value - lr) <= 2e-2 * lr
    for a, b, name in ((R.oW1, R.oB1, "W1"), (R.oB1, R.oW2, "b1"), (R.oW2, R.oB2, "W2"),
                       (R.oB2, R.oW3, "b2"), (R.oW3, R.oB3, "w3"), (R.oB3, R.oB3 + 1, "b3")):
        g, gr = G[a:b].astype(np.float64), Gr[a:b]
        rel = np.linalg.norm(g - gr) / max(np.linalg.norm(gr), 1e-12)
        assert rel < 2e-2, (name, rel)
    assert np.all(G[R.NTRAIN:] == 0)


@pytest.mark.gpu
def test_gpu_training_reduces_loss(pkg):
    import ctypes
    X, T = _data(1 << 14, 7)
    model = np.empty(R.NPARAM, np.float32)
    losses = np.empty(6, np.float32)
    L = pkg.lib()
    rc = L.emst_mlp_train(0, X.ctypes.data_as(ctypes.c_void_p), T.ctypes.data_as(ctypes.c_void_p), len(X), 6,
                          256, ctypes.c_float(3e-3), 11, model.ctypes.data_as(ctypes.c_void_p),
                          losses.ctypes.data_as(ctypes.c_void_p))
    assert rc == 0
    var = float(np.var(T))
    assert losses[-1] < 0.3 * var, (losses, var)
    assert losses[-1] < losses[0]
    pred, _ = R.forward(model, X, True)
    assert np.mean((pred - T) ** 2) < 0.3 * var


@pytest.mark.gpu
@pytest.mark.parametrize("dim,kind", [(3, 1), (2, 2)])
@pytest.mark.parametrize("offset", [-6.0, 0.0, 6.0])
def test_any_seed_gives_the_same_tree(pkg, dim, kind, offset):
    """Seeds from a model (random weights, offsets from tiny to huge radii)
    change how many passes the cone search makes, never its candidates or the
    tree."""
    pts = pkg.synth(kind, 200_000, dim, 9)
    with pkg.Calculateur(dim=dim) as c:
        c.initialize(pts)
        c.run()
        u0, v0, _ = c.getLinks()
        cand0, d0 = c.candidates()
        X, _ = c.radius_samples()
        model = R.random_model(13, X)
        model[R.oOff] = offset
        c.set_radius_model(model)
        c.initialize(pts)
        c.run()
        u1, v1, _ = c.getLinks()
        cand1, d1 = c.candidates()
    assert np.array_equal(u0, u1) and np.array_equal(v0, v1)
    assert np.array_equal(cand0, cand1) and np.array_equal(d0.view(np.uint32), d1.view(np.uint32))
