"""numpy reference of the seed-radius MLP (8 -> 64 -> 64 -> 1, ReLU).

`bf16=True` rounds operands and activations to bfloat16 exactly where the GPU
kernels do (normalised inputs, weights, hidden activations, backward
deltas), with float32 accumulation; `bf16=False` is the same network in
float64, used to pin the backward formulas by finite differences."""
import numpy as np

F, H = 8, 64
oW1 = 0
oB1 = oW1 + F * H
oW2 = oB1 + H
oB2 = oW2 + H * H
oW3 = oB2 + H
oB3 = oW3 + H
oMean = oB3 + 1
oIstd = oMean + F
oOff = oIstd + F
NPARAM = oOff + 1
NTRAIN = oMean
XCLAMP = 4.0      # normalised features are clamped to [-4, 4]


def bf16(a):
    """Round float32 to bfloat16 (nearest even), returned as float32."""
    a = np.ascontiguousarray(a, dtype=np.float32)
    u = a.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) >> 16
    return (u.astype(np.uint32) << 16).view(np.float32)


def unpack(P):
    P = np.asarray(P)
    return (P[oW1:oB1].reshape(F, H), P[oB1:oW2], P[oW2:oB2].reshape(H, H), P[oB2:oW3],
            P[oW3:oB3], P[oB3], P[oMean:oIstd], P[oIstd:oOff], P[oOff])


def forward(P, X, bf16_on=True):
    W1, b1, W2, b2, w3, b3, mean, istd, _ = unpack(P)
    r = bf16 if bf16_on else (lambda a: a)
    dt = np.float32 if bf16_on else np.float64
    X = np.asarray(X, dt)
    xn = r(np.clip((X - mean.astype(dt)) * istd.astype(dt), -XCLAMP, XCLAMP).astype(dt))
    z1 = xn @ r(W1.astype(dt)) + b1.astype(dt)
    h1 = r(np.maximum(z1, 0).astype(dt))
    z2 = h1 @ r(W2.astype(dt)) + b2.astype(dt)
    h2 = r(np.maximum(z2, 0).astype(dt))
    y = h2 @ r(w3.astype(dt)) + dt(b3)
    return y, (xn, h1, h2)


def grad(P, X, T, bf16_on=True):
    """Gradient of sum (y - t)^2 over the trainable parameters, and that sum."""
    W1, b1, W2, b2, w3, b3, mean, istd, _ = unpack(P)
    r = bf16 if bf16_on else (lambda a: a)
    dt = np.float32 if bf16_on else np.float64
    y, (xn, h1, h2) = forward(P, X, bf16_on)
    e = y - np.asarray(T, dt)
    dy = 2 * e
    G = np.zeros(NPARAM, np.float64)
    G[oB3] = dy.sum()
    G[oW3:oB3] = h2.T @ dy
    dz2 = r((dy[:, None] * r(w3.astype(dt))[None, :] * (h2 > 0)).astype(dt))
    G[oB2:oW3] = dz2.sum(0)
    G[oW2:oB2] = (h1.T @ dz2).ravel()
    dh1 = dz2 @ r(W2.astype(dt)).T
    dz1 = r((dh1 * (h1 > 0)).astype(dt))
    G[oB1:oW2] = dz1.sum(0)
    G[oW1:oB1] = (xn.T @ dz1).ravel()
    return G, float((e.astype(np.float64) ** 2).sum())


def random_model(seed, X=None):
    rng = np.random.default_rng(seed)
    P = np.zeros(NPARAM, np.float32)
    P[oW1:oB1] = rng.uniform(-1, 1, F * H) * np.sqrt(6 / F)
    P[oB1:oW2] = rng.uniform(-0.1, 0.1, H)
    P[oW2:oB2] = rng.uniform(-1, 1, H * H) * np.sqrt(6 / H)
    P[oB2:oW3] = rng.uniform(-0.1, 0.1, H)
    P[oW3:oB3] = rng.uniform(-1, 1, H) * np.sqrt(3 / H)
    P[oB3] = 0.3
    if X is None:
        P[oMean:oIstd] = 0
        P[oIstd:oOff] = 1
    else:
        P[oMean:oIstd] = X.mean(0)
        P[oIstd:oOff] = 1 / (X.std(0) + 1e-6)
    return P
