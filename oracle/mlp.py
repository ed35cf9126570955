"""ORACLE -- test infrastructure only.

FP64 numpy restatement of the selector network in libgraphhip.so
(graphkokkos_amd/csrc/selector.hip): 16 -> 64 -> 64 -> 4, ReLU, inputs
standardised by the stored mean/std, loss = mean over valid outputs of the
squared error.  Used to check the GPU forward pass and gradients; pinned by
central finite differences (tests/test_oracle_mlp.py).
"""
import numpy as np

N_IN, N_H, N_OUT = 16, 64, 4
SIZES = [("W1", (N_IN, N_H)), ("b1", (N_H,)), ("W2", (N_H, N_H)), ("b2", (N_H,)), ("W3", (N_H, N_OUT)),
         ("b3", (N_OUT,)), ("mean", (N_IN,)), ("std", (N_IN,))]


def unpack(p):
    out, o = {}, 0
    for name, shape in SIZES:
        n = int(np.prod(shape))
        out[name] = np.asarray(p[o:o + n], np.float64).reshape(shape)
        o += n
    return out


def n_params():
    return sum(int(np.prod(s)) for _, s in SIZES)


def n_trainable():
    return n_params() - 2 * N_IN


def forward(p, X):
    w = unpack(p)
    x = (np.asarray(X, np.float64) - w["mean"]) / w["std"]
    a1 = x @ w["W1"] + w["b1"]
    h1 = np.maximum(a1, 0)
    a2 = h1 @ w["W2"] + w["b2"]
    h2 = np.maximum(a2, 0)
    return h2 @ w["W3"] + w["b3"], (x, a1, h1, a2, h2)


def loss_grad(p, X, T, V):
    """Loss and gradient w.r.t. W1, b1, W2, b2, W3, b3 (flattened in order)."""
    w = unpack(p)
    y, (x, a1, h1, a2, h2) = forward(p, X)
    V = np.asarray(V, bool)
    cnt = max(int(V.sum()), 1)
    d = np.where(V, y - np.asarray(T, np.float64), 0.0)
    loss = float((d * d).sum() / cnt)
    dy = 2 * d / cnt
    gW3 = h2.T @ dy
    gb3 = dy.sum(0)
    dh2 = (dy @ w["W3"].T) * (a2 > 0)
    gW2 = h1.T @ dh2
    gb2 = dh2.sum(0)
    dh1 = (dh2 @ w["W2"].T) * (a1 > 0)
    gW1 = x.T @ dh1
    gb1 = dh1.sum(0)
    return loss, np.concatenate([gW1.ravel(), gb1, gW2.ravel(), gb2, gW3.ravel(), gb3])


def random_params(rng, scale=0.3):
    p = rng.normal(0, scale, n_params())
    p[-2 * N_IN:-N_IN] = rng.normal(0, 1, N_IN)        # mean
    p[-N_IN:] = rng.uniform(0.5, 2.0, N_IN)            # std
    return p
