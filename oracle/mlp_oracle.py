"""TEST INFRASTRUCTURE ONLY (the checker of the learned screen's network, never the product path).

numpy (FP64) restatement of the network that lightspeed_amd/csrc/hip/mlp.hip evaluates and trains:
ni inputs -> 3 hidden layers of 64 ReLU units -> one linear output.  Parameter vector layout:
W1 (64 x ni, row-major), b1 (64), W2 (64 x 64), b2, W3 (64 x 64), b3, w4 (64), b4 (1).

Loss of a batch: L = (1/nb) sum_s sw_s r_s^2 (times wu where r_s < 0), r_s = f(x_s) - y_s (an
asymmetric squared error: predicting below the target costs wu times more).
The gradient is pinned by central finite differences (tests/test_learned_screen.py).
"""
import numpy as np

H = 64


def offsets(ni):
    o = {}
    k = 0
    for name, size in (("W1", H * ni), ("b1", H), ("W2", H * H), ("b2", H), ("W3", H * H), ("b3", H),
                       ("w4", H), ("b4", 1)):
        o[name] = (k, k + size)
        k += size
    o["n"] = k
    return o


def unpack(p, ni):
    o = offsets(ni)
    g = {k: np.asarray(p[v[0]:v[1]], dtype=np.float64) for k, v in o.items() if k != "n"}
    g["W1"] = g["W1"].reshape(H, ni)
    g["W2"] = g["W2"].reshape(H, H)
    g["W3"] = g["W3"].reshape(H, H)
    return g


def forward(p, X, ni):
    """Outputs f (n,) and the activations (x, h1, h2, h3)."""
    w = unpack(p, ni)
    X = np.asarray(X, dtype=np.float64).reshape(-1, ni)
    h1 = np.maximum(X @ w["W1"].T + w["b1"], 0.0)
    h2 = np.maximum(h1 @ w["W2"].T + w["b2"], 0.0)
    h3 = np.maximum(h2 @ w["W3"].T + w["b3"], 0.0)
    f = h3 @ w["w4"] + w["b4"][0]
    return f, (X, h1, h2, h3)


def loss(p, X, y, ni, sw=None, wu=1.0):
    f, _ = forward(p, X, ni)
    r = f - np.asarray(y, dtype=np.float64)
    s = np.ones_like(r) if sw is None else np.asarray(sw, dtype=np.float64)
    return float(np.sum(s * np.where(r < 0, wu, 1.0) * r * r) / len(r))


def grad(p, X, y, ni, sw=None, wu=1.0):
    """(loss, dL/dp) by backpropagation."""
    w = unpack(p, ni)
    f, (x, h1, h2, h3) = forward(p, X, ni)
    y = np.asarray(y, dtype=np.float64)
    n = len(y)
    r = f - y
    s = np.ones_like(r) if sw is None else np.asarray(sw, dtype=np.float64)
    wgt = s * np.where(r < 0, wu, 1.0) / n
    L = float(np.sum(wgt * r * r))
    g = 2.0 * wgt * r  # dL/df
    d3 = np.outer(g, w["w4"]) * (h3 > 0)
    d2 = (d3 @ w["W3"]) * (h2 > 0)
    d1 = (d2 @ w["W2"]) * (h1 > 0)
    out = {
        "W1": d1.T @ x, "b1": d1.sum(0),
        "W2": d2.T @ h1, "b2": d2.sum(0),
        "W3": d3.T @ h2, "b3": d3.sum(0),
        "w4": h3.T @ g, "b4": np.array([g.sum()]),
    }
    o = offsets(ni)
    G = np.zeros(o["n"])
    for k, v in o.items():
        if k != "n":
            G[v[0]:v[1]] = out[k].reshape(-1)
    return L, G


def adam(p, g, m, v, lr, b1, b2, eps, t):
    """One Adam step (Kingma & Ba 2015, bias-corrected); returns new (p, m, v)."""
    m = b1 * m + (1 - b1) * g
    v = b2 * v + (1 - b2) * g * g
    p = p - lr * (m / (1 - b1 ** t)) / (np.sqrt(v / (1 - b2 ** t)) + eps)
    return p, m, v


def init(ni, seed):
    """He-normal weights, zero biases (the training driver's initialisation)."""
    rng = np.random.default_rng(seed)
    o = offsets(ni)
    p = np.zeros(o["n"])
    for name, fan in (("W1", ni), ("W2", H), ("W3", H), ("w4", H)):
        a, b = o[name]
        p[a:b] = rng.standard_normal(b - a) * np.sqrt(2.0 / fan)
    return p
