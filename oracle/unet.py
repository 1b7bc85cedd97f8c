"""FP64 numpy reference of the learned preconditioner's network (forward and
backward), used to check the GPU conv kernels and their gradients.

TEST INFRASTRUCTURE ONLY (tests/ and the checker legs of smoke/bench).  The
network has no counterpart in the reference code base; this module is pinned
by finite differences in tests/test_unet_oracle.py.

Network (3 levels, channels 16/32/64, 3x3x3 convolutions, zero padding,
channels-last activations X[z, y, x, c]):
    h0 = relu(conv(x, 8 -> 16))                       level 0
    h1 = relu(conv(pool(h0), 16 -> 32))               level 1
    h2 = relu(conv(pool(h1), 32 -> 64))               level 2
    g1 = relu(conv(cat(up(h2), h1), 96 -> 32))        level 1
    g0 = relu(conv(cat(up(g1), h0), 48 -> 16))        level 0
    out = conv1x1(g0, 16 -> 3)                        level 0, linear
pool = 2x2x2 average, up = nearest-neighbour x2.  Weights W[tap, co, ci] with
tap = 9*(dz+1) + 3*(dy+1) + (dx+1), biases b[co].
"""
import numpy as np

LAYERS = [  # (name, cin, cout, taps)
    ("c0", 8, 16, 27),
    ("c1", 16, 32, 27),
    ("c2", 32, 64, 27),
    ("c3", 96, 32, 27),
    ("c4", 48, 16, 27),
    ("c5", 16, 3, 1),
]

OFFS = [(dz, dy, dx) for dz in (-1, 0, 1) for dy in (-1, 0, 1) for dx in (-1, 0, 1)]


def init_params(seed=0, dtype=np.float64):
    rng = np.random.default_rng(seed)
    P = {}
    for name, ci, co, taps in LAYERS:
        std = np.sqrt(2.0 / (ci * taps))
        P[name + ".w"] = (rng.standard_normal((taps, co, ci)) * std).astype(dtype)
        P[name + ".b"] = np.zeros(co, dtype=dtype)
    return P


def _shift(x, off):
    """y[v] = x[v + off] with zero outside."""
    dz, dy, dx = off
    Z, Y, X = x.shape[:3]
    y = np.zeros_like(x)
    zs, ze = max(0, -dz), min(Z, Z - dz)
    ys, ye = max(0, -dy), min(Y, Y - dy)
    xs, xe = max(0, -dx), min(X, X - dx)
    y[zs:ze, ys:ye, xs:xe] = x[zs + dz:ze + dz, ys + dy:ye + dy, xs + dx:xe + dx]
    return y


def conv(x, w, b):
    taps = w.shape[0]
    offs = OFFS if taps == 27 else [(0, 0, 0)]
    out = np.zeros(x.shape[:3] + (w.shape[1],))
    for t, off in enumerate(offs):
        out += _shift(x, off) @ w[t].T
    return out + b


def conv_backward(x, w, dy):
    taps = w.shape[0]
    offs = OFFS if taps == 27 else [(0, 0, 0)]
    dx = np.zeros_like(x)
    dw = np.zeros_like(w)
    for t, off in enumerate(offs):
        xs = _shift(x, off)
        dw[t] = dy.reshape(-1, dy.shape[-1]).T @ xs.reshape(-1, xs.shape[-1])
        g = dy @ w[t]
        dx += _shift(g, (-off[0], -off[1], -off[2]))
    db = dy.reshape(-1, dy.shape[-1]).sum(axis=0)
    return dx, dw, db


def pool(x):
    Z, Y, X, C = x.shape
    return x.reshape(Z // 2, 2, Y // 2, 2, X // 2, 2, C).mean(axis=(1, 3, 5))


def pool_backward(dy):
    return np.repeat(np.repeat(np.repeat(dy, 2, 0), 2, 1), 2, 2) / 8.0


def up(x):
    return np.repeat(np.repeat(np.repeat(x, 2, 0), 2, 1), 2, 2)


def up_backward(dy):
    Z, Y, X, C = dy.shape
    return dy.reshape(Z // 2, 2, Y // 2, 2, X // 2, 2, C).sum(axis=(1, 3, 5))


def forward(P, x):
    """x: (Z, Y, X, 8) with Z, Y, X divisible by 4.  Returns out and the cache."""
    c = {"x": x}
    c["a0"] = conv(x, P["c0.w"], P["c0.b"]); c["h0"] = np.maximum(c["a0"], 0)
    c["p0"] = pool(c["h0"])
    c["a1"] = conv(c["p0"], P["c1.w"], P["c1.b"]); c["h1"] = np.maximum(c["a1"], 0)
    c["p1"] = pool(c["h1"])
    c["a2"] = conv(c["p1"], P["c2.w"], P["c2.b"]); c["h2"] = np.maximum(c["a2"], 0)
    c["k1"] = np.concatenate([up(c["h2"]), c["h1"]], axis=-1)
    c["a3"] = conv(c["k1"], P["c3.w"], P["c3.b"]); c["g1"] = np.maximum(c["a3"], 0)
    c["k0"] = np.concatenate([up(c["g1"]), c["h0"]], axis=-1)
    c["a4"] = conv(c["k0"], P["c4.w"], P["c4.b"]); c["g0"] = np.maximum(c["a4"], 0)
    out = conv(c["g0"], P["c5.w"], P["c5.b"])
    return out, c


def loss(out, target):
    """Relative squared error 0.5 * ||out - t||^2 / ||t||^2 and its gradient."""
    n2 = float(np.sum(target * target))
    d = out - target
    return 0.5 * float(np.sum(d * d)) / n2, d / n2


def backward(P, c, dout):
    G = {}
    dg0, G["c5.w"], G["c5.b"] = conv_backward(c["g0"], P["c5.w"], dout)
    da4 = dg0 * (c["a4"] > 0)
    dk0, G["c4.w"], G["c4.b"] = conv_backward(c["k0"], P["c4.w"], da4)
    dh0 = dk0[..., 32:]
    dg1 = up_backward(dk0[..., :32])
    da3 = dg1 * (c["a3"] > 0)
    dk1, G["c3.w"], G["c3.b"] = conv_backward(c["k1"], P["c3.w"], da3)
    dh1 = dk1[..., 64:]
    dh2 = up_backward(dk1[..., :64])
    da2 = dh2 * (c["a2"] > 0)
    dp1, G["c2.w"], G["c2.b"] = conv_backward(c["p1"], P["c2.w"], da2)
    dh1 = dh1 + pool_backward(dp1)
    da1 = dh1 * (c["a1"] > 0)
    dp0, G["c1.w"], G["c1.b"] = conv_backward(c["p0"], P["c1.w"], da1)
    dh0 = dh0 + pool_backward(dp0)
    da0 = dh0 * (c["a0"] > 0)
    _, G["c0.w"], G["c0.b"] = conv_backward(c["x"], P["c0.w"], da0)
    return G


def adam(P, G, state, lr=1e-3, b1=0.9, b2=0.999, eps=1e-8):
    """One Adam step (Kingma & Ba 2015) on every parameter, in place."""
    state["t"] = state.get("t", 0) + 1
    t = state["t"]
    for k in P:
        m = state.setdefault("m." + k, np.zeros_like(P[k]))
        v = state.setdefault("v." + k, np.zeros_like(P[k]))
        m *= b1
        m += (1 - b1) * G[k]
        v *= b2
        v += (1 - b2) * G[k] * G[k]
        mh = m / (1 - b1 ** t)
        vh = v / (1 - b2 ** t)
        P[k] -= lr * mh / (np.sqrt(vh) + eps)
