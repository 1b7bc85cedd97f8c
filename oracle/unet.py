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


# ---------------------------------------------------------------------------
# The preconditioner around the network: node lattice <-> cell grid.
# Cell (I, J, K) covers elements 2I..2I+1 (and likewise in y, z); its centre is
# node (2I+1, 2J+1, 2K+1).  Restriction is full weighting over the 3x3x3 nodes
# around the centre (1-D weights 1/2, 1, 1/2, normalised by 8); prolongation
# is its transpose without the 1/8 (trilinear interpolation from the centres).

def cell_grid(nx, ny, nz):
    """Cell grid (X, Y, Z) for nx x ny x nz elements: ceil-half rounded up to 4."""
    r4 = lambda v: (v + 3) // 4 * 4
    return r4((nx + 1) // 2), r4((ny + 1) // 2), r4((nz + 1) // 2)


def _weights_1d(nn, ncell):
    """W[I, i] = weight of node i in cell I along one axis (nn nodes)."""
    W = np.zeros((ncell, nn))
    for I in range(ncell):
        c = 2 * I + 1
        for d, w in ((-1, 0.5), (0, 1.0), (1, 0.5)):
            if 0 <= c + d < nn:
                W[I, c + d] = w
    return W


def restrict(v, grid):
    """v: (NZ, NY, NX, C) nodal -> (Z, Y, X, C) per cell (full weighting / 8)."""
    X, Y, Z = grid
    NZ, NY, NX = v.shape[:3]
    Wx, Wy, Wz = _weights_1d(NX, X), _weights_1d(NY, Y), _weights_1d(NZ, Z)
    return np.einsum("Kk,Jj,Ii,kjic->KJIc", Wz, Wy, Wx, v) / 8.0


def prolong(o, nodes):
    """o: (Z, Y, X, C) per cell -> (NZ, NY, NX, C) nodal (transpose of restrict x 8)."""
    NX, NY, NZ = nodes
    Z, Y, X = o.shape[:3]
    Wx, Wy, Wz = _weights_1d(NX, X), _weights_1d(NY, Y), _weights_1d(NZ, Z)
    return np.einsum("Kk,Jj,Ii,KJIc->kjic", Wz, Wy, Wx, o)


def stiffness_feature(young, grid):
    """young: (nz, ny, nx) per element -> (Z, Y, X) clamp((log10 mean + 9) / 9, 0, 1)
    over the elements of each cell."""
    X, Y, Z = grid
    nz, ny, nx = young.shape
    f = np.zeros((Z, Y, X))
    for K in range(Z):
        for J in range(Y):
            for I in range(X):
                blk = young[2 * K:2 * K + 2, 2 * J:2 * J + 2, 2 * I:2 * I + 2]
                if blk.size and blk.mean() > 0:
                    f[K, J, I] = (np.log10(blk.mean()) + 9.0) / 9.0
    return np.clip(f, 0.0, 1.0)


def bf16_round(a):
    """Round to bfloat16 (nearest, ties to even) and back to float64."""
    f = np.ascontiguousarray(a, dtype=np.float32)
    u = f.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) >> 16 << 16
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


def precondition(P, r, dinv, free, young):
    """z = D^-1 r + s P(net(feature, R(D^-1 r) / s)) on the node lattice.
    r, dinv: (NZ, NY, NX, 3); free: (NZ, NY, NX, 3) bool; young: (nz, ny, nx).
    s = RMS of R(D^-1 r) over all cell components; network inputs rounded to
    bfloat16 as the device kernels store them.  Returns z and the correction."""
    NZ, NY, NX = r.shape[:3]
    grid = cell_grid(NX - 1, NY - 1, NZ - 1)
    s3 = restrict(dinv * r, grid)
    sig = np.sqrt(np.mean(s3 * s3))
    x = np.zeros(s3.shape[:3] + (8,))
    x[..., 0] = stiffness_feature(young, grid)
    x[..., 1:4] = s3 / sig if sig > 0 else 0.0
    out, _ = forward(P, bf16_round(x))
    corr = sig * prolong(out, (NX, NY, NZ))
    z = np.where(free, dinv * r + corr, 0.0)
    return z, np.where(free, corr, 0.0)
