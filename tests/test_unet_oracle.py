"""Pins the FP64 numpy network reference (oracle/unet.py) by central finite
differences of the loss with respect to every layer's weights, biases and the
input, and by closed forms of pool/up and their adjoints."""
import numpy as np


def test_pool_up_adjoints(oracle):
    import unet as U
    rng = np.random.default_rng(0)
    a = rng.standard_normal((4, 6, 8, 3))
    b = rng.standard_normal((2, 3, 4, 3))
    # <pool(a), b> == <a, pool^T(b)>, <up(b), a> == <b, up^T(a)>
    assert abs(np.sum(U.pool(a) * b) - np.sum(a * U.pool_backward(b))) < 1e-12
    assert abs(np.sum(U.up(b) * a) - np.sum(b * U.up_backward(a))) < 1e-12
    assert np.allclose(U.pool(U.up(b)), b)


def test_conv_matches_direct_sum(oracle):
    import unet as U
    rng = np.random.default_rng(1)
    x = rng.standard_normal((3, 4, 5, 2))
    w = rng.standard_normal((27, 3, 2))
    b = rng.standard_normal(3)
    y = U.conv(x, w, b)
    v = (1, 2, 0)
    s = b.copy()
    for t, (dz, dy, dx) in enumerate(U.OFFS):
        z, yy, xx = v[0] + dz, v[1] + dy, v[2] + dx
        if 0 <= z < 3 and 0 <= yy < 4 and 0 <= xx < 5:
            s += w[t] @ x[z, yy, xx]
    assert np.allclose(y[v], s)


def test_gradients_finite_differences(oracle):
    import unet as U
    rng = np.random.default_rng(2)
    P = U.init_params(3)
    for k in P:
        if k.endswith(".b"):
            P[k] = rng.standard_normal(P[k].shape) * 0.1
    x = rng.standard_normal((4, 4, 8, 8))
    t = rng.standard_normal((4, 4, 8, 3))
    out, c = U.forward(P, x)
    L, dout = U.loss(out, t)
    G = U.backward(P, c, dout)
    h = 1e-6
    for k in P:
        flat = P[k].reshape(-1)
        for idx in rng.choice(flat.size, 3, replace=False):
            old = flat[idx]
            flat[idx] = old + h
            Lp, _ = U.loss(U.forward(P, x)[0], t)
            flat[idx] = old - h
            Lm, _ = U.loss(U.forward(P, x)[0], t)
            flat[idx] = old
            fd = (Lp - Lm) / (2 * h)
            g = G[k].reshape(-1)[idx]
            assert abs(fd - g) < 1e-6 * max(1.0, abs(g)) + 1e-8, (k, idx, fd, g)


def test_adam_reduces_loss(oracle):
    import unet as U
    rng = np.random.default_rng(4)
    P = U.init_params(5)
    x = rng.standard_normal((4, 4, 4, 8))
    t = rng.standard_normal((4, 4, 4, 3)) * 0.1
    st = {}
    L0 = U.loss(U.forward(P, x)[0], t)[0]
    for _ in range(30):
        out, c = U.forward(P, x)
        _, d = U.loss(out, t)
        U.adam(P, U.backward(P, c, d), st, lr=3e-3)
    assert U.loss(U.forward(P, x)[0], t)[0] < 0.5 * L0


def test_transfers_closed_forms(oracle):
    """Restriction keeps a constant in interior cells, prolongation reproduces
    a linear field at interior nodes (trilinear interpolation), and the two are
    adjoint up to the factor 8."""
    import unet as U
    rng = np.random.default_rng(6)
    nx, ny, nz = 14, 10, 6
    g = U.cell_grid(nx, ny, nz)
    assert g == (8, 8, 4)
    NX, NY, NZ = nx + 1, ny + 1, nz + 1
    v = np.full((NZ, NY, NX, 3), 2.5)
    s = U.restrict(v, g)
    assert np.allclose(s[1, 1:4, 1:6], 2.5)
    # a linear function of the cell centres (2I+1, 2J+1, 2K+1) ...
    K, J, I = np.meshgrid(np.arange(g[2]), np.arange(g[1]), np.arange(g[0]), indexing="ij")
    lin = lambda x, y, z: 0.3 * x - 0.2 * y + 0.7 * z + 1.0
    o = lin(2 * I + 1, 2 * J + 1, 2 * K + 1)[..., None]
    p = U.prolong(o, (NX, NY, NZ))[..., 0]
    # ... is interpolated exactly at nodes with centres on both sides
    z, y, x = np.meshgrid(np.arange(NZ), np.arange(NY), np.arange(NX), indexing="ij")
    inner = (x >= 1) & (x <= 2 * g[0] - 1) & (y >= 1) & (y <= 2 * g[1] - 1) & (z >= 1) & (z <= 2 * g[2] - 1)
    inner &= (x < NX) & (y < NY) & (z < NZ)
    assert np.allclose(p[inner], lin(x, y, z)[inner])
    # adjointness
    v = rng.standard_normal((NZ, NY, NX, 3))
    o = rng.standard_normal((g[2], g[1], g[0], 3))
    assert abs(8 * np.sum(U.restrict(v, g) * o) - np.sum(v * U.prolong(o, (NX, NY, NZ)))) < 1e-10


def test_bf16_round_known_values(oracle):
    import unet as U
    a = np.array([1.0, 1.0 + 2 ** -8, 1.0 + 3 * 2 ** -8, -2.0 - 2 ** -7, 3.0e-3, 0.0])
    r = U.bf16_round(a)
    assert r[0] == 1.0 and r[1] == 1.0            # tie to even (down)
    assert r[2] == 1.0 + 2 ** -6                  # tie to even (up)
    assert r[3] == -2.0                           # tie between -2 and -2 - 2^-6: even mantissa
    assert abs(r[4] - 3.0e-3) <= 3.0e-3 * 2 ** -8 and r[5] == 0.0
