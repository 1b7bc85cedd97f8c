"""The learned preconditioner's network on the GPU (BF16 MFMA convolutions,
FP32 accumulation, BF16 activations) against the FP64 numpy network
(oracle/unet.py, pinned by finite differences).  Tolerances, normwise:
  forward output ............... 1e-2 of ||out|| (BF16 inputs, weights, activations;
                                 measured 2.8e-3)
  parameter gradient per tensor  7e-2 of ||g||, cosine >= 0.998 (measured at most
                                 4.8e-2 and 0.9989: the backward pass rounds the
                                 stored activations and each layer's output
                                 gradient to BF16, about 2e-3 per rounding)
  first Adam step .............. >= 95 % of the gradient mass moves with the oracle's sign
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

GRID = (8, 8, 4)  # X, Y, Z


def _oracle_params(U, flat):
    P, off = {}, 0
    for name, ci, co, taps in U.LAYERS:
        n = taps * co * ci
        P[name + ".w"] = flat[off:off + n].astype(np.float64).reshape(taps, co, ci)
        off += n
        P[name + ".b"] = flat[off:off + co].astype(np.float64)
        off += co
    assert off == flat.size
    return P


def _flat(U, G):
    return np.concatenate([np.concatenate([G[n + ".w"].reshape(-1), G[n + ".b"]])
                           for n, *_ in U.LAYERS])


def _setup(gpu, seed):
    import unet as U
    rng = np.random.default_rng(seed)
    X, Y, Z = GRID
    net = gpu.UNet(X, Y, Z, seed=seed)
    flat = net.get_params()
    assert flat.size == gpu.unet_param_count()
    # non-zero biases so that the bias paths are exercised
    P = _oracle_params(U, flat)
    for n, *_ in U.LAYERS:
        P[n + ".b"] = rng.standard_normal(P[n + ".b"].shape) * 0.05
    flat = _flat(U, P).astype(np.float32)
    net.set_params(flat)
    assert np.array_equal(net.get_params(), flat)
    P = _oracle_params(U, flat)
    x4 = rng.standard_normal((Z, Y, X, 4)).astype(np.float32)
    x8 = np.zeros((Z, Y, X, 8))
    x8[..., :4] = U.bf16_round(x4)
    return U, net, P, x4, x8, rng


def _nrel(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def test_forward_matches_oracle(gpu, oracle):
    U, net, P, x4, x8, _ = _setup(gpu, 11)
    out = net.forward(x4)
    ref, _ = U.forward(P, x8)
    e = _nrel(out, ref)
    print(f"forward normwise error {e:.2e}")
    assert e < 1e-2


def test_gradients_match_oracle(gpu, oracle):
    U, net, P, x4, x8, rng = _setup(gpu, 12)
    t = rng.standard_normal(x8.shape[:3] + (3,)).astype(np.float32)
    loss, g = net.grads(x4, t)
    out, c = U.forward(P, x8)
    L, dout = U.loss(out, t.astype(np.float64))
    G = U.backward(P, c, dout)
    assert abs(loss - L) < 2e-2 * L
    Gg = _oracle_params(U, g)
    for n, *_ in U.LAYERS:
        for k in (n + ".w", n + ".b"):
            e = _nrel(Gg[k], G[k])
            cos = float(np.sum(Gg[k] * G[k]) / (np.linalg.norm(Gg[k]) * np.linalg.norm(G[k])))
            print(f"{k}: normwise {e:.2e} cosine {cos:.5f}")
            assert e < 7e-2 and cos > 0.998, k


def test_adam_step_matches_oracle(gpu, oracle):
    U, net, P, x4, x8, rng = _setup(gpu, 13)
    t = rng.standard_normal(x8.shape[:3] + (3,)).astype(np.float32)
    before = net.get_params().astype(np.float64)
    net.train_step(x4, t, lr=1e-3)
    d_gpu = net.get_params().astype(np.float64) - before
    out, c = U.forward(P, x8)
    _, dout = U.loss(out, t.astype(np.float64))
    G = U.backward(P, c, dout)
    P1 = {k: v.copy() for k, v in P.items()}
    U.adam(P1, G, {}, lr=1e-3)
    d_orc = _flat(U, P1) - _flat(U, P)
    g = np.abs(_flat(U, G))
    agree = np.sum(g[np.sign(d_gpu) == np.sign(d_orc)]) / np.sum(g)
    print(f"first Adam step: gradient mass with the oracle's sign {agree:.4f}")
    assert agree > 0.95
    assert np.max(np.abs(d_gpu)) <= 1e-3 * 1.001


def test_training_reduces_loss(gpu, oracle):
    U, net, P, x4, x8, rng = _setup(gpu, 14)
    # a target the network can represent: the oracle's output under other weights
    Q = U.init_params(99)
    t, _ = U.forward(Q, x8)
    t = t.astype(np.float32)
    l0 = net.train_step(x4, t, lr=3e-3)
    for _ in range(60):
        l = net.train_step(x4, t, lr=3e-3)
    print(f"loss {l0:.4f} -> {l:.4f}")
    assert l < 0.5 * l0
