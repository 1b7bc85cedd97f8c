"""MMASolver(n, m) with m constraints on the device (icarat_hip_mmasolver_*).

- The reference's own known-answer test (test_opt_solver.cpp:17-73: n = 3,
  m = 2, after 10 updates f < 8.78 and |g| < 1e-6), with the design after
  every update against the compiled reference's trajectory
  (tests/golden/mma_toy.npz).
- A 400-variable, three-constraint run with per-variable bounds and
  non-default parameters against the compiled reference
  (tests/golden/mma_m3.npz): designs after every update.
- m = 1, 4 and 8 (the largest supported) with 5,000 variables (multi-block
  sums) against the oracle's restatement: designs and asymptotes.
"""
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _golden(name):
    return np.load(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", name))


def _toy(s):
    # test_opt_solver.cpp:17-73: minimize |s|^2 subject to two spheres
    f = s[0] ** 2 + s[1] ** 2 + s[2] ** 2
    df = 2.0 * s
    g = np.array([(s[0] - 5) ** 2 + (s[1] - 2) ** 2 + (s[2] - 1) ** 2 - 9.0,
                  (s[0] - 3) ** 2 + (s[1] - 4) ** 2 + (s[2] - 3) ** 2 - 9.0])
    # dgdx[i m + j]
    dg = np.array([2 * (s[0] - 5), 2 * (s[0] - 3), 2 * (s[1] - 2), 2 * (s[1] - 4),
                   2 * (s[2] - 1), 2 * (s[2] - 3)])
    return f, df, g, dg


def test_mmasolver_reference_kat(gpu):
    d = _golden("mma_toy.npz")
    mma = gpu.MMASolver(3, 2)
    mma.set_params(*d["params"])  # SetMoveLimit(0.5), SetOscillation(0.1), SetAsymptotes(0.5, 0.7, 1.2)
    s = d["x0"].copy()
    for k in range(10):
        f, df, g, dg = _toy(s)
        outer = mma.update(s, df, g, dg, np.zeros(3), np.full(3, 5.0))
        assert outer > 0
        err = np.max(np.abs(s - d["traj"][k]))
        print(f"update {k}: f {f:.10f} g {g} |s - s_ref| {err:.2e}")
        assert err < 1e-9, k
    # the KAT's assertions on the last evaluated objective and constraints
    assert f < 8.78
    assert np.all(np.abs(g) < 1e-6)


def test_mmasolver_m3_matches_reference_fixture(gpu):
    d = _golden("mma_m3.npz")
    n, m = d["x0"].size, d["gx"].shape[1]
    mma = gpu.MMASolver(n, m)
    mma.set_params(*d["params"])
    x = d["x0"].copy()
    for k in range(d["x"].shape[0]):
        mma.update(x, d["dfdx"][k], d["gx"][k], d["dgdx"][k], d["xmin"], d["xmax"])
        err = np.max(np.abs(x - d["x"][k]))
        print(f"update {k}: max |x - x_ref| {err:.2e}")
        assert err < 1e-9, k


@pytest.mark.parametrize("m", [1, 4, 8])
def test_mmasolver_matches_oracle(gpu, oracle, m):
    rng = np.random.default_rng(40 + m)
    n = 5000
    x0 = rng.uniform(0.2, 0.8, n)
    xmin, xmax = np.full(n, 1e-3), np.ones(n)
    w = rng.uniform(0.5, 1.5, (m, n)) / n
    cap = rng.uniform(0.35, 0.5, m)
    gm = gpu.MMASolver(n, m)
    om = oracle.MMA(n, m)
    xg, xo = x0.copy(), x0.copy()
    for it in range(6):
        df = -rng.uniform(0.1, 3.0, n) * (1.0 + xo)
        gx = w @ xo - cap
        dg = np.ascontiguousarray(w.T).reshape(-1)
        gm.update(xg, df, gx, dg, xmin, xmax)
        om.update(xo, df, gx, dg, xmin, xmax)
        lg, ug = gm.asymptotes()
        lo, uo = om.state()
        err = np.max(np.abs(xg - xo))
        print(f"m {m} update {it}: max |x - x_oracle| {err:.2e}, active {np.sum(w @ xo - cap > -1e-6)}")
        assert err < 1e-9, it
        assert np.max(np.abs(lg - lo)) < 1e-12 and np.max(np.abs(ug - uo)) < 1e-12
        xg[:] = xo  # the next update from the same design
