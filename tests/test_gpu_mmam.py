"""MMASolver(n, m) and GCMMASolver(n, m) with m constraints on the device
(icarat_hip_mmasolver_*, icarat_hip_gcmmasolver_*).

- The reference's own known-answer test (test_opt_solver.cpp:17-73: n = 3,
  m = 2, after 10 updates f < 8.78 and |g| < 1e-6), with the design after
  every update against the compiled reference's trajectory
  (tests/golden/mma_toy.npz).
- A 400-variable, three-constraint run with per-variable bounds and
  non-default parameters against the compiled reference
  (tests/golden/mma_m3.npz): designs after every update.
- m = 1, 4 and 8 (the largest supported) with 5,000 variables (multi-block
  sums) against the oracle's restatement: designs and asymptotes.
- GCMMA: the reference's known-answer test (test_opt_solver.cpp:75-161, n = 3,
  m = 2) with every returned design and inner count against the compiled
  reference (tests/golden/gcmma_toy.npz); a 300-variable three-constraint
  loop against the compiled reference (tests/golden/gcmma_m3.npz); m = 2 and
  5 with 3,000 variables and forced inner updates against the oracle (designs
  and the raa / approximation state).
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


# ---- GCMMASolver(n, m) -------------------------------------------------------

def test_gcmmasolver_reference_kat(gpu):
    """The reference's GCMMA known-answer test (test_opt_solver.cpp:75-161) as
    written -- the first ConCheck of each outer iteration gets the objective
    of the zero vector -- with every returned design and the inner counts
    against the compiled reference (tests/golden/gcmma_toy.npz)."""
    d = _golden("gcmma_toy.npz")
    g = gpu.GCMMASolver(3, 2)
    g.set_params(*d["params"])  # SetMoveLimit(0.5), SetOscillation(0.1), SetAsymptotes(0.5, 0.7, 1.2)
    lo, hi = np.zeros(3), np.full(3, 5.0)
    s = d["x0"].copy()
    calls, inners = [], []
    for _ in range(10):
        f, df, gx, dg = _toy(s)
        fnew = 0.0  # objectivenew of the zero snew
        snew = g.outer(s, f, df, gx, dg, lo, hi)
        calls.append(snew)
        gnew = _toy(snew)[2]
        ok = g.concheck(fnew, gnew)
        k = 0
        while not ok and k < 15:
            snew = g.inner(snew, fnew, gnew, s, f, df, gx, dg, lo, hi)
            calls.append(snew)
            fnew, _, gnew, _ = _toy(snew)
            ok = g.concheck(fnew, gnew)
            k += 1
        inners.append(k)
        s = snew
    print("inner iterations", inners, "reference", d["inners"].tolist())
    assert inners == d["inners"].tolist()
    err = max(np.max(np.abs(a - b)) for a, b in zip(calls, d["xmma"]))
    print(f"max |xmma - xmma_ref| over {len(calls)} updates: {err:.2e}")
    assert err < 1e-9
    # the KAT's assertions on the last evaluated objective and constraints
    assert f < 8.78
    assert np.all(np.abs(gx) < 1e-6)


def test_gcmmasolver_m3_matches_reference_fixture(gpu):
    d = _golden("gcmma_m3.npz")
    cw, w, cap, xmin, xmax = d["cw"], d["w"], d["cap"], d["xmin"], d["xmax"]
    n, m = cw.size, cap.size

    def fg(v):
        return float(np.sum(cw / v)), w @ v - cap

    g = gpu.GCMMASolver(n, m)
    g.set_params(*d["params"])
    dg = np.ascontiguousarray(w.T).reshape(-1)
    x = d["x0"].copy()
    calls, inners = [], []
    for _ in range(d["inners"].size):
        f, gx = fg(x)
        df = -cw / x ** 2
        xm = g.outer(x, f, df, gx, dg, xmin, xmax)
        calls.append(xm)
        fnew, gnew = fg(xm)
        ok = g.concheck(fnew, gnew)
        k = 0
        while not ok and k < 5:
            xm = g.inner(xm, fnew, gnew, x, f, df, gx, dg, xmin, xmax)
            calls.append(xm)
            fnew, gnew = fg(xm)
            ok = g.concheck(fnew, gnew)
            k += 1
        inners.append(k)
        x = xm
    assert inners == d["inners"].tolist()
    err = max(np.max(np.abs(a - b)) for a, b in zip(calls, d["xmma"]))
    print(f"inner iterations {inners}; max |xmma - xmma_ref| over {len(calls)} updates: {err:.2e}")
    assert err < 1e-9


@pytest.mark.parametrize("m", [2, 5])
def test_gcmmasolver_matches_oracle(gpu, oracle, m):
    """Outer updates and forced inner updates (raa0 and raa raised) against the
    oracle's restatement: designs and the state {raa0, raa, f0app, fapp}."""
    rng = np.random.default_rng(60 + m)
    n = 3000
    cw = rng.uniform(0.5, 2.0, n) / n
    w = rng.uniform(0.5, 1.5, (m, n)) / n
    cap = rng.uniform(0.3, 0.45, m)
    xmin, xmax = np.full(n, 1e-3), np.ones(n)
    dg = np.ascontiguousarray(w.T).reshape(-1)
    gg, go = gpu.GCMMASolver(n, m), oracle.GCMMA(n, m)

    def same_state():
        # raa0, raa to 1e-9; f0app and fapp are sums of n terms that scale with
        # raa0 (objective) and raa_j (constraints) minus the same sums at xval
        # (the constants r0, r): each inner update above the approximation
        # raises raa up to tenfold, so another summation order leaves rounding
        # of the order 1e-16 raa n.  The error is measured in units of 1e-13
        # raa n and must stay below 1
        a, b = gg.state(), go.state()
        raa = b[:1 + m]
        d_raa = np.max(np.abs(a[:1 + m] - raa) / np.abs(raa))
        d_app = np.max(np.abs(a[1 + m:] - b[1 + m:]) / (1e-13 * n * np.maximum(1.0, raa)))
        return d_raa < 1e-9 and d_app < 1.0, (d_raa, d_app)

    x = rng.uniform(0.3, 0.7, n)
    for it in range(4):
        f, gx = float(np.sum(cw / x)), w @ x - cap
        df = -cw / x ** 2
        xg, xo = gg.outer(x, f, df, gx, dg, xmin, xmax), go.outer(x, f, df, gx, dg, xmin, xmax)
        assert np.max(np.abs(xg - xo)) < 1e-9, it
        ok, d = same_state()
        assert ok, d
        # two inner updates with values above the approximations
        for k in range(2):
            st = go.state()
            fnew, gnew = st[1 + m] + 0.1 * abs(st[1 + m]), st[2 + m:] + 0.01
            xg = gg.inner(xo, fnew, gnew, x, f, df, gx, dg, xmin, xmax)
            xo = go.inner(xo, fnew, gnew, x, f, df, gx, dg, xmin, xmax)
            err = np.max(np.abs(xg - xo))
            ok, d = same_state()
            print(f"m {m} outer {it} inner {k}: max |x - x_oracle| {err:.2e}, raa {d[0]:.2e}, "
                  f"approximations {d[1]:.2e} x 1e-13 raa n")
            assert err < 1e-9, (it, k)
            assert ok, (it, k, d)
            assert gg.concheck(fnew, gnew) == go.concheck(fnew, gnew)
        x = xo
