#!/usr/bin/env python3
"""Writes the optimizer fixtures of tests/golden from the reference's own
optimizer sources compiled into oracle/_ref (MMASolver.cpp, OCSolver.hpp; built
by oracle/Makefile where the reference tree is present).  The fixtures are
inputs and outputs only:
  mma_toy.npz   the reference's MMA toy problem (test_opt_solver.cpp:17-73):
                design after each of 10 updates
  mma_m1.npz    a 500-variable, one-constraint MMA run: per update the inputs
                (dfdx, g, dgdx) and the design after the update
  mma_m3.npz    a 400-variable, three-constraint MMA run with per-variable
                bounds and non-default parameters: per update the inputs
                (dfdx, gx, dgdx) and the design after it
  gcmma_toy.npz the reference's GCMMA toy problem (test_opt_solver.cpp:75-150,
                its loop as written): every xmma returned, the inner count
                per outer iteration, the final design
  gcmma_m3.npz  a 300-variable, three-constraint GCMMA loop (convex
                separable objective, weighted-volume constraints,
                per-variable bounds): every xmma returned and the inner
                count per outer iteration
  oc_filter.npz OC on an 8x6x5 mesh whose volume callback is the cone filter
                (r = 1.5) summed, as the SIMP drivers pass it: inputs and the
                updated design
Run from the repository root: python tests/golden/make_golden.py
"""
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, os.path.join(ROOT, "oracle"))
import oracle as O  # noqa: E402

HERE = os.path.dirname(os.path.abspath(__file__))


def toy(s):
    # the reference's toy problem (test_opt_solver.cpp:17-73)
    f = s[0] ** 2 + s[1] ** 2 + s[2] ** 2
    df = 2.0 * s
    g = np.array([(s[0] - 5) ** 2 + (s[1] - 2) ** 2 + (s[2] - 1) ** 2 - 9.0,
                  (s[0] - 3) ** 2 + (s[1] - 4) ** 2 + (s[2] - 3) ** 2 - 9.0])
    dg = np.array([2 * (s[0] - 5), 2 * (s[1] - 2), 2 * (s[2] - 1),
                   2 * (s[0] - 3), 2 * (s[1] - 4), 2 * (s[2] - 3)])
    # dgdx[i * m + j]
    dgx = np.array([dg[0], dg[3], dg[1], dg[4], dg[2], dg[5]])
    return f, df, g, dgx


def main():
    if O.ref_lib() is None:
        sys.exit("oracle/_ref is not built: run make -C oracle where the reference tree is present")
    # MMA toy
    m = O.MMA(3, 2, reference=True)
    m.set(0.5, 0.1, 0.5, 0.7, 1.2)
    s = np.array([4.0, 3.0, 2.0])
    traj = []
    for _ in range(10):
        f, df, g, dg = toy(s)
        m.update(s, df, g, dg, np.zeros(3), np.full(3, 5.0))
        traj.append(s.copy())
    np.savez(os.path.join(HERE, "mma_toy.npz"), x0=np.array([4.0, 3.0, 2.0]), traj=np.array(traj),
             params=np.array([0.5, 0.1, 0.5, 0.7, 1.2]))
    # MMA, n = 500, m = 1
    rng = np.random.default_rng(6)
    n = 500
    x0 = rng.uniform(0.1, 0.9, n)
    x = x0.copy()
    mm = O.MMA(n, 1, reference=True)
    dfs, gs, dgs, xs = [], [], [], []
    for _ in range(6):
        df = -rng.uniform(0.1, 3.0, n)
        dg = np.full(n, 1.0 / n)
        gval = np.array([x.mean() - 0.4])
        mm.update(x, df, gval, dg, np.full(n, 1e-4), np.ones(n))
        dfs.append(df)
        gs.append(gval[0])
        dgs.append(dg)
        xs.append(x.copy())
    np.savez(os.path.join(HERE, "mma_m1.npz"), x0=x0, dfdx=np.array(dfs), g=np.array(gs),
             dgdx=np.array(dgs), x=np.array(xs))
    # MMA, n = 400, m = 3: weighted-volume constraints, per-variable bounds
    rng = np.random.default_rng(8)
    n, m = 400, 3
    x0 = rng.uniform(0.2, 0.8, n)
    xmin = rng.uniform(0.0, 0.1, n)
    xmax = rng.uniform(0.9, 1.0, n)
    w = rng.uniform(0.5, 1.5, (m, n)) / n
    cap = np.array([0.45, 0.5, 0.4])
    x = x0.copy()
    mm = O.MMA(n, m, reference=True)
    mm.set(0.3, 0.15, 0.3, 0.65, 1.1)
    dfs, gs, dgs, xs = [], [], [], []
    for _ in range(8):
        df = -rng.uniform(0.1, 3.0, n) * (1.0 + x)
        gx = w @ x - cap
        dg = np.ascontiguousarray(w.T).reshape(-1)  # dgdx[i m + j]
        mm.update(x, df, gx, dg, xmin, xmax)
        dfs.append(df)
        gs.append(gx)
        dgs.append(dg)
        xs.append(x.copy())
    np.savez(os.path.join(HERE, "mma_m3.npz"), x0=x0, xmin=xmin, xmax=xmax, dfdx=np.array(dfs),
             gx=np.array(gs), dgdx=np.array(dgs), x=np.array(xs),
             params=np.array([0.3, 0.15, 0.3, 0.65, 1.1]))
    # GCMMA toy (test_opt_solver.cpp:75-150 as written: objectivenew is
    # evaluated before the outer update, on the zero vector)
    def toy_fg(v):
        return toy(v)[0], toy(v)[2]
    g = O.GCMMA(3, 2, reference=True)
    g.set(0.5, 0.1, 0.5, 0.7, 1.2)
    s = np.array([4.0, 3.0, 2.0])
    xs_all, inners = [], []
    for _ in range(10):
        f, df, gx, dg = toy(s)
        snew = np.zeros(3)
        fnew = float(np.sum(snew ** 2))
        snew = g.outer(s, f, df, gx, dg, np.zeros(3), np.full(3, 5.0))
        xs_all.append(snew.copy())
        _, gnew = toy_fg(snew)
        ok = g.concheck(fnew, gnew)
        k = 0
        while not ok and k < 15:
            snew = g.inner(snew, fnew, gnew, s, f, df, gx, dg, np.zeros(3), np.full(3, 5.0))
            xs_all.append(snew.copy())
            fnew, gnew = toy_fg(snew)
            ok = g.concheck(fnew, gnew)
            k += 1
        inners.append(k)
        s = snew.copy()
    np.savez(os.path.join(HERE, "gcmma_toy.npz"), x0=np.array([4.0, 3.0, 2.0]),
             xmma=np.array(xs_all), inners=np.array(inners), x=s,
             params=np.array([0.5, 0.1, 0.5, 0.7, 1.2]))
    # GCMMA, n = 300, m = 3
    rng = np.random.default_rng(9)
    n, m = 300, 3
    cw = rng.uniform(0.5, 2.0, n) / n
    w = rng.uniform(0.5, 1.5, (m, n)) / n
    cap = np.array([0.4, 0.45, 0.35])
    xmin = rng.uniform(0.01, 0.05, n)
    xmax = rng.uniform(0.95, 1.0, n)

    def fg(v):
        return float(np.sum(cw / v)), w @ v - cap

    x = rng.uniform(0.3, 0.7, n)
    x0 = x.copy()
    g = O.GCMMA(n, m, reference=True)
    g.set(0.2, 0.2, 0.3, 0.7, 1.2)
    xs_all, inners = [], []
    dg = np.ascontiguousarray(w.T).reshape(-1)
    for _ in range(6):
        f, gx = fg(x)
        df = -cw / x ** 2
        xm = g.outer(x, f, df, gx, dg, xmin, xmax)
        xs_all.append(xm.copy())
        fnew, gnew = fg(xm)
        ok = g.concheck(fnew, gnew)
        k = 0
        while not ok and k < 5:
            xm = g.inner(xm, fnew, gnew, x, f, df, gx, dg, xmin, xmax)
            xs_all.append(xm.copy())
            fnew, gnew = fg(xm)
            ok = g.concheck(fnew, gnew)
            k += 1
        inners.append(k)
        x = xm.copy()
    np.savez(os.path.join(HERE, "gcmma_m3.npz"), x0=x0, cw=cw, w=w, cap=cap, xmin=xmin, xmax=xmax,
             xmma=np.array(xs_all), inners=np.array(inners),
             params=np.array([0.2, 0.2, 0.3, 0.7, 1.2]))
    # OC with the filtered-volume callback
    rng = np.random.default_rng(7)
    pr = O.Problem(8, 6, 5)
    pr.make_dof()
    F = O.Filter(pr, 1.5)
    n = pr.nelx
    x = rng.uniform(0.2, 0.6, n)
    dc = -rng.uniform(0.01, 2.0, n)
    dv = np.ones(n)
    g = 0.35 * n

    def vol(ptr, nn, user):
        arr = np.ctypeslib.as_array(ptr, shape=(nn,)).copy()
        return float(F.density_filter(arr, False).sum())

    cb = O.REF_VOLUME_FN(vol)
    x2 = x.copy()
    O.ref_lib().ref_oc_update(n, O.P(x2), O.P(dc.copy()), g, O.P(dv.copy()),
                              O.P(np.full(n, 1e-4)), O.P(np.ones(n)), cb, None)
    np.savez(os.path.join(HERE, "oc_filter.npz"), dims=np.array([8, 6, 5]), radius=1.5, x=x,
             dfdx=dc, dgdx=dv, g=g, x_new=x2)
    print("wrote", sorted(f for f in os.listdir(HERE) if f.endswith(".npz")))


if __name__ == "__main__":
    main()
