"""The design loop's convergence stop (Optimize::isConvergence, reference
optimization/opt_base.hpp:28-44, called after the compliance at
opt_macro_oc.cpp:137) and history file (addPlot, io/export_graph.hpp:56-70,
opt_macro_oc.cpp:134) on the GPU."""
import os

import numpy as np
import pytest

from test_history_cpu import _relative_rule

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("precond", ["jacobi", "mg"])
def test_simp_convergence_stop_and_history(gpu, oracle, tmp_path, precond):
    pre = gpu.PRECOND_MG if precond == "mg" else gpu.PRECOND_JACOBI
    c = gpu.cantilever(24, 8, 8)
    c.filter_setup(1.5)
    x0 = np.full(c.nelx, 0.3)
    path = str(tmp_path / "function.csv")
    eps, minstep = 2e-3, 5
    x, hist, st = c.simp_run(x0, 150, tol=1e-10, precond=pre, conv_eps=eps,
                             conv_minstep=minstep, history_csv=path)
    print(f"{precond}: converged {st.converged} after {st.steps} iterations")
    assert st.converged == 1 and minstep < st.steps < 150 and hist.shape == (st.steps, 4)
    flags = _relative_rule(hist[:, 0], eps, minstep)
    assert flags[-1] == 0 and all(f == 1 for f in flags[:-1])
    # the converged iteration's update ran: the same design as a fixed-length
    # loop (a fresh context, so the multigrid hierarchy starts the same way)
    c2 = gpu.cantilever(24, 8, 8)
    c2.filter_setup(1.5)
    x2, h2, _ = c2.simp_run(x0, st.steps, tol=1e-10, precond=pre)
    assert np.array_equal(x, x2) and np.array_equal(hist, h2)
    lines = open(path).read().splitlines()
    assert len(lines) == st.steps and lines[0].split()[0] == "1"
    R = oracle.ref_lib()
    if R is not None:
        ref = str(tmp_path / "ref.csv")
        for k in range(st.steps):
            R.ref_add_plot(os.fsencode(ref), k + 1, float(hist[k, 0]), k + 1)
        assert open(ref).read() == open(path).read()
