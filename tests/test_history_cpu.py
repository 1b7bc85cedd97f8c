"""The drivers' history file and convergence test (reference io/export_graph.hpp,
optimization/opt_base.hpp:28-44) without a GPU: the product's writers against
the reference's own headers compiled into oracle/_ref (text bit for bit), and
the convergence rule the design loop applies against the compiled
Optimize::isConvergence."""
import os

import numpy as np
import pytest


def _ref(oracle):
    R = oracle.ref_lib()
    if R is None:
        pytest.skip("oracle/_ref not built (reference tree absent)")
    return R


def test_add_plot_matches_reference(capi, oracle, tmp_path):
    R = _ref(oracle)
    vals = [1234.56789, 0.1, 1.0e-9, 3.0, 98765432.1, -2.5e-5, 1.0 / 3.0]
    a, b = str(tmp_path / "ours.csv"), str(tmp_path / "ref.csv")
    for k, v in enumerate(vals, start=1):
        capi.add_plot(a, k, v, k)
        R.ref_add_plot(os.fsencode(b), k, v, k)
    assert open(a).read() == open(b).read()
    # is_first == 1 truncates: a second run leaves only its own lines
    capi.add_plot(a, 1, 7.0, 1)
    R.ref_add_plot(os.fsencode(b), 1, 7.0, 1)
    assert open(a).read() == open(b).read() == "1 7\n"


def test_export_graph_matches_reference(capi, oracle, tmp_path):
    import ctypes as C
    R = _ref(oracle)
    rng = np.random.default_rng(4)
    yy = rng.standard_normal(17) * 10.0 ** rng.integers(-8, 8, 17)
    xx = np.cumsum(rng.uniform(0.1, 2.0, 17))
    dp = C.POINTER(C.c_double)
    for x in (None, xx):
        a, b = str(tmp_path / "ours.txt"), str(tmp_path / "ref.txt")
        capi.export_graph(a, yy, x)
        R.ref_export_graph(os.fsencode(b), None if x is None else x.ctypes.data_as(dp),
                           yy.ctypes.data_as(dp), yy.size)
        assert open(a).read() == open(b).read()


def _relative_rule(f, eps, minstep):
    """The rule of the design loop: converged at step k (1-based) when
    |f_old - f_k| / f_k < eps and k > minstep, f_old the last step that did
    not pass (the stated rule of Optimize::isConvergence)."""
    out, old = [], 0.0
    for k, v in enumerate(f, start=1):
        if k > minstep and abs(old - v) / v < eps:
            out.append(0)
        else:
            out.append(1)
            old = v
    return out


def test_convergence_rule_against_compiled_reference(oracle):
    import ctypes as C
    R = _ref(oracle)

    def ref(f, eps, minstep):
        f = np.ascontiguousarray(f, dtype=np.float64)
        out = (C.c_int * f.size)()
        R.ref_is_convergence(f.ctypes.data_as(C.POINTER(C.c_double)), f.size, eps,
                             float(minstep), out)
        return list(out)

    # changes of at least 1 per step: the compiled header and the rule agree
    f = 1.0e6 * np.exp(-0.05 * np.arange(80)) + 5.0e4
    for eps, minstep in ((1e-5, 50), (1e-3, 10), (5e-2, 0)):
        assert ref(f, eps, minstep) == _relative_rule(f, eps, minstep)
    # below 1: compiled on its own, the header's unqualified abs is ::abs(int),
    # which truncates |f_old - f| (0.5 -> 0) and declares convergence that the
    # relative test (5e-4 > 1e-5) does not; the design loop follows the
    # relative test
    g = [1000.0] * 60 + [999.5]
    assert ref(g, 1e-5, 50)[-1] == 0
    assert _relative_rule(g, 1e-5, 50)[-1] == 1
