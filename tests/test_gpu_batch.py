"""Load cases and batched problems.

Load cases: the design loop sums compliance and sensitivities over the cases
(reference opt_multi_load.cpp:118-146); pinned against separate single-case
solves of the same density and against the oracle's compliance.
Batch: one captured graph per PCG round over problems of uneven size; every
problem must give what it gives alone (same kernels, same order: iteration
counts equal, solutions and designs to 1e-12), keep its FP64 verified stop,
and survive the autotuner's reconfiguration.  Both batched preconditioners
(Jacobi and the multigrid V-cycle) are held to that.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _cantilever_two_cases(gpu, dims):
    nx, ny, nz = dims
    c = gpu.cantilever(*dims)  # case 0: the edge point loads
    assert c.new_load_case() == 1
    c.add_neumann([0.0, 1.0 / (ny * nz), 0.0], [[nx, nx], [0, ny], [0, nz]])  # face load
    c.select_load_case(0)
    assert c.load_case_count() == 2
    return c


def test_load_cases_sum_compliance(gpu, oracle):
    dims = (24, 8, 8)
    c = _cantilever_two_cases(gpu, dims)
    c.filter_setup(1.5)
    x0 = np.full(c.nelx, 0.3)
    # compliance of each case on the first design (x_phys = filter(x0) = x0 inside)
    xf = c.density_filter(x0)
    c.set_density(xf)
    cs = []
    for lc in range(2):
        c.select_load_case(lc)
        u, st = c.solve(tol=1e-10)
        f = c.fext()
        o = oracle.Problem(*dims)
        o.add_dirichlet([1, 1, 1, 0, 0, 0], [0.0] * 6, [[0, 0], [0, dims[1]], [0, dims[2]]])
        o.make_dof()
        o.set_density(xf)
        assert np.allclose(o.spmv(u), f, atol=1e-9 * np.abs(f).max())  # K u = f_lc
        cs.append(f @ u)
    c.select_load_case(0)
    _, hist, st = c.simp_run(x0, 2, tol=1e-10)
    print(f"case compliances {cs}, loop step-1 compliance {hist[0, 0]}")
    assert abs(hist[0, 0] - sum(cs)) < 1e-8 * sum(cs)
    assert np.all(hist[:, 3] <= 1e-10)


def test_zero_load_case_converges(gpu):
    """A load case that lands entirely on clamped DOFs has f = 0: u = 0 exactly."""
    c = gpu.cantilever(12, 6, 6)
    c.new_load_case()
    c.add_neumann([0.0, 0.0, -1.0], [[0, 0], [0, 6], [0, 6]])  # the clamped face
    c.set_density(np.full(c.nelx, 0.5))
    u, st = c.solve(tol=1e-8)
    assert st.status == 0 and np.all(u == 0.0)


def _batch_specs():
    from icarat2_amd import workloads as W
    # uneven sizes, 1..4 cases, including a zero-load case (seed 106, case 2)
    return [W.batch_problem(s) for s in (108, 113, 106, 123)]


def _scaled(spec, factor):
    s = dict(spec)
    s["dims"] = tuple(max(8, d // factor) for d in spec["dims"])
    # loads are tractions of unit total force on the face
    loads = []
    for face, val in spec["loads"]:
        ax = "xyz".index(face[0])
        area = 1.0
        for a in range(3):
            if a != ax:
                area *= s["dims"][a]
        v = [0.0, 0.0, 0.0]
        d = int(np.flatnonzero(val)[0])
        v[d] = np.sign(val[d]) / area
        loads.append((face, v))
    s["loads"] = loads
    return s


@pytest.mark.parametrize("mg", [False, True])
def test_batch_solve_equals_individual_solves(gpu, mg):
    from icarat2_amd import workloads as W
    pre = gpu.PRECOND_MG if mg else gpu.PRECOND_JACOBI
    specs = [_scaled(s, 2) for s in _batch_specs()]
    ctxs = [W.build_context(gpu, s) for s in specs]
    alone = []
    for c, s in zip(ctxs, specs):
        x = W.initial_design(s) * (1.0 + 0.3 * np.sin(np.arange(c.nelx)))
        c.set_density(np.clip(x, 1e-3, 1.0))
        u, st = c.solve(tol=1e-9, precond=pre)
        alone.append((c.fext() @ u, st))
    b = gpu.Batch(ctxs)
    st = b.solve(tol=1e-9, precond=pre)
    n, launches = b.info()
    assert n == len(ctxs) and launches > 0
    for c, (c0, s0), s1 in zip(ctxs, alone, st):
        # the batched solution is the one the context holds now
        c1, _, _ = c.compliance_sensitivity()
        print(f"iterations alone {s0.iters} batched {s1.iters}; compliance {c0:.15g} {c1:.15g}")
        assert s1.status == 0 and s1.true_relres <= 1e-9
        assert s1.iters == s0.iters
        assert abs(c1 - c0) <= 1e-12 * abs(c0)


@pytest.mark.parametrize("mg", [False, True])
def test_batch_simp_equals_individual_loops(gpu, mg):
    from icarat2_amd import workloads as W
    pre = gpu.PRECOND_MG if mg else gpu.PRECOND_JACOBI
    specs = [_scaled(s, 2) for s in _batch_specs()]
    single = []
    for s in specs:
        c = W.build_context(gpu, s)
        single.append(c.simp_run(W.initial_design(s), 3, tol=1e-9, precond=pre))
    ctxs = [W.build_context(gpu, s) for s in specs]
    b = gpu.Batch(ctxs)
    xs, hists, sts = b.simp_run([W.initial_design(s) for s in specs], 3, tol=1e-9, precond=pre)
    for (x1, h1, s1), x2, h2, s2 in zip(single, xs, hists, sts):
        print(f"iters alone {s1.pcg_iters} batched {s2.pcg_iters}; max design diff "
              f"{np.max(np.abs(x1 - x2)):.2e}")
        assert s1.pcg_iters == s2.pcg_iters
        assert np.max(np.abs(x1 - x2)) < 1e-12
        assert np.allclose(h1, h2, rtol=1e-12, atol=0)
        assert np.all(h2[:, 3] <= 1e-9)


def test_batch_autotune_keeps_results(gpu):
    from icarat2_amd import workloads as W
    specs = [_scaled(s, 2) for s in _batch_specs()]
    ctxs = [W.build_context(gpu, s) for s in specs]
    for c, s in zip(ctxs, specs):
        c.set_density(W.initial_design(s))
    b = gpu.Batch(ctxs)
    st0 = b.solve(tol=1e-9)
    cfg, ms = b.autotune(reps=3)
    print(f"autotune: config {cfg}, ms per launch {ms}")
    assert 0 <= cfg <= 3 and all(t > 0 for t in ms)
    st1 = b.solve(tol=1e-9)
    for a, c in zip(st0, st1):
        assert c.status == 0 and c.true_relres <= 1e-9
        assert abs(a.iters - c.iters) <= max(2, 0.01 * a.iters)


@pytest.mark.parametrize("budget", [0, 1])
def test_batch_budget_does_not_change_results(gpu, budget):
    """Admission order and concurrency change only the schedule."""
    from icarat2_amd import workloads as W
    specs = [_scaled(s, 2) for s in _batch_specs()]
    ctxs = [W.build_context(gpu, s) for s in specs]
    b = gpu.Batch(ctxs)
    x0s = [W.initial_design(s) for s in specs]
    xa, ha, _ = b.simp_run(x0s, 2, tol=1e-9)
    b.set_budget(budget)
    xb, hb, _ = b.simp_run(x0s, 2, tol=1e-9)
    for x1, x2, h1, h2 in zip(xa, xb, ha, hb):
        assert np.max(np.abs(x1 - x2)) < 1e-12
        assert np.allclose(h1, h2, rtol=1e-12, atol=0)
