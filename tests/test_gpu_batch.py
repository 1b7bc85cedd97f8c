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
        # multigrid: the batched loop solves a problem's load cases together,
        # the standalone loops here one after another (the same results)
        single.append(c.simp_run(W.initial_design(s), 3, tol=1e-9, precond=pre,
                                 cases_together=False if mg else None))
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


def _oracle_problem(oracle, spec):
    """The oracle's problem of one batched spec: the clamped face, then every
    load case's face load (face_range and the unit tractions of workloads)."""
    from icarat2_amd import workloads as W
    dims = spec["dims"]
    o = oracle.Problem(*dims)
    o.add_dirichlet([1, 1, 1, 0, 0, 0], [0.0] * 6, W.face_range(spec["clamp"], dims))
    o.make_dof()
    return o


@pytest.mark.slow
def test_full_batched_workload(gpu, oracle):
    """All 24 problems of the batched workload (seeds 100..123, 32^3 to
    160x80x80, 1-4 face load cases each) in one batch, multigrid PCG, 2 design
    iterations: every solve at the verified FP64 residual, four problems equal
    to their standalone loops, and the smallest problem's first compliance
    (summed over its cases) against the oracle's own solves."""
    from icarat2_amd import workloads as W
    specs = W.batch_workload()
    assert len(specs) == 24
    ctxs = [W.build_context(gpu, s) for s in specs]
    x0s = [W.initial_design(s) for s in specs]
    b = gpu.Batch(ctxs)
    xs, hists, sts = b.simp_run(x0s, 2, tol=1e-8, maxit=100000, precond=gpu.PRECOND_MG)
    for s, h, st in zip(specs, hists, sts):
        assert np.all(h[:, 3] <= 1e-8), (s["seed"], h[:, 3])
        assert st.max_true_relres <= 1e-8
    print("batched: PCG iterations per problem", [int(st.pcg_iters) for st in sts])
    # four problems alone: smallest, largest, most load cases, and the first
    order = sorted(range(24), key=lambda i: np.prod(specs[i]["dims"]))
    picks = [order[0], order[-1], max(range(24), key=lambda i: len(specs[i]["loads"]))]
    for i in range(24):
        if len(set(picks)) >= 4:
            break
        picks.append(i)
    picks = sorted(set(picks))
    assert len(picks) == 4
    for i in picks:
        c = W.build_context(gpu, specs[i])
        x1, h1, s1 = c.simp_run(x0s[i], 2, tol=1e-8, maxit=100000, precond=gpu.PRECOND_MG)
        print(f"problem {specs[i]['seed']} {specs[i]['dims']}: alone {s1.pcg_iters} "
              f"batched {sts[i].pcg_iters} iterations")
        assert s1.pcg_iters == sts[i].pcg_iters
        assert np.max(np.abs(x1 - xs[i])) < 1e-12
        assert np.allclose(h1, hists[i], rtol=1e-12, atol=0)
    # the oracle's compliance of the smallest problem's first design
    i = order[0]
    s = specs[i]
    c = ctxs[i]
    xf = c.density_filter(x0s[i])
    o = _oracle_problem(oracle, s)
    o.set_density(xf)
    comp = 0.0
    for lc, (face, val) in enumerate(s["loads"]):
        q = _oracle_problem(oracle, s)
        q.add_neumann(val, W.face_range(face, s["dims"]))
        q.set_density(xf)
        if not np.any(q.fext()):
            continue  # a face load entirely on clamped DOFs: u = 0, no compliance
        u, it, _ = q.scaling_cg(tol=1e-12, maxit=200000)
        assert q.true_residual(u) < 1e-10
        comp += q.compliance(u)
    print(f"problem {s['seed']}: first compliance batched {hists[i][0, 0]:.15g} oracle {comp:.15g}")
    assert abs(hists[i][0, 0] - comp) < 1e-8 * abs(comp)
