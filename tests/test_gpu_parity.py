"""GPU parity: the HIP path through the C ABI against the CPU oracle on the same
seeded inputs.  Tolerances (FP64 path):
  K.u, diag(K) ............... 1e-13 of max |y| (summation order only)
  filter (both modes) ........ bit-exact (same order, no contraction)
  threshold .................. 1e-15 relative
  compliance ................. 1e-8 relative (solves at the configs' 1e-8)
  sensitivities .............. 1e-8 relative (solves at 1e-11), 1e-12 given the same u
  OC / MMA update ............ 1e-12 / 1e-9 given the same inputs
  SIMP design after 20 its ... 1e-6 (the tolerance on design iterates)
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _pair(capi, oracle, dims, seed=None, dens=None):
    c = capi.cantilever(*dims)
    o = oracle.cantilever(*dims)
    if dens is None:
        rng = np.random.default_rng(seed)
        dens = rng.uniform(1e-3, 1.0, c.nelx)
    c.set_density(dens)
    o.set_density(dens)
    return c, o, dens


def _rel(a, b):
    return np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300)


@pytest.mark.parametrize("dims", [(60, 20, 20), (7, 5, 3), (33, 9, 5), (1, 1, 1)])
def test_apply_k_matches_oracle(gpu, oracle, dims):
    c, o, _ = _pair(gpu, oracle, dims, seed=11)
    rng = np.random.default_rng(12)
    for _ in range(2):
        u = rng.standard_normal(c.numeq)
        assert _rel(c.apply_K(u), o.spmv(u)) < 1e-13
    assert _rel(c.diag(), o.diag()) < 1e-14


@pytest.mark.parametrize("kernel", [0, 1])
@pytest.mark.parametrize("sym_max", [0, 1, 2])
@pytest.mark.parametrize("lengths", [None, (40.0, 9.0, 21.0)])
def test_apply_k_variants(gpu, oracle, kernel, sym_max, lengths):
    """Every FP64 K.u kernel (z-marching / tiles) and KE0 source (memory, box
    orbits, cube orbits) on a cubic-voxel and a box-voxel mesh."""
    dims = (20, 9, 7)
    L = lengths or tuple(float(d) for d in dims)
    c = gpu.Context(*dims, *L)
    o = oracle.Problem(*dims, *L)
    for p in (c, o):
        p.add_dirichlet([1, 1, 1, 0, 0, 0], [0.0] * 6, [[0, 0], [0, dims[1]], [0, dims[2]]])
        p.make_dof()
    rng = np.random.default_rng(14)
    x = rng.uniform(1e-3, 1.0, c.nelx)
    c.set_density(x)
    o.set_density(x)
    used = c.set_stencil(kernel, sym_max)
    assert used == (min(sym_max, 2) if lengths is None else min(sym_max, 1))
    u = rng.standard_normal(c.numeq)
    assert _rel(c.apply_K(u), o.spmv(u)) < 1e-13


def test_apply_k_partial_dirichlet(gpu, oracle):
    args = [([1, 0, 1], [0.0] * 3, [[0, 0], [0, 6], [0, 5]]),
            ([0, 1, 0], [0.0] * 3, [[9, 9], [0, 0], [0, 5]])]
    c = gpu.Context(9, 6, 5)
    o = oracle.Problem(9, 6, 5)
    for a in args:
        c.add_dirichlet(*a)
        o.add_dirichlet(*a)
    c.make_dof()
    o.make_dof()
    rng = np.random.default_rng(13)
    x = rng.uniform(0.01, 1, c.nelx)
    c.set_density(x)
    o.set_density(x)
    u = rng.standard_normal(c.numeq)
    assert _rel(c.apply_K(u), o.spmv(u)) < 1e-13


def test_solve_config1_uniform(gpu, oracle):
    c, o, x = _pair(gpu, oracle, (60, 20, 20), dens=np.full(24000, 0.3))
    u, st = c.solve(tol=1e-8)
    uo, it, _ = o.scaling_cg(tol=1e-8)
    assert st.status == 0
    assert st.true_relres < 1e-8
    assert abs(st.iters - it) <= max(3, 0.02 * it)
    cc, df, dg = c.compliance_sensitivity()
    co = o.compliance(uo)
    assert abs(cc - co) < 1e-8 * abs(co)
    assert abs(cc - c.fext() @ u) < 1e-12 * abs(cc)
    # the FP64 true residual of the GPU solution, recomputed by the oracle
    assert o.true_residual(u) < 1e-8


def test_sensitivity_matches_oracle(gpu, oracle):
    c, o, x = _pair(gpu, oracle, (24, 10, 8), seed=14)
    u, st = c.solve(tol=1e-11)
    uo, _, _ = o.scaling_cg(tol=1e-11)
    cc, df, dg = c.compliance_sensitivity()
    dfo, dgo = o.sensitivity(o.node_values(uo))
    assert _rel(df, dfo) < 1e-8
    assert np.max(np.abs(dg - dgo)) < 1e-15
    # same u on both sides: only rounding differs
    dfo2, _ = o.sensitivity(o.node_values(u))
    assert _rel(df, dfo2) < 1e-12


def test_nonzero_dirichlet_solve(gpu, oracle):
    nx, ny, nz = 10, 4, 4
    c = gpu.Context(nx, ny, nz)
    o = oracle.Problem(nx, ny, nz)
    for P in (c, o):
        P.add_dirichlet([1, 1, 1], [0.0, 0.0, 0.0], [[0, 0], [0, ny], [0, nz]])
        P.add_dirichlet([0, 0, 1], [0.0, 0.0, -0.01], [[nx, nx], [0, ny], [0, nz]])
        P.make_dof()
    rng = np.random.default_rng(15)
    x = rng.uniform(0.2, 1.0, c.nelx)
    c.set_density(x)
    o.set_density(x)
    u, st = c.solve(tol=1e-12)
    rhs = o.fext() - o.apply_k_elementwise(np.zeros(o.numeq), o.dirichlet_values())
    uo, _, _ = o.scaling_cg(rhs=rhs, tol=1e-12)
    assert _rel(u, uo) < 1e-8
    # equilibrium of the free rows with the prescribed displacements included
    r = o.apply_k_elementwise(u, o.dirichlet_values())
    assert np.max(np.abs(r)) < 1e-9 * np.max(np.abs(rhs))


def test_solve_cap_is_reported(gpu):
    c = gpu.cantilever(20, 8, 8)
    c.set_density(np.full(c.nelx, 0.5))
    with pytest.raises(gpu.IcaratError) as e:
        c.solve(tol=1e-14, maxit=5)
    assert e.value.code == gpu.NOT_CONVERGED
    u, st = c.solve(tol=1e-14, maxit=5, allow_cap=True)
    assert st.iters == 6 and st.status == gpu.NOT_CONVERGED


@pytest.mark.parametrize("radius", [1.2, 1.5, 2.0, 2.5, 3.0])
def test_filter_bit_exact(gpu, oracle, radius):
    dims = (37, 11, 9)
    c = gpu.Context(*dims)
    c.make_dof()
    o = oracle.Problem(*dims)
    o.make_dof()
    c.filter_setup(radius)
    F = oracle.Filter(o, radius, brute=True)
    rng = np.random.default_rng(16)
    x = rng.uniform(1e-3, 1.0, c.nelx)
    assert np.array_equal(c.density_filter(x, False), F.density_filter(x, False))
    v = rng.standard_normal(c.nelx)
    assert np.array_equal(c.density_filter(v, True), F.density_filter(v, True))


def test_filter_reference_kat_gpu(gpu):
    c = gpu.Context(2, 2, 1)
    c.make_dof()
    c.filter_setup(2.0)
    s = np.array([1e-4, 1.0, 1e-4, 1.0])
    ans = np.array([0.34587, 0.65423, 0.34587, 0.65423])
    assert np.sum((c.density_filter(s) - ans) ** 2) < 1e-10
    assert np.sum((c.density_filter(s, True) - ans) ** 2) < 1e-10
    sf = c.density_filter(s)
    st = c.threshold(sf, 16.0)
    assert st[0] < 0.01 and st[1] > 0.99
    assert np.sum((c.threshold(sf, 1.0, deriv=True) - 1.0) ** 2) < 0.015


def test_sensitivity_filter_gpu(gpu, oracle):
    """Filter::sensitivityFilter: the KAT of test_filter.cpp:63-89, then bit
    equality with the oracle on a 40x20x10 lattice with random densities."""
    c = gpu.Context(2, 2, 1)
    c.make_dof()
    c.filter_setup(2.0)
    answer = np.array([0.34587, 0.65423, 0.34587, 0.65423])
    out = c.sensitivity_filter(np.full(4, 0.5), np.array([1e-4, 1.0, 1e-4, 1.0]))
    assert np.sum((out - answer) ** 2) < 1e-10
    dims = (40, 20, 10)
    rng = np.random.default_rng(23)
    for radius in (1.5, 2.5):
        c = gpu.Context(*dims, 40.0, 20.0, 10.0)
        c.make_dof()
        c.filter_setup(radius)
        o = oracle.Problem(*dims, 40.0, 20.0, 10.0)
        o.make_dof()
        F = oracle.Filter(o, radius)
        x = rng.uniform(1e-3, 1.0, c.nelx)
        dfds = -rng.uniform(0.0, 3.0, c.nelx)
        assert np.array_equal(c.sensitivity_filter(x, dfds), F.sensitivity_filter(x, dfds))


def test_threshold_matches_oracle(gpu, oracle):
    c = gpu.Context(20, 10, 5)
    c.make_dof()
    x = np.random.default_rng(17).uniform(0, 1, c.nelx)
    for beta in (1.0, 4.0, 16.0):
        assert _rel(c.threshold(x, beta), oracle.threshold(x, beta)) < 1e-15
        assert _rel(c.threshold(x, beta, deriv=True), oracle.threshold_deriv(x, beta)) < 1e-14


def test_oc_matches_oracle(gpu, oracle):
    dims = (30, 10, 10)
    c = gpu.Context(*dims)
    c.make_dof()
    o = oracle.Problem(*dims)
    o.make_dof()
    c.filter_setup(1.5)
    F = oracle.Filter(o, 1.5)
    rng = np.random.default_rng(18)
    n = c.nelx
    x = rng.uniform(0.05, 0.8, n)
    dc = F.density_filter(-rng.uniform(0.01, 5.0, n), True)
    dv = F.density_filter(np.ones(n), True)
    g = 0.3 * n
    for proj, beta in ((False, 1.0), (True, 2.0)):
        xg, steps = c.oc_update(x, dc, g, dv, use_projection=proj, beta=beta)
        xo, cnt, lam = F.oc_update(x, dc, g, dv, use_projection=proj, beta=beta)
        assert steps == cnt
        assert np.max(np.abs(xg - xo)) < 1e-12


def test_mma_matches_oracle(gpu, oracle):
    dims = (20, 10, 6)
    c = gpu.Context(*dims)
    c.make_dof()
    n = c.nelx
    rng = np.random.default_rng(19)
    x0 = rng.uniform(0.1, 0.9, n)
    xg, xo = x0.copy(), x0.copy()
    mo = oracle.MMA(n, 1)
    c.mma_reset()
    for it in range(6):
        df = -rng.uniform(0.1, 3.0, n)
        dg = np.full(n, 1.0 / (0.4 * n))
        gval = xg.mean() / 0.4 - 1.0
        xg, outer = c.mma_update(xg, df, gval, dg)
        mo.update(xo, df, np.array([gval]), dg, np.full(n, 1e-4), np.ones(n))
        assert np.max(np.abs(xg - xo)) < 1e-9, it


def _golden(name):
    import os
    return np.load(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", name))


def test_oc_matches_reference_fixture(gpu):
    """The device OC against the compiled reference OCSolver.hpp (tests/golden)."""
    d = _golden("oc_filter.npz")
    c = gpu.Context(*[int(v) for v in d["dims"]])
    c.make_dof()
    c.filter_setup(float(d["radius"]))
    x1, steps = c.oc_update(d["x"], d["dfdx"], float(d["g"]), d["dgdx"])
    assert np.max(np.abs(x1 - d["x_new"])) < 1e-12


def test_mma_matches_reference_fixture(gpu):
    """The device MMA (m = 1) against the compiled reference MMASolver.cpp (tests/golden)."""
    d = _golden("mma_m1.npz")
    c = gpu.Context(10, 10, 5)
    c.make_dof()
    assert c.nelx == d["x0"].size
    c.mma_reset()
    x = d["x0"].copy()
    for k in range(d["x"].shape[0]):
        x, _ = c.mma_update(x, d["dfdx"][k], float(d["g"][k]), d["dgdx"][k])
        assert np.max(np.abs(x - d["x"][k])) < 1e-9, k


def test_simp_oc_config1_20_iterations(gpu, oracle):
    """Config 1: 60x20x20 cantilever, rmin 1.5, x0 = 0.3, OC, 20 iterations."""
    c = gpu.cantilever(60, 20, 20)
    o = oracle.cantilever(60, 20, 20)
    c.filter_setup(1.5)
    F = oracle.Filter(o, 1.5)
    x0 = np.full(c.nelx, 0.3)
    xg, hg, sg = c.simp_run(x0, 20, tol=1e-8)
    xo, ho, so = F.simp_run(x0, 20, cg_tol=1e-8)
    assert np.max(np.abs(xg - xo)) < 1e-6
    assert np.max(np.abs(hg[:, 0] - ho[:, 0]) / ho[:, 0]) < 1e-7
    assert np.max(np.abs(hg[:, 1] - ho[:, 1]) / ho[:, 1]) < 1e-9
    assert np.all(hg[:, 3] < 1e-8)
    assert hg[-1, 0] < 0.7 * hg[0, 0]  # the design improves


def test_simp_mma_small(gpu, oracle):
    c = gpu.cantilever(24, 8, 8)
    o = oracle.cantilever(24, 8, 8)
    c.filter_setup(1.5)
    F = oracle.Filter(o, 1.5)
    x0 = np.full(c.nelx, 0.3)
    xg, hg, _ = c.simp_run(x0, 8, optimizer=1, tol=1e-10)
    xo, ho, _ = F.simp_run(x0, 8, optimizer=1, cg_tol=1e-10)
    assert np.max(np.abs(xg - xo)) < 1e-6
    assert np.max(np.abs(hg[:, 0] - ho[:, 0]) / ho[:, 0]) < 1e-7


@pytest.mark.slow
def test_config2_one_solve_sens_filter(gpu, oracle):
    """Config 2 at full size (192x64x64, Gaussian-field density seed 2): K.u and
    the filter against the oracle, the solve judged by the FP64 true residual,
    the sensitivities against the oracle on the same u."""
    dims = (192, 64, 64)
    x, m = gpu.gaussian_field(*dims, 2)
    c = gpu.cantilever(*dims)
    c.set_density(x)
    o = oracle.cantilever(*dims)
    o.set_density(x)
    rng = np.random.default_rng(20)
    u = rng.standard_normal(c.numeq)
    assert _rel(c.apply_K(u), o.spmv(u)) < 1e-13
    v = rng.standard_normal(c.numeq)
    a, b = v @ c.apply_K(u), u @ c.apply_K(v)
    assert abs(a - b) < 1e-11 * abs(a)
    ug, st = c.solve(tol=1e-8)
    assert st.status == 0 and st.true_relres < 1e-8
    # the oracle recomputes the true residual of the GPU solution with its
    # assembled K (analysis/linear.hpp:93-103)
    r_or = o.true_residual(ug)
    print(f"C2 true residual: GPU {st.true_relres:.4e}, oracle {r_or:.4e}")
    assert r_or < 1e-8
    assert abs(r_or - st.true_relres) < 1e-3 * r_or
    cc, df, dg = c.compliance_sensitivity()
    dfo, _ = o.sensitivity(o.node_values(ug))
    assert _rel(df, dfo) < 1e-12
    assert abs(cc - o.compliance(ug)) < 1e-12 * abs(cc)
    c.filter_setup(1.5)
    F = oracle.Filter(o, 1.5)
    assert np.array_equal(c.density_filter(x), F.density_filter(x))
    assert np.array_equal(c.density_filter(df, True), F.density_filter(df, True))
