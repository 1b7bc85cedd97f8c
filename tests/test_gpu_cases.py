"""Load cases solved together (icarat_hip_solve_cases).

The reference solves each load case in turn and sums the compliances
(sample/software/opt-multi-load/opt_multi_load.cpp:118-146).  The FP64
Jacobi-PCG and multigrid-PCG paths run every case in one multi-column
iteration; each case keeps
its own scalars and verified stop and does the arithmetic of its own solve, so
the results must equal the single-case solves bit for bit (solutions,
iteration counts, verification counts), and each case's solution must solve
K u = f_case by the oracle's assembled K.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _spec(seed, factor):
    from icarat2_amd import workloads as W
    s = dict(W.batch_problem(seed))
    s["dims"] = tuple(max(8, d // factor) for d in s["dims"])
    loads = []
    for face, val in s["loads"]:
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


def _design(c, s, phase=0.0):
    from icarat2_amd import workloads as W
    x = W.initial_design(s) * (1.0 + 0.3 * np.sin(np.arange(c.nelx) + phase))
    return np.clip(x, 1e-3, 1.0)


def _per_case(c, **kw):
    us, sts = [], []
    for lc in range(c.load_case_count()):
        c.select_load_case(lc)
        u, st = c.solve(**kw)
        us.append(u)
        sts.append(st)
    c.select_load_case(0)
    return np.array(us), sts


# seed 106: four cases on a 20^3 lattice, one of them a load on the clamped face
# (f = 0); seed 108: four cases on a 16^3 lattice; seed 105: three cases on a box
@pytest.mark.parametrize("seed,factor", [(106, 4), (108, 2), (105, 2)])
def test_cases_together_equal_single_solves(gpu, oracle, seed, factor):
    from icarat2_amd import workloads as W
    s = _spec(seed, factor)
    a = W.build_context(gpu, s)
    b = W.build_context(gpu, s)
    x = _design(a, s)
    a.set_density(x)
    b.set_density(x)
    nc = a.load_case_count()
    assert nc == len(s["loads"]) >= 3
    ua, sa = a.solve_cases(tol=1e-9)
    ub, sb = _per_case(b, tol=1e-9)
    for lc in range(nc):
        print(f"case {lc}: iterations together {sa[lc].iters} alone {sb[lc].iters}, "
              f"verifications {sa[lc].refine_steps} {sb[lc].refine_steps}, "
              f"true relres {sa[lc].true_relres:.3e}")
        assert sa[lc].status == 0 and sa[lc].true_relres <= 1e-9
        assert sa[lc].iters == sb[lc].iters
        assert sa[lc].refine_steps == sb[lc].refine_steps
        assert sa[lc].true_relres == sb[lc].true_relres
        assert np.array_equal(ua[lc], ub[lc])
    # each case against the oracle's assembled K
    o = oracle.Problem(*s["dims"])
    o.add_dirichlet([1, 1, 1, 0, 0, 0], [0.0] * 6, W.face_range(s["clamp"], s["dims"]))
    o.make_dof()
    o.set_density(x)
    for lc in range(nc):
        a.select_load_case(lc)
        f = a.fext()
        if not np.any(f):
            assert np.all(ua[lc] == 0.0)  # a load on clamped DOFs only
            continue
        r = np.linalg.norm(f - o.spmv(ua[lc])) / np.linalg.norm(f)
        assert r <= 1.5e-9, (lc, r)
    # the context holds the selected case's solution afterwards
    a.select_load_case(0)
    a.solve_cases(tol=1e-9, want_u=False)
    c0, _, _ = a.compliance_sensitivity()
    ref = a.fext() @ ua[0]  # (host summation order)
    assert abs(c0 - ref) <= 1e-12 * abs(ref)


def test_cases_together_warm_and_cap(gpu):
    """Warm starts from each case's last solution, and the iteration cap, as
    the single-case solves have them."""
    from icarat2_amd import workloads as W
    s = _spec(117, 2)
    a = W.build_context(gpu, s)
    b = W.build_context(gpu, s)
    for k, phase in enumerate((0.0, 0.7)):
        x = _design(a, s, phase)
        a.set_density(x)
        b.set_density(x)
        ua, sa = a.solve_cases(tol=1e-9, warm=k > 0)
        ub, sb = _per_case(b, tol=1e-9, warm=k > 0)
        assert [t.iters for t in sa] == [t.iters for t in sb]
        assert np.array_equal(ua, ub)
    ua, sa = a.solve_cases(tol=1e-12, maxit=40, allow_cap=True)
    ub, sb = _per_case(b, tol=1e-12, maxit=40, allow_cap=True)
    assert all(t.status != 0 for t in sa)
    assert [t.status for t in sa] == [t.status for t in sb]
    assert [t.iters for t in sa] == [t.iters for t in sb]
    assert np.array_equal(ua, ub)


def test_simp_cases_together_equals_one_after_another(gpu):
    from icarat2_amd import workloads as W
    s = _spec(120, 2)
    a = W.build_context(gpu, s)
    b = W.build_context(gpu, s)
    x0 = W.initial_design(s)
    xa, ha, sa = a.simp_run(x0, 4, tol=1e-9, warm=True, cases_together=True)
    xb, hb, sb = b.simp_run(x0, 4, tol=1e-9, warm=True, cases_together=False)
    print(f"PCG iterations {sa.pcg_iters} {sb.pcg_iters}; solve s {sa.solve_s:.3f} {sb.solve_s:.3f}")
    assert sa.pcg_iters == sb.pcg_iters
    assert np.array_equal(xa, xb)
    assert np.array_equal(ha, hb)
    assert np.all(ha[:, 3] <= 1e-9)


def _same(sa, sb, ua, ub):
    for lc, (p, q) in enumerate(zip(sa, sb)):
        print(f"case {lc}: status {p.status} {q.status}, iterations together {p.iters} alone "
              f"{q.iters}, verifications {p.refine_steps} {q.refine_steps}, true relres "
              f"{p.true_relres:.3e}")
        assert p.status == q.status
        assert p.iters == q.iters and p.refine_steps == q.refine_steps
        assert p.true_relres == q.true_relres
    assert np.array_equal(ua, ub)


# The multigrid solve of every case together (the z-marching PCG stencil, the
# cycle's transfers, sweeps and dense solve over the stacked cases), against
# each case solved on its own, bit for bit.  Hierarchies: the default one of
# a 32^3 problem (its fine level smoothed by the direct kernel); an 80^3
# problem with z-marching fine and level-1 sweeps (direct_nodes 20,000) and
# the W-cycle on levels 1-2; a 64x32x32 box with the V-cycle, plain updates
# and a smoothed coarsest level (no dense solve).
MG_CASES = [(108, {}),
            (106, {"direct_nodes": 20000, "w_levels": 2}),
            (105, {"w_levels": 0, "compensated": 0, "dense_dofs": 0, "nu_coarse": 8})]


@pytest.mark.parametrize("seed,mg", MG_CASES)
def test_cases_multigrid_together_equal_single_solves(gpu, oracle, seed, mg):
    from icarat2_amd import workloads as W
    s = _spec(seed, 1)
    a = W.build_context(gpu, s)
    b = W.build_context(gpu, s)
    x = _design(a, s)
    a.set_density(x)
    b.set_density(x)
    a.mg_setup(**mg)
    b.mg_setup(**mg)
    print(f"seed {seed}: hierarchy {a.mg_info()}")
    ua, sa = a.solve_cases(tol=1e-9, precond=gpu.PRECOND_MG)
    ub, sb = _per_case(b, tol=1e-9, precond=gpu.PRECOND_MG)
    _same(sa, sb, ua, ub)
    assert all(t.status == 0 and t.true_relres <= 1e-9 for t in sa)
    if seed == 105:
        # each case against the oracle's assembled K
        o = oracle.Problem(*s["dims"])
        o.add_dirichlet([1, 1, 1, 0, 0, 0], [0.0] * 6, W.face_range(s["clamp"], s["dims"]))
        o.make_dof()
        o.set_density(x)
        for lc in range(a.load_case_count()):
            a.select_load_case(lc)
            f = a.fext()
            if not np.any(f):
                assert np.all(ua[lc] == 0.0)  # a load on clamped DOFs only
                continue
            r = np.linalg.norm(f - o.spmv(ua[lc])) / np.linalg.norm(f)
            assert r <= 1.5e-9, (lc, r)


def test_cases_multigrid_warm_cap_and_stall(gpu):
    """Warm starts from each case's last solution, the iteration cap, and the
    stall guard at a tolerance below the FP64 floor (the best verified iterate
    returned, not converged), as the single-case solves have them."""
    from icarat2_amd import workloads as W
    s = _spec(117, 1)
    a = W.build_context(gpu, s)
    b = W.build_context(gpu, s)
    for k, phase in enumerate((0.0, 0.7)):
        x = _design(a, s, phase)
        a.set_density(x)
        b.set_density(x)
        ua, sa = a.solve_cases(tol=1e-9, precond=gpu.PRECOND_MG, warm=k > 0)
        ub, sb = _per_case(b, tol=1e-9, precond=gpu.PRECOND_MG, warm=k > 0)
        _same(sa, sb, ua, ub)
    kw = dict(precond=gpu.PRECOND_MG, allow_cap=True)
    ua, sa = a.solve_cases(tol=1e-12, maxit=10, **kw)
    ub, sb = _per_case(b, tol=1e-12, maxit=10, **kw)
    assert all(t.status != 0 for t in sa)
    _same(sa, sb, ua, ub)
    ua, sa = a.solve_cases(tol=1e-17, maxit=4000, **kw)
    ub, sb = _per_case(b, tol=1e-17, maxit=4000, **kw)
    assert all(t.status != 0 for t in sa)
    _same(sa, sb, ua, ub)


def test_simp_multigrid_cases_together_equals_one_after_another(gpu):
    from icarat2_amd import workloads as W
    s = _spec(120, 1)
    a = W.build_context(gpu, s)
    b = W.build_context(gpu, s)
    x0 = W.initial_design(s)
    kw = dict(tol=1e-9, warm=True, precond=gpu.PRECOND_MG)
    xa, ha, sa = a.simp_run(x0, 3, cases_together=True, **kw)
    xb, hb, sb = b.simp_run(x0, 3, cases_together=False, **kw)
    print(f"PCG iterations {sa.pcg_iters} {sb.pcg_iters}; solve s {sa.solve_s:.3f} {sb.solve_s:.3f}")
    assert sa.pcg_iters == sb.pcg_iters
    assert np.array_equal(xa, xb)
    assert np.array_equal(ha, hb)
    assert np.all(ha[:, 3] <= 1e-9)
