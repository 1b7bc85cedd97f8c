"""Pins the CPU oracle against the reference's own known-answer tests, closed
forms, invariants and independent formulations (no GPU)."""
import os
from fractions import Fraction

import numpy as np
import pytest


# ---------------------------------------------------------------- mesh / DOF
def test_counts_config1(oracle):
    pr = oracle.cantilever(60, 20, 20)
    # io/mesh_icarat.hpp:106-111 closed forms
    assert pr.numnp == 61 * 21 * 21 == 26901
    assert pr.nelx == 60 * 20 * 20 == 24000
    assert pr.neq == 3 * 26901 == 80703
    assert pr.numeq == 80703 - 3 * 21 * 21 == 79380


def test_node_and_element_numbering(oracle):
    nx, ny, nz = 3, 2, 2
    pr = oracle.Problem(nx, ny, nz, 1.5, 2.0, 1.0)
    X = pr.coords()
    ids = np.arange(pr.numnp)
    k, j, i = ids % (nx + 1), (ids // (nx + 1)) % (ny + 1), ids // ((nx + 1) * (ny + 1))
    assert np.array_equal(X[:, 0], (1.5 / nx) * k)
    assert np.array_equal(X[:, 1], (2.0 / ny) * j)
    assert np.array_equal(X[:, 2], (1.0 / nz) * i)
    conn = pr.connectivity()
    e = np.arange(pr.nelx)
    ex, ey, ez = e % nx, (e // nx) % ny, e // (nx * ny)
    base = ex + (nx + 1) * ey + (nx + 1) * (ny + 1) * ez
    sx, sxy = nx + 1, (nx + 1) * (ny + 1)
    expect = np.stack([base, base + 1, base + 1 + sx, base + sx, base + sxy, base + 1 + sxy,
                       base + 1 + sx + sxy, base + sx + sxy], axis=1)
    assert np.array_equal(conn, expect)


def test_dof_numbering_clamped_face(oracle):
    nx, ny, nz = 5, 3, 2
    pr = oracle.cantilever(nx, ny, nz)
    dof = pr.dof()
    n = np.arange(pr.numnp)
    clamped = (n % (nx + 1)) == 0
    nclamped_before = np.cumsum(clamped) - clamped
    free = ~clamped
    # free DOFs: 3*(n - #clamped nodes before n) + d
    for d in range(3):
        assert np.array_equal(dof[free, d], 3 * (n[free] - nclamped_before[free]) + d)
    # constrained DOFs follow in node order
    numeq = pr.numeq
    assert np.array_equal(dof[clamped].ravel(), numeq + np.arange(3 * clamped.sum()))


def test_dof_numbering_partial_flags_bruteforce(oracle):
    """Partial Dirichlet flags: free first in (node, comp) order, then the
    constrained in (node, comp) order (io/boundary_condition_base.hpp:35-95)."""
    pr = oracle.Problem(4, 3, 2)
    pr.add_dirichlet([1, 0, 1], [0.0, 0.0, 0.0], [[0, 0], [0, 3], [0, 2]])
    pr.add_dirichlet([0, 1, 0], [0.0, 0.5, 0.0], [[4, 4], [0, 0], [0, 2]])
    pr.add_dirichlet([0, 0, 0], [0.0, 0.0, 0.0], [[2, 2], [1, 1], [1, 1]])  # flags 0: all free
    pr.make_dof()
    X = pr.coords()
    flags = np.zeros((pr.numnp, 3), dtype=int)
    on0 = np.isclose(X[:, 0], 0)
    flags[on0, 0] = 1
    flags[on0, 2] = 1
    on4 = np.isclose(X[:, 0], 4) & np.isclose(X[:, 1], 0)
    flags[on4, 1] = 1
    expect = -np.ones((pr.numnp, 3), dtype=int)
    q = 0
    for nd in range(pr.numnp):
        for c in range(3):
            if not flags[nd, c]:
                expect[nd, c] = q
                q += 1
    numeq = q
    for nd in range(pr.numnp):
        for c in range(3):
            if flags[nd, c]:
                expect[nd, c] = q
                q += 1
    assert pr.numeq == numeq
    assert np.array_equal(pr.dof(), expect)
    dv = pr.dirichlet_values().reshape(-1, 3)
    assert np.all(dv[on4, 1] == 0.5)


# ------------------------------------------------------------------- loads
def test_point_loads_config1(oracle):
    pr = oracle.cantilever(60, 20, 20)
    f = pr.fext()
    dof = pr.dof()
    nz_idx = np.nonzero(f)[0]
    assert len(nz_idx) == 21
    assert abs(f.sum() + 1.0) < 1e-15
    for j in range(21):
        n = 60 + 61 * j
        expect = -1 / 40 if j in (0, 20) else -1 / 20
        assert f[dof[n, 2]] == expect  # bit-exact
        assert f[dof[n, 0]] == 0.0 and f[dof[n, 1]] == 0.0


def test_face_load_unit_faces(oracle):
    """Consistent nodal loads of a unit-voxel face load: val/4 per node per
    face (corners 1/4, edges 1/2, interior 1 per unit traction)."""
    nx, ny, nz = 3, 2, 2
    pr = oracle.Problem(nx, ny, nz)
    pr.add_dirichlet([1, 1, 1], [0, 0, 0], [[0, 0], [0, ny], [0, nz]])
    pr.make_dof()
    pr.add_neumann([0.0, 0.0, -2.0], [[nx, nx], [0, ny], [0, nz]])  # the x = nx face
    f = pr.fext()
    dof = pr.dof()
    assert abs(f.sum() + 2.0 * ny * nz) < 1e-13
    for j in range(ny + 1):
        for k in range(nz + 1):
            n = nx + (nx + 1) * (j + (ny + 1) * k)
            nfaces = (1 + (0 < j < ny)) * (1 + (0 < k < nz))
            assert abs(f[dof[n, 2]] - (-2.0) * nfaces / 4) < 1e-15
    # the other faces through the same element sides
    for side, rng, val in [("zmin", [[0, nx], [0, ny], [0, 0]], [1.0, 0, 0]),
                           ("ymax", [[0, nx], [ny, ny], [0, nz]], [0, 3.0, 0]),
                           ("zmax", [[0, nx], [0, ny], [nz, nz]], [0, 0, 1.0])]:
        p2 = oracle.Problem(nx, ny, nz)
        p2.make_dof()
        p2.add_neumann(val, rng)
        f2 = p2.fext()
        area = {"zmin": nx * ny, "ymax": nx * nz, "zmax": nx * ny}[side]
        assert abs(f2.sum() - sum(val) * area) < 1e-12, side


# --------------------------------------------------------------------- KE0
def _exact_h8_stiffness(nu):
    """Exact stiffness of the unit trilinear hexahedron from products of 1-D
    integrals (independent of any quadrature), at the exact value of the
    double nu, each entry correctly rounded to FP64."""
    nu = Fraction(nu)
    lam = nu / ((1 + nu) * (1 - 2 * nu))
    mu = Fraction(1, 2) / (1 + nu)
    corners = [(0, 0, 0), (1, 0, 0), (1, 1, 0), (0, 1, 0), (0, 0, 1), (1, 0, 1), (1, 1, 1), (0, 1, 1)]
    # 1-D integrals over [0,1] of phi_a^(da) phi_b^(db), phi_0 = 1 - t, phi_1 = t
    def i1(a, b, da, db):
        sa = -1 if a == 0 else 1
        sb = -1 if b == 0 else 1
        if da and db:
            return Fraction(sa * sb)
        if da:
            return Fraction(sa, 2)
        if db:
            return Fraction(sb, 2)
        return Fraction(1, 3) if a == b else Fraction(1, 6)

    def g(a, b, p, q):  # integral of dN_a/dx_p dN_b/dx_q
        v = Fraction(1)
        for d in range(3):
            v *= i1(corners[a][d], corners[b][d], d == p, d == q)
        return v

    K = np.zeros((24, 24))
    for a in range(8):
        for b in range(8):
            for i in range(3):
                for j in range(3):
                    v = lam * g(a, b, i, j) + mu * g(a, b, j, i)
                    if i == j:
                        v += mu * sum(g(a, b, k, k) for k in range(3))
                    K[3 * a + i, 3 * b + j] = float(v)
    return K


def test_ke0_equals_exact_integral(oracle):
    """KE0 is the FP64 rounding of the exact integral, entry by entry (the
    Gauss sums run in extended precision): equal exact entries are equal
    doubles, so the operator does not depend on a summation order."""
    for nu in (0.3, 0.25, 0.45):
        pr = oracle.Problem(2, 2, 2, poisson=nu)
        K = pr.ke0()
        Kx = _exact_h8_stiffness(nu)
        assert np.array_equal(K, Kx), np.max(np.abs(K - Kx))


def test_ke0_invariants(oracle):
    pr = oracle.Problem(2, 2, 2)
    K = pr.ke0()
    assert np.max(np.abs(K - K.T)) < 1e-16 * 10
    w = np.linalg.eigvalsh(K)
    assert np.sum(np.abs(w) < 1e-12) == 6      # six rigid modes
    assert w.min() > -1e-14                     # positive semi-definite
    X = np.array([[0, 0, 0], [1, 0, 0], [1, 1, 0], [0, 1, 0], [0, 0, 1], [1, 0, 1], [1, 1, 1], [0, 1, 1]], float)
    for d in range(3):                          # translations
        t = np.zeros(24)
        t[d::3] = 1.0
        assert np.max(np.abs(K @ t)) < 1e-15 * 10
    for (a, b) in [(0, 1), (1, 2), (0, 2)]:     # infinitesimal rotations
        r = np.zeros((8, 3))
        r[:, a] = -X[:, b]
        r[:, b] = X[:, a]
        assert np.max(np.abs(K @ r.ravel())) < 1e-14
    # reference KAT (test_bmatrix.cpp:123-146): unit hexa8 volume == 1
    _, vol = pr.element_ke(0, 1.0)
    assert abs(vol - 1.0) < 1e-10


def test_element_ke_translation_and_scaling(oracle):
    pr = oracle.Problem(60, 20, 20)
    K0 = pr.ke0()
    Ke, vol = pr.element_ke(57 + 60 * (13 + 20 * 9), 0.37)
    assert abs(vol - 1.0) < 1e-12
    assert np.max(np.abs(Ke - 0.37 * K0)) < 1e-14


# ----------------------------------------------------------- K, CG, c, dc
def test_csr_equals_elementwise(oracle):
    rng = np.random.default_rng(1)
    pr = oracle.cantilever(7, 4, 3)
    x = rng.uniform(1e-3, 1.0, pr.nelx)
    pr.set_density(x)
    u = rng.standard_normal(pr.numeq)
    y1 = pr.spmv(u)
    y2 = pr.apply_k_elementwise(u)
    assert np.max(np.abs(y1 - y2)) <= 1e-13 * np.max(np.abs(y1))
    # symmetry of the assembled operator
    v = rng.standard_normal(pr.numeq)
    assert abs(v @ pr.spmv(u) - u @ pr.spmv(v)) < 1e-12 * abs(v @ y1)


def test_nodal_apply_equals_csr_and_elementwise(oracle):
    """The node-by-node K.u (used where the CSR does not fit) against the
    assembled K and the element-by-element scatter, with nonzero prescribed
    displacements, box voxels and the matrix-free density setter."""
    rng = np.random.default_rng(11)
    dims, L = (6, 5, 4), (3.0, 7.5, 2.0)
    pr = oracle.Problem(*dims, *L)
    pr.add_dirichlet([1, 1, 1, 0, 0, 0], [0.0] * 6, [[0, 0], [0, L[1]], [0, L[2]]])
    pr.add_dirichlet([1, 0, 1, 0, 0, 0], [0.02, 0.0, -0.01, 0, 0, 0], [[L[0], L[0]], [0, L[1]], [0, L[2]]])
    pr.make_dof()
    pr.add_neumann([0.0, 1.0, 0.0], [[L[0], L[0]], [0, L[1]], [0, L[2]]])
    x = rng.uniform(1e-3, 1.0, pr.nelx)
    pr.set_density(x)
    u = rng.standard_normal(pr.numeq)
    dv = pr.dirichlet_values()
    y_csr = pr.spmv(u)
    y_el = pr.apply_k_elementwise(u, dv)
    y_el0 = pr.apply_k_elementwise(u)
    scale = np.max(np.abs(y_el))
    assert np.max(np.abs(pr.apply_k_nodal(u, dv) - y_el)) <= 1e-13 * scale
    assert np.max(np.abs(pr.apply_k_nodal(u) - y_csr)) <= 1e-13 * np.max(np.abs(y_csr))
    assert np.max(np.abs(y_el0 - y_csr)) <= 1e-13 * np.max(np.abs(y_csr))
    # prescribed displacements enter the nodal apply (the term K_fc u_c)
    assert np.max(np.abs(y_el - y_el0)) > 1e-6 * scale
    r_ref = pr.true_residual(u)
    q = oracle.Problem(*dims, *L)
    q.add_dirichlet([1, 1, 1, 0, 0, 0], [0.0] * 6, [[0, 0], [0, L[1]], [0, L[2]]])
    q.add_dirichlet([1, 0, 1, 0, 0, 0], [0.02, 0.0, -0.01, 0, 0, 0], [[L[0], L[0]], [0, L[1]], [0, L[2]]])
    q.make_dof()
    q.add_neumann([0.0, 1.0, 0.0], [[L[0], L[0]], [0, L[1]], [0, L[2]]])
    q.set_density_matrix_free(x)
    assert q.nnz() == 0
    assert np.array_equal(q.apply_k_nodal(u, dv), pr.apply_k_nodal(u, dv))
    assert abs(q.true_residual_nodal(u) - r_ref) <= 1e-12 * r_ref


def _dense_K(pr):
    n = pr.numeq
    K = np.zeros((n, n))
    for i in range(n):
        e = np.zeros(n)
        e[i] = 1.0
        K[:, i] = pr.spmv(e)
    return K


def test_scaling_cg_against_dense_solve(oracle):
    rng = np.random.default_rng(2)
    pr = oracle.cantilever(6, 3, 3)
    x = rng.uniform(0.05, 1.0, pr.nelx)
    pr.set_density(x)
    K = _dense_K(pr)
    f = pr.fext()
    ustar = np.linalg.solve(K, f)
    u, it, rn = pr.scaling_cg(tol=1e-12, maxit=10000)
    assert np.linalg.norm(u - ustar) < 1e-9 * np.linalg.norm(ustar)
    assert rn < 1e-12 * np.linalg.norm(f)
    assert np.allclose(pr.diag(), np.diag(K), rtol=0, atol=1e-15)
    # a capped solve stops after maxit + 1 iterations (CG.cpp:90-91)
    _, it2, _ = pr.scaling_cg(tol=1e-30, maxit=5)
    assert it2 == 6
    # true residual of the converged solution (linear.hpp:93-103)
    assert pr.true_residual(u) < 1e-11


def test_compliance_identity(oracle):
    rng = np.random.default_rng(3)
    pr = oracle.cantilever(8, 4, 4)
    pr.set_density(rng.uniform(0.1, 1.0, pr.nelx))
    u, _, _ = pr.scaling_cg(tol=1e-12)
    c = pr.compliance(u)
    assert abs(c - u @ pr.spmv(u)) < 1e-10 * c


def test_sensitivity_finite_differences(oracle):
    """The reference's own check (sample/algorithm/opt-mma-HelmholtzFilter/fdm.cpp):
    the analytic dc/dx_e against a difference quotient of re-solved compliances
    (central differences here for a tighter bound)."""
    rng = np.random.default_rng(4)
    pr = oracle.cantilever(5, 3, 2)
    x = rng.uniform(0.2, 1.0, pr.nelx)
    pr.set_density(x)
    u, _, _ = pr.scaling_cg(tol=1e-14)
    nv = pr.node_values(u)
    df, dg = pr.sensitivity(nv)
    assert np.all(np.abs(dg - 1.0) < 1e-14)  # element volume (sum of Gauss jac)

    def comp(xx):
        pr.set_density(xx)
        uu, _, _ = pr.scaling_cg(tol=1e-14)
        return pr.compliance(uu)

    h = 1e-5
    for e in rng.choice(pr.nelx, 6, replace=False):
        xp, xm = x.copy(), x.copy()
        xp[e] += h
        xm[e] -= h
        fd = (comp(xp) - comp(xm)) / (2 * h)
        assert abs(fd - df[e]) < 1e-6 * abs(df[e]), (e, fd, df[e])


# ---------------------------------------------------------------- filter
def test_filter_reference_kat(oracle):
    """test_filter.cpp:63-124 restated on a 2x2x1 hexa8 slab (the same four
    centre distances as the reference's 2x2 quad4 mesh), radius 2."""
    pr = oracle.Problem(2, 2, 1)
    pr.make_dof()
    F = oracle.Filter(pr, 2.0, brute=True)
    s = np.array([1e-4, 1.0, 1e-4, 1.0])
    answer = np.array([0.34587, 0.65423, 0.34587, 0.65423])
    sf = F.density_filter(s, False)
    assert np.sum((sf - answer) ** 2) < 1e-10
    # closed form of the same numbers
    w = np.array([2.0, 1.0, 1.0, 2.0 - np.sqrt(2.0)])
    exact0 = (w[0] * 1e-4 + w[1] * 1.0 + w[2] * 1e-4 + w[3] * 1.0) / w.sum()
    assert abs(sf[0] - exact0) < 1e-15
    dfds = np.array([1e-4, 1.0, 1e-4, 1.0])
    ds = F.density_filter(dfds, True)
    assert np.sum((ds - answer) ** 2) < 1e-10
    st = oracle.threshold(sf, 16.0, 0.5)
    assert st[0] < 0.01 and st[1] > 0.99
    d1 = oracle.threshold_deriv(sf, 1.0, 0.5)
    assert np.sum((d1 - 1.0) ** 2) < 0.015


def test_sensitivity_filter_reference_kat(oracle):
    """test_filter.cpp:63-89 (Filter::sensitivityFilter) restated on the 2x2x1
    slab: s = 0.5 everywhere, dfds = [1e-4, 1, 1e-4, 1], radius 2."""
    pr = oracle.Problem(2, 2, 1)
    pr.make_dof()
    F = oracle.Filter(pr, 2.0, brute=True)
    s = np.full(4, 0.5)
    dfds = np.array([1e-4, 1.0, 1e-4, 1.0])
    answer = np.array([0.34587, 0.65423, 0.34587, 0.65423])
    out = F.sensitivity_filter(s, dfds)
    assert np.sum((out - answer) ** 2) < 1e-10
    # closed form: with uniform densities the filter is the cone average of dfds
    w = np.array([2.0, 1.0, 1.0, 2.0 - np.sqrt(2.0)])
    exact0 = (w[0] * 1e-4 + w[1] * 1.0 + w[2] * 1e-4 + w[3] * 1.0) / w.sum()
    assert abs(out[0] - exact0) < 1e-15


@pytest.mark.parametrize("radius", [1.5, 2.5])
def test_sensitivity_filter_cancels_densities(oracle, radius):
    """filter.hpp:93-100 divides each term by x_j V and multiplies the sum by
    x_i V / (x_i sum w): the densities cancel, so for any positive x the result
    equals the cone average sum w dfds / sum w (the density filter of dfds) up
    to rounding.  A term with x_i for x_j (or a missing V) fails this."""
    pr = oracle.Problem(9, 6, 5, 18.0, 12.0, 10.0)
    pr.make_dof()
    F = oracle.Filter(pr, radius * 2.0)
    rng = np.random.default_rng(5)
    x = rng.uniform(1e-3, 1.0, pr.nelx)
    dfds = -rng.uniform(0.0, 3.0, pr.nelx)
    out = F.sensitivity_filter(x, dfds)
    avg = F.density_filter(dfds, False)
    assert np.max(np.abs(out - avg)) <= 1e-14 * np.max(np.abs(avg))
    assert not np.array_equal(out, dfds)


@pytest.mark.parametrize("dims,h", [((7, 5, 4), 1.0), ((6, 4, 5), 0.5)])
@pytest.mark.parametrize("radius", [1.2, 1.5, 2.0, 2.5])
def test_structured_neighbours_equal_bruteforce(oracle, dims, h, radius):
    nx, ny, nz = dims
    pr = oracle.Problem(nx, ny, nz, nx * h, ny * h, nz * h)
    pr.make_dof()
    rad = radius * h
    A = oracle.Filter(pr, rad, brute=True).lists()
    B = oracle.Filter(pr, rad, brute=False).lists()
    for a, b in zip(A, B):
        assert np.array_equal(a, b)


def test_filter_interior_neighbour_counts(oracle):
    pr = oracle.Problem(9, 9, 9)
    pr.make_dof()
    centre = 4 + 9 * (4 + 9 * 4)
    for r, cnt in [(1.2, 7), (1.5, 19), (2.0, 33), (2.5, 81)]:
        rp, ids, dis = oracle.Filter(pr, r).lists()
        assert rp[centre + 1] - rp[centre] == cnt


def test_filter_chain_rule_is_transpose(oracle):
    """densityFilter(., true) is the transpose of the linear map x -> x_tilde."""
    rng = np.random.default_rng(5)
    pr = oracle.Problem(6, 5, 4)
    pr.make_dof()
    F = oracle.Filter(pr, 1.5)
    n = pr.nelx
    M = np.zeros((n, n))
    for i in range(n):
        e = np.zeros(n)
        e[i] = 1
        M[:, i] = F.density_filter(e, False)
    v = rng.standard_normal(n)
    assert np.allclose(F.density_filter(v, True), M.T @ v, atol=1e-14)
    assert np.allclose(M.sum(axis=1), 1.0, atol=1e-14)  # rows are averages


# -------------------------------------------------------------- OC / MMA
def _toy(s):
    f = s[0] ** 2 + s[1] ** 2 + s[2] ** 2
    df = 2 * s
    g = np.array([(s[0] - 5) ** 2 + (s[1] - 2) ** 2 + (s[2] - 1) ** 2 - 9,
                  (s[0] - 3) ** 2 + (s[1] - 4) ** 2 + (s[2] - 3) ** 2 - 9])
    dg = np.array([2 * (s[0] - 5), 2 * (s[0] - 3), 2 * (s[1] - 2), 2 * (s[1] - 4),
                   2 * (s[2] - 1), 2 * (s[2] - 3)])
    return f, df, g, dg


def _run_toy(oracle, reference):
    m = oracle.MMA(3, 2, reference=reference)
    m.set(0.5, 0.1, 0.5, 0.7, 1.2)
    s = np.array([4.0, 3.0, 2.0])
    traj = []
    for _ in range(10):
        f, df, g, dg = _toy(s)
        m.update(s, df, g, dg, np.zeros(3), np.full(3, 5.0))
        traj.append(s.copy())
    f, df, g, dg = _toy(s)
    return f, g, np.array(traj)


def test_mma_reference_kat(oracle):
    """test_opt_solver.cpp:17-73: f < 8.78, |g_j| < 1e-6 after 10 updates."""
    f, g, _ = _run_toy(oracle, reference=False)
    assert f < 8.78
    assert np.all(np.abs(g) < 1e-6)


def test_mma_matches_compiled_reference(oracle):
    if oracle.ref_lib() is None:
        pytest.skip("oracle/_ref not built (reference tree absent)")
    _, _, t1 = _run_toy(oracle, reference=False)
    _, _, t2 = _run_toy(oracle, reference=True)
    assert np.max(np.abs(t1 - t2)) < 1e-10   # rounding order only (measured 6.9e-12)
    # a SIMP-like m = 1 problem
    rng = np.random.default_rng(6)
    n = 500
    x0 = rng.uniform(0.1, 0.9, n)
    xs = [x0.copy(), x0.copy()]
    mm = [oracle.MMA(n, 1, reference=r) for r in (False, True)]
    for it in range(6):
        df = -rng.uniform(0.1, 3.0, n)
        dg = np.full(n, 1.0 / n)
        gval = np.array([xs[0].mean() - 0.4])
        for k in range(2):
            mm[k].update(xs[k], df, gval, dg, np.full(n, 1e-4), np.ones(n))
        assert np.max(np.abs(xs[0] - xs[1])) < 1e-11


def test_oc_matches_compiled_reference(oracle):
    if oracle.ref_lib() is None:
        pytest.skip("oracle/_ref not built (reference tree absent)")
    rng = np.random.default_rng(7)
    pr = oracle.Problem(8, 6, 5)
    pr.make_dof()
    F = oracle.Filter(pr, 1.5)
    n = pr.nelx
    x = rng.uniform(0.2, 0.6, n)
    dc = -rng.uniform(0.01, 2.0, n)
    dv = np.ones(n)
    g = 0.35 * n
    x1, cnt, lam = F.oc_update(x, dc, g, dv)

    def vol(ptr, nn, user):
        arr = np.ctypeslib.as_array(ptr, shape=(nn,)).copy()
        return float(F.density_filter(arr, False).sum())

    cb = oracle.REF_VOLUME_FN(vol)
    x2 = x.copy()
    oracle.ref_lib().ref_oc_update(n, oracle.P(x2), oracle.P(dc.copy()), g, oracle.P(dv.copy()),
                                   oracle.P(np.full(n, 1e-4)), oracle.P(np.ones(n)), cb, None)
    assert np.max(np.abs(x1 - x2)) < 1e-14
    assert 10 < cnt < 100
    # volume constraint active at the converged multiplier
    assert abs(F.density_filter(x1, False).sum() - g) < 1e-3 * g


def test_gaussian_field_recipe(oracle):
    x, m = oracle.gaussian_field(24, 12, 10, seed=2)
    assert x.min() >= 1e-3 and x.max() <= 1.0
    assert abs(m - x.mean()) < 1e-15
    x2, _ = oracle.gaussian_field(24, 12, 10, seed=2)
    assert np.array_equal(x, x2)
    x3, _ = oracle.gaussian_field(24, 12, 10, seed=3)
    assert not np.array_equal(x, x3)


# ---- committed fixtures from the compiled reference optimizers (tests/golden) ----
GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def test_mma_toy_matches_golden(oracle):
    d = np.load(os.path.join(GOLDEN, "mma_toy.npz"))
    _, _, t = _run_toy(oracle, reference=False)
    assert np.max(np.abs(t - d["traj"])) < 1e-12


def test_mma_m1_matches_golden(oracle):
    d = np.load(os.path.join(GOLDEN, "mma_m1.npz"))
    n = d["x0"].size
    x = d["x0"].copy()
    m = oracle.MMA(n, 1)
    for k in range(d["x"].shape[0]):
        m.update(x, d["dfdx"][k], np.array([d["g"][k]]), d["dgdx"][k], np.full(n, 1e-4), np.ones(n))
        assert np.max(np.abs(x - d["x"][k])) < 1e-11, k


def test_oc_matches_golden(oracle):
    d = np.load(os.path.join(GOLDEN, "oc_filter.npz"))
    pr = oracle.Problem(*[int(v) for v in d["dims"]])
    pr.make_dof()
    F = oracle.Filter(pr, float(d["radius"]))
    x1, cnt, _ = F.oc_update(d["x"], d["dfdx"], float(d["g"]), d["dgdx"])
    assert np.max(np.abs(x1 - d["x_new"])) < 1e-14


# ------------------------------------------------------------ Mises stress
def _d0(nu):
    lam = nu / ((1 + nu) * (1 - 2 * nu))
    mu = 1 / (2 * (1 + nu))
    D = np.full((3, 3), lam) + 2 * mu * np.eye(3)
    return np.block([[D, np.zeros((3, 3))], [np.zeros((3, 3)), mu * np.eye(3)]])


def _vm(s):
    return np.sqrt(s[:, 0] ** 2 + s[:, 1] ** 2 + s[:, 2] ** 2 - s[:, 0] * s[:, 1] -
                   s[:, 1] * s[:, 2] - s[:, 2] * s[:, 0] +
                   3 * (s[:, 3] ** 2 + s[:, 4] ** 2 + s[:, 5] ** 2))


def test_mises_affine_field_closed_form(oracle):
    """problem/elasticity.cpp:166-209: for u = G X the trilinear interpolation is
    exact, the strain is constant eps = [Gxx, Gyy, Gzz, Gxy+Gyx, Gyz+Gzy, Gzx+Gxz],
    stress = E(x) D0 eps and mises = sqrt(s^T V s)."""
    pr = oracle.Problem(4, 3, 2, 2.0, 4.5, 1.0, young1=1e-3, young2=2.0, poisson=0.3,
                        penalty=3.0)
    pr.make_dof()
    rng = np.random.default_rng(5)
    x = rng.uniform(0.05, 1.0, pr.nelx)
    pr.set_density(x)
    G = rng.standard_normal((3, 3))
    X = pr.coords()
    nodeval = (X @ G.T).reshape(-1)
    m, st = pr.mises(nodeval)
    eps = np.array([G[0, 0], G[1, 1], G[2, 2], G[0, 1] + G[1, 0], G[1, 2] + G[2, 1],
                    G[2, 0] + G[0, 2]])
    E = (1 - x ** 3) * 1e-3 + x ** 3 * 2.0
    ref = E[:, None] * (_d0(0.3) @ eps)[None, :]
    assert np.max(np.abs(st - ref)) < 1e-13 * np.max(np.abs(ref))
    assert np.max(np.abs(m - _vm(ref)) / _vm(ref)) < 1e-13


def test_mises_rigid_motion_is_stress_free(oracle):
    pr = oracle.cantilever(5, 3, 3)
    pr.set_density(np.full(pr.nelx, 0.7))
    X = pr.coords()
    w = np.array([0.3, -0.2, 0.5])          # infinitesimal rotation + translation
    u = np.cross(w, X) + np.array([1.0, -2.0, 0.5])
    m, st = pr.mises(u.reshape(-1))
    assert np.max(m) < 1e-13 and np.max(np.abs(st)) < 1e-13


def test_mises_gauss_mean_is_centroid_value(oracle):
    """The mean of B over the 2x2x2 points equals B at the element centre
    (B is multilinear), so the stress of a random field is E D0 B(0) u_e."""
    nx, ny, nz, hx, hy, hz = 3, 2, 2, 0.5, 1.0, 2.0
    pr = oracle.Problem(nx, ny, nz, nx * hx, ny * hy, nz * hz)
    pr.make_dof()
    x = np.full(pr.nelx, 0.6)
    pr.set_density(x)
    u = np.random.default_rng(6).standard_normal(3 * pr.numnp)
    m, st = pr.mises(u)
    conn = pr.connectivity()
    sgn = np.array([[-1, -1, -1], [1, -1, -1], [1, 1, -1], [-1, 1, -1],
                    [-1, -1, 1], [1, -1, 1], [1, 1, 1], [-1, 1, 1]], dtype=float)
    dN = sgn / (4 * np.array([hx, hy, hz]))  # dN_a/dX at the centre
    E = (1 - 0.6 ** 3) * 1e-9 + 0.6 ** 3 * 1.0
    for e in range(pr.nelx):
        ue = u.reshape(-1, 3)[conn[e]]
        Gr = ue.T @ dN                       # du_i/dX_j
        eps = np.array([Gr[0, 0], Gr[1, 1], Gr[2, 2], Gr[0, 1] + Gr[1, 0],
                        Gr[1, 2] + Gr[2, 1], Gr[2, 0] + Gr[0, 2]])
        ref = E * (_d0(0.3) @ eps)
        assert np.max(np.abs(st[e] - ref)) < 1e-13 * np.max(np.abs(ref))


# ------------------------------------------------------------------ GCMMA
def _gcmma_toy(G):
    """sample/test/test-optimization/test_opt_solver.cpp:75-150, step for step
    (including its use of the zero-initialised new design's objective in the
    first conservativeness check)."""
    n = 3
    G.set(0.5, 0.1, 0.5, 0.7, 1.2)
    s = np.array([4.0, 3.0, 2.0])
    smin, smax = np.zeros(n), np.full(n, 5.0)

    def cons(v):
        return np.array([(v[0] - 5) ** 2 + (v[1] - 2) ** 2 + (v[2] - 1) ** 2 - 9.0,
                         (v[0] - 3) ** 2 + (v[1] - 4) ** 2 + (v[2] - 3) ** 2 - 9.0])

    traj = []
    for _ in range(10):
        obj = float(np.sum(s ** 2))
        dobj = 2 * s
        g = cons(s)
        dg = np.array([2 * (s[0] - 5), 2 * (s[0] - 3), 2 * (s[1] - 2), 2 * (s[1] - 4),
                       2 * (s[2] - 1), 2 * (s[2] - 3)])
        objnew = 0.0
        snew = G.outer(s, obj, dobj, g, dg, smin, smax)
        gnew = cons(snew)
        ok = G.concheck(objnew, gnew)
        k = 0
        while not ok and k < 15:
            snew = G.inner(snew, objnew, gnew, s, obj, dobj, g, dg, smin, smax)
            objnew = float(np.sum(snew ** 2))
            gnew = cons(snew)
            ok = G.concheck(objnew, gnew)
            k += 1
        s = snew
        traj.append(s.copy())
    return obj, g, np.array(traj)


def test_gcmma_reference_kat(oracle):
    obj, g, _ = _gcmma_toy(oracle.GCMMA(3, 2))
    assert obj < 8.78
    assert abs(g[0]) < 1e-6 and abs(g[1]) < 1e-6


def test_gcmma_matches_compiled_reference(oracle):
    """The restatement against GCMMASolver.cpp compiled from the reference
    (oracle/_ref): the KAT trajectory, then seeded SIMP-like problems with one
    constraint through outer and inner updates."""
    if oracle.ref_lib() is None:
        pytest.skip("reference tree absent: oracle/_ref not built")
    _, _, t1 = _gcmma_toy(oracle.GCMMA(3, 2))
    _, _, t2 = _gcmma_toy(oracle.GCMMA(3, 2, reference=True))
    assert np.max(np.abs(t1 - t2)) < 1e-10   # rounding order only (measured 6.9e-12)
    rng = np.random.default_rng(12)
    n = 200
    G1, G2 = oracle.GCMMA(n, 1), oracle.GCMMA(n, 1, reference=True)
    x = rng.uniform(0.2, 0.5, n)
    lo, hi = np.full(n, 1e-4), np.ones(n)
    for it in range(6):
        f0 = 10.0 + rng.uniform()
        df0 = -rng.uniform(0.1, 2.0, n)
        g = np.array([x.sum() - 0.3 * n])
        dg = np.ones(n)
        xa = G1.outer(x, f0, df0, g, dg, lo, hi)
        xb = G2.outer(x, f0, df0, g, dg, lo, hi)
        assert np.max(np.abs(xa - xb)) < 1e-12
        for _ in range(2):      # non-conservative new values force inner updates
            fn, gn = f0 + 1.0, np.array([xa.sum() - 0.3 * n + 0.5])
            assert G1.concheck(fn, gn) == G2.concheck(fn, gn)
            xa = G1.inner(xa, fn, gn, x, f0, df0, g, dg, lo, hi)
            xb = G2.inner(xb, fn, gn, x, f0, df0, g, dg, lo, hi)
            assert np.max(np.abs(xa - xb)) < 1e-12
        x = xa
