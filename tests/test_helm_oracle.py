"""Pins the oracle's Helmholtz PDE filter (reference optimization/filter_helm.hpp)
against the reference's known-answer test, closed forms and invariants (no GPU)."""
import itertools

import numpy as np
import pytest


# ------------------------------------------------------- reference KAT (2-D)
def test_reference_kat_quad4(oracle):
    """sample/test/test-optimization/test_filter.cpp:126-175: the 2x2 quad4 mesh,
    r = 2, s = [1e-4, 1, 1e-4, 1]; both modes give the listed answer
    (sum of squared differences < 1e-10), projection beta 16 is nearly a step,
    beta 1 nearly linear."""
    H = oracle.HelmFilter((2, 2), 2.0, dim=2, tol=1e-10)
    s = np.array([1e-4, 1.0, 1e-4, 1.0])
    answer = np.array([0.312569, 0.687531, 0.312569, 0.687531])
    s_new = H.density_filter(s, False)
    deriv = H.density_filter(s, True)
    assert np.sum((s_new - answer) ** 2) < 1e-10
    assert np.sum((deriv - answer) ** 2) < 1e-10
    thre = oracle.threshold(s_new, 16.0, 0.5)
    assert thre[0] < 0.01 and thre[1] > 0.99
    der = oracle.threshold_deriv(s_new, 1.0, 0.5)
    assert np.sum((der - 1.0) ** 2) < 0.015


# ------------------------------------------------- element matrix (closed form)
def _closed_form_ke(h, r, dim=3):
    """M_e + c L_e of a box element from 1-D linear-element integrals:
    mass m = h [[1/3, 1/6], [1/6, 1/3]], stiffness k = [[1, -1], [-1, 1]] / h,
    M_e = m_x (x) m_y (x) m_z, L_e = k_x m_y m_z + m_x k_y m_z + m_x m_y k_z."""
    c = (r / (2 * np.sqrt(3.0))) ** 2
    # local node -> (ix, iy, iz) corner, bottom face counter-clockwise then top
    corners = [(0, 0, 0), (1, 0, 0), (1, 1, 0), (0, 1, 0),
               (0, 0, 1), (1, 0, 1), (1, 1, 1), (0, 1, 1)][: 4 if dim == 2 else 8]
    m = [hd * np.array([[1 / 3, 1 / 6], [1 / 6, 1 / 3]]) for hd in h]
    k = [np.array([[1.0, -1.0], [-1.0, 1.0]]) / hd for hd in h]
    n = len(corners)
    K = np.zeros((n, n))
    for a, b in itertools.product(range(n), range(n)):
        ca, cb = corners[a], corners[b]
        mass = np.prod([m[d][ca[d], cb[d]] for d in range(dim)])
        lap = 0.0
        for d in range(dim):
            t = k[d][ca[d], cb[d]]
            for d2 in range(dim):
                if d2 != d:
                    t *= m[d2][ca[d2], cb[d2]]
            lap += t
        K[a, b] = mass + c * lap
    return K


@pytest.mark.parametrize("h,r", [((1.0, 1.0, 1.0), 1.5), ((1.0, 0.5, 2.0), 2.5),
                                 ((0.25, 0.25, 0.25), 0.6)])
def test_element_matrix_closed_form(oracle, h, r):
    H = oracle.HelmFilter((3, 2, 2), r, h=h)
    for e in (0, 5, H.nelx - 1):
        ke, _ = H.element(e)
        ref = _closed_form_ke(h, r)
        assert np.max(np.abs(ke - ref)) < 1e-14 * np.max(np.abs(ref))


def test_element_matrix_closed_form_quad4(oracle):
    H = oracle.HelmFilter((2, 3), 1.7, h=(0.5, 2.0, 1.0), dim=2)
    ke, _ = H.element(4)
    ref = _closed_form_ke((0.5, 2.0), 1.7, dim=2)
    assert np.max(np.abs(ke - ref)) < 1e-14 * np.max(np.abs(ref))


# ------------------------------------------------------ assembled system / solve
def _dense(H):
    n = H.numnp
    A = np.zeros((n, n))
    for e in range(H.nelx):
        ke, conn = H.element(e)
        A[np.ix_(conn, conn)] += ke
    return A


def test_apply_equals_assembled_matrix(oracle):
    H = oracle.HelmFilter((4, 3, 2), 1.5)
    A = _dense(H)
    v = np.random.default_rng(1).standard_normal(H.numnp)
    assert np.max(np.abs(H.apply(v) - A @ v)) < 1e-13 * np.max(np.abs(A @ v))
    # no boundary condition: K 1 = M 1 = nodal volume shares (sum = domain volume)
    assert abs(np.sum(A @ np.ones(H.numnp)) - H.nelx) < 1e-12 * H.nelx


def test_filter_equals_dense_solve(oracle):
    """out = (1/8) S A^-1 G x with A the assembled matrix, G_ae = int_e N_a = V/8,
    S_ea = 1 for the nodes of e (Gauss sums of N_a are 1 at every node)."""
    dims, h, r = (5, 3, 4), (1.0, 0.5, 2.0), 2.0
    H = oracle.HelmFilter(dims, r, h=h)
    A = _dense(H)
    S = np.zeros((H.nelx, H.numnp))
    for e in range(H.nelx):
        S[e, H.element(e)[1]] = 1.0
    V = h[0] * h[1] * h[2]
    x = np.random.default_rng(2).uniform(1e-3, 1.0, H.nelx)
    ref = S @ np.linalg.solve(A, (V / 8) * (S.T @ x)) / 8
    out = H.density_filter(x)
    assert np.max(np.abs(out - ref)) < 1e-12
    # sensitivity mode: fe / 8 and no final / 8 -- the same symmetric operator
    outs = H.density_filter(x, True)
    assert np.max(np.abs(outs - ref)) < 1e-12


def test_invariants(oracle):
    H = oracle.HelmFilter((6, 4, 3), 1.5)
    rng = np.random.default_rng(3)
    # constants are preserved (the Laplacian part annihilates them)
    assert np.max(np.abs(H.density_filter(np.full(H.nelx, 0.37)) - 0.37)) < 1e-12
    assert np.max(np.abs(H.density_filter(np.full(H.nelx, 0.37), True) - 0.37)) < 1e-12
    # volume is preserved: sum out = sum in (natural boundary condition)
    x = rng.uniform(0.0, 1.0, H.nelx)
    assert abs(H.density_filter(x).sum() - x.sum()) < 1e-11 * x.sum()
    # the operator is symmetric: <W x, y> = <x, W y>
    y = rng.standard_normal(H.nelx)
    assert abs(H.density_filter(x) @ y - x @ H.density_filter(y)) < 1e-11 * np.abs(x) @ np.abs(y)
    # a larger radius smooths more (total variation along x decreases)
    step = np.where(np.arange(H.nelx) % 6 < 3, 1e-3, 1.0)
    tv = [np.abs(np.diff(oracle.HelmFilter((6, 4, 3), rr).density_filter(step))).sum()
          for rr in (0.5, 1.5, 3.0)]
    assert tv[0] > tv[1] > tv[2]


def test_cg_tolerance_semantics(oracle):
    """Jacobi-PCG from zero stops at ||r|| <= tol ||b||: a loose tolerance is
    farther from the tight solution than a strict one."""
    H = oracle.HelmFilter((8, 4, 4), 2.0)
    x = np.random.default_rng(4).uniform(1e-3, 1.0, H.nelx)
    exact = H.density_filter(x, tol=1e-14)
    loose = H.density_filter(x, tol=1e-3)
    it_loose = H.iters
    tight = H.density_filter(x, tol=1e-10)
    assert H.iters > it_loose
    assert np.max(np.abs(tight - exact)) < np.max(np.abs(loose - exact))
    assert np.max(np.abs(tight - exact)) < 1e-8
