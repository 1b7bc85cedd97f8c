"""GPU parity of the Helmholtz PDE filter (reference optimization/filter_helm.hpp)
and of the SIMP loop that uses it, through the C-ABI, against the oracle."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("dims,L,r", [
    ((12, 6, 5), None, 1.5),
    ((16, 8, 8), None, 2.5),
    ((9, 7, 4), (4.5, 14.0, 2.0), 1.2),   # box voxels 0.5 x 2 x 0.5
    ((20, 10, 6), None, 4.0),
    ((2, 2, 2), None, 2.0),
    ((4, 90, 90), None, 1.5),             # long y and z axes, short x
])
def test_helm_filter_matches_oracle(gpu, oracle, dims, L, r):
    """Both modes against the oracle's assembled system solved by Jacobi-PCG to
    1e-14: the fast-diagonalisation solve is exact to FP64 rounding."""
    kw = {} if L is None else dict(lx=L[0], ly=L[1], lz=L[2])
    c = gpu.Context(*dims, **kw)
    h = (1.0, 1.0, 1.0) if L is None else tuple(L[d] / dims[d] for d in range(3))
    H = oracle.HelmFilter(dims, r, h=h, tol=1e-14)
    c.helm_setup(r)
    rng = np.random.default_rng(sum(dims))
    for sens in (False, True):
        x = rng.uniform(1e-3, 1.0, c.nelx) if not sens else -rng.uniform(0.0, 5.0, c.nelx)
        g = c.helm_filter(x, sens)
        o = H.density_filter(x, sens)
        assert np.max(np.abs(g - o)) < 1e-12 * np.max(np.abs(o)), (sens, np.max(np.abs(g - o)))


def test_helm_filter_invariants_config2(gpu):
    """192x64x64 (the C2 grid): constants preserved, volume preserved, the
    operator symmetric, both modes equal (size-independent properties)."""
    c = gpu.Context(192, 64, 64)
    c.helm_setup(1.5)
    n = c.nelx
    assert np.max(np.abs(c.helm_filter(np.full(n, 0.3)) - 0.3)) < 1e-12
    x, _ = gpu.gaussian_field(192, 64, 64, 2)
    fx = c.helm_filter(x)
    assert abs(fx.sum() - x.sum()) < 1e-11 * x.sum()
    assert np.max(np.abs(c.helm_filter(x, True) - fx)) < 1e-12
    y = np.random.default_rng(7).standard_normal(n)
    a, b = c.helm_filter(x) @ y, x @ c.helm_filter(y)
    assert abs(a - b) < 1e-11 * (np.abs(x) @ np.abs(y))


def test_helm_filter_errors(gpu):
    c = gpu.Context(4, 4, 4)
    with pytest.raises(gpu.IcaratError):
        c.helm_filter(np.ones(c.nelx))      # before helm_setup
    with pytest.raises(gpu.IcaratError):
        c.helm_setup(0.0)


@pytest.mark.parametrize("optimizer,proj", [(0, False), (1, True), (0, True)])
def test_simp_helmholtz_matches_oracle(gpu, oracle, optimizer, proj):
    """The SIMP loop with the Helmholtz filter (opt_mma_Helmholtz.cpp:84-150:
    filter -> projection -> solve -> sensitivity -> chain rule -> MMA; and the
    OC variant whose bisection re-filters every candidate) against the oracle."""
    dims = (24, 8, 8)
    c = gpu.cantilever(*dims)
    o = oracle.cantilever(*dims)
    c.helm_setup(2.0)
    H = oracle.HelmFilter(dims, 2.0, tol=1e-14)
    x0 = np.full(c.nelx, 0.3)
    n_iter = 6
    xg, hg, _ = c.simp_run(x0, n_iter, optimizer=optimizer, use_projection=proj, tol=1e-11,
                           helmholtz=True)
    xo, ho, _ = H.simp_run(o, x0, n_iter, optimizer=optimizer, use_projection=proj,
                           cg_tol=1e-11)
    assert np.max(np.abs(xg - xo)) < 1e-6
    assert np.max(np.abs(hg[:, 0] - ho[:, 0]) / ho[:, 0]) < 1e-7
    assert np.max(np.abs(hg[:, 1] - ho[:, 1]) / ho[:, 1]) < 1e-9
    assert hg[-1, 0] < hg[0, 0]

