"""GPU parity of the GCMMA design update (reference optimization/opt-solver/
GCMMASolver.cpp) and of the GCMMA driver's loop (opt-macro-gcmma/
opt_macro_gcmma.cpp:84-212) against the oracle, through the C-ABI."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def test_gcmma_steps_match_oracle(gpu, oracle):
    """Outer updates, forced inner updates (non-conservative new values) and
    the conservativeness checks on the same inputs."""
    c = gpu.cantilever(12, 6, 4)
    n = c.nelx
    G = oracle.GCMMA(n, 1)
    rng = np.random.default_rng(21)
    x = rng.uniform(0.2, 0.5, n)
    lo, hi = np.full(n, 1e-4), np.ones(n)
    for it in range(6):
        f0 = 10.0 + rng.uniform()
        df0 = -rng.uniform(0.1, 2.0, n)
        g = x.sum() - 0.3 * n
        dg = np.ones(n)
        xa = c.gcmma_outer(x, f0, df0, g, dg)
        xb = G.outer(x, f0, df0, np.array([g]), dg, lo, hi)
        assert np.max(np.abs(xa - xb)) < 1e-10, it
        for k in range(2):
            fn, gn = f0 + 1.0 - 0.4 * k, xa.sum() - 0.3 * n + 0.5
            assert c.gcmma_concheck(fn, gn) == G.concheck(fn, np.array([gn]))
            xa = c.gcmma_inner(xa, fn, gn, x, f0, df0, g, dg)
            xb = G.inner(xb, fn, np.array([gn]), x, f0, df0, np.array([g]), dg, lo, hi)
            assert np.max(np.abs(xa - xb)) < 1e-10, (it, k)
        x = xa


@pytest.mark.parametrize("proj", [True, False])
def test_gcmma_loop_matches_oracle(gpu, oracle, proj):
    dims = (24, 8, 8)
    c = gpu.cantilever(*dims)
    o = oracle.cantilever(*dims)
    c.filter_setup(1.5)
    F = oracle.Filter(o, 1.5)
    x0 = np.full(c.nelx, 0.3)
    xg, hg, sg = c.simp_run(x0, 6, optimizer=2, use_projection=proj, tol=1e-11)
    xo, ho, so = oracle.simp_run_gcmma(o, F, x0, 6, use_projection=proj, cg_tol=1e-11)
    assert np.max(np.abs(xg - xo)) < 1e-6
    assert np.max(np.abs(hg[:, 0] - ho[:, 0]) / ho[:, 0]) < 1e-7
    assert np.max(np.abs(hg[:, 1] - ho[:, 1]) / ho[:, 1]) < 1e-9
    assert hg[-1, 0] < hg[0, 0]


def test_gcmma_not_in_batch(gpu):
    c = gpu.cantilever(8, 4, 4)
    c.filter_setup(1.5)
    b = gpu.Batch([c])
    with pytest.raises(gpu.IcaratError):
        b.simp_run([np.full(c.nelx, 0.3)], 1, optimizer=2)
