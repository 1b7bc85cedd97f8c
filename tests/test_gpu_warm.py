"""Warm starts: a solve may start from its load case's last solution instead of
0.  The stop test is the same (FP64 verified residual <= tol), so warm and
cold solves agree to the solve tolerance (1e-6 of max |u| at 1e-9), and a warm
solve after a small density change needs fewer iterations.  Design loops with
warm starts follow the cold ones (designs 1e-6, compliance 1e-8), and the
batch engine's warm starts equal the standalone ones bit for bit."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("mg", [False, True])
def test_warm_solve_after_density_change(gpu, mg):
    pre = gpu.PRECOND_MG if mg else gpu.PRECOND_JACOBI
    dims = (48, 16, 16)
    c = gpu.cantilever(*dims)
    x, _ = gpu.gaussian_field(*dims, seed=13, smin=2.0, smax=6.0)
    c.set_density(x)
    c.solve(tol=1e-9, precond=pre)
    x2 = np.clip(x * (1.0 + 0.05 * np.sin(np.arange(x.size))), 1e-3, 1.0)
    c.set_density(x2)
    uw, sw = c.solve(tol=1e-9, precond=pre, warm=True)
    uc, sc = c.solve(tol=1e-9, precond=pre)
    print(f"{'multigrid' if mg else 'Jacobi'}: cold {sc.iters} iterations, warm {sw.iters}")
    assert sw.status == 0 and sw.true_relres <= 1e-9 and sc.true_relres <= 1e-9
    assert sw.iters < sc.iters
    assert np.max(np.abs(uw - uc)) < 1e-6 * np.max(np.abs(uc))


def test_warm_solve_of_the_same_system_is_immediate(gpu):
    c = gpu.cantilever(24, 8, 8)
    c.set_density(np.full(c.nelx, 0.4))
    u0, s0 = c.solve(tol=1e-9, precond=gpu.PRECOND_MG)
    u1, s1 = c.solve(tol=1e-9, precond=gpu.PRECOND_MG, warm=True)
    assert s1.status == 0 and s1.iters <= 1 and s1.true_relres <= 1e-9
    assert np.max(np.abs(u1 - u0)) < 1e-9 * np.max(np.abs(u0))


@pytest.mark.parametrize("mg", [False, True])
def test_warm_simp_follows_cold(gpu, mg):
    pre = gpu.PRECOND_MG if mg else gpu.PRECOND_JACOBI
    dims = (48, 16, 16)
    x0 = np.full(48 * 16 * 16, 0.3)
    out = []
    for warm in (False, True):
        c = gpu.cantilever(*dims)
        c.filter_setup(1.5)
        out.append(c.simp_run(x0, 6, tol=1e-10, precond=pre, warm=warm))
    (xc, hc, sc), (xw, hw, sw) = out
    print(f"{'multigrid' if mg else 'Jacobi'}: PCG iterations cold {sc.pcg_iters}, warm {sw.pcg_iters}")
    assert sw.pcg_iters < sc.pcg_iters
    assert np.max(np.abs(xw - xc)) < 1e-6
    assert np.max(np.abs(hw[:, 0] - hc[:, 0]) / hc[:, 0]) < 1e-8
    assert np.all(hw[:, 3] <= 1e-10)


def test_batch_warm_equals_standalone_warm(gpu):
    from icarat2_amd import workloads as W
    specs = [W.batch_problem(s) for s in (108, 113)]
    single = []
    for s in specs:
        c = W.build_context(gpu, s)
        single.append(c.simp_run(W.initial_design(s), 3, tol=1e-9, precond=gpu.PRECOND_MG,
                                 warm=True))
    ctxs = [W.build_context(gpu, s) for s in specs]
    b = gpu.Batch(ctxs)
    xs, hists, sts = b.simp_run([W.initial_design(s) for s in specs], 3, tol=1e-9,
                                precond=gpu.PRECOND_MG, warm=True)
    for (x1, h1, s1), x2, h2, s2 in zip(single, xs, hists, sts):
        print(f"iters alone {s1.pcg_iters} batched {s2.pcg_iters}")
        assert s1.pcg_iters == s2.pcg_iters
        assert np.max(np.abs(x1 - x2)) < 1e-12
        assert np.allclose(h1, h2, rtol=1e-12, atol=0)
