"""Mixed precision: the FP32 / BF16 MFMA element contraction against the FP64
oracle, and FP64 iterative refinement around it, judged by the FP64 true
residual.  Tolerances:
  K.u with FP32 operands ........ 1e-5 of max |y| (FP32 rounding of U, KE0, young)
  K.u with BF16 operands ........ 3e-2 of max |y| (8-bit mantissa operands)
  reliable-update solve ......... FP64 true residual <= tol, compliance 1e-8
                                  relative to the FP64 solve (both at 1e-10)
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _rel(a, b):
    return np.max(np.abs(a - b)) / np.max(np.abs(b))


def _smooth_u(c, rng):
    """A smooth nodal field (low-order polynomial) restricted to free DOFs."""
    dof = c.dof()
    NX, NY = c.nx + 1, c.ny + 1
    n = np.arange(c.numnp)
    x, y, z = n % NX, (n // NX) % NY, n // (NX * NY)
    u = np.zeros(c.numeq)
    for d in range(3):
        a = rng.standard_normal(4)
        v = a[0] * x + a[1] * y * 0.5 + a[2] * z + a[3] * x * z * 0.1
        m = dof[:, d] < c.numeq
        u[dof[m, d]] = v[m]
    return u


@pytest.mark.parametrize("dims", [(60, 20, 20), (33, 9, 17), (7, 5, 3)])
def test_apply_k_fp32_bf16(gpu, oracle, dims):
    c = gpu.cantilever(*dims)
    o = oracle.cantilever(*dims)
    rng = np.random.default_rng(31)
    x = rng.uniform(1e-3, 1.0, c.nelx)
    c.set_density(x)
    o.set_density(x)
    for u in (rng.standard_normal(c.numeq), _smooth_u(c, rng)):
        y = o.spmv(u)
        e32 = _rel(c.apply_K(u, gpu.FP32), y)
        e16 = _rel(c.apply_K(u, gpu.BF16), y)
        print(f"dims={dims} fp32 err={e32:.2e} bf16 err={e16:.2e}")
        assert e32 < 1e-5
        assert e16 < 3e-2


@pytest.mark.parametrize("prec", ["FP32", "BF16"])
def test_refined_solve_reaches_fp64_residual(gpu, oracle, prec):
    c = gpu.cantilever(48, 16, 16)
    x, _ = gpu.gaussian_field(48, 16, 16, seed=7, smin=2.0, smax=6.0)
    c.set_density(x)
    u64, s64 = c.solve(tol=1e-10)
    um, sm = c.solve(tol=1e-10, precision=getattr(gpu, prec), maxit=40000)
    print(f"{prec}: iters={sm.iters} refine={sm.refine_steps} fallback={sm.fallback} "
          f"relres={sm.true_relres:.2e}  fp64 iters={s64.iters}")
    assert sm.status == 0
    assert sm.true_relres < 1e-10
    o = oracle.cantilever(48, 16, 16)
    o.set_density(x)
    assert o.true_residual(um) < 1e-10  # the oracle agrees on the FP64 residual
    f = c.fext()
    assert abs(f @ um - f @ u64) < 1e-8 * abs(f @ u64)


def test_simp_loop_fp32_matches_fp64(gpu):
    c = gpu.cantilever(60, 20, 20)
    c.filter_setup(1.5)
    x0 = np.full(c.nelx, 0.3)
    x64, h64, _ = c.simp_run(x0, 10, tol=1e-8)
    x32, h32, s32 = c.simp_run(x0, 10, tol=1e-8, precision=gpu.FP32)
    assert np.all(h32[:, 3] < 1e-8)
    assert np.max(np.abs(x32 - x64)) < 1e-5
    assert np.max(np.abs(h32[:, 0] - h64[:, 0]) / h64[:, 0]) < 1e-6
