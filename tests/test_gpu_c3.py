"""The 256x128x128 mixed-precision configuration at full size (4,194,304
elements, 12,780,288 free DOFs; Gaussian-field density seed 3, rmin 1.5): the
K.u contraction of every precision against the oracle's node-by-node apply,
the FP64 multigrid solve and the solve with the FP32 fine level judged by the
oracle's own recomputation of the true residual (analysis/linear.hpp:93-103),
and the FP64-oracle error of u, compliance and sensitivities per precision.

The tolerance is 2e-8: this field leaves 47% of the filtered densities below
0.01, |u| reaches 3.9e7 with |f| = 1, and rounding u to FP64 alone leaves a
relative residual of 1.4-1.9e-8 on this design (a random half-ulp perturbation
adds 1.8e-8 in quadrature), so no FP64 iterate verifies at 1e-8; 2e-8 is the
tightest tolerance the multigrid solve reaches on every design of the loop.
"""
import numpy as np
import pytest

pytestmark = [pytest.mark.gpu, pytest.mark.slow]

DIMS = (256, 128, 128)
TOL = 2e-8


@pytest.fixture(scope="module")
def c3(gpu, oracle):
    c = gpu.cantilever(*DIMS)
    x0, _ = gpu.gaussian_field(*DIMS, 3)
    c.filter_setup(1.5)
    xf = c.density_filter(x0)
    c.set_density(xf)
    o = oracle.cantilever(*DIMS)
    o.set_density_matrix_free(xf)
    assert o.numeq == c.numeq == 12780288
    return c, o


def _rel(a, b):
    return np.max(np.abs(a - b)) / np.max(np.abs(b))


def test_c3_contraction_per_precision(gpu, c3):
    c, o = c3
    u = np.random.default_rng(30).standard_normal(c.numeq)
    yo = o.apply_k_nodal(u)
    for prec, bound in ((gpu.FP64, 1e-13), (gpu.FP32, 1e-5), (gpu.BF16, 3e-2)):
        err = _rel(c.apply_K(u, prec), yo)
        print(f"C3 K.u precision {prec}: max error / max |Ku| = {err:.3e}")
        assert err < bound


def test_c3_solves_against_oracle(gpu, c3):
    c, o = c3
    out = {}
    for prec in (gpu.FP64, gpu.FP32):
        u, st = c.solve(tol=TOL, maxit=40000, precision=prec, precond=gpu.PRECOND_MG)
        r_or = o.true_residual_nodal(u)
        cc, df, _ = c.compliance_sensitivity()
        c_or = o.compliance(u)
        df_or, _ = o.sensitivity(o.node_values(u))
        print(f"C3 multigrid, fine level precision {prec}: {st.iters} iterations, "
              f"{st.refine_steps} verifications, {st.seconds:.2f} s; true residual GPU "
              f"{st.true_relres:.4e} oracle {r_or:.4e}")
        assert st.status == 0 and st.true_relres <= TOL
        # The oracle evaluates the same u with another summation order (element
        # by element) and its own rounding.  At this field the residual sits at
        # the FP64 floor of u, where the two evaluations differ by a few
        # percent, and the solve stops when its own evaluation first passes, so
        # the oracle's reads 2-6% higher (measured over runs): bounded at 10%.
        assert r_or <= 1.1 * TOL
        assert abs(r_or - st.true_relres) < 0.1 * r_or
        # compliance and sensitivities of this u, GPU vs oracle
        assert abs(cc - c_or) < 1e-12 * abs(c_or)
        assert _rel(df, df_or) < 1e-12
        out[prec] = (u, c_or, df_or)
    # FP64-oracle error per precision: the oracle's compliance and
    # sensitivities of each solution against those of the FP64 solution
    u64, c64, df64 = out[gpu.FP64]
    for prec, (u, cp, dfp) in out.items():
        eu = np.linalg.norm(u - u64) / np.linalg.norm(u64)
        ec = abs(cp - c64) / abs(c64)
        ed = _rel(dfp, df64)
        print(f"C3 precision {prec}: |u - u64|/|u64| {eu:.3e}, |c - c64|/c64 {ec:.3e}, "
              f"max|dc - dc64|/max|dc64| {ed:.3e}")
        assert eu < 1e-7 and ec < 1e-8 and ed < 1e-6


def test_c3_compensated_update_reaches_tolerance(gpu, c3):
    """The first design of this field with the default hierarchy (16x8x8
    solved exactly at this size): with x += alpha p in plain FP64 the updates
    of the near-void regions' large displacements (|u| up to 3.9e7) fall below
    half an ulp near the tolerance and are lost while the recursive residual
    counts them, so the verified residual stalled near 8e-8 (1,457 iterations
    at a 3e-8 tolerance); accumulated as a compensated sum they are kept and
    the solve passes."""
    c, o = c3
    c.mg_setup()
    assert c.mg_info()[-1] == (16, 8, 8)
    res = {}
    for comp in (0, 1):
        c.mg_setup(compensated=comp)
        _, st = c.solve(tol=TOL, maxit=40000, precond=gpu.PRECOND_MG, want_u=False,
                        allow_cap=True)
        res[comp] = st
        print(f"C3 design 1, compensated={comp}: status {st.status}, {st.iters} iterations, "
              f"{st.refine_steps} verifications, true residual {st.true_relres:.3e}")
    assert res[1].status == 0 and res[1].true_relres <= TOL
    assert res[1].iters <= res[0].iters
    c.mg_setup()
