"""FP64 K.u on the matrix cores (v_mfma_f64_4x4x4 element contraction,
icarat_hip_apply_K with ICARAT_HIP_FP64_MFMA) against the FP64 VALU K.u and
the oracle's assembled K (reference analysis/linear.hpp:117-166).  Different
summation order, so the bar is FP64 rounding: 1e-13 of max |K u|."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

FP64_MFMA = 3


@pytest.mark.parametrize("dims", [(12, 6, 5), (40, 20, 10), (70, 9, 17)])
def test_mfma_apply_matches_valu_and_oracle(gpu, oracle, dims):
    nx, ny, nz = dims
    c = gpu.cantilever(*dims)
    rng = np.random.default_rng(sum(dims))
    x = rng.uniform(1e-3, 1.0, c.nelx)
    c.set_density(x)
    u = rng.standard_normal(c.numeq)
    y0 = c.apply_K(u, gpu.FP64)
    y3 = c.apply_K(u, FP64_MFMA)
    scale = np.max(np.abs(y0))
    err = np.max(np.abs(y3 - y0)) / scale
    print(f"{dims}: max |mfma - valu| / max |y| = {err:.2e}")
    assert err < 1e-13
    o = oracle.cantilever(*dims)
    o.set_density(x)
    assert np.max(np.abs(o.spmv(u) - y3)) / scale < 1e-13


def test_mfma_apply_full_size_c2(gpu):
    dims = (192, 64, 64)
    c = gpu.cantilever(*dims)
    x, _ = gpu.gaussian_field(*dims, 2)
    c.set_density(x)
    u = np.random.default_rng(2).standard_normal(c.numeq)
    y0 = c.apply_K(u, gpu.FP64)
    y3 = c.apply_K(u, FP64_MFMA)
    err = np.max(np.abs(y3 - y0)) / np.max(np.abs(y0))
    print(f"C2: max |mfma - valu| / max |y| = {err:.2e}")
    assert err < 1e-13
    # K is linear: K(2u) = 2 K(u) exactly (a power-of-two scale)
    assert np.array_equal(c.apply_K(2.0 * u, FP64_MFMA), 2.0 * y3)
