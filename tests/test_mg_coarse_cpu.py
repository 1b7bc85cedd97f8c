"""Conditioning of the multigrid hierarchy's coarsest matrix, on the CPU.

The multigrid preconditioner solves its coarsest grid (12x4x4 voxels of 16^3
for a 192x64x64 mesh) by a dense Cholesky inverse; the coarse Young's moduli
are the arithmetic means of the 8 children, level by level.  Here the oracle
assembles that coarsest matrix from fine densities and its spectrum is
measured: for a stiff block floating in near-void material (a rigid motion
held only by Emin = 1e-9) the smallest eigenvalue falls to about 1e-12 of the
largest, within a few times the n eps backward error of a Cholesky
factorization in FP64 (n = 900: 2e-13).  A factorization can therefore break
down for such designs; the relative diagonal shift of 1e-10 that the
preconditioner then applies lifts the ratio to about 20 times that bound.
The benchmark field at 192x64x64 is far from that limit (1.5e-7)."""
import numpy as np


def _coarse_young(E, dims, times):
    nx, ny, nz = dims
    E = E.reshape(nz, ny, nx)
    for _ in range(times):
        E = 0.125 * sum(E[a::2, b::2, c::2] for a in (0, 1) for b in (0, 1) for c in (0, 1))
    return E.ravel()


def _coarsest_matrix(oracle, dims, x, times):
    E = 1e-9 + x ** 3 * (1.0 - 1e-9)  # SIMP, young = [1e-9, 1], p = 3
    Ec = _coarse_young(E, dims, times)
    cd = tuple(d >> times for d in dims)
    h = 2 ** times
    pr = oracle.Problem(*cd, cd[0] * h, cd[1] * h, cd[2] * h)
    pr.add_dirichlet([1, 1, 1, 0, 0, 0], [0.0] * 6, [[0, 0], [0, cd[1] * h], [0, cd[2] * h]])
    pr.make_dof()
    # the densities whose SIMP young is the coarse young
    pr.set_density(np.clip((Ec - 1e-9) / (1.0 - 1e-9), 0.0, None) ** (1.0 / 3.0))
    n = pr.numeq
    K = np.zeros((n, n))
    for i in range(n):
        e = np.zeros(n)
        e[i] = 1.0
        K[:, i] = pr.spmv(e)
    return K


def test_coarsest_matrix_conditioning(oracle):
    dims = (192, 64, 64)
    x_bench, _ = oracle.gaussian_field(*dims, 2)
    block = np.full(dims[::-1], 1e-3)
    block[20:40, 20:40, 150:180] = 1.0  # a stiff block away from the clamp
    ratios = {}
    for label, x in (("bench", x_bench), ("floating block", block.ravel())):
        K = _coarsest_matrix(oracle, dims, x, 4)
        n = K.shape[0]
        w = np.linalg.eigvalsh(K)
        ratios[label] = w[0] / w[-1]
        print(f"{label}: n {n}, lambda [{w[0]:.3e}, {w[-1]:.3e}], ratio {ratios[label]:.2e}, "
              f"n eps {n * np.finfo(float).eps:.1e}")
        assert w[0] > 0.0
        if label == "floating block":
            assert ratios[label] < 10 * n * np.finfo(float).eps
            # the shifted matrix of the retry: A + 1e-10 diag(A)
            Ks = K + 1e-10 * np.diag(np.diag(K))
            ws = np.linalg.eigvalsh(Ks)
            print(f"shifted: ratio {ws[0] / ws[-1]:.2e}")
            assert ws[0] / ws[-1] > 10 * n * np.finfo(float).eps
            np.linalg.cholesky(Ks)
    assert ratios["bench"] > 1e-8
