"""Geometric multigrid preconditioner: hierarchy shape, solves that reach the
FP64 verified residual and agree with Jacobi-PCG, and the design loop against
the FP64 Jacobi loop.  Tolerances: every solve <= tol (verified stop);
solutions 1e-6 of max |u| (both at 1e-9); designs 1e-6, compliance 1e-8."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def test_hierarchy_levels(gpu):
    c = gpu.cantilever(48, 16, 16)
    c.mg_setup(dense_dofs=0)
    assert c.mg_info() == [(48, 16, 16), (24, 8, 8), (12, 4, 4), (6, 2, 2)]
    # coarsening stops at the first grid with at most dense_dofs lattice DOFs:
    # 12x4x4 has 3 * 13 * 5 * 5 = 975
    c.mg_setup()
    assert c.mg_info() == [(48, 16, 16), (24, 8, 8), (12, 4, 4)]
    c.mg_setup(dense_dofs=600)
    assert c.mg_info() == [(48, 16, 16), (24, 8, 8), (12, 4, 4), (6, 2, 2)]
    c = gpu.cantilever(60, 20, 20)
    c.mg_setup(max_levels=2)
    assert c.mg_info() == [(60, 20, 20), (30, 10, 10)]


@pytest.mark.parametrize("dims", [(8, 4, 4), (6, 4, 2)])
def test_dense_inverse_is_exact(gpu, dims):
    """One level solved by the dense inverse: the preconditioner is K^-1 of the
    assembled free-DOF system, so PCG stops after one or two iterations.  This
    pins the assembly (element scatter, constrained DOFs) to the matrix-free K."""
    c = gpu.cantilever(*dims)
    x, _ = gpu.gaussian_field(*dims, seed=3, smin=1.0, smax=3.0)
    c.set_density(x)
    c.mg_setup(max_levels=1)
    assert c.mg_info() == [dims]
    uj, sj = c.solve(tol=1e-12)
    um, sm = c.solve(tol=1e-12, precond=gpu.PRECOND_MG)
    print(f"{dims}: Jacobi {sj.iters} iterations, exact {sm.iters}")
    assert sm.status == 0 and sm.iters <= 2 and sm.true_relres <= 1e-12
    assert np.max(np.abs(um - uj)) < 1e-9 * np.max(np.abs(uj))


# (40, 20, 20) coarsens to 10 x 5 x 5, which cannot be halved: its 1188 lattice
# DOFs exceed the default dense limit (1024) but not 1.5 times it, so that
# blocked coarsest grid is solved by its dense inverse
@pytest.mark.parametrize("dims", [(48, 16, 16), (32, 24, 16), (40, 20, 20)])
def test_mg_solve_agrees_with_jacobi(gpu, dims):
    c = gpu.cantilever(*dims)
    x, _ = gpu.gaussian_field(*dims, seed=13, smin=2.0, smax=6.0)
    c.set_density(x)
    uj, sj = c.solve(tol=1e-9)
    um, sm = c.solve(tol=1e-9, precond=gpu.PRECOND_MG)
    print(f"{dims}: Jacobi {sj.iters} iterations {sj.seconds * 1e3:.1f} ms; multigrid "
          f"{sm.iters} iterations {sm.seconds * 1e3:.1f} ms, relres {sm.true_relres:.2e}, "
          f"levels {c.mg_info()}")
    assert sm.status == 0 and sm.true_relres <= 1e-9
    assert np.max(np.abs(um - uj)) < 1e-6 * np.max(np.abs(uj))
    assert sm.iters < sj.iters
    ms = c.mg_status()
    assert ms.levels == len(c.mg_info()) and ms.factorizations >= 1 and ms.shifted == 0
    if dims == (40, 20, 20):
        # the blocked 10x5x5 coarsest grid (1188 DOFs > 1024) is solved densely
        assert c.mg_info() == [(40, 20, 20), (20, 10, 10), (10, 5, 5)]
        assert ms.dense == 1
        # a new density: its inverse is factored again before the solve (one
        # factorization more), and the solve still matches
        x2 = np.clip(x * (1.0 + 0.5 * np.cos(np.arange(c.nelx) * 0.01)), 1e-3, 1.0)
        c.set_density(x2)
        uj2, sj2 = c.solve(tol=1e-9)
        um2, sm2 = c.solve(tol=1e-9, precond=gpu.PRECOND_MG)
        ms2 = c.mg_status()
        print(f"second density: Jacobi {sj2.iters}, multigrid {sm2.iters} iterations, "
              f"factorizations {ms2.factorizations}")
        assert sm2.status == 0 and sm2.true_relres <= 1e-9
        assert ms2.factorizations == ms.factorizations + 1 and ms2.shifted == 0
        assert np.max(np.abs(um2 - uj2)) < 1e-6 * np.max(np.abs(uj2))


@pytest.mark.parametrize("dims", [(48, 16, 16), (60, 20, 20)])
def test_mg_kernel_paths_agree(gpu, dims):
    """The smoothing kernels (z-marching K.u + vector kernels, or the fused
    one-thread-per-node sweep) run the same V-cycle: either gives the same
    iterations (+-2, summation order only) and the same solution (1e-9 of max
    |u|, all solves at 1e-10).  With the exact coarsest solve (dense inverse:
    6x2x2 or 12x4x4 for (48, 16, 16)) the iterations do not grow."""
    c = gpu.cantilever(*dims)
    x, _ = gpu.gaussian_field(*dims, seed=5, smin=2.0, smax=6.0)
    c.set_density(x)
    res = {}
    for dense, direct in ((0, 0), (0, 1 << 30), (600, 0), (600, 1 << 30), (1024, 1 << 30)):
        c.mg_setup(dense_dofs=dense, direct_nodes=direct)
        u, st = c.solve(tol=1e-10, precond=gpu.PRECOND_MG)
        assert st.status == 0 and st.true_relres <= 1e-10
        res[(dense, direct)] = (u, st.iters)
        print(f"{dims} dense {dense} direct {direct}: {st.iters} iterations, "
              f"{st.seconds * 1e3:.1f} ms")
    u0, it0 = res[(0, 0)]
    for k, (u, it) in res.items():
        if k[0] == 0:
            assert abs(it - it0) <= 2, k
        else:
            assert it <= it0 + 2, k
        assert np.max(np.abs(u - u0)) < 1e-9 * np.max(np.abs(u0)), k


@pytest.mark.parametrize("shape", [(0, 1), (0, 2), (1, 2)])
def test_mg_nonsymmetric_cycles(gpu, shape):
    """V(nu_pre, nu_post) cycles without pre-smoothing or with unequal sweeps
    are not symmetric; the flexible PCG still reaches the verified residual and
    the Jacobi-PCG solution (1e-6 of max |u|, both solves at 1e-9)."""
    dims = (48, 16, 16)
    c = gpu.cantilever(*dims)
    x, _ = gpu.gaussian_field(*dims, seed=13, smin=2.0, smax=6.0)
    c.set_density(x)
    uj, _ = c.solve(tol=1e-9)
    c.mg_setup(nu_pre=shape[0], nu_post=shape[1])
    um, sm = c.solve(tol=1e-9, precond=gpu.PRECOND_MG)
    print(f"V{shape}: {sm.iters} iterations")
    assert sm.status == 0 and sm.true_relres <= 1e-9
    assert np.max(np.abs(um - uj)) < 1e-6 * np.max(np.abs(uj))


@pytest.mark.parametrize("avg", [1, 2, 3, 4, 5])
def test_mg_coarse_averaging_rules(gpu, avg):
    """Coarse Young's moduli by the geometric mean, the SIMP law of the mean
    density, the quadratic mean or the maximum of the children (the default is
    the arithmetic mean), or Galerkin operators R A P (5): other coarse
    operators, the same verified solve (1e-6 of max |u| of Jacobi-PCG, both at
    1e-9)."""
    dims = (48, 16, 16)
    c = gpu.cantilever(*dims)
    x, _ = gpu.gaussian_field(*dims, seed=13, smin=2.0, smax=6.0)
    c.set_density(x)
    uj, _ = c.solve(tol=1e-9)
    c.mg_setup(coarse_avg=avg)
    um, sm = c.solve(tol=1e-9, precond=gpu.PRECOND_MG)
    print(f"coarse_avg {avg}: {sm.iters} iterations")
    assert sm.status == 0 and sm.true_relres <= 1e-9
    assert np.max(np.abs(um - uj)) < 1e-6 * np.max(np.abs(uj))


def test_mg_simp_matches_jacobi(gpu):
    dims = (60, 20, 20)
    x0 = np.full(24000, 0.3)
    out = []
    for pre in (gpu.PRECOND_JACOBI, gpu.PRECOND_MG):
        c = gpu.cantilever(*dims)
        c.filter_setup(1.5)
        out.append(c.simp_run(x0, 6, tol=1e-10, precond=pre))
    (xj, hj, sj), (xm, hm, sm) = out
    print(f"PCG iterations over 6 design iterations: Jacobi {sj.pcg_iters}, multigrid "
          f"{sm.pcg_iters}; solve time {sj.solve_s:.3f} s vs {sm.solve_s:.3f} s")
    assert np.max(np.abs(xm - xj)) < 1e-6
    assert np.max(np.abs(hm[:, 0] - hj[:, 0]) / hj[:, 0]) < 1e-8
    assert np.all(hm[:, 3] <= 1e-10)


def _prolong(v, dims):
    """Trilinear interpolation of a coarse nodal field (node-major, 3 comps) to
    the 2x finer lattice."""
    nx, ny, nz = dims
    vc = v.reshape(nz // 2 + 1, ny // 2 + 1, nx // 2 + 1, 3)
    out = np.zeros((nz + 1, ny + 1, nx + 1, 3))
    for k in range(nz + 1):
        for j in range(ny + 1):
            for i in range(nx + 1):
                acc = 0.0
                for (K, wk) in ([(k // 2, 1.0)] if k % 2 == 0 else [(k // 2, .5), (k // 2 + 1, .5)]):
                    for (J, wj) in ([(j // 2, 1.0)] if j % 2 == 0 else [(j // 2, .5), (j // 2 + 1, .5)]):
                        for (I, wi) in ([(i // 2, 1.0)] if i % 2 == 0 else [(i // 2, .5), (i // 2 + 1, .5)]):
                            acc = acc + wk * wj * wi * vc[K, J, I]
                out[k, j, i] = acc
    return out.reshape(-1, 3)


@pytest.mark.parametrize("avg,exact", [(5, True), (0, False)])
def test_galerkin_two_grid_is_exact_on_coarse_solutions(gpu, avg, exact):
    """Two levels, no pre-smoothing, the coarse level solved exactly: with the
    Galerkin operator R A P the coarse correction reproduces any solution in
    the range of P, so PCG stops after one iteration; with a rediscretised
    coarse operator (arithmetic mean of E) on a high-contrast field it does not.
    This pins the element-by-element Galerkin matrices (weights, masks,
    assembly) against the matrix-free fine operator and the transfers."""
    dims = (8, 4, 4)
    c = gpu.cantilever(*dims)
    rng = np.random.default_rng(31)
    c.set_density(rng.uniform(1e-3, 1.0, c.nelx) ** 0.5)
    c.mg_setup(max_levels=2, nu_pre=0, nu_post=1, dense_dofs=4096, coarse_avg=avg)
    assert c.mg_info() == [(8, 4, 4), (4, 2, 2)]
    cn = (dims[0] // 2 + 1) * (dims[1] // 2 + 1) * (dims[2] // 2 + 1)
    vc = rng.standard_normal((cn, 3))
    vc[::dims[0] // 2 + 1] = 0.0            # coarse nodes on the clamped face x = 0
    uf = _prolong(vc.reshape(-1), dims)
    dof = c.dof().reshape(-1, 3)
    ustar = np.zeros(c.numeq)
    free = dof < c.numeq
    ustar[dof[free]] = uf[free]
    c.set_fext(c.apply_K(ustar))
    u, st = c.solve(tol=1e-10, precond=gpu.PRECOND_MG)
    assert st.status == 0 and st.true_relres <= 1e-10
    assert np.max(np.abs(u - ustar)) < 1e-8 * np.max(np.abs(ustar))
    if exact:
        assert st.iters <= 1, st.iters
    else:
        assert st.iters > 1


@pytest.mark.parametrize("dims,direct", [((48, 16, 16), 1 << 30), ((60, 20, 20), 1 << 30),
                                         ((64, 32, 32), 4096), ((30, 10, 14), 1 << 30)])
def test_fused_correction_sweep_matches_separate_kernels(gpu, monkeypatch, dims, direct):
    """On the one-thread-per-node levels the coarse correction x + P e and the
    first post-sweep run as one kernel.  It forms the same values in the same
    order as the prolongation and sweep kernels it replaces, so the solve takes
    the same iterations and the solutions agree to rounding (1e-12 of max |u|).
    Grids include tiles cut by the lattice edge in every axis (30 x 10 x 14:
    nodes 31 x 11 x 15 against 16 x 4 x 4 tiles) and a hierarchy with both
    kernel shapes (direct_nodes 4096)."""
    c = gpu.cantilever(*dims)
    x, _ = gpu.gaussian_field(*dims, seed=21, smin=2.0, smax=6.0)
    c.set_density(x)
    res = {}
    for fuse in ("0", "1"):
        monkeypatch.setenv("ICARAT_HIP_MG_FUSE", fuse)
        c.mg_setup(direct_nodes=direct)
        u, st = c.solve(tol=1e-10, precond=gpu.PRECOND_MG)
        assert st.status == 0 and st.true_relres <= 1e-10
        res[fuse] = (u, st.iters)
        print(f"{dims} fuse {fuse}: {st.iters} iterations, {st.seconds * 1e3:.2f} ms, "
              f"levels {c.mg_info()}")
    (u0, it0), (u1, it1) = res["0"], res["1"]
    assert it0 == it1
    assert np.max(np.abs(u1 - u0)) <= 1e-12 * np.max(np.abs(u0))


def test_multigrid_loop_repeats_bit_for_bit(gpu):
    """Two fresh contexts run the same multigrid SIMP loop (192x64x64, seeded
    field, 6 design iterations, each density's coarsest inverse refactored):
    designs and histories equal bit for bit, also after other contexts have
    run in the process.  (The overlapped refactorization, ICARAT_HIP_MG_OVERLAP=1,
    does not meet this and is off by default.)"""
    dims = (192, 64, 64)
    x0, _ = gpu.gaussian_field(*dims, 2)
    out = []
    for _ in range(2):
        c = gpu.cantilever(*dims)
        c.filter_setup(1.5)
        out.append(c.simp_run(x0, 6, tol=1e-8, maxit=100000, precond=gpu.PRECOND_MG))
        del c
    (x1, h1, _), (x2, h2, _) = out
    assert np.array_equal(h1, h2)
    assert np.array_equal(x1, x2)


def test_w_cycle_solves_and_switches(gpu):
    """The W-cycle at the top of the hierarchy (coarse levels 1..w cycled twice;
    to the end by default, or with w_switch the V-cycle from w_switch x tol on):
    the verified residual, the Jacobi solution to 1e-6 of max |u|, fewer
    iterations than the V-cycle; the automatic depth by mesh size."""
    dims = (96, 48, 48)
    c = gpu.cantilever(*dims)
    x, _ = gpu.gaussian_field(*dims, 6)
    c.set_density(x)
    uj, sj = c.solve(tol=1e-9)
    its = {}
    for w in (0, 1, 2):
        c.mg_setup(w_levels=w)
        um, sm = c.solve(tol=1e-9, precond=gpu.PRECOND_MG)
        assert c.mg_status().w_levels == w
        print(f"w_levels {w}: {sm.iters} iterations, relres {sm.true_relres:.2e}")
        assert sm.status == 0 and sm.true_relres <= 1e-9
        assert np.max(np.abs(um - uj)) < 1e-6 * np.max(np.abs(uj))
        its[w] = sm.iters
    assert its[1] < its[0] and its[2] <= its[1]
    # the switch to the V-cycle near the tolerance keeps the verified stop
    c.mg_setup(w_levels=2, w_switch=30.0)
    um, sm = c.solve(tol=1e-9, precond=gpu.PRECOND_MG)
    print(f"w_levels 2, w_switch 30: {sm.iters} iterations, relres {sm.true_relres:.2e}")
    assert sm.status == 0 and sm.true_relres <= 1e-9
    assert np.max(np.abs(um - uj)) < 1e-6 * np.max(np.abs(uj))
    # automatic: none below 500,000 elements, one level up to 2,000,000
    c.mg_setup()
    c.solve(tol=1e-9, precond=gpu.PRECOND_MG)
    assert c.mg_status().w_levels == 0
    big = gpu.cantilever(192, 64, 64)
    big.set_density(np.full(big.nelx, 0.5))
    big.solve(tol=1e-6, precond=gpu.PRECOND_MG)
    assert big.mg_status().w_levels == 1


def test_w_cycle_batched_equals_standalone(gpu):
    from icarat2_amd import workloads as W
    specs = [W.batch_problem(s) for s in (108, 113)]
    alone, ctxs = [], []
    for s in specs:
        c = W.build_context(gpu, s)
        c.mg_setup(w_levels=2)
        c.set_density(W.initial_design(s))
        u, st = c.solve(tol=1e-9, precond=gpu.PRECOND_MG)
        alone.append((c.fext() @ u, st.iters))
        c2 = W.build_context(gpu, s)
        c2.mg_setup(w_levels=2)
        c2.set_density(W.initial_design(s))
        ctxs.append(c2)
    st = gpu.Batch(ctxs).solve(tol=1e-9, precond=gpu.PRECOND_MG)
    for c, (c0, i0), s1 in zip(ctxs, alone, st):
        c1, _, _ = c.compliance_sensitivity()
        assert s1.iters == i0 and s1.true_relres <= 1e-9
        assert abs(c1 - c0) <= 1e-12 * abs(c0)


def test_w_cycle_breakdown_goes_on_with_v_cycle(gpu):
    """160x80x80 with its 4 face-load cases (batched workload seed 122): the
    W-cycle on the rediscretised first coarse level is not always positive
    definite there (z.r <= 0 in some solves); those solves go on with the
    V-cycle from their iterate and still pass the verified 1e-8."""
    from icarat2_amd import workloads as W
    s = W.batch_problem(122)
    c = W.build_context(gpu, s)
    x, h, st = c.simp_run(W.initial_design(s), 12, tol=1e-8, maxit=100000,
                          precond=gpu.PRECOND_MG)
    m = c.mg_status()
    print(f"w_levels {m.w_levels}, W phases ended by a breakdown {m.w_fallbacks}, "
          f"iterations {h[:, 2]}, relres {h[:, 3]}")
    assert m.w_levels == 1
    assert m.w_fallbacks >= 1  # the recovery ran (the loop is deterministic)
    assert np.all(h[:, 3] <= 1e-8)
