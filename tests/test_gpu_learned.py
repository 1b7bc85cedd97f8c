"""The learned preconditioner: z = M(r) against the numpy restatement
(oracle/unet.py precondition), flexible PCG with a zero network against
Jacobi-PCG, online training, the ablation (same solve with and without the
network) and the design loop against the FP64 Jacobi loop.  The network is used only
while its relative squared error over the stored samples is below the gate
(0.8); a network that fails the gate leaves flexible PCG at z = D^-1 r.
Tolerances:
  M(r) with a zero network ...... D^-1 r to 1e-15 relative (no network term)
  M(r) network correction ....... 5e-2 normwise (BF16 network vs FP64 oracle)
  every solve ................... FP64 true residual <= tol (verified stop)
  design loop ................... designs within 1e-6, compliance 1e-8 relative
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _nodal(c, v, fill=0.0):
    """numeq vector -> (NZ, NY, NX, 3) nodal array (constrained DOFs = fill)."""
    dof = c.dof()
    out = np.full((c.numnp, 3), fill)
    m = dof < c.numeq
    out[m] = v[dof[m]]
    return out.reshape(c.nz + 1, c.ny + 1, c.nx + 1, 3)


def _problem(gpu, dims, seed):
    c = gpu.cantilever(*dims)
    x, _ = gpu.gaussian_field(*dims, seed=seed, smin=2.0, smax=6.0)
    c.set_density(x)
    return c, x


def test_apply_zero_network_is_jacobi(gpu):
    c, _ = _problem(gpu, (20, 12, 8), 3)
    c.learned_setup(seed=5)
    c.learned_set_params(np.zeros(gpu.unet_param_count(), dtype=np.float32))
    r = np.random.default_rng(4).standard_normal(c.numeq)
    z = c.learned_apply(r)
    ref = r / c.diag()
    assert np.max(np.abs(z - ref) / np.abs(ref)) < 1e-15


def test_apply_matches_oracle(gpu, oracle):
    import unet as U
    dims = (22, 12, 10)
    c, x = _problem(gpu, dims, 5)
    c.learned_setup(seed=7)
    info = c.learned_info()
    assert info["grid"] == U.cell_grid(*dims)
    flat = c.learned_get_params()
    rng = np.random.default_rng(8)
    # random biases: the default initialisation has zero biases
    off = 0
    for name, ci, co, taps in U.LAYERS:
        off += taps * co * ci
        flat[off:off + co] = rng.standard_normal(co) * 0.05
        off += co
    c.learned_set_params(flat)
    P = {}
    off = 0
    for name, ci, co, taps in U.LAYERS:
        P[name + ".w"] = flat[off:off + taps * co * ci].astype(np.float64).reshape(taps, co, ci)
        off += taps * co * ci
        P[name + ".b"] = flat[off:off + co].astype(np.float64)
        off += co
    r = rng.standard_normal(c.numeq)
    d = c.diag()
    z = c.learned_apply(r)
    free = _nodal(c, np.ones(c.numeq)) > 0
    rn = _nodal(c, r)
    dinv = _nodal(c, 1.0 / d)
    young = ((1.0 - x ** 3) * 1e-9 + x ** 3 * 1.0).reshape(dims[2], dims[1], dims[0])
    zo, corr = U.precondition(P, rn, dinv, free, young)
    zg = _nodal(c, z)
    e = float(np.linalg.norm((zg - dinv * rn) - corr) / np.linalg.norm(corr))
    print(f"network correction normwise error {e:.2e}, |corr|/|D^-1 r| = "
          f"{np.linalg.norm(corr) / np.linalg.norm(dinv * rn):.3f}")
    assert e < 5e-2


def test_flexible_pcg_zero_network_matches_jacobi(gpu):
    c, _ = _problem(gpu, (32, 16, 16), 9)
    u0, s0 = c.solve(tol=1e-8)
    c.learned_setup(seed=5)
    c.learned_set_params(np.zeros(gpu.unet_param_count(), dtype=np.float32))
    u1, s1 = c.solve(tol=1e-8, precond=gpu.PRECOND_LEARNED)
    print(f"Jacobi {s0.iters} iterations, flexible PCG with zero network {s1.iters} "
          f"(fallback {s1.fallback}), relres {s1.true_relres:.2e}")
    assert s1.status == 0 and s1.fallback == 0
    assert s1.true_relres <= 1e-8
    assert abs(s1.iters - s0.iters) <= 0.05 * s0.iters + 16
    assert np.max(np.abs(u1 - u0)) < 1e-6 * np.max(np.abs(u0))


def test_training_and_ablation(gpu):
    """Harvest Jacobi solves of nearby designs, train, then solve a new design
    with and without the network (the ablation); both must converge."""
    dims = (48, 16, 16)
    c = gpu.cantilever(*dims)
    c.learned_setup(seed=3, max_samples=64, snapshot_every=32)
    base, _ = gpu.gaussian_field(*dims, seed=21, smin=2.0, smax=6.0)
    rng = np.random.default_rng(22)
    for k in range(4):
        x = np.clip(base * (1.0 + 0.1 * rng.standard_normal(base.size)), 1e-3, 1.0)
        c.set_density(x)
        _, st = c.learned_harvest(tol=1e-8)
        assert st.status == 0
    assert c.learned_info()["samples"] > 4
    loss = c.learned_train(300)
    info = c.learned_info()
    print(f"samples {info['samples']}, gain {info['gain']:.3g}, relative squared error: last "
          f"steps {loss:.3f}, all samples {info['val_loss']:.3f}, network active {info['active']}")
    assert loss >= 0.0 and info["val_loss"] >= 0.0 and info["steps"] == 300
    assert info["active"] == (info["val_loss"] < 0.8)
    x = np.clip(base * (1.0 + 0.1 * rng.standard_normal(base.size)), 1e-3, 1.0)
    c.set_density(x)
    uj, sj = c.solve(tol=1e-8)
    ul, sl = c.solve(tol=1e-8, precond=gpu.PRECOND_LEARNED)
    print(f"ablation: Jacobi {sj.iters} iterations {sj.seconds * 1e3:.1f} ms; learned "
          f"{sl.iters} iterations {sl.seconds * 1e3:.1f} ms (fallback {sl.fallback}, "
          f"relres {sl.true_relres:.2e})")
    assert sl.status == 0 and sl.true_relres <= 1e-8
    assert np.max(np.abs(ul - uj)) < 1e-6 * np.max(np.abs(uj))
    if not info["active"]:
        # a network that failed the gate is not used: the solve is plain Jacobi-PCG
        assert sl.fallback == 0 and sl.iters == sj.iters


def test_simp_learned_matches_fp64(gpu):
    dims = (24, 8, 8)
    x0 = np.full(dims[0] * dims[1] * dims[2], 0.3)
    res = []
    for pre in (gpu.PRECOND_JACOBI, gpu.PRECOND_LEARNED):
        c = gpu.cantilever(*dims)
        c.filter_setup(1.5)
        if pre == gpu.PRECOND_LEARNED:
            c.learned_setup(seed=2, train_iters=3, train_steps=60, snapshot_every=16)
        res.append(c.simp_run(x0, 8, tol=1e-10, precond=pre))
    (xj, hj, sj), (xl, hl, sl) = res
    print(f"PCG iterations: Jacobi {sj.pcg_iters}, learned {sl.pcg_iters} "
          f"(fallbacks {sl.fallbacks}, training {sl.train_s:.3f} s)")
    assert np.max(np.abs(xl - xj)) < 1e-6
    assert np.max(np.abs(hl[:, 0] - hj[:, 0]) / hj[:, 0]) < 1e-8
    assert np.all(hl[:, 3] <= 1e-10)
    # the gate never lets the network make the loop slower than Jacobi by more than noise
    assert sl.pcg_iters <= 1.05 * sj.pcg_iters
