"""GPU parity of the element-mean and von Mises stress (reference
problem/elasticity.cpp:166-209) against the oracle, on the same solution."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("dims,L", [((24, 8, 8), None), ((12, 6, 4), (6.0, 9.0, 2.0))])
def test_mises_matches_oracle(gpu, oracle, dims, L):
    kw = {} if L is None else dict(lx=L[0], ly=L[1], lz=L[2])
    Lx, Ly, Lz = L if L is not None else dims
    c = gpu.Context(*dims, **kw)
    o = oracle.Problem(*dims, Lx, Ly, Lz)
    for p in (c, o):
        p.add_dirichlet([1, 1, 1, 0, 0, 0], [0.0] * 6, [[0, 0], [0, Ly], [0, Lz]])
        p.make_dof()
        # unit -z traction on the free end face (face-load path)
        p.add_neumann([0.0, 0.0, -1.0], [[Lx, Lx], [0, Ly], [0, Lz]])
    x, _ = gpu.gaussian_field(*dims, 7, smin=2.0, smax=4.0)
    c.set_density(x)
    o.set_density(x)
    ug, st = c.solve(tol=1e-10)
    assert st.status == 0
    mg, sg = c.mises()
    mo, so = o.mises(o.node_values(ug))
    assert np.max(np.abs(sg - so)) < 1e-12 * np.max(np.abs(so))
    assert np.max(np.abs(mg - mo)) < 1e-12 * np.max(mo)
    assert np.all(mg >= 0.0)


def test_mises_before_solve_is_an_error(gpu):
    c = gpu.cantilever(4, 2, 2)
    c.set_density(np.full(c.nelx, 0.5))
    with pytest.raises(gpu.IcaratError):
        c.mises()
