"""The entry point that allocates no device memory (icarat_hip_mises_noalloc)
against icarat_hip_mises, through the C ABI: the same element stresses and
von Mises stresses on two calls in a row, the second on a smaller problem, and
the size limit of its fixed scratch array."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _solved(gpu, dims, seed, L=None):
    kw = {} if L is None else dict(lx=L[0], ly=L[1], lz=L[2])
    Ly, Lz = (L[1], L[2]) if L is not None else (dims[1], dims[2])
    Lx = L[0] if L is not None else dims[0]
    c = gpu.Context(*dims, **kw)
    c.add_dirichlet([1, 1, 1, 0, 0, 0], [0.0] * 6, [[0, 0], [0, Ly], [0, Lz]])
    c.make_dof()
    c.add_neumann([0.0, 0.0, -1.0], [[Lx, Lx], [0, Ly], [0, Lz]])
    x, _ = gpu.gaussian_field(*dims, seed, smin=2.0, smax=4.0)
    c.set_density(x)
    _, st = c.solve(tol=1e-10)
    assert st.status == 0
    return c


def test_noalloc_matches_allocating_entry_point(gpu):
    big = _solved(gpu, (40, 16, 16), 11)
    small = _solved(gpu, (12, 6, 4), 12, L=(6.0, 9.0, 2.0))
    # the no-allocation calls first, so the allocating path cannot have
    # prepared any state they rely on; then the bigger problem once more after
    # the smaller one (the scratch still holds the smaller problem's values)
    results = []
    for c in (big, small, big):
        mn, sn = c.mises_noalloc()
        results.append((mn, sn))
    for c, (mn, sn) in zip((big, small, big), results):
        ma, sa = c.mises()
        assert sn.shape == sa.shape == (c.nelx, 6)
        assert np.array_equal(mn, ma)
        assert np.array_equal(sn, sa)
        assert np.max(np.abs(sa)) > 0.0


def test_noalloc_limit_and_state_errors(gpu):
    # beyond the scratch array's size: rejected before any device work
    c = gpu.Context(128, 96, 86)  # 1,056,768 elements > 2^20
    assert c.nelx > 1 << 20
    with pytest.raises(gpu.IcaratError) as e:
        c.mises_noalloc()
    assert e.value.code == gpu.INVALID_ARGUMENT
    assert "maximum" in str(e.value)
    # no solution yet: not ready, as the allocating entry point
    d = gpu.cantilever(4, 2, 2)
    d.set_density(np.full(d.nelx, 0.5))
    with pytest.raises(gpu.IcaratError) as e:
        d.mises_noalloc()
    assert e.value.code == gpu.NOT_READY
