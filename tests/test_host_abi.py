"""The product's host logic and C ABI without a GPU: the library loads, exports
every declared entry point, numbers DOFs and builds loads and KE0 exactly as
the oracle does, and reports errors as status codes."""
import os
import re

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    names = set()
    for f in os.listdir(os.path.join(ROOT, "include")):
        if f.endswith(".h"):
            txt = open(os.path.join(ROOT, "include", f)).read()
            names |= set(re.findall(r"\b(icarat_hip_\w+)\s*\(", txt))
    return sorted(names)


def test_library_exports_every_declared_symbol(capi):
    L = capi.lib()
    decl = declared_symbols()
    assert len(decl) >= 25
    for name in decl:
        assert hasattr(L, name), name
    # the binding covers the header
    assert set(decl) == set(capi.exported_symbols())


def test_library_is_gfx950_code_object():
    import subprocess
    path = os.path.join(ROOT, "icarat2_amd", "libicarat_hip.so")
    out = subprocess.run(["/opt/rocm/llvm/bin/llvm-readobj", "--sections", path],
                         capture_output=True, text=True).stdout
    assert ".hip_fatbin" in out
    blob = open(path, "rb").read()
    assert b"gfx950" in blob


@pytest.mark.parametrize("dims", [(60, 20, 20), (7, 5, 3)])
def test_dof_and_loads_match_oracle(capi, oracle, dims):
    nx, ny, nz = dims
    c = capi.cantilever(nx, ny, nz)
    o = oracle.cantilever(nx, ny, nz)
    assert (c.numnp, c.nelx, c.numeq) == (o.numnp, o.nelx, o.numeq)
    assert np.array_equal(c.dof(), o.dof())
    assert np.array_equal(c.fext(), o.fext())  # point loads: bit-exact


def test_partial_dirichlet_and_face_loads_match_oracle(capi, oracle):
    args_d = [([1, 0, 1], [0.0, 0.0, 0.0], [[0, 0], [0, 3], [0, 2]]),
              ([0, 1, 0], [0.0, 0.5, 0.0], [[4, 4], [0, 0], [0, 2]])]
    args_n = [([0.0, 0.0, -2.0], [[4, 4], [0, 3], [0, 2]]),
              ([1.0, 0.5, 0.0], [[0, 4], [0, 3], [2, 2]]),
              ([0.0, 3.0, 0.0], [[0, 4], [3, 3], [0, 2]]),
              ([0.0, 0.0, 0.7], [[2, 2], [1, 1], [1, 1]])]
    c = capi.Context(4, 3, 2)
    o = oracle.Problem(4, 3, 2)
    for a in args_d:
        c.add_dirichlet(*a)
        o.add_dirichlet(*a)
    assert c.make_dof() == o.make_dof()
    for a in args_n:
        c.add_neumann(*a)
        o.add_neumann(*a)
    assert np.array_equal(c.dof(), o.dof())
    assert np.max(np.abs(c.fext() - o.fext())) < 1e-15


def test_ke0_matches_oracle(capi, oracle):
    # both round the exact element stiffness: bit-identical
    for nu in (0.3, 0.25, 0.45):
        c = capi.Context(2, 2, 2, poisson=nu)
        o = oracle.Problem(2, 2, 2, poisson=nu)
        assert np.array_equal(c.ke0(), o.ke0())
    # non-unit cubic voxels scale KE by h; box voxels
    for L in ((8.0, 8.0, 8.0), (6.0, 9.0, 2.0)):
        c = capi.Context(4, 4, 4, *L)
        o = oracle.Problem(4, 4, 4, *L)
        assert np.max(np.abs(c.ke0() - o.ke0())) <= 2.3e-16 * np.max(np.abs(o.ke0()))


def test_gaussian_field_matches_oracle(capi, oracle):
    for dims, seed in [((24, 12, 10), 2), ((16, 8, 8), 4)]:
        a, ma = capi.gaussian_field(*dims, seed)
        b, mb = oracle.gaussian_field(*dims, seed)
        assert np.array_equal(a, b)
        assert ma == mb


def test_errors_are_status_codes(capi):
    c = capi.Context(4, 2, 2)
    # loads before numbering
    with pytest.raises(capi.IcaratError) as e:
        c.add_neumann([0, 0, -1.0], [[4, 4], [0, 0], [0, 0]])
    assert e.value.code == capi.NOT_READY
    c.add_dirichlet([1, 1, 1], [0, 0, 0], [[0, 0], [0, 2], [0, 2]])
    c.make_dof()
    # density of the wrong length
    with pytest.raises(capi.IcaratError) as e:
        c.set_density(np.ones(3))
    assert e.value.code == capi.INVALID_ARGUMENT
    # null pointers through the raw ABI
    L = capi.lib()
    assert L.icarat_hip_apply_K(c.h, None, None, 0) == capi.INVALID_ARGUMENT
    assert L.icarat_hip_destroy(None) == capi.INVALID_ARGUMENT
    # a bad grid
    import ctypes as C
    d = capi.GridDesc(0, 2, 2, 1.0, 1.0, 1.0, 1e-9, 1.0, 0.3, 3.0)
    h = C.c_void_p()
    assert L.icarat_hip_create(C.byref(d), C.byref(h)) == capi.INVALID_ARGUMENT
    d = capi.GridDesc(2, 2, 2, 2.0, 2.0, 2.0, 1.0, 1e-9, 0.3, 3.0)  # young2 < young1
    assert L.icarat_hip_create(C.byref(d), C.byref(h)) == capi.INVALID_ARGUMENT
    # MMA with m constraints: 1 <= m <= 8, n >= 1 (checked before any device call)
    for n, m in ((10, 9), (10, 0), (0, 2)):
        assert L.icarat_hip_mmasolver_create(n, m, 0.0, 1e5, 0.0, C.byref(h)) == capi.INVALID_ARGUMENT
    assert L.icarat_hip_mmasolver_update(None, None, None, None, None, None, None, None) == \
        capi.INVALID_ARGUMENT
    for n, m in ((10, 9), (10, 0), (0, 2)):
        assert L.icarat_hip_gcmmasolver_create(n, m, 0.0, 1e3, 0.0, C.byref(h)) == \
            capi.INVALID_ARGUMENT
    assert L.icarat_hip_gcmmasolver_state(None, None) == capi.INVALID_ARGUMENT


def test_device_calls_without_gpu_fail_loudly(capi):
    if capi.device_available():
        pytest.skip("a GPU is present")
    c = capi.cantilever(4, 2, 2)
    with pytest.raises(capi.IcaratError) as e:
        c.set_density(np.full(c.nelx, 0.5))
    assert e.value.code == capi.HIP_ERROR


FACADE_DRIVER = r'''
#include <cstdio>
#include "icarat_hip.hpp"
int main(int argc, char **argv) {
  icarat_hip::Device d(6, 3, 2, 6.0, 3.0, 2.0, 1e-9, 1.0, 0.3, 3.0);
  const int32_t flag[6] = {1, 1, 1, 0, 0, 0};
  const double val[6] = {0, 0, 0, 0, 0, 0};
  const double clamp[6] = {0, 0, 0, 3, 0, 2};
  d.addDirichlet(flag, val, clamp);
  const int numeq = d.makeDOF();
  const double load[3] = {0.0, 0.0, -1.0};
  const double tip[6] = {6, 6, 0, 0, 0, 0};
  d.addNeumann(load, tip);
  double s = 0.0;
  for (double v : d.fext()) s += v;
  std::printf("%d %lld %lld %.17g\n", numeq, d.numnp(), d.nelx(), s);
  std::fflush(stdout);
  if (argc > 1) {
    // the first call that needs the device
    std::vector<double> x(d.nelx(), 0.5);
    d.setDesign(x);
    icarat_hip::Filter f(d, 1.5);
    std::vector<double> xf = f.densityFilter(x, false);
    std::printf("filtered %.17g\n", xf[0]);
  }
  return 0;
}
'''


def test_cpp_facade_compiles_and_keeps_reference_error_behaviour(capi, tmp_path):
    """The header-only facade (reference class names over the C ABI) builds with
    the host compiler, runs the host-side set-up, and turns a failing call into
    the reference's cerr + exit(1)."""
    import subprocess
    src = tmp_path / "driver.cpp"
    src.write_text(FACADE_DRIVER)
    exe = tmp_path / "driver"
    libdir = os.path.join(ROOT, "icarat2_amd")
    subprocess.run(["g++", "-std=c++17", "-O1", "-I", os.path.join(ROOT, "include"), str(src),
                    "-L", libdir, "-licarat_hip", f"-Wl,-rpath,{libdir}", "-o", str(exe)],
                   check=True)
    r = subprocess.run([str(exe)], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    numeq, numnp, nelx, s = r.stdout.split()
    assert (int(numeq), int(numnp), int(nelx)) == (3 * (7 * 4 * 3 - 4 * 3), 7 * 4 * 3, 36)
    assert float(s) == -1.0  # one point load of -1 at the single tip node (6, 0, 0)
    r = subprocess.run([str(exe), "device"], capture_output=True, text=True)
    if capi.device_available():
        assert r.returncode == 0 and "filtered" in r.stdout
    else:
        assert r.returncode == 1 and r.stderr.startswith("Error : setDesign")
