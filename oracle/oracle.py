"""ctypes binding of the CPU FP64 oracle (oracle/oracle.cpp).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py.  The product (icarat2_amd) never imports it.
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None
_REF = None

dp = C.POINTER(C.c_double)
ip = C.POINTER(C.c_int)


class SimpOpts(C.Structure):
    _fields_ = [("n_iter", C.c_int), ("optimizer", C.c_int), ("use_projection", C.c_int),
                ("cg_tol", C.c_double), ("cg_maxit", C.c_int), ("xmin", C.c_double),
                ("xmax", C.c_double)]


class SimpStats(C.Structure):
    _fields_ = [("assembly_s", C.c_double), ("pcg_s", C.c_double), ("filter_s", C.c_double),
                ("sens_s", C.c_double), ("update_s", C.c_double), ("total_s", C.c_double),
                ("pcg_iters", C.c_longlong)]


def build():
    import subprocess
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(HERE, "liboracle.so")
        if not os.path.exists(path):
            build()
        L = C.CDLL(path)
        vp = C.c_void_p
        sig = {
            "orc_create": (vp, [C.c_int, C.c_int, C.c_int, C.c_double, C.c_double, C.c_double,
                                C.c_double, C.c_double, C.c_double, C.c_double]),
            "orc_destroy": (None, [vp]),
            "orc_numnp": (C.c_longlong, [vp]), "orc_nelx": (C.c_longlong, [vp]),
            "orc_neq": (C.c_longlong, [vp]), "orc_numeq": (C.c_int, [vp]),
            "orc_coords": (None, [vp, dp]), "orc_connectivity": (None, [vp, ip]),
            "orc_add_dirichlet": (None, [vp, ip, dp, dp]), "orc_make_dof": (C.c_int, [vp]),
            "orc_dof": (None, [vp, ip]), "orc_dirichlet_values": (None, [vp, dp]),
            "orc_add_neumann": (None, [vp, dp, dp]), "orc_fext": (None, [vp, dp]),
            "orc_ke0": (None, [vp, dp]),
            "orc_element_ke": (C.c_double, [vp, C.c_longlong, C.c_double, dp]),
            "orc_set_density": (None, [vp, dp]), "orc_nnz": (C.c_longlong, [vp]),
            "orc_spmv": (None, [vp, dp, dp]), "orc_diag": (None, [vp, dp]),
            "orc_apply_k_elementwise": (None, [vp, dp, dp, dp]),
            "orc_set_density_mf": (None, [vp, dp]),
            "orc_apply_k_nodal": (None, [vp, dp, dp, dp]),
            "orc_true_residual_nodal": (C.c_double, [vp, dp]),
            "orc_scaling_cg": (C.c_int, [vp, dp, dp, C.c_double, C.c_int, dp]),
            "orc_true_residual": (C.c_double, [vp, dp]),
            "orc_node_values": (None, [vp, dp, dp]),
            "orc_compliance": (C.c_double, [vp, dp]),
            "orc_sensitivity": (None, [vp, dp, C.c_double, dp, dp]),
            "orc_filter_create": (vp, [vp, C.c_double, C.c_int]),
            "orc_filter_destroy": (None, [vp]), "orc_filter_nnz": (C.c_longlong, [vp]),
            "orc_filter_lists": (None, [vp, C.POINTER(C.c_longlong), ip, dp]),
            "orc_density_filter": (C.c_int, [vp, dp, dp, C.c_int]),
            "orc_threshold": (None, [C.c_longlong, dp, C.c_double, C.c_double, dp]),
            "orc_threshold_deriv": (None, [C.c_longlong, dp, C.c_double, C.c_double, dp]),
            "orc_oc_update": (C.c_int, [vp, dp, dp, C.c_double, dp, dp, dp, C.c_int,
                                        C.c_double, dp]),
            "orc_mma_create": (vp, [C.c_int, C.c_int, C.c_double, C.c_double, C.c_double]),
            "orc_mma_destroy": (None, [vp]),
            "orc_mma_set": (None, [vp] + [C.c_double] * 5),
            "orc_mma_update": (None, [vp, dp, dp, dp, dp, dp, dp]),
            "orc_mma_state": (None, [vp, dp, dp]),
            "orc_simp_run": (C.c_int, [vp, vp, C.POINTER(SimpOpts), dp, dp,
                                       C.POINTER(SimpStats)]),
            "orc_gaussian_field": (C.c_double, [C.c_int, C.c_int, C.c_int, C.c_double,
                                                C.c_ulonglong, C.c_int, C.c_double, C.c_double,
                                                C.c_double, C.c_double, C.c_double, dp]),
            "orc_num_threads": (C.c_int, []),
            "orc_mises": (None, [vp, dp, dp, dp]),
            "orc_gcmma_create": (vp, [C.c_int, C.c_int, C.c_double, C.c_double, C.c_double]),
            "orc_gcmma_destroy": (None, [vp]),
            "orc_gcmma_set": (None, [vp] + [C.c_double] * 5),
            "orc_gcmma_outer": (None, [vp, dp, dp, C.c_double, dp, dp, dp, dp, dp]),
            "orc_gcmma_inner": (None, [vp, dp, C.c_double, dp, dp, C.c_double, dp, dp, dp, dp,
                                       dp]),
            "orc_gcmma_concheck": (C.c_int, [vp, C.c_double, dp]),
            "orc_gcmma_state": (None, [vp, dp]),
            "orc_simp_run_gcmma": (C.c_int, [vp, vp, C.POINTER(SimpOpts), dp, dp,
                                             C.POINTER(SimpStats)]),
            "orc_helm_create": (vp, [C.c_int, C.c_int, C.c_int, C.c_int, C.c_double, C.c_double,
                                     C.c_double, C.c_double]),
            "orc_helm_destroy": (None, [vp]), "orc_helm_numnp": (C.c_longlong, [vp]),
            "orc_sensitivity_filter": (None, [vp, dp, dp, dp]),
            "orc_helm_element": (None, [vp, C.c_longlong, dp, ip]),
            "orc_helm_apply": (None, [vp, dp, dp]),
            "orc_helm_filter": (C.c_int, [vp, dp, dp, C.c_int, C.c_double, C.c_int]),
            "orc_helm_oc_update": (C.c_int, [vp, vp, dp, dp, C.c_double, dp, dp, dp, C.c_int,
                                             C.c_double, C.c_double]),
            "orc_simp_run_helm": (C.c_int, [vp, vp, C.POINTER(SimpOpts), C.c_double, dp, dp,
                                            C.POINTER(SimpStats)]),
        }
        for name, (res, args) in sig.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _LIB = L
    return _LIB


def ref_lib():
    """oracle/_ref/libref_optim.so (the reference's MMASolver.cpp, GCMMASolver.cpp,
    OCSolver.hpp, opt_base.hpp and export_graph.hpp), or None."""
    global _REF
    if _REF is None:
        path = os.path.join(HERE, "_ref", "libref_optim.so")
        if not os.path.exists(path):
            return None
        L = C.CDLL(path)
        vp = C.c_void_p
        L.ref_mma_create.restype = vp
        L.ref_mma_create.argtypes = [C.c_int, C.c_int, C.c_double, C.c_double, C.c_double]
        L.ref_mma_destroy.argtypes = [vp]
        L.ref_mma_set.argtypes = [vp] + [C.c_double] * 5
        L.ref_mma_update.argtypes = [vp, dp, dp, dp, dp, dp, dp]
        L.ref_oc_update.argtypes = [C.c_int, dp, dp, C.c_double, dp, dp, dp, C.c_void_p,
                                    C.c_void_p]
        L.ref_gcmma_create.restype = vp
        L.ref_gcmma_create.argtypes = [C.c_int, C.c_int, C.c_double, C.c_double, C.c_double]
        L.ref_gcmma_destroy.argtypes = [vp]
        L.ref_gcmma_set.argtypes = [vp] + [C.c_double] * 5
        L.ref_gcmma_outer.argtypes = [vp, dp, dp, C.c_double, dp, dp, dp, dp, dp]
        L.ref_gcmma_inner.argtypes = [vp, dp, C.c_double, dp, dp, C.c_double, dp, dp, dp, dp, dp]
        L.ref_gcmma_concheck.restype = C.c_int
        L.ref_gcmma_concheck.argtypes = [vp, C.c_double, dp]
        L.ref_add_plot.argtypes = [C.c_char_p, C.c_int, C.c_double, C.c_int]
        L.ref_export_graph.argtypes = [C.c_char_p, C.c_void_p, dp, C.c_int]
        L.ref_is_convergence.argtypes = [dp, C.c_int, C.c_double, C.c_double, C.POINTER(C.c_int)]
        _REF = L
    return _REF


REF_VOLUME_FN = C.CFUNCTYPE(C.c_double, dp, C.c_int, C.c_void_p)


def P(a):
    """double* of a contiguous float64 numpy array."""
    assert a.dtype == np.float64 and a.flags.c_contiguous
    return a.ctypes.data_as(dp)


def PI(a):
    assert a.dtype == np.int32 and a.flags.c_contiguous
    return a.ctypes.data_as(ip)


class Problem:
    """Voxel hexa8 box (reference MeshIcarat + BCIcarat + LinearElastic)."""

    def __init__(self, nx, ny, nz, lx=None, ly=None, lz=None, young1=1e-9, young2=1.0,
                 poisson=0.3, penalty=3.0):
        L = lib()
        self.L = L
        self.nx, self.ny, self.nz = nx, ny, nz
        self.h = L.orc_create(nx, ny, nz, float(lx if lx is not None else nx),
                              float(ly if ly is not None else ny),
                              float(lz if lz is not None else nz), young1, young2, poisson,
                              penalty)
        self.numnp = L.orc_numnp(self.h)
        self.nelx = L.orc_nelx(self.h)
        self.neq = L.orc_neq(self.h)
        self.numeq = -1

    def __del__(self):
        if getattr(self, "h", None):
            self.L.orc_destroy(self.h)
            self.h = None

    def coords(self):
        x = np.zeros(self.numnp * 3)
        self.L.orc_coords(self.h, P(x))
        return x.reshape(-1, 3)

    def connectivity(self):
        c = np.zeros(self.nelx * 8, dtype=np.int32)
        self.L.orc_connectivity(self.h, PI(c))
        return c.reshape(-1, 8)

    def add_dirichlet(self, flag, val, rng):
        f = np.array(list(flag) + [0] * (6 - len(flag)), dtype=np.int32)
        v = np.array(list(val) + [0.0] * (6 - len(val)), dtype=np.float64)
        r = np.array(rng, dtype=np.float64).reshape(6)
        self.L.orc_add_dirichlet(self.h, PI(f), P(v), P(r))

    def make_dof(self):
        self.numeq = self.L.orc_make_dof(self.h)
        return self.numeq

    def dof(self):
        d = np.zeros(self.numnp * 3, dtype=np.int32)
        self.L.orc_dof(self.h, PI(d))
        return d.reshape(-1, 3)

    def dirichlet_values(self):
        d = np.zeros(self.numnp * 3)
        self.L.orc_dirichlet_values(self.h, P(d))
        return d

    def add_neumann(self, val, rng):
        v = np.array(val, dtype=np.float64)
        r = np.array(rng, dtype=np.float64).reshape(6)
        self.L.orc_add_neumann(self.h, P(v), P(r))

    def fext(self):
        f = np.zeros(self.numeq)
        self.L.orc_fext(self.h, P(f))
        return f

    def ke0(self):
        k = np.zeros(576)
        self.L.orc_ke0(self.h, P(k))
        return k.reshape(24, 24)

    def element_ke(self, e, young):
        k = np.zeros(576)
        vol = self.L.orc_element_ke(self.h, e, young, P(k))
        return k.reshape(24, 24), vol

    def set_density(self, x):
        self.x = np.ascontiguousarray(x, dtype=np.float64)
        self.L.orc_set_density(self.h, P(self.x))

    def set_density_matrix_free(self, x):
        """Densities without assembling K (for meshes whose CSR does not fit);
        spmv / scaling_cg / diag need set_density instead."""
        self.x = np.ascontiguousarray(x, dtype=np.float64)
        self.L.orc_set_density_mf(self.h, P(self.x))

    def apply_k_nodal(self, u, dirval=None):
        """K.u row by row (node-centric gather over the <= 8 elements of each
        node), in parallel; needs only the densities."""
        u = np.ascontiguousarray(u, dtype=np.float64)
        y = np.zeros(self.numeq)
        dv = None if dirval is None else np.ascontiguousarray(dirval, dtype=np.float64)
        self.L.orc_apply_k_nodal(self.h, P(u), None if dv is None else P(dv), P(y))
        return y

    def true_residual_nodal(self, sol):
        return self.L.orc_true_residual_nodal(self.h, P(np.ascontiguousarray(sol)))

    def nnz(self):
        return self.L.orc_nnz(self.h)

    def spmv(self, u):
        u = np.ascontiguousarray(u, dtype=np.float64)
        y = np.zeros(self.numeq)
        self.L.orc_spmv(self.h, P(u), P(y))
        return y

    def diag(self):
        d = np.zeros(self.numeq)
        self.L.orc_diag(self.h, P(d))
        return d

    def apply_k_elementwise(self, u, dirval=None):
        u = np.ascontiguousarray(u, dtype=np.float64)
        y = np.zeros(self.numeq)
        dv = None if dirval is None else np.ascontiguousarray(dirval, dtype=np.float64)
        self.L.orc_apply_k_elementwise(self.h, P(u), None if dv is None else P(dv), P(y))
        return y

    def scaling_cg(self, rhs=None, tol=1e-8, maxit=10000):
        rhs = self.fext() if rhs is None else np.ascontiguousarray(rhs, dtype=np.float64)
        sol = np.zeros(self.numeq)
        rn = C.c_double(0.0)
        it = self.L.orc_scaling_cg(self.h, P(rhs), P(sol), tol, maxit, C.byref(rn))
        return sol, it, rn.value

    def true_residual(self, sol):
        return self.L.orc_true_residual(self.h, P(np.ascontiguousarray(sol)))

    def node_values(self, sol):
        nv = np.zeros(self.numnp * 3)
        self.L.orc_node_values(self.h, P(np.ascontiguousarray(sol)), P(nv))
        return nv

    def compliance(self, sol):
        return self.L.orc_compliance(self.h, P(np.ascontiguousarray(sol)))

    def sensitivity(self, nodeval, vnorm=1.0):
        df = np.zeros(self.nelx)
        dg = np.zeros(self.nelx)
        self.L.orc_sensitivity(self.h, P(np.ascontiguousarray(nodeval)), vnorm, P(df), P(dg))
        return df, dg

    def mises(self, nodeval):
        """Element-mean stress (nelx x 6) and von Mises stress of the nodal
        values, densities as last set (Elastic::makeMisesStress)."""
        nodeval = np.ascontiguousarray(nodeval, dtype=np.float64)
        m = np.zeros(self.nelx)
        st = np.zeros(6 * self.nelx)
        self.L.orc_mises(self.h, P(nodeval), P(m), P(st))
        return m, st.reshape(self.nelx, 6)

    def gaussian_field(self, seed, nblob=32, smin=4.0, smax=12.0, mean=0.3, lo=1e-3, hi=1.0):
        return gaussian_field(self.nx, self.ny, self.nz, seed, nblob, smin, smax, mean, lo, hi)


def gaussian_field(nx, ny, nz, seed, nblob=32, smin=4.0, smax=12.0, mean=0.3, lo=1e-3, hi=1.0,
                   h=1.0):
    out = np.zeros(nx * ny * nz)
    m = lib().orc_gaussian_field(nx, ny, nz, h, seed, nblob, smin, smax, mean, lo, hi, P(out))
    return out, m


class Filter:
    """Cone-weight density filter (reference optimization/filter.hpp)."""

    def __init__(self, prob, radius, brute=False):
        self.prob = prob
        self.L = prob.L
        self.radius = radius
        self.h = self.L.orc_filter_create(prob.h, radius, 1 if brute else 0)

    def __del__(self):
        if getattr(self, "h", None):
            self.L.orc_filter_destroy(self.h)
            self.h = None

    def lists(self):
        nnz = self.L.orc_filter_nnz(self.h)
        rp = np.zeros(self.prob.nelx + 1, dtype=np.int64)
        ids = np.zeros(nnz, dtype=np.int32)
        dis = np.zeros(nnz)
        self.L.orc_filter_lists(self.h, rp.ctypes.data_as(C.POINTER(C.c_longlong)), PI(ids),
                                P(dis))
        return rp, ids, dis

    def density_filter(self, x, is_sens=False):
        x = np.ascontiguousarray(x, dtype=np.float64)
        out = np.zeros(self.prob.nelx)
        rc = self.L.orc_density_filter(self.h, P(x), P(out), 1 if is_sens else 0)
        if rc != 0:
            raise RuntimeError("coefficient aa is too small in the density filter")
        return out

    def sensitivity_filter(self, x, dfds):
        """Filter::sensitivityFilter(design_s, dfds) (filter.hpp:76-104)."""
        x = np.ascontiguousarray(x, dtype=np.float64)
        d = np.ascontiguousarray(dfds, dtype=np.float64)
        out = np.zeros(self.prob.nelx)
        self.L.orc_sensitivity_filter(self.h, P(x), P(d), P(out))
        return out

    def oc_update(self, x, dfdx, g, dgdx, xmin=1e-4, xmax=1.0, use_projection=False, beta=1.0):
        n = self.prob.nelx
        x = np.array(x, dtype=np.float64)
        lo = np.full(n, xmin)
        hi = np.full(n, xmax)
        lam = C.c_double(0.0)
        cnt = self.L.orc_oc_update(self.h, P(x), P(np.ascontiguousarray(dfdx)), g,
                                   P(np.ascontiguousarray(dgdx)), P(lo), P(hi),
                                   1 if use_projection else 0, beta, C.byref(lam))
        return x, cnt, lam.value

    def simp_run(self, x0, n_iter, optimizer=0, use_projection=False, cg_tol=1e-8,
                 cg_maxit=10000, xmin=1e-4, xmax=1.0):
        o = SimpOpts(n_iter, optimizer, 1 if use_projection else 0, cg_tol, cg_maxit, xmin, xmax)
        x = np.array(x0, dtype=np.float64)
        hist = np.zeros(4 * n_iter)
        st = SimpStats()
        self.L.orc_simp_run(self.prob.h, self.h, C.byref(o), P(x), P(hist), C.byref(st))
        return x, hist.reshape(n_iter, 4), st


class HelmFilter:
    """Helmholtz PDE filter (reference optimization/filter_helm.hpp) on a
    structured grid: dim 2 = quad4 (nz ignored), dim 3 = hexa8.  `tol` is the
    relative residual of its Jacobi-PCG solve (the reference's Eigen_CG
    default is 1e-10)."""

    def __init__(self, dims, radius, h=(1.0, 1.0, 1.0), dim=3, tol=1e-13):
        self.L = lib()
        nx, ny = dims[0], dims[1]
        nz = dims[2] if dim == 3 else 1
        self.dim, self.dims, self.radius, self.tol = dim, (nx, ny, nz), radius, tol
        self.nelx = nx * ny * nz
        self.ne = 4 if dim == 2 else 8
        self.h = self.L.orc_helm_create(dim, nx, ny, nz, h[0], h[1], h[2], radius)
        self.numnp = self.L.orc_helm_numnp(self.h)
        self.iters = 0

    def __del__(self):
        if getattr(self, "h", None):
            self.L.orc_helm_destroy(self.h)
            self.h = None

    def element(self, e):
        ke = np.zeros(self.ne * self.ne)
        conn = np.zeros(self.ne, dtype=np.int32)
        self.L.orc_helm_element(self.h, e, P(ke), PI(conn))
        return ke.reshape(self.ne, self.ne), conn

    def apply(self, v):
        v = np.ascontiguousarray(v, dtype=np.float64)
        y = np.zeros(self.numnp)
        self.L.orc_helm_apply(self.h, P(v), P(y))
        return y

    def density_filter(self, x, is_sens=False, tol=None, maxit=100000):
        x = np.ascontiguousarray(x, dtype=np.float64)
        out = np.zeros(self.nelx)
        it = self.L.orc_helm_filter(self.h, P(x), P(out), 1 if is_sens else 0,
                                    self.tol if tol is None else tol, maxit)
        if it < 0:
            raise RuntimeError("Helmholtz filter CG did not converge")
        self.iters = it
        return out

    def oc_update(self, prob, x, dfdx, g, dgdx, xmin=1e-4, xmax=1.0, use_projection=False,
                  beta=1.0):
        n = self.nelx
        x = np.array(x, dtype=np.float64)
        cnt = self.L.orc_helm_oc_update(self.h, prob.h, P(x), P(np.ascontiguousarray(dfdx)), g,
                                        P(np.ascontiguousarray(dgdx)), P(np.full(n, xmin)),
                                        P(np.full(n, xmax)), 1 if use_projection else 0, beta,
                                        self.tol)
        return x, cnt

    def simp_run(self, prob, x0, n_iter, optimizer=0, use_projection=False, cg_tol=1e-8,
                 cg_maxit=10000, xmin=1e-4, xmax=1.0):
        o = SimpOpts(n_iter, optimizer, 1 if use_projection else 0, cg_tol, cg_maxit, xmin, xmax)
        x = np.array(x0, dtype=np.float64)
        hist = np.zeros(4 * n_iter)
        st = SimpStats()
        self.L.orc_simp_run_helm(prob.h, self.h, C.byref(o), self.tol, P(x), P(hist),
                                 C.byref(st))
        return x, hist.reshape(n_iter, 4), st


def threshold(x, beta, T=0.5):
    x = np.ascontiguousarray(x, dtype=np.float64)
    out = np.zeros_like(x)
    lib().orc_threshold(x.size, P(x), beta, T, P(out))
    return out


def threshold_deriv(x, beta, T=0.5):
    x = np.ascontiguousarray(x, dtype=np.float64)
    out = np.zeros_like(x)
    lib().orc_threshold_deriv(x.size, P(x), beta, T, P(out))
    return out


class MMA:
    def __init__(self, n, m, a=0.0, c=1e5, d=0.0, reference=False):
        self.ref = reference
        self.L = ref_lib() if reference else lib()
        self.n, self.m = n, m
        self.h = (self.L.ref_mma_create if reference else self.L.orc_mma_create)(n, m, a, c, d)

    def set(self, move, albefa, asyminit, asymdec, asyminc):
        (self.L.ref_mma_set if self.ref else self.L.orc_mma_set)(self.h, move, albefa, asyminit,
                                                                 asymdec, asyminc)

    def update(self, x, dfdx, gx, dgdx, xmin, xmax):
        f = self.L.ref_mma_update if self.ref else self.L.orc_mma_update
        args = [np.ascontiguousarray(a, dtype=np.float64) for a in (dfdx, gx, dgdx, xmin, xmax)]
        f(self.h, P(x), *[P(a) for a in args])

    def state(self):
        """the asymptotes (low, upp) of the last update (oracle only)"""
        low, upp = np.zeros(self.n), np.zeros(self.n)
        self.L.orc_mma_state(self.h, P(low), P(upp))
        return low, upp

    def __del__(self):
        if getattr(self, "h", None):
            (self.L.ref_mma_destroy if self.ref else self.L.orc_mma_destroy)(self.h)
            self.h = None


class GCMMA:
    """GCMMASolver(n, m, a, c, d) (reference optimization/opt-solver/GCMMASolver.hpp);
    reference=True drives the compiled reference in oracle/_ref."""

    def __init__(self, n, m, a=0.0, c=1000.0, d=0.0, reference=False):
        self.ref = reference
        self.L = ref_lib() if reference else lib()
        self.n, self.m = n, m
        self.pre = "ref_gcmma_" if reference else "orc_gcmma_"
        self.h = getattr(self.L, self.pre + "create")(n, m, a, c, d)

    def _f(self, name):
        return getattr(self.L, self.pre + name)

    def set(self, move, albefa, asyminit, asymdec, asyminc):
        self._f("set")(self.h, move, albefa, asyminit, asymdec, asyminc)

    def outer(self, xval, f0x, df0dx, fx, dfdx, xmin, xmax):
        xm = np.zeros(self.n)
        a = [np.ascontiguousarray(v, dtype=np.float64) for v in (xval, df0dx, fx, dfdx, xmin, xmax)]
        self._f("outer")(self.h, P(xm), P(a[0]), f0x, P(a[1]), P(a[2]), P(a[3]), P(a[4]), P(a[5]))
        return xm

    def inner(self, xmma, f0xnew, fxnew, xval, f0x, df0dx, fx, dfdx, xmin, xmax):
        xm = np.array(xmma, dtype=np.float64)
        a = [np.ascontiguousarray(v, dtype=np.float64)
             for v in (fxnew, xval, df0dx, fx, dfdx, xmin, xmax)]
        self._f("inner")(self.h, P(xm), f0xnew, P(a[0]), P(a[1]), f0x, P(a[2]), P(a[3]),
                         P(a[4]), P(a[5]), P(a[6]))
        return xm

    def concheck(self, f0xnew, fxnew):
        return bool(self._f("concheck")(self.h, f0xnew,
                                        P(np.ascontiguousarray(fxnew, dtype=np.float64))))

    def state(self):
        """{raa0, raa[m], f0app, fapp[m]} (oracle only)."""
        out = np.zeros(2 + 2 * self.m)
        self.L.orc_gcmma_state(self.h, P(out))
        return out

    def __del__(self):
        if getattr(self, "h", None):
            self._f("destroy")(self.h)
            self.h = None


def simp_run_gcmma(prob, filt, x0, n_iter, use_projection=True, cg_tol=1e-8, cg_maxit=10000,
                   xmin=1e-4, xmax=1.0):
    """The GCMMA driver's loop (opt-macro-gcmma/opt_macro_gcmma.cpp:84-212) with
    the cone filter `filt`; hist rows {compliance, volume, PCG iterations,
    inner iterations}."""
    o = SimpOpts(n_iter, 2, 1 if use_projection else 0, cg_tol, cg_maxit, xmin, xmax)
    x = np.array(x0, dtype=np.float64)
    hist = np.zeros(4 * n_iter)
    st = SimpStats()
    lib().orc_simp_run_gcmma(prob.h, filt.h, C.byref(o), P(x), P(hist), C.byref(st))
    return x, hist.reshape(n_iter, 4), st


def cantilever(nx, ny, nz, edge_load=True):
    """Cantilever of the SIMP configs: x=0 clamped; unit total -z load on the
    bottom edge (y = 0..ny, z = 0) of the free end x = nx, as point loads
    (-1/(2 ny) at the two ends, -1/ny inside)."""
    pr = Problem(nx, ny, nz)
    pr.add_dirichlet([1, 1, 1, 0, 0, 0], [0.0] * 6, [[0, 0], [0, ny], [0, nz]])
    pr.make_dof()
    for j in range(ny + 1):
        v = -1.0 / (2 * ny) if j in (0, ny) else -1.0 / ny
        pr.add_neumann([0.0, 0.0, v], [[nx, nx], [j, j], [0, 0]])
    return pr
