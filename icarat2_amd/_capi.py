"""ctypes binding of libicarat_hip.so (C ABI of include/icarat_hip.h).

The library is the product: this module only marshals arguments.  Loading
fails loudly when the built library is missing; there is no CPU fallback."""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libicarat_hip.so")

OK, INVALID_ARGUMENT, HIP_ERROR, NOT_READY, OUT_OF_MEMORY, NOT_CONVERGED, NUMERICAL, INTERNAL = range(8)
FP64, FP32, BF16 = 0, 1, 2
PRECOND_JACOBI, PRECOND_LEARNED, PRECOND_MG = 0, 1, 2

dp = C.POINTER(C.c_double)
i32p = C.POINTER(C.c_int32)


class GridDesc(C.Structure):
    _fields_ = [("nx", C.c_int32), ("ny", C.c_int32), ("nz", C.c_int32),
                ("lx", C.c_double), ("ly", C.c_double), ("lz", C.c_double),
                ("young1", C.c_double), ("young2", C.c_double),
                ("poisson", C.c_double), ("penalty", C.c_double)]


class SolveOpts(C.Structure):
    _fields_ = [("tol", C.c_double), ("maxit", C.c_int32), ("precision", C.c_int32),
                ("precond", C.c_int32), ("inner_tol", C.c_double), ("max_refine", C.c_int32),
                ("warm", C.c_int32)]


class SolveStats(C.Structure):
    _fields_ = [("iters", C.c_int32), ("refine_steps", C.c_int32), ("rec_relres", C.c_double),
                ("true_relres", C.c_double), ("seconds", C.c_double), ("fallback", C.c_int32),
                ("status", C.c_int32)]


class TomlInfo(C.Structure):
    _fields_ = [("nx", C.c_int32), ("ny", C.c_int32), ("nz", C.c_int32),
                ("lx", C.c_double), ("ly", C.c_double), ("lz", C.c_double),
                ("designs_0", C.c_double), ("filter_radius", C.c_double),
                ("n_dirichlet", C.c_int32), ("n_neumann", C.c_int32)]


class SimpOpts(C.Structure):
    _fields_ = [("n_iter", C.c_int32), ("optimizer", C.c_int32), ("use_projection", C.c_int32),
                ("filter", C.c_int32), ("xmin", C.c_double), ("xmax", C.c_double),
                ("solve", SolveOpts), ("conv_eps", C.c_double), ("conv_minstep", C.c_int32),
                ("history_csv", C.c_char_p), ("case_mode", C.c_int32)]


class SimpStats(C.Structure):
    _fields_ = [("seconds", C.c_double), ("solve_s", C.c_double), ("filter_s", C.c_double),
                ("sens_s", C.c_double), ("update_s", C.c_double), ("pcg_iters", C.c_int64),
                ("max_true_relres", C.c_double), ("train_s", C.c_double),
                ("fallbacks", C.c_int64), ("inner_iters", C.c_int64), ("steps", C.c_int32),
                ("converged", C.c_int32)]


class LearnedOpts(C.Structure):
    _fields_ = [("train_iters", C.c_int32), ("train_steps", C.c_int32),
                ("snapshot_every", C.c_int32), ("max_samples", C.c_int32),
                ("batch_iters", C.c_int32), ("stall_iters", C.c_int32), ("lr", C.c_float),
                ("gate", C.c_float), ("seed", C.c_uint64), ("base", C.c_int32),
                ("correction_scale", C.c_float)]


class LearnedInfo(C.Structure):
    _fields_ = [("samples", C.c_int32), ("active", C.c_int32), ("steps", C.c_int64),
                ("loss", C.c_double), ("val_loss", C.c_double), ("gain", C.c_double),
                ("n_params", C.c_int64), ("grid", C.c_int32 * 3), ("harvested", C.c_int32),
                ("fallback_iter", C.c_int32), ("trial_iters", C.c_int32),
                ("trial_base_iters", C.c_int32)]


class MgOpts(C.Structure):
    _fields_ = [("max_levels", C.c_int32), ("nu_pre", C.c_int32), ("nu_post", C.c_int32),
                ("nu_coarse", C.c_int32), ("omega", C.c_double), ("batch_iters", C.c_int32),
                ("dense_dofs", C.c_int32), ("direct_nodes", C.c_int32), ("coarse_avg", C.c_int32),
                ("w_levels", C.c_int32), ("w_switch", C.c_double), ("compensated", C.c_int32)]


class MgStatus(C.Structure):
    _fields_ = [("levels", C.c_int32), ("dense", C.c_int32), ("direct_fine", C.c_int32),
                ("fine_fp32", C.c_int32), ("factorizations", C.c_int64), ("shifted", C.c_int64),
                ("w_levels", C.c_int32), ("w_fallbacks", C.c_int64)]


fp = C.POINTER(C.c_float)
i64p = C.POINTER(C.c_int64)


_SIGS = {
    "icarat_hip_device_available": (C.c_int, []),
    "icarat_hip_create": (C.c_int, [C.POINTER(GridDesc), C.POINTER(C.c_void_p)]),
    "icarat_hip_destroy": (C.c_int, [C.c_void_p]),
    "icarat_hip_last_error": (C.c_char_p, [C.c_void_p]),
    "icarat_hip_counts": (C.c_int, [C.c_void_p, C.POINTER(C.c_int64), C.POINTER(C.c_int64),
                                    C.POINTER(C.c_int64), i32p]),
    "icarat_hip_add_dirichlet": (C.c_int, [C.c_void_p, i32p, dp, dp]),
    "icarat_hip_make_dof": (C.c_int, [C.c_void_p, i32p]),
    "icarat_hip_add_neumann": (C.c_int, [C.c_void_p, dp, dp]),
    "icarat_hip_new_load_case": (C.c_int, [C.c_void_p, i32p]),
    "icarat_hip_select_load_case": (C.c_int, [C.c_void_p, C.c_int32]),
    "icarat_hip_load_case_count": (C.c_int, [C.c_void_p, i32p]),
    "icarat_hip_get_dof": (C.c_int, [C.c_void_p, i32p]),
    "icarat_hip_get_fext": (C.c_int, [C.c_void_p, dp]),
    "icarat_hip_set_fext": (C.c_int, [C.c_void_p, dp]),
    "icarat_hip_get_ke0": (C.c_int, [C.c_void_p, dp]),
    "icarat_hip_set_density": (C.c_int, [C.c_void_p, dp, C.c_int64]),
    "icarat_hip_apply_K": (C.c_int, [C.c_void_p, dp, dp, C.c_int32]),
    "icarat_hip_get_diag": (C.c_int, [C.c_void_p, dp]),
    "icarat_hip_solve": (C.c_int, [C.c_void_p, C.POINTER(SolveOpts), dp, C.POINTER(SolveStats)]),
    "icarat_hip_true_residual": (C.c_int, [C.c_void_p, dp, dp]),
    "icarat_hip_compliance_sensitivity": (C.c_int, [C.c_void_p, C.c_double, dp, dp, dp]),
    "icarat_hip_filter_setup": (C.c_int, [C.c_void_p, C.c_double]),
    "icarat_hip_filter": (C.c_int, [C.c_void_p, dp, dp, C.c_int32]),
    "icarat_hip_solve_cases": (C.c_int, [C.c_void_p, C.c_void_p, dp, C.c_void_p]),
    "icarat_hip_sensitivity_filter": (C.c_int, [C.c_void_p, dp, dp, dp]),
    "icarat_hip_helm_setup": (C.c_int, [C.c_void_p, C.c_double]),
    "icarat_hip_mises": (C.c_int, [C.c_void_p, dp, dp]),
    "icarat_hip_mises_noalloc": (C.c_int, [C.c_void_p, dp, dp]),
    "icarat_hip_create_from_toml": (C.c_int, [C.c_char_p, C.POINTER(C.c_void_p),
                                              C.c_void_p, C.c_char_p, C.c_int32]),
    "icarat_hip_write_vtk": (C.c_int, [C.c_void_p, C.c_char_p, dp]),
    "icarat_hip_add_plot": (C.c_int, [C.c_char_p, C.c_double, C.c_double, C.c_int32]),
    "icarat_hip_export_graph": (C.c_int, [C.c_char_p, dp, dp, C.c_int64]),
    "icarat_hip_gcmma_set_params": (C.c_int, [C.c_void_p] + [C.c_double] * 5),
    "icarat_hip_gcmma_reset": (C.c_int, [C.c_void_p]),
    "icarat_hip_gcmma_outer": (C.c_int, [C.c_void_p, dp, dp, C.c_double, dp, C.c_double, dp,
                                         C.c_double, C.c_double]),
    "icarat_hip_gcmma_inner": (C.c_int, [C.c_void_p, dp, C.c_double, C.c_double, dp,
                                         C.c_double, dp, C.c_double, dp, C.c_double,
                                         C.c_double]),
    "icarat_hip_gcmma_concheck": (C.c_int, [C.c_void_p, C.c_double, C.c_double,
                                            C.POINTER(C.c_int32)]),
    "icarat_hip_helm_filter": (C.c_int, [C.c_void_p, dp, dp, C.c_int32]),
    "icarat_hip_threshold": (C.c_int, [C.c_void_p, dp, C.c_double, C.c_double, dp, C.c_int32]),
    "icarat_hip_oc_update": (C.c_int, [C.c_void_p, dp, dp, C.c_double, dp, C.c_double,
                                       C.c_double, C.c_int32, C.c_double, i32p]),
    "icarat_hip_mma_set_params": (C.c_int, [C.c_void_p] + [C.c_double] * 5),
    "icarat_hip_mma_reset": (C.c_int, [C.c_void_p]),
    "icarat_hip_mma_update": (C.c_int, [C.c_void_p, dp, dp, C.c_double, dp, C.c_double,
                                        C.c_double, i32p]),
    "icarat_hip_simp_run": (C.c_int, [C.c_void_p, C.POINTER(SimpOpts), dp, dp,
                                      C.POINTER(SimpStats)]),
    "icarat_hip_learned_default_opts": (C.c_int, [C.POINTER(LearnedOpts)]),
    "icarat_hip_learned_setup": (C.c_int, [C.c_void_p, C.POINTER(LearnedOpts)]),
    "icarat_hip_learned_harvest": (C.c_int, [C.c_void_p, C.POINTER(SolveOpts), dp,
                                             C.POINTER(SolveStats)]),
    "icarat_hip_learned_train": (C.c_int, [C.c_void_p, C.c_int32, dp]),
    "icarat_hip_learned_info": (C.c_int, [C.c_void_p, C.POINTER(LearnedInfo)]),
    "icarat_hip_learned_get_params": (C.c_int, [C.c_void_p, fp]),
    "icarat_hip_learned_set_params": (C.c_int, [C.c_void_p, fp]),
    "icarat_hip_learned_apply": (C.c_int, [C.c_void_p, dp, dp]),
    "icarat_hip_mg_default_opts": (C.c_int, [C.POINTER(MgOpts)]),
    "icarat_hip_mg_setup": (C.c_int, [C.c_void_p, C.POINTER(MgOpts)]),
    "icarat_hip_mg_info": (C.c_int, [C.c_void_p, i32p, i32p]),
    "icarat_hip_unet_create": (C.c_int, [C.c_int32, C.c_int32, C.c_int32, C.c_uint64,
                                         C.POINTER(C.c_void_p)]),
    "icarat_hip_unet_destroy": (C.c_int, [C.c_void_p]),
    "icarat_hip_unet_last_error": (C.c_char_p, [C.c_void_p]),
    "icarat_hip_unet_param_count": (C.c_int, [i64p]),
    "icarat_hip_unet_set_params": (C.c_int, [C.c_void_p, fp]),
    "icarat_hip_unet_get_params": (C.c_int, [C.c_void_p, fp]),
    "icarat_hip_unet_forward": (C.c_int, [C.c_void_p, fp, fp]),
    "icarat_hip_unet_grads": (C.c_int, [C.c_void_p, fp, fp, dp, fp]),
    "icarat_hip_unet_train_step": (C.c_int, [C.c_void_p, fp, fp, C.c_float, dp]),
    "icarat_hip_set_stencil": (C.c_int, [C.c_void_p, C.c_int32, C.c_int32, i32p]),
    "icarat_hip_batch_create": (C.c_int, [C.POINTER(C.c_void_p)]),
    "icarat_hip_batch_destroy": (C.c_int, [C.c_void_p]),
    "icarat_hip_batch_last_error": (C.c_char_p, [C.c_void_p]),
    "icarat_hip_batch_add": (C.c_int, [C.c_void_p, C.c_void_p]),
    "icarat_hip_batch_set_budget": (C.c_int, [C.c_void_p, C.c_int64]),
    "icarat_hip_batch_info": (C.c_int, [C.c_void_p, i32p, i64p]),
    "icarat_hip_batch_solve": (C.c_int, [C.c_void_p, C.POINTER(SolveOpts), C.POINTER(SolveStats)]),
    "icarat_hip_batch_simp_run": (C.c_int, [C.c_void_p, C.POINTER(SimpOpts), C.POINTER(dp),
                                            C.POINTER(dp), C.POINTER(SimpStats)]),
    "icarat_hip_batch_autotune": (C.c_int, [C.c_void_p, C.c_int32, i32p, dp]),
    "icarat_hip_mmasolver_create": (C.c_int, [C.c_int32, C.c_int32, C.c_double, C.c_double,
                                              C.c_double, C.POINTER(C.c_void_p)]),
    "icarat_hip_mmasolver_destroy": (C.c_int, [C.c_void_p]),
    "icarat_hip_mmasolver_last_error": (C.c_char_p, [C.c_void_p]),
    "icarat_hip_mmasolver_set_params": (C.c_int, [C.c_void_p] + [C.c_double] * 5),
    "icarat_hip_mmasolver_reset": (C.c_int, [C.c_void_p]),
    "icarat_hip_mmasolver_update": (C.c_int, [C.c_void_p, dp, dp, dp, dp, dp, dp, i32p]),
    "icarat_hip_mmasolver_asymptotes": (C.c_int, [C.c_void_p, dp, dp]),
    "icarat_hip_gcmmasolver_create": (C.c_int, [C.c_int32, C.c_int32, C.c_double, C.c_double,
                                                C.c_double, C.POINTER(C.c_void_p)]),
    "icarat_hip_gcmmasolver_destroy": (C.c_int, [C.c_void_p]),
    "icarat_hip_gcmmasolver_last_error": (C.c_char_p, [C.c_void_p]),
    "icarat_hip_gcmmasolver_set_params": (C.c_int, [C.c_void_p] + [C.c_double] * 5),
    "icarat_hip_gcmmasolver_reset": (C.c_int, [C.c_void_p]),
    "icarat_hip_gcmmasolver_outer": (C.c_int, [C.c_void_p, dp, dp, C.c_double, dp, dp, dp, dp,
                                               dp]),
    "icarat_hip_gcmmasolver_inner": (C.c_int, [C.c_void_p, dp, C.c_double, dp, dp, C.c_double,
                                               dp, dp, dp, dp, dp]),
    "icarat_hip_gcmmasolver_concheck": (C.c_int, [C.c_void_p, C.c_double, dp, i32p]),
    "icarat_hip_gcmmasolver_state": (C.c_int, [C.c_void_p, dp]),
    "icarat_hip_mg_status": (C.c_int, [C.c_void_p, C.POINTER(MgStatus)]),
    "icarat_hip_time_kernel": (C.c_int, [C.c_void_p, C.c_int32, C.c_int32, dp, dp, dp]),
    "icarat_hip_gaussian_field": (C.c_int, [C.c_int32, C.c_int32, C.c_int32, C.c_double,
                                            C.c_uint64, C.c_int32, C.c_double, C.c_double,
                                            C.c_double, C.c_double, C.c_double, dp, dp]),
}

_LIB = None


def lib():
    global _LIB
    if _LIB is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is not built: run `python -m icarat2_amd.build`")
        L = C.CDLL(LIB_PATH)
        for name, (res, args) in _SIGS.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _LIB = L
    return _LIB


def exported_symbols():
    return sorted(_SIGS)


class IcaratError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"[{code}] {msg}")
        self.code = code


def _p(a):
    assert a.dtype == np.float64 and a.flags.c_contiguous
    return a.ctypes.data_as(dp)


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def _f32(a):
    a = np.ascontiguousarray(a, dtype=np.float32)
    return a, a.ctypes.data_as(fp)


def learned_default_opts(**kw):
    """LearnedOpts with the library defaults, fields overridden by keyword."""
    o = LearnedOpts()
    rc = lib().icarat_hip_learned_default_opts(C.byref(o))
    if rc:
        raise IcaratError(rc, "learned_default_opts")
    for k, v in kw.items():
        setattr(o, k, v)
    return o


def unet_param_count():
    n = C.c_int64()
    lib().icarat_hip_unet_param_count(C.byref(n))
    return n.value


class UNet:
    """The learned preconditioner's network alone on an X x Y x Z cell grid
    (multiples of 4): x (Z, Y, X, 4) -> out (Z, Y, X, 3), FP32 host arrays."""

    def __init__(self, X, Y, Z, seed=1):
        self.L = lib()
        self.shape = (Z, Y, X)
        h = C.c_void_p()
        rc = self.L.icarat_hip_unet_create(X, Y, Z, seed, C.byref(h))
        if rc:
            raise IcaratError(rc, "unet_create")
        self.h = h

    def close(self):
        if getattr(self, "h", None):
            self.L.icarat_hip_unet_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def _chk(self, rc, what):
        if rc:
            msg = self.L.icarat_hip_unet_last_error(self.h)
            raise IcaratError(rc, f"{what}: {msg.decode() if msg else ''}")

    def set_params(self, flat):
        a, p = _f32(flat)
        assert a.size == unet_param_count()
        self._chk(self.L.icarat_hip_unet_set_params(self.h, p), "unet_set_params")

    def get_params(self):
        out = np.zeros(unet_param_count(), dtype=np.float32)
        self._chk(self.L.icarat_hip_unet_get_params(self.h, out.ctypes.data_as(fp)),
                  "unet_get_params")
        return out

    def forward(self, x):
        a, p = _f32(x)
        assert a.shape == self.shape + (4,)
        out = np.zeros(self.shape + (3,), dtype=np.float32)
        self._chk(self.L.icarat_hip_unet_forward(self.h, p, out.ctypes.data_as(fp)), "unet_forward")
        return out

    def grads(self, x, target):
        a, p = _f32(x)
        t, tp = _f32(target)
        g = np.zeros(unet_param_count(), dtype=np.float32)
        loss = C.c_double()
        self._chk(self.L.icarat_hip_unet_grads(self.h, p, tp, C.byref(loss), g.ctypes.data_as(fp)),
                  "unet_grads")
        return loss.value, g

    def train_step(self, x, target, lr=1e-3):
        a, p = _f32(x)
        t, tp = _f32(target)
        loss = C.c_double()
        self._chk(self.L.icarat_hip_unet_train_step(self.h, p, tp, lr, C.byref(loss)),
                  "unet_train_step")
        return loss.value


def add_plot(filename, x, y, is_first):
    """addPlot(x, y, filename, isFirst) of the reference's export_graph.hpp."""
    rc = lib().icarat_hip_add_plot(os.fsencode(filename), float(x), float(y), int(is_first))
    if rc:
        raise IcaratError(rc, f"add_plot: cannot write {filename}")


def export_graph(filename, yy, xx=None):
    """exportGraph(yy, filename) or exportGraph(xx, yy, filename)."""
    yy = _f64(yy)
    xp = None if xx is None else _p(_f64(xx))
    rc = lib().icarat_hip_export_graph(os.fsencode(filename), xp, _p(yy), yy.size)
    if rc:
        raise IcaratError(rc, f"export_graph: cannot write {filename}")


def device_available():
    return bool(lib().icarat_hip_device_available())


def gaussian_field(nx, ny, nz, seed, nblob=32, smin=4.0, smax=12.0, mean=0.3, lo=1e-3, hi=1.0,
                   h=1.0):
    out = np.zeros(nx * ny * nz)
    m = C.c_double(0.0)
    rc = lib().icarat_hip_gaussian_field(nx, ny, nz, h, seed, nblob, smin, smax, mean, lo, hi,
                                         _p(out), C.byref(m))
    if rc:
        raise IcaratError(rc, "gaussian_field")
    return out, m.value


class Context:
    """One voxel hexa8 SIMP problem on the GPU (mesh, BCs, K, filter, OC/MMA)."""

    def __init__(self, nx, ny, nz, lx=None, ly=None, lz=None, young1=1e-9, young2=1.0,
                 poisson=0.3, penalty=3.0):
        self.L = lib()
        d = GridDesc(nx, ny, nz, float(nx if lx is None else lx), float(ny if ly is None else ly),
                     float(nz if lz is None else lz), young1, young2, poisson, penalty)
        h = C.c_void_p()
        rc = self.L.icarat_hip_create(C.byref(d), C.byref(h))
        if rc:
            raise IcaratError(rc, "icarat_hip_create failed")
        self.h = h
        self.nx, self.ny, self.nz = nx, ny, nz
        a, b, c, e = C.c_int64(), C.c_int64(), C.c_int64(), C.c_int32()
        self.L.icarat_hip_counts(self.h, C.byref(a), C.byref(b), C.byref(c), C.byref(e))
        self.numnp, self.nelx, self.neq, self.numeq = a.value, b.value, c.value, e.value

    @classmethod
    def from_toml(cls, path):
        """A context from one of ICARAT's TOML configuration files; returns
        (context, TomlInfo)."""
        L = lib()
        h = C.c_void_p()
        info = TomlInfo()
        err = C.create_string_buffer(512)
        rc = L.icarat_hip_create_from_toml(os.fsencode(path), C.byref(h), C.byref(info), err, 512)
        if rc:
            raise IcaratError(rc, "create_from_toml: " + err.value.decode())
        self = cls.__new__(cls)
        self.L, self.h = L, h
        self.nx, self.ny, self.nz = info.nx, info.ny, info.nz
        a, b, c, e = C.c_int64(), C.c_int64(), C.c_int64(), C.c_int32()
        L.icarat_hip_counts(h, C.byref(a), C.byref(b), C.byref(c), C.byref(e))
        self.numnp, self.nelx, self.neq, self.numeq = a.value, b.value, c.value, e.value
        return self, info

    def write_vtk(self, path, density=None):
        """The drivers' result file (legacy VTK) of the current state."""
        d = None if density is None else _p(_f64(density))
        self._chk(self.L.icarat_hip_write_vtk(self.h, os.fsencode(path), d), "write_vtk")

    def close(self):
        if getattr(self, "h", None):
            self.L.icarat_hip_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def _chk(self, rc, what):
        if rc:
            raise IcaratError(rc, f"{what}: {self.L.icarat_hip_last_error(self.h).decode()}")

    # --- boundary conditions
    def add_dirichlet(self, flag, val, rng):
        f = np.array(list(flag) + [0] * (6 - len(flag)), dtype=np.int32)
        v = np.array(list(val) + [0.0] * (6 - len(val)), dtype=np.float64)
        r = np.array(rng, dtype=np.float64).reshape(6)
        self._chk(self.L.icarat_hip_add_dirichlet(self.h, f.ctypes.data_as(i32p), _p(v), _p(r)),
                  "add_dirichlet")

    def make_dof(self):
        n = C.c_int32()
        self._chk(self.L.icarat_hip_make_dof(self.h, C.byref(n)), "make_dof")
        self.numeq = n.value
        return n.value

    def add_neumann(self, val, rng):
        v = _f64(val)
        r = np.array(rng, dtype=np.float64).reshape(6)
        self._chk(self.L.icarat_hip_add_neumann(self.h, _p(v), _p(r)), "add_neumann")

    def new_load_case(self):
        i = C.c_int32()
        self._chk(self.L.icarat_hip_new_load_case(self.h, C.byref(i)), "new_load_case")
        return i.value

    def select_load_case(self, i):
        self._chk(self.L.icarat_hip_select_load_case(self.h, i), "select_load_case")

    def load_case_count(self):
        n = C.c_int32()
        self._chk(self.L.icarat_hip_load_case_count(self.h, C.byref(n)), "load_case_count")
        return n.value

    def dof(self):
        d = np.zeros(self.numnp * 3, dtype=np.int32)
        self._chk(self.L.icarat_hip_get_dof(self.h, d.ctypes.data_as(i32p)), "get_dof")
        return d.reshape(-1, 3)

    def fext(self):
        f = np.zeros(self.numeq)
        self._chk(self.L.icarat_hip_get_fext(self.h, _p(f)), "get_fext")
        return f

    def set_fext(self, f):
        self._chk(self.L.icarat_hip_set_fext(self.h, _p(_f64(f))), "set_fext")

    def ke0(self):
        k = np.zeros(576)
        self._chk(self.L.icarat_hip_get_ke0(self.h, _p(k)), "get_ke0")
        return k.reshape(24, 24)

    # --- analysis
    def set_density(self, x):
        x = _f64(x)
        self._chk(self.L.icarat_hip_set_density(self.h, _p(x), x.size), "set_density")

    def apply_K(self, u, precision=FP64):
        y = np.zeros(self.numeq)
        self._chk(self.L.icarat_hip_apply_K(self.h, _p(_f64(u)), _p(y), precision), "apply_K")
        return y

    def diag(self):
        d = np.zeros(self.numeq)
        self._chk(self.L.icarat_hip_get_diag(self.h, _p(d)), "get_diag")
        return d

    def solve(self, tol=1e-8, maxit=10000, precision=FP64, precond=PRECOND_JACOBI,
              inner_tol=0.1, max_refine=200, want_u=True, allow_cap=False, warm=False):
        o = SolveOpts(tol, maxit, precision, precond, inner_tol, max_refine, 1 if warm else 0)
        st = SolveStats()
        u = np.zeros(self.numeq) if want_u else None
        rc = self.L.icarat_hip_solve(self.h, C.byref(o), None if u is None else _p(u), C.byref(st))
        if rc and not (allow_cap and rc == NOT_CONVERGED):
            self._chk(rc, "solve")
        return u, st

    def solve_cases(self, tol=1e-8, maxit=10000, precision=FP64, precond=PRECOND_JACOBI,
                    want_u=True, allow_cap=False, warm=False):
        """Every load case (icarat_hip_solve_cases): (ncase x numeq solutions or
        None, list of per-case stats)."""
        nc = self.load_case_count()
        o = SolveOpts(tol, maxit, precision, precond, 0.1, 200, 1 if warm else 0)
        st = (SolveStats * nc)()
        u = np.zeros((nc, self.numeq)) if want_u else None
        rc = self.L.icarat_hip_solve_cases(self.h, C.byref(o), None if u is None else _p(u), st)
        if rc and not (allow_cap and rc == NOT_CONVERGED):
            self._chk(rc, "solve_cases")
        return u, list(st)

    def true_residual(self, u):
        r = C.c_double()
        self._chk(self.L.icarat_hip_true_residual(self.h, _p(_f64(u)), C.byref(r)), "true_residual")
        return r.value

    def compliance_sensitivity(self, vnorm=1.0):
        c = C.c_double()
        df = np.zeros(self.nelx)
        dg = np.zeros(self.nelx)
        self._chk(self.L.icarat_hip_compliance_sensitivity(self.h, vnorm, C.byref(c), _p(df), _p(dg)),
                  "compliance_sensitivity")
        return c.value, df, dg

    # --- optimization
    def filter_setup(self, radius):
        self._chk(self.L.icarat_hip_filter_setup(self.h, radius), "filter_setup")

    def density_filter(self, x, is_sens=False):
        out = np.zeros(self.nelx)
        self._chk(self.L.icarat_hip_filter(self.h, _p(_f64(x)), _p(out), 1 if is_sens else 0),
                  "filter")
        return out

    def sensitivity_filter(self, x, dfds):
        """Filter::sensitivityFilter(design_s, dfds) (optimization/filter.hpp:76-104)."""
        out = np.zeros(self.nelx)
        self._chk(self.L.icarat_hip_sensitivity_filter(self.h, _p(_f64(x)), _p(_f64(dfds)),
                                                       _p(out)), "sensitivity_filter")
        return out

    def mises(self):
        """Element-mean stress (nelx x 6) and von Mises stress of the last solve."""
        m = np.zeros(self.nelx)
        st = np.zeros(6 * self.nelx)
        self._chk(self.L.icarat_hip_mises(self.h, _p(m), _p(st)), "mises")
        return m, st.reshape(self.nelx, 6)

    def mises_noalloc(self):
        """As mises(), through the entry point that allocates no device memory."""
        m = np.zeros(self.nelx)
        st = np.zeros(6 * self.nelx)
        self._chk(self.L.icarat_hip_mises_noalloc(self.h, _p(m), _p(st)), "mises_noalloc")
        return m, st.reshape(self.nelx, 6)

    def gcmma_set_params(self, move=0.15, albefa=0.2, asyminit=0.2, asymdec=0.7, asyminc=1.2):
        self._chk(self.L.icarat_hip_gcmma_set_params(self.h, move, albefa, asyminit, asymdec,
                                                     asyminc), "gcmma_set_params")

    def gcmma_reset(self):
        self._chk(self.L.icarat_hip_gcmma_reset(self.h), "gcmma_reset")

    def gcmma_outer(self, xval, f0x, df0dx, fx, dfdx, xmin=1e-4, xmax=1.0):
        xm = np.zeros(self.nelx)
        self._chk(self.L.icarat_hip_gcmma_outer(self.h, _p(xm), _p(_f64(xval)), f0x,
                                                _p(_f64(df0dx)), fx, _p(_f64(dfdx)), xmin, xmax),
                  "gcmma_outer")
        return xm

    def gcmma_inner(self, xmma, f0xnew, fxnew, xval, f0x, df0dx, fx, dfdx, xmin=1e-4,
                    xmax=1.0):
        xm = np.array(xmma, dtype=np.float64)
        self._chk(self.L.icarat_hip_gcmma_inner(self.h, _p(xm), f0xnew, fxnew, _p(_f64(xval)),
                                                f0x, _p(_f64(df0dx)), fx, _p(_f64(dfdx)), xmin,
                                                xmax), "gcmma_inner")
        return xm

    def gcmma_concheck(self, f0xnew, fxnew):
        ok = C.c_int32()
        self._chk(self.L.icarat_hip_gcmma_concheck(self.h, f0xnew, fxnew, C.byref(ok)),
                  "gcmma_concheck")
        return bool(ok.value)

    def helm_setup(self, radius):
        """FilterHelmholtz(fem, element, node, radius) on this box."""
        self._chk(self.L.icarat_hip_helm_setup(self.h, radius), "helm_setup")

    def helm_filter(self, x, is_sens=False):
        """FilterHelmholtz::densityFilter(x, solver, log, is_sens)."""
        out = np.zeros(self.nelx)
        self._chk(self.L.icarat_hip_helm_filter(self.h, _p(_f64(x)), _p(out),
                                                1 if is_sens else 0), "helm_filter")
        return out

    def threshold(self, x, beta, T=0.5, deriv=False):
        out = np.zeros(self.nelx)
        self._chk(self.L.icarat_hip_threshold(self.h, _p(_f64(x)), beta, T, _p(out),
                                              1 if deriv else 0), "threshold")
        return out

    def oc_update(self, x, dfdx, g, dgdx, xmin=1e-4, xmax=1.0, use_projection=False, beta=1.0):
        x = np.array(x, dtype=np.float64)
        steps = C.c_int32()
        self._chk(self.L.icarat_hip_oc_update(self.h, _p(x), _p(_f64(dfdx)), g, _p(_f64(dgdx)),
                                              xmin, xmax, 1 if use_projection else 0, beta,
                                              C.byref(steps)), "oc_update")
        return x, steps.value

    def mma_set_params(self, move=0.15, albefa=0.2, asyminit=0.2, asymdec=0.7, asyminc=1.2):
        self._chk(self.L.icarat_hip_mma_set_params(self.h, move, albefa, asyminit, asymdec,
                                                   asyminc), "mma_set_params")

    def mma_reset(self):
        self._chk(self.L.icarat_hip_mma_reset(self.h), "mma_reset")

    def mma_update(self, x, dfdx, g, dgdx, xmin=1e-4, xmax=1.0):
        x = np.array(x, dtype=np.float64)
        outer = C.c_int32()
        self._chk(self.L.icarat_hip_mma_update(self.h, _p(x), _p(_f64(dfdx)), g, _p(_f64(dgdx)),
                                               xmin, xmax, C.byref(outer)), "mma_update")
        return x, outer.value

    # --- learned preconditioner
    def learned_setup(self, opts=None, **kw):
        o = opts if opts is not None else learned_default_opts(**kw)
        self._chk(self.L.icarat_hip_learned_setup(self.h, C.byref(o)), "learned_setup")

    def learned_harvest(self, tol=1e-8, maxit=10000, precision=FP64, want_u=False):
        o = SolveOpts(tol, maxit, precision, PRECOND_JACOBI, 0.1, 200, 0)
        st = SolveStats()
        u = np.zeros(self.numeq) if want_u else None
        self._chk(self.L.icarat_hip_learned_harvest(self.h, C.byref(o),
                                                    None if u is None else _p(u), C.byref(st)),
                  "learned_harvest")
        return u, st

    def learned_train(self, steps):
        loss = C.c_double()
        self._chk(self.L.icarat_hip_learned_train(self.h, steps, C.byref(loss)), "learned_train")
        return loss.value

    def learned_info(self):
        d = LearnedInfo()
        self._chk(self.L.icarat_hip_learned_info(self.h, C.byref(d)), "learned_info")
        out = {k: getattr(d, k) for k, _ in LearnedInfo._fields_}
        out["grid"] = tuple(d.grid)
        return out

    def learned_get_params(self):
        out = np.zeros(unet_param_count(), dtype=np.float32)
        self._chk(self.L.icarat_hip_learned_get_params(self.h, out.ctypes.data_as(fp)),
                  "learned_get_params")
        return out

    def learned_set_params(self, flat):
        a, p = _f32(flat)
        self._chk(self.L.icarat_hip_learned_set_params(self.h, p), "learned_set_params")

    def learned_apply(self, r):
        z = np.zeros(self.numeq)
        self._chk(self.L.icarat_hip_learned_apply(self.h, _p(_f64(r)), _p(z)), "learned_apply")
        return z

    def set_stencil(self, kernel=0, sym_max=2):
        """FP64 K.u kernel: 0 z-marching, 1 3-D tiles; highest KE0 symmetry
        class allowed (0 memory, 1 box, 2 cube).  Returns the class used."""
        used = C.c_int32()
        self._chk(self.L.icarat_hip_set_stencil(self.h, kernel, sym_max, C.byref(used)),
                  "set_stencil")
        return used.value

    # --- multigrid preconditioner
    def mg_setup(self, **kw):
        o = MgOpts()
        self._chk(self.L.icarat_hip_mg_default_opts(C.byref(o)), "mg_default_opts")
        for k, v in kw.items():
            setattr(o, k, v)
        self._chk(self.L.icarat_hip_mg_setup(self.h, C.byref(o)), "mg_setup")

    def mg_info(self):
        n = C.c_int32()
        dims = (C.c_int32 * 60)()
        self._chk(self.L.icarat_hip_mg_info(self.h, C.byref(n), dims), "mg_info")
        return [tuple(dims[3 * l:3 * l + 3]) for l in range(n.value)]

    def mg_status(self):
        st = MgStatus()
        self._chk(self.L.icarat_hip_mg_status(self.h, C.byref(st)), "mg_status")
        return st

    def time_kernel(self, kernel, reps=20):
        ms, fl, by = C.c_double(), C.c_double(), C.c_double()
        self._chk(self.L.icarat_hip_time_kernel(self.h, kernel, reps, C.byref(ms), C.byref(fl),
                                                C.byref(by)), "time_kernel")
        return ms.value, fl.value, by.value

    def simp_run(self, x0, n_iter, optimizer=0, use_projection=False, xmin=1e-4, xmax=1.0,
                 tol=1e-8, maxit=10000, precision=FP64, precond=PRECOND_JACOBI, inner_tol=0.1,
                 max_refine=200, warm=False, helmholtz=False, conv_eps=0.0, conv_minstep=0,
                 history_csv=None, cases_together=None):
        """The design loop; with conv_eps > 0 it may end before n_iter
        (isConvergence): hist then has st.steps rows.  cases_together: None
        automatic, True together, False one load case after another."""
        o = SimpOpts(n_iter, optimizer, 1 if use_projection else 0, 1 if helmholtz else 0,
                     xmin, xmax,
                     SolveOpts(tol, maxit, precision, precond, inner_tol, max_refine,
                               1 if warm else 0), conv_eps, conv_minstep,
                     None if history_csv is None else os.fsencode(history_csv),
                     0 if cases_together is None else (2 if cases_together else 1))
        x = np.array(x0, dtype=np.float64)
        hist = np.zeros(4 * max(n_iter, 1))
        st = SimpStats()
        self._chk(self.L.icarat_hip_simp_run(self.h, C.byref(o), _p(x), _p(hist), C.byref(st)),
                  "simp_run")
        steps = st.steps if conv_eps > 0 else n_iter
        return x, hist[:4 * steps].reshape(steps, 4), st


def cantilever(nx, ny, nz, **kw):
    """The SIMP configs' cantilever: x = 0 clamped; unit total -z load on the
    bottom edge (y = 0..ny, z = 0) of the free end x = nx as point loads
    (-1/(2 ny) at the two ends, -1/ny inside)."""
    c = Context(nx, ny, nz, **kw)
    c.add_dirichlet([1, 1, 1, 0, 0, 0], [0.0] * 6, [[0, 0], [0, ny], [0, nz]])
    c.make_dof()
    for j in range(ny + 1):
        v = -1.0 / (2 * ny) if j in (0, ny) else -1.0 / ny
        c.add_neumann([0.0, 0.0, v], [[nx, nx], [j, j], [0, 0]])
    return c


class MMASolver:
    """MMASolver(n, m, a, c, d) on the device (icarat_hip_mmasolver_*): any n,
    1 <= m <= 8 constraints, host arrays as the reference's Update."""

    def __init__(self, n, m, a=0.0, c=100000.0, d=0.0):
        self.L = lib()
        self.n, self.m = int(n), int(m)
        h = C.c_void_p()
        rc = self.L.icarat_hip_mmasolver_create(self.n, self.m, a, c, d, C.byref(h))
        if rc:
            raise IcaratError(rc, "mmasolver_create")
        self.h = h

    def close(self):
        if getattr(self, "h", None):
            self.L.icarat_hip_mmasolver_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def _chk(self, rc, what):
        if rc:
            msg = self.L.icarat_hip_mmasolver_last_error(self.h)
            raise IcaratError(rc, f"{what}: {msg.decode() if msg else ''}")

    def set_params(self, move=0.15, albefa=0.2, asyminit=0.2, asymdec=0.7, asyminc=1.2):
        self._chk(self.L.icarat_hip_mmasolver_set_params(self.h, move, albefa, asyminit, asymdec,
                                                         asyminc), "mmasolver_set_params")

    def reset(self):
        self._chk(self.L.icarat_hip_mmasolver_reset(self.h), "mmasolver_reset")

    def update(self, x, dfdx, gx, dgdx, xmin, xmax):
        """x (n) updated in place; dgdx (n x m, element-major, or flat n m);
        returns the interior-point outer loops."""
        n, m = self.n, self.m
        arr = [_f64(v).reshape(-1) for v in (dfdx, gx, dgdx, xmin, xmax)]
        if x.dtype != np.float64 or not x.flags.c_contiguous or x.size != n:
            raise ValueError("x must be a contiguous float64 array of n values")
        if [a.size for a in arr] != [n, m, n * m, n, n]:
            raise ValueError("array sizes do not match n and m")
        outer = C.c_int32()
        self._chk(self.L.icarat_hip_mmasolver_update(self.h, _p(x), *[_p(a) for a in arr],
                                                     C.byref(outer)), "mmasolver_update")
        return outer.value

    def asymptotes(self):
        low, upp = np.zeros(self.n), np.zeros(self.n)
        self._chk(self.L.icarat_hip_mmasolver_asymptotes(self.h, _p(low), _p(upp)),
                  "mmasolver_asymptotes")
        return low, upp


class GCMMASolver:
    """GCMMASolver(n, m, a, c, d) on the device (icarat_hip_gcmmasolver_*): any n,
    1 <= m <= 8 constraints, host arrays as the reference's OuterUpdate /
    InnerUpdate / ConCheck."""

    def __init__(self, n, m, a=0.0, c=1000.0, d=0.0):
        self.L = lib()
        self.n, self.m = int(n), int(m)
        h = C.c_void_p()
        rc = self.L.icarat_hip_gcmmasolver_create(self.n, self.m, a, c, d, C.byref(h))
        if rc:
            raise IcaratError(rc, "gcmmasolver_create")
        self.h = h

    def close(self):
        if getattr(self, "h", None):
            self.L.icarat_hip_gcmmasolver_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def _chk(self, rc, what):
        if rc:
            msg = self.L.icarat_hip_gcmmasolver_last_error(self.h)
            raise IcaratError(rc, f"{what}: {msg.decode() if msg else ''}")

    def _arrays(self, fx, dfdx, rest):
        n, m = self.n, self.m
        a = [_f64(v).reshape(-1) for v in rest]
        if any(v.size != n for v in a):
            raise ValueError("n-arrays of the wrong size")
        fx, dfdx = _f64(fx).reshape(-1), _f64(dfdx).reshape(-1)
        if fx.size != m or dfdx.size != n * m:
            raise ValueError("constraint arrays of the wrong size")
        return fx, dfdx, a

    def set_params(self, move=0.15, albefa=0.2, asyminit=0.2, asymdec=0.7, asyminc=1.2):
        self._chk(self.L.icarat_hip_gcmmasolver_set_params(self.h, move, albefa, asyminit,
                                                           asymdec, asyminc),
                  "gcmmasolver_set_params")

    def reset(self):
        self._chk(self.L.icarat_hip_gcmmasolver_reset(self.h), "gcmmasolver_reset")

    def outer(self, xval, f0x, df0dx, fx, dfdx, xmin, xmax):
        """OuterUpdate: returns xmma"""
        fx, dfdx, (xv, df, lo, hi) = self._arrays(fx, dfdx, (xval, df0dx, xmin, xmax))
        xm = np.zeros(self.n)
        self._chk(self.L.icarat_hip_gcmmasolver_outer(self.h, _p(xm), _p(xv), float(f0x), _p(df),
                                                      _p(fx), _p(dfdx), _p(lo), _p(hi)),
                  "gcmmasolver_outer")
        return xm

    def inner(self, xmma, f0xnew, fxnew, xval, f0x, df0dx, fx, dfdx, xmin, xmax):
        """InnerUpdate from xmma: returns the new xmma"""
        fx, dfdx, (xv, df, lo, hi) = self._arrays(fx, dfdx, (xval, df0dx, xmin, xmax))
        fxn = _f64(fxnew).reshape(-1)
        xm = np.array(xmma, dtype=np.float64)
        if xm.size != self.n or fxn.size != self.m:
            raise ValueError("xmma or fxnew of the wrong size")
        self._chk(self.L.icarat_hip_gcmmasolver_inner(self.h, _p(xm), float(f0xnew), _p(fxn),
                                                      _p(xv), float(f0x), _p(df), _p(fx),
                                                      _p(dfdx), _p(lo), _p(hi)),
                  "gcmmasolver_inner")
        return xm

    def concheck(self, f0xnew, fxnew):
        ok = C.c_int32()
        self._chk(self.L.icarat_hip_gcmmasolver_concheck(self.h, float(f0xnew),
                                                         _p(_f64(fxnew).reshape(-1)),
                                                         C.byref(ok)), "gcmmasolver_concheck")
        return bool(ok.value)

    def state(self):
        """{raa0, raa[m], f0app, fapp[m]}"""
        out = np.zeros(2 + 2 * self.m)
        self._chk(self.L.icarat_hip_gcmmasolver_state(self.h, _p(out)), "gcmmasolver_state")
        return out


class Batch:
    """Independent problems (Context objects, kept alive by the batch) solved
    together: one captured graph per PCG round with a branch per problem."""

    def __init__(self, contexts=()):
        self.L = lib()
        h = C.c_void_p()
        rc = self.L.icarat_hip_batch_create(C.byref(h))
        if rc:
            raise IcaratError(rc, "batch_create")
        self.h = h
        self.ctx = []
        for c in contexts:
            self.add(c)

    def close(self):
        if getattr(self, "h", None):
            self.L.icarat_hip_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def _chk(self, rc, what):
        if rc:
            msg = self.L.icarat_hip_batch_last_error(self.h)
            raise IcaratError(rc, f"{what}: {msg.decode() if msg else ''}")

    def add(self, c):
        self._chk(self.L.icarat_hip_batch_add(self.h, c.h), "batch_add")
        self.ctx.append(c)

    def set_budget(self, elements):
        self._chk(self.L.icarat_hip_batch_set_budget(self.h, int(elements)), "batch_set_budget")

    def info(self):
        n, k = C.c_int32(), C.c_int64()
        self._chk(self.L.icarat_hip_batch_info(self.h, C.byref(n), C.byref(k)), "batch_info")
        return n.value, k.value

    def solve(self, tol=1e-8, maxit=10000, allow_cap=False, precond=PRECOND_JACOBI, warm=False):
        o = SolveOpts(tol, maxit, FP64, precond, 0.1, 200, 1 if warm else 0)
        st = (SolveStats * len(self.ctx))()
        rc = self.L.icarat_hip_batch_solve(self.h, C.byref(o), st)
        if rc and not (allow_cap and rc == NOT_CONVERGED):
            self._chk(rc, "batch_solve")
        return list(st)

    def simp_run(self, x0s, n_iter, optimizer=0, use_projection=False, xmin=1e-4, xmax=1.0,
                 tol=1e-8, maxit=10000, precond=PRECOND_JACOBI, warm=False):
        o = SimpOpts(n_iter, optimizer, 1 if use_projection else 0, 0, xmin, xmax,
                     SolveOpts(tol, maxit, FP64, precond, 0.1, 200, 1 if warm else 0))
        xs = [np.array(x, dtype=np.float64) for x in x0s]
        hs = [np.zeros(4 * max(n_iter, 1)) for _ in xs]
        xp = (dp * len(xs))(*[_p(x) for x in xs])
        hp = (dp * len(xs))(*[_p(h) for h in hs])
        st = (SimpStats * len(xs))()
        self._chk(self.L.icarat_hip_batch_simp_run(self.h, C.byref(o), xp, hp, st), "batch_simp_run")
        return xs, [h[:4 * n_iter].reshape(n_iter, 4) for h in hs], list(st)

    def autotune(self, reps=5):
        cfg = C.c_int32()
        ms = (C.c_double * 4)()
        self._chk(self.L.icarat_hip_batch_autotune(self.h, reps, C.byref(cfg), ms), "batch_autotune")
        return cfg.value, list(ms)
