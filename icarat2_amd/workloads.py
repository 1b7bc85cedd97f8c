"""Seeded problem generators of the benchmark configurations (host side).

Batched-serving workload: 24 independent cantilever-like problems, seeds
100..123.  Each problem draws from its own numpy generator:
  - dimensions from SIZES (cubes and boxes, 32^3 ... 160x80x80, unit voxels);
  - volume fraction U[0.15, 0.4] (the uniform initial design; OC keeps the volume);
  - filter radius U[1.2, 2.5];
  - 1..4 load cases, each a unit total face load on a random boundary face
    (face-load semantics of the reference's SurfaceLoad: a traction val over
    the face, integrated with the quad4 2x2 rule), in a random direction
    tangential to the face;
  - supports: the face opposite the first load's face is clamped.
"""
import numpy as np

SIZES = [(32, 32, 32), (48, 48, 48), (64, 32, 32), (64, 64, 64), (96, 48, 48),
         (80, 80, 80), (128, 64, 64), (96, 96, 96), (160, 80, 80), (64, 48, 32)]
FACES = ["x0", "x1", "y0", "y1", "z0", "z1"]
OPPOSITE = {"x0": "x1", "x1": "x0", "y0": "y1", "y1": "y0", "z0": "z1", "z1": "z0"}


def face_range(face, dims):
    """Coordinate box [[x0,x1],[y0,y1],[z0,z1]] of a boundary face (unit voxels)."""
    nx, ny, nz = dims
    r = [[0, nx], [0, ny], [0, nz]]
    ax = "xyz".index(face[0])
    v = 0 if face[1] == "0" else dims[ax]
    r[ax] = [v, v]
    return r


def batch_problem(seed):
    """Specification of one problem of the batched workload."""
    rng = np.random.default_rng(seed)
    dims = SIZES[int(rng.integers(len(SIZES)))]
    volfrac = float(rng.uniform(0.15, 0.4))
    rmin = float(rng.uniform(1.2, 2.5))
    ncase = int(rng.integers(1, 5))
    loads = []
    for _ in range(ncase):
        face = FACES[int(rng.integers(6))]
        ax = "xyz".index(face[0])
        tangential = [a for a in range(3) if a != ax]
        d = tangential[int(rng.integers(2))]
        sign = 1.0 if rng.integers(2) else -1.0
        area = 1.0
        for a in range(3):
            if a != ax:
                area *= dims[a]
        val = [0.0, 0.0, 0.0]
        val[d] = sign / area          # unit total force
        loads.append((face, val))
    clamp = OPPOSITE[loads[0][0]]
    return {"seed": seed, "dims": dims, "volfrac": volfrac, "rmin": rmin, "loads": loads,
            "clamp": clamp}


def batch_workload(n=24, first_seed=100):
    return [batch_problem(first_seed + i) for i in range(n)]


def build_context(A, spec):
    """A prepared context (supports, load cases, filter) for one problem spec;
    A is the ctypes binding module (icarat2_amd._capi)."""
    dims = spec["dims"]
    c = A.Context(*dims)
    c.add_dirichlet([1, 1, 1, 0, 0, 0], [0.0] * 6, face_range(spec["clamp"], dims))
    c.make_dof()
    for k, (face, val) in enumerate(spec["loads"]):
        if k > 0:
            c.new_load_case()
        c.add_neumann(val, face_range(face, dims))
    c.select_load_case(0)
    c.filter_setup(spec["rmin"])
    return c


def initial_design(spec):
    n = spec["dims"][0] * spec["dims"][1] * spec["dims"][2]
    return np.full(n, spec["volfrac"])
