"""ctypes binding of the FP64 CPU oracle (``fem_oracle.cpp``).

TEST INFRASTRUCTURE ONLY (see the package docstring).
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")
_lib = None


def build() -> None:
    """Compile the oracle with make (g++)."""
    subprocess.run(["make", "-s", "-B", "-C", _HERE, "liboracle.so"], check=True)


class OracleParams(C.Structure):
    _fields_ = [
        ("alpha", C.c_double),
        ("beta", C.c_double),
        ("cg_min_iterations", C.c_int),
        ("cg_max_iterations", C.c_int),
        ("tol_mode", C.c_int),
        ("tol", C.c_double),
        ("tiny", C.c_double),
        ("rotation_mode", C.c_int),
    ]


class OracleStats(C.Structure):
    _fields_ = [
        ("iterations", C.c_int),
        ("converged", C.c_int),
        ("rr0", C.c_double),
        ("rr", C.c_double),
        ("bb", C.c_double),
    ]


def reference_params() -> OracleParams:
    """The reference's hard-coded solver settings (fem_simulate.h:104-109,
    fem_conjugate_gradients.h:43-44)."""
    return OracleParams(2.0, 0.0, 20, 20, 0, 1e-3, 1e-10, 0)


def config_params(alpha=0.1, beta=0.001, tol=1e-6, cap=200, tiny=0.0, rotation_mode=0):
    """Relative-residual settings: stop when ||b - A v|| <= tol ||b||.

    The reference's absolute 1e-10 clamps on r.r and p.Ap are off here by
    default: they are scale-dependent and stall CG once r.r < 1e-10."""
    return OracleParams(alpha, beta, 0, cap, 1, tol, tiny, rotation_mode)


def load_library(path=None):
    """The oracle library (built on first use), or the library at `path`
    (a variant of the oracle built elsewhere, e.g. with a planted mistake)."""
    global _lib
    if path is None and _lib is not None:
        return _lib
    if path is None and not os.path.exists(_LIB_PATH):
        build()
    lib = C.CDLL(_LIB_PATH if path is None else path)
    P = C.c_void_p
    dp = C.POINTER(C.c_double)
    ip = C.POINTER(C.c_int)
    lp = C.POINTER(C.c_long)
    up = C.POINTER(C.c_ubyte)
    lib.orc_create.restype = P
    lib.orc_create.argtypes = [C.c_int, C.c_int, dp, dp, dp, up, ip]
    lib.orc_destroy.argtypes = [P]
    lib.orc_init.restype = C.c_int
    lib.orc_init.argtypes = [P] + [C.c_double] * 6
    lib.orc_init_materials.restype = C.c_int
    lib.orc_init_materials.argtypes = [P, dp, dp, dp, C.c_double, C.c_double, C.c_double]
    lib.orc_simulate.restype = C.c_int
    lib.orc_simulate.argtypes = [P, C.c_double, C.c_int, C.POINTER(OracleParams), C.POINTER(OracleStats)]
    lib.orc_update_original_coord.argtypes = [P]
    lib.orc_set_state.argtypes = [P, dp, dp]
    lib.orc_get_state.argtypes = [P, dp, dp]
    lib.orc_set_fixed.argtypes = [P, up]
    lib.orc_set_external_forces.argtypes = [P, dp]
    lib.orc_get_external_forces.argtypes = [P, dp]
    lib.orc_get_mass.argtypes = [P, dp]
    lib.orc_get_volume.argtypes = [P, dp]
    lib.orc_get_rotations.argtypes = [P, dp]
    lib.orc_get_Ke.argtypes = [P, C.c_int, dp]
    lib.orc_get_plastic.argtypes = [P, dp]
    lib.orc_get_rhs.argtypes = [P, dp]
    lib.orc_get_f0.argtypes = [P, dp]
    lib.orc_row_blocks.restype = C.c_long
    lib.orc_row_blocks.argtypes = [P, C.c_int]
    lib.orc_export_rows.argtypes = [P, C.c_int, lp, ip, dp]
    lib.orc_export_K_dense.argtypes = [P, dp]
    lib.orc_compute_Ke.argtypes = [dp, C.c_double, C.c_double, dp]
    lib.orc_compute_Ke_BD.argtypes = [dp, C.c_double, C.c_double, dp]
    lib.orc_ortonormalize.argtypes = [dp, dp]
    lib.orc_polar_decomposition.restype = C.c_int
    lib.orc_polar_decomposition.argtypes = [dp, dp, dp]
    lib.orc_default_params.argtypes = [C.POINTER(OracleParams)]
    if path is None:
        _lib = lib
    return lib


def _d(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def _c64(a, shape=None):
    a = np.ascontiguousarray(a, dtype=np.float64)
    if shape is not None:
        a = a.reshape(shape)
    return a


class Oracle:
    """One mesh in the oracle.  Mirrors fem::init / fem::simulate."""

    def __init__(self, coords, tets, fixed=None, velocity=None, model_coords=None, lib=None):
        self.lib = load_library() if lib is None else lib
        coords = _c64(coords)
        self.N = coords.shape[0]
        tets = np.ascontiguousarray(tets, dtype=np.int32)
        self.T = tets.shape[0]
        mc = coords if model_coords is None else _c64(model_coords)
        vel = np.zeros_like(coords) if velocity is None else _c64(velocity)
        fx = np.zeros(self.N, np.uint8) if fixed is None else np.ascontiguousarray(fixed, dtype=np.uint8)
        self._keep = (coords, tets, mc, vel, fx)
        self.h = self.lib.orc_create(
            self.N, self.T, _d(coords), _d(mc), _d(vel),
            fx.ctypes.data_as(C.POINTER(C.c_ubyte)), tets.ctypes.data_as(C.POINTER(C.c_int)))

    def __del__(self):
        if getattr(self, "h", None):
            self.lib.orc_destroy(self.h)
            self.h = None

    def init(self, young=5e5, poisson=0.33, density=1000.0, c_yield=0.0, c_creep=0.0, c_max=0.0):
        return self.lib.orc_init(self.h, young, poisson, density, c_yield, c_creep, c_max)

    def init_materials(self, young, poisson, density, c_yield=0.0, c_creep=0.0, c_max=0.0):
        E = _c64(np.broadcast_to(young, (self.T,)))
        nu = _c64(np.broadcast_to(poisson, (self.T,)))
        rho = _c64(np.broadcast_to(density, (self.T,)))
        return self.lib.orc_init_materials(self.h, _d(E), _d(nu), _d(rho), c_yield, c_creep, c_max)

    def set_external_forces(self, f):
        f = _c64(f, (self.N, 3))
        self.lib.orc_set_external_forces(self.h, _d(f))

    def external_forces(self):
        f = np.empty((self.N, 3))
        self.lib.orc_get_external_forces(self.h, _d(f))
        return f

    def set_fixed(self, fixed):
        fx = np.ascontiguousarray(fixed, dtype=np.uint8)
        self.lib.orc_set_fixed(self.h, fx.ctypes.data_as(C.POINTER(C.c_ubyte)))

    def set_state(self, x=None, v=None):
        xp = _c64(x, (self.N, 3)) if x is not None else None
        vp = _c64(v, (self.N, 3)) if v is not None else None
        self.lib.orc_set_state(self.h, _d(xp) if xp is not None else None,
                               _d(vp) if vp is not None else None)

    def simulate(self, dt, warp=True, params=None):
        P = reference_params() if params is None else params
        st = OracleStats()
        self.lib.orc_simulate(self.h, dt, 1 if warp else 0, C.byref(P), C.byref(st))
        return st

    def update_original_coord(self):
        self.lib.orc_update_original_coord(self.h)

    def state(self):
        x = np.empty((self.N, 3))
        v = np.empty((self.N, 3))
        self.lib.orc_get_state(self.h, _d(x), _d(v))
        return x, v

    def mass(self):
        m = np.empty(self.N)
        self.lib.orc_get_mass(self.h, _d(m))
        return m

    def volumes(self):
        V = np.empty(self.T)
        self.lib.orc_get_volume(self.h, _d(V))
        return V

    def rotations(self):
        R = np.empty((self.T, 3, 3))
        self.lib.orc_get_rotations(self.h, _d(R))
        return R

    def Ke(self, t):
        K = np.empty((4, 4, 3, 3))
        self.lib.orc_get_Ke(self.h, int(t), _d(K))
        return K

    def plastic(self):
        p = np.empty((self.T, 6))
        self.lib.orc_get_plastic(self.h, _d(p))
        return p

    def rhs(self):
        b = np.empty((self.N, 3))
        self.lib.orc_get_rhs(self.h, _d(b))
        return b

    def f0(self):
        f = np.empty((self.N, 3))
        self.lib.orc_get_f0(self.h, _d(f))
        return f

    def rows(self, which=1):
        """Block-CSR of the stored K' (which=0) or A (which=1) rows."""
        nb = self.lib.orc_row_blocks(self.h, which)
        rp = np.empty(self.N + 1, dtype=np.int64)
        cols = np.empty(nb, dtype=np.int32)
        blocks = np.empty((nb, 3, 3))
        self.lib.orc_export_rows(self.h, which, rp.ctypes.data_as(C.POINTER(C.c_long)),
                                 cols.ctypes.data_as(C.POINTER(C.c_int)), _d(blocks))
        return rp, cols, blocks

    def dense(self, which=1):
        """Dense 3N x 3N copy of the stored rows (all rows and columns)."""
        rp, cols, blocks = self.rows(which)
        A = np.zeros((3 * self.N, 3 * self.N))
        for i in range(self.N):
            for k in range(rp[i], rp[i + 1]):
                j = cols[k]
                A[3 * i:3 * i + 3, 3 * j:3 * j + 3] = blocks[k]
        return A

    def export_K_dense(self):
        K = np.empty((3 * self.N, 3 * self.N))
        self.lib.orc_export_K_dense(self.h, _d(K))
        return K


def compute_Ke(points, E, nu):
    lib = load_library()
    p = _c64(points, (12,))
    K = np.empty((4, 4, 3, 3))
    lib.orc_compute_Ke(_d(p), E, nu, _d(K))
    return K


def compute_Ke_BD(points, E, nu):
    lib = load_library()
    p = _c64(points, (12,))
    K = np.empty((4, 4, 3, 3))
    lib.orc_compute_Ke_BD(_d(p), E, nu, _d(K))
    return K


def ortonormalize(A):
    lib = load_library()
    a = _c64(A, (9,))
    R = np.empty((3, 3))
    lib.orc_ortonormalize(_d(a), _d(R))
    return R


def polar_decomposition(A):
    lib = load_library()
    a = _c64(A, (9,))
    R = np.empty((3, 3))
    S = np.empty((3, 3))
    ok = lib.orc_polar_decomposition(_d(a), _d(R), _d(S))
    return bool(ok), R, S
