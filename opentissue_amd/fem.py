"""Python mirror of the OpenTissue::fem API over the otfem C ABI.

Two layers:

* :class:`Solver` — a handle on the device solver (``otfem_create`` ...),
  device-resident state, asynchronous steps.  This is the fast path.
* :class:`Mesh` with :func:`init`, :func:`simulate`,
  :func:`update_original_coord` and :func:`export_K` — the reference's
  ``fem::`` call shapes (fem_init.h:45-54, fem_simulate.h:49-54,
  fem_update_original_coord.h:28-29, fem_export_k.h:23-24): the caller edits
  node arrays on the host between steps, and every call marshals them through
  the ABI, exactly as a C++ caller of ``<otfem/fem.hpp>`` does.

The shared library is ``opentissue_amd/libotfem.so`` (built by
``__graft_entry__.build()``).  There is no fallback: if it is missing or the
device is not a gfx950, these calls raise.
"""
from __future__ import annotations

import ctypes as C
import os
from dataclasses import dataclass

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libotfem.so")
_lib = None

OK, INVALID_ARG, INVERTED_TET, NONFINITE, NOT_CONVERGED, HIP_ERROR, NO_DEVICE, STATE, OUT_OF_MEMORY = range(9)


class OtfemError(RuntimeError):
    def __init__(self, status, message):
        super().__init__(f"otfem status {status}: {message}")
        self.status = status


class Params(C.Structure):
    """otfem_params (include/otfem/otfem.h)."""
    _fields_ = [
        ("alpha", C.c_double),
        ("beta", C.c_double),
        ("cg_min_iterations", C.c_int),
        ("cg_max_iterations", C.c_int),
        ("tol_mode", C.c_int),
        ("tol", C.c_double),
        ("tiny", C.c_double),
        ("rotation_mode", C.c_int),
        ("warm_start", C.c_int),
        ("store_rotations", C.c_int),
        ("subspace", C.c_int),
    ]


class Stats(C.Structure):
    _fields_ = [
        ("iterations", C.c_int),
        ("bodies_converged", C.c_int),
        ("bodies_capped", C.c_int),
        ("flags", C.c_int),
        ("rr0", C.c_double),
        ("rr", C.c_double),
        ("bb", C.c_double),
        ("elapsed_ms", C.c_double),
    ]


class MeshDesc(C.Structure):
    _fields_ = [
        ("num_nodes", C.c_int),
        ("num_tets", C.c_int),
        ("coord", C.POINTER(C.c_double)),
        ("model_coord", C.POINTER(C.c_double)),
        ("velocity", C.POINTER(C.c_double)),
        ("fixed", C.POINTER(C.c_ubyte)),
        ("tets", C.POINTER(C.c_int32)),
        ("num_bodies", C.c_int),
        ("body_node_offsets", C.POINTER(C.c_int64)),
        ("body_tet_offsets", C.POINTER(C.c_int64)),
        ("target_tile_nodes", C.c_int),
    ]


JitterFn = C.CFUNCTYPE(C.c_int, C.c_void_p, C.c_int64, C.POINTER(C.c_double))


class BlocksDesc(C.Structure):
    """otfem_blocks_desc (include/otfem/otfem.h)."""
    _fields_ = [
        ("num_bodies", C.c_int64),
        ("dims", C.POINTER(C.c_int32)),
        ("width", C.c_double),
        ("height", C.c_double),
        ("depth", C.c_double),
        ("jitter", C.c_double),
        ("min_volume", C.c_double),
        ("max_rounds", C.c_int32),
        ("seed", C.c_uint64),
        ("draw", JitterFn),
        ("draw_user", C.c_void_p),
        ("fixed_axis", C.c_int32),
        ("fixed_side", C.c_int32),
        ("load_axis", C.c_int32),
        ("load_side", C.c_int32),
        ("traction", C.POINTER(C.c_double)),
        ("traction_per_body", C.c_int32),
    ]


# Every symbol of include/otfem/otfem.h, with (restype, argtypes).
_dp = C.POINTER(C.c_double)
_fp = C.POINTER(C.c_float)
_H = C.c_void_p
_M = C.c_void_p

WARM_PREVIOUS_VELOCITY, WARM_LEARNED = 0, 1
SPMV_F32, SPMV_F32_ACC64, SPMV_F64, SPMV_F64_STORE64 = 0, 1, 2, 3
MLP_INPUTS, MLP_OUTPUTS, MLP_PARAMS = 12, 3, 35075
SIGNATURES = {
    "otfem_version": (C.c_char_p, []),
    "otfem_status_string": (C.c_char_p, [C.c_int]),
    "otfem_default_params": (C.c_int, [C.POINTER(Params), C.c_int]),
    "otfem_device_count": (C.c_int, [C.POINTER(C.c_int)]),
    "otfem_create": (C.c_int, [C.POINTER(MeshDesc), C.POINTER(C.c_void_p)]),
    "otfem_destroy": (C.c_int, [_H]),
    "otfem_last_error": (C.c_char_p, [_H]),
    "otfem_create_last_error": (C.c_char_p, []),
    "otfem_init": (C.c_int, [_H] + [C.c_double] * 6),
    "otfem_init_materials": (C.c_int, [_H, _dp, _dp, _dp, C.c_double, C.c_double, C.c_double]),
    "otfem_set_external_forces": (C.c_int, [_H, _dp]),
    "otfem_get_external_forces": (C.c_int, [_H, _dp]),
    "otfem_set_state": (C.c_int, [_H, _dp, _dp]),
    "otfem_get_state": (C.c_int, [_H, _dp, _dp]),
    "otfem_set_fixed": (C.c_int, [_H, C.POINTER(C.c_ubyte)]),
    "otfem_get_mass": (C.c_int, [_H, _dp]),
    "otfem_get_volumes": (C.c_int, [_H, _dp]),
    "otfem_get_rotations": (C.c_int, [_H, _dp]),
    "otfem_get_plastic": (C.c_int, [_H, _dp]),
    "otfem_step": (C.c_int, [_H, C.c_double, C.c_int, C.POINTER(Params), C.POINTER(Stats)]),
    "otfem_synchronize": (C.c_int, [_H, C.POINTER(Stats)]),
    "otfem_set_timing": (C.c_int, [_H, C.c_int]),
    "otfem_get_timing": (C.c_int, [_H, _dp, _dp, _dp, C.POINTER(C.c_longlong)]),
    "otfem_get_timing_detail": (C.c_int, [_H, _dp, C.POINTER(C.c_longlong)]),
    "otfem_subspace_coefficients": (C.c_int, [_H, _dp]),
    "otfem_step_noalloc": (C.c_int, [_H, C.c_double, C.c_int, C.POINTER(Params)]),
    "otfem_iteration_totals": (C.c_int, [_H, C.POINTER(C.c_int64), C.c_int]),
    "otfem_unconverged_totals": (C.c_int, [_H, C.POINTER(C.c_int64), C.c_int]),
    "otfem_set_kernel_config": (C.c_int, [_H, C.c_int, C.c_int, C.c_int]),
    "otfem_set_deterministic": (C.c_int, [_H, C.c_int]),
    "otfem_autotune": (C.c_int, [_H, C.c_double, C.c_int, C.POINTER(Params), C.c_int,
                                 C.POINTER(C.c_int)]),
    "otfem_body_stats": (C.c_int, [_H, C.POINTER(C.c_int), _dp]),
    "otfem_body_start_residuals": (C.c_int, [_H, _dp]),
    "otfem_update_original_coord": (C.c_int, [_H]),
    "otfem_export_K": (C.c_int, [_H, C.POINTER(C.c_int64), C.POINTER(C.c_int64),
                                 C.POINTER(C.c_int32), _dp]),
    "otfem_layout_info": (C.c_int, [_H] + [C.POINTER(C.c_int64)] * 4),
    "otfem_model_create": (C.c_int, [C.c_uint64, C.POINTER(C.c_void_p)]),
    "otfem_model_destroy": (C.c_int, [_M]),
    "otfem_model_last_error": (C.c_char_p, [_M]),
    "otfem_model_get_params": (C.c_int, [_M, _fp]),
    "otfem_model_set_params": (C.c_int, [_M, _fp]),
    "otfem_model_forward": (C.c_int, [_M, _fp, C.c_int64, _fp]),
    "otfem_model_gradient": (C.c_int, [_M, _fp, _fp, C.c_int64, _dp, _fp]),
    "otfem_model_reserve": (C.c_int, [_M, C.c_int64]),
    "otfem_model_samples": (C.c_int, [_M, C.POINTER(C.c_int64), C.POINTER(C.c_int64)]),
    "otfem_model_clear_samples": (C.c_int, [_M]),
    "otfem_model_get_samples": (C.c_int, [_M, C.c_int64, C.c_int64, _fp, _fp]),
    "otfem_model_add_samples": (C.c_int, [_M, _fp, _fp, C.c_int64]),
    "otfem_model_train": (C.c_int, [_M, C.c_int, C.c_int, C.c_double, C.c_uint64, _fp]),
    "otfem_model_train_residual": (C.c_int, [_M, _H, C.c_double, C.POINTER(Params), C.c_int, C.c_int,
                                             C.c_double, C.c_uint64, _fp]),
    "otfem_model_residual_gradient": (C.c_int, [_M, _H, C.c_double, C.POINTER(Params),
                                                C.POINTER(C.c_int32), C.c_int, _dp, _fp]),
    "otfem_set_warmstart_model": (C.c_int, [_H, _M]),
    "otfem_set_recording": (C.c_int, [_H, C.c_int]),
    "otfem_save_state": (C.c_int, [_H, C.c_char_p]),
    "otfem_load_state": (C.c_int, [_H, C.c_char_p]),
    "otfem_model_save": (C.c_int, [_M, C.c_char_p]),
    "otfem_model_load": (C.c_int, [_M, C.c_char_p]),
    "otfem_surface_bind": (C.c_int, [_H, C.c_int64, _dp, _dp, C.POINTER(C.c_int32)]),
    "otfem_surface_create": (C.c_int, [_H, C.c_int64, _dp, _dp, C.POINTER(C.c_int32),
                                       C.POINTER(C.c_void_p)]),
    "otfem_surface_update": (C.c_int, [C.c_void_p, _dp, _dp]),
    "otfem_surface_set_normals": (C.c_int, [C.c_void_p, _dp]),
    "otfem_surface_update_normals": (C.c_int, [C.c_void_p, _dp]),
    "otfem_surface_last_error": (C.c_char_p, [C.c_void_p]),
    "otfem_surface_destroy": (None, [C.c_void_p]),
    "otfem_blocks_count": (C.c_int, [C.POINTER(BlocksDesc), C.POINTER(C.c_int64), C.POINTER(C.c_int64)]),
    "otfem_blocks_build": (C.c_int, [C.POINTER(BlocksDesc), _dp, C.POINTER(C.c_int32), C.POINTER(C.c_ubyte),
                                     _dp, C.POINTER(C.c_int64), C.POINTER(C.c_int64), C.POINTER(C.c_int32)]),
    "otfem_blocks_last_error": (C.c_char_p, []),
    "otfem_autotune_layout": (C.c_int, [C.POINTER(MeshDesc), _dp, _dp, _dp, _dp, C.c_double, C.c_int,
                                        C.POINTER(Params), C.c_int, C.POINTER(C.c_int32), C.c_int,
                                        C.POINTER(C.c_int32), C.POINTER(C.c_int), _dp]),
    "otfem_autotune_layout_last_error": (C.c_char_p, []),
}


def load_library():
    """Load libotfem.so; raises if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise OtfemError(NO_DEVICE, f"{LIB_PATH} is missing: run __graft_entry__.build()")
    lib = C.CDLL(LIB_PATH)
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def default_params(mode=0) -> Params:
    p = Params()
    load_library().otfem_default_params(C.byref(p), mode)
    return p


def reference_params() -> Params:
    """The reference's hard-coded settings: alpha 2, beta 0, 20 CG iterations."""
    return default_params(0)


def config_params(alpha=0.1, beta=0.001, tol=1e-6, cap=200, rotation_mode=0, store_rotations=0,
                  subspace=0) -> Params:
    """Relative-residual settings.  `subspace` (0..7) turns on the subspace
    CG start over that many past velocities (0: start from v^n)."""
    p = default_params(1)
    p.alpha, p.beta, p.tol, p.cg_max_iterations = alpha, beta, tol, cap
    p.rotation_mode, p.store_rotations = rotation_mode, store_rotations
    p.subspace = subspace
    return p


def _dptr(a):
    return a.ctypes.data_as(_dp)


def _f64(a, shape):
    return np.ascontiguousarray(np.asarray(a, dtype=np.float64).reshape(shape))


def _mesh_desc(coords, tets, fixed=None, velocity=None, model_coords=None, body_node_offsets=None,
               body_tet_offsets=None, target_tile_nodes=0):
    """(otfem_mesh_desc, arrays it points into) for a mesh or a CSR-of-bodies
    batch; keep the arrays alive while the description is used."""
    coords = _f64(coords, (-1, 3))
    N = coords.shape[0]
    tets = np.ascontiguousarray(tets, dtype=np.int32).reshape(-1, 4)
    mc = None if model_coords is None else _f64(model_coords, (N, 3))
    vel = None if velocity is None else _f64(velocity, (N, 3))
    fx = None if fixed is None else np.ascontiguousarray(fixed, dtype=np.uint8)
    d = MeshDesc()
    d.num_nodes, d.num_tets = N, tets.shape[0]
    d.coord = _dptr(coords)
    d.model_coord = _dptr(mc) if mc is not None else None
    d.velocity = _dptr(vel) if vel is not None else None
    d.fixed = fx.ctypes.data_as(C.POINTER(C.c_ubyte)) if fx is not None else None
    d.tets = tets.ctypes.data_as(C.POINTER(C.c_int32))
    bno = bto = None
    if body_node_offsets is not None:
        bno = np.ascontiguousarray(body_node_offsets, dtype=np.int64)
        bto = np.ascontiguousarray(body_tet_offsets, dtype=np.int64)
        d.num_bodies = len(bno) - 1
        d.body_node_offsets = bno.ctypes.data_as(C.POINTER(C.c_int64))
        d.body_tet_offsets = bto.ctypes.data_as(C.POINTER(C.c_int64))
    d.target_tile_nodes = int(target_tile_nodes)
    return d, (coords, tets, mc, fx, vel, bno, bto)


def autotune_layout(coords, tets, fixed, young, poisson, density, forces, dt, params=None, warp=True,
                    steps=2, candidates=None, body_node_offsets=None, body_tet_offsets=None):
    """``otfem_autotune_layout``: (target_tile_nodes, cg_block, ms) of the
    fastest candidate layout, each timed on a throw-away solver with the
    block-size sweep of :meth:`Solver.autotune`."""
    lib = load_library()
    d, keep = _mesh_desc(coords, tets, fixed, None, None, body_node_offsets, body_tet_offsets)
    T = d.num_tets
    E = _f64(np.broadcast_to(np.asarray(young, dtype=np.float64), (T,)), (T,))
    nu = _f64(np.broadcast_to(np.asarray(poisson, dtype=np.float64), (T,)), (T,))
    rho = _f64(np.broadcast_to(np.asarray(density, dtype=np.float64), (T,)), (T,))
    f = None if forces is None else _f64(forces, (d.num_nodes, 3))
    P = params if params is not None else default_params(0)
    cand = None if candidates is None else np.ascontiguousarray(candidates, dtype=np.int32)
    tile, block, ms = C.c_int32(), C.c_int(), C.c_double()
    st = lib.otfem_autotune_layout(C.byref(d), _dptr(E), _dptr(nu), _dptr(rho),
                                   _dptr(f) if f is not None else None, float(dt), 1 if warp else 0,
                                   C.byref(P), int(steps),
                                   cand.ctypes.data_as(C.POINTER(C.c_int32)) if cand is not None else None,
                                   0 if cand is None else len(cand), C.byref(tile), C.byref(block),
                                   C.byref(ms))
    if st != OK:
        raise OtfemError(st, lib.otfem_autotune_layout_last_error().decode())
    return tile.value, block.value, ms.value


class Solver:
    """Device solver for one mesh or a CSR-of-bodies batch."""

    def __init__(self, coords, tets, fixed=None, velocity=None, model_coords=None,
                 body_node_offsets=None, body_tet_offsets=None, target_tile_nodes=0):
        self.lib = load_library()
        d, keep = _mesh_desc(coords, tets, fixed, velocity, model_coords, body_node_offsets,
                             body_tet_offsets, target_tile_nodes)
        self.N, self.T, self.B = d.num_nodes, d.num_tets, max(1, d.num_bodies)
        fx = keep[3]
        self.n_free = self.N if fx is None else int((fx == 0).sum())
        h = C.c_void_p()
        st = self.lib.otfem_create(C.byref(d), C.byref(h))  # copies what d points to
        del keep
        if st != OK:
            raise OtfemError(st, "otfem_create failed: "
                             + (self.lib.otfem_create_last_error() or b"").decode())
        self.h = h

    def close(self):
        if getattr(self, "h", None):
            self.lib.otfem_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _check(self, st):
        if st not in (OK,):
            msg = self.lib.otfem_last_error(self.h)
            raise OtfemError(st, msg.decode() if msg else "")
        return st

    # fem::init
    def init(self, young=5e5, poisson=0.33, density=1000.0, c_yield=0.0, c_creep=0.0, c_max=0.0):
        self._check(self.lib.otfem_init(self.h, young, poisson, density, c_yield, c_creep, c_max))

    def init_materials(self, young, poisson, density, c_yield=0.0, c_creep=0.0, c_max=0.0):
        E = _f64(np.broadcast_to(young, (self.T,)), (self.T,))
        nu = _f64(np.broadcast_to(poisson, (self.T,)), (self.T,))
        rho = _f64(np.broadcast_to(density, (self.T,)), (self.T,))
        self._check(self.lib.otfem_init_materials(self.h, _dptr(E), _dptr(nu), _dptr(rho),
                                                  c_yield, c_creep, c_max))

    def set_external_forces(self, f):
        f = _f64(f, (self.N, 3))
        self._check(self.lib.otfem_set_external_forces(self.h, _dptr(f)))

    def external_forces(self):
        f = np.empty((self.N, 3))
        self._check(self.lib.otfem_get_external_forces(self.h, _dptr(f)))
        return f

    def set_state(self, x=None, v=None):
        xp = _f64(x, (self.N, 3)) if x is not None else None
        vp = _f64(v, (self.N, 3)) if v is not None else None
        self._check(self.lib.otfem_set_state(self.h, _dptr(xp) if xp is not None else None,
                                             _dptr(vp) if vp is not None else None))

    def state(self):
        x = np.empty((self.N, 3))
        v = np.empty((self.N, 3))
        self._check(self.lib.otfem_get_state(self.h, _dptr(x), _dptr(v)))
        return x, v

    def set_fixed(self, fixed):
        fx = np.ascontiguousarray(fixed, dtype=np.uint8)
        self.n_free = int((fx == 0).sum())
        self._check(self.lib.otfem_set_fixed(self.h, fx.ctypes.data_as(C.POINTER(C.c_ubyte))))

    def mass(self):
        m = np.empty(self.N)
        self._check(self.lib.otfem_get_mass(self.h, _dptr(m)))
        return m

    def volumes(self):
        V = np.empty(self.T)
        self._check(self.lib.otfem_get_volumes(self.h, _dptr(V)))
        return V

    def rotations(self):
        R = np.empty((self.T, 3, 3))
        self._check(self.lib.otfem_get_rotations(self.h, _dptr(R)))
        return R

    def plastic(self):
        p = np.empty((self.T, 6))
        self._check(self.lib.otfem_get_plastic(self.h, _dptr(p)))
        return p

    # fem::simulate
    def simulate(self, dt, warp=True, params=None):
        P = reference_params() if params is None else params
        st = Stats()
        self._check(self.lib.otfem_step(self.h, dt, 1 if warp else 0, C.byref(P), C.byref(st)))
        return st

    def step_async(self, dt, warp=True, params=None):
        P = reference_params() if params is None else params
        self._check(self.lib.otfem_step(self.h, dt, 1 if warp else 0, C.byref(P), None))

    def synchronize(self):
        st = Stats()
        self._check(self.lib.otfem_synchronize(self.h, C.byref(st)))
        return st

    def set_timing(self, enable=True):
        self._check(self.lib.otfem_set_timing(self.h, 1 if enable else 0))

    def timing(self):
        """(element ms, CG ms, steps) accumulated since set_timing."""
        e, c, w, n = self.timing4()
        return e, c, n

    def timing4(self):
        """(element ms, CG ms, learned-start network ms, steps)."""
        e = C.c_double()
        c = C.c_double()
        w = C.c_double()
        n = C.c_longlong()
        self._check(self.lib.otfem_get_timing(self.h, C.byref(e), C.byref(c), C.byref(w), C.byref(n)))
        return e.value, c.value, w.value, n.value

    def subspace_coefficients(self):
        """[B, 12] coefficients of the last subspace start (v^n's first, then
        the differences from the newest, then the network's direction)."""
        c = np.zeros((self.B, 12))
        self._check(self.lib.otfem_subspace_coefficients(self.h, _dptr(c)))
        return c

    def timing_detail(self):
        """dict of accumulated ms (network, element, subspace start, CG) and steps."""
        ms = np.zeros(4)
        n = C.c_longlong()
        self._check(self.lib.otfem_get_timing_detail(self.h, _dptr(ms), C.byref(n)))
        return dict(network=ms[0], element=ms[1], subspace=ms[2], cg=ms[3], steps=n.value)

    def iteration_totals(self, reset=False):
        tot = np.empty(self.B, dtype=np.int64)
        self._check(self.lib.otfem_iteration_totals(self.h, tot.ctypes.data_as(C.POINTER(C.c_int64)),
                                                    1 if reset else 0))
        return tot

    def unconverged_totals(self, reset=False):
        tot = np.empty(self.B, dtype=np.int64)
        self._check(self.lib.otfem_unconverged_totals(self.h, tot.ctypes.data_as(C.POINTER(C.c_int64)),
                                                      1 if reset else 0))
        return tot

    def set_kernel_config(self, cg_block=256, spmv_mode=SPMV_F64, use_graphs=True):
        """Threads per tile in the CG kernels (256, 512, 1024) and the SpMV
        precision (SPMV_*; the library's default is SPMV_F64)."""
        self._check(self.lib.otfem_set_kernel_config(self.h, cg_block, spmv_mode, 1 if use_graphs else 0))

    def set_deterministic(self, on=True):
        """Per-node sums in a fixed order: the same bits on every run
        (``otfem_set_deterministic``)."""
        self._check(self.lib.otfem_set_deterministic(self.h, 1 if on else 0))

    def autotune(self, dt, warp=True, params=None, steps=2):
        """Pick the fastest CG block size for this mesh (state is restored)."""
        P = reference_params() if params is None else params
        best = C.c_int()
        self._check(self.lib.otfem_autotune(self.h, dt, 1 if warp else 0, C.byref(P), steps,
                                            C.byref(best)))
        return best.value

    def body_stats(self):
        it = np.empty(self.B, dtype=np.int32)
        rel = np.empty(self.B)
        self._check(self.lib.otfem_body_stats(self.h, it.ctypes.data_as(C.POINTER(C.c_int)), _dptr(rel)))
        return it, rel

    def start_residuals(self):
        """Per body, ||b - A v_start|| / ||b|| of the last step's CG start."""
        rel = np.empty(self.B)
        self._check(self.lib.otfem_body_start_residuals(self.h, _dptr(rel)))
        return rel

    def update_original_coord(self):
        self._check(self.lib.otfem_update_original_coord(self.h))

    def export_K(self):
        """Block-CSR (row_ptr, cols, blocks[nnz,3,3]) of fem::export_K."""
        nnz = C.c_int64()
        self._check(self.lib.otfem_export_K(self.h, C.byref(nnz), None, None, None))
        rp = np.empty(self.N + 1, dtype=np.int64)
        cols = np.empty(nnz.value, dtype=np.int32)
        blocks = np.empty((nnz.value, 3, 3))
        self._check(self.lib.otfem_export_K(
            self.h, C.byref(nnz), rp.ctypes.data_as(C.POINTER(C.c_int64)),
            cols.ctypes.data_as(C.POINTER(C.c_int32)), _dptr(blocks)))
        return rp, cols, blocks

    def save_state(self, path):
        """Checkpoint of the simulation state (``otfem_save_state``)."""
        self._check(self.lib.otfem_save_state(self.h, os.fsencode(path)))

    def load_state(self, path):
        """Restore a checkpoint of this mesh (``otfem_load_state``)."""
        self._check(self.lib.otfem_load_state(self.h, os.fsencode(path)))

    def surface_bind(self, vertices):
        """t4mesh_mesh_coupling.h:100-200: (barycentric [n, 3], tet_index [n])
        of surface vertices [n, 3] at the current node positions; index -1
        where no tetrahedron contains the vertex."""
        v = _f64(vertices, (-1, 3))
        n = v.shape[0]
        w = np.zeros((n, 3))
        idx = np.zeros(n, dtype=np.int32)
        self._check(self.lib.otfem_surface_bind(self.h, n, _dptr(v), _dptr(w),
                                                idx.ctypes.data_as(C.POINTER(C.c_int32))))
        return w, idx

    def layout_info(self):
        vals = [C.c_int64() for _ in range(4)]
        self._check(self.lib.otfem_layout_info(self.h, *[C.byref(v) for v in vals]))
        return dict(tiles=vals[0].value, slots=vals[1].value, halo=vals[2].value,
                    max_tile_nodes=vals[3].value)

    # learned warm start
    def set_warmstart_model(self, model):
        """Attach a :class:`Model` (None detaches).  Keeps a reference."""
        self._check(self.lib.otfem_set_warmstart_model(self.h, model.h if model is not None else None))
        self._model = model

    def set_recording(self, on=True):
        self._check(self.lib.otfem_set_recording(self.h, 1 if on else 0))


def _f32(a, shape):
    return np.ascontiguousarray(np.asarray(a, dtype=np.float32).reshape(shape))


class Surface:
    """A bound surface on the device (``otfem_surface_*``): update() returns
    the vertex positions from the solver's current node positions
    (t4mesh_mesh_coupling.h:202-238).  Close it before the solver."""

    def __init__(self, solver, vertices, barycentric, tet_index):
        self.lib = load_library()
        self.solver = solver
        v = _f64(vertices, (-1, 3))
        self.n = v.shape[0]
        w = _f64(barycentric, (self.n, 3))
        idx = np.ascontiguousarray(np.asarray(tet_index, dtype=np.int32).reshape(self.n))
        h = C.c_void_p()
        solver._check(self.lib.otfem_surface_create(solver.h, self.n, _dptr(v), _dptr(w),
                                                    idx.ctypes.data_as(C.POINTER(C.c_int32)),
                                                    C.byref(h)))
        self.h = h
        self.kernel_ms = 0.0

    def update(self):
        out = np.zeros((self.n, 3))
        ms = C.c_double()
        st = self.lib.otfem_surface_update(self.h, _dptr(out), C.byref(ms))
        if st != OK:
            msg = self.lib.otfem_surface_last_error(self.h)
            raise OtfemError(st, msg.decode() if msg else "")
        self.kernel_ms = ms.value
        return out

    def _surface_check(self, st):
        if st != OK:
            msg = self.lib.otfem_surface_last_error(self.h)
            raise OtfemError(st, msg.decode() if msg else "")

    def set_normals(self, normals):
        """Rest normals n0 [n, 3] of the vertices."""
        n0 = _f64(normals, (self.n, 3))
        self._surface_check(self.lib.otfem_surface_set_normals(self.h, _dptr(n0)))

    def update_normals(self):
        """n = Re n0 per bound vertex (t4mesh_mesh_coupling.h:171-176), Re of
        its tetrahedron from the last warped step with store_rotations."""
        out = np.zeros((self.n, 3))
        self._surface_check(self.lib.otfem_surface_update_normals(self.h, _dptr(out)))
        return out

    def close(self):
        if getattr(self, "h", None):
            self.lib.otfem_surface_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


@dataclass
class Blocks:
    """Lattice bodies from :func:`build_blocks` (CSR of bodies)."""
    coords: np.ndarray              # [N, 3] float64
    tets: np.ndarray                # [T, 4] int32, global node ids
    fixed: np.ndarray               # [N] uint8
    f_ext: np.ndarray               # [N, 3] float64 traction loads
    body_node_offsets: np.ndarray   # [B+1] int64
    body_tet_offsets: np.ndarray    # [B+1] int64
    rounds: int                     # jitter rounds taken


_AXIS = {None: -1, "x": 0, "y": 1, "z": 2}


def _face(face):
    """(axis, side) of a face name such as "x-", "x+", "z-"; None: no face."""
    if face is None:
        return -1, 0
    if len(face) != 2 or face[0] not in "xyz" or face[1] not in "+-":
        raise ValueError("face must be one of x-, x+, y-, y+, z-, z+ or None")
    return _AXIS[face[0]], 1 if face[1] == "+" else 0


def build_blocks(dims, width, height=None, depth=None, jitter=0.0, min_volume=0.0, max_rounds=200,
                 seed=0, draw=None, fixed_face=None, load_face=None, traction=None) -> Blocks:
    """Lattice bodies built on the device (``otfem_blocks_build``):
    ``t4mesh::generate_blocks`` (t4mesh_block_generator.h:32-110) for every
    row (I, J, K) of ``dims``, packed as CSR of bodies, with optional jitter
    of interior nodes (volume floor ``min_volume``, redrawn round after
    round), a fixed face and a traction load face (``"x-"``, ``"x+"``, ...).

    ``draw(n)`` returns the [n, 3] displacements of the n nodes a round
    moves, in ascending node order; without it the library's own
    counter-hash source (``seed``, amplitude ``jitter`` times the edge) is
    used.  ``traction`` is one 3-vector for all bodies or one per body."""
    lib = load_library()
    dims = np.ascontiguousarray(np.asarray(dims, dtype=np.int32).reshape(-1, 3))
    d = BlocksDesc()
    d.num_bodies = dims.shape[0]
    d.dims = dims.ctypes.data_as(C.POINTER(C.c_int32))
    d.width = float(width)
    d.height = float(width if height is None else height)
    d.depth = float(width if depth is None else depth)
    d.jitter = float(jitter)
    d.min_volume = float(min_volume)
    d.max_rounds = int(max_rounds)
    d.seed = int(seed)
    d.fixed_axis, d.fixed_side = _face(fixed_face)
    d.load_axis, d.load_side = _face(load_face)
    tr = None
    if traction is not None:
        tr = np.ascontiguousarray(np.asarray(traction, dtype=np.float64))
        if tr.size not in (3, 3 * dims.shape[0]):
            raise ValueError("traction must be one 3-vector or one per body")
        d.traction = _dptr(tr)
        d.traction_per_body = 1 if tr.size != 3 else 0
    failure = []

    def _draw(_user, n, out):
        try:
            v = np.ascontiguousarray(np.asarray(draw(int(n)), dtype=np.float64).reshape(int(n), 3))
            C.memmove(out, v.ctypes.data, v.nbytes)
            return 0
        except Exception as e:  # reported after the call returns
            failure.append(e)
            return 1

    cb = JitterFn(_draw) if draw is not None else JitterFn()
    d.draw = cb
    N, T = C.c_int64(), C.c_int64()
    st = lib.otfem_blocks_count(C.byref(d), C.byref(N), C.byref(T))
    if st != OK:
        raise OtfemError(st, lib.otfem_blocks_last_error().decode())
    N, T, B = N.value, T.value, dims.shape[0]
    coords = np.empty((N, 3))
    tets = np.empty((T, 4), dtype=np.int32)
    fixed = np.empty(N, dtype=np.uint8)
    f = np.empty((N, 3))
    no = np.empty(B + 1, dtype=np.int64)
    to = np.empty(B + 1, dtype=np.int64)
    rounds = C.c_int32()
    st = lib.otfem_blocks_build(C.byref(d), _dptr(coords), tets.ctypes.data_as(C.POINTER(C.c_int32)),
                                fixed.ctypes.data_as(C.POINTER(C.c_ubyte)), _dptr(f),
                                no.ctypes.data_as(C.POINTER(C.c_int64)),
                                to.ctypes.data_as(C.POINTER(C.c_int64)), C.byref(rounds))
    if failure:
        raise failure[0]
    if st != OK:
        raise OtfemError(st, lib.otfem_blocks_last_error().decode())
    return Blocks(coords, tets, fixed, f, no, to, rounds.value)


class Model:
    """Per-node MLP of the learned CG warm start (``otfem_model_*``).

    Parameters are FP32 masters in the order W1[128,12], b1, W2[128,128], b2,
    W3[128,128], b3, W4[3,128], b4; the device kernels run BF16 copies."""

    def __init__(self, seed=0):
        self.lib = load_library()
        h = C.c_void_p()
        st = self.lib.otfem_model_create(seed, C.byref(h))
        if st != OK:
            raise OtfemError(st, "otfem_model_create failed")
        self.h = h

    def close(self):
        if getattr(self, "h", None):
            self.lib.otfem_model_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _check(self, st):
        if st != OK:
            msg = self.lib.otfem_model_last_error(self.h)
            raise OtfemError(st, msg.decode() if msg else "")
        return st

    def save(self, path):
        """Parameters, Adam moments and step count (``otfem_model_save``)."""
        self._check(self.lib.otfem_model_save(self.h, os.fsencode(path)))

    def load(self, path):
        self._check(self.lib.otfem_model_load(self.h, os.fsencode(path)))

    def params(self):
        th = np.empty(MLP_PARAMS, dtype=np.float32)
        self._check(self.lib.otfem_model_get_params(self.h, th.ctypes.data_as(_fp)))
        return th

    def set_params(self, theta):
        th = _f32(theta, (MLP_PARAMS,))
        self._check(self.lib.otfem_model_set_params(self.h, th.ctypes.data_as(_fp)))

    def forward(self, features):
        X = _f32(features, (-1, MLP_INPUTS))
        Y = np.empty((X.shape[0], MLP_OUTPUTS), dtype=np.float32)
        self._check(self.lib.otfem_model_forward(self.h, X.ctypes.data_as(_fp), X.shape[0],
                                                 Y.ctypes.data_as(_fp)))
        return Y

    def gradient(self, features, targets):
        """(loss, grad[MLP_PARAMS]) from the device training kernels."""
        X = _f32(features, (-1, MLP_INPUTS))
        T = _f32(targets, (X.shape[0], 4))
        g = np.empty(MLP_PARAMS, dtype=np.float32)
        loss = C.c_double()
        self._check(self.lib.otfem_model_gradient(self.h, X.ctypes.data_as(_fp), T.ctypes.data_as(_fp),
                                                  X.shape[0], C.byref(loss), g.ctypes.data_as(_fp)))
        return loss.value, g

    def reserve(self, capacity):
        self._check(self.lib.otfem_model_reserve(self.h, int(capacity)))

    def num_samples(self):
        n = C.c_int64()
        self._check(self.lib.otfem_model_samples(self.h, C.byref(n), None))
        return n.value

    def clear_samples(self):
        self._check(self.lib.otfem_model_clear_samples(self.h))

    def samples(self, offset=0, n=None):
        if n is None:
            n = self.num_samples() - offset
        X = np.empty((n, MLP_INPUTS), dtype=np.float32)
        T = np.empty((n, 4), dtype=np.float32)
        self._check(self.lib.otfem_model_get_samples(self.h, offset, n, X.ctypes.data_as(_fp),
                                                     T.ctypes.data_as(_fp)))
        return X, T

    def add_samples(self, features, targets):
        X = _f32(features, (-1, MLP_INPUTS))
        T = _f32(targets, (X.shape[0], 4))
        self._check(self.lib.otfem_model_add_samples(self.h, X.ctypes.data_as(_fp), T.ctypes.data_as(_fp),
                                                     X.shape[0]))

    def train_residual(self, solver, dt, params=None, steps=1000, steps_per_batch=None, lr=1e-3,
                       seed=0):
        """Adam on the residual loss over steps recorded by `solver`; the
        default minibatch is the whole steps closest to 65,536 nodes."""
        if steps_per_batch is None:
            steps_per_batch = max(1, round(65536 / max(1, int(solver.n_free))))
        P = config_params() if params is None else params
        loss = np.empty(max(steps, 1), dtype=np.float32)
        st = self.lib.otfem_model_train_residual(self.h, solver.h, dt, C.byref(P), steps,
                                                 steps_per_batch, lr, seed, loss.ctypes.data_as(_fp))
        if st != OK:
            raise OtfemError(st, (self.lib.otfem_model_last_error(self.h) or b"").decode())
        return loss[:steps]

    def residual_gradient(self, solver, dt, steps, params=None):
        P = config_params() if params is None else params
        st_idx = np.ascontiguousarray(steps, dtype=np.int32)
        g = np.empty(MLP_PARAMS, dtype=np.float32)
        loss = C.c_double()
        st = self.lib.otfem_model_residual_gradient(self.h, solver.h, dt, C.byref(P),
                                                    st_idx.ctypes.data_as(C.POINTER(C.c_int32)),
                                                    len(st_idx), C.byref(loss), g.ctypes.data_as(_fp))
        if st != OK:
            raise OtfemError(st, (self.lib.otfem_model_last_error(self.h) or b"").decode())
        return loss.value, g

    def train(self, steps, batch=65536, lr=1e-3, seed=0):
        loss = np.empty(max(steps, 1), dtype=np.float32)
        self._check(self.lib.otfem_model_train(self.h, steps, batch, lr, seed, loss.ctypes.data_as(_fp)))
        return loss[:steps]


# ---------------------------------------------------------------------------
# Reference-shaped API: fem::Mesh + free functions.


@dataclass
class Mesh:
    """Host mirror of ``fem::Mesh`` node/tet trait fields used by a caller
    (fem_node_traits.h:37-52, fem_tetrahedron_traits.h:29-35)."""
    coord: np.ndarray            # m_coord       [N,3]
    tets: np.ndarray             # i, j, k, m    [T,4]
    model_coord: np.ndarray = None   # m_model_coord
    velocity: np.ndarray = None      # m_velocity
    fixed: np.ndarray = None         # m_fixed
    f_external: np.ndarray = None    # m_f_external
    young: np.ndarray = None         # m_young (per tet), for a custom init
    poisson: np.ndarray = None
    density: np.ndarray = None
    mass: np.ndarray = None          # m_mass (filled by init)
    Re: np.ndarray = None            # m_Re (filled by simulate when warping)
    solver: Solver = None

    def __post_init__(self):
        self.coord = np.array(self.coord, dtype=np.float64).reshape(-1, 3)
        self.tets = np.ascontiguousarray(self.tets, dtype=np.int32).reshape(-1, 4)
        n = self.coord.shape[0]
        if self.model_coord is None:
            self.model_coord = self.coord.copy()
        if self.velocity is None:
            self.velocity = np.zeros((n, 3))
        if self.fixed is None:
            self.fixed = np.zeros(n, dtype=bool)
        if self.f_external is None:
            self.f_external = np.zeros((n, 3))

    def size_nodes(self):
        return self.coord.shape[0]

    def size_tetrahedra(self):
        return self.tets.shape[0]


def init(mesh: Mesh, young, poisson, density, c_yield=0.0, c_creep=0.0, c_max=0.0):
    """fem::init (fem_init.h:45-65); per-tet materials if mesh.young etc. are set
    and young is None (the user-written init of fem_init.h:32-35)."""
    if mesh.solver is not None:
        mesh.solver.close()
    mesh.solver = Solver(mesh.coord, mesh.tets, mesh.fixed, mesh.velocity, mesh.model_coord)
    if young is None:
        mesh.solver.init_materials(mesh.young, mesh.poisson, mesh.density, c_yield, c_creep, c_max)
    else:
        mesh.solver.init(young, poisson, density, c_yield, c_creep, c_max)
    mesh.mass = mesh.solver.mass()


def simulate(mesh: Mesh, time_step, use_stiffness_warping=True, params=None):
    """fem::simulate (fem_simulate.h:49-113).  Host fields are uploaded, one
    step is taken, and coord / velocity / f_external / Re are read back."""
    if mesh.solver is None:
        raise OtfemError(STATE, "fem.init must be called before fem.simulate")
    s = mesh.solver
    s.set_fixed(mesh.fixed)
    s.set_state(mesh.coord, mesh.velocity)
    s.set_external_forces(mesh.f_external)
    P = reference_params() if params is None else params
    P.store_rotations = 1
    st = s.simulate(time_step, use_stiffness_warping, P)
    mesh.coord[:], mesh.velocity[:] = s.state()
    mesh.f_external[:] = s.external_forces()
    mesh.Re = s.rotations()
    return st


def update_original_coord(mesh: Mesh):
    """fem::update_original_coord (fem_update_original_coord.h:28-33)."""
    mesh.model_coord[:] = mesh.coord
    if mesh.solver is not None:
        mesh.solver.set_state(mesh.coord, None)
        mesh.solver.update_original_coord()


def export_K(mesh: Mesh):
    """fem::export_K (fem_export_k.h:23-68) as a dense 3N x 3N array."""
    rp, cols, blocks = mesh.solver.export_K()
    N = mesh.size_nodes()
    K = np.zeros((3 * N, 3 * N))
    for i in range(N):
        for k in range(rp[i], rp[i + 1]):
            j = cols[k]
            K[3 * i:3 * i + 3, 3 * j:3 * j + 3] = blocks[k]
    return K
