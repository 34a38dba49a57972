"""Concrete, seeded workloads for the FEM path.

Each builder returns a :class:`Workload` with the mesh (possibly a
CSR-of-bodies batch), fixed mask, per-tet materials, external forces and the
stepping parameters.  Unspecified quantities default to the single-lattice
case's values: edge h = 0.01 m, rho = 1000 kg/m^3, Rayleigh alpha = 0.1,
beta = 0.001, traction direction -z.

* ``cantilever``  10x10x40 cubes (I=10, J=40, K=10 so the long axis is world
  z), E 5e5, nu 0.33, gravity, nodes at z = 0 fixed, dt 0.01.
* ``cube64``      64^3 cubes, interior jitter +/-0.2 h, E log-uniform per tet
  in [1e5, 1e6], nu 0.3, face x = 0 fixed, 1e3 Pa traction on x = max,
  dt 0.005.
* ``bodies``      many jittered lattice bodies with per-axis dims uniform in
  [lo, hi] cubes, per-body E log-uniform [1e4, 1e6] and nu uniform
  [0.25, 0.45], fixed face + traction per body, dt 0.01.

``cube64`` and ``bodies`` take ``builder``: ``"numpy"`` builds the lattice,
jitter, fixed face and loads on the host (``mesh``), ``"device"`` on the GPU
(``fem.build_blocks``) with the same numpy random streams, so both give the
same coordinates, tetrahedra and fixed masks bit for bit and the same loads
to rounding (their per-node sums run in another order); ``"auto"`` (the
default) takes the device when there is one.
"""
from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np

from . import mesh as M


def _use_device(builder):
    if builder == "numpy":
        return False
    if builder == "device":
        return True
    if builder != "auto":
        raise ValueError("builder must be 'auto', 'numpy' or 'device'")
    import ctypes as C

    from . import fem
    try:
        lib = fem.load_library()
    except (OSError, fem.OtfemError):  # library not built: host builders
        return False
    c = C.c_int()
    lib.otfem_device_count(C.byref(c))
    return c.value > 0


@dataclass
class Workload:
    name: str
    coords: np.ndarray
    tets: np.ndarray
    fixed: np.ndarray
    young: np.ndarray
    poisson: np.ndarray
    density: np.ndarray
    forces: np.ndarray
    dt: float
    steps: int
    body_node_offsets: np.ndarray = None
    body_tet_offsets: np.ndarray = None
    meta: dict = field(default_factory=dict)

    @property
    def num_nodes(self):
        return self.coords.shape[0]

    @property
    def num_tets(self):
        return self.tets.shape[0]

    @property
    def num_bodies(self):
        return 1 if self.body_node_offsets is None else len(self.body_node_offsets) - 1


def cantilever(I=10, J=40, K=10, h=0.01, E=5e5, nu=0.33, rho=1000.0, g=-9.81):
    c, t = M.generate_blocks(I, J, K, h, h, h)
    fixed = (np.abs(c[:, 2]) < 1e-12 * max(h, 1.0)).astype(np.uint8)
    T = t.shape[0]
    mass = M.lumped_mass(c, t, rho)
    f = M.gravity_forces(np.where(fixed == 1, 0.0, mass), fixed, (0.0, 0.0, g))
    return Workload("cantilever_%dx%dx%d" % (I, J, K), c, t, fixed, np.full(T, E), np.full(T, nu),
                    np.full(T, rho), f, 0.01, 100, meta=dict(h=h, gravity=g))


def _jitter_draw(seed, amplitude):
    """The displacements mesh.jitter_interior / jitter_bodies draw, round by
    round, as a random source for fem.build_blocks."""
    rng = np.random.default_rng(seed)
    return lambda n: rng.uniform(-amplitude, amplitude, size=(n, 3))


def cube64(n=64, h=0.01, jitter=0.2, seed=0, traction=1e3, direction=(0.0, 0.0, -1.0),
           rho=1000.0, nu=0.3, E_range=(1e5, 1e6), builder="auto"):
    load = traction * np.asarray(direction, dtype=np.float64)
    if _use_device(builder):
        from . import fem
        blk = fem.build_blocks([(n, n, n)], h, jitter=jitter, min_volume=0.05 * h ** 3,
                               draw=_jitter_draw(seed, jitter * h) if jitter > 0 else None,
                               fixed_face="x-", load_face="x+", traction=load)
        c, t, fixed, f = blk.coords, blk.tets, blk.fixed, blk.f_ext
    else:
        c0, t = M.generate_blocks(n, n, n, h, h, h)
        c = M.jitter_interior(c0, t, h, jitter, seed) if jitter > 0 else c0
        fixed = M.face_nodes(c, 0, 0).astype(np.uint8)
        f = M.traction_forces(c, t, 0, 1, load)
        f[fixed == 1] = 0.0
    rng = np.random.default_rng(seed + 1)
    T = t.shape[0]
    E = np.exp(rng.uniform(np.log(E_range[0]), np.log(E_range[1]), size=T))
    return Workload("cube%d_jitter%.2f" % (n, jitter), c, t, fixed, E, np.full(T, nu),
                    np.full(T, rho), f, 0.005, 200,
                    meta=dict(h=h, seed=seed, traction=traction, direction=list(direction)))


def bodies(count=512, lo=4, hi=24, h=0.01, jitter=0.2, seed=0, traction=1e3, rho=1000.0,
           random_traction=False, traction_range=(0.5e3, 2e3), builder="auto"):
    rng = np.random.default_rng(seed)
    dims = rng.integers(lo, hi + 1, size=(count, 3))
    Es = np.exp(rng.uniform(np.log(1e4), np.log(1e6), size=count))
    nus = rng.uniform(0.25, 0.45, size=count)
    # per-body tractions, drawn in body order after the materials
    trs = []
    for b in range(count):
        if random_traction:
            mag = rng.uniform(*traction_range)
            d = rng.normal(size=3)
            d /= np.linalg.norm(d)
            trs.append(mag * d)
        else:
            trs.append(traction * np.array([0.0, 0.0, -1.0]))
    if _use_device(builder):
        from . import fem
        blk = fem.build_blocks(dims, h, jitter=jitter, min_volume=0.05 * h ** 3,
                               draw=_jitter_draw(seed * 100003 + 1, jitter * h) if jitter > 0 else None,
                               fixed_face="x-", load_face="x+", traction=np.stack(trs))
        coords, tets, no, to = blk.coords, blk.tets, blk.body_node_offsets, blk.body_tet_offsets
        fixed, forces = blk.fixed, blk.f_ext
    else:
        parts = [M.generate_blocks(int(I), int(J), int(K), h, h, h) for I, J, K in dims]
        coords0, tets, no, to = M.pack_bodies(parts)
        coords = M.jitter_bodies(coords0, tets, no, h, jitter, seed * 100003 + 1) if jitter > 0 else coords0
        fixed_l, f_l = [], []
        for b in range(count):
            c = coords[no[b]:no[b + 1]]
            t = tets[to[b]:to[b + 1]] - no[b]
            fx = M.face_nodes(c, 0, 0).astype(np.uint8)
            f = M.traction_forces(c, t, 0, 1, trs[b])
            f[fx == 1] = 0.0
            fixed_l.append(fx)
            f_l.append(f)
        fixed, forces = np.concatenate(fixed_l), np.concatenate(f_l, 0)
    bt = np.diff(to)
    T = tets.shape[0]
    return Workload("bodies%d_dims%d-%d" % (count, lo, hi), coords, tets,
                    fixed, np.repeat(Es, bt), np.repeat(nus, bt),
                    np.full(T, rho), forces, 0.01, 100, no, to,
                    meta=dict(h=h, seed=seed, dims=dims.tolist(), jitter=jitter,
                              random_traction=random_traction))


def combined(count=2048, lo=4, hi=32, jitter=0.25, seed=0, builder="auto"):
    """Uneven bodies with per-axis dims uniform in [lo, hi] cubes, jitter
    +/-0.25 h (with resampling), per-body materials and a per-body random
    traction (0.5e3 .. 2e3 Pa, random direction) on the face opposite the
    fixed one; dt 0.01, 100 steps."""
    w = bodies(count, lo, hi, jitter=jitter, seed=seed, random_traction=True, builder=builder)
    w.name = "combined%d_dims%d-%d" % (count, lo, hi)
    return w
