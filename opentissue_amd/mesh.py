"""Mesh builders for the FEM path: lattices, jitter, boundary faces, loads.

These are the caller-side helpers that sit in front of ``fem::init``:

* :func:`generate_blocks` follows the index formulas of
  ``OpenTissue/core/containers/t4mesh/util/t4mesh_block_generator.h:32-110``
  (node order x, then J, then K; world coordinates ``x = w*i``,
  ``coord(2) = d*j``, ``coord(1) = h*k``; five tetrahedra per cube with the
  parity switch of ``:91-106``).
* :func:`boundary_faces` extracts faces that belong to exactly one tetrahedron,
  with the per-tet face orientation of
  ``OpenTissue/core/containers/t4mesh/t4mesh_t4boundary_faces.h:130-150``
  (jkm, ijm, kim, ikj).
* :func:`jitter_interior` perturbs interior nodes and resamples any node whose
  incident tetrahedra would fall below a minimum volume.
* :func:`traction_forces`, :func:`gravity_forces` and :func:`lumped_mass`
  build the nodal loads the configurations use.
* :func:`pack_bodies` flattens many bodies into one CSR-of-bodies mesh.

Everything here is plain numpy on the host; none of it is on the timed path.
"""
from __future__ import annotations

import numpy as np

__all__ = [
    "generate_blocks",
    "tet_volumes",
    "boundary_faces",
    "jitter_interior",
    "jitter_bodies",
    "plane_faces",
    "lumped_mass",
    "gravity_forces",
    "traction_forces",
    "face_nodes",
    "pack_bodies",
    "lattice_body",
]


def generate_blocks(I: int, J: int, K: int, width: float, height: float, depth: float):
    """Lattice of I*J*K cubes, five tetrahedra each.

    Returns ``(coords[N,3] float64, tets[T,4] int32)`` with
    N = (I+1)(J+1)(K+1) and T = 5*I*J*K, in the reference generator's order
    (t4mesh_block_generator.h:53-106).
    """
    I, J, K = int(I), int(J), int(K)
    xs, ys, zs = np.meshgrid(np.arange(I + 1), np.arange(J + 1), np.arange(K + 1), indexing="ij")
    coords = np.empty(((I + 1) * (J + 1) * (K + 1), 3), dtype=np.float64)
    coords[:, 0] = width * xs.ravel()
    coords[:, 2] = depth * ys.ravel()   # J runs along world z (:60)
    coords[:, 1] = height * zs.ravel()  # K runs along world y (:61)

    ii, jj, kk = np.meshgrid(np.arange(I), np.arange(J), np.arange(K), indexing="ij")
    ii, jj, kk = ii.ravel(), jj.ravel(), kk.ravel()
    p0 = (ii * (J + 1) + jj) * (K + 1) + kk
    p1 = p0 + 1
    p3 = ((ii + 1) * (J + 1) + jj) * (K + 1) + kk
    p2 = p3 + 1
    p7 = ((ii + 1) * (J + 1) + (jj + 1)) * (K + 1) + kk
    p6 = p7 + 1
    p4 = (ii * (J + 1) + (jj + 1)) * (K + 1) + kk
    p5 = p4 + 1
    odd = ((ii + jj + kk) % 2) == 1
    odd_t = np.stack([
        np.stack([p1, p2, p6, p3], 1),
        np.stack([p3, p6, p4, p7], 1),
        np.stack([p1, p4, p6, p5], 1),
        np.stack([p1, p3, p4, p0], 1),
        np.stack([p1, p6, p4, p3], 1),
    ], 1)
    even_t = np.stack([
        np.stack([p2, p0, p5, p1], 1),
        np.stack([p2, p7, p0, p3], 1),
        np.stack([p2, p5, p7, p6], 1),
        np.stack([p0, p7, p5, p4], 1),
        np.stack([p2, p0, p7, p5], 1),
    ], 1)
    tets = np.where(odd[:, None, None], odd_t, even_t).reshape(-1, 4).astype(np.int32)
    return coords, tets


def generate_blocks_from_voxels(voxels, min_coord, max_coord):
    """Tetrahedra from a voxel grid (t4mesh_block_generator.h:119-243).

    ``voxels`` is a boolean array indexed [i, j, k] (shape (I, J, K), each at
    least 2); the grid spans ``min_coord`` .. ``max_coord``, so grid point
    (i, j, k) sits at min + (i dx, j dy, k dz) with dx = (max - min)/(I - 1)
    (grid_idx2coord.h:29-46, grid.h:171-173).  Every set voxel becomes a node,
    numbered in the grid's storage order (i fastest); each cell contributes
    those of its five tetrahedra whose four corners are all set.  Nodes that
    end up in no tetrahedron are kept, as in the reference (which reports
    them); the solver rejects them unless they are fixed.

    Returns ``(coords[N,3] float64, tets[T,4] int32)``.
    """
    v = np.asarray(voxels, dtype=bool)
    if v.ndim != 3 or min(v.shape) < 2:
        raise ValueError("voxels must be a 3-D array with at least 2 points per axis")
    I, J, K = v.shape
    lo = np.asarray(min_coord, dtype=np.float64).reshape(3)
    hi = np.asarray(max_coord, dtype=np.float64).reshape(3)
    d = (hi - lo) / (np.array([I, J, K]) - 1)
    flat = v.transpose(2, 1, 0).ravel()          # storage order p = (k J + j) I + i
    node_of = np.full(flat.size, -1, dtype=np.int64)
    node_of[flat] = np.arange(int(flat.sum()))
    kk, jj, ii = np.meshgrid(np.arange(K), np.arange(J), np.arange(I), indexing="ij")
    sel = flat
    coords = np.stack([ii.ravel()[sel] * d[0] + lo[0], jj.ravel()[sel] * d[1] + lo[1],
                       kk.ravel()[sel] * d[2] + lo[2]], axis=1)
    # cells (i, j, k) in the reference's loop order: i outer, then j, then k
    ci, cj, ck = np.meshgrid(np.arange(I - 1), np.arange(J - 1), np.arange(K - 1), indexing="ij")
    ci, cj, ck = ci.ravel(), cj.ravel(), ck.ravel()
    p0 = (ck * J + cj) * I + ci
    p1 = (ck * J + cj) * I + ci + 1
    p2 = (ck * J + cj + 1) * I + ci + 1
    p3 = (ck * J + cj + 1) * I + ci
    p4 = ((ck + 1) * J + cj) * I + ci
    p5 = ((ck + 1) * J + cj) * I + ci + 1
    p6 = ((ck + 1) * J + cj + 1) * I + ci + 1
    p7 = ((ck + 1) * J + cj + 1) * I + ci
    odd = ((ci + cj + ck) % 2) == 1
    odd_t = np.stack([np.stack(q, 1) for q in ((p1, p2, p3, p6), (p3, p6, p7, p4), (p1, p4, p5, p6),
                                                (p1, p3, p0, p4), (p1, p6, p3, p4))], 1)
    even_t = np.stack([np.stack(q, 1) for q in ((p2, p0, p1, p5), (p2, p7, p3, p0), (p2, p5, p6, p7),
                                                 (p0, p7, p4, p5), (p2, p0, p5, p7))], 1)
    cand = np.where(odd[:, None, None], odd_t, even_t).reshape(-1, 4)
    keep = flat[cand].all(axis=1)
    tets = node_of[cand[keep]].astype(np.int32)
    return coords, tets


def tet_volumes(coords: np.ndarray, tets: np.ndarray) -> np.ndarray:
    """Signed volumes e10.(e20 x e30)/6 (fem_compute_volume.h:27-39)."""
    p = coords[tets]
    e1 = p[:, 1] - p[:, 0]
    e2 = p[:, 2] - p[:, 0]
    e3 = p[:, 3] - p[:, 0]
    return np.einsum("ij,ij->i", e1, np.cross(e2, e3)) / 6.0


def boundary_faces(tets: np.ndarray) -> np.ndarray:
    """Faces with exactly one incident tetrahedron, as (F,3) node triples in the
    reference's per-tet orientation (t4mesh_t4boundary_faces.h:137-148)."""
    t = np.asarray(tets, dtype=np.int64)
    i, j, k, m = t[:, 0], t[:, 1], t[:, 2], t[:, 3]
    faces = np.concatenate([
        np.stack([j, k, m], 1),
        np.stack([i, j, m], 1),
        np.stack([k, i, m], 1),
        np.stack([i, k, j], 1),
    ], 0)
    key = np.sort(faces, axis=1)
    n = int(t.max()) + 1 if t.size else 1
    if n < (1 << 21):  # one int64 per face: a 1-D unique instead of a row unique
        code = (key[:, 0] << 42) | (key[:, 1] << 21) | key[:, 2]
        _, inv, counts = np.unique(code, return_inverse=True, return_counts=True)
    else:
        _, inv, counts = np.unique(key, axis=0, return_inverse=True, return_counts=True)
    inv = inv.ravel()
    once = counts[inv] == 1
    # keep the reference's emission order: tet by tet, faces jkm, ijm, kim, ikj
    T = t.shape[0]
    order = np.arange(4 * T).reshape(4, T).T.ravel()
    sel = order[once[order]]
    return faces[sel]


def jitter_interior(coords, tets, h, amplitude, seed, min_volume_frac=0.05, max_rounds=200):
    """Perturb interior nodes uniformly by +/- amplitude*h per axis.

    Any node incident to a tetrahedron whose volume falls below
    ``min_volume_frac * h**3`` is redrawn (all of that tet's interior nodes),
    round after round, until every tetrahedron passes.  Deterministic for a
    given seed.  Boundary nodes (on the lattice's bounding box) are not moved,
    so faces stay planar for fixed-face and traction conditions.
    """
    rng = np.random.default_rng(seed)
    x = np.array(coords, dtype=np.float64, copy=True)
    lo = x.min(0)
    hi = x.max(0)
    tol = 1e-9 * h
    interior = np.all((x > lo + tol) & (x < hi - tol), axis=1)
    idx = np.nonzero(interior)[0]
    x[idx] += rng.uniform(-amplitude * h, amplitude * h, size=(idx.size, 3))
    vmin = min_volume_frac * h ** 3
    for _ in range(max_rounds):
        V = tet_volumes(x, tets)
        bad = np.nonzero(V < vmin)[0]
        if bad.size == 0:
            return x
        nodes = np.unique(tets[bad].ravel())
        nodes = nodes[interior[nodes]]
        if nodes.size == 0:
            raise RuntimeError("jitter: a boundary-only tetrahedron is below the volume floor")
        x[nodes] = coords[nodes] + rng.uniform(-amplitude * h, amplitude * h, size=(nodes.size, 3))
    raise RuntimeError("jitter: resampling did not converge")


def jitter_bodies(coords, tets, node_offsets, h, amplitude, seed, min_volume_frac=0.05,
                  max_rounds=200, chunk=1 << 22):
    """jitter_interior for a CSR-of-bodies mesh in one pass: each body's
    bounding-box nodes stay put, interior nodes move uniformly by +/-
    amplitude*h per axis, and nodes of tets below ``min_volume_frac * h**3``
    are redrawn until every tet passes (volumes checked in chunks)."""
    rng = np.random.default_rng(seed)
    x0 = np.asarray(coords, dtype=np.float64)
    x = x0.copy()
    no = np.asarray(node_offsets, dtype=np.int64)
    counts = np.diff(no)
    lo = np.repeat(np.minimum.reduceat(x0, no[:-1], axis=0), counts, axis=0)
    hi = np.repeat(np.maximum.reduceat(x0, no[:-1], axis=0), counts, axis=0)
    tol = 1e-9 * h
    interior = np.all((x0 > lo + tol) & (x0 < hi - tol), axis=1)
    del lo, hi
    idx = np.nonzero(interior)[0]
    x[idx] += rng.uniform(-amplitude * h, amplitude * h, size=(idx.size, 3))
    vmin = min_volume_frac * h ** 3
    check = np.arange(len(tets))
    for _ in range(max_rounds):
        bad = []
        for a in range(0, len(check), chunk):
            sel = check[a:a + chunk]
            V = tet_volumes(x, tets[sel])
            bad.append(sel[V < vmin])
        bad = np.concatenate(bad) if bad else np.zeros(0, dtype=np.int64)
        if bad.size == 0:
            return x
        nodes = np.unique(tets[bad].ravel())
        nodes = nodes[interior[nodes]]
        if nodes.size == 0:
            raise RuntimeError("jitter: a boundary-only tetrahedron is below the volume floor")
        x[nodes] = x0[nodes] + rng.uniform(-amplitude * h, amplitude * h, size=(nodes.size, 3))
        # only tets touching a redrawn node can change
        touched = np.zeros(len(x), dtype=bool)
        touched[nodes] = True
        check = np.nonzero(touched[tets].any(axis=1))[0]
    raise RuntimeError("jitter: resampling did not converge")


def lumped_mass(coords, tets, density, fixed=None):
    """m_i = sum over incident tets of rho_e V_e / 4 (fem_compute_mass.h:164-171).
    Fixed nodes keep the value the reference would give them before the
    DBL_MAX start (the caller decides what to do with them)."""
    V = tet_volumes(coords, tets)
    rho = np.broadcast_to(np.asarray(density, dtype=np.float64), V.shape)
    m = np.zeros(coords.shape[0])
    np.add.at(m, tets.ravel(), np.repeat(rho * V * 0.25, 4))
    return m


def gravity_forces(mass, fixed, g=(0.0, 0.0, -9.81)):
    """f_ext,i = m_i g on free nodes, zero on fixed nodes."""
    f = np.outer(mass, np.asarray(g, dtype=np.float64))
    f[np.asarray(fixed, dtype=bool)] = 0.0
    return f


def face_nodes(coords, axis, side, tol=1e-9):
    """Boolean mask of nodes on the bounding-box face ``axis`` (0/1/2) at the
    minimum (side=0) or maximum (side=1)."""
    c = coords[:, axis]
    ref = c.min() if side == 0 else c.max()
    span = max(c.max() - c.min(), 1.0)
    return np.abs(c - ref) <= tol * span


def plane_faces(tets, on):
    """Tet faces whose three nodes all satisfy ``on``, in boundary_faces'
    emission order.  When ``on`` marks a bounding plane of the mesh every such
    face is a boundary face (two tets sharing a face lie on opposite sides of
    it, and no node lies beyond a bounding plane), so no face matching is
    needed."""
    t = np.asarray(tets, dtype=np.int64)
    hit = on[t]                                   # [T, 4]
    cand = hit.sum(1) >= 3
    tc = t[cand]
    i, j, k, m = tc[:, 0], tc[:, 1], tc[:, 2], tc[:, 3]
    faces = np.stack([np.stack([j, k, m], 1), np.stack([i, j, m], 1),
                      np.stack([k, i, m], 1), np.stack([i, k, j], 1)], 1)   # [Tc, 4, 3]
    faces = faces.reshape(-1, 3)
    return faces[on[faces].all(axis=1)]


def traction_forces(coords, tets, axis, side, traction, faces=None):
    """Constant traction ``traction`` (3-vector, Pa) on the bounding-box face
    (axis, side), converted to nodal forces t*A_tri/3 per boundary triangle."""
    on = face_nodes(coords, axis, side)
    if faces is None:
        faces = plane_faces(tets, on)
    sel = faces[on[faces].all(axis=1)]
    p = coords[sel]
    area = 0.5 * np.linalg.norm(np.cross(p[:, 1] - p[:, 0], p[:, 2] - p[:, 0]), axis=1)
    f = np.zeros_like(coords)
    t = np.asarray(traction, dtype=np.float64)
    contrib = (area / 3.0)[:, None] * t[None, :]
    for c in range(3):
        np.add.at(f, sel[:, c], contrib)
    return f


def lattice_body(dims, h, jitter=0.0, seed=0, min_volume_frac=0.05):
    """One lattice body with ``dims=(I,J,K)`` cubes of edge h, optionally
    jittered.  Returns ``(coords, tets)``."""
    I, J, K = dims
    coords, tets = generate_blocks(I, J, K, h, h, h)
    if jitter > 0.0:
        coords = jitter_interior(coords, tets, h, jitter, seed, min_volume_frac)
    return coords, tets


def pack_bodies(bodies):
    """Flatten ``[(coords, tets), ...]`` into one mesh plus CSR-of-bodies
    offsets: body b owns nodes [node_off[b], node_off[b+1]) and tetrahedra
    [tet_off[b], tet_off[b+1]); tetrahedra reference global node indices."""
    node_off = [0]
    tet_off = [0]
    cs, ts = [], []
    for c, t in bodies:
        ts.append(np.asarray(t, dtype=np.int64) + node_off[-1])
        cs.append(np.asarray(c, dtype=np.float64))
        node_off.append(node_off[-1] + c.shape[0])
        tet_off.append(tet_off[-1] + t.shape[0])
    coords = np.concatenate(cs, 0)
    tets = np.concatenate(ts, 0).astype(np.int32)
    return coords, tets, np.asarray(node_off, dtype=np.int64), np.asarray(tet_off, dtype=np.int64)
