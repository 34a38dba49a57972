"""FP64 CPU restatement of the surface-to-volume mesh coupling.  TEST
INFRASTRUCTURE ONLY: the product runs include/otfem/mesh_coupling.hpp (host
binding) and the device update of otfem_surface_update; this module is the
checker.

Follows core/containers/t4mesh/util/t4mesh_mesh_coupling.h:
    bind    :100-200 with the collision policy's test :80-98 (all four
            barycentric coordinates in (-1e-4, 1 + 1e-4); stored are the
            coordinates of corners j, k, m) -- by brute force over the
            tetrahedra in index order, so the first hit is the lowest index
    update  :202-238  (w0 = 1 - w1 - w2 - w3; p0 w0 + p1 w1 + p2 w2 + p3 w3)
and core/geometry/geometry_barycentric.h:339-372 for the coordinates
(w = P^-1 (p - x4), P = [x1 - x4, x2 - x4, x3 - x4]), and the normal update
the coupling notes describe (:171-176): n = Re n0 with Re the bound
tetrahedron's rotation.

Pins (tests/test_coupling.py): corners and centroids of tetrahedra have the
closed-form coordinates, a bound point is reproduced by the update at the
binding positions, an affine motion of the nodes moves every bound point by
the same affine map, and points outside the mesh stay unbound; for normals,
identity rotations keep n0 and a rotation Q gives Q n0 (a closed form the
GPU test also checks on a rigidly rotated mesh).
"""
from __future__ import annotations

import numpy as np

TOL = 1e-4


def barycentric_all(nodes, tets, p):
    """Coordinates [T, 4] of point p in every tetrahedron (NaN rows for
    degenerate ones)."""
    x = np.asarray(nodes, dtype=np.float64)[np.asarray(tets)]
    x4 = x[:, 3]
    P = np.stack([x[:, 0] - x4, x[:, 1] - x4, x[:, 2] - x4], axis=2)  # columns
    r = np.asarray(p, dtype=np.float64)[None, :] - x4
    det = np.linalg.det(P)
    ok = det != 0.0
    w = np.full((len(tets), 4), np.nan)
    if ok.any():
        w[ok, :3] = np.linalg.solve(P[ok], r[ok][:, :, None])[:, :, 0]
        w[ok, 3] = 1.0 - w[ok, :3].sum(axis=1)
    return w


def bind(nodes, tets, points, tol=TOL):
    """(w123 [n, 3], tet_index [n]) with index -1 for points in no
    tetrahedron."""
    pts = np.asarray(points, dtype=np.float64).reshape(-1, 3)
    w123 = np.zeros((len(pts), 3))
    idx = np.full(len(pts), -1, dtype=np.int64)
    for i, p in enumerate(pts):
        w = barycentric_all(nodes, tets, p)
        inside = np.all((w > -tol) & (w < 1.0 + tol), axis=1)
        hit = np.flatnonzero(inside)
        if hit.size:
            t = hit[0]
            idx[i] = t
            w123[i] = w[t, 1:4]
    return w123, idx


def update(nodes, tets, w123, idx, points):
    """Positions of the bound points at node positions `nodes`; unbound
    points keep `points`."""
    nodes = np.asarray(nodes, dtype=np.float64)
    out = np.array(points, dtype=np.float64).reshape(-1, 3)
    for i, t in enumerate(np.asarray(idx)):
        if t < 0:
            continue
        w1, w2, w3 = w123[i]
        w0 = 1.0 - w1 - w2 - w3
        c = np.asarray(tets)[t]
        out[i] = ((nodes[c[0]] * w0 + nodes[c[1]] * w1) + nodes[c[2]] * w2) + nodes[c[3]] * w3
    return out


def normals(Re, idx, n0):
    """n = Re[idx] n0 per bound vertex (Re [T, 3, 3] row-major), n0 for
    unbound ones (idx -1)."""
    Re = np.asarray(Re, dtype=np.float64).reshape(-1, 3, 3)
    n0 = np.asarray(n0, dtype=np.float64).reshape(-1, 3)
    out = n0.copy()
    b = np.asarray(idx) >= 0
    out[b] = np.einsum("nij,nj->ni", Re[np.asarray(idx)[b]], n0[b])
    return out
