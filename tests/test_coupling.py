"""Surface-to-volume mesh coupling on the CPU: the oracle (oracle/coupling.py)
pinned by closed forms, and the C++ header (include/otfem/mesh_coupling.hpp,
through tests/cpp/coupling_check) against the oracle on the reference's
irregular blocky.xml mesh."""
import os
import subprocess

import numpy as np
import pytest

from oracle import coupling as OC

HERE = os.path.dirname(os.path.abspath(__file__))
BLOCKY = os.path.join(HERE, "golden", "blocky.xml")
CHECK = os.path.join(HERE, "cpp", "build", "coupling_check")
A = np.array([[1.1, 0.2, -0.1], [0.05, 0.9, 0.3], [-0.2, 0.1, 1.3]])
SHIFT = np.array([0.3, -0.7, 0.25])


def _blocky():
    d = np.load(os.path.join(HERE, "golden", "blocky.npz"))
    return d["coords"], d["tets"]


def _points(coords, tets, seed=0, inside=300, outside=20):
    rng = np.random.default_rng(seed)
    t = rng.integers(0, len(tets), inside)
    w = rng.dirichlet(np.ones(4), inside)
    pin = np.einsum("nk,nkd->nd", w, coords[tets[t]])
    lo, hi = coords.min(0), coords.max(0)
    pout = hi + rng.uniform(0.1, 1.0, (outside, 3))
    return np.concatenate([pin, coords[::7], pout, lo[None] - 0.5])


def test_closed_forms_on_one_tet():
    nodes = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0], [0, 0, 1.0]])
    tets = np.array([[0, 1, 2, 3]])
    for k in range(4):
        w = OC.barycentric_all(nodes, tets, nodes[k])[0]
        assert np.allclose(w, np.eye(4)[k], atol=1e-15)
    assert np.allclose(OC.barycentric_all(nodes, tets, nodes.mean(0))[0], 0.25, atol=1e-15)
    w123, idx = OC.bind(nodes, tets, [[0.2, 0.3, 0.1], [0.6, 0.6, 0.6], [-1e-5, 0.5, 0.2]])
    assert idx.tolist() == [0, -1, 0]  # the last one lies within the 1e-4 tolerance
    assert np.allclose(w123[0], [0.2, 0.3, 0.1], atol=1e-15)


def test_oracle_update_reproduces_points_and_follows_affine_motion():
    c, t = _blocky()
    pts = _points(c, t)
    w123, idx = OC.bind(c, t, pts)
    bound = idx >= 0
    assert bound.sum() >= 300 + len(c[::7]) and (~bound).sum() == 21
    rest = OC.update(c, t, w123, idx, pts)
    assert np.abs(rest - pts).max() < 1e-13
    moved = OC.update(c @ A.T + SHIFT, t, w123, idx, pts)
    assert np.abs(moved[bound] - (pts[bound] @ A.T + SHIFT)).max() < 1e-12
    assert np.array_equal(moved[~bound], pts[~bound])


def test_cpp_header_matches_oracle(tmp_path):
    assert os.path.exists(CHECK), "build() compiles tests/cpp/coupling_check.cpp"
    c, t = _blocky()
    pts = _points(c, t, seed=1)
    pf, of = str(tmp_path / "p.txt"), str(tmp_path / "o.txt")
    with open(pf, "w") as f:
        f.write(f"{len(pts)}\n" + "".join(f"{x!r} {y!r} {z!r}\n" for x, y, z in pts.tolist()))
    r = subprocess.run([CHECK, BLOCKY, pf, of], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    out = np.loadtxt(of).reshape(-1, 7)
    w123, idx = OC.bind(c, t, pts)
    assert np.array_equal(out[:, 0].astype(np.int64), idx)
    assert np.abs(out[:, 1:4] - w123).max() < 1e-12
    bound = idx >= 0
    assert np.abs(out[bound, 4:] - (pts[bound] @ A.T + SHIFT)).max() < 1e-12
    assert np.array_equal(out[~bound, 4:], pts[~bound])


def test_oracle_normals_closed_forms():
    rng = np.random.default_rng(4)
    n0 = rng.normal(size=(50, 3))
    idx = rng.integers(-1, 6, 50)
    eye = np.broadcast_to(np.eye(3), (6, 3, 3))
    assert np.array_equal(OC.normals(eye, idx, n0), n0)
    Q, _ = np.linalg.qr(rng.normal(size=(3, 3)))
    out = OC.normals(np.broadcast_to(Q, (6, 3, 3)), idx, n0)
    b = idx >= 0
    assert np.allclose(out[b], n0[b] @ Q.T, atol=1e-14)
    assert np.array_equal(out[~b], n0[~b])
