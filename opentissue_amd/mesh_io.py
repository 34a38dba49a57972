"""Tetrahedral mesh files: T4MESH XML read / write and the TetGen reader.

Python counterparts of <otfem/t4mesh_io.hpp>, over numpy arrays
(coords [N, 3] float64, tets [T, 4] int32, zero-based).

Reference interfaces mirrored (paths relative to the OpenTissue tree):
    xml_read     core/containers/t4mesh/io/t4mesh_xml_read.h:42-200
    xml_write    core/containers/t4mesh/io/t4mesh_xml_write.h:82-115, 134-149
    tetgen_read  core/containers/t4mesh/io/t4mesh_tetgen_read.h:29-136

Where the reference returns false (a missing attribute, an index out of
range) or asserts (no T4MESH element, a node count that is not positive, a
coordinate that is not a finite number), these raise ValueError.  Optional
`tag` attributes are ignored.  xml_write prints coordinates with 17
significant digits (the reference: 15), so write + read is exact.
"""
from __future__ import annotations

import xml.etree.ElementTree as ET

import numpy as np


def _int(el, key, what):
    v = el.get(key)
    if v is None:
        raise ValueError(f"t4mesh::xml_read(): Missing {key} {what}")
    try:
        return int(v.strip())
    except ValueError:
        raise ValueError(f"t4mesh::xml_read(): {key}={v!r} is not an integer") from None


def _coord(text):
    """"[x,y,z]" as math_vector3.h:252-263 reads it: one separator character
    before, between and after the three numbers."""
    s = text.strip()
    if len(s) < 2:
        raise ValueError(f"t4mesh::xml_read(): bad coord {text!r}")
    parts = s[1:-1].split(",") if s[0] == "[" and s[-1] == "]" else None
    if parts is None or len(parts) != 3:
        raise ValueError(f"t4mesh::xml_read(): bad coord {text!r}")
    v = np.array([float(p) for p in parts])
    if not np.all(np.isfinite(v)):
        raise ValueError("t4mesh::xml_read(): coordinate is not a finite number")
    return v


def xml_read(path):
    """(coords, tets) of a T4MESH XML file.  Nodes without a POINT element
    keep the coordinate (0, 0, 0), as the reference's default-constructed
    points do."""
    root = ET.parse(path).getroot()
    if root.tag != "T4MESH":
        raise ValueError("t4mesh::xml_read(): no T4MESH element")
    try:
        n = int(root.get("nodes", "0"))
    except ValueError:
        n = 0
    if n <= 0:
        raise ValueError("t4mesh::xml_read(): node count was not positive")
    tets = []
    for el in root.findall("TETRAHEDRON"):
        idx = [_int(el, k, "index on tetrahedron") for k in ("i", "j", "k", "m")]
        if min(idx) < 0 or max(idx) >= n:
            raise ValueError("t4mesh::xml_read(): Illegal node index on tetrahedron")
        tets.append(idx)
    coords = np.zeros((n, 3))
    for el in root.findall("POINT"):
        i = _int(el, "idx", "index on point")
        if i < 0 or i >= n:
            raise ValueError("t4mesh::xml_read(): Illegal index on point")
        c = el.get("coord")
        if c is None:
            raise ValueError("t4mesh::xml_read(): Missing coord on point")
        coords[i] = _coord(c)
    return coords, np.asarray(tets, dtype=np.int32).reshape(-1, 4)


def xml_write(path, coords, tets):
    """Writes a T4MESH XML file (TETRAHEDRON elements, then POINT elements)."""
    coords = np.asarray(coords, dtype=np.float64).reshape(-1, 3)
    tets = np.asarray(tets).reshape(-1, 4)
    n = coords.shape[0]
    if tets.size and (tets.min() < 0 or tets.max() >= n):
        raise ValueError("t4mesh::xml_write(): node index out of range")
    lines = ['<?xml version="1.0" ?>', f'<T4MESH nodes="{n}">']
    lines += [f'    <TETRAHEDRON i="{a}" j="{b}" k="{c}" m="{d}" />' for a, b, c, d in tets.tolist()]
    lines += [f'    <POINT idx="{i}" coord="[{x:.17g},{y:.17g},{z:.17g}]" />'
              for i, (x, y, z) in enumerate(coords.tolist())]
    lines.append("</T4MESH>")
    with open(path, "w") as f:
        f.write("\n".join(lines) + "\n")


def _tokens(path):
    out = []
    with open(path) as f:
        for line in f:
            out += line.split("#", 1)[0].split()
    return out


def tetgen_read(basename):
    """(coords, tets) of `basename`.node and `basename`.ele as TetGen writes
    them (any index base; node attributes, boundary markers and one region
    attribute per tetrahedron are skipped).  Nodes are numbered in file
    order."""
    tok = _tokens(basename + ".node")
    try:
        n, dim, nattr, bnd = (int(t) for t in tok[:4])
    except ValueError:
        raise ValueError(f"t4mesh::tetgen_read(): bad header in {basename}.node") from None
    if n < 0 or dim != 3 or nattr < 0:
        raise ValueError(f"t4mesh::tetgen_read(): bad header in {basename}.node")
    width = 4 + nattr + (1 if bnd else 0)
    body = tok[4:4 + n * width]
    if len(body) < n * width:
        raise ValueError(f"t4mesh::tetgen_read(): {basename}.node ends early")
    rows = np.array(body, dtype=object).reshape(n, width) if n else np.zeros((0, width), object)
    ext = [int(r[0]) for r in rows]
    coords = np.array([[float(r[1]), float(r[2]), float(r[3])] for r in rows]).reshape(-1, 3)
    internal = {e: i for i, e in enumerate(ext)}
    tok = _tokens(basename + ".ele")
    try:
        t, npt, nattr = (int(v) for v in tok[:3])
    except ValueError:
        raise ValueError(f"t4mesh::tetgen_read(): bad header in {basename}.ele") from None
    if npt != 4:
        raise ValueError("We only support tetrahedra with 4 points!")
    if nattr > 1:
        raise ValueError("We only support tetrahedra with 1 region!")
    width = 5 + nattr
    body = tok[3:3 + t * width]
    if len(body) < t * width:
        raise ValueError(f"t4mesh::tetgen_read(): {basename}.ele ends early")
    tets = np.zeros((t, 4), dtype=np.int32)
    for k in range(t):
        row = body[k * width:(k + 1) * width]
        for c in range(4):
            e = int(row[1 + c])
            if e not in internal:
                raise ValueError(f"t4mesh::tetgen_read(): tetrahedron {row[0]} names unknown node {e}")
            tets[k, c] = internal[e]
    return coords, tets
