"""Writes tests/golden/blocky.npz from the reference's T4MESH fixture
demos/data/xml/blocky.xml (438 nodes, 1831 tetrahedra).  Data only: node
coordinates and the i, j, k, m corner indices as the file lists them."""
import os
import re
import sys

import numpy as np

SRC = "/root/reference/demos/data/xml/blocky.xml"
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "blocky.npz")


def main(src=SRC):
    text = open(src).read()
    n = int(re.search(r'<T4MESH nodes="(\d+)"', text).group(1))
    tets = np.array([[int(a), int(b), int(c), int(d)] for a, b, c, d in re.findall(
        r'<TETRAHEDRON i="(\d+)" j="(\d+)" k="(\d+)" m="(\d+)"/>', text)], dtype=np.int32)
    coords = np.zeros((n, 3))
    seen = np.zeros(n, dtype=bool)
    for idx, c in re.findall(r'<POINT idx="(\d+)" coord="\[([^\]]+)\]"/>', text):
        i = int(idx)
        coords[i] = [float(v) for v in c.split(",")]
        seen[i] = True
    assert seen.all() and tets.shape == (1831, 4)
    np.savez(OUT, coords=coords, tets=tets)
    print(f"wrote {OUT}: {n} nodes, {tets.shape[0]} tets")


if __name__ == "__main__":
    main(*sys.argv[1:])
