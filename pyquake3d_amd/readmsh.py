"""Mesh input and element frames (mirror of src/readmsh.py).

* ``read_mshV2(fname)``  -- gmsh 2.2 ASCII reader; node coordinates are scaled
  km -> m (x1e3); triangles are element lines with at least 8 tokens, the last
  three tokens being the 1-based node ids (readmsh.py:12-68).
* ``get_eleVec(nodelst, elelst, jud_ele_order)`` -- per-element frame rows
  [e1, e2, n] and centroids (readmsh.py:75-134).
"""
import numpy as np


def read_mshV2(fname):
    nodes = []
    eles = []
    mode = None
    count_line = False
    with open(fname, "r") as f:
        for raw in f:
            line = raw.rstrip("\n")
            if line == "$Nodes":
                mode, count_line = "n", True
                continue
            if line == "$Elements":
                mode, count_line = "e", True
                continue
            if mode is None:
                continue
            if count_line:  # the count line right after the section header
                count_line = False
                continue
            tok = line.split()
            if mode == "n":
                try:
                    nodes.append([float(tok[1]), float(tok[2]), float(tok[3])])
                except (IndexError, ValueError):
                    mode = None
            else:
                try:
                    if len(tok) >= 8:
                        eles.append([int(tok[-3]), int(tok[-2]), int(tok[-1])])
                    elif line.startswith("$"):
                        mode = None
                except ValueError:
                    mode = None
    return np.array(nodes) * 1e3, np.array(eles)


def get_eleVec(nodelst, elelst, jud_ele_order=False):
    nodes = np.asarray(nodelst, dtype=np.float64)
    e = np.asarray(elelst, dtype=np.int64) - 1
    xa, xb, xc = nodes[e[:, 0]], nodes[e[:, 1]], nodes[e[:, 2]]
    xg = np.stack([np.mean(np.stack([xa[:, k], xb[:, k], xc[:, k]], axis=1), axis=1) for k in range(3)], axis=1)
    vba = xb - xa
    vca = xc - xa
    if jud_ele_order:  # clockwise node order: n = (c - a) x (b - a)
        n = np.cross(vca, vba)
    else:              # counter-clockwise: n = (b - a) x (c - a)
        n = np.cross(vba, vca)
    n = n / np.sqrt((n * n).sum(axis=1))[:, None]
    e1 = np.zeros_like(n)
    steep = np.abs(n[:, 2]) < 1
    e1[:, 0] = np.where(steep, n[:, 1], 1.0)
    e1[:, 1] = np.where(steep, -n[:, 0], 0.0)
    rr = np.sqrt(e1[:, 0] ** 2 + e1[:, 1] ** 2)
    e1[steep, 0] /= rr[steep]
    e1[steep, 1] /= rr[steep]
    e2 = np.stack([n[:, 1] * e1[:, 2] - n[:, 2] * e1[:, 1],
                   n[:, 2] * e1[:, 0] - n[:, 0] * e1[:, 2],
                   n[:, 0] * e1[:, 1] - n[:, 1] * e1[:, 0]], axis=1)
    return np.concatenate([e1, e2, n], axis=1), xg
