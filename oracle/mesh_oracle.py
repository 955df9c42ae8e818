"""Oracle (test infrastructure only): scalar restatements of the mesh
contract, used to check the product's vectorised mesh code bit for bit.

* ``csr_fortran``: the cell->face incidence exactly as load_case assembles it
  (solver/cut_cells/cutcell_solver.f90:5093-5110): count incidences, prefix
  sum, then one ascending pass over faces inserting the owner (+1) and the
  neighbour (-1) at each cell's cursor.
* ``boundary_fortran``: the per-boundary face lists (:5111-5126).
* ``hilbert_d``: the classic xy->d walk of the Hilbert curve for one point
  (rotate/flip per quadrant), quantised over the bounding box.
"""
import numpy as np


def csr_fortran(nc, owner, neighbour):
    counts = [0] * nc
    for f in range(len(owner)):
        counts[owner[f]] += 1
        if neighbour[f] >= 0:
            counts[neighbour[f]] += 1
    offset = [0] * (nc + 1)
    for c in range(nc):
        offset[c + 1] = offset[c] + counts[c]
    cursor = offset[:nc]
    faces = [0] * offset[nc]
    signs = [0] * offset[nc]
    for f in range(len(owner)):
        c = owner[f]
        faces[cursor[c]] = f; signs[cursor[c]] = 1; cursor[c] += 1
        c = neighbour[f]
        if c >= 0:
            faces[cursor[c]] = f; signs[cursor[c]] = -1; cursor[c] += 1
    return np.array(offset), np.array(faces), np.array(signs)


def boundary_fortran(nb, bid):
    lists = [[] for _ in range(nb)]
    for f, b in enumerate(bid):
        if b >= 0:
            lists[b].append(f)
    offset = [0]
    for b in range(nb):
        offset.append(offset[-1] + len(lists[b]))
    return np.array(offset), np.array([f for b in range(nb) for f in lists[b]], dtype=np.int64)


def hilbert_d(xi, yi, bits):
    n = 1 << bits
    d = 0
    s = n >> 1
    x, y = int(xi), int(yi)
    while s > 0:
        rx = 1 if (x & s) > 0 else 0
        ry = 1 if (y & s) > 0 else 0
        d += s * s * ((3 * rx) ^ ry)
        if ry == 0:
            if rx == 1:
                x = s - 1 - x
                y = s - 1 - y
            x, y = y, x
        s >>= 1
    return d


def hilbert_keys_scalar(x, y, bits=16):
    n = 1 << bits
    x0, x1 = float(np.min(x)), float(np.max(x))
    y0, y1 = float(np.min(y)), float(np.max(y))
    span = max(x1 - x0, y1 - y0, 1e-300)
    out = []
    for px, py in zip(x, y):
        xi = min(int((px - x0) / span * (n - 1)), n - 1)
        yi = min(int((py - y0) / span * (n - 1)), n - 1)
        out.append(hilbert_d(xi, yi, bits))
    return np.array(out, dtype=np.int64)
