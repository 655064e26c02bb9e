"""Systems whose pairs sit within a few ulps of the list radius.

IMD puts a pair in the neighbour list when r^2 < cellsz, strictly
(src/imd_forces_nbl.c:258), with r^2 evaluated from the atom whose half
stencil (src/imd_geom_3d.c:895-899) holds the other atom's cell: d = (image
of the partner) - (owner), the image built by adding box vectors z, then y,
then x (copy_cell, src/imd_comm_force_3d.c:754-756), and r^2 = d0*d0 + d1*d1
+ d2*d2 left to right, every operation rounded on its own.

``ref_pairs`` is that decision by brute force in Python floats (IEEE
doubles, no contraction) for grids of 3 cells per direction, where every
cell is the neighbour of every other through exactly one offset.  The
builders place owner/partner pairs at prescribed offsets of ulps around
cellsz: across every periodic face, edge and corner image, inside one cell,
with the owner being the lower or the higher atom number, in a cubic and in
a sheared box.
"""
from __future__ import annotations

import math

import numpy as np

from imd_amd import workloads as W

G3 = 3  # cells per direction of these boxes


def frac(box, x):
    """Fractional coordinates x . tbox (tbox = dual vectors, make_box)."""
    return np.linalg.solve(np.asarray(box, dtype=np.float64).T, x)


def cell_of(box, x):
    f = frac(box, np.asarray(x, dtype=np.float64))
    return tuple(int(math.floor(G3 * v)) for v in f)


def image(box, x, s):
    """x + s . box, box vectors added z, then y, then x."""
    y = [float(x[0]), float(x[1]), float(x[2])]
    for ax in (2, 1, 0):
        if s[ax]:
            sv = float(s[ax])
            y = [y[e] + sv * float(box[ax][e]) for e in range(3)]
    return y


def r2_from(box, xo, xp, s):
    """(image of partner xp shifted by s) - owner xo, squared, left to right."""
    y = image(box, xp, s)
    d0 = y[0] - float(xo[0])
    d1 = y[1] - float(xo[1])
    d2 = y[2] - float(xo[2])
    return d0 * d0 + d1 * d1 + d2 * d2


def offset_and_shift(ci, cj):
    """Cell offset o of cj as seen from ci on a 3-cell periodic grid, and the
    image shift s of the atom in cj that makes it that neighbour."""
    o, s = [], []
    for a, b in zip(ci, cj):
        d = b - a
        if d > 1:
            o.append(d - G3)
            s.append(-1)
        elif d < -1:
            o.append(d + G3)
            s.append(1)
        else:
            o.append(d)
            s.append(0)
    return tuple(o), tuple(s)


def pair_r2(box, xi, xj, ci, cj):
    """IMD's r^2 of the unordered pair (i, j) and the shift of j relative
    to i."""
    o, s = offset_and_shift(ci, cj)
    if o > (0, 0, 0):
        return r2_from(box, xi, xj, s), s
    if o < (0, 0, 0):
        return r2_from(box, xj, xi, tuple(-v for v in s)), s
    return r2_from(box, xi, xj, (0, 0, 0)), s


def ref_pairs(system, cellsz):
    """The reference's pair set (id_lo, id_hi, sx, sy, sz) by brute force."""
    box = system.box
    cells = [cell_of(box, system.x[i]) for i in range(system.n)]
    out = set()
    for i in range(system.n):
        for j in range(i + 1, system.n):
            r2, s = pair_r2(box, system.x[i], system.x[j], cells[i], cells[j])
            if r2 < cellsz:
                a, b = int(system.id[i]), int(system.id[j])
                if a < b:
                    out.add((a, b) + tuple(s))
                else:
                    out.add((b, a) + tuple(-v for v in s))
    return out


# (owner position as fractions of the box, direction to the partner image,
# owner gets the lower atom number)
GEOMS = [
    ((0.20, 0.50, 0.50), (1, 0, 0), True),      # interior, cell 0 -> 1
    ((0.50, 0.20, 0.55), (0, 1, 0), False),     # interior, owner higher
    ((0.45, 0.55, 0.20), (0, 0, 1), True),      # interior along z
    ((0.90, 0.35, 0.40), (1, 0, 0), True),      # periodic x face
    ((0.35, 0.90, 0.60), (0, 1, 0), False),     # periodic y face
    ((0.60, 0.40, 0.90), (0, 0, 1), True),      # periodic z face
    ((0.90, 0.90, 0.90), (1, 1, 1), False),     # corner image (1, 1, 1)
    ((0.90, 0.10, 0.50), (1, -1, 0), True),     # edge image (1, -1, 0)
    ((0.50, 0.90, 0.10), (0, 1, -1), False),    # edge image (0, 1, -1)
    ((0.02, 0.15, 0.70), (1, 0, 0.05), True),   # both atoms in one cell
    ((0.88, 0.12, 0.88), (1, -1, 1), True),     # corner image (1, -1, 1)
]


def _place(box, cellsz, xo, u, k):
    """Partner position for owner xo: its image lies along u at a distance
    whose r^2 is the k-th representable step (of the partner's coordinate
    along u's largest component) from the first at or above cellsz."""
    box = np.asarray(box, dtype=np.float64)
    u = np.asarray(u, dtype=np.float64)
    u = u / np.linalg.norm(u)
    y = xo + math.sqrt(cellsz) * u
    # the partner inside the box and the shift that maps it to y
    fy = frac(box, y)
    s = np.floor(fy).astype(int)
    xp = y - s @ box
    ax = int(np.argmax(np.abs(u)))
    sign = 1.0 if u[ax] > 0 else -1.0

    def at(n):
        v = xp.copy()
        step = np.inf if sign * n > 0 else -np.inf
        for _ in range(abs(n)):
            v[ax] = np.nextafter(v[ax], step)
        return v

    # r^2 grows with n: find the first n with r^2 >= cellsz
    n = 0
    while r2_from(box, xo, at(n), tuple(s)) >= cellsz:
        n -= 1
    while r2_from(box, xo, at(n), tuple(s)) < cellsz:
        n += 1
    return at(n + k)


def lj_cellsz(margin=0.3):
    """init_cells' list radius squared for the LJ table: (sqrt(end) +
    nbl_margin)^2 (src/imd_geom_3d.c:123-126)."""
    return (math.sqrt(float(W.lj_table().end[0])) + margin) ** 2


def radius_system(k, cellsz, sheared=False, L=10.0):
    """All geometries of GEOMS at offset k around cellsz (k < 0: inside the
    list radius, k >= 0: outside), in a box of 3 cells per direction."""
    box = np.eye(3) * L
    if sheared:
        box[1] = (1.5, L, 0.0)
        box[2] = (0.7, -1.1, L)
    xs, ids = [], []
    nid = 0
    for fo, u, owner_low in GEOMS:
        xo = np.asarray(fo, dtype=np.float64) @ box
        u = np.asarray(u, dtype=np.float64) @ (box / L) if sheared else \
            np.asarray(u, dtype=np.float64)
        xp = _place(box, cellsz, xo, u, k)
        a, b = (nid, nid + 1) if owner_low else (nid + 1, nid)
        xs += [xo, xp]
        ids += [a, b]
        nid += 2
    x = np.array(xs)
    n = len(x)
    return W.System(x=x, p=np.zeros((n, 3)), type=np.zeros(n, np.int32),
                    mass=np.ones(n), id=np.array(ids, dtype=np.int64),
                    box=box)


KS = (-3, -2, -1, 0, 1, 2)
