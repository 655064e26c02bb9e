"""Pins for the oracle's grid sizing, binning, wrapping and neighbour-pair
set (init_cells, cell_coord, do_boundaries, make_nblist with buffer-cell
images).

The pair set is checked against an O(N^2) brute force over the 27 periodic
images (minimum image is not assumed), on random configurations chosen so
that no pair distance lies within 1e-9 of the list radius.
"""
import math

import numpy as np
import pytest

import oracle as O
from imd_amd import workloads as W


def one_col_table(end_r2):
    step, xs = W._grid(1.0, end_r2, 50)
    return W.Table(np.array([1.0]), np.array([end_r2]), np.array([step]),
                   [np.zeros(50) + 1e-3 * (end_r2 - xs)], True)


def make(system, rc=2.5, margin=0.3):
    pair = O.OracleTable.from_table(one_col_table(rc * rc))
    return O.OracleSystem(system, pair, None, None, nbl_margin=margin,
                          timestep=0.001, ntypes=1), pair


def random_system(box, n, pbc=(1, 1, 1), seed=0):
    rng = np.random.default_rng(seed)
    frac = rng.random((n, 3))
    x = frac @ np.asarray(box)
    return W.System(x=x, p=np.zeros_like(x), type=np.zeros(n, np.int32),
                    mass=np.ones(n), id=np.arange(n, dtype=np.int64),
                    box=np.asarray(box, dtype=float), pbc=pbc)


def brute_pairs(x, box, pbc, cellsz):
    n = x.shape[0]
    out, near = set(), 0
    shifts = [(a, b, c) for a in (-1, 0, 1) for b in (-1, 0, 1) for c in (-1, 0, 1)
              if (pbc[0] or a == 0) and (pbc[1] or b == 0) and (pbc[2] or c == 0)]
    for (a, b, c) in shifts:
        S = a * box[0] + b * box[1] + c * box[2]
        d = x[None, :, :] + S - x[:, None, :]
        r2 = np.einsum("ijk,ijk->ij", d, d)
        near += int(np.sum(np.abs(r2 - cellsz) < 1e-9))
        ii, jj = np.nonzero(r2 < cellsz)
        for i, j in zip(ii.tolist(), jj.tolist()):
            if i < j:
                out.add((i, j, a, b, c))
            elif i == j and (a, b, c) > (0, 0, 0):
                out.add((i, j, a, b, c))
    return out, near


@pytest.mark.parametrize("L,rc,margin", [(15.496, 2.5, 0.3), (20.0, 2.5, 0.4),
                                         (9.0, 2.5, 0.3)])
def test_grid_sizing_closed_form(L, rc, margin):
    s = random_system(np.eye(3) * L, 20)
    orc, _ = make(s, rc, margin)
    expect = int(L / (rc + margin))
    assert orc.grid() == (expect,) * 3
    assert orc.scalar("cellsz") == (rc + margin) ** 2


def test_cell_assignment_and_wrap():
    L, g = 16.8, 6  # cell edge 2.8 = 2.5 + 0.3
    centres = (np.stack(np.meshgrid(np.arange(g), np.arange(g), np.arange(g),
                                    indexing="ij"), -1).reshape(-1, 3) + 0.5) * (L / g)
    # move some atoms out of the box by whole box vectors
    x = centres.copy()
    x[::7] += L
    x[3::11, 1] -= 2 * L
    s = W.System(x=x, p=np.zeros_like(x), type=np.zeros(len(x), np.int32),
                 mass=np.ones(len(x)), id=np.arange(len(x)), box=np.eye(3) * L)
    orc, _ = make(s)
    orc.calc_forces()
    assert orc.grid() == (g, g, g)
    expect = np.floor(centres / (L / g)).astype(int)
    assert np.array_equal(orc.cells(), expect)
    assert np.allclose(orc.atoms()["x"], centres, atol=1e-12)


@pytest.mark.parametrize("case", ["cubic", "slab", "tall", "triclinic"])
def test_pair_set_matches_brute_force(case):
    rc, margin = 2.5, 0.3
    cs = (rc + margin) ** 2
    if case == "cubic":
        box, pbc, n = np.eye(3) * 9.5, (1, 1, 1), 300
    elif case == "slab":
        box, pbc, n = np.diag([10.0, 9.0, 14.0]), (1, 1, 0), 300
    elif case == "tall":
        box, pbc, n = np.diag([8.5, 9.0, 23.0]), (1, 1, 1), 400
    else:
        box, pbc, n = np.array([[10.0, 0, 0], [1.7, 9.8, 0], [0.9, -1.1, 10.2]]), (1, 1, 1), 350
    for seed in range(5):
        s = random_system(box, n, pbc, seed)
        pairs_bf, near = brute_pairs(s.x, box, pbc, cs)
        if near == 0:
            break
    assert near == 0
    orc, _ = make(s, rc, margin)
    orc.calc_forces()
    got = orc.pairs()
    gotset = set(map(tuple, got.tolist()))
    assert len(gotset) == len(got)          # every pair once (half list)
    # positions were wrapped into the box before the list was built, so the
    # image shift of a pair can differ from the brute force's by the wrap
    xw = orc.atoms()["x"]
    pairs_w, near_w = brute_pairs(xw, box, pbc, cs)
    assert near_w == 0
    assert gotset == pairs_w
    assert len(pairs_w) == len(pairs_bf)


def test_open_direction_has_no_images():
    box, pbc = np.diag([9.0, 9.0, 9.0]), (1, 1, 0)
    x = np.array([[1.0, 1.0, 0.2], [1.0, 1.0, 8.9]])   # 0.3 apart through z
    s = W.System(x=x, p=np.zeros_like(x), type=np.zeros(2, np.int32),
                 mass=np.ones(2), id=np.arange(2), box=box, pbc=pbc)
    orc, _ = make(s)
    orc.calc_forces()
    assert len(orc.pairs()) == 0
    s.pbc = (1, 1, 1)
    orc2, _ = make(s)
    orc2.calc_forces()
    assert set(map(tuple, orc2.pairs().tolist())) == {(0, 1, 0, 0, -1)}
