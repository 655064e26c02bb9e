"""The list decision at the radius: IMD's strict r^2 < cellsz on images
built z, then y, then x (src/imd_forces_nbl.c:249-261,
src/imd_comm_force_3d.c:754-756), for pairs placed within a few ulps of
cellsz across every periodic face, edge and corner image, inside one cell,
and with either atom as the owner, in a cubic and a sheared box.

* CPU: the oracle's pair sets equal the brute-force reference expression
  (radius_cases.ref_pairs), and the designated pairs are in the list
  exactly when their r^2 lies below cellsz;
* CPU: the FP32 r^2 the GPU list build screens with (cell-relative records
  and offsets rounded to FP32) stays within 2e-6 relative of the FP64 r^2
  on random pairs near the radius, far inside the 1e-4 band that the FP64
  expression decides;
* GPU: the pair sets of both precisions equal the oracle's, and the FP64
  decision ran (imdg_info.exact_pairs > 0).
"""
import math

import numpy as np
import pytest

from imd_amd import imdg as G

from helpers import canon_pairs_gpu, canon_pairs_oracle, lj, oracle_lj
from radius_cases import (GEOMS, KS, cell_of, lj_cellsz, offset_and_shift,
                          pair_r2, radius_system, ref_pairs)


def _cases(sheared):
    cellsz = lj_cellsz()
    out = []
    for k in KS:
        s = radius_system(k, cellsz, sheared)
        orc = oracle_lj(s, nbl_margin=0.3, timestep=0.002)
        assert orc.scalar("cellsz") == cellsz
        orc.calc_forces()
        out.append((k, s, orc))
    return cellsz, out


@pytest.mark.parametrize("sheared", [False, True])
def test_oracle_pairs_at_radius(sheared):
    cellsz, cases = _cases(sheared)
    for k, s, orc in cases:
        assert orc.grid() == (3, 3, 3)
        cells = orc.cells()
        for i in range(s.n):
            assert tuple(cells[i]) == cell_of(s.box, s.x[i])
        got = canon_pairs_oracle(orc.pairs())
        assert got == ref_pairs(s, cellsz)
        for g in range(len(GEOMS)):
            i, j = 2 * g, 2 * g + 1
            r2, sh = pair_r2(s.box, s.x[i], s.x[j], cell_of(s.box, s.x[i]),
                             cell_of(s.box, s.x[j]))
            assert abs(r2 / cellsz - 1.0) < 2e-14
            assert (r2 < cellsz) == (k < 0)
            a, b = int(s.id[i]), int(s.id[j])
            key = ((a, b) + tuple(sh)) if a < b else \
                ((b, a) + tuple(-v for v in sh))
            assert (key in got) == (k < 0)


def _fp32_r2(box, cellsz_g, xi, xj, ci, cj, fma):
    """r^2 as the list build screens it: u = x - centre(cell) and the cell
    offset o . cvec rounded to FP32, then w = o - u_i, d = u_j + w and
    d0*d0 + d1*d1 + d2*d2 in FP32 (with fma: contracted as a compiler may)."""
    f32 = np.float32
    cvec = np.asarray(box, dtype=np.float64) / cellsz_g

    def centre(c):
        return sum((c[d] + 0.5) * cvec[d] for d in range(3))

    o, _ = offset_and_shift(ci, cj)
    ui = (np.asarray(xi) - centre(ci)).astype(f32)
    uj = (np.asarray(xj) - centre(cj)).astype(f32)
    off = sum(o[d] * cvec[d] for d in range(3)).astype(f32)
    w = (off - ui).astype(f32)
    d = (uj + w).astype(f32)
    if not fma:
        return float(f32(f32(d[0] * d[0] + d[1] * d[1]) + d[2] * d[2]))
    a = f32(np.float64(d[0]) * d[0])
    a = f32(np.float64(d[1]) * d[1] + a)
    return float(f32(np.float64(d[2]) * d[2] + a))


@pytest.mark.parametrize("fma", [False, True])
def test_fp32_screen_error_bound(fma):
    cellsz = lj_cellsz()
    L = 10.0
    box = np.eye(3) * L
    box[1] = (1.5, L, 0.0)
    rng = np.random.default_rng(5)
    worst = 0.0
    for _ in range(4000):
        xi = rng.uniform(0.0, 1.0, 3) @ box
        u = rng.normal(size=3)
        u /= np.linalg.norm(u)
        y = xi + math.sqrt(cellsz) * u * rng.uniform(0.999, 1.001)
        xj = y - np.floor(np.linalg.solve(box.T, y)) @ box
        ci, cj = cell_of(box, xi), cell_of(box, xj)
        o, _ = offset_and_shift(ci, cj)
        r2, _ = pair_r2(box, xi, xj, ci, cj)
        # the screen evaluates the pair from atom i's list; seen from j the
        # roles swap, so check both directions
        for a, b, ca, cb in ((xi, xj, ci, cj), (xj, xi, cj, ci)):
            r32 = _fp32_r2(box, 3, a, b, ca, cb, fma)
            worst = max(worst, abs(r32 - r2) / cellsz)
    assert worst < 2e-6, worst
    assert worst > 1e-9          # FP32 rounding is visible at all


def _by_id(pairs, s):
    """GPU pairs name atoms by input index, the oracle by atom number (the
    ids of these systems are swapped within some pairs): rename, then put
    the lower number first with the shift of the higher relative to it."""
    out = set()
    for a, b, *sh in pairs:
        ia, ib = int(s.id[a]), int(s.id[b])
        out.add((ia, ib, *sh) if ia < ib else (ib, ia, *(-v for v in sh)))
    return out


@pytest.mark.gpu
@pytest.mark.parametrize("precision", [G.FP64, G.MIXED])
@pytest.mark.parametrize("sheared", [False, True])
def test_gpu_pairs_at_radius(precision, sheared):
    cellsz, cases = _cases(sheared)
    systems = [s for _, s, _ in cases]
    with G.pair_context(lj(), systems, precision=precision) as gpu:
        gpu.calc_forces()
        pairs = gpu.pairs()
        for r, (k, s, orc) in enumerate(cases):
            assert gpu.grid(r) == (3, 3, 3)
            assert _by_id(canon_pairs_gpu(pairs, r), s) == \
                canon_pairs_oracle(orc.pairs()), k
        info = gpu.info()
        # every designated pair is a band candidate from both of its atoms
        assert info["exact_pairs"] >= 2 * len(GEOMS) * len(KS)
