"""Per-atom stress tensor (IMD's STRESS_TENS) of the HIP path against the
oracle's PRESSTENS on the same seeded inputs.

The tensor of a step is -1/2 sum_j g_ij d d^T over each atom's neighbours
plus p p^T / m with the momenta after the step (components xx yy zz yz zx
xy).  Bars: FP64 within 1e-9 of the largest component (the sums run in a
different order: full list vs half list); mixed precision (FP32 pair
terms, FP64 kinetic part) within 2e-5 of it.  The replica sums are the sums
of the per-atom tensors.
"""
import os

import numpy as np
import pytest

from imd_amd import imdg as G
from imd_amd import workloads as W

from helpers import eam_tables, lj, oracle_eam, oracle_lj

pytestmark = pytest.mark.gpu

TOL = {G.FP64: 1e-9, G.MIXED: 2e-5}


def _case(kind):
    if kind == "eam":
        eam = eam_tables()
        s = W.eam_crystal(n_cells=6, temp_k=600.0, seed=17, displace=0.05)
        orc = oracle_eam(s, eam, nbl_margin=0.4, timestep=W.FS)
        return s, orc, (lambda prec, **kw: G.eam_context(eam, [s], precision=prec, **kw))
    if kind == "alloy":
        eam = eam_tables(nspecies=2, rc=5.5)
        s = W.fcc_system((6, 6, 4), 3.615, eam.masses, W.kT(500.0), 41,
                         type_fractions=(0.7, 0.3), pbc=(1, 1, 0),
                         vacuum_z=20.0, displace=0.03)
        orc = oracle_eam(s, eam, nbl_margin=0.4, timestep=W.FS)
        return s, orc, (lambda prec, **kw: G.eam_context(eam, [s], precision=prec, **kw))
    s = W.fcc_system((10, 10, 10), 1.5496, [1.0], 0.7, 1000, displace=0.03)
    orc = oracle_lj(s, nbl_margin=0.3, timestep=0.002)
    return s, orc, (lambda prec, **kw: G.pair_context(lj(), [s], precision=prec, **kw))


def _check(per, tot, ref, tol):
    scale = np.abs(ref).max()
    err = np.abs(per - ref).max()
    assert err <= tol * scale, (err, scale)
    # a displaced, hot lattice: the off-diagonal terms are not small
    assert np.abs(ref[:, 3:]).max() > 1e-2 * scale
    assert np.allclose(tot[0], per.sum(0), rtol=1e-12, atol=1e-12 * scale)
    assert np.all(np.abs(tot[0] - ref.sum(0)) <= tol * np.abs(ref).sum(0))


@pytest.mark.parametrize("precision", [G.FP64, G.MIXED])
@pytest.mark.parametrize("kind", ["eam", "alloy", "lj"])
def test_stress_after_fused_run(kind, precision):
    """imdg_run's last step (list rebuilds on the way)."""
    s, orc, make = _case(kind)
    nsteps = 15
    orc.set_press(True)
    orc.run(nsteps)
    with make(precision) as gpu:
        gpu.set_stress(True)
        gpu.run(nsteps)
        per, tot = gpu.stress()
    _check(per, tot, orc.press(), TOL[precision])


def test_stress_unfused_steps_and_state_errors():
    """calc_forces + move_atoms (IMD's separate entry points) give the
    tensor of that step; with the tensor off, or after calc_forces alone,
    there is none to read."""
    s, orc, make = _case("eam")
    orc.set_press(True)
    with make(G.FP64) as gpu:
        gpu.calc_forces()
        gpu.move_atoms()
        with pytest.raises(G.ImdgError):
            gpu.stress()            # not enabled
        gpu.set_stress(True)
        for step in range(3):
            orc.calc_forces(step)
            orc.move_atoms()
            orc.check_nblist()
        gpu.check_nblist()
        for step in range(1, 3):
            gpu.calc_forces(step)
            with pytest.raises(G.ImdgError):
                gpu.stress()        # forces of a step not yet moved
            gpu.move_atoms()
            gpu.check_nblist()
        per, tot = gpu.stress()
    _check(per, tot, orc.press(), TOL[G.FP64])


def test_stress_replica_sums_and_graph():
    """Several replicas in one context, the captured step (IMDG_GRAPH=1):
    the sums are per replica, each matching its own oracle."""
    eam = eam_tables()
    systems = [W.fcc_system(d, 3.615, [W.MASS_CU], W.kT(t), seed,
                            vacancy_fraction=v, displace=0.03)
               for d, t, seed, v in [((5, 5, 5), 400.0, 1, 0.0),
                                     ((6, 5, 7), 900.0, 2, 0.01)]]
    nsteps = 12
    old = os.environ.get("IMDG_GRAPH")
    os.environ["IMDG_GRAPH"] = "1"
    try:
        gpu = G.eam_context(eam, systems, precision=G.FP64)
    finally:
        if old is None:
            del os.environ["IMDG_GRAPH"]
        else:
            os.environ["IMDG_GRAPH"] = old
    with gpu:
        gpu.set_stress(True)
        gpu.run(5)
        gpu.run(nsteps - 5)   # graph replays, the tensor after the last
        per, tot = gpu.stress()
    off = 0
    for r, s in enumerate(systems):
        orc = oracle_eam(s, eam, nbl_margin=0.4, timestep=W.FS)
        orc.set_press(True)
        orc.run(nsteps)
        ref = orc.press()
        sl = slice(off, off + s.n)
        scale = np.abs(ref).max()
        assert np.abs(per[sl] - ref).max() <= 1e-9 * scale
        assert np.allclose(tot[r], per[sl].sum(0), rtol=1e-12,
                           atol=1e-12 * scale)
        off += s.n
