"""The 4-point and spline table interpolations (IMD's FOURPOINT and SPLINE
build options, src/potaccess.h:365-457) on the HIP path against the oracle
with the same interpolation.

The library stores both as cubic segments (two knots each) derived in FP64
from the table rows; the oracle evaluates the reference's own formulas, so
the FP64 bars (1e-10) also pin the segment coefficients, and the mixed bars
are those of the 3-point path.
"""
import numpy as np
import pytest

from imd_amd import imdg as G
from imd_amd import workloads as W

from helpers import (eam_tables, lj, max_abs, oracle_eam, oracle_lj,
                     rel_rms)

pytestmark = pytest.mark.gpu

MODES = [G.INTERP_4POINT, G.INTERP_SPLINE]


def _case(kind, interp):
    if kind == "eam":
        eam = eam_tables()
        s = W.eam_crystal(n_cells=6, temp_k=600.0, seed=23, displace=0.05)
        orc = oracle_eam(s, eam, interp=interp, nbl_margin=0.4, timestep=W.FS)
        make = lambda prec, **kw: G.eam_context(eam, [s], precision=prec,
                                                interpolation=interp, **kw)
        return s, orc, make, True
    if kind == "alloy":
        eam = eam_tables(nspecies=2, rc=5.5)
        s = W.fcc_system((6, 6, 4), 3.615, eam.masses, W.kT(500.0), 43,
                         type_fractions=(0.7, 0.3), pbc=(1, 1, 0),
                         vacuum_z=20.0, displace=0.03)
        orc = oracle_eam(s, eam, interp=interp, nbl_margin=0.4, timestep=W.FS)
        make = lambda prec, **kw: G.eam_context(eam, [s], precision=prec,
                                                interpolation=interp, **kw)
        return s, orc, make, True
    s = W.fcc_system((10, 10, 10), 1.5496, [1.0], 0.7, 1000, displace=0.03)
    orc = oracle_lj(s, interp=interp, nbl_margin=0.3, timestep=0.002)
    make = lambda prec, **kw: G.pair_context(lj(), [s], precision=prec,
                                             interpolation=interp, **kw)
    return s, orc, make, False


@pytest.mark.parametrize("precision", [G.FP64, G.MIXED])
@pytest.mark.parametrize("kind", ["eam", "alloy", "lj"])
@pytest.mark.parametrize("interp", MODES)
def test_interp_static_parity(interp, kind, precision):
    s, orc, make, eam = _case(kind, interp)
    orc.calc_forces()
    with make(precision) as gpu:
        gpu.calc_forces()
        ga, oa = gpu.atoms(), orc.atoms()
        tf, ts = (1e-10, 1e-10) if precision == G.FP64 else (1e-5, 1e-6)
        assert rel_rms(ga["f"], oa["f"]) <= tf, rel_rms(ga["f"], oa["f"])
        assert rel_rms(ga["epot"], oa["epot"]) <= ts
        if eam:
            assert rel_rms(ga["rho"], oa["rho"]) <= ts
            assert rel_rms(ga["df"], oa["df"]) <= ts
        th = gpu.thermo()
        e_o = orc.scalar("epot")
        assert abs(th["epot"] - e_o) <= ts * abs(e_o)
        assert abs(th["virial"] - orc.scalar("virial")) <= ts * orc.scalar("virial_abs")


@pytest.mark.parametrize("interp", MODES)
def test_interp_differs_from_threepoint(interp):
    """The mode reaches the kernels: forces differ from the 3-point ones by
    far more than the parity bar."""
    s, orc, make, _ = _case("eam", interp)
    res = []
    for mode in (G.INTERP_3POINT, interp):
        eam = eam_tables()
        with G.eam_context(eam, [s], precision=G.FP64,
                           interpolation=mode) as gpu:
            gpu.calc_forces()
            res.append(gpu.atoms()["f"])
    assert rel_rms(res[1], res[0]) > 1e-8


@pytest.mark.parametrize("interp", MODES)
def test_interp_trajectory_and_stress_fp64(interp):
    """20 fused steps (rebuilds included) and the stress tensor after them."""
    s, orc, make, _ = _case("eam", interp)
    nsteps = 20
    orc.set_press(True)
    th_o = orc.run(nsteps)
    with make(G.FP64) as gpu:
        gpu.set_stress(True)
        gpu.run(nsteps)
        hist = gpu.thermo_history(0, nsteps)[:, 0, :]
        ga, oa = gpu.atoms(), orc.atoms()
        per, _ = gpu.stress()
    d = ga["x"] - oa["x"]
    L = np.diag(s.box)
    d -= L * np.round(d / L)
    assert np.max(np.abs(d)) < 1e-9
    assert max_abs(ga["p"], oa["p"]) < 1e-9
    assert np.allclose(hist[:, 1], th_o[:, 0], rtol=1e-10, atol=0)
    ref = orc.press()
    assert np.abs(per - ref).max() <= 1e-9 * np.abs(ref).max()
