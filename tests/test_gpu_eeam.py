"""EEAM on the HIP path against the oracle (src/imd_forces_nbl.c:591-608,
1090-1095, 1180-1207): the second density P_i = sum_j rho_ij^2, the energy
term E(P) and the force term 2E'(P_i) rho_j rho'_j + 2E'(P_j) rho_i rho'_i.
Bars as for plain EAM: FP64 1e-10, mixed 1e-5 RMS forces and 1e-6 for the
per-atom scalars."""
import numpy as np
import pytest

import oracle as O
from imd_amd import imdg as G
from imd_amd import workloads as W

from helpers import eam_tables, emod_table, max_abs, oracle_eam, rel_rms

pytestmark = pytest.mark.gpu


def _case(nt, interp=G.INTERP_3POINT):
    if nt == 1:
        eam = eam_tables()
        s = W.eam_crystal(n_cells=6, temp_k=600.0, seed=29, displace=0.05)
    else:
        eam = eam_tables(nspecies=2, rc=5.5)
        s = W.fcc_system((6, 6, 4), 3.615, eam.masses, W.kT(500.0), 47,
                         type_fractions=(0.7, 0.3), pbc=(1, 1, 0),
                         vacuum_z=20.0, displace=0.03)
    et = emod_table(nt)
    orc = oracle_eam(s, eam, interp=interp, nbl_margin=0.4, timestep=W.FS)
    emod = O.OracleTable.from_table(et, interp=interp)
    orc.set_eeam(emod)
    orc._keep_emod = emod
    make = lambda prec: G.eam_context(eam, [s], precision=prec, emod=et,
                                      interpolation=interp)
    return s, orc, make


@pytest.mark.parametrize("precision", [G.FP64, G.MIXED])
@pytest.mark.parametrize("nt", [1, 2])
def test_eeam_static_parity(nt, precision):
    s, orc, make = _case(nt)
    orc.calc_forces()
    with make(precision) as gpu:
        gpu.calc_forces()
        ga, oa = gpu.atoms(), orc.atoms()
        P, dm = gpu.eeam()
    oP, odm = orc.eeam()
    tf, ts = (1e-10, 1e-10) if precision == G.FP64 else (1e-5, 1e-6)
    assert rel_rms(ga["f"], oa["f"]) <= tf, rel_rms(ga["f"], oa["f"])
    for a, b in ((ga["epot"], oa["epot"]), (ga["rho"], oa["rho"]),
                 (ga["df"], oa["df"]), (P, oP)):
        assert rel_rms(a, b) <= ts, rel_rms(a, b)
    # E'(P) of this table is a difference of two terms about 20 times its
    # size: its FP32 error is 1e-6 of those terms
    assert rel_rms(dm, odm) <= (ts if precision == G.FP64 else 2e-5)
    # the EEAM term changes the forces well beyond the bar
    orc2 = oracle_eam(s, eam_tables() if nt == 1 else
                      eam_tables(nspecies=2, rc=5.5), nbl_margin=0.4,
                      timestep=W.FS)
    orc2.calc_forces()
    assert rel_rms(orc2.atoms()["f"], oa["f"]) > 1e-3


def test_eeam_trajectory_and_stress_fp64():
    s, orc, make = _case(1)
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


@pytest.mark.parametrize("precision", [G.FP64, G.MIXED])
def test_eeam_with_spline_tables(precision):
    s, orc, make = _case(2, interp=G.INTERP_SPLINE)
    orc.calc_forces()
    with make(precision) as gpu:
        gpu.calc_forces()
        ga, oa = gpu.atoms(), orc.atoms()
    tf, ts = (1e-10, 1e-10) if precision == G.FP64 else (1e-5, 1e-6)
    assert rel_rms(ga["f"], oa["f"]) <= tf
    assert rel_rms(ga["epot"], oa["epot"]) <= ts
