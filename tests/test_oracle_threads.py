"""The oracle's threaded timing variant (orc_set_threads), used for the CPU
baseline on several host cores, against its serial run (IMD's order, the
checker): the same neighbour list entry for entry, forces and energies equal
up to the reordering of sums, and a short NVT trajectory that stays on the
serial one."""
import numpy as np
import pytest

from imd_amd import workloads as W
from helpers import eam_tables, oracle_eam, oracle_lj


def _alloy(seed=21):
    eam = eam_tables(nspecies=2, rc=5.5)
    s = W.fcc_system((6, 6, 5), 3.615, eam.masses, W.kT(500.0), seed,
                     type_fractions=(0.7, 0.3), pbc=(1, 1, 0), vacuum_z=20.0,
                     displace=0.02)
    return eam, s


@pytest.mark.parametrize("threads", [2, 5])
def test_threaded_list_and_forces(threads):
    eam, s = _alloy()
    a = oracle_eam(s, eam, nbl_margin=0.4, timestep=W.FS)
    b = oracle_eam(s, eam, nbl_margin=0.4, timestep=W.FS)
    b.set_threads(threads)
    a.calc_forces()
    b.calc_forces()
    assert np.array_equal(a.pairs(), b.pairs())
    fa, fb = a.atoms(), b.atoms()
    scale = np.abs(fa["f"]).max()
    assert np.abs(fa["f"] - fb["f"]).max() <= 1e-12 * scale
    assert np.allclose(fa["rho"], fb["rho"], rtol=1e-13, atol=0)
    assert np.allclose(fa["epot"], fb["epot"], rtol=1e-12, atol=1e-14)
    for k in ("epot", "virial"):
        assert abs(a.scalar(k) - b.scalar(k)) <= 1e-12 * abs(a.scalar(k))
    # the per-thread copies are cleared on every call
    b.calc_forces()
    assert np.abs(b.atoms()["f"] - fb["f"]).max() <= 1e-15 * scale


def test_threaded_nvt_trajectory():
    eam, s = _alloy(seed=22)
    kw = dict(nbl_margin=0.4, timestep=W.FS, ensemble=1,
              temperature=W.kT(500.0), eta=0.0,
              isq_tau_eta=1.0 / (50 * W.FS) ** 2)
    a = oracle_eam(s, eam, **kw)
    b = oracle_eam(s, eam, **kw)
    b.set_threads(4)
    ta, tb = a.run(30), b.run(30)
    assert np.array_equal(ta[:, 4], tb[:, 4])  # same rebuild steps
    assert np.allclose(ta[:, :4], tb[:, :4], rtol=1e-10, atol=1e-14)
    assert np.abs(a.atoms()["x"] - b.atoms()["x"]).max() < 1e-10


def test_threaded_pair_only():
    s = W.lj_crystal(n_cells=6, temp=0.7, seed=5)
    a = oracle_lj(s, nbl_margin=0.3, timestep=0.002)
    b = oracle_lj(s, nbl_margin=0.3, timestep=0.002)
    b.set_threads(3)
    ta, tb = a.run(10), b.run(10)
    assert np.allclose(ta[:, :3], tb[:, :3], rtol=1e-11, atol=0)
