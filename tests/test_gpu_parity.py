"""Parity of the HIP path (through the C ABI) with the oracle on the same
seeded inputs.

Bars:
* cell assignment and the neighbour-pair set: bit-exact as sets;
* FP64 path: forces, rho, 2F', per-atom energies, Epot and virial within
  1e-10 relative (summation order differs: full list vs half list);
* mixed path (FP32 forces/tables, FP64 integration): RMS relative force
  error <= 1e-5, rho/2F'/per-atom energy <= 1e-6 relative, Epot and virial
  <= 1e-6 relative;
* short trajectories: FP64 GPU and oracle agree to 1e-9 in positions after
  40 steps (the reference operation order is kept in the integrator).
"""
import numpy as np
import pytest

from imd_amd import imdg as G
from imd_amd import workloads as W

from helpers import (canon_pairs_gpu, canon_pairs_oracle, eam_tables, lj,
                     max_abs, oracle_eam, oracle_lj, rel_rms)

pytestmark = pytest.mark.gpu


def _eam_system(n_cells=6, seed=7, temp_k=600.0, displace=0.05):
    return W.eam_crystal(n_cells=n_cells, temp_k=temp_k, seed=seed,
                         displace=displace)


def _compare_static(gpu, orc, n, tol_f, tol_s, has_eam=True):
    ga = gpu.atoms()
    oa = orc.atoms()
    assert rel_rms(ga["f"], oa["f"]) <= tol_f, rel_rms(ga["f"], oa["f"])
    assert rel_rms(ga["epot"], oa["epot"]) <= tol_s
    if has_eam:
        assert rel_rms(ga["rho"], oa["rho"]) <= tol_s
        assert rel_rms(ga["df"], oa["df"]) <= tol_s
    th = gpu.thermo()
    e_o, v_o = orc.scalar("epot"), orc.scalar("virial")
    assert abs(th["epot"] - e_o) <= tol_s * abs(e_o)
    # the virial is a sum of pair terms of both signs that mostly cancel
    # (sum |terms| / |sum| is about 1e3 in a crystal); its error is bounded
    # relative to the sum of |terms|, the scale its rounding works on
    assert abs(th["virial"] - v_o) <= tol_s * orc.scalar("virial_abs")


@pytest.mark.parametrize("precision", [G.FP64, G.MIXED])
def test_eam_static_parity(precision):
    eam = eam_tables()
    s = _eam_system()
    orc = oracle_eam(s, eam, nbl_margin=0.4, timestep=W.FS)
    orc.calc_forces()
    with G.eam_context(eam, [s], precision=precision) as gpu:
        gpu.calc_forces()
        assert gpu.grid() == orc.grid()
        # cell assignment: bit-exact
        assert np.array_equal(gpu.cells(), orc.cells())
        # neighbour pairs: bit-exact as sets
        gp = canon_pairs_gpu(gpu.pairs())
        op = canon_pairs_oracle(orc.pairs())
        assert len(op) == len(orc.pairs())
        assert gp == op
        if precision == G.FP64:
            _compare_static(gpu, orc, s.n, 1e-10, 1e-10)
        else:
            _compare_static(gpu, orc, s.n, 1e-5, 1e-6)
        assert np.allclose(gpu.atoms()["x"], orc.atoms()["x"], rtol=0,
                           atol=0)


@pytest.mark.parametrize("precision", [G.FP64, G.MIXED])
def test_lj_pair_only_parity(precision):
    # displaced lattice: on the perfect lattice every force vanishes by
    # symmetry and a relative error is meaningless
    s = W.fcc_system((10, 10, 10), 1.5496, [1.0], 0.7, 1000, displace=0.03)
    orc = oracle_lj(s, nbl_margin=0.3, timestep=0.002)
    orc.calc_forces()
    with G.pair_context(lj(), [s], precision=precision) as gpu:
        gpu.calc_forces()
        assert gpu.grid() == orc.grid() == (5, 5, 5)
        assert np.array_equal(gpu.cells(), orc.cells())
        assert canon_pairs_gpu(gpu.pairs()) == canon_pairs_oracle(orc.pairs())
        tol = (1e-10, 1e-10) if precision == G.FP64 else (1e-5, 1e-6)
        _compare_static(gpu, orc, s.n, *tol, has_eam=False)


def test_eam_trajectory_fp64_nve():
    eam = eam_tables()
    s = _eam_system(n_cells=6, seed=11, temp_k=900.0, displace=0.02)
    nsteps = 40
    orc = oracle_eam(s, eam, nbl_margin=0.4, timestep=W.FS)
    th_o = orc.run(nsteps)
    with G.eam_context(eam, [s], precision=G.FP64) as gpu:
        gpu.run(nsteps)
        hist = gpu.thermo_history(0, nsteps)[:, 0, :]
        ga, oa = gpu.atoms(), orc.atoms()
        # positions may differ by a box vector where one side wrapped at a
        # rebuild; compare minimum-image differences
        d = ga["x"] - oa["x"]
        L = np.diag(s.box)
        d -= L * np.round(d / L)
        assert np.max(np.abs(d)) < 1e-9
        assert max_abs(ga["p"], oa["p"]) < 1e-9
        assert np.allclose(hist[:, 1], th_o[:, 0], rtol=1e-10, atol=0)
        assert np.allclose(hist[:, 2], th_o[:, 1], rtol=1e-10, atol=0)
        assert np.array_equal(hist[:, 6] > 0, th_o[:, 4] > 0)


def test_eam_trajectory_fp64_nvt_unfused():
    """IMD's separate entry points: calc_forces, move_atoms, check_nblist."""
    eam = eam_tables()
    s = _eam_system(n_cells=6, seed=12, temp_k=300.0, displace=0.02)
    T = W.kT(800.0)
    isq = 1.0 / (100 * W.FS) ** 2
    orc = oracle_eam(s, eam, nbl_margin=0.4, timestep=W.FS,
                     ensemble=1, temperature=T, eta=0.0, isq_tau_eta=isq)
    with G.eam_context(eam, [s], precision=G.FP64, ensemble=G.NVT) as gpu:
        gpu.set_nvt(0, T, 0.0, isq)
        for step in range(25):
            orc.calc_forces(step)
            orc.move_atoms()
            orc.check_nblist()
            gpu.calc_forces(step)
            gpu.move_atoms()
            gpu.check_nblist()
            th = gpu.thermo()
            assert abs(th["eta"] - orc.scalar("eta")) <= 1e-10 * max(1.0, abs(orc.scalar("eta")))
            assert abs(th["ekin"] - orc.scalar("ekin")) <= 1e-10 * abs(orc.scalar("ekin"))
        ga, oa = gpu.atoms(), orc.atoms()
        assert max_abs(ga["p"], oa["p"]) < 1e-9


def test_mixed_force_error_vs_fp64():
    """Mixed vs FP64 GPU on the same configuration: the accuracy target of
    <= 1e-5 relative RMS force error."""
    eam = eam_tables()
    s = _eam_system(n_cells=8, seed=21, temp_k=800.0, displace=0.08)
    res = {}
    for prec in (G.FP64, G.MIXED):
        with G.eam_context(eam, [s], precision=prec) as gpu:
            gpu.calc_forces()
            res[prec] = (gpu.atoms(), gpu.thermo())
    f64, fm = res[G.FP64][0]["f"], res[G.MIXED][0]["f"]
    assert rel_rms(fm, f64) <= 1e-5
    assert abs(res[G.MIXED][1]["epot"] - res[G.FP64][1]["epot"]) <= 1e-6 * abs(res[G.FP64][1]["epot"])


def test_alloy_slab_open_z():
    """Two species, three pair columns, open z with vacuum."""
    eam = eam_tables(nspecies=2, rc=5.5)
    s = W.fcc_system((6, 6, 4), 3.615, eam.masses, W.kT(500.0), 31,
                     type_fractions=(0.7, 0.3), pbc=(1, 1, 0), vacuum_z=20.0,
                     displace=0.03)
    orc = oracle_eam(s, eam, nbl_margin=0.4, timestep=W.FS)
    orc.calc_forces()
    with G.eam_context(eam, [s], precision=G.FP64) as gpu:
        gpu.calc_forces()
        assert gpu.grid() == orc.grid()
        assert np.array_equal(gpu.cells(), orc.cells())
        assert canon_pairs_gpu(gpu.pairs()) == canon_pairs_oracle(orc.pairs())
        _compare_static(gpu, orc, s.n, 1e-10, 1e-10)


def test_batched_replicas_match_individual_oracles():
    eam = eam_tables()
    systems = [W.fcc_system(d, 3.615, [W.MASS_CU], W.kT(t), seed,
                            vacancy_fraction=v, displace=0.02)
               for d, t, seed, v in [((5, 5, 5), 300.0, 1, 0.0),
                                     ((6, 5, 7), 700.0, 2, 0.01),
                                     ((5, 6, 5), 1100.0, 3, 0.005)]]
    nsteps = 20
    with G.eam_context(eam, systems, precision=G.FP64) as gpu:
        gpu.calc_forces()
        ga = gpu.atoms()
        pairs = gpu.pairs()
        off = 0
        for r, s in enumerate(systems):
            orc = oracle_eam(s, eam, nbl_margin=0.4, timestep=W.FS)
            orc.calc_forces()
            sl = slice(off, off + s.n)
            assert canon_pairs_gpu(pairs, r) == canon_pairs_oracle(orc.pairs())
            assert rel_rms(ga["f"][sl], orc.atoms()["f"]) <= 1e-10
            assert abs(gpu.thermo(r)["epot"] - orc.scalar("epot")) <= 1e-10 * abs(orc.scalar("epot"))
            off += s.n
        gpu.run(nsteps)
        hist = gpu.thermo_history(0, nsteps)
        for r, s in enumerate(systems):
            orc = oracle_eam(s, eam, nbl_margin=0.4, timestep=W.FS)
            th = orc.run(nsteps)
            assert np.allclose(hist[:, r, 1], th[:, 0], rtol=1e-10)
            assert np.allclose(hist[:, r, 2], th[:, 1], rtol=1e-9)


def test_list_overflow_grows_and_matches():
    eam = eam_tables()
    s = _eam_system(n_cells=5, seed=5)
    with G.eam_context(eam, [s], precision=G.FP64, nbl_size=0.3) as small:
        k0 = small.info()["kmax"]
        small.run(10)
        info = small.info()
        assert info["kmax"] > k0 and info["kmax"] >= info["kmax_seen"]
        a_small = small.atoms()
    with G.eam_context(eam, [s], precision=G.FP64) as big:
        big.run(10)
        a_big = big.atoms()
    assert max_abs(a_small["x"], a_big["x"]) == 0.0


@pytest.mark.parametrize("lanes", [1, 2, 4])
def test_alloy_mixed_shared_grid(lanes):
    """Two species whose pair and density columns share one grid, mixed
    precision: the force loop reads one record per type pair and segment
    and the density loop one (the several-species shared-grid path).  Same
    bars as the one-species mixed path; the generic several-species loop
    (IMDG_JOINT_TABLE=0) agrees with it to the FP32 rounding."""
    import os
    eam = eam_tables(nspecies=2, rc=5.5)
    s = W.fcc_system((6, 6, 4), 3.615, eam.masses, W.kT(500.0), 37,
                     type_fractions=(0.6, 0.4), pbc=(1, 1, 0), vacuum_z=20.0,
                     displace=0.04)
    orc = oracle_eam(s, eam, nbl_margin=0.4, timestep=W.FS)
    orc.calc_forces()
    res = {}
    for joint in ("1", "0"):
        old = os.environ.get("IMDG_JOINT_TABLE")
        os.environ["IMDG_JOINT_TABLE"] = joint
        try:
            gpu = G.eam_context(eam, [s], precision=G.MIXED)
        finally:
            if old is None:
                del os.environ["IMDG_JOINT_TABLE"]
            else:
                os.environ["IMDG_JOINT_TABLE"] = old
        with gpu:
            gpu.set_tuning(lanes, 512, lanes, 512)
            gpu.calc_forces()
            _compare_static(gpu, orc, s.n, 1e-5, 1e-6)
            res[joint] = gpu.atoms()
    assert rel_rms(res["1"]["f"], res["0"]["f"]) <= 2e-6
    assert rel_rms(res["1"]["rho"], res["0"]["rho"]) <= 1e-6


def test_alloy_mixed_trajectory():
    """20 fused NVT steps of the two-species slab in mixed precision follow
    the FP64 oracle (energies to 1e-6, momenta to the mixed force error)."""
    eam = eam_tables(nspecies=2, rc=5.5)
    s = W.fcc_system((6, 6, 4), 3.615, eam.masses, W.kT(500.0), 38,
                     type_fractions=(0.7, 0.3), pbc=(1, 1, 0), vacuum_z=20.0,
                     displace=0.02)
    T, isq = W.kT(500.0), 1.0 / (50 * W.FS) ** 2
    orc = oracle_eam(s, eam, nbl_margin=0.4, timestep=W.FS, ensemble=1,
                     temperature=T, eta=0.0, isq_tau_eta=isq)
    th = orc.run(20)
    with G.eam_context(eam, [s], precision=G.MIXED, ensemble=G.NVT) as gpu:
        gpu.set_nvt(0, T, 0.0, isq)
        gpu.run(20)
        hist = gpu.thermo_history(0, 20)[:, 0, :]
        p = gpu.atoms()["p"]
    assert np.allclose(hist[:, 1], th[:, 0], rtol=1e-6, atol=0)
    assert rel_rms(p, orc.atoms()["p"]) <= 1e-5
