"""Launch shapes of the neighbour loops (lanes per atom x threads per
block) and the autotuner.  Every shape must meet the same parity bars
against the oracle as the default one (results differ between shapes only
by the order of floating-point sums); the autotuner must not advance the
state."""
import numpy as np
import pytest

from imd_amd import imdg as G
from imd_amd import workloads as W

from helpers import (canon_pairs_gpu, canon_pairs_oracle, eam_tables, lj,
                     max_abs, oracle_eam, oracle_lj, rel_rms)

pytestmark = pytest.mark.gpu

TOL = {G.FP64: (1e-10, 1e-10), G.MIXED: (1e-5, 1e-6)}


def _check_static(gpu, orc, tol_f, tol_s, has_eam=True):
    ga, oa = gpu.atoms(), orc.atoms()
    assert rel_rms(ga["f"], oa["f"]) <= tol_f, rel_rms(ga["f"], oa["f"])
    assert rel_rms(ga["epot"], oa["epot"]) <= tol_s
    if has_eam:
        assert rel_rms(ga["rho"], oa["rho"]) <= tol_s
        assert rel_rms(ga["df"], oa["df"]) <= tol_s
    th = gpu.thermo()
    assert abs(th["epot"] - orc.scalar("epot")) <= tol_s * abs(orc.scalar("epot"))
    assert abs(th["virial"] - orc.scalar("virial")) <= tol_s * orc.scalar("virial_abs")


def _max_threads(precision):
    return 768 if precision == G.MIXED else 512


@pytest.mark.parametrize("precision", [G.FP64, G.MIXED])
@pytest.mark.parametrize("lanes", [1, 2, 4])
def test_eam_shapes_match_oracle(precision, lanes):
    eam = eam_tables()
    s = W.eam_crystal(n_cells=6, temp_k=600.0, seed=41, displace=0.05)
    orc = oracle_eam(s, eam, nbl_margin=0.4, timestep=W.FS)
    orc.calc_forces()
    with G.eam_context(eam, [s], precision=precision) as gpu:
        for dthr, fthr in ((256, 256), (1024, _max_threads(precision))):
            gpu.set_tuning(lanes, dthr, lanes, fthr)
            gpu.calc_forces()
            t = gpu.tuning()
            assert (t["density_lanes"], t["force_lanes"]) == (lanes, lanes)
            assert gpu.info()["lanes_per_atom"] == lanes
            _check_static(gpu, orc, *TOL[precision])
        assert canon_pairs_gpu(gpu.pairs()) == canon_pairs_oracle(orc.pairs())


@pytest.mark.parametrize("lanes", [2, 4])
def test_alloy_and_pair_only_shapes(lanes):
    # several species (per-pair column lookups) and a pair-only table
    eam = eam_tables(nspecies=2, rc=5.5)
    s = W.fcc_system((6, 6, 4), 3.615, eam.masses, W.kT(500.0), 31,
                     type_fractions=(0.7, 0.3), pbc=(1, 1, 0), vacuum_z=20.0,
                     displace=0.03)
    orc = oracle_eam(s, eam, nbl_margin=0.4, timestep=W.FS)
    orc.calc_forces()
    with G.eam_context(eam, [s], precision=G.FP64) as gpu:
        gpu.set_tuning(lanes, 512, lanes, 512)
        gpu.calc_forces()
        _check_static(gpu, orc, 1e-10, 1e-10)
    s = W.fcc_system((10, 10, 10), 1.5496, [1.0], 0.7, 1000, displace=0.03)
    orc = oracle_lj(s, nbl_margin=0.3, timestep=0.002)
    orc.calc_forces()
    for prec in (G.FP64, G.MIXED):
        with G.pair_context(lj(), [s], precision=prec) as gpu:
            gpu.set_tuning(lanes, 512, lanes, 512)
            gpu.calc_forces()
            _check_static(gpu, orc, *TOL[prec], has_eam=False)


def test_fused_trajectory_four_lanes_fp64():
    # the fused step integrates slot epi_slot(lane) after the lane passes:
    # a 40-step FP64 trajectory must still follow the oracle to 1e-9
    eam = eam_tables()
    s = W.eam_crystal(n_cells=6, temp_k=900.0, seed=11, displace=0.02)
    nsteps = 40
    orc = oracle_eam(s, eam, nbl_margin=0.4, timestep=W.FS)
    th_o = orc.run(nsteps)
    with G.eam_context(eam, [s], precision=G.FP64, lanes_per_atom=4) as gpu:
        assert gpu.tuning()["force_lanes"] == 4
        gpu.run(nsteps)
        hist = gpu.thermo_history(0, nsteps)[:, 0, :]
        ga, oa = gpu.atoms(), orc.atoms()
        d = ga["x"] - oa["x"]
        L = np.diag(s.box)
        d -= L * np.round(d / L)
        assert np.max(np.abs(d)) < 1e-9
        assert max_abs(ga["p"], oa["p"]) < 1e-9
        assert np.allclose(hist[:, 1], th_o[:, 0], rtol=1e-10, atol=0)


def test_batched_replicas_four_lanes():
    eam = eam_tables()
    systems = [W.fcc_system(d, 3.615, [W.MASS_CU], W.kT(t), seed,
                            vacancy_fraction=v, displace=0.02)
               for d, t, seed, v in [((5, 5, 5), 300.0, 1, 0.0),
                                     ((6, 5, 7), 700.0, 2, 0.01),
                                     ((5, 6, 5), 1100.0, 3, 0.005)]]
    nsteps = 20
    with G.eam_context(eam, systems, precision=G.FP64,
                       lanes_per_atom=4) as gpu:
        gpu.run(nsteps)
        hist = gpu.thermo_history(0, nsteps)
        for r, s in enumerate(systems):
            orc = oracle_eam(s, eam, nbl_margin=0.4, timestep=W.FS)
            th = orc.run(nsteps)
            assert np.allclose(hist[:, r, 1], th[:, 0], rtol=1e-10)
            assert np.allclose(hist[:, r, 2], th[:, 1], rtol=1e-9)


@pytest.mark.parametrize("precision", [G.FP64, G.MIXED])
def test_autotune_keeps_state(precision):
    eam = eam_tables()
    s = W.eam_crystal(n_cells=8, temp_k=700.0, seed=43, displace=0.03)
    with G.eam_context(eam, [s], precision=precision) as gpu:
        gpu.run(15)
        before = gpu.atoms()
        steps = gpu.info()["steps_done"]
        t = gpu.autotune(reps=2)
        assert t["density_lanes"] in (1, 2, 4) and t["force_lanes"] in (1, 2, 4)
        assert t["density_ms"] > 0 and t["force_ms"] > 0
        assert gpu.tuning()["force_lanes"] == t["force_lanes"]
        after = gpu.atoms()
        assert gpu.info()["steps_done"] == steps
        assert np.array_equal(after["x"], before["x"])
        assert np.array_equal(after["p"], before["p"])
        # forces and energies as calc_forces leaves them with that shape
        gpu.calc_forces()
        again = gpu.atoms()
        for k in ("f", "epot", "rho", "df"):
            assert np.array_equal(again[k], after[k]), k
        # and the run continues on the chosen shape like an untuned one
        gpu.run(10)
        assert gpu.info()["steps_done"] == steps + 10


def test_set_tuning_rejects_bad_shapes():
    eam = eam_tables()
    s = W.eam_crystal(n_cells=5, temp_k=300.0, seed=3, displace=0.02)
    with G.eam_context(eam, [s], precision=G.MIXED) as gpu:
        old = gpu.tuning()
        for bad in ((3, 512, 1, 512), (1, 100, 1, 512), (1, 512, 8, 512),
                    (1, 512, 1, 2048)):
            with pytest.raises(G.ImdgError):
                gpu.set_tuning(*bad)
        assert gpu.tuning() == old
