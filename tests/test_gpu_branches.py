"""Branches the ordinary parity cases do not reach.

* The list build's fallback for blocks whose neighbour tile does not fit in
  LDS (their candidates are read from HBM): a lowered tile cap
  (IMDG_TILE_LIMIT) sends blocks there; the pair sets and forces still
  match the oracle, and a fused run gives bit-identical results with and
  without the fallback (the list is the same).
* The neighbour loops' kernel instances that load the list words
  non-temporally (chosen for lists larger than half the last-level cache;
  IMDG_LIST_NT=1 forces them): the cache policy changes no result, so
  fused runs with and without them are bit-identical (one species, tables
  in LDS; two species on a shared grid, tables in HBM/L2).
* Two species on one grid whose density columns depend on the
  neighbour's type only (the usual EAM form): the loops read per-type
  density segments from LDS, with the same FP32 values as the records in
  HBM/L2 (IMDG_RHO_BY_TYPE=0), so runs with and without them are
  bit-identical; density columns that depend on both types keep the
  records, and match the oracle; with three species only the force loop's
  per-type segments fit in LDS.
* Pairs closer than a table's begin: IMD clamps r^2 to begin and sets
  is_short (src/potaccess.h:329-334, src/imd_forces_nbl.c:982, 1323); the
  forces at the clamped argument and the flag must match the oracle, and
  imdg_get_info reports the flag once.
"""
import copy

import numpy as np
import pytest

from imd_amd import imdg as G
from imd_amd import workloads as W

from helpers import (canon_pairs_gpu, canon_pairs_oracle, eam_tables, lj,
                     oracle_eam, oracle_lj, rel_rms)

pytestmark = pytest.mark.gpu

BARS = {G.FP64: (1e-10, 1e-10), G.MIXED: (1e-5, 1e-6)}


def _static(gpu, orc, precision, has_eam=True):
    tol_f, tol_s = BARS[precision]
    ga, oa = gpu.atoms(), orc.atoms()
    assert rel_rms(ga["f"], oa["f"]) <= tol_f
    assert rel_rms(ga["epot"], oa["epot"]) <= tol_s
    if has_eam:
        assert rel_rms(ga["rho"], oa["rho"]) <= tol_s
        assert rel_rms(ga["df"], oa["df"]) <= tol_s
    th = gpu.thermo()
    assert abs(th["epot"] - orc.scalar("epot")) <= tol_s * abs(orc.scalar("epot"))
    assert abs(th["virial"] - orc.scalar("virial")) <= \
        tol_s * orc.scalar("virial_abs")


@pytest.mark.parametrize("precision", [G.FP64, G.MIXED])
def test_tile_fallback(precision, monkeypatch):
    eam = eam_tables()
    s = W.eam_crystal(n_cells=8, temp_k=900.0, seed=41, displace=0.04)
    orc = oracle_eam(s, eam, nbl_margin=0.4, timestep=W.FS)
    orc.calc_forces()
    monkeypatch.setenv("IMDG_TILE_LIMIT", "400")
    low = G.eam_context(eam, [s], precision=precision)
    monkeypatch.delenv("IMDG_TILE_LIMIT")
    with low:
        low.calc_forces()
        info = low.info()
        assert info["tiles_fallback"] > 0
        assert canon_pairs_gpu(low.pairs()) == canon_pairs_oracle(orc.pairs())
        _static(low, orc, precision)
        low.run(30)
        a_low = low.atoms()
        low.make_nblist()
        n_fb = low.info()["tiles_fallback"]
    monkeypatch.setenv("IMDG_TILE_LIMIT", "2048")     # the largest cap
    full = G.eam_context(eam, [s], precision=precision)
    monkeypatch.delenv("IMDG_TILE_LIMIT")
    with full:
        full.calc_forces()
        assert full.info()["tiles_fallback"] == 0
        full.run(30)
        a_full = full.atoms()
    # a later rebuild goes through the fallback again
    assert n_fb > info["tiles_fallback"]
    for k in ("x", "p", "f", "epot"):
        assert np.array_equal(a_low[k], a_full[k]), k


@pytest.mark.parametrize("nspecies", [1, 2])
def test_list_nt_instances(nspecies, monkeypatch):
    if nspecies == 1:
        eam = eam_tables()
        s = W.eam_crystal(n_cells=7, temp_k=900.0, seed=43, displace=0.04)
    else:
        eam = eam_tables(nspecies=2, rc=5.5)
        s = W.fcc_system((7, 7, 5), 3.615, eam.masses, W.kT(700.0), 44,
                         type_fractions=(0.7, 0.3), pbc=(1, 1, 0),
                         vacuum_z=20.0, displace=0.04)
    res = {}
    for nt in ("1", "0"):
        monkeypatch.setenv("IMDG_LIST_NT", nt)
        ctx = G.eam_context(eam, [s], precision=G.MIXED,
                            ensemble=G.NVT if nspecies == 1 else G.NVE,
                            isq_tau_eta=1.0 / (100 * W.FS) ** 2)
        monkeypatch.delenv("IMDG_LIST_NT")
        with ctx:
            ctx.calc_forces()
            first = ctx.atoms()
            ctx.run(60)
            res[nt] = (first, ctx.atoms(), ctx.thermo(),
                       ctx.info()["nbl_count"])
    assert res["1"][3] == res["0"][3] > 1     # rebuilds inside the run
    for k in ("f", "epot", "rho", "df"):
        assert np.array_equal(res["1"][0][k], res["0"][0][k]), k
    for k in ("x", "p", "f", "epot"):
        assert np.array_equal(res["1"][1][k], res["0"][1][k]), k
    for k in ("epot", "ekin", "virial"):
        assert res["1"][2][k] == res["0"][2][k], k


def _alloy(seed):
    eam = eam_tables(nspecies=2, rc=5.5)
    s = W.fcc_system((7, 7, 5), 3.615, eam.masses, W.kT(700.0), seed,
                     type_fractions=(0.7, 0.3), pbc=(1, 1, 0), vacuum_z=20.0,
                     displace=0.04)
    return eam, s


@pytest.mark.parametrize("lanes", [1, 2, 4])
def test_rho_by_type_in_lds(lanes, monkeypatch):
    eam, s = _alloy(45)
    res = {}
    for name, by_type in (("lds", "1"), ("records", "0")):
        monkeypatch.setenv("IMDG_RHO_BY_TYPE", by_type)
        ctx = G.eam_context(eam, [s], precision=G.MIXED)
        monkeypatch.delenv("IMDG_RHO_BY_TYPE")
        with ctx:
            ctx.set_tuning(lanes, 512, lanes, 512)
            ctx.calc_forces()
            first = ctx.atoms()
            th0 = ctx.thermo()
            ctx.run(60)
            res[name] = (first, ctx.atoms(), th0, ctx.info())
    assert res["lds"][3]["tables_in_lds"] == 3
    assert res["records"][3]["tables_in_lds"] == 0
    assert res["lds"][3]["nbl_count"] == res["records"][3]["nbl_count"] > 1
    for k in ("f", "epot", "rho", "df"):
        assert np.array_equal(res["lds"][0][k], res["records"][0][k]), k
    for k in ("x", "p", "f", "epot"):
        assert np.array_equal(res["lds"][1][k], res["records"][1][k]), k
    # the step sums combine per-block partials (the grids may differ)
    for k in ("epot", "virial"):
        assert abs(res["lds"][2][k] - res["records"][2][k]) <= \
            1e-13 * abs(res["records"][2][k]), k


def test_rho_by_both_types_keeps_records():
    """A density column that depends on the type of the host atom too
    (column 1*nt + 0 scaled): no per-type LDS table; the mixed path still
    matches the oracle."""
    eam, s = _alloy(46)
    eam = copy.deepcopy(eam)      # the builder's tables are shared (cached)
    eam.rho.values[2] = 1.1 * np.asarray(eam.rho.values[2])
    orc = oracle_eam(s, eam, nbl_margin=0.4, timestep=W.FS)
    orc.calc_forces()
    with G.eam_context(eam, [s], precision=G.MIXED) as gpu:
        gpu.calc_forces()
        assert gpu.info()["tables_in_lds"] == 0
        _static(gpu, orc, G.MIXED)


def test_three_species_split_placement(monkeypatch):
    """Three species on one grid: the force loop's per-type density
    segments (3 x 5,000 x 8 B) fit in LDS, the density loop's (3 x 5,000 x
    12 B) do not and it reads the records; parity with the oracle, and
    bit-identical to the records everywhere (IMDG_RHO_BY_TYPE=0)."""
    third = W.JohnsonParams(r_e=2.60, f_e=0.95, phi_e=0.57, E_c=3.9,
                            alpha=5.0, beta=5.9, gamma=8.2, mass=60.0,
                            a0=3.6)
    eam = W.johnson_eam((W.CU, W.NI, third), rc=5.5)
    s = W.fcc_system((6, 6, 5), 3.6, eam.masses, W.kT(600.0), 47,
                     type_fractions=(0.5, 0.3, 0.2), displace=0.04)
    orc = oracle_eam(s, eam, nbl_margin=0.4, timestep=W.FS)
    orc.calc_forces()
    res = {}
    for name, by_type in (("lds", "1"), ("records", "0")):
        monkeypatch.setenv("IMDG_RHO_BY_TYPE", by_type)
        ctx = G.eam_context(eam, [s], precision=G.MIXED)
        monkeypatch.delenv("IMDG_RHO_BY_TYPE")
        with ctx:
            ctx.calc_forces()
            if name == "lds":
                _static(ctx, orc, G.MIXED)
            first = ctx.atoms()
            ctx.run(30)
            res[name] = (first, ctx.atoms(), ctx.info())
    assert res["lds"][2]["tables_in_lds"] == 2
    assert res["records"][2]["tables_in_lds"] == 0
    for k in ("f", "epot", "rho", "df"):
        assert np.array_equal(res["lds"][0][k], res["records"][0][k]), k
    for k in ("x", "p", "f", "epot"):
        assert np.array_equal(res["lds"][1][k], res["records"][1][k]), k


def _bring_close(s, i, dist):
    """Move the nearest neighbour of atom i (minimum image) to distance dist
    along the line that joins them."""
    L = np.diag(s.box)
    d = s.x - s.x[i]
    d -= L * np.round(d / L)
    r = np.sqrt(np.sum(d * d, axis=1))
    r[i] = np.inf
    j = int(np.argmin(r))
    x = s.x.copy()
    x[j] = s.x[i] + d[j] * (dist / r[j])
    x[j] -= L * np.floor(x[j] / L)
    s.x = x
    return j


@pytest.mark.parametrize("precision", [G.FP64, G.MIXED])
def test_short_distance_eam(precision):
    eam = eam_tables()
    s = W.eam_crystal(n_cells=5, temp_k=300.0, seed=43, displace=0.02)
    _bring_close(s, 5, 0.95)          # r^2 = 0.9025 below begin 1.0
    orc = oracle_eam(s, eam, nbl_margin=0.4, timestep=W.FS)
    orc.calc_forces()
    assert orc.scalar("is_short") == 1
    with G.eam_context(eam, [s], precision=precision) as gpu:
        gpu.calc_forces()
        assert canon_pairs_gpu(gpu.pairs()) == canon_pairs_oracle(orc.pairs())
        _static(gpu, orc, precision)
        assert gpu.info()["is_short"] == 1
        assert gpu.info()["is_short"] == 0      # reported once
        gpu.calc_forces()
        assert gpu.info()["is_short"] == 1


@pytest.mark.parametrize("precision", [G.FP64, G.MIXED])
def test_short_distance_pair_only(precision):
    s = W.fcc_system((6, 6, 6), 1.5496, [1.0], 0.5, 44, displace=0.02)
    _bring_close(s, 10, 0.45)         # r^2 = 0.2025 below begin 0.25
    orc = oracle_lj(s, nbl_margin=0.3, timestep=0.002)
    orc.calc_forces()
    assert orc.scalar("is_short") == 1
    with G.pair_context(lj(), [s], precision=precision) as gpu:
        gpu.calc_forces()
        _static(gpu, orc, precision, has_eam=False)
        assert gpu.info()["is_short"] == 1


def test_no_short_distance_flag_in_a_crystal():
    eam = eam_tables()
    s = W.eam_crystal(n_cells=5, temp_k=300.0, seed=45, displace=0.02)
    with G.eam_context(eam, [s], precision=G.MIXED) as gpu:
        gpu.run(5)
        assert gpu.info()["is_short"] == 0
