"""Parity at the sizes the path is measured on, against the oracle on the
same inputs (threaded for speed: the same list, sums regrouped over threads,
which moves results by ~1e-15 relative, far below every bar):

* 4,000-atom LJ crystal: 1,000 NVE steps, FP64 and mixed;
* 256,000-atom EAM crystal (40^3 fcc cells): static parity and 20 fused NVE
  steps, FP64 and mixed;
* 2,048,000-atom EAM crystal (80^3 cells), NVT 800 K: static parity and 5
  fused NVT steps, FP64 and mixed;
* 64 replicas of 500-50,000 atoms: per-replica static parity, then 20 NVE
  steps of the captured (hipGraph) step with autotuned launch shapes;
* 8,000,000-atom two-species slab (200x200x50 cells, open z), mixed, NVT
  500 K: cells bit-exact, the pair set by count and an order-independent
  digest, per-atom quantities at the mixed bars, Newton's third law, and the
  mixed forces against the FP64 GPU path.

Pair sets: sorted 64-bit keys (id_lo, id_hi, image shift) compared for
equality; the full list must hold every pair from both atoms (its entry
count is twice the pair count).  Bars as in test_gpu_parity.py (IMD
semantics: src/imd_forces_nbl.c:136-273, 281-1323).
"""
import os

import numpy as np
import pytest

from imd_amd import imdg as G
from imd_amd import workloads as W

from helpers import eam_tables, lj, max_abs, oracle_eam, oracle_lj, rel_rms

pytestmark = pytest.mark.gpu

THREADS = max(1, min(16, os.cpu_count() or 1))
BARS = {G.FP64: (1e-10, 1e-10), G.MIXED: (1e-5, 1e-6)}


def pair_keys(p5, n, sort=True):
    """(id_lo, id_hi, sx, sy, sz) rows -> int64 keys ((lo n + hi) 27 +
    shift code), sorted unless sort is False (in-place arithmetic: the
    8M-atom slab has 3e8 pairs)."""
    k = p5[:, 0] * n
    k += p5[:, 1]
    for c in (2, 3, 4):
        k *= 3
        k += p5[:, c]
        k += 1
    if sort:
        k.sort()
    return k


def digest(keys):
    """Order-independent digest of a key set: count, wrapping sum and xor of
    a 64-bit mix of every key."""
    x = keys.view(np.uint64)
    x ^= x >> np.uint64(31)
    x *= np.uint64(0xBF58476D1CE4E5B9)
    x ^= x >> np.uint64(29)
    return (len(x), int(np.sum(x, dtype=np.uint64)),
            int(np.bitwise_xor.reduce(x)))


def assert_same_pairs(gpu, orc, n, rep=0, exact=True):
    gp = gpu.pairs()
    if gpu.nrep > 1:
        gp = gp[gp[:, 0] == rep]
    kg = pair_keys(gp[:, 1:], n, exact)
    del gp
    op = orc.pairs()
    ko = pair_keys(op, n, exact)
    del op
    assert len(kg) == len(ko)
    if exact:
        assert np.array_equal(kg, ko)
    else:
        assert digest(kg) == digest(ko)
    return len(ko)


def assert_static(gpu, orc, precision, rep=0, sl=None, has_eam=True):
    tol_f, tol_s = BARS[precision]
    ga, oa = gpu.atoms(), orc.atoms()
    if sl is not None:
        ga = {k: v[sl] for k, v in ga.items()}
    assert rel_rms(ga["f"], oa["f"]) <= tol_f, rel_rms(ga["f"], oa["f"])
    assert rel_rms(ga["epot"], oa["epot"]) <= tol_s
    if has_eam:
        assert rel_rms(ga["rho"], oa["rho"]) <= tol_s
        assert rel_rms(ga["df"], oa["df"]) <= tol_s
    th = gpu.thermo(rep)
    e_o = orc.scalar("epot")
    assert abs(th["epot"] - e_o) <= tol_s * abs(e_o)
    assert abs(th["virial"] - orc.scalar("virial")) <= \
        tol_s * orc.scalar("virial_abs")


def assert_symmetric_list(gpu, npairs_total):
    info = gpu.info()
    assert round(info["mean_k"] * info["natoms"]) == 2 * npairs_total


def _oracle(s, eam, **kw):
    orc = oracle_eam(s, eam, nbl_margin=0.4, timestep=W.FS, **kw)
    orc.set_threads(THREADS)
    return orc


def _traj_bars(precision, hist, th, ga, oa, box, p_bar=1e-5):
    """FP64: the oracle's trajectory to rounding.  Mixed: energies to 1e-6
    and momenta to p_bar (the FP32 force error, ~1e-6 relative, separates
    the trajectories slowly: 1e-5 over tens of steps, more over 1,000)."""
    if precision == G.FP64:
        d = ga["x"] - oa["x"]
        L = np.diag(box)
        d -= L * np.round(d / L)
        assert np.max(np.abs(d)) < 1e-9
        assert max_abs(ga["p"], oa["p"]) < 1e-9
        assert np.allclose(hist[:, 1], th[:, 0], rtol=1e-10, atol=0)
        assert np.allclose(hist[:, 2], th[:, 1], rtol=1e-10, atol=0)
    else:
        assert np.allclose(hist[:, 1], th[:, 0], rtol=1e-6, atol=0)
        assert np.allclose(hist[:, 2], th[:, 1], rtol=1e-6, atol=0)
        assert rel_rms(ga["p"], oa["p"]) <= p_bar
    assert np.array_equal(hist[:, 6] > 0, th[:, 4] > 0)


@pytest.mark.parametrize("precision", [G.FP64, G.MIXED])
def test_lj_4000_atoms_1000_steps(precision):
    s = W.lj_crystal()
    orc = oracle_lj(s, nbl_margin=0.3, timestep=0.002)
    th = orc.run(1000)
    with G.pair_context(lj(), [s], precision=precision) as gpu:
        gpu.run(1000)
        hist = np.concatenate([gpu.thermo_history(k, 250)[:, 0, :]
                               for k in range(0, 1000, 250)])
        ga = gpu.atoms()
    _traj_bars(precision, hist, th, ga, orc.atoms(), s.box, p_bar=1e-4)


@pytest.mark.parametrize("precision", [G.FP64, G.MIXED])
def test_eam_256k_static_and_20_steps(precision):
    eam = eam_tables()
    s = W.eam_crystal(n_cells=40, temp_k=600.0, seed=1001)
    orc = _oracle(s, eam)
    orc.calc_forces()
    with G.eam_context(eam, [s], precision=precision) as gpu:
        gpu.calc_forces()
        assert gpu.grid() == orc.grid() == (26, 26, 26)
        assert np.array_equal(gpu.cells(), orc.cells())
        npairs = assert_same_pairs(gpu, orc, s.n)
        assert_symmetric_list(gpu, npairs)
        assert_static(gpu, orc, precision)
    orc = _oracle(s, eam)
    th = orc.run(20)
    with G.eam_context(eam, [s], precision=precision) as gpu:
        gpu.run(20)
        hist = gpu.thermo_history(0, 20)[:, 0, :]
        _traj_bars(precision, hist, th, gpu.atoms(), orc.atoms(), s.box)


@pytest.mark.parametrize("precision", [G.FP64, G.MIXED])
def test_eam_2M_nvt_static_and_5_steps(precision):
    eam = eam_tables()
    T, isq = W.kT(800.0), 1.0 / (100 * W.FS) ** 2
    s = W.eam_crystal(n_cells=80, temp_k=800.0, seed=1002)
    nvt = dict(ensemble=1, temperature=T, eta=0.0, isq_tau_eta=isq)
    orc = _oracle(s, eam, **nvt)
    orc.calc_forces()
    with G.eam_context(eam, [s], precision=precision, ensemble=G.NVT,
                       isq_tau_eta=isq) as gpu:
        gpu.calc_forces()
        assert gpu.grid() == orc.grid() == (53, 53, 53)
        assert np.array_equal(gpu.cells(), orc.cells())
        npairs = assert_same_pairs(gpu, orc, s.n)
        assert_symmetric_list(gpu, npairs)
        assert_static(gpu, orc, precision)
        info = gpu.info()
    del orc
    orc = _oracle(s, eam, **nvt)
    th = orc.run(5)
    with G.eam_context(eam, [s], precision=precision, ensemble=G.NVT,
                       isq_tau_eta=isq) as gpu:
        gpu.run(5)
        hist = gpu.thermo_history(0, 5)[:, 0, :]
        _traj_bars(precision, hist, th, gpu.atoms(), orc.atoms(), s.box)
        assert np.allclose(hist[:, 4], th[:, 3], rtol=1e-9 if precision ==
                           G.FP64 else 1e-5, atol=1e-12)
    print("2M: tiles_fallback %d exact_pairs %d" % (info["tiles_fallback"],
                                                   info["exact_pairs"]))


@pytest.mark.parametrize("precision", [G.FP64, G.MIXED])
def test_64_replicas_static_and_20_graph_steps(precision, monkeypatch):
    eam = eam_tables()
    systems = W.replica_set(64, seed=1003)
    monkeypatch.setenv("IMDG_GRAPH", "1")
    gpu = G.eam_context(eam, systems, precision=precision)
    monkeypatch.delenv("IMDG_GRAPH")
    offs = np.cumsum([0] + [s.n for s in systems])
    with gpu:
        gpu.calc_forces()
        ga_all = gpu.atoms()
        gp = gpu.pairs()
        total = 0
        for r, s in enumerate(systems):
            orc = _oracle(s, eam)
            orc.calc_forces()
            sl = slice(offs[r], offs[r + 1])
            assert gpu.grid(r) == orc.grid()
            assert np.array_equal(gpu.cells()[sl], orc.cells())
            mine = gp[gp[:, 0] == r][:, 1:]
            assert np.array_equal(pair_keys(mine, s.n),
                                  pair_keys(orc.pairs(), s.n))
            total += len(mine)
            tol_f, tol_s = BARS[precision]
            oa = orc.atoms()
            # the replicas start on the perfect lattice (vacancies aside),
            # where the net forces cancel to ~1e-14 eV/A: forces are
            # compared on the scale of a thermal force (0.1 eV/A) there
            fscale = max(float(np.sqrt(np.mean(oa["f"] ** 2))), 0.1)
            assert np.sqrt(np.mean((ga_all["f"][sl] - oa["f"]) ** 2)) <= \
                tol_f * fscale
            assert rel_rms(ga_all["rho"][sl], oa["rho"]) <= tol_s
            assert rel_rms(ga_all["epot"][sl], oa["epot"]) <= tol_s
            th = gpu.thermo(r)
            assert abs(th["epot"] - orc.scalar("epot")) <= \
                tol_s * abs(orc.scalar("epot"))
            assert abs(th["virial"] - orc.scalar("virial")) <= \
                tol_s * orc.scalar("virial_abs")
        assert_symmetric_list(gpu, total)
        gpu.autotune(2)
        s0 = gpu.info()["steps_done"]
        gpu.run(20)
        hist = gpu.thermo_history(s0, 20)
        ga_all = gpu.atoms()
    for r, s in enumerate(systems):
        orc = _oracle(s, eam)
        th = orc.run(20)
        sl = slice(offs[r], offs[r + 1])
        ga = {k: v[sl] for k, v in ga_all.items()}
        _traj_bars(precision, hist[:, r, :], th, ga, orc.atoms(), s.box)


@pytest.mark.timeout(300)
def test_alloy_slab_8M_mixed():
    eam = eam_tables(nspecies=2, rc=5.5)
    T, isq = W.kT(500.0), 1.0 / (100 * W.FS) ** 2
    s = W.alloy_slab()
    assert s.n == 8_000_000
    orc = _oracle(s, eam, ensemble=1, temperature=T, eta=0.0, isq_tau_eta=isq)
    orc.calc_forces()
    with G.eam_context(eam, [s], precision=G.MIXED, ensemble=G.NVT,
                       isq_tau_eta=isq) as gpu:
        gpu.calc_forces()
        assert gpu.grid() == orc.grid() == (122, 122, 34)
        assert np.array_equal(gpu.cells(), orc.cells())
        npairs = assert_same_pairs(gpu, orc, s.n, exact=False)
        assert_symmetric_list(gpu, npairs)
        assert_static(gpu, orc, G.MIXED)
        fm = gpu.atoms()["f"]
    del orc
    # Newton's third law: the forces of the whole slab sum to zero, up to
    # the FP32 rounding of each atom's sum (a random walk over the atoms)
    assert np.max(np.abs(fm.sum(axis=0))) <= 1e-5 * np.sqrt(np.sum(fm ** 2))
    with G.eam_context(eam, [s], precision=G.FP64, ensemble=G.NVT,
                       isq_tau_eta=isq) as gpu:
        gpu.calc_forces()
        f64 = gpu.atoms()["f"]
    assert rel_rms(fm, f64) <= 1e-5
