"""Pins for the oracle's move_atoms_nve / move_atoms_nvt / check_nblist.

Free particles (no partner inside the cutoff) have the closed-form
trajectory x_n = x_0 + n dt p / m, the skin test then fires at a step that
is known in advance, NVT with eta = 0 and 1/tau^2 = 0 is NVE, the
Nose-Hoover thermostat drives the temperature to its target, and leapfrog's
energy error shrinks as dt^2.
"""
import numpy as np
import pytest

import oracle as O
from imd_amd import workloads as W


def short_table():
    # a pair table that is zero beyond r = 1.5 (r^2 in [1, 2.25])
    step, xs = W._grid(1.0, 2.25, 100)
    return W.Table(np.array([1.0]), np.array([2.25]), np.array([step]),
                   [(2.25 - xs) ** 2], True)


def free_system(n=27, L=30.0, seed=3, vmax=0.05):
    g = int(round(n ** (1 / 3)))
    pts = (np.stack(np.meshgrid(*[np.arange(g)] * 3, indexing="ij"), -1)
           .reshape(-1, 3) + 0.5) * (L / g)
    rng = np.random.default_rng(seed)
    mass = rng.uniform(1.0, 3.0, len(pts))
    v = rng.uniform(-vmax, vmax, (len(pts), 3))
    return W.System(x=pts, p=v * mass[:, None], type=np.zeros(len(pts), np.int32),
                    mass=mass, id=np.arange(len(pts)), box=np.eye(3) * L)


def oracle_for(s, ensemble=O.ORC_NVE, dt=0.5, margin=0.4, **kw):
    pair = O.OracleTable.from_table(short_table())
    return O.OracleSystem(s, pair, None, None, nbl_margin=margin, timestep=dt,
                          ensemble=ensemble, ntypes=1, **kw), pair


def test_free_particles_closed_form():
    s = free_system()
    orc, _ = oracle_for(s, dt=0.5)
    n = 40
    th = orc.run(n)
    a = orc.atoms()
    x_exact = s.x + n * 0.5 * s.p / s.mass[:, None]
    L = 30.0
    d = a["x"] - x_exact
    d -= L * np.round(d / L)
    assert np.max(np.abs(d)) < 1e-11
    assert np.allclose(a["p"], s.p, rtol=0, atol=0)
    ek = 0.5 * np.sum(s.p ** 2 / s.mass[:, None])
    assert np.allclose(th[:, 1], ek, rtol=1e-14)
    assert np.all(th[:, 0] == 0.0)


def test_skin_test_fires_at_predicted_step():
    s = free_system()
    dt, margin = 0.5, 0.4
    orc, _ = oracle_for(s, dt=dt, margin=margin)
    th = orc.run(60)
    speed = np.sqrt(np.sum((s.p / s.mass[:, None]) ** 2, axis=1)).max()
    # displacement after k steps is k*dt*speed; the list is rebuilt at the
    # step after the first k with displacement > margin/2
    k1 = int(np.floor((0.5 * margin) / (dt * speed))) + 1
    rebuilt = np.nonzero(th[:, 4])[0]
    assert rebuilt[0] == 0           # first call builds the list
    assert rebuilt[1] == k1          # check after step k1-1 -> rebuild at k1


def test_nvt_with_zero_friction_is_nve():
    eam = W.johnson_eam()
    s = W.eam_crystal(n_cells=4, temp_k=500.0, seed=8, displace=0.02)
    tabs = [O.OracleTable.from_table(eam.pair), O.OracleTable.from_table(eam.rho),
            O.OracleTable.from_table(eam.embed, radial=False)]
    a = O.OracleSystem(s, *tabs, nbl_margin=0.4, timestep=W.FS)
    b = O.OracleSystem(s, *tabs, nbl_margin=0.4, timestep=W.FS,
                       ensemble=O.ORC_NVT, temperature=W.kT(500.0), eta=0.0,
                       isq_tau_eta=0.0)
    ta, tb = a.run(30), b.run(30)
    assert np.array_equal(a.atoms()["x"], b.atoms()["x"])
    assert np.array_equal(a.atoms()["p"], b.atoms()["p"])
    assert np.allclose(ta[:, 1], tb[:, 1], rtol=1e-14)
    assert np.all(tb[:, 3] == 0.0)


def test_nose_hoover_reaches_target_temperature():
    eam = W.johnson_eam()
    s = W.eam_crystal(n_cells=4, temp_k=100.0, seed=9, displace=0.0)
    target = W.kT(600.0)
    tabs = [O.OracleTable.from_table(eam.pair), O.OracleTable.from_table(eam.rho),
            O.OracleTable.from_table(eam.embed, radial=False)]
    orc = O.OracleSystem(s, *tabs, nbl_margin=0.4, timestep=2 * W.FS,
                         ensemble=O.ORC_NVT, temperature=target, eta=0.0,
                         isq_tau_eta=1.0 / (40 * W.FS) ** 2)
    th = orc.run(1500)
    T = 2 * th[:, 1] / (3 * s.n)
    assert abs(T[-500:].mean() / target - 1.0) < 0.1
    assert th[:50, 3].max() < 0          # eta < 0 heats while T < target


def test_leapfrog_energy_error_scales_as_dt_squared():
    eam = W.johnson_eam()
    s = W.eam_crystal(n_cells=4, temp_k=400.0, seed=10, displace=0.02)
    tabs = [O.OracleTable.from_table(eam.pair), O.OracleTable.from_table(eam.rho),
            O.OracleTable.from_table(eam.embed, radial=False)]
    spread = []
    for dt, n in ((2 * W.FS, 200), (1 * W.FS, 400)):
        orc = O.OracleSystem(s, *tabs, nbl_margin=0.4, timestep=dt)
        th = orc.run(n)
        e = th[:, 0] + th[:, 1]
        spread.append(e.std())
    ratio = spread[0] / spread[1]
    assert 2.5 < ratio < 6.0
