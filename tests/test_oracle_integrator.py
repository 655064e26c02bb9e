"""Pins for the oracle's move_atoms_nve / move_atoms_nvt / check_nblist.

Free particles (no partner inside the cutoff) have the closed-form
trajectory x_n = x_0 + n dt p / m, the skin test then fires at a step that
is known in advance, NVT with eta = 0 and 1/tau^2 = 0 is NVE, the
Nose-Hoover thermostat drives the temperature to its target, leapfrog's
energy error shrinks as dt^2, and for free particles with eta != 0 the
Nose-Hoover recurrence (friction form, Ekin, eta update) is followed step by
step in numpy, tight enough that an explicit friction step, a missing
1/(1 + eta dt/2) or an eta update from E1 fail.
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


def nose_hoover_free(s, dt, T, eta0, isq, n, friction="imd", eta_from="e2"):
    """Nose-Hoover for free particles (F = 0) in numpy, step by step:
    reib = 1 - eta dt/2, einsd = 1/(1 + eta dt/2), p <- (p reib + dt F)
    einsd, x <- x + (dt/m) p, Ekin = (E1 + E2)/4 and eta <- eta + dt (E2 /
    (3N T) - 1) / tau^2, E1 and E2 = sum |p|^2/m before and after the
    momentum update (src/imd_integrate.c:907-908, 1020-1027, 1103,
    1140-1141).  friction="euler" and eta_from="e1" are the plausible
    mistakes the pin must reject."""
    m = s.mass[:, None]
    p = s.p.astype(np.float64).copy()
    x = s.x.astype(np.float64).copy()
    eta = eta0
    ek, etas = [], []
    for _ in range(n):
        reib = 1.0 - eta * dt / 2.0
        einsd = 1.0 / (1.0 + eta * dt / 2.0)
        e1 = np.sum(p * p / m)
        if friction == "imd":
            p = (p * reib + dt * 0.0) * einsd
        elif friction == "euler":
            p = p * (1.0 - eta * dt) + dt * 0.0
        else:                       # the division by 1 + eta dt/2 left out
            p = p * reib + dt * 0.0
        e2 = np.sum(p * p / m)
        x = x + (dt / m) * p
        ek.append((e1 + e2) / 4.0)
        eta += dt * ((e2 if eta_from == "e2" else e1) / (3 * s.n * T) - 1.0) * isq
        etas.append(eta)
    return x, p, np.array(ek), np.array(etas)


NH_FREE = dict(dt=0.5, eta0=0.3, tau=6.0)


def nh_free_case():
    s = free_system(n=64, vmax=0.05, seed=21)
    # target between the initial temperature and zero: the friction cools,
    # eta relaxes through zero and changes sign within the run
    T = 0.5 * np.sum(s.p ** 2 / s.mass[:, None]) / (3 * s.n)
    return s, T


def test_nose_hoover_free_particles_closed_form():
    s, T = nh_free_case()
    dt, eta0 = NH_FREE["dt"], NH_FREE["eta0"]
    isq = 1.0 / NH_FREE["tau"] ** 2
    n = 50
    orc, _ = oracle_for(s, ensemble=O.ORC_NVT, dt=dt, temperature=T,
                        eta=eta0, isq_tau_eta=isq)
    th = orc.run(n)
    a = orc.atoms()
    x, p, ek, etas = nose_hoover_free(s, dt, T, eta0, isq, n)
    # the recurrence moves eta and Ekin by O(1) relative amounts
    assert etas.min() < 0.0 < etas.max()
    assert ek[-1] < 0.6 * ek[0]
    assert np.allclose(th[:, 3], etas, rtol=1e-14, atol=1e-15)
    assert np.allclose(th[:, 1], ek, rtol=1e-14, atol=0)
    assert np.allclose(a["p"], p, rtol=1e-13, atol=1e-18)
    L = 30.0
    d = a["x"] - x
    d -= L * np.round(d / L)
    assert np.max(np.abs(d)) < 1e-12
    assert np.all(th[:, 0] == 0.0)
    # each plausible mistake misses the bar by orders of magnitude
    for kw in (dict(friction="euler"), dict(friction="no_einsd"),
               dict(eta_from="e1")):
        _, pw, ekw, etaw = nose_hoover_free(s, dt, T, eta0, isq, n, **kw)
        assert np.max(np.abs(etaw - th[:, 3])) > 1e-6
        assert np.max(np.abs(ekw - th[:, 1]) / ek[0]) > 1e-6
