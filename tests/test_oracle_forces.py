"""Pins for the oracle's pair + two-pass EAM force loop.

Tables sampled from quadratics (phi and rho quadratic in r^2, F quadratic in
rho) are reproduced exactly by IMD's 3-point interpolation, so energies,
densities, 2F', forces and the virial have closed forms that an independent
brute-force numpy computation evaluates.  With two species and asymmetric
density columns this also pins the column convention (an it-atom takes its
density from column it*nt+jt; the force uses rho'_i from jt*nt+it and rho'_j
from it*nt+jt).  Finite differences pin the smooth Johnson tables; Newton-3,
the perfect lattice and the volume derivative of the energy pin the rest.
"""
import numpy as np
import pytest

import oracle as O
from imd_amd import workloads as W

RC = 4.0


def poly_tables(nt, A, B, C, D, npts=400, rho_max=None):
    rc2 = RC * RC
    step, xs = W._grid(1.0, rc2, npts)
    pair_cols, rho_cols = [], []
    for it in range(nt):
        for jt in range(nt):
            a = 0.5 * (A[it] + A[jt])
            pair_cols.append(a * (xs - rc2) ** 2)
            rho_cols.append(B[jt] * (rc2 - xs) ** 2)
    nc = nt * nt
    pair = W.Table(np.full(nc, 1.0), np.full(nc, rc2), np.full(nc, step),
                   pair_cols, True)
    rho = W.Table(np.full(nc, 1.0), np.full(nc, rc2), np.full(nc, step),
                  rho_cols, True)
    rm = rho_max
    fstep, fx = W._grid(0.0, rm, npts)
    emb = W.Table(np.zeros(nt), np.full(nt, rm), np.full(nt, fstep),
                  [C[t] * fx * fx + D[t] * fx for t in range(nt)], False)
    return pair, rho, emb


def analytic(x, types, box, A, B, C, D, G=None, H=None):
    """Closed forms of the polynomial EAM; with G, H also EEAM's term
    E(P) = G P^2 + H P of P_i = sum_j rho_ij^2."""
    n = x.shape[0]
    rc2 = RC * RC
    L = np.diag(box)
    d = x[None, :, :] - x[:, None, :]          # x_j - x_i
    d -= L * np.round(d / L)
    r2 = np.einsum("ijk,ijk->ij", d, d)
    m = (r2 < rc2) & ~np.eye(n, dtype=bool)
    ti, tj = types[:, None], types[None, :]
    a = 0.5 * (np.asarray(A)[ti] + np.asarray(A)[tj])
    phi = np.where(m, a * (r2 - rc2) ** 2, 0.0)
    dphi = np.where(m, 2 * a * (r2 - rc2), 0.0)            # d phi / d r2
    Bj = np.asarray(B)[tj]
    Bi = np.asarray(B)[ti]
    f_j = np.where(m, Bj * (rc2 - r2) ** 2, 0.0)           # density at i from j
    df_j = np.where(m, -2 * Bj * (rc2 - r2), 0.0)
    df_i = np.where(m, -2 * Bi * (rc2 - r2), 0.0)          # density at j from i
    rho = f_j.sum(1)
    Ct, Dt = np.asarray(C)[types], np.asarray(D)[types]
    F = Ct * rho ** 2 + Dt * rho
    Fp = 2 * Ct * rho + Dt
    g = 2 * dphi + 2 * (Fp[:, None] * df_j + Fp[None, :] * df_i)
    P = (f_j ** 2).sum(1)
    Ep = np.zeros_like(P)
    E = np.zeros_like(P)
    if G is not None:
        f_i = np.where(m, Bi * (rc2 - r2) ** 2, 0.0)       # density at j from i
        Gt, Ht = np.asarray(G)[types], np.asarray(H)[types]
        E = Gt * P ** 2 + Ht * P
        Ep = 2 * Gt * P + Ht
        g = g + 4 * (Ep[:, None] * f_j * df_j + Ep[None, :] * f_i * df_i)
    f = np.einsum("ij,ijk->ik", g, d)
    epot_i = 0.5 * phi.sum(1) + F + E
    epot = epot_i.sum()
    virial = -0.5 * np.sum(r2 * g)
    # per-atom stress tensor: -1/2 sum_j g d_a d_b, xx yy zz yz zx xy
    ab = [(0, 0), (1, 1), (2, 2), (1, 2), (2, 0), (0, 1)]
    press = np.stack([-0.5 * np.einsum("ij,ij,ij->i", g, d[:, :, a], d[:, :, b])
                      for a, b in ab], axis=1)
    return dict(f=f, rho=rho, df=2 * Fp, epot_i=epot_i, epot=epot,
                virial=virial, press=press, P=P, dm=2 * Ep)


def run_oracle(s, tabs, nt, press=False):
    pair, rho, emb = [O.OracleTable.from_table(t) for t in tabs]
    orc = O.OracleSystem(s, pair, rho, emb, nbl_margin=0.3, timestep=0.001,
                         ntypes=nt)
    orc.set_press(press)
    orc.calc_forces()
    return orc, (pair, rho, emb)


@pytest.mark.parametrize("nt", [1, 2])
def test_polynomial_eam_closed_form(nt):
    rng = np.random.default_rng(5 + nt)
    box = np.eye(3) * 13.0
    s = W.fcc_system((4, 4, 4), 3.25, [1.0] * nt, 0.0, 9, displace=0.15,
                     type_fractions=None if nt == 1 else (0.6, 0.4))
    A = [0.05, 0.08][:nt]
    B = [0.02, 0.035][:nt]          # asymmetric density columns
    C = [0.3, 0.2][:nt]
    D = [-1.0, -0.7][:nt]
    ref = analytic(s.x, s.type, s.box, A, B, C, D)
    tabs = poly_tables(nt, A, B, C, D, rho_max=1.3 * ref["rho"].max())
    orc, _ = run_oracle(s, tabs, nt)
    a = orc.atoms()
    scale = np.abs(ref["f"]).max()
    assert np.max(np.abs(a["f"] - ref["f"])) < 1e-10 * scale
    assert np.allclose(a["rho"], ref["rho"], rtol=1e-11, atol=0)
    assert np.allclose(a["df"], ref["df"], rtol=1e-10, atol=1e-13)
    assert np.allclose(a["epot"], ref["epot_i"], rtol=1e-10, atol=1e-12)
    assert abs(orc.scalar("epot") - ref["epot"]) < 1e-10 * abs(ref["epot"])
    assert abs(orc.scalar("virial") - ref["virial"]) < 1e-9 * abs(ref["virial"])


def test_swapped_density_columns_would_fail():
    """The two-species closed form distinguishes rho_{it jt} from rho_{jt it}:
    swapping the density prefactors between species changes the forces."""
    s = W.fcc_system((4, 4, 4), 3.25, [1.0, 1.0], 0.0, 9, displace=0.15,
                     type_fractions=(0.6, 0.4))
    A, B, C, D = [0.05, 0.08], [0.02, 0.035], [0.3, 0.2], [-1.0, -0.7]
    ref = analytic(s.x, s.type, s.box, A, B, C, D)
    swapped = analytic(s.x, s.type, s.box, A, B[::-1], C, D)
    assert np.max(np.abs(ref["f"] - swapped["f"])) > 1e-3 * np.abs(ref["f"]).max()


def _energy(s, tabs, nt=1):
    orc, keep = run_oracle(s, tabs, nt)
    return orc.scalar("epot"), orc


def test_johnson_forces_are_energy_gradients():
    eam = W.johnson_eam()
    tabs = (eam.pair, eam.rho, eam.embed)
    s = W.eam_crystal(n_cells=4, temp_k=300.0, seed=4, displace=0.08)
    e0, orc = _energy(s, tabs)
    f = orc.atoms()["f"]
    h = 1e-5
    rng = np.random.default_rng(1)
    for atom in rng.choice(s.n, 6, replace=False):
        for k in range(3):
            sp = W.System(**{**s.__dict__})
            sp.x = s.x.copy()
            sp.x[atom, k] += h
            sm = W.System(**{**s.__dict__})
            sm.x = s.x.copy()
            sm.x[atom, k] -= h
            ep, _ = _energy(sp, tabs)
            em, _ = _energy(sm, tabs)
            fd = -(ep - em) / (2 * h)
            assert abs(fd - f[atom, k]) < 2e-5 + 1e-4 * abs(f[atom, k])


def test_newton3_and_perfect_lattice():
    eam = W.johnson_eam()
    tabs = (eam.pair, eam.rho, eam.embed)
    s = W.eam_crystal(n_cells=5, temp_k=0.0, seed=2, displace=0.05)
    _, orc = _energy(s, tabs)
    f = orc.atoms()["f"]
    assert np.max(np.abs(f.sum(0))) < 1e-11 * np.abs(f).max() * s.n
    lat = W.eam_crystal(n_cells=5, temp_k=0.0, seed=2, displace=0.0)
    _, orc2 = _energy(lat, tabs)
    a = orc2.atoms()
    assert np.max(np.abs(a["f"])) < 1e-11
    # every site has the same density: the shell sum through the table
    assert np.ptp(a["rho"]) < 1e-11 * a["rho"].mean()
    rho_closed = W.fcc_shell_density(lambda r: np.interp(r * r, *_table_xy(eam.rho)), 3.615, 5.0)
    assert abs(a["rho"].mean() - rho_closed) < 1e-3 * rho_closed


def _table_xy(t):
    xs = t.begin[0] + t.step[0] * np.arange(len(t.values[0]))
    return xs, t.values[0]


def test_virial_is_volume_derivative():
    eam = W.johnson_eam()
    tabs = (eam.pair, eam.rho, eam.embed)
    s = W.eam_crystal(n_cells=4, temp_k=0.0, seed=6, displace=0.05)
    _, orc = _energy(s, tabs)
    vir = orc.scalar("virial")
    h = 1e-6
    es = []
    for lam in (1 + h, 1 - h):
        t = W.System(**{**s.__dict__})
        t.x = s.x * lam
        t.box = s.box * lam
        es.append(_energy(t, tabs)[0])
    dEdl = (es[0] - es[1]) / (2 * h)
    assert abs(-dEdl - vir) < 1e-5 * abs(vir)


# ---------------------------------------------------------------------------
# per-atom stress tensor (STRESS_TENS)
# ---------------------------------------------------------------------------

@pytest.mark.parametrize("nt", [1, 2])
def test_stress_closed_form(nt):
    """PRESSTENS after calc_forces: -1/2 sum_j g_ij d_a d_b over all
    neighbours (src/imd_forces_nbl.c:558-581, 1283-1318, buffer copies added
    back), against the brute-force closed form; its trace summed over the
    atoms is the virial."""
    s = W.fcc_system((4, 4, 4), 3.25, [1.0] * nt, 0.0, 9, displace=0.15,
                     type_fractions=None if nt == 1 else (0.6, 0.4))
    A = [0.05, 0.08][:nt]
    B = [0.02, 0.035][:nt]
    C = [0.3, 0.2][:nt]
    D = [-1.0, -0.7][:nt]
    ref = analytic(s.x, s.type, s.box, A, B, C, D)
    tabs = poly_tables(nt, A, B, C, D, rho_max=1.3 * ref["rho"].max())
    orc, _ = run_oracle(s, tabs, nt, press=True)
    P = orc.press()
    scale = np.abs(ref["press"]).max()
    assert np.max(np.abs(P - ref["press"])) < 1e-10 * scale
    # off-diagonal terms are not small here (displaced lattice)
    assert np.abs(ref["press"][:, 3:]).max() > 1e-2 * scale
    tr = P[:, :3].sum()
    assert abs(tr - orc.scalar("virial")) < 1e-10 * abs(orc.scalar("virial"))


def test_stress_is_strain_derivative():
    """sum_i S_aa = -dE/d eps_aa for the stretch x_a -> (1 + eps) x_a of
    positions and box (Johnson EAM, finite differences)."""
    eam = W.johnson_eam()
    tabs = (eam.pair, eam.rho, eam.embed)
    s = W.eam_crystal(n_cells=4, temp_k=0.0, seed=6, displace=0.05)
    orc, _ = run_oracle(s, tabs, 1, press=True)
    tot = orc.press().sum(0)
    h = 1e-6
    for a in range(3):
        es = []
        for sign in (1, -1):
            t = W.System(**{**s.__dict__})
            t.x = s.x.copy()
            t.x[:, a] *= 1 + sign * h
            t.box = s.box.copy()
            t.box[a, a] *= 1 + sign * h
            es.append(_energy(t, tabs)[0])
        dE = (es[0] - es[1]) / (2 * h)
        assert abs(-dE - tot[a]) < 1e-5 * abs(tot[a]), (a, -dE, tot[a])


@pytest.mark.parametrize("ensemble", ["nve", "nvt"])
def test_stress_kinetic_part(ensemble):
    """move_atoms adds p (x) p / m with the momenta after the step
    (src/imd_integrate.c:411-431, 1085-1094)."""
    eam = W.johnson_eam()
    s = W.eam_crystal(n_cells=4, temp_k=600.0, seed=8, displace=0.03)
    pair, rho, emb = [O.OracleTable.from_table(t)
                      for t in (eam.pair, eam.rho, eam.embed)]
    kw = {}
    if ensemble == "nvt":
        kw = dict(ensemble=O.ORC_NVT, temperature=W.kT(300.0), eta=0.01,
                  isq_tau_eta=1e-2)
    orc = O.OracleSystem(s, pair, rho, emb, nbl_margin=0.3, timestep=W.FS,
                         ntypes=1, **kw)
    orc.set_press(True)
    orc.calc_forces()
    pot = orc.press()
    orc.move_atoms()
    p = orc.atoms()["p"]
    m = s.mass[:, None]
    kin = np.stack([p[:, 0] * p[:, 0], p[:, 1] * p[:, 1], p[:, 2] * p[:, 2],
                    p[:, 1] * p[:, 2], p[:, 2] * p[:, 0], p[:, 0] * p[:, 1]],
                   axis=1) / m
    assert np.allclose(orc.press() - pot, kin, rtol=1e-12, atol=1e-15)
    # the diagonal carries 2 Ekin of the new momenta
    assert abs(kin[:, :3].sum() - (p * p / m).sum()) < 1e-12 * kin[:, :3].sum()


# ---------------------------------------------------------------------------
# EEAM (second density P = sum rho^2 and its energy term E(P))
# ---------------------------------------------------------------------------

def emod_table(nt, G, H, p_max, npts=400):
    fstep, px = W._grid(0.0, p_max, npts)
    return W.Table(np.zeros(nt), np.full(nt, p_max), np.full(nt, fstep),
                   [G[t] * px * px + H[t] * px for t in range(nt)], False)


@pytest.mark.parametrize("nt", [1, 2])
def test_eeam_closed_form(nt):
    """EEAM on quadratic tables: P_i, 2E'(P_i), energies, forces (with the
    DM_i rho_j rho'_j + DM_j rho_i rho'_i term of src/imd_forces_nbl.c:
    1204-1207), virial and stress against the brute-force closed form."""
    s = W.fcc_system((4, 4, 4), 3.25, [1.0] * nt, 0.0, 9, displace=0.15,
                     type_fractions=None if nt == 1 else (0.6, 0.4))
    A = [0.05, 0.08][:nt]
    B = [0.02, 0.035][:nt]
    C = [0.3, 0.2][:nt]
    D = [-1.0, -0.7][:nt]
    Gq = [0.01, -0.02][:nt]
    Hq = [-0.05, 0.03][:nt]
    ref = analytic(s.x, s.type, s.box, A, B, C, D, Gq, Hq)
    tabs = poly_tables(nt, A, B, C, D, rho_max=1.3 * ref["rho"].max())
    pair, rho, emb = [O.OracleTable.from_table(t) for t in tabs]
    emod = O.OracleTable.from_table(emod_table(nt, Gq, Hq, 1.3 * ref["P"].max()))
    orc = O.OracleSystem(s, pair, rho, emb, nbl_margin=0.3, timestep=0.001,
                         ntypes=nt)
    orc.set_eeam(emod)
    orc.set_press(True)
    orc.calc_forces()
    a = orc.atoms()
    P, dm = orc.eeam()
    assert np.allclose(P, ref["P"], rtol=1e-11, atol=0)
    assert np.allclose(dm, ref["dm"], rtol=1e-10, atol=1e-13)
    scale = np.abs(ref["f"]).max()
    assert np.max(np.abs(a["f"] - ref["f"])) < 1e-10 * scale
    assert np.allclose(a["epot"], ref["epot_i"], rtol=1e-10, atol=1e-12)
    assert abs(orc.scalar("epot") - ref["epot"]) < 1e-10 * abs(ref["epot"])
    assert abs(orc.scalar("virial") - ref["virial"]) < 1e-9 * abs(ref["virial"])
    ps = np.abs(ref["press"]).max()
    assert np.max(np.abs(orc.press() - ref["press"])) < 1e-10 * ps
    # the EEAM term is not negligible in this case
    plain = analytic(s.x, s.type, s.box, A, B, C, D)
    assert np.abs(plain["f"] - ref["f"]).max() > 1e-3 * scale


def test_eeam_forces_are_energy_gradients():
    """Johnson EAM plus a smooth E(P) table: forces are minus the finite-
    difference gradient of the energy."""
    eam = W.johnson_eam()
    s = W.eam_crystal(n_cells=4, temp_k=300.0, seed=4, displace=0.08)
    pmax = 40.0
    fstep, px = W._grid(0.0, pmax, 4000)
    et = W.Table(np.zeros(1), np.full(1, pmax), np.full(1, fstep),
                 [-0.4 * np.sqrt(px + 0.5) + 0.05 * px], False)

    def energy(sys_):
        pair, rho, emb = [O.OracleTable.from_table(t)
                          for t in (eam.pair, eam.rho, eam.embed)]
        emod = O.OracleTable.from_table(et)
        orc = O.OracleSystem(sys_, pair, rho, emb, nbl_margin=0.3,
                             timestep=0.001, ntypes=1)
        orc.set_eeam(emod)
        orc.calc_forces()
        return orc.scalar("epot"), orc

    e0, orc = energy(s)
    assert orc.eeam()[0].max() < pmax
    f = orc.atoms()["f"]
    h = 1e-5
    rng = np.random.default_rng(5)
    for atom in rng.choice(s.n, 5, replace=False):
        for k in range(3):
            sp = W.System(**{**s.__dict__})
            sp.x = s.x.copy()
            sp.x[atom, k] += h
            sm = W.System(**{**s.__dict__})
            sm.x = s.x.copy()
            sm.x[atom, k] -= h
            fd = -(energy(sp)[0] - energy(sm)[0]) / (2 * h)
            assert abs(fd - f[atom, k]) < 2e-5 + 1e-4 * abs(f[atom, k])
