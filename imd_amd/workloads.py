"""Synthetic inputs for the neighbour-list pair/EAM path.

Everything here writes or returns data in IMD's own conventions so that the
GPU library, the C host and the oracle read identical bytes:

* fcc sites at ``(x + 0.5) * a / 2`` for integer ``x, y, z`` with even sum,
  as ``generate_fcc`` places them (src/imd_generate.c:306-486, site formula
  at 453-455);
* potential tables in IMD format 2 (``#F 2 ncols`` header, one
  ``begin end step`` line per column, then the values column after column;
  src/imd_potential.c:394-462), tabulated in r^2 for radial functions;
* configuration files in IMD's ASCII format
  (``#F A 1 1 1 3 3 0``, rows ``number type mass x y z vx vy vz``;
  src/imd_io_3d.c:240-380, header src/imd_io.c:3369-3497).

Units are IMD's implicit ones: Angstrom, eV, amu; time unit
sqrt(amu*A^2/eV) = 10.1805 fs; temperature as k_B*T in eV
(src/imd_maxwell.c:146, src/imd_io.c:2344).
"""
from __future__ import annotations

import dataclasses
import math
import os
from typing import Optional, Sequence

import numpy as np

KB_EV = 8.617333262e-5                       # eV / K
TIME_UNIT_FS = 10.180505671156723            # sqrt(amu A^2 / eV) in fs
FS = 1.0 / TIME_UNIT_FS                      # 1 fs in IMD time units
MASS_CU = 63.546
MASS_NI = 58.6934


def kT(kelvin: float) -> float:
    return KB_EV * kelvin


# --------------------------------------------------------------------------
# lattices
# --------------------------------------------------------------------------

def fcc_positions(nx: int, ny: int, nz: int, a: float) -> np.ndarray:
    """fcc block of nx*ny*nz conventional cells (4 atoms each), sites at
    (x+0.5)*a/2 for integer x,y,z with x+y+z even, sorted x-major."""
    ix, iy, iz = np.meshgrid(np.arange(2 * nx), np.arange(2 * ny),
                             np.arange(2 * nz), indexing="ij")
    keep = ((ix + iy + iz) % 2) == 0
    pts = np.stack([ix[keep], iy[keep], iz[keep]], axis=1).astype(np.float64)
    return (pts + 0.5) * (a / 2.0)


def ortho_box(lx: float, ly: float, lz: float) -> np.ndarray:
    return np.array([[lx, 0.0, 0.0], [0.0, ly, 0.0], [0.0, 0.0, lz]])


def maxwell_momenta(rng: np.random.Generator, mass: np.ndarray,
                    temp: float) -> np.ndarray:
    """p ~ N(0, sqrt(kT m)) per component, then the mean momentum removed,
    as maxwell() does (src/imd_maxwell.c:146-160, 274-299)."""
    sig = np.sqrt(temp * mass)[:, None]
    p = rng.standard_normal((mass.size, 3)) * sig
    p -= p.mean(axis=0, keepdims=True)
    return p


# --------------------------------------------------------------------------
# tables
# --------------------------------------------------------------------------

@dataclasses.dataclass
class Table:
    """One IMD format-2 table: per column begin/end/step and values."""
    begin: np.ndarray
    end: np.ndarray
    step: np.ndarray
    values: list          # list of 1-D arrays, one per column
    radial: bool = True

    @property
    def ncols(self) -> int:
        return len(self.values)

    def write(self, path: str) -> str:
        with open(path, "w") as f:
            f.write("#F 2 %d\n" % self.ncols)
            f.write("#C generated for the MI355X neighbour-list path\n")
            f.write("#E\n")
            for c in range(self.ncols):
                f.write("%.17g %.17g %.17g\n" % (self.begin[c], self.end[c],
                                                 self.step[c]))
            for c in range(self.ncols):
                np.savetxt(f, self.values[c], fmt="%.17g")
                f.write("\n")
        return path


def _grid(begin: float, end: float, npts: int):
    step = (end - begin) / (npts - 1)
    # IMD recomputes the length as (int)(1 + (end-begin)/step + 0.49)
    n = int(1 + (end - begin) / step + 0.49)
    assert n == npts, (n, npts)
    return step, begin + step * np.arange(npts)


def table_from_functions(funcs: Sequence, begin: float, end: float,
                         npts: int, radial: bool = True,
                         in_r2: bool = True) -> Table:
    """Sample callables f(r) (radial tables, argument r = sqrt(r^2)) or
    f(x) (non-radial) on an equidistant grid."""
    step, xs = _grid(begin, end, npts)
    arg = np.sqrt(xs) if (radial and in_r2) else xs
    vals = [np.asarray(fn(arg), dtype=np.float64) for fn in funcs]
    nc = len(funcs)
    return Table(np.full(nc, begin), np.full(nc, end), np.full(nc, step),
                 vals, radial)


def taper(r, rs: float, rc: float):
    """C2 polynomial switch: 1 below rs, 0 above rc."""
    x = np.clip((np.asarray(r, dtype=np.float64) - rs) / (rc - rs), 0.0, 1.0)
    return 1.0 - x ** 3 * (10.0 - 15.0 * x + 6.0 * x * x)


@dataclasses.dataclass
class JohnsonParams:
    """Johnson's analytic EAM form (PRB 37, 3924 (1988); PRB 39, 12554
    (1989)): phi = phi_e exp(-gamma (r/r_e-1)), f = f_e exp(-beta (r/r_e-1)),
    F(rho) = -E_c (1 - (alpha/beta) ln s) s^(alpha/beta)
             - 6 phi_e s^(gamma/beta),  s = rho/rho_e."""
    r_e: float = 2.556
    f_e: float = 1.0
    phi_e: float = 0.59
    E_c: float = 3.54
    alpha: float = 5.09
    beta: float = 5.85
    gamma: float = 8.00
    mass: float = MASS_CU
    a0: float = 3.615


CU = JohnsonParams()
NI = JohnsonParams(r_e=2.489, f_e=1.10, phi_e=0.61, E_c=4.45, alpha=4.98,
                   beta=6.10, gamma=8.40, mass=MASS_NI, a0=3.52)


def _phi(p: JohnsonParams, rs, rc):
    return lambda r: p.phi_e * np.exp(-p.gamma * (r / p.r_e - 1.0)) * taper(r, rs, rc)


def _dens(p: JohnsonParams, rs, rc):
    return lambda r: p.f_e * np.exp(-p.beta * (r / p.r_e - 1.0)) * taper(r, rs, rc)


def _embed(p: JohnsonParams, rho_e: float):
    def F(rho):
        s = np.asarray(rho, dtype=np.float64) / rho_e
        out = np.zeros_like(s)
        m = s > 0
        ab, gb = p.alpha / p.beta, p.gamma / p.beta
        out[m] = (-p.E_c * (1.0 - ab * np.log(s[m])) * s[m] ** ab
                  - 6.0 * p.phi_e * s[m] ** gb)
        return out
    return F


def fcc_shell_density(dens, a: float, rc: float) -> float:
    """Host density of a perfect fcc lattice (closed form over shells)."""
    n = int(math.ceil(rc / a)) + 1
    rng_ = np.arange(-2 * n, 2 * n + 1)
    ix, iy, iz = np.meshgrid(rng_, rng_, rng_, indexing="ij")
    m = ((ix + iy + iz) % 2 == 0) & ~((ix == 0) & (iy == 0) & (iz == 0))
    r = np.sqrt(ix[m] ** 2 + iy[m] ** 2 + iz[m] ** 2) * (a / 2.0)
    r = r[r < rc]
    return float(np.sum(dens(r)))


@dataclasses.dataclass
class EamTables:
    pair: Table
    rho: Table
    embed: Table
    rc: float
    masses: list

    def write(self, directory: str, prefix: str = "eam") -> dict:
        os.makedirs(directory, exist_ok=True)
        return {
            "core_potential_file": self.pair.write(os.path.join(directory, prefix + "_phi.pt")),
            "atomic_e-density_file": self.rho.write(os.path.join(directory, prefix + "_rho.pt")),
            "embedding_energy_file": self.embed.write(os.path.join(directory, prefix + "_F.pt")),
        }


def johnson_eam(species: Sequence[JohnsonParams] = (CU,), rc: float = 5.0,
                rs: Optional[float] = None, npts: int = 5000,
                r2_begin: float = 1.0) -> EamTables:
    """Johnson-form EAM sampled into IMD tables: phi and rho on
    r^2 in [r2_begin, rc^2], F on rho in [0, 1.5 rho_eq], npts each.
    Pair columns are it*nt+jt (cross term = mean of the pure ones), density
    columns it*nt+jt hold f_jt (density at an it-atom due to a jt-atom, the
    column convention of src/imd_forces_nbl.c:1156-1160)."""
    rs = rc - 0.5 if rs is None else rs
    nt = len(species)
    phis = [_phi(p, rs, rc) for p in species]
    dens = [_dens(p, rs, rc) for p in species]
    pair_f, rho_f = [], []
    for it in range(nt):
        for jt in range(nt):
            if it == jt:
                pair_f.append(phis[it])
            else:
                pa, pb = phis[it], phis[jt]
                pair_f.append(lambda r, pa=pa, pb=pb: 0.5 * (pa(r) + pb(r)))
            rho_f.append(dens[jt])
    pair = table_from_functions(pair_f, r2_begin, rc * rc, npts)
    rho = table_from_functions(rho_f, r2_begin, rc * rc, npts)
    rho_eq = [fcc_shell_density(dens[t], species[t].a0, rc) for t in range(nt)]
    emb_cols, emb_end = [], 1.5 * max(rho_eq)
    for t in range(nt):
        emb_cols.append(_embed(species[t], rho_eq[t]))
    embed = table_from_functions(emb_cols, 0.0, emb_end, npts, radial=False)
    return EamTables(pair, rho, embed, rc, [p.mass for p in species])


def lj_table(r2_begin: float = 0.25, r2_end: float = 6.25, npts: int = 2000,
             eps: float = 1.0, sigma: float = 1.0) -> Table:
    """Lennard-Jones 4 eps ((s/r)^12 - (s/r)^6) tabulated in r^2; IMD's
    radial shift makes it zero at the cutoff when read."""
    def lj(r):
        sr6 = (sigma / r) ** 6
        return 4.0 * eps * (sr6 * sr6 - sr6)
    return table_from_functions([lj], r2_begin, r2_end, npts)


# --------------------------------------------------------------------------
# configurations
# --------------------------------------------------------------------------

@dataclasses.dataclass
class System:
    """One periodic (or slab) system in IMD conventions."""
    x: np.ndarray          # (n, 3) positions
    p: np.ndarray          # (n, 3) momenta p = m v
    type: np.ndarray       # (n,) int32
    mass: np.ndarray       # (n,) float64
    id: np.ndarray         # (n,) int64 atom numbers
    box: np.ndarray        # (3, 3) rows box_x, box_y, box_z
    pbc: tuple = (1, 1, 1)
    temperature: float = 0.0

    @property
    def n(self) -> int:
        return int(self.x.shape[0])

    def write_config(self, path: str) -> str:
        """IMD ASCII configuration with velocities (v = p / m)."""
        v = self.p / self.mass[:, None]
        with open(path, "w") as f:
            f.write("#F A 1 1 1 3 3 0\n")
            f.write("#C number type mass x y z vx vy vz\n")
            for k, lab in enumerate("XYZ"):
                f.write("#%s %.17g %.17g %.17g\n" % ((lab,) + tuple(self.box[k])))
            f.write("#E\n")
            rows = np.column_stack([self.mass, self.x, v])
            for i in range(self.n):
                f.write("%d %d " % (self.id[i], self.type[i]))
                f.write(" ".join("%.17g" % val for val in rows[i]))
                f.write("\n")
        return path


def fcc_system(ncells: Sequence[int], a: float, masses: Sequence[float],
               temp: float, seed: int, displace: float = 0.0,
               types: Optional[np.ndarray] = None,
               type_fractions: Optional[Sequence[float]] = None,
               vacancy_fraction: float = 0.0, pbc=(1, 1, 1),
               vacuum_z: float = 0.0) -> System:
    nx, ny, nz = ncells
    rng = np.random.default_rng(seed)
    x = fcc_positions(nx, ny, nz, a)
    if vacancy_fraction > 0.0:
        keep = rng.random(x.shape[0]) >= vacancy_fraction
        x = x[keep]
    n = x.shape[0]
    if types is None:
        if type_fractions is None:
            types = np.zeros(n, dtype=np.int32)
        else:
            types = rng.choice(len(type_fractions), size=n,
                               p=np.asarray(type_fractions)).astype(np.int32)
    masses = np.asarray(masses, dtype=np.float64)
    mass = masses[types]
    if displace > 0.0:
        x = x + rng.normal(0.0, displace, size=x.shape)
    box = ortho_box(nx * a, ny * a, nz * a + vacuum_z)
    if vacuum_z > 0.0:
        x[:, 2] += 0.5 * vacuum_z
    p = maxwell_momenta(rng, mass, temp) if temp > 0 else np.zeros_like(x)
    ids = np.arange(n, dtype=np.int64)
    return System(x=x, p=p, type=types.astype(np.int32), mass=mass, id=ids,
                  box=box, pbc=tuple(pbc), temperature=temp)


# --------------------------------------------------------------------------
# the configurations the path is measured on
# --------------------------------------------------------------------------

def lj_crystal(n_cells: int = 10, a: float = 1.5496, temp: float = 0.7,
               seed: int = 1000) -> System:
    """4,000-atom fcc LJ crystal at a = 1.5496 sigma, Maxwell at kT = 0.7."""
    return fcc_system((n_cells,) * 3, a, [1.0], temp, seed)


def eam_crystal(n_cells: int = 40, temp_k: float = 600.0, seed: int = 1001,
                displace: float = 0.02, a: float = 3.615) -> System:
    """fcc Cu-like EAM block: n_cells^3 cells at a = 3.615 A, displacements
    N(0, 0.02 A), Maxwell velocities at temp_k."""
    return fcc_system((n_cells,) * 3, a, [MASS_CU], kT(temp_k), seed,
                      displace=displace)


def alloy_slab(nxyz=(200, 200, 50), temp_k: float = 500.0, seed: int = 1004,
               a: float = 3.615, vacuum: float = 20.0) -> System:
    """Binary fcc slab, 70/30 types, periodic in x/y, open z with vacuum."""
    return fcc_system(nxyz, a, [MASS_CU, MASS_NI], kT(temp_k), seed,
                      type_fractions=(0.7, 0.3), pbc=(1, 1, 0),
                      vacuum_z=vacuum)


def replica_set(nrep: int = 64, seed: int = 1003, nmin: int = 500,
                nmax: int = 50000, a: float = 3.615) -> list:
    """Independent fcc blocks: N log-uniform in [nmin, nmax], block of
    cells closest to N/4 (at least 3 per side), 0-1% vacancies,
    temperatures uniform in [300, 1200] K."""
    rng = np.random.default_rng(seed)
    out = []
    for r in range(nrep):
        n_s = math.exp(rng.uniform(math.log(nmin), math.log(nmax)))
        cells = max(3, int(round((n_s / 4.0) ** (1.0 / 3.0))))
        dims = [cells, cells, cells]
        # stretch one side so the block is closest to n_s / 4 cells
        best = None
        for dx in range(-1, 2):
            for dy in range(-1, 2):
                cand = (max(3, cells + dx), max(3, cells + dy), cells)
                err = abs(cand[0] * cand[1] * cand[2] * 4 - n_s)
                if best is None or err < best[0]:
                    best = (err, cand)
        dims = best[1]
        vac = float(rng.uniform(0.0, 0.01))
        temp = float(rng.uniform(300.0, 1200.0))
        sub_seed = int(rng.integers(1 << 30))
        out.append(fcc_system(dims, a, [MASS_CU], kT(temp), sub_seed,
                              vacancy_fraction=vac))
    return out
