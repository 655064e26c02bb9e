"""Shared builders for the parity tests: the same tables and systems go to
the oracle (through IMD-format files and its read_pot_table restatement)
and to the HIP library (through the C ABI)."""
from __future__ import annotations

import functools
import os
import tempfile

import numpy as np

from imd_amd import workloads as W


@functools.lru_cache(maxsize=None)
def eam_tables(nspecies=1, rc=5.0, npts=5000):
    species = (W.CU,) if nspecies == 1 else (W.CU, W.NI)
    return W.johnson_eam(species, rc=rc, npts=npts)


@functools.lru_cache(maxsize=None)
def lj():
    return W.lj_table()


def emod_table(nt, pmax=40.0, npts=4000):
    """A smooth EEAM energy term E(P) per type (not radial) on [0, pmax]."""
    step, px = W._grid(0.0, pmax, npts)
    cols = [(-0.4 - 0.1 * t) * np.sqrt(px + 0.5) + (0.05 + 0.02 * t) * px
            for t in range(nt)]
    return W.Table(np.zeros(nt), np.full(nt, pmax), np.full(nt, step), cols,
                   False)


def oracle_tables(eam: "W.EamTables", interp=0):
    import oracle as O
    pair = O.OracleTable.from_table(eam.pair, interp=interp)
    rho = O.OracleTable.from_table(eam.rho, interp=interp)
    emb = O.OracleTable.from_table(eam.embed, radial=False, interp=interp)
    return pair, rho, emb


def oracle_eam(system, eam, interp=0, **kw):
    """interp: oracle.INTERP_* (the same numbers as imdg.INTERP_*)"""
    import oracle as O
    pair, rho, emb = oracle_tables(eam, interp)
    return O.OracleSystem(system, pair, rho, emb, ntypes=len(eam.masses), **kw)


def oracle_lj(system, interp=0, **kw):
    import oracle as O
    pair = O.OracleTable.from_table(lj(), interp=interp)
    return O.OracleSystem(system, pair, None, None, ntypes=1, **kw)


def canon_pairs_gpu(pairs6, rep=0):
    """GPU pairs (rep, i_lo, i_hi, sx, sy, sz) of one replica as a set of
    (id_lo, id_hi, sx, sy, sz) tuples (ids are input indices)."""
    p = pairs6[pairs6[:, 0] == rep][:, 1:]
    return set(map(tuple, p.tolist()))


def canon_pairs_oracle(pairs5):
    return set(map(tuple, pairs5.tolist()))


def rel_rms(a, b):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    den = np.sqrt(np.mean(b * b))
    return float(np.sqrt(np.mean((a - b) ** 2)) / (den if den > 0 else 1.0))


def max_abs(a, b):
    return float(np.max(np.abs(np.asarray(a) - np.asarray(b))))
