"""ctypes binding of the oracle (test infrastructure only).

The oracle is a plain-C FP64 restatement of IMD's neighbour-list pair/EAM
path (imd_oracle.c).  Only tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py import this module; the product never does.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")

ORC_NVE, ORC_NVT = 0, 1
INTERP_3POINT, INTERP_4POINT, INTERP_SPLINE = 0, 1, 2


def build(force: bool = True) -> str:
    if force or not os.path.exists(LIB_PATH):
        subprocess.run(["make", "-B", "-C", HERE, "liboracle.so"], check=True,
                       stdout=subprocess.DEVNULL)
    return LIB_PATH


class _Table(C.Structure):
    _fields_ = [("ncols", C.c_int), ("maxsteps", C.c_int),
                ("begin", C.POINTER(C.c_double)), ("end", C.POINTER(C.c_double)),
                ("step", C.POINTER(C.c_double)), ("invstep", C.POINTER(C.c_double)),
                ("len", C.POINTER(C.c_int)), ("table", C.POINTER(C.c_double)),
                ("interp", C.c_int), ("table2", C.POINTER(C.c_double))]


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build(force=False)
        L = C.CDLL(LIB_PATH)
        dp, ip, lp = C.POINTER(C.c_double), C.POINTER(C.c_int), C.POINTER(C.c_long)
        L.orc_table_read.argtypes = [C.POINTER(_Table), C.c_char_p, C.c_int,
                                     C.c_int, C.c_int, dp]
        L.orc_table_free.argtypes = [C.POINTER(_Table)]
        L.orc_table_set_interp.argtypes = [C.POINTER(_Table), C.c_int, C.c_int]
        L.orc_pair_int2.argtypes = [C.POINTER(_Table), C.c_int, C.c_int,
                                    C.c_double, dp, dp, ip]
        L.orc_val_func2.argtypes = [C.POINTER(_Table), C.c_int, C.c_int,
                                    C.c_double, ip]
        L.orc_val_func2.restype = C.c_double
        L.orc_deriv_func2.argtypes = L.orc_val_func2.argtypes
        L.orc_deriv_func2.restype = C.c_double
        L.orc_new.argtypes = [C.c_int, dp, ip, C.c_double, C.c_double, C.c_int,
                              C.c_double]
        L.orc_new.restype = C.c_void_p
        L.orc_free.argtypes = [C.c_void_p]
        L.orc_set_tables.argtypes = [C.c_void_p] + [C.POINTER(_Table)] * 3
        L.orc_set_nvt.argtypes = [C.c_void_p, C.c_double, C.c_double, C.c_double]
        L.orc_set_atoms.argtypes = [C.c_void_p, C.c_int, lp, ip, dp, dp, dp]
        L.orc_calc_forces.argtypes = [C.c_void_p, C.c_int]
        L.orc_move_atoms.argtypes = [C.c_void_p]
        L.orc_check_nblist.argtypes = [C.c_void_p]
        L.orc_invalidate_nblist.argtypes = [C.c_void_p]
        L.orc_run.argtypes = [C.c_void_p, C.c_int, dp]
        L.orc_get_atoms.argtypes = [C.c_void_p] + [dp] * 6
        L.orc_get_scalar.argtypes = [C.c_void_p, C.c_int]
        L.orc_get_scalar.restype = C.c_double
        L.orc_get_grid.argtypes = [C.c_void_p, ip]
        L.orc_get_cells.argtypes = [C.c_void_p, ip]
        L.orc_set_press.argtypes = [C.c_void_p, C.c_int]
        L.orc_set_threads.argtypes = [C.c_void_p, C.c_int]
        L.orc_set_eeam.argtypes = [C.c_void_p, C.POINTER(_Table)]
        L.orc_get_eeam.argtypes = [C.c_void_p, dp, dp]
        L.orc_get_press.argtypes = [C.c_void_p, dp]
        L.orc_get_pairs.argtypes = [C.c_void_p, lp, C.c_long]
        L.orc_get_pairs.restype = C.c_long
        _lib = L
    return _lib


def _dp(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


class OracleTable:
    """An IMD function table read through read_pot_table semantics."""

    def __init__(self, path: str, ncols: int, radial: bool,
                 default_format: int = 2, interp: int = INTERP_3POINT):
        self._t = _Table()
        self.cellsz = C.c_double(0.0)
        rc = lib().orc_table_read(C.byref(self._t), path.encode(), ncols,
                                  int(radial), default_format,
                                  C.byref(self.cellsz))
        if rc != 0:
            raise ValueError("orc_table_read(%s) failed with %d" % (path, rc))
        self.ncols = ncols
        if interp != INTERP_3POINT:
            rc = lib().orc_table_set_interp(C.byref(self._t), interp,
                                            int(radial))
            if rc != 0:
                raise ValueError("orc_table_set_interp failed with %d" % rc)

    @classmethod
    def from_table(cls, table, ncols=None, radial=None, interp=INTERP_3POINT):
        """Write a workloads.Table to a temporary file and read it back."""
        fd, path = tempfile.mkstemp(suffix=".pt")
        os.close(fd)
        try:
            table.write(path)
            return cls(path, ncols or table.ncols,
                       table.radial if radial is None else radial,
                       interp=interp)
        finally:
            os.unlink(path)

    def __del__(self):
        try:
            lib().orc_table_free(C.byref(self._t))
        except Exception:
            pass

    def info(self):
        t = self._t
        n = t.ncols
        return dict(begin=np.ctypeslib.as_array(t.begin, (n,)).copy(),
                    end=np.ctypeslib.as_array(t.end, (n,)).copy(),
                    step=np.ctypeslib.as_array(t.step, (n,)).copy(),
                    invstep=np.ctypeslib.as_array(t.invstep, (n,)).copy(),
                    len=np.ctypeslib.as_array(t.len, (n,)).copy(),
                    maxsteps=t.maxsteps,
                    table=np.ctypeslib.as_array(
                        t.table, ((t.maxsteps + 2) * n,)).copy().reshape(-1, n),
                    table2=(np.ctypeslib.as_array(
                        t.table2, ((t.maxsteps + 2) * n,)).copy().reshape(-1, n)
                        if t.table2 else None))

    def pair_int2(self, col, r2):
        pot, grad, short = C.c_double(), C.c_double(), C.c_int(0)
        lib().orc_pair_int2(C.byref(self._t), col, self._t.ncols, r2,
                            C.byref(pot), C.byref(grad), C.byref(short))
        return pot.value, grad.value, short.value

    def val_func2(self, col, r2):
        s = C.c_int(0)
        return lib().orc_val_func2(C.byref(self._t), col, self._t.ncols, r2,
                                   C.byref(s))

    def deriv_func2(self, col, r2):
        s = C.c_int(0)
        return lib().orc_deriv_func2(C.byref(self._t), col, self._t.ncols, r2,
                                     C.byref(s))


class OracleSystem:
    """One IMD system run through the oracle's calc_forces/move_atoms."""

    def __init__(self, sysobj, pair=None, rho=None, embed=None,
                 nbl_margin=0.4, timestep=0.001, ensemble=ORC_NVE,
                 temperature=0.0, eta=0.0, isq_tau_eta=0.0, ntypes=None):
        ntypes = ntypes or int(sysobj.type.max()) + 1
        cellsz = max([t.cellsz.value for t in (pair, rho) if t is not None])
        box = np.ascontiguousarray(sysobj.box, dtype=np.float64).ravel()
        pbc = (C.c_int * 3)(*sysobj.pbc)
        L = lib()
        self._s = L.orc_new(ntypes, _dp(box), pbc, nbl_margin, cellsz,
                            ensemble, timestep)
        if not self._s:
            raise ValueError("orc_new failed (box too small for the cutoff?)")
        self._tabs = (pair, rho, embed)
        null = C.POINTER(_Table)()
        rc = L.orc_set_tables(self._s,
                              C.byref(pair._t) if pair else null,
                              C.byref(rho._t) if rho else null,
                              C.byref(embed._t) if embed else null)
        if rc:
            raise ValueError("orc_set_tables failed %d" % rc)
        L.orc_set_nvt(self._s, temperature, eta, isq_tau_eta)
        self.n = sysobj.n
        ids = np.ascontiguousarray(sysobj.id, dtype=np.int64)
        ty = np.ascontiguousarray(sysobj.type, dtype=np.int32)
        ms = np.ascontiguousarray(sysobj.mass, dtype=np.float64)
        x = np.ascontiguousarray(sysobj.x, dtype=np.float64)
        p = np.ascontiguousarray(sysobj.p, dtype=np.float64)
        rc = L.orc_set_atoms(self._s, self.n,
                             ids.ctypes.data_as(C.POINTER(C.c_long)),
                             ty.ctypes.data_as(C.POINTER(C.c_int)),
                             _dp(ms), _dp(x), _dp(p))
        if rc:
            raise ValueError("orc_set_atoms failed %d" % rc)
        self.ids = ids

    def __del__(self):
        try:
            lib().orc_free(self._s)
        except Exception:
            pass

    def calc_forces(self, steps=0):
        rc = lib().orc_calc_forces(self._s, steps)
        if rc:
            raise RuntimeError("orc_calc_forces %d" % rc)

    def move_atoms(self):
        lib().orc_move_atoms(self._s)

    def check_nblist(self):
        lib().orc_check_nblist(self._s)

    def run(self, nsteps):
        th = np.zeros((nsteps, 5))
        rc = lib().orc_run(self._s, nsteps, _dp(th))
        if rc:
            raise RuntimeError("orc_run %d" % rc)
        return th

    def atoms(self):
        n = self.n
        out = {k: np.zeros((n, 3)) for k in ("x", "p", "f")}
        out.update({k: np.zeros(n) for k in ("epot", "rho", "df")})
        lib().orc_get_atoms(self._s, _dp(out["x"]), _dp(out["p"]),
                            _dp(out["f"]), _dp(out["epot"]), _dp(out["rho"]),
                            _dp(out["df"]))
        return out

    def set_eeam(self, emod):
        """EEAM energy term E(P) (an OracleTable, ntypes columns)."""
        self._emod = emod
        rc = lib().orc_set_eeam(self._s, C.byref(emod._t) if emod else
                                C.POINTER(_Table)())
        if rc:
            raise ValueError("orc_set_eeam failed %d" % rc)

    def eeam(self):
        """P_i = sum_j rho_ij^2 and 2E'(P_i), caller order."""
        P, dm = np.zeros(self.n), np.zeros(self.n)
        lib().orc_get_eeam(self._s, _dp(P), _dp(dm))
        return P, dm

    def set_press(self, on=True):
        lib().orc_set_press(self._s, int(on))

    def set_threads(self, n):
        """Threads for timing runs (1 = IMD's serial order, the checker)."""
        lib().orc_set_threads(self._s, int(n))

    def press(self):
        """Per-atom stress tensor (n, 6): xx yy zz yz zx xy."""
        out = np.zeros((self.n, 6))
        lib().orc_get_press(self._s, _dp(out))
        return out

    def scalar(self, which):
        names = dict(epot=0, virial=1, ekin=2, eta=3, cellsz=4, volume=5,
                     is_short=6, nbl_count=7, natoms=8, virial_abs=9)
        return lib().orc_get_scalar(self._s, names.get(which, which))

    def grid(self):
        g = (C.c_int * 3)()
        lib().orc_get_grid(self._s, g)
        return tuple(g)

    def cells(self):
        c = np.zeros((self.n, 3), dtype=np.int32)
        lib().orc_get_cells(self._s, c.ctypes.data_as(C.POINTER(C.c_int)))
        return c

    def pairs(self):
        L = lib()
        npairs = L.orc_get_pairs(self._s, C.POINTER(C.c_long)(), 0)
        out = np.zeros((max(npairs, 1), 5), dtype=np.int64)
        L.orc_get_pairs(self._s, out.ctypes.data_as(C.POINTER(C.c_long)),
                        npairs)
        return out[:npairs]
