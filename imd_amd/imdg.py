"""ctypes binding of libimdg.so (the C ABI in include/imdg.h).

The library is the product: every force, list and integration step runs in
its HIP kernels.  Loading fails loudly if libimdg.so is missing; there is no
CPU fallback.
"""
from __future__ import annotations

import ctypes as C
import os
from typing import Optional, Sequence

import numpy as np

from . import workloads as W

PKG = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(PKG, "libimdg.so")

FP64, MIXED = 0, 1
NVE, NVT = 0, 1
OK, ERR_ARG, ERR_HIP, ERR_NOMEM, ERR_GEOM, ERR_NOTABLE, ERR_STATE = range(7)

SYMBOLS = [
    "imdg_ctx_create", "imdg_ctx_destroy", "imdg_last_error",
    "imdg_set_tables", "imdg_set_atoms", "imdg_set_nvt", "imdg_calc_forces",
    "imdg_move_atoms", "imdg_check_nblist", "imdg_make_nblist",
    "imdg_deallocate_nblist", "imdg_run", "imdg_sync", "imdg_get_atoms",
    "imdg_get_thermo", "imdg_get_thermo_history", "imdg_get_cells",
    "imdg_get_pairs", "imdg_get_grid", "imdg_get_info", "imdg_profile",
    "imdg_kernel_times", "imdg_struct_size", "imdg_get_tuning",
    "imdg_set_tuning", "imdg_autotune", "imdg_set_stress", "imdg_get_stress",
    "imdg_set_eeam_table", "imdg_get_eeam", "imdg_check_nblist_noalloc",
]


class Config(C.Structure):
    _fields_ = [("precision", C.c_int), ("ensemble", C.c_int),
                ("ntypes", C.c_int), ("timestep", C.c_double),
                ("nbl_margin", C.c_double), ("nbl_size", C.c_double),
                ("device", C.c_int), ("lanes_per_atom", C.c_int),
                ("interpolation", C.c_int)]


# table interpolation (IMD's build options: 3-point, FOURPOINT, SPLINE)
INTERP_3POINT, INTERP_4POINT, INTERP_SPLINE = 0, 1, 2


class Table(C.Structure):
    _fields_ = [("ncols", C.c_int), ("maxsteps", C.c_int),
                ("begin", C.POINTER(C.c_double)), ("end", C.POINTER(C.c_double)),
                ("step", C.POINTER(C.c_double)),
                ("invstep", C.POINTER(C.c_double)),
                ("len", C.POINTER(C.c_int)), ("table", C.POINTER(C.c_double))]


class Replica(C.Structure):
    _fields_ = [("box_x", C.c_double * 3), ("box_y", C.c_double * 3),
                ("box_z", C.c_double * 3), ("pbc", C.c_int * 3),
                ("temperature", C.c_double), ("eta", C.c_double),
                ("isq_tau_eta", C.c_double)]


class Thermo(C.Structure):
    _fields_ = [("step", C.c_long), ("epot", C.c_double), ("ekin", C.c_double),
                ("virial", C.c_double), ("eta", C.c_double),
                ("max_disp2", C.c_double), ("rebuilt", C.c_int),
                ("natoms", C.c_int)]


class Info(C.Structure):
    _fields_ = [("nrep", C.c_int), ("natoms", C.c_long), ("nslots", C.c_long),
                ("kmax", C.c_int), ("kmax_seen", C.c_int),
                ("nbl_count", C.c_long), ("steps_done", C.c_long),
                ("is_short", C.c_int), ("lanes_per_atom", C.c_int),
                ("list_bytes", C.c_double), ("mean_k", C.c_double),
                ("tiles_fallback", C.c_int), ("exact_pairs", C.c_long),
                ("tables_in_lds", C.c_int)]


class Tuning(C.Structure):
    _fields_ = [("density_lanes", C.c_int), ("density_threads", C.c_int),
                ("force_lanes", C.c_int), ("force_threads", C.c_int),
                ("density_ms", C.c_double), ("force_ms", C.c_double)]


_lib = None


def lib():
    """Load libimdg.so (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError("libimdg.so is not built (%s); run "
                               "python -c 'from imd_amd import build; "
                               "build.build()'" % LIB_PATH)
        L = C.CDLL(LIB_PATH)
        vp, dp, ip, lp = C.c_void_p, C.POINTER(C.c_double), C.POINTER(C.c_int), C.POINTER(C.c_long)
        L.imdg_ctx_create.argtypes = [C.POINTER(Config), C.POINTER(vp)]
        L.imdg_ctx_destroy.argtypes = [vp]
        L.imdg_last_error.argtypes = [vp]
        L.imdg_last_error.restype = C.c_char_p
        L.imdg_set_tables.argtypes = [vp] + [C.POINTER(Table)] * 3
        L.imdg_set_atoms.argtypes = [vp, C.c_int, C.POINTER(Replica), ip, dp,
                                     dp, ip, dp]
        L.imdg_set_nvt.argtypes = [vp, C.c_int, C.c_double, C.c_double,
                                   C.c_double]
        L.imdg_calc_forces.argtypes = [vp, C.c_int]
        L.imdg_move_atoms.argtypes = [vp]
        L.imdg_check_nblist.argtypes = [vp, ip]
        L.imdg_make_nblist.argtypes = [vp]
        L.imdg_deallocate_nblist.argtypes = [vp]
        L.imdg_run.argtypes = [vp, C.c_long]
        L.imdg_sync.argtypes = [vp]
        L.imdg_get_atoms.argtypes = [vp] + [dp] * 6
        L.imdg_get_thermo.argtypes = [vp, C.c_int, C.POINTER(Thermo)]
        L.imdg_get_thermo_history.argtypes = [vp, C.c_long, C.c_long,
                                              C.POINTER(Thermo)]
        L.imdg_get_cells.argtypes = [vp, ip]
        L.imdg_get_pairs.argtypes = [vp, lp, C.c_long, lp]
        L.imdg_get_grid.argtypes = [vp, C.c_int, ip]
        L.imdg_set_stress.argtypes = [vp, C.c_int]
        L.imdg_set_eeam_table.argtypes = [vp, C.POINTER(Table)]
        L.imdg_get_eeam.argtypes = [vp, dp, dp]
        L.imdg_get_stress.argtypes = [vp, dp, dp]
        L.imdg_get_info.argtypes = [vp, C.POINTER(Info)]
        L.imdg_profile.argtypes = [vp, C.c_int]
        L.imdg_kernel_times.argtypes = [vp, C.POINTER(C.c_char_p), dp, lp,
                                        C.c_int]
        _lib = L
    return _lib


def _dp(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


class ImdgError(RuntimeError):
    pass


class TableArrays:
    """A read IMD table (pot_table_t contents) held as numpy arrays: the
    result of read_pot_table semantics (radial shift, two guard rows)."""

    def __init__(self, begin, end, step, invstep, length, maxsteps, table):
        self.begin = np.ascontiguousarray(begin, dtype=np.float64)
        self.end = np.ascontiguousarray(end, dtype=np.float64)
        self.step = np.ascontiguousarray(step, dtype=np.float64)
        self.invstep = np.ascontiguousarray(invstep, dtype=np.float64)
        self.len = np.ascontiguousarray(length, dtype=np.int32)
        self.maxsteps = int(maxsteps)
        self.table = np.ascontiguousarray(table, dtype=np.float64).ravel()
        self.ncols = self.begin.size

    @classmethod
    def from_format2(cls, t: "W.Table") -> "TableArrays":
        """Apply read_pot_table2 + init_threepoint (src/imd_potential.c:
        394-462, 1256-1272) to a generated format-2 table, exactly as the
        C host reader does; used where tests bypass files."""
        nc = t.ncols
        begin = np.asarray(t.begin, dtype=np.float64)
        end = np.asarray(t.end, dtype=np.float64)
        step = np.asarray(t.step, dtype=np.float64)
        invstep = 1.0 / step
        length = np.array([int(1 + (end[c] - begin[c]) / step[c] + 0.49)
                           for c in range(nc)], dtype=np.int32)
        maxsteps = int(length.max())
        tab = np.zeros((maxsteps + 2, nc))
        for c in range(nc):
            tab[:length[c], c] = t.values[c][:length[c]]
            if t.radial:
                delta = tab[length[c] - 1, c]
                if delta != 0.0:
                    tab[:length[c], c] -= delta
            n = length[c]
            y = tab[:, c]
            y[n] = 3 * y[n - 1] - 3 * y[n - 2] + y[n - 3]
            y[n + 1] = 6 * y[n - 1] - 8 * y[n - 2] + 3 * y[n - 3]
        return cls(begin, end, step, invstep, length, maxsteps, tab)

    def cstruct(self) -> Table:
        t = Table()
        t.ncols = self.ncols
        t.maxsteps = self.maxsteps
        t.begin = _dp(self.begin)
        t.end = _dp(self.end)
        t.step = _dp(self.step)
        t.invstep = _dp(self.invstep)
        t.len = self.len.ctypes.data_as(C.POINTER(C.c_int))
        t.table = _dp(self.table)
        return t


class Context:
    """One imdg context (one HIP stream, all device memory)."""

    def __init__(self, precision=MIXED, ensemble=NVE, ntypes=1,
                 timestep=0.001, nbl_margin=0.4, nbl_size=1.1, device=0,
                 lanes_per_atom=0, interpolation=INTERP_3POINT):
        L = lib()
        cfg = Config(precision, ensemble, ntypes, timestep, nbl_margin,
                     nbl_size, device, lanes_per_atom, interpolation)
        h = C.c_void_p()
        rc = L.imdg_ctx_create(C.byref(cfg), C.byref(h))
        if rc != 0:
            raise ImdgError("imdg_ctx_create failed with code %d" % rc)
        self._h = h
        self.ntypes = ntypes
        self._keep = []
        self.nrep = 0
        self.natoms = []

    def _chk(self, rc, what):
        if rc != 0:
            msg = lib().imdg_last_error(self._h)
            raise ImdgError("%s failed (%d): %s" % (what, rc,
                                                   msg.decode() if msg else ""))

    def close(self):
        if getattr(self, "_h", None):
            lib().imdg_ctx_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def set_tables(self, pair: Optional[TableArrays],
                   rho: Optional[TableArrays] = None,
                   embed: Optional[TableArrays] = None):
        structs = [t.cstruct() if t is not None else None
                   for t in (pair, rho, embed)]
        self._keep = [pair, rho, embed, structs]
        ptrs = [C.byref(s) if s is not None else C.POINTER(Table)()
                for s in structs]
        self._chk(lib().imdg_set_tables(self._h, *ptrs), "imdg_set_tables")

    def set_eeam_table(self, emod: Optional[TableArrays]):
        """EEAM energy term E(P) (ntypes columns); before set_systems."""
        st = emod.cstruct() if emod is not None else None
        self._keep_eeam = (emod, st)
        self._chk(lib().imdg_set_eeam_table(
            self._h, C.byref(st) if st is not None else C.POINTER(Table)()),
            "imdg_set_eeam_table")

    def eeam(self):
        """EEAM P_i and 2E'(P_i) per atom (input order)."""
        n = sum(self.natoms)
        P, dm = np.zeros(n), np.zeros(n)
        self._chk(lib().imdg_get_eeam(self._h, _dp(P), _dp(dm)),
                  "imdg_get_eeam")
        return P, dm

    def set_systems(self, systems: Sequence["W.System"], eta=0.0,
                    isq_tau_eta=0.0):
        n = len(systems)
        reps = (Replica * n)()
        for r, s in enumerate(systems):
            for k, name in enumerate(("box_x", "box_y", "box_z")):
                getattr(reps[r], name)[:] = list(map(float, s.box[k]))
            reps[r].pbc[:] = list(map(int, s.pbc))
            reps[r].temperature = float(s.temperature)
            reps[r].eta = eta
            reps[r].isq_tau_eta = isq_tau_eta
        nat = np.array([s.n for s in systems], dtype=np.int32)
        x = np.ascontiguousarray(np.concatenate([s.x for s in systems]))
        p = np.ascontiguousarray(np.concatenate([s.p for s in systems]))
        ty = np.ascontiguousarray(np.concatenate([s.type for s in systems]).astype(np.int32))
        ms = np.ascontiguousarray(np.concatenate([s.mass for s in systems]))
        self._chk(lib().imdg_set_atoms(self._h, n, reps,
                                       nat.ctypes.data_as(C.POINTER(C.c_int)),
                                       _dp(x), _dp(p),
                                       ty.ctypes.data_as(C.POINTER(C.c_int)),
                                       _dp(ms)), "imdg_set_atoms")
        self.nrep = n
        self.natoms = list(map(int, nat))

    def set_nvt(self, rep, temperature, eta, isq_tau_eta):
        self._chk(lib().imdg_set_nvt(self._h, rep, temperature, eta,
                                     isq_tau_eta), "imdg_set_nvt")

    def calc_forces(self, steps=0):
        self._chk(lib().imdg_calc_forces(self._h, steps), "imdg_calc_forces")

    def move_atoms(self):
        self._chk(lib().imdg_move_atoms(self._h), "imdg_move_atoms")

    def check_nblist(self):
        inv = C.c_int(0)
        self._chk(lib().imdg_check_nblist(self._h, C.byref(inv)),
                  "imdg_check_nblist")
        return bool(inv.value)

    def make_nblist(self):
        self._chk(lib().imdg_make_nblist(self._h), "imdg_make_nblist")

    def deallocate_nblist(self):
        self._chk(lib().imdg_deallocate_nblist(self._h),
                  "imdg_deallocate_nblist")

    def run(self, nsteps, sync=True):
        self._chk(lib().imdg_run(self._h, nsteps), "imdg_run")
        if sync:
            self.sync()

    def sync(self):
        self._chk(lib().imdg_sync(self._h), "imdg_sync")

    def atoms(self):
        n = sum(self.natoms)
        out = {k: np.zeros((n, 3)) for k in ("x", "p", "f")}
        out.update({k: np.zeros(n) for k in ("epot", "rho", "df")})
        self._chk(lib().imdg_get_atoms(self._h, _dp(out["x"]), _dp(out["p"]),
                                       _dp(out["f"]), _dp(out["epot"]),
                                       _dp(out["rho"]), _dp(out["df"])),
                  "imdg_get_atoms")
        return out

    def thermo(self, rep=0):
        t = Thermo()
        self._chk(lib().imdg_get_thermo(self._h, rep, C.byref(t)),
                  "imdg_get_thermo")
        return {k: getattr(t, k) for k, _ in Thermo._fields_}

    def thermo_history(self, first, count):
        arr = (Thermo * (count * self.nrep))()
        self._chk(lib().imdg_get_thermo_history(self._h, first, count, arr),
                  "imdg_get_thermo_history")
        out = np.zeros((count, self.nrep, 7))
        for s in range(count):
            for r in range(self.nrep):
                t = arr[s * self.nrep + r]
                out[s, r] = (t.step, t.epot, t.ekin, t.virial, t.eta,
                             t.max_disp2, t.rebuilt)
        return out

    def cells(self):
        c = np.zeros((sum(self.natoms), 3), dtype=np.int32)
        self._chk(lib().imdg_get_cells(self._h,
                                       c.ctypes.data_as(C.POINTER(C.c_int))),
                  "imdg_get_cells")
        return c

    def pairs(self):
        n = C.c_long(0)
        self._chk(lib().imdg_get_pairs(self._h, C.POINTER(C.c_long)(), 0,
                                       C.byref(n)), "imdg_get_pairs")
        out = np.zeros((max(n.value, 1), 6), dtype=np.int64)
        self._chk(lib().imdg_get_pairs(self._h,
                                       out.ctypes.data_as(C.POINTER(C.c_long)),
                                       n.value, C.byref(n)), "imdg_get_pairs")
        return out[:n.value]

    def grid(self, rep=0):
        g = (C.c_int * 3)()
        self._chk(lib().imdg_get_grid(self._h, rep, g), "imdg_get_grid")
        return tuple(g)

    def set_stress(self, enable=True):
        self._chk(lib().imdg_set_stress(self._h, int(enable)),
                  "imdg_set_stress")

    def stress(self):
        """Per-atom stress tensor (n, 6) in input order and the per-replica
        sums (nrep, 6); components xx yy zz yz zx xy."""
        per = np.zeros((sum(self.natoms), 6))
        tot = np.zeros((self.nrep, 6))
        self._chk(lib().imdg_get_stress(self._h, _dp(per), _dp(tot)),
                  "imdg_get_stress")
        return per, tot

    def info(self):
        i = Info()
        self._chk(lib().imdg_get_info(self._h, C.byref(i)), "imdg_get_info")
        return {k: getattr(i, k) for k, _ in Info._fields_}

    def tuning(self):
        """Launch shape of the density and force loops (lanes per atom,
        threads per block)."""
        t = Tuning()
        self._chk(lib().imdg_get_tuning(self._h, C.byref(t)),
                  "imdg_get_tuning")
        return {k: getattr(t, k) for k, _ in Tuning._fields_}

    def set_tuning(self, density_lanes, density_threads, force_lanes,
                   force_threads):
        t = Tuning(density_lanes, density_threads, force_lanes,
                   force_threads, 0.0, 0.0)
        self._chk(lib().imdg_set_tuning(self._h, C.byref(t)),
                  "imdg_set_tuning")

    def autotune(self, reps=5):
        """Time every launch shape on the current state, keep the fastest;
        returns the chosen shape with its mean launch times (ms)."""
        t = Tuning()
        self._chk(lib().imdg_autotune(self._h, int(reps), C.byref(t)),
                  "imdg_autotune")
        return {k: getattr(t, k) for k, _ in Tuning._fields_}

    KERNEL_CLASSES = ("rebuild", "density", "force", "move", "finalize")

    def profile(self, enable=True, classes=None):
        """Time launches with HIP events: every kernel class, or only the
        named ones (KERNEL_CLASSES)."""
        if not enable:
            mask = 0
        elif classes is None:
            mask = 1
        else:
            mask = sum(1 << self.KERNEL_CLASSES.index(k) for k in classes)
            if mask == 1:  # only "rebuild": 1 would mean every class
                mask = 1 | (1 << 31)
        self._chk(lib().imdg_profile(self._h, int(np.int32(np.uint32(mask)))),
                  "imdg_profile")

    def kernel_times(self):
        n = 8
        names = (C.c_char_p * n)()
        ms = np.zeros(n)
        cnt = np.zeros(n, dtype=np.int64)
        k = lib().imdg_kernel_times(self._h, names, _dp(ms),
                                    cnt.ctypes.data_as(C.POINTER(C.c_long)), n)
        return {names[q].decode(): (float(ms[q]), int(cnt[q]))
                for q in range(k)}


def eam_context(tables: "W.EamTables", systems, precision=MIXED,
                ensemble=NVE, timestep=W.FS, nbl_margin=0.4, eta=0.0,
                isq_tau_eta=0.0, emod=None, **kw) -> Context:
    """Context with IMD-read EAM tables and the given systems; emod: an
    EEAM energy-term table (workloads.Table, ntypes columns)."""
    ntypes = len(tables.masses)
    ctx = Context(precision=precision, ensemble=ensemble, ntypes=ntypes,
                  timestep=timestep, nbl_margin=nbl_margin, **kw)
    ctx.set_tables(TableArrays.from_format2(tables.pair),
                   TableArrays.from_format2(tables.rho),
                   TableArrays.from_format2(tables.embed))
    if emod is not None:
        ctx.set_eeam_table(TableArrays.from_format2(emod))
    ctx.set_systems(systems, eta=eta, isq_tau_eta=isq_tau_eta)
    return ctx


def pair_context(table: "W.Table", systems, precision=MIXED, ensemble=NVE,
                 timestep=0.002, nbl_margin=0.3, **kw) -> Context:
    ctx = Context(precision=precision, ensemble=ensemble, ntypes=1,
                  timestep=timestep, nbl_margin=nbl_margin, **kw)
    ctx.set_tables(TableArrays.from_format2(table), None, None)
    ctx.set_systems(systems)
    return ctx
