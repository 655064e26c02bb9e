"""The C host (imd_amd/host): IMD parameter files, table reader, config
reader.  CPU only: nothing here calls into the GPU."""
import ctypes as C
import os
import subprocess
import sys
import tempfile

import numpy as np
import pytest

import oracle as O
from imd_amd import build as B
from imd_amd import workloads as W

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


class PotTable(C.Structure):
    _fields_ = [("begin", C.POINTER(C.c_double)), ("end", C.POINTER(C.c_double)),
                ("step", C.POINTER(C.c_double)), ("invstep", C.POINTER(C.c_double)),
                ("len", C.POINTER(C.c_int)), ("ncols", C.c_int),
                ("maxsteps", C.c_int), ("table", C.POINTER(C.c_double))]


@pytest.fixture(scope="module")
def host():
    if not os.path.exists(B.HOSTLIB):
        B.build_host()
    L = C.CDLL(B.HOSTLIB)
    L.read_pot_table.argtypes = [C.POINTER(PotTable), C.c_char_p, C.c_int, C.c_int]
    L.read_parameters.argtypes = [C.c_char_p, C.c_int]
    L.read_atoms.argtypes = [C.c_char_p]
    return L


def table_arrays(pt):
    n = pt.ncols
    rows = pt.maxsteps + 2
    return dict(begin=np.ctypeslib.as_array(pt.begin, (n,)).copy(),
                len=np.ctypeslib.as_array(pt.len, (n,)).copy(),
                invstep=np.ctypeslib.as_array(pt.invstep, (n,)).copy(),
                maxsteps=pt.maxsteps,
                table=np.ctypeslib.as_array(pt.table, (rows * n,)).copy().reshape(rows, n))


@pytest.mark.parametrize("which", ["pair", "rho", "embed"])
def test_host_table_reader_matches_oracle(host, which, tmp_path):
    eam = W.johnson_eam((W.CU, W.NI), rc=5.5, npts=800)
    tab = getattr(eam, which)
    path = str(tmp_path / ("t_%s.pt" % which))
    tab.write(path)
    radial = which != "embed"
    C.c_int.in_dll(host, "have_eam").value = 1
    C.c_double.in_dll(host, "cellsz").value = 0.0
    pt = PotTable()
    assert host.read_pot_table(C.byref(pt), path.encode(), tab.ncols, int(radial)) == 0
    mine = table_arrays(pt)
    ref = O.OracleTable(path, tab.ncols, radial).info()
    assert np.array_equal(mine["len"], ref["len"])
    assert np.array_equal(mine["begin"], ref["begin"])
    assert np.array_equal(mine["invstep"], ref["invstep"])
    assert mine["maxsteps"] == ref["maxsteps"]
    assert np.array_equal(mine["table"], ref["table"])
    if radial:
        assert C.c_double.in_dll(host, "cellsz").value == 5.5 ** 2


def test_host_format1_reader_matches_oracle(host, tmp_path):
    rows = ["%.17g %.17g %.17g" % (1.0 + 0.25 * k, 7.0 - k, (3.0 - 0.5 * k) if k < 5 else 0.0)
            for k in range(12)]
    path = str(tmp_path / "f1.pt")
    open(path, "w").write("#F 1 2\n#E\n" + "\n".join(rows) + "\n")
    pt = PotTable()
    assert host.read_pot_table(C.byref(pt), path.encode(), 2, 1) == 0
    mine = table_arrays(pt)
    ref = O.OracleTable(path, 2, True).info()
    assert np.array_equal(mine["len"], ref["len"])
    # rows past a column's guard rows are unused (and unset in format 1)
    for c in range(2):
        n = mine["len"][c] + 2
        assert np.array_equal(mine["table"][:n, c], ref["table"][:n, c])


def test_parameter_file(host, tmp_path):
    p = tmp_path / "run.param"
    p.write_text("""# comment
ensemble   nvt
ntypes     2
masses     63.546 58.69
timestep   0.098227
maxsteps   100 # trailing comment
eng_int    10
starttemp  0.0689387
tau_eta    9.8227
nbl_margin 0.4
pbc_dirs   1 1 0
box_x 10 0 0
box_y = 0 11 0
box_z 0 0 12
core_potential_file a.pt
embedding_energy_file b.pt
atomic_e-density_file c.pt
gpu_precision fp64
gpu_lanes 2
gpu_autotune 1
simulation 2
maxsteps 999
""")
    host.init_parameters()
    host.read_parameters(str(p).encode(), 1)
    g = lambda t, n: t.in_dll(host, n).value
    assert g(C.c_int, "ensemble") == 1 and g(C.c_int, "ntypes") == 2
    assert g(C.c_int, "steps_max") == 100          # phase 2 not read
    assert g(C.c_int, "eng_int") == 10
    assert abs(g(C.c_double, "isq_tau_eta") - 1 / 9.8227 ** 2) < 1e-15
    assert g(C.c_int, "gpu_precision") == 0
    assert g(C.c_int, "gpu_lanes") == 2 and g(C.c_int, "gpu_autotune") == 1
    masses = (C.c_double * 8).in_dll(host, "masses")
    assert masses[0] == 63.546 and masses[1] == 58.69
    assert list((C.c_int * 3).in_dll(host, "pbc_dirs")) == [1, 1, 0]
    assert list((C.c_double * 3).in_dll(host, "box_y")) == [0.0, 11.0, 0.0]


def test_config_reader(host, tmp_path):
    s = W.fcc_system((3, 3, 3), 3.615, [63.546, 58.69], W.kT(300), 5,
                     type_fractions=(0.5, 0.5))
    path = str(tmp_path / "c.cfg")
    s.write_config(path)
    host.init_parameters()
    C.c_int.in_dll(host, "ntypes").value = 2
    assert host.read_atoms(path.encode()) == 0
    n = C.c_long.in_dll(host, "natoms").value
    assert n == s.n
    xp = C.POINTER(C.c_double).in_dll(host, "atom_x")
    pp = C.POINTER(C.c_double).in_dll(host, "atom_p")
    tp = C.POINTER(C.c_int).in_dll(host, "atom_type")
    x = np.ctypeslib.as_array(xp, (n, 3))
    p = np.ctypeslib.as_array(pp, (n, 3))
    t = np.ctypeslib.as_array(tp, (n,))
    assert np.array_equal(x, s.x)
    assert np.allclose(p, s.p, rtol=1e-15, atol=1e-18)
    assert np.array_equal(t, s.type)
    assert (C.c_double * 3).in_dll(host, "box_x")[0] == s.box[0, 0]


def test_missing_ensemble_is_fatal(tmp_path):
    p = tmp_path / "bad.param"
    p.write_text("ntypes 1\ntimestep 0.1\ncore_potential_file x.pt\n")
    exe = os.path.join(ROOT, "imd_amd", "host", "imd_gpu")
    if not os.path.exists(exe):
        B.build_host()
    r = subprocess.run([exe, "-p", str(p)], capture_output=True, text=True,
                       timeout=60)
    assert r.returncode == 2
    assert "ensemble" in r.stderr
