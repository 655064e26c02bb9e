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
from imd_amd import imdg as G
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
interpolation spline
outfiles %s
simulation 2
maxsteps 999
""" % str(tmp_path / "run"))
    host.init_parameters()
    assert host.read_parameters(str(p).encode(), 1) == 0  # phase 2 follows
    g = lambda t, n: t.in_dll(host, n).value
    assert g(C.c_int, "ensemble") == 1 and g(C.c_int, "ntypes") == 2
    assert g(C.c_int, "steps_max") == 100          # phase 2 not read
    assert g(C.c_int, "eng_int") == 10
    assert abs(g(C.c_double, "isq_tau_eta") - 1 / 9.8227 ** 2) < 1e-15
    assert g(C.c_int, "gpu_precision") == 0
    assert g(C.c_int, "gpu_lanes") == 2 and g(C.c_int, "gpu_autotune") == 1
    assert g(C.c_int, "interpolation") == G.INTERP_SPLINE
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
    r = subprocess.run([_exe(), "-p", str(p)], capture_output=True, text=True,
                       timeout=60)
    assert r.returncode == 2
    assert "ensemble" in r.stderr


# ---------------------------------------------------------------------------
# checkpoints, iteration files, restart and phases
# ---------------------------------------------------------------------------

def _set_str(host, name, val):
    (C.c_char * 256).in_dll(host, name).value = val.encode()


def _get_str(host, name):
    return (C.c_char * 256).in_dll(host, name).value.decode()


def _param(tmp_path, out, extra=()):
    p = tmp_path / "run.param"
    p.write_text("\n".join([
        "ensemble nvt", "ntypes 1", "masses 63.546", "timestep 0.098227",
        "startstep 0", "maxsteps 100", "starttemp 0.05", "tau_eta 9.8227",
        "core_potential_file a.pt", "embedding_energy_file b.pt",
        "atomic_e-density_file c.pt", "outfiles %s" % out, *extra]) + "\n")
    return str(p).encode()


def _nvt_state(host, out):
    """Globals a checkpoint carries, with values that need all 17 digits."""
    host.init_parameters()
    _set_str(host, "outfilename", out)
    C.c_int.in_dll(host, "ensemble").value = 1
    C.c_double.in_dll(host, "temperature").value = 0.1 / 3
    C.c_double.in_dll(host, "eta").value = -2.0 / 7
    C.c_double.in_dll(host, "isq_tau_eta").value = 1 / 9.8227 ** 2
    box = {"box_x": [289.2 / 3, 0.0, 0.0], "box_y": [1e-3 / 3, 96.4, 0.0],
           "box_z": [0.0, 0.0, 100.0 / 7]}
    for k, v in box.items():
        (C.c_double * 3).in_dll(host, k)[:] = v
    return box


def _box_back(host):
    return {k: list((C.c_double * 3).in_dll(host, k))
            for k in ("box_x", "box_y", "box_z")}


def test_itr_file_format_and_restart_lookup(host, tmp_path, capfd):
    """write_itr_file (src/imd_io_3d.c:949-1095) writes a parameter file with
    startstep = step + 1, the box and, for NVT, starttemp and eta; -r n
    (src/imd_param.c:3907-3941) reads it and takes <out>.%05d.chkpt as the
    configuration."""
    out = str(tmp_path / "run")
    box = _nvt_state(host, out)
    host.write_itr_file.argtypes = [C.c_int, C.c_int, C.c_char_p]
    assert host.write_itr_file(3, 41, b"") == 0
    text = open(out + ".00003.itr").read().splitlines()
    assert text[0] == "# checkpoint 3"
    assert text[1] == "startstep \t42"
    assert [l.split()[0] for l in text[2:]] == ["box_x", "box_y", "box_z",
                                                "starttemp", "eta"]
    open(out + ".00003.chkpt", "w").write("")
    param = _param(tmp_path, out)
    host.init_parameters()
    C.c_int.in_dll(host, "imdrestart").value = 3
    try:
        host.read_parameters(param, 1)
    finally:
        C.c_int.in_dll(host, "imdrestart").value = 0
    g = lambda t, n: t.in_dll(host, n).value
    assert g(C.c_int, "steps_min") == 42 and g(C.c_int, "steps_max") == 100
    assert g(C.c_double, "temperature") == 0.1 / 3
    assert g(C.c_double, "eta") == -2.0 / 7
    # %.16f keeps 16 decimals of the box: within 1e-16 absolute
    for k, v in _box_back(host).items():
        assert np.allclose(v, box[k], rtol=0, atol=1e-15), k
    assert _get_str(host, "infilename") == out + ".00003.chkpt"
    C.CDLL(None).fflush(None)  # the C library's stdout buffer
    assert "Restarting from %s.00003.chkpt." % out in capfd.readouterr().out


def test_restart_prefers_unpadded_names(host, tmp_path):
    out = str(tmp_path / "run")
    _nvt_state(host, out)
    host.write_itr_file.argtypes = [C.c_int, C.c_int, C.c_char_p]
    host.write_itr_file(7, 69, b"")
    os.rename(out + ".00007.itr", out + ".7.itr")
    for name in (".7.chkpt", ".00007.chkpt"):
        open(out + name, "w").write("")
    host.init_parameters()
    C.c_int.in_dll(host, "imdrestart").value = 7
    try:
        host.read_parameters(_param(tmp_path, out), 1)
    finally:
        C.c_int.in_dll(host, "imdrestart").value = 0
    assert C.c_int.in_dll(host, "steps_min").value == 70
    assert _get_str(host, "infilename") == out + ".7.chkpt"


def test_checkpoint_round_trip(host, tmp_path):
    """write_config(n, steps) writes IMD's checkpoint (header
    src/imd_io.c:3152-3300, rows src/imd_io_3d.c:696-948) and its
    iteration file; read_atoms reads the same positions back bit for bit
    and p = (p/m) m to the last bit or so."""
    s = W.fcc_system((3, 3, 3), 3.615, [63.546, 58.69], W.kT(300), 5,
                     type_fractions=(0.5, 0.5), displace=0.05)
    path = str(tmp_path / "c.cfg")
    s.write_config(path)
    out = str(tmp_path / "ck")
    host.init_parameters()
    C.c_int.in_dll(host, "ntypes").value = 2
    C.c_int.in_dll(host, "ensemble").value = 0
    C.c_int.in_dll(host, "have_eam").value = 1
    assert host.read_atoms(path.encode()) == 0
    n = s.n
    ep = np.ctypeslib.as_array(C.POINTER(C.c_double).in_dll(host, "atom_epot"), (n,))
    rh = np.ctypeslib.as_array(C.POINTER(C.c_double).in_dll(host, "atom_rho"), (n,))
    rng = np.random.default_rng(3)
    ep[:] = -3.5 + rng.random(n) / 3
    rh[:] = 0.25 + rng.random(n) / 7
    epot, rho = ep.copy(), rh.copy()
    _set_str(host, "outfilename", out)
    host.write_config.argtypes = [C.c_int, C.c_int]
    assert host.write_config(2, 20) == 0
    lines = open(out + ".00002.chkpt").read().splitlines()
    assert lines[0] == "#F A 1 1 1 3 3 2"
    assert lines[1] == "#C number type mass x y z vx vy vz Epot eam_rho"
    assert lines[2].startswith("#X") and lines[5] == "##PBC 1 1 1"
    end = lines.index("#E")
    rows = np.array([[float(v) for v in l.split()] for l in lines[end + 1:]])
    assert rows.shape == (n, 11)
    assert np.array_equal(rows[:, 0], s.id) and np.array_equal(rows[:, 1], s.type)
    assert np.array_equal(rows[:, 9], epot) and np.array_equal(rows[:, 10], rho)
    assert open(out + ".00002.itr").read().splitlines()[1] == "startstep \t21"
    for name, fz in (("-final", -1), ("-interm", -2)):
        host.write_config(fz, 20)
        assert os.path.exists(out + name + ".chkpt")
        assert os.path.exists(out + name + ".itr")
    # read the checkpoint back
    host.init_parameters()
    C.c_int.in_dll(host, "ntypes").value = 2
    assert host.read_atoms((out + ".00002.chkpt").encode()) == 0
    assert C.c_long.in_dll(host, "natoms").value == n
    assert C.c_int.in_dll(host, "do_maxwell").value == 0  # velocities given
    x = np.ctypeslib.as_array(C.POINTER(C.c_double).in_dll(host, "atom_x"), (n, 3))
    p = np.ctypeslib.as_array(C.POINTER(C.c_double).in_dll(host, "atom_p"), (n, 3))
    assert np.array_equal(x, s.x)
    assert np.allclose(p, s.p, rtol=4e-16, atol=0)
    assert list((C.c_double * 3).in_dll(host, "box_x")) == list(s.box[0])


def test_phases_carry_state(host, tmp_path):
    """read_parameters(file, n > 1) (src/imd_param.c:3880-3903): writes
    <out>-final.itr of the phase that ended, reads the file through phase n
    and then that iteration file; returns 1 at the last phase."""
    out = str(tmp_path / "ph")
    p = tmp_path / "ph.param"
    p.write_text("\n".join([
        "ensemble nvt", "ntypes 1", "masses 63.546", "timestep 0.098227",
        "maxsteps 20", "starttemp 0.05", "tau_eta 9.8227",
        "core_potential_file a.pt", "embedding_energy_file b.pt",
        "atomic_e-density_file c.pt", "outfiles %s" % out,
        "simulation 2", "ensemble nve", "maxsteps 40", "timestep 0.05",
        "simulation 3", "maxsteps 60"]) + "\n")
    g = lambda t, n: t.in_dll(host, n).value
    host.init_parameters()
    assert host.read_parameters(str(p).encode(), 1) == 0
    assert g(C.c_int, "ensemble") == 1 and g(C.c_int, "steps_max") == 20
    C.c_double.in_dll(host, "eta").value = 0.125  # as left by phase 1
    assert host.read_parameters(str(p).encode(), 2) == 0
    assert os.path.exists(out + "-final.itr")
    assert g(C.c_int, "ensemble") == 0
    assert g(C.c_int, "steps_min") == 21 and g(C.c_int, "steps_max") == 40
    assert g(C.c_double, "timestep") == 0.05
    assert host.read_parameters(str(p).encode(), 3) == 1
    assert g(C.c_int, "steps_min") == 41 and g(C.c_int, "steps_max") == 60


def _exe():
    exe = os.path.join(ROOT, "imd_amd", "host", "imd_gpu")
    if not os.path.exists(exe):
        B.build_host()
    return exe


def test_command_line(tmp_path):
    """-x is an illegal option (exit -1); -r n without its iteration file is
    a fatal error, raised before any device work."""
    r = subprocess.run([_exe(), "-x"], capture_output=True, text=True,
                       timeout=60)
    assert r.returncode == 255 and "Illegal option -x" in r.stdout
    out = str(tmp_path / "run")
    param = _param(tmp_path, out).decode()
    for args in (["-r", "7", "-p", param], ["-r7", "-p" + param]):
        r = subprocess.run([_exe(), *args], capture_output=True, text=True,
                           timeout=60, cwd=str(tmp_path))
        assert r.returncode == 2, r.stderr
        assert "run.00007.itr not found" in r.stderr


def test_press_keys_and_file_format(host, tmp_path):
    """press_int turns the stress tensor on (STRESS_TENS); write_config_press
    writes <out>.%05d.press with IMD's header and rows "number type mass
    x y z P_xx P_yy P_zz P_yz P_zx P_xy" (src/imd_io.c:1047-1240)."""
    out = str(tmp_path / "run")
    host.init_parameters()
    host.read_parameters(_param(tmp_path, out, ["press_int 10"]), 1)
    assert C.c_int.in_dll(host, "press_int").value == 10
    assert C.c_int.in_dll(host, "stress_tens").value == 1
    s = W.fcc_system((2, 2, 2), 3.615, [63.546], W.kT(300), 5, displace=0.05)
    path = str(tmp_path / "c.cfg")
    s.write_config(path)
    C.c_int.in_dll(host, "ntypes").value = 1
    assert host.read_atoms(path.encode()) == 0
    n = s.n
    ap = np.ctypeslib.as_array(C.POINTER(C.c_double).in_dll(host, "atom_press"),
                               (n, 6))
    ap[:] = np.random.default_rng(4).normal(size=(n, 6)) / 3
    want = ap.copy()
    _set_str(host, "outfilename", out)
    host.write_config_press.argtypes = [C.c_int]
    assert host.write_config_press(3) == 0
    lines = open(out + ".00003.press").read().splitlines()
    assert lines[0] == "#F A 1 1 1 3 0 6"
    assert lines[1] == "#C number type mass x y z P_xx P_yy P_zz P_yz P_zx P_xy"
    end = lines.index("#E")
    rows = np.array([[float(v) for v in l.split()] for l in lines[end + 1:]])
    assert rows.shape == (n, 12)
    assert np.array_equal(rows[:, 3:6], s.x)
    assert np.array_equal(rows[:, 6:], want)


def _polar(rand, sigma):
    """Knuth's polar method (vol. 2, 3rd ed., p. 122) on a uniform stream."""
    while True:
        x = -1 + 2 * rand()
        y = -1 + 2 * rand()
        r2 = x * x + y * y
        if not (r2 > 1.0 or r2 == 0):
            return sigma * y * np.sqrt(-2.0 * np.log(r2) / r2)


def test_maxwell_cell_order_and_stream(host, tmp_path):
    """maxwell (src/imd_maxwell.c:44-323) draws three normal variates per
    atom in IMD's cell traversal order from the drand48 stream that
    read_atoms seeded (src/imd_io_3d.c:87), then removes the mean momentum.
    The stream is libc's own drand48, the traversal order is the oracle's
    (its binning is pinned by the grid tests), and the atoms are written
    shuffled so that input order and cell order differ."""
    eam = W.johnson_eam((W.CU, W.NI), rc=5.0, npts=400)
    s = W.fcc_system((4, 5, 4), 3.615, eam.masses, W.kT(300.0), 13,
                     type_fractions=(0.6, 0.4), displace=0.05)
    rng = np.random.default_rng(3)
    perm = rng.permutation(s.n)
    pos = s.x[perm].copy()
    pos[::7, 0] += s.box[0, 0]  # outside the box: back_into_box before binning
    types, masses = s.type[perm], s.mass[perm]
    path = str(tmp_path / "nov.conf")
    with open(path, "w") as f:
        f.write("#F A 1 1 1 3 0 0\n#C number type mass x y z\n")
        for d, v in zip("XYZ", s.box):
            f.write("#%s %.17g %.17g %.17g\n" % (d, *v))
        f.write("#E\n")
        for k in range(s.n):
            f.write("%d %d %.17g %.17g %.17g %.17g\n" % (
                k, types[k], masses[k], *pos[k]))
    host.init_parameters()
    C.c_int.in_dll(host, "ntypes").value = 2
    C.c_long.in_dll(host, "seed").value = -77
    C.c_double.in_dll(host, "nbl_margin").value = 0.4
    rc_list = max(eam.pair.end.max(), eam.rho.end.max())
    C.c_double.in_dll(host, "cellsz").value = rc_list
    assert host.read_atoms(path.encode()) == 0
    assert C.c_int.in_dll(host, "do_maxwell").value == 1
    T = W.kT(300.0)
    host.maxwell.argtypes = [C.c_double]
    host.maxwell(T)
    n = s.n
    p = np.ctypeslib.as_array(C.POINTER(C.c_double).in_dll(host, "atom_p"),
                              (n, 3)).copy()
    # the cell of each atom as IMD bins it at read time (oracle), then the
    # inner cells x-major, y, z, and input order within a cell
    sysobj = W.System(x=pos, p=np.zeros_like(pos), type=types.astype(np.int32),
                      mass=masses, box=s.box, pbc=(1, 1, 1),
                      id=np.arange(n, dtype=np.int64))
    pair = O.OracleTable.from_table(eam.pair)
    rho = O.OracleTable.from_table(eam.rho)
    emb = O.OracleTable.from_table(eam.embed, radial=False)
    orc = O.OracleSystem(sysobj, pair, rho, emb, nbl_margin=0.4, ntypes=2)
    orc.calc_forces()
    g = orc.grid()
    c = orc.cells()
    key = (c[:, 0] * g[1] + c[:, 1]) * g[2] + c[:, 2]
    order = np.argsort(key, kind="stable")
    assert not np.array_equal(order, np.arange(n))
    libc = C.CDLL(None)
    libc.drand48.restype = C.c_double
    libc.srand48.argtypes = [C.c_long]
    libc.srand48(-77)
    want = np.zeros((n, 3))
    for i in order:
        sig = np.sqrt(T * masses[i])
        for d in range(3):
            want[i, d] = _polar(libc.drand48, sig)
    tot = np.zeros(3)
    for i in order:
        tot += want[i]
    want -= tot / n
    assert np.abs(p - want).max() <= 1e-15 * np.abs(want).max()
    assert np.abs(p.sum(axis=0)).max() < 1e-12 * np.abs(p).sum()
    # the variance is kT m per component (n = 320 atoms: 3 sigma bands)
    z = p / np.sqrt(T * masses)[:, None]
    assert abs(z.var() - 1.0) < 3 * np.sqrt(2.0 / z.size)
