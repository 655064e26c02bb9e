"""The IMD-style C host on the GPU: `imd_gpu -p file.param` reads IMD's
parameter keys, format-2 tables and an ASCII configuration, runs the fused
main loop through the C ABI and writes the `.eng` file; its rows must match
the oracle's thermo at the same steps (FP64 path).  Checkpoints, `-r n`
restarts, phases and the stop/write files must continue the same
trajectory.  The executable runs as a child process."""
import dataclasses
import os
import shutil
import subprocess

import numpy as np
import pytest

from imd_amd import build as B
from imd_amd import workloads as W
from helpers import eam_tables, oracle_eam

pytestmark = pytest.mark.gpu

EXE = os.path.join(B.HOST, "imd_gpu")


def _read_eng(path):
    rows = [l.split() for l in open(path) if l.strip() and not l.startswith("#")]
    # an NVT phase adds the eta*tau_eta column: keep the common columns
    w = min(len(r) for r in rows)
    return np.array([[float(v) for v in r[:w]] for r in rows])


@pytest.mark.parametrize("ensemble,extra", [
    ("nve", []), ("nvt", []),
    # launch shape from the parameter file and the autotuner before the
    # main loop: same trajectory
    ("nve", ["gpu_lanes 2", "gpu_autotune 1"])])
def test_imd_gpu_eng_matches_oracle(tmp_path, ensemble, extra):
    if not os.path.exists(EXE):
        pytest.fail("imd_gpu is not built (build() first)")
    eam = eam_tables()
    s = W.eam_crystal(n_cells=5, temp_k=700.0, seed=21, displace=0.03)
    # the host reads v = p/m from the file and forms p = v m: give the
    # oracle the same momenta
    s = dataclasses.replace(s, p=(s.p / s.mass[:, None]) * s.mass[:, None])
    files = eam.write(str(tmp_path / "pot"))
    cfg = s.write_config(str(tmp_path / "start.chkpt"))
    nsteps, eng_int = 40, 10
    kT = W.kT(700.0)
    tau = 50 * W.FS
    lines = [
        "ensemble %s" % ensemble,
        "ntypes 1",
        "masses %.17g" % s.mass[0],
        "timestep %.17g" % W.FS,
        "startstep 0",
        "maxsteps %d" % nsteps,
        "eng_int %d" % eng_int,
        "coordname %s" % cfg,
        "outfiles %s" % str(tmp_path / "run"),
        "core_potential_file %s" % files["core_potential_file"],
        "embedding_energy_file %s" % files["embedding_energy_file"],
        "atomic_e-density_file %s" % files["atomic_e-density_file"],
        "nbl_margin 0.4",
        "gpu_precision fp64",
        "box_from_header 1",
    ]
    if ensemble == "nvt":
        lines += ["starttemp %.17g" % kT, "tau_eta %.17g" % tau]
    lines += extra
    param = tmp_path / "run.param"
    param.write_text("\n".join(lines) + "\n")
    r = subprocess.run([EXE, "-p", str(param)], capture_output=True,
                       text=True, timeout=120, cwd=str(tmp_path))
    assert r.returncode == 0, r.stderr[-2000:]
    if "gpu_autotune 1" in extra:
        assert "gpu_autotune: density" in r.stdout
    eng = _read_eng(str(tmp_path / "run.eng"))
    assert eng.shape[0] == nsteps // eng_int + 1

    import oracle as O
    kw = {}
    if ensemble == "nvt":
        kw = dict(ensemble=O.ORC_NVT, temperature=kT, eta=0.0,
                  isq_tau_eta=1.0 / tau ** 2)
    orc = oracle_eam(s, eam, nbl_margin=0.4, timestep=W.FS, **kw)
    th = orc.run(nsteps + 1)  # steps 0..nsteps
    n = s.n
    for row in eng:
        step = int(round(row[0] / W.FS))
        epot = th[step, 0] / n
        temp = 2.0 * th[step, 1] / (3 * n)
        assert abs(row[1] - epot) <= 1e-9 * abs(epot), (step, row[1], epot)
        # the temperature column is printed with 6 significant digits
        assert abs(row[2] - temp) <= 1e-5 * abs(temp), (step, row[2], temp)
    # the phase ends with the iteration file of its last step
    itr = open(str(tmp_path / "run-final.itr")).read().splitlines()
    assert itr[1] == "startstep \t%d" % (nsteps + 1)


def _setup(tmp_path, temp_k=700.0):
    eam = eam_tables()
    s = W.eam_crystal(n_cells=5, temp_k=temp_k, seed=21, displace=0.03)
    s = dataclasses.replace(s, p=(s.p / s.mass[:, None]) * s.mass[:, None])
    files = eam.write(str(tmp_path / "pot"))
    cfg = s.write_config(str(tmp_path / "start.chkpt"))
    return eam, s, files, cfg


def _lines(s, files, cfg, out, ensemble, maxsteps, extra=()):
    lines = [
        "ensemble %s" % ensemble, "ntypes 1", "masses %.17g" % s.mass[0],
        "timestep %.17g" % W.FS, "startstep 0", "maxsteps %d" % maxsteps,
        "eng_int 10", "coordname %s" % cfg, "outfiles %s" % out,
        "core_potential_file %s" % files["core_potential_file"],
        "embedding_energy_file %s" % files["embedding_energy_file"],
        "atomic_e-density_file %s" % files["atomic_e-density_file"],
        "nbl_margin 0.4", "gpu_precision fp64", "box_from_header 1"]
    if ensemble == "nvt":
        lines += ["starttemp %.17g" % W.kT(700.0), "tau_eta %.17g" % (50 * W.FS)]
    return lines + list(extra)


def _run(dirpath, lines, args=()):
    param = os.path.join(str(dirpath), "run.param")
    open(param, "w").write("\n".join(lines) + "\n")
    r = subprocess.run([EXE, *args, "-p", param], capture_output=True,
                       text=True, timeout=120, cwd=str(dirpath))
    assert r.returncode == 0, r.stderr[-2000:]
    return r


def _rows(path):
    """number type mass x y z vx vy vz Epot eam_rho rows of a checkpoint"""
    lines = open(path).read().splitlines()
    end = lines.index("#E")
    return np.array([[float(v) for v in l.split()] for l in lines[end + 1:]])


def test_checkpoint_restart_continues_trajectory(tmp_path):
    """A 40-step NVT run with checkpt_int 20 writes run.00001/2.chkpt and
    .itr; `imd_gpu -r 1` in a fresh directory with the same parameter file
    continues from step 21 (startstep, box, starttemp and eta from the
    iteration file; x and v from the checkpoint) and must reach the same
    state and .eng rows as the uninterrupted run and the oracle."""
    eam, s, files, cfg = _setup(tmp_path)
    a = tmp_path / "a"
    b = tmp_path / "b"
    a.mkdir()
    b.mkdir()
    lines = _lines(s, files, cfg, "run", "nvt", 40, ["checkpt_int 20"])
    _run(a, lines)
    for k in (1, 2):
        for ext in ("chkpt", "itr"):
            assert os.path.exists(str(a / ("run.%05d.%s" % (k, ext))))
    itr = open(str(a / "run.00001.itr")).read().splitlines()
    assert itr[0] == "# checkpoint 1" and itr[1] == "startstep \t21"
    assert [l.split()[0] for l in itr[5:]] == ["starttemp", "eta"]
    for ext in ("chkpt", "itr"):
        shutil.copy(str(a / ("run.00001.%s" % ext)), str(b))
    r = _run(b, lines, ["-r", "1"])
    assert "Restarting from run.00001.chkpt." in r.stdout
    # the restart appends to .eng without a header: rows 30 and 40
    eng_a = _read_eng(str(a / "run.eng"))
    eng_b = _read_eng(str(b / "run.eng"))
    assert eng_a.shape[0] == 5 and eng_b.shape[0] == 2
    assert np.allclose(eng_b[:, 0], eng_a[3:, 0], rtol=1e-12)
    assert np.allclose(eng_b[:, 1], eng_a[3:, 1], rtol=1e-11, atol=0)
    assert np.allclose(eng_b[:, 5], eng_a[3:, 5], rtol=1e-5)  # eta*tau, %e
    ca = _rows(str(a / "run.00002.chkpt"))
    cb = _rows(str(b / "run.00002.chkpt"))
    assert np.array_equal(ca[:, :3], cb[:, :3])
    # positions are wrapped at list rebuilds, which the two runs do at
    # different steps: compare them modulo the (orthogonal) box
    L = np.diag(s.box)
    d = ca[:, 3:6] - cb[:, 3:6]
    d -= L * np.round(d / L)
    assert np.max(np.abs(d)) < 1e-10
    vmax = np.max(np.abs(ca[:, 6:9]))
    assert np.max(np.abs(ca[:, 6:9] - cb[:, 6:9])) < 1e-10 * vmax
    assert np.allclose(cb[:, 9:], ca[:, 9:], rtol=1e-10, atol=1e-12)
    # and the oracle's uninterrupted run
    import oracle as O
    kT, tau = W.kT(700.0), 50 * W.FS
    orc = oracle_eam(s, eam, nbl_margin=0.4, timestep=W.FS,
                     ensemble=O.ORC_NVT, temperature=kT, eta=0.0,
                     isq_tau_eta=1.0 / tau ** 2)
    th = orc.run(41)
    for row in eng_b:
        step = int(round(row[0] / W.FS))
        ep = th[step, 0] / s.n
        assert abs(row[1] - ep) <= 1e-9 * abs(ep), (step, row[1], ep)
    # per-atom Epot and eam_rho of the checkpoint (input order): the
    # oracle's from the force evaluation of step 40
    o = orc.atoms()
    assert np.allclose(cb[:, 9], o["epot"], rtol=1e-9, atol=1e-12)
    assert np.allclose(cb[:, 10], o["rho"], rtol=1e-9, atol=1e-12)


def test_stop_and_write_files(tmp_path):
    """An empty "write" file found at a watch_int step writes
    run-interm.chkpt and is removed; a "stop" file found at a stop_int step
    writes run-final.chkpt / .itr and ends the run
    (src/imd_main_3d.c:1892-1940).  Step 0 is such a step, so files present
    at the start are found after step 0."""
    eam, s, files, cfg = _setup(tmp_path)
    open(str(tmp_path / "write"), "w").close()
    r = _run(tmp_path, _lines(s, files, cfg, "run", "nve", 40,
                              ["watch_int 10"]))
    assert "Write file found after 0 steps" in r.stdout
    assert not os.path.exists(str(tmp_path / "write"))
    assert os.path.exists(str(tmp_path / "run-interm.chkpt"))
    assert _read_eng(str(tmp_path / "run.eng")).shape[0] == 5
    open(str(tmp_path / "stop"), "w").close()
    r = _run(tmp_path, _lines(s, files, cfg, "run", "nve", 40,
                              ["stop_int 20"]))
    assert "Stop file found after 0 steps" in r.stdout
    assert "Did 1 steps" in r.stdout
    assert not os.path.exists(str(tmp_path / "stop"))
    assert _read_eng(str(tmp_path / "run.eng")).shape[0] == 1
    itr = open(str(tmp_path / "run-final.itr")).read().splitlines()
    assert itr[1] == "startstep \t1"
    c = _rows(str(tmp_path / "run-final.chkpt"))
    assert c.shape == (s.n, 11)
    # maxwalltime is tested after every fused run
    r = _run(tmp_path, _lines(s, files, cfg, "run", "nve", 40,
                              ["maxwalltime 1e-9"]))
    assert "Maximal allowed walltime reached after 0 steps" in r.stdout


def test_phases_nvt_then_nve(tmp_path):
    """Phase 1 NVT through step 20, phase 2 NVE through step 40: the state
    (x, p) carries over, startstep comes from run-final.itr, and the .eng
    rows of phase 2 match an oracle that switches at the same step."""
    eam, s, files, cfg = _setup(tmp_path)
    lines = _lines(s, files, cfg, "run", "nvt", 20) + [
        "simulation 2", "ensemble nve", "maxsteps 40"]
    r = _run(tmp_path, lines)
    assert "Starting simulation 2" in r.stdout
    eng = _read_eng(str(tmp_path / "run.eng"))
    assert eng.shape[0] == 5
    import oracle as O
    kT, tau = W.kT(700.0), 50 * W.FS
    o1 = oracle_eam(s, eam, nbl_margin=0.4, timestep=W.FS,
                    ensemble=O.ORC_NVT, temperature=kT, eta=0.0,
                    isq_tau_eta=1.0 / tau ** 2)
    th1 = o1.run(21)
    st = o1.atoms()
    s2 = dataclasses.replace(s, x=st["x"], p=st["p"])
    o2 = oracle_eam(s2, eam, nbl_margin=0.4, timestep=W.FS)
    th2 = o2.run(20)  # steps 21..40
    for row in eng:
        step = int(round(row[0] / W.FS))
        ep = (th1[step, 0] if step <= 20 else th2[step - 21, 0]) / s.n
        assert abs(row[1] - ep) <= 1e-9 * abs(ep), (step, row[1], ep)


def test_stress_outputs_match_oracle(tmp_path):
    """press_int 10: the .eng rows carry tot_presstens / volume (6 columns,
    IMD's STRESS_TENS layout) and run.%05d.press the per-atom tensors, both
    matching the oracle's PRESSTENS at the same steps (FP64)."""
    eam, s, files, cfg = _setup(tmp_path)
    _run(tmp_path, _lines(s, files, cfg, "run", "nve", 20, ["press_int 10"]))
    hdr = open(str(tmp_path / "run.eng")).readline().split()
    assert hdr[-6:] == ["Press_xx", "Press_yy", "Press_zz", "Press_yz",
                        "Press_xz", "Press_xy"]
    eng = _read_eng(str(tmp_path / "run.eng"))
    assert eng.shape == (3, 11)
    orc = oracle_eam(s, eam, nbl_margin=0.4, timestep=W.FS)
    orc.set_press(True)
    vol = abs(np.linalg.det(s.box))
    sums = {}
    for step in range(21):
        orc.calc_forces(step)
        orc.move_atoms()
        orc.check_nblist()
        if step % 10 == 0:
            sums[step] = orc.press().sum(0) / vol
    scale = max(np.abs(v).max() for v in sums.values())
    for row in eng:
        step = int(round(row[0] / W.FS))
        # printed with %e: 7 significant digits
        assert np.all(np.abs(row[5:] - sums[step]) <= 2e-6 * scale), step
    for k in range(3):
        assert os.path.exists(str(tmp_path / ("run.%05d.press" % k)))
    rows = _rows(str(tmp_path / "run.00002.press"))
    ref = orc.press()
    assert rows.shape == (s.n, 12)
    assert np.abs(rows[:, 6:] - ref).max() <= 1e-9 * np.abs(ref).max()


def test_eeam_run_matches_oracle(tmp_path):
    """eeam_energy_file (EEAM): the .eng rows follow the oracle with the
    same E(P) table, and checkpoints carry the eam_p column."""
    from helpers import emod_table
    import oracle as O
    eam, s, files, cfg = _setup(tmp_path)
    et = emod_table(1)
    epath = str(tmp_path / "emod.pt")
    et.write(epath)
    _run(tmp_path, _lines(s, files, cfg, "run", "nve", 20,
                          ["eeam_energy_file %s" % epath, "checkpt_int 20"]))
    eng = _read_eng(str(tmp_path / "run.eng"))
    orc = oracle_eam(s, eam, nbl_margin=0.4, timestep=W.FS)
    emod = O.OracleTable.from_table(et)
    orc.set_eeam(emod)
    th = orc.run(21)
    for row in eng:
        step = int(round(row[0] / W.FS))
        ep = th[step, 0] / s.n
        assert abs(row[1] - ep) <= 1e-9 * abs(ep), (step, row[1], ep)
    lines = open(str(tmp_path / "run.00001.chkpt")).read().splitlines()
    assert lines[0] == "#F A 1 1 1 3 3 3"
    assert lines[1].endswith("Epot eam_rho eam_p")
    rows = _rows(str(tmp_path / "run.00001.chkpt"))
    P, _ = orc.eeam()
    assert np.allclose(rows[:, 11], P, rtol=1e-9)


def test_imd_gpu_maxwell_start_matches_oracle(tmp_path):
    """A configuration without velocities: imd_gpu draws Maxwell momenta
    (src/imd_maxwell.c:44-301: three variates per atom from the drand48
    stream seeded at read time, cells in IMD's traversal order, atoms of a
    cell in input order, then the mean momentum removed) and runs 30 NVE
    steps; the oracle starts from the same draw, made here with libc's
    drand48, and its thermo must match the .eng rows."""
    import ctypes as C
    from test_host import _polar
    eam = eam_tables()
    s = W.eam_crystal(n_cells=5, temp_k=0.0, seed=27, displace=0.04)
    perm = np.random.default_rng(4).permutation(s.n)
    x = s.x[perm].copy()
    files = eam.write(str(tmp_path / "pot"))
    cfg = str(tmp_path / "nov.conf")
    with open(cfg, "w") as f:
        f.write("#F A 1 1 1 3 0 0\n#C number type mass x y z\n")
        for d, v in zip("XYZ", s.box):
            f.write("#%s %.17g %.17g %.17g\n" % (d, *v))
        f.write("#E\n")
        for k in range(s.n):
            f.write("%d 0 %.17g %.17g %.17g %.17g\n" % (k, s.mass[0], *x[k]))
    kT = W.kT(500.0)
    nsteps = 30
    lines = _lines(s, files, cfg, str(tmp_path / "run"), "nve", nsteps,
                   extra=["starttemp %.17g" % kT, "seed -4242"])
    _run(tmp_path, lines)
    eng = _read_eng(str(tmp_path / "run.eng"))
    # the same draw on the CPU: the oracle's cells (pinned against the
    # grid closed forms) give the traversal order
    sysobj = W.System(x=x, p=np.zeros_like(x), type=np.zeros(s.n, np.int32),
                      mass=s.mass.copy(), box=s.box, pbc=(1, 1, 1),
                      id=np.arange(s.n, dtype=np.int64))
    probe = oracle_eam(sysobj, eam, nbl_margin=0.4, timestep=W.FS)
    probe.calc_forces()
    g, c = probe.grid(), probe.cells()
    order = np.argsort((c[:, 0] * g[1] + c[:, 1]) * g[2] + c[:, 2],
                       kind="stable")
    libc = C.CDLL(None)
    libc.drand48.restype = C.c_double
    libc.srand48.argtypes = [C.c_long]
    libc.srand48(-4242)
    p = np.zeros((s.n, 3))
    for i in order:
        sig = np.sqrt(kT * s.mass[i])
        for d in range(3):
            p[i, d] = _polar(libc.drand48, sig)
    tot = np.zeros(3)
    for i in order:
        tot += p[i]
    p -= tot / s.n
    sysobj.p = p
    orc = oracle_eam(sysobj, eam, nbl_margin=0.4, timestep=W.FS)
    th = orc.run(nsteps + 1)
    assert eng.shape[0] == nsteps // 10 + 1
    for row in eng:
        step = int(round(row[0] / W.FS))
        epot = th[step, 0] / s.n
        temp = 2.0 * th[step, 1] / (3 * s.n)
        assert abs(row[1] - epot) <= 1e-9 * abs(epot), (step, row[1], epot)
        assert abs(row[2] - temp) <= 1e-5 * abs(temp), (step, row[2], temp)
    # about the requested temperature at the start (320 atoms)
    assert abs(2.0 * th[0, 1] / (3 * s.n) / kT - 1.0) < 0.25


def test_imd_gpu_short_distance_warning_and_summary(tmp_path):
    """A pair closer than the tables' begin: imd_gpu prints IMD's
    'Short distance' warnings (src/imd_forces_nbl.c:982, 1323) and the
    closing summary lines of src/imd.c:318-364."""
    eam, s, files, cfg = _setup(tmp_path)
    L = np.diag(s.box)
    d = s.x - s.x[0]
    d -= L * np.round(d / L)
    r = np.sqrt(np.sum(d * d, axis=1))
    r[0] = np.inf
    j = int(np.argmin(r))
    x = s.x.copy()
    x[j] = s.x[0] + d[j] * (0.95 / r[j])
    x[j] -= L * np.floor(x[j] / L)
    s.x = x
    cfg = s.write_config(str(tmp_path / "short.chkpt"))
    lines = _lines(s, files, cfg, str(tmp_path / "run"), "nve", 20)
    r = _run(tmp_path, lines)
    assert "Short distance, pair, step" in r.stderr
    assert "Short distance, EAM, step" in r.stderr
    assert "Neighbor list update every" in r.stdout
    assert "Did 21 steps with %d atoms." % s.n in r.stdout
    assert "seconds per step and atom" in r.stdout
