#!/usr/bin/env python3
"""Throughput of the neighbour-list EAM path on one MI355X.

Workload (default): 2,048,000-atom fcc EAM crystal (80^3 cells, a = 3.615 A),
Johnson-form EAM tables sampled into IMD format-2 tables (5,000 points,
r_c = 5.0 A), Nose-Hoover NVT at 800 K, dt = 1 fs, tau_eta = 100 fs,
skin 0.4 A, displacements N(0, 0.02 A), Maxwell velocities, seed 1002.
A step is one main_loop iteration: list rebuild when the skin test asks for
one, host density + embedding, pair + EAM forces, leapfrog, skin test and
the compensated energy/virial reductions.  Inputs are resident in HBM before
the timed region starts.

Prints one JSON line (the bench contract).
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

from imd_amd import workloads as W  # noqa: E402

HBM_PEAK_GBS = 8000.0
METRIC = "atom-steps/s, 2M-atom EAM: mixed FP32/FP64 vs plain FP64 on 1 GPU"


def algorithmic_bytes(kind: str, k: float, mixed: bool) -> float:
    """Algorithmic bytes per atom of one launch: every per-atom array the
    kernel must touch once, the int32 full-list entries once (k = mean list
    length); neighbour record gathers count as on-chip reuse.

    mixed: density 4k+24 (16 B FP32 xyz+type in, 4 B 2F' out, 4 B energy
      out); force 4k+36 (16 B record, 4 B 2F', 12 B FP32 force, 4 B
      energy); integrate + skin test 152 (12 B force, 48 B x read/write,
      48 B p read/write, 4 B type, 16 B FP32 record out, 24 B x at the last
      build).
    FP64: density 4k+48, force 4k+72, integrate + skin test 148.
    "fused": the force kernel of a timed step, which integrates in the same
    launch and keeps the forces in registers (per-atom forces and energies
    are stored on a run's last step only): 4k+20 of the force part plus the
    integration without its force read, 4k+160 (mixed) | 4k+164 (FP64).
    "step": one whole step as the metric defines it, density + force +
    integrate: 8k+212 (mixed) | 8k+268 (FP64).
    """
    if kind == "density":
        return 4 * k + (24 if mixed else 48)
    if kind == "force":
        return 4 * k + (36 + 152 if mixed else 72 + 148)
    if kind == "fused":
        return 4 * k + (20 + 140 if mixed else 40 + 124)
    if kind == "step":
        return 8 * k + (212 if mixed else 268)
    raise ValueError(kind)


def make_workload(cells: int, seed: int, temp_k: float):
    eam = W.johnson_eam()
    s = W.eam_crystal(n_cells=cells, temp_k=temp_k, seed=seed, displace=0.02)
    return eam, s


def cpu_threads() -> int:
    """Host cores for the CPU baseline: the process's CPU share, at most 16
    (one GPU's share of a host)."""
    try:
        n = len(os.sched_getaffinity(0))
    except AttributeError:
        n = os.cpu_count() or 1
    env = os.environ.get("OMP_NUM_THREADS")
    if env and env.isdigit() and int(env) > 0:
        n = min(n, int(env))
    return max(1, min(n, 16))


def cpu_baseline(eam, s, steps: int, tau_fs: float, temp_k: float,
                 threads: int):
    """The oracle (C restatement of IMD's NBL/EAM path, FP64, half list) on
    the same workload, timed over a bounded number of steps after its first
    (list-building) step, on `threads` host cores (OpenMP by atom, per-thread
    scatter accumulators; 1 = IMD's serial loop)."""
    import oracle as O
    pair = O.OracleTable.from_table(eam.pair)
    rho = O.OracleTable.from_table(eam.rho)
    emb = O.OracleTable.from_table(eam.embed, radial=False)
    orc = O.OracleSystem(s, pair, rho, emb, nbl_margin=0.4, timestep=W.FS,
                         ensemble=O.ORC_NVT, temperature=W.kT(temp_k),
                         eta=0.0, isq_tau_eta=1.0 / (tau_fs * W.FS) ** 2)
    orc.set_threads(threads)
    orc.run(1)
    t0 = time.perf_counter()
    orc.run(steps)
    dt = time.perf_counter() - t0
    return dict(value=s.n * steps / dt, unit="atom-steps/s", cores=threads,
                kind="port",
                sample="%d NVT steps of the same %d-atom workload after one "
                       "warm-up step, oracle (FP64 C, half list), %d "
                       "thread(s)" % (steps, s.n, threads))


def copy_probe(mib: int) -> int:
    """One device-to-device copy of a known size through the HIP runtime
    (a plain copy kernel): the calibration run for the HBM traffic counters
    (FETCH_SIZE / WRITE_SIZE of the copy against its known byte count)."""
    import ctypes as C
    hip = C.CDLL("libamdhip64.so")
    n = mib << 20
    a, b = C.c_void_p(), C.c_void_p()
    for ptr in (a, b):
        if hip.hipMalloc(C.byref(ptr), C.c_size_t(n)) != 0:
            raise RuntimeError("hipMalloc failed")
    hip.hipMemset(a, 1, C.c_size_t(n))
    hip.hipDeviceSynchronize()
    rc = hip.hipMemcpy(b, a, C.c_size_t(n), 3)  # device to device
    hip.hipDeviceSynchronize()
    hip.hipFree(a)
    hip.hipFree(b)
    print(json.dumps({"copy_probe_bytes": n, "rc": rc}))
    return rc


def _pmc_pass(counter: str, outdir: str, args_list, timeout: int = 240):
    """One rocprofv3 counter pass over a child bench process; returns
    {kernel name: mean counter value per dispatch} or None."""
    import csv
    import shutil
    import subprocess
    exe = shutil.which("rocprofv3")
    if exe is None:
        return None
    cmd = [exe, "--pmc", counter, "--output-format", "csv", "-d", outdir,
           "-o", "run", "--", sys.executable, "-u",
           os.path.join(ROOT, "bench.py")] + args_list
    try:
        subprocess.run(cmd, stdout=subprocess.DEVNULL, stderr=subprocess.DEVNULL,
                       timeout=timeout, check=False)
    except subprocess.TimeoutExpired:
        return None
    path = os.path.join(outdir, "run_counter_collection.csv")
    if not os.path.exists(path):
        return None
    acc = {}
    for row in csv.DictReader(open(path)):
        if row["Counter_Name"] != counter:
            continue
        k = row["Kernel_Name"]
        acc.setdefault(k, []).append(float(row["Counter_Value"]))
    return {k: sum(v) / len(v) for k, v in acc.items()}


def measure_traffic(args, kernel_key: str = "k_force"):
    """HBM bytes per launch of the force kernel from FETCH_SIZE and
    WRITE_SIZE (KiB per dispatch, one rocprofv3 pass each), each scaled by
    what the counter reads for a 256 MiB device-to-device copy."""
    import tempfile
    base = tempfile.mkdtemp(prefix="imdg_pmc_")
    run = ["--steps", "5", "--warmup", "2", "--no-cpu-baseline",
           "--no-profile", "--no-traffic", "--precision", args.precision,
           "--cells", str(args.cells), "--seed", str(args.seed)]
    probe = ["--copy-probe", "256"]
    res = {}
    for counter in ("FETCH_SIZE", "WRITE_SIZE"):
        k = _pmc_pass(counter, os.path.join(base, counter), run)
        c = _pmc_pass(counter, os.path.join(base, "copy_" + counter), probe,
                      timeout=120)
        if not k or not c:
            return None
        kv = [v for name, v in k.items() if kernel_key in name]
        cv = [v for name, v in c.items() if "copyBuffer" in name]
        if not kv or not cv or cv[0] <= 0:
            return None
        res[counter] = kv[0] * (256 * 1024 / cv[0]) * 1024  # bytes
    return dict(bytes=res["FETCH_SIZE"] + res["WRITE_SIZE"],
                fetch_bytes=res["FETCH_SIZE"], write_bytes=res["WRITE_SIZE"])


DRIFT_CONFIGS = {
    "eam256k": "fcc EAM crystal 40^3 cells (256,000 atoms), 600 K, dt 1 fs, "
               "skin 0.4 A, NVE",
    "eam32k": "fcc EAM crystal 20^3 cells (32,000 atoms, a subset of the "
              "256,000-atom workload), 600 K, dt 1 fs, skin 0.4 A, NVE",
    "lj4000": "fcc LJ crystal 10^3 cells (4,000 atoms), kT 0.7, dt 0.002, "
              "skin 0.3, NVE",
}


def drift_run(args):
    """NVE energy drift of one input in each precision over args.drift
    steps.  E_tot per atom = (Epot + Ekin) / N every 10 steps; drift =
    least-squares slope x 1000 steps / |<E_tot>|.  The 32,000- and
    4,000-atom inputs are the ones tests/golden/drift_oracle.py runs through
    the oracle (the reference algorithm); its drift is printed beside."""
    from imd_amd import imdg as G
    name = args.drift_config
    if name == "lj4000":
        s = W.lj_crystal()
        make = lambda prec: G.pair_context(W.lj_table(), [s], precision=prec,
                                           ensemble=G.NVE, timestep=0.002,
                                           nbl_margin=0.3)
    else:
        eam = W.johnson_eam()
        s = W.eam_crystal(n_cells=40 if name == "eam256k" else 20,
                          temp_k=600.0, seed=1001, displace=0.02)
        make = lambda prec: G.eam_context(eam, [s], precision=prec,
                                          ensemble=G.NVE, timestep=W.FS,
                                          nbl_margin=0.4)
    out = {"metric": "NVE total-energy drift per 1000 steps (relative)",
           "config": {"workload": DRIFT_CONFIGS[name], "name": name,
                      "steps": args.drift, "sample_every": 10}}
    for prec, label in ((G.MIXED, "mixed"), (G.FP64, "fp64")):
        ctx = make(prec)
        e, done, t0 = [], 0, time.perf_counter()
        while done < args.drift:
            n = min(1000, args.drift - done)
            ctx.run(n)
            h = ctx.thermo_history(done, n)[:, 0, :]
            for k in range(0, n, 10):
                e.append((done + k, (h[k, 1] + h[k, 2]) / s.n))
            done += n
        el = time.perf_counter() - t0
        ctx.close()
        t = np.array([v[0] for v in e], dtype=np.float64)
        y = np.array([v[1] for v in e], dtype=np.float64)
        slope, icept = np.polyfit(t, y, 1)
        resid = y - (slope * t + icept)
        out[label] = {"drift_per_1000_steps": float(slope * 1000 / abs(y.mean())),
                      "etot_per_atom_mean": float(y.mean()),
                      "etot_per_atom_span": float(y.max() - y.min()),
                      "etot_resid_rms": float(np.sqrt(np.mean(resid ** 2))),
                      "atom_steps_per_s": s.n * args.drift / el}
    ref = os.path.join(ROOT, "tests", "golden", "drift_oracle.json")
    if name != "eam256k" and os.path.exists(ref):
        out["oracle"] = json.load(open(ref)).get(name)
    print(json.dumps(out))
    return 0


def state_errors(G, eam, s):
    """Mixed against FP64 on one state, each quantity from the GPU: forces,
    rho, 2F'(rho) and the per-atom energy from one force evaluation with the
    full tables; F(rho) per atom from an evaluation without the pair table
    (the per-atom energy is then the embedding term alone); the pressure per
    atom (trace of the per-atom stress tensor / 3, energy units) and the
    total pressure P = sum of the traces / (3 V), after one NVE step from
    the state (the stress is computed after a step)."""
    res = {}
    vol = abs(np.linalg.det(np.asarray(s.box, dtype=np.float64)))
    for prec in (G.FP64, G.MIXED):
        ctx = G.eam_context(eam, [s], precision=prec, ensemble=G.NVE,
                            timestep=W.FS, nbl_margin=0.4)
        ctx.calc_forces()
        a, th = ctx.atoms(), ctx.thermo()
        ctx.set_stress(True)
        ctx.move_atoms()
        per, tot = ctx.stress()
        ctx.close()
        emb = G.Context(precision=prec, ensemble=G.NVE, ntypes=1,
                        timestep=W.FS, nbl_margin=0.4)
        emb.set_tables(None, G.TableArrays.from_format2(eam.rho),
                       G.TableArrays.from_format2(eam.embed))
        emb.set_systems([s])
        emb.calc_forces()
        femb = emb.atoms()["epot"]
        emb.close()
        res[prec] = dict(f=a["f"], rho=a["rho"], df=a["df"], epot=a["epot"],
                         femb=femb, press=per[:, :3].sum(axis=1) / 3.0,
                         ptot=tot[0, :3].sum() / (3.0 * vol),
                         virial=th["virial"], etot=th["epot"])
    r64, r32 = res[G.FP64], res[G.MIXED]
    out = {}
    for key, name in (("f", "force"), ("rho", "rho"), ("femb", "F(rho)"),
                      ("df", "2F'(rho)"), ("epot", "energy per atom"),
                      ("press", "pressure per atom")):
        ref = np.asarray(r64[key], dtype=np.float64)
        d = np.asarray(r32[key], dtype=np.float64) - ref
        scale = np.sqrt(np.mean(ref * ref))
        out[name] = {"rms_rel": float(np.sqrt(np.mean(d * d)) / scale),
                     "max_abs": float(np.max(np.abs(d))),
                     "max_rel": float(np.max(np.abs(d)) / scale),
                     "rms_ref": float(scale)}
    out["total energy"] = {"rel": abs(r32["etot"] - r64["etot"]) / abs(r64["etot"])}
    out["virial"] = {"rel": abs(r32["virial"] - r64["virial"]) / abs(r64["virial"])}
    out["pressure"] = {"rel": abs(r32["ptot"] - r64["ptot"]) / abs(r64["ptot"]),
                       "value": r64["ptot"]}
    return out


def precision_errors(G, eam, s):
    """Per-quantity error of the mixed path against the FP64 path on the
    initial configuration (state_errors)."""
    return state_errors(G, eam, s)


def rebuild_probe(args):
    """Host-timed cost of the list rebuild on the 2M-atom crystal: steps of
    the crystal at 1 K (no skin test fires, so the rebuild launches of every
    step return at once) and forced full rebuilds (make_nblist: the whole
    chain, then a host synchronisation).  One JSON line."""
    from imd_amd import imdg as G
    eam, s = make_workload(args.cells, args.seed, 1.0)
    ctx = G.eam_context(eam, [s], precision=G.MIXED, ensemble=G.NVE,
                        timestep=W.FS, nbl_margin=0.4)
    ctx.run(args.warmup)
    tuning = ctx.tuning() if args.no_autotune else ctx.autotune(reps=10)
    ctx.run(2)
    nb0 = ctx.info()["nbl_count"]
    t0 = time.perf_counter()
    ctx.run(args.steps)
    t_step = (time.perf_counter() - t0) / args.steps
    nb1 = ctx.info()["nbl_count"]
    ctx.make_nblist()
    reps = 10
    t0 = time.perf_counter()
    for _ in range(reps):
        ctx.make_nblist()
    t_build = (time.perf_counter() - t0) / reps
    ctx.close()
    print(json.dumps({"metric": "list rebuild cost (host-timed)",
                      "config": {"workload": "fcc EAM crystal %d^3 cells "
                                 "(%d atoms) at 1 K, mixed" % (args.cells, s.n),
                                 "steps": args.steps},
                      "launch_shape": _shape(tuning),
                      "step_us_without_rebuild": 1e6 * t_step,
                      "rebuilds_during_steps": int(nb1 - nb0),
                      "full_rebuild_us": 1e6 * t_build}))
    return 0


def precision_study(args):
    """Per-quantity error of the mixed path against the FP64 path along the
    2M-atom NVT run: the mixed run advances args.precision_study steps, and
    every args.sample_every steps its state (positions, momenta) is
    evaluated by both precisions (state_errors).  Prints one JSON line with
    every sample and the largest value of each error over the samples."""
    import dataclasses
    from imd_amd import imdg as G
    eam, s = make_workload(args.cells, args.seed, args.temp)
    isq = 1.0 / (args.tau_fs * W.FS) ** 2
    ctx = G.eam_context(eam, [s], precision=G.MIXED, ensemble=G.NVT,
                        timestep=W.FS, nbl_margin=0.4)
    ctx.set_nvt(0, W.kT(args.temp), 0.0, isq)
    samples, done = [], 0
    while True:
        a = ctx.atoms()
        st = dataclasses.replace(s, x=a["x"], p=a["p"])
        err = state_errors(G, eam, st)
        samples.append({"step": done, "errors": err})
        print(json.dumps({"sample": done, "force_rms_rel":
                          err["force"]["rms_rel"]}), file=sys.stderr,
              flush=True)
        if done >= args.precision_study:
            break
        n = min(args.sample_every, args.precision_study - done)
        ctx.run(n)
        done += n
    ctx.close()
    worst = {}
    for smp in samples:
        for q, e in smp["errors"].items():
            w = worst.setdefault(q, {})
            for k, v in e.items():
                if k in ("rms_ref", "value"):
                    continue
                w[k] = max(w.get(k, 0.0), v)
    print(json.dumps({"metric": "mixed vs FP64 per-quantity error along the "
                                "NVT run",
                      "config": {"workload": "fcc EAM crystal %d^3 cells (%d "
                                             "atoms), NVT %g K, dt 1 fs"
                                             % (args.cells, s.n, args.temp),
                                 "steps": args.precision_study,
                                 "sample_every": args.sample_every},
                      "max_over_samples": worst, "samples": samples}))
    return 0



CONFIGS = {
    "c0": "4,000-atom fcc LJ crystal (10^3 cells, a = 1.5496 sigma), LJ table "
          "r^2 0.25-6.25 with 2,000 points, kT 0.7, NVE, dt 0.002, skin 0.3, FP64",
    "c1": "256,000-atom fcc EAM crystal (40^3 cells), Johnson-form tables "
          "5,000 points r_c 5.0 A, 600 K, NVE, dt 1 fs, skin 0.4 A",
    "c3": "64 independent fcc EAM replicas, N log-uniform in [500, 50,000], "
          "0-1% vacancies, 300-1,200 K, NVE, dt 1 fs, skin 0.4 A, mixed",
    "c4": "8,000,000-atom binary fcc slab (200x200x50 cells, 70/30 types, "
          "open z with 20 A vacuum), two-species Johnson EAM r_c 5.5 A, "
          "NVT 500 K, dt 1 fs, skin 0.4 A, mixed",
}


def config_run(args):
    """Throughput of one of the other configurations (one JSON line; the
    default run measures the 2M-atom metric workload)."""
    from imd_amd import imdg as G
    name = args.config
    runs = []
    if name == "c0":
        s = W.lj_crystal()
        systems = [s]
        make = lambda prec: G.pair_context(W.lj_table(), systems, precision=prec,
                                           ensemble=G.NVE, timestep=0.002,
                                           nbl_margin=0.3)
        runs = [("fp64", G.FP64), ("mixed", G.MIXED)]
    elif name == "c1":
        eam = W.johnson_eam()
        systems = [W.eam_crystal(n_cells=40, temp_k=600.0, seed=1001)]
        make = lambda prec: G.eam_context(eam, systems, precision=prec,
                                          ensemble=G.NVE, timestep=W.FS,
                                          nbl_margin=0.4)
        runs = [("mixed", G.MIXED), ("fp64", G.FP64)]
    elif name == "c3":
        eam = W.johnson_eam()
        systems = W.replica_set(64, seed=1003)
        make = lambda prec: G.eam_context(eam, systems, precision=prec,
                                          ensemble=G.NVE, timestep=W.FS,
                                          nbl_margin=0.4)
        runs = [("mixed", G.MIXED), ("fp64", G.FP64)]
    elif name == "c4":
        eam = W.johnson_eam((W.CU, W.NI), rc=5.5)
        systems = [W.alloy_slab()]
        isq = 1.0 / (args.tau_fs * W.FS) ** 2

        def make(prec):
            ctx = G.eam_context(eam, systems, precision=prec, ensemble=G.NVT,
                                timestep=W.FS, nbl_margin=0.4)
            ctx.set_nvt(0, W.kT(500.0), 0.0, isq)
            return ctx
        runs = [("mixed", G.MIXED)]
    else:
        raise SystemExit("unknown config %s" % name)
    natoms = sum(x.n for x in systems)
    # the batched replicas replay one captured step per step (hipGraph), and
    # are also measured with plain stream launches (faster on MI355X)
    graph_default = name == "c3" if args.graph == "auto" else args.graph == "on"
    runs = [(label, prec, graph_default) for label, prec in runs]
    if name == "c3" and args.graph == "auto":
        runs += [(label + "_stream", prec, False) for label, prec, _ in runs]
    res = {}
    for label, prec, graph in runs:
        old = os.environ.get("IMDG_GRAPH")
        os.environ["IMDG_GRAPH"] = "1" if graph else "0"
        try:
            ctx = make(prec)
        finally:
            if old is None:
                del os.environ["IMDG_GRAPH"]
            else:
                os.environ["IMDG_GRAPH"] = old
        ctx.run(args.warmup)
        tuning = ctx.tuning() if args.no_autotune else ctx.autotune(reps=10)
        ctx.run(2)  # capture the step with the chosen shapes
        nb0 = ctx.info()["nbl_count"]
        ctx.sync()
        t0 = time.perf_counter()
        ctx.run(args.steps, sync=False)
        ctx.sync()
        el = time.perf_counter() - t0
        info = ctx.info()
        ctx.close()
        res[label] = {"atom_steps_per_s": natoms * args.steps / el,
                      "ms_per_step": 1e3 * el / args.steps,
                      "mean_neighbours": info["mean_k"],
                      "list_rebuilds_timed": int(info["nbl_count"] - nb0),
                      "launch_shape": _shape(tuning),
                      "step": "hipGraph replay" if graph else "stream launches"}
    out = {"metric": "atom-steps/s", "config": {"workload": CONFIGS[name],
                                                 "name": name,
                                                 "natoms": natoms,
                                                 "replicas": len(systems)},
           "steps": args.steps, "warmup": args.warmup, "results": res}
    if "mixed" in res and "fp64" in res:
        out["mixed_over_fp64"] = (res["mixed"]["atom_steps_per_s"]
                                  / res["fp64"]["atom_steps_per_s"])
    print(json.dumps(out))
    return 0


def _shape(t):
    """Launch shape of the two neighbour loops as reported by the library
    (atoms per block = threads / lanes)."""
    return {"density_lanes_per_atom": t["density_lanes"],
            "density_threads": t["density_threads"],
            "force_lanes_per_atom": t["force_lanes"],
            "force_threads": t["force_threads"],
            "autotuned": t["force_ms"] > 0}


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1000)
    ap.add_argument("--warmup", type=int, default=50)
    ap.add_argument("--precision", choices=["mixed", "fp64"], default="mixed")
    ap.add_argument("--cells", type=int, default=80)
    ap.add_argument("--seed", type=int, default=1002)
    ap.add_argument("--temp", type=float, default=800.0)
    ap.add_argument("--tau-fs", type=float, default=100.0)
    ap.add_argument("--cpu-steps", type=int, default=40)
    ap.add_argument("--cpu-threads", type=int, default=0,
                    help="host threads of the CPU baseline (0: the "
                         "process's CPU share, at most 16)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-profile", action="store_true")
    ap.add_argument("--no-traffic", action="store_true",
                    help="skip the rocprofv3 FETCH_SIZE/WRITE_SIZE passes")
    ap.add_argument("--copy-probe", type=int, default=0, metavar="MIB",
                    help="only run one device-to-device copy of MIB MiB "
                         "(counter calibration) and exit")
    ap.add_argument("--drift", type=int, default=0, metavar="STEPS",
                    help="only measure the NVE energy drift over STEPS steps "
                         "(both precisions) and exit")
    ap.add_argument("--drift-config", choices=sorted(DRIFT_CONFIGS),
                    default="eam256k", help="input of the --drift run")
    ap.add_argument("--precision-study", type=int, default=0, metavar="STEPS",
                    help="only sample the mixed-vs-FP64 errors along STEPS "
                         "steps of the NVT run and exit")
    ap.add_argument("--sample-every", type=int, default=1000)
    ap.add_argument("--rebuild-probe", action="store_true",
                    help="host-timed steps without rebuilds and forced full "
                         "rebuilds of the 2M-atom crystal")
    ap.add_argument("--config", choices=sorted(CONFIGS), default=None,
                    help="measure one of the other configurations instead "
                         "of the 2M-atom metric workload")
    ap.add_argument("--graph", choices=["auto", "on", "off"], default="auto",
                    help="--config runs: replay one captured step per step "
                         "(auto: for the batched replicas only)")
    ap.add_argument("--no-autotune", action="store_true",
                    help="keep the default launch shapes instead of timing "
                         "every lanes-per-atom x block-size shape first")
    ap.add_argument("--no-compare", action="store_true",
                    help="skip the timed run of the other precision and the "
                         "per-quantity errors of mixed vs FP64")
    args = ap.parse_args(argv)
    if args.copy_probe:
        return copy_probe(args.copy_probe)
    if args.drift:
        return drift_run(args)
    if args.precision_study:
        return precision_study(args)
    if args.rebuild_probe:
        return rebuild_probe(args)
    if args.config:
        return config_run(args)

    if args.gpus != 1:
        print("bench.py measures one GPU only: this build runs the path on "
              "a single GPU (several GPUs are a design, not code); run with "
              "--gpus 1", file=sys.stderr)
        return 2

    mixed = args.precision == "mixed"
    eam, s = make_workload(args.cells, args.seed, args.temp)
    isq = 1.0 / (args.tau_fs * W.FS) ** 2

    # the counter passes run as child processes before this process touches
    # the GPU (a process that has initialised the GPU starts no programs)
    traffic = None
    if not args.no_traffic and not args.no_profile:
        traffic = measure_traffic(args)

    from imd_amd import imdg as G

    def run_one(prec_mixed, profile):
        ctx = G.eam_context(eam, [s], precision=G.MIXED if prec_mixed else G.FP64,
                            ensemble=G.NVT, timestep=W.FS, nbl_margin=0.4)
        ctx.set_nvt(0, W.kT(args.temp), 0.0, isq)
        ctx.run(args.warmup)
        # launch shapes: timed on the warmed-up state (which it leaves as it
        # is), before the timed region
        tuning = ctx.tuning() if args.no_autotune else ctx.autotune(reps=10)
        nb0 = ctx.info()["nbl_count"]
        # the timed region records HIP events around the force launches only
        # (the roofline kernel); the other classes are timed afterwards
        if profile:
            # every 10th force launch of the timed region gets an event
            # pair (each pair makes the stream wait at both ends of the
            # launch: timing all of them would cost ~6 us per step)
            os.environ["IMDG_PROFILE_STRIDE"] = "10"
            ctx.profile(True, classes=("force",))
            del os.environ["IMDG_PROFILE_STRIDE"]
        ctx.sync()
        t0 = time.perf_counter()
        ctx.run(args.steps, sync=False)
        ctx.sync()
        el = time.perf_counter() - t0
        info = ctx.info()
        th = ctx.thermo()
        kt = {}
        if profile:
            kt = {"force": ctx.kernel_times()["force"]}
            ctx.profile(True)
            ctx.run(min(args.steps, 200))
            other = ctx.kernel_times()
            for k in ("density", "rebuild", "finalize"):
                kt[k] = other[k]
            ctx.profile(False)
        ctx.close()
        return el, info, kt, th, info["nbl_count"] - nb0, tuning

    el, info, kt, th, rebuilds, tuning = run_one(mixed, not args.no_profile)
    natoms = s.n
    value = natoms * args.steps / el
    ms_per_step = 1e3 * el / args.steps
    k_mean = info["mean_k"]

    roofline = None
    if kt and kt.get("force", (0, 0))[1] > 0:
        ms, n = kt["force"]
        avg_s = ms / n / 1e3
        byts = algorithmic_bytes("fused", k_mean, mixed) * natoms
        ach = byts / avg_s / 1e9
        step_ach = value * algorithmic_bytes("step", k_mean, mixed) / 1e9
        roofline = {"bound": "hbm", "achieved": round(ach, 1),
                    "peak": HBM_PEAK_GBS, "unit": "GB/s",
                    "frac": round(ach / HBM_PEAK_GBS, 4), "traffic": None,
                    "kernel": "k_force (pair+EAM force, leapfrog, skin test)",
                    "avg_launch_us": round(avg_s * 1e6, 2),
                    "bytes_per_launch": int(byts),
                    "step_achieved": round(step_ach, 1),
                    "step_frac": round(step_ach / HBM_PEAK_GBS, 4),
                    "step_bytes_per_atom": round(
                        algorithmic_bytes("step", k_mean, mixed), 1)}
        if "density" in kt and kt["density"][1] > 0:
            dms, dn = kt["density"]
            roofline["density_avg_launch_us"] = round(dms / dn * 1e3, 2)
            roofline["density_achieved"] = round(
                algorithmic_bytes("density", k_mean, mixed) * natoms
                / (dms / dn / 1e3) / 1e9, 1)
        if "rebuild" in kt and kt["rebuild"][1] > 0:
            roofline["rebuild_ms_per_step"] = round(kt["rebuild"][0] / kt["rebuild"][1], 4)
        if "finalize" in kt and kt["finalize"][1] > 0:
            roofline["finalize_avg_launch_us"] = round(
                kt["finalize"][0] / kt["finalize"][1] * 1e3, 2)

    if roofline is not None and traffic is not None:
        # HBM bytes per launch: calibrated FETCH_SIZE + WRITE_SIZE
        roofline["traffic"] = int(traffic["bytes"])
        roofline["traffic_over_algorithmic"] = round(
            traffic["bytes"] / roofline["bytes_per_launch"], 3)

    cpu = None
    if not args.no_cpu_baseline:
        cpu = cpu_baseline(eam, s, args.cpu_steps, args.tau_fs, args.temp,
                           args.cpu_threads or cpu_threads())

    compare = None
    if not args.no_compare:
        el2, _, _, _, _, tun2 = run_one(not mixed, False)
        other = natoms * args.steps / el2
        compare = {"other_precision": "fp64" if mixed else "mixed",
                   "other_value": round(other, 1),
                   "other_tuning": _shape(tun2),
                   "mixed_over_fp64": round(value / other if mixed
                                            else other / value, 3),
                   "errors_mixed_vs_fp64": precision_errors(G, eam, s)}

    out = {
        "metric": METRIC,
        "value": round(value, 1),
        "unit": "atom-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 forces/tables + f64 integration" if mixed else "f64",
        "data": "synthetic",
        "config": {"workload": "fcc EAM crystal %d^3 cells (%d atoms), "
                               "Johnson-form EAM tables 5000 pts r_c 5.0 A, "
                               "NVT 800 K, dt 1 fs, skin 0.4 A"
                               % (args.cells, natoms),
                   "natoms": natoms, "precision": args.precision,
                   "mean_neighbours": round(k_mean, 2),
                   "list_rebuilds_timed": int(rebuilds),
                   "epot_per_atom": th["epot"] / natoms,
                   "launch_shape": _shape(tuning),
                   "parallelism": "1 GPU"},
        "roofline": roofline,
        "cpu_baseline": cpu,
    }
    if compare is not None:
        out["precision_compare"] = compare
    print(json.dumps(out))
    return 0


if __name__ == "__main__":
    sys.exit(main())
