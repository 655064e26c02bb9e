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

from imd_amd import imdg as G  # noqa: E402
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
    The fused force kernel carries force + integrate: 4k+188 | 4k+220.
    """
    if kind == "density":
        return 4 * k + (24 if mixed else 48)
    if kind == "force":
        return 4 * k + (36 + 152 if mixed else 72 + 148)
    raise ValueError(kind)


def make_workload(cells: int, seed: int, temp_k: float):
    eam = W.johnson_eam()
    s = W.eam_crystal(n_cells=cells, temp_k=temp_k, seed=seed, displace=0.02)
    return eam, s


def cpu_baseline(eam, s, steps: int, tau_fs: float, temp_k: float):
    """The oracle (single-threaded C restatement of IMD's NBL/EAM path) on
    the same workload, timed over a bounded number of steps after its first
    (list-building) step."""
    import oracle as O
    pair = O.OracleTable.from_table(eam.pair)
    rho = O.OracleTable.from_table(eam.rho)
    emb = O.OracleTable.from_table(eam.embed, radial=False)
    orc = O.OracleSystem(s, pair, rho, emb, nbl_margin=0.4, timestep=W.FS,
                         ensemble=O.ORC_NVT, temperature=W.kT(temp_k),
                         eta=0.0, isq_tau_eta=1.0 / (tau_fs * W.FS) ** 2)
    orc.run(1)
    t0 = time.perf_counter()
    orc.run(steps)
    dt = time.perf_counter() - t0
    return dict(value=s.n * steps / dt, unit="atom-steps/s", cores=1,
                kind="port",
                sample="%d NVT steps of the same %d-atom workload after one "
                       "warm-up step, oracle (FP64 C, half list), 1 thread"
                       % (steps, s.n))


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
    ap.add_argument("--cpu-steps", type=int, default=5)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-profile", action="store_true")
    ap.add_argument("--copy-probe", type=int, default=0, metavar="MIB",
                    help="only run one device-to-device copy of MIB MiB "
                         "(counter calibration) and exit")
    ap.add_argument("--compare", action="store_true",
                    help="also time the other precision (ratio in stderr)")
    args = ap.parse_args(argv)
    if args.copy_probe:
        return copy_probe(args.copy_probe)

    if args.gpus != 1:
        print("bench.py measures one GPU only: this build has no multi-GPU "
              "path (the halo-exchange design is described in DESIGN.md); "
              "run with --gpus 1", file=sys.stderr)
        return 2

    mixed = args.precision == "mixed"
    eam, s = make_workload(args.cells, args.seed, args.temp)
    isq = 1.0 / (args.tau_fs * W.FS) ** 2

    def run_one(prec_mixed, profile):
        ctx = G.eam_context(eam, [s], precision=G.MIXED if prec_mixed else G.FP64,
                            ensemble=G.NVT, timestep=W.FS, nbl_margin=0.4)
        ctx.set_nvt(0, W.kT(args.temp), 0.0, isq)
        ctx.run(args.warmup)
        nb0 = ctx.info()["nbl_count"]
        if profile:
            ctx.profile(True)
        ctx.sync()
        t0 = time.perf_counter()
        ctx.run(args.steps, sync=False)
        ctx.sync()
        el = time.perf_counter() - t0
        info = ctx.info()
        kt = ctx.kernel_times() if profile else {}
        th = ctx.thermo()
        ctx.close()
        return el, info, kt, th, info["nbl_count"] - nb0

    el, info, kt, th, rebuilds = run_one(mixed, not args.no_profile)
    natoms = s.n
    value = natoms * args.steps / el
    ms_per_step = 1e3 * el / args.steps
    k_mean = info["mean_k"]

    roofline = None
    if kt and kt.get("force", (0, 0))[1] > 0:
        ms, n = kt["force"]
        avg_s = ms / n / 1e3
        byts = algorithmic_bytes("force", k_mean, mixed) * natoms
        ach = byts / avg_s / 1e9
        roofline = {"bound": "hbm", "achieved": round(ach, 1),
                    "peak": HBM_PEAK_GBS, "unit": "GB/s",
                    "frac": round(ach / HBM_PEAK_GBS, 4), "traffic": None,
                    "kernel": "k_force (pair+EAM force, leapfrog, skin test)",
                    "avg_launch_us": round(avg_s * 1e6, 2),
                    "bytes_per_launch": int(byts)}
        if "density" in kt and kt["density"][1] > 0:
            dms, dn = kt["density"]
            roofline["density_avg_launch_us"] = round(dms / dn * 1e3, 2)
            roofline["density_achieved"] = round(
                algorithmic_bytes("density", k_mean, mixed) * natoms
                / (dms / dn / 1e3) / 1e9, 1)
        if "rebuild" in kt and kt["rebuild"][1] > 0:
            roofline["rebuild_ms_per_step"] = round(kt["rebuild"][0] / kt["rebuild"][1], 4)

    cpu = None
    if not args.no_cpu_baseline:
        cpu = cpu_baseline(eam, s, args.cpu_steps, args.tau_fs, args.temp)

    if args.compare:
        el2, _, _, _, _ = run_one(not mixed, False)
        other = natoms * args.steps / el2
        ratio = value / other if mixed else other / value
        print("mixed/fp64 atom-steps/s ratio: %.3f" % ratio, file=sys.stderr)

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
                   "parallelism": "1 GPU"},
        "roofline": roofline,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out))
    return 0


if __name__ == "__main__":
    sys.exit(main())
