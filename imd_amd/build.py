"""Build the native parts in-tree.

* ``imd_amd/libimdg.so`` -- the HIP library behind include/imdg.h, built for
  gfx950 with hipcc.  Packed-FP32 instructions are disabled
  (``-target-feature -packed-fp32-ops``); FP32 throughput comes from the
  scalar VALU rate instead.
* ``imd_amd/host/imd_gpu`` and ``imd_amd/host/libimdhost.so`` -- the C host
  (IMD parameter files, tables, configs, calc_forces/move_atoms entry
  points) linked against libimdg.so.
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
HOST = os.path.join(PKG, "host")
INCLUDE = os.path.join(ROOT, "include")
LIB = os.path.join(PKG, "libimdg.so")
HOSTLIB = os.path.join(HOST, "libimdhost.so")
HOSTEXE = os.path.join(HOST, "imd_gpu")

# (source, extra flags, object): the step kernels are compiled once per
# precision so that the two halves build in parallel
HIP_UNITS = [
    ("imdg_build.hip", [], "imdg_build.o"),
    ("imdg_step.hip", ["-DIMDG_STEP_FP32"], "imdg_step_f32.o"),
    ("imdg_step.hip", ["-DIMDG_STEP_FP64"], "imdg_step_f64.o"),
    ("imdg_api.hip", [], "imdg_api.o"),
    ("imdg_noalloc.hip", [], "imdg_noalloc.o"),
]
HIP_SOURCES = sorted({u[0] for u in HIP_UNITS})
HIPCC_FLAGS = [
    "--offload-arch=gfx950", "-O3", "-fPIC", "-std=c++17",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function",
]


def hipcc() -> str:
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc",
                 shutil.which("hipcc")):
        if cand and os.path.exists(cand):
            return cand
    raise RuntimeError("hipcc not found")


def _run(cmd, cwd=None):
    r = subprocess.run(cmd, cwd=cwd, stdout=subprocess.PIPE,
                       stderr=subprocess.STDOUT, text=True)
    out = "\n".join(l for l in r.stdout.splitlines()
                    if "packed-fp32-ops" not in l or "error" in l)
    if r.returncode != 0:
        sys.stderr.write(out + "\n")
        raise RuntimeError("build step failed: %s" % " ".join(cmd))
    return out


def build_lib(verbose: bool = False) -> str:
    from concurrent.futures import ThreadPoolExecutor
    objs = [os.path.join(CSRC, o) for _, _, o in HIP_UNITS]
    cmds = [[hipcc(), *HIPCC_FLAGS, *flags, "-I", INCLUDE, "-c",
             os.path.join(CSRC, src), "-o", obj]
            for (src, flags, _), obj in zip(HIP_UNITS, objs)]
    with ThreadPoolExecutor(max_workers=len(cmds)) as ex:
        for out in ex.map(_run, cmds):
            if verbose and out.strip():
                print(out)
    tmp = LIB + ".tmp"
    _run([hipcc(), "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp,
          *objs])
    os.replace(tmp, LIB)
    for o in objs:
        os.unlink(o)
    return LIB


def build_host() -> str:
    srcs = sorted(f for f in os.listdir(HOST) if f.endswith(".c"))
    if not srcs:
        return ""
    lib_srcs = [os.path.join(HOST, f) for f in srcs if f != "imd_main.c"]
    cflags = ["-O2", "-std=c11", "-Wall", "-Wextra", "-Wno-unused-parameter",
              "-fPIC", "-I", INCLUDE, "-I", HOST]
    ldflags = ["-L", PKG, "-limdg", "-Wl,-rpath,$ORIGIN/..", "-lm"]
    _run(["gcc", *cflags, "-shared", "-o", HOSTLIB, *lib_srcs, *ldflags])
    if "imd_main.c" in srcs:
        _run(["gcc", *cflags, "-o", HOSTEXE, os.path.join(HOST, "imd_main.c"),
              *lib_srcs, *ldflags])
    return HOSTLIB


def build(verbose: bool = False) -> None:
    build_lib(verbose)
    build_host()


if __name__ == "__main__":
    build(verbose=True)
    print("built", LIB)
