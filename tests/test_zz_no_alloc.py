"""imdg_check_nblist_noalloc (no allocation: a file-scope device scratch
array) against the context's imdg_check_nblist on the same states, through
the C ABI: the positions and the positions at the last list build in device
buffers the test owns, the maximum squared displacement and the rebuild
flag written to device memory.  Two systems in a row, the second smaller;
each stepped (unfused move_atoms + check_nblist) until the skin test fires.
The maximum must equal the context's own (the thermo record of the step)
bit for bit, and the flag the context's decision."""
import ctypes as C

import numpy as np
import pytest

from imd_amd import imdg as G
from imd_amd import workloads as W

from helpers import eam_tables

pytestmark = pytest.mark.gpu


class Hip:
    def __init__(self):
        self.h = C.CDLL("libamdhip64.so")
        self.h.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        self.h.hipFree.argtypes = [C.c_void_p]
        self.h.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t,
                                     C.c_int]
        self.h.hipDeviceSynchronize.argtypes = []

    def alloc(self, nbytes):
        p = C.c_void_p()
        assert self.h.hipMalloc(C.byref(p), C.c_size_t(max(nbytes, 8))) == 0
        return p

    def put(self, dst, arr):
        arr = np.ascontiguousarray(arr)
        assert self.h.hipMemcpy(dst, arr.ctypes.data_as(C.c_void_p),
                                arr.nbytes, 1) == 0

    def get(self, src, arr):
        assert self.h.hipMemcpy(arr.ctypes.data_as(C.c_void_p), src,
                                arr.nbytes, 2) == 0
        return arr


def entry():
    L = G.lib()
    f = L.imdg_check_nblist_noalloc
    f.argtypes = [C.c_int, C.c_void_p, C.c_void_p, C.c_long, C.c_double,
                  C.c_void_p, C.c_void_p, C.c_void_p]
    f.restype = C.c_int
    L.imdg_last_error.restype = C.c_char_p
    L.imdg_last_error.argtypes = [C.c_void_p]
    return f, lambda: L.imdg_last_error(None)


def _walk(system, margin, hip, f, dx, dr, dmax, dinv):
    """Step the context until its skin test fires; at every step compare the
    no-allocation entry point with the context's result."""
    eam = eam_tables()
    n = system.n
    with G.eam_context(eam, [system], precision=G.FP64,
                       nbl_margin=margin) as gpu:
        gpu.calc_forces()
        x0 = gpu.atoms()["x"]            # positions of the list build
        hip.put(dr, x0)
        for step in range(400):
            gpu.move_atoms()
            inv = gpu.check_nblist()
            th = gpu.thermo()
            hip.put(dx, gpu.atoms()["x"])
            hip.put(dmax, np.array([-1.0]))
            hip.put(dinv, np.array([-1], dtype=np.int32))
            rc = f(0, dx, dr, n, margin, dmax, dinv, None)
            assert rc == 0
            assert hip.h.hipDeviceSynchronize() == 0
            m = hip.get(dmax, np.zeros(1))[0]
            flag = hip.get(dinv, np.zeros(1, dtype=np.int32))[0]
            assert m == th["max_disp2"], (step, m, th["max_disp2"])
            assert bool(flag) == inv, step
            if inv:
                return step
    raise AssertionError("the skin test never fired")


def test_noalloc_check_matches_context_twice():
    hip = Hip()
    f, err = entry()
    big = W.eam_crystal(n_cells=10, temp_k=1500.0, seed=61, displace=0.02)
    small = W.eam_crystal(n_cells=5, temp_k=1500.0, seed=62, displace=0.02)
    nmax = big.n
    dx, dr = hip.alloc(24 * nmax), hip.alloc(24 * nmax)
    dmax, dinv = hip.alloc(8), hip.alloc(4)
    try:
        s1 = _walk(big, 0.2, hip, f, dx, dr, dmax, dinv)
        s2 = _walk(small, 0.2, hip, f, dx, dr, dmax, dinv)
        assert s1 > 0 and s2 > 0
        # no atoms: the maximum is 0 and the list stays valid
        assert f(0, dx, dr, 0, 0.2, dmax, dinv, None) == 0
        assert hip.h.hipDeviceSynchronize() == 0
        assert hip.get(dmax, np.zeros(1))[0] == 0.0
        assert hip.get(dinv, np.zeros(1, dtype=np.int32))[0] == 0
        # beyond the fixed scratch: a clear error, nothing launched
        assert f(0, dx, dr, 1 << 40, 0.2, dmax, dinv, None) == G.ERR_ARG
        assert b"exceed" in err()
    finally:
        for p in (dx, dr, dmax, dinv):
            hip.h.hipFree(p)
