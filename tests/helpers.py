"""Shared test helpers: golden fixture loading."""
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

GOLDEN = os.path.join(ROOT, "tests", "golden")
CASES = ("water_tiny", "water_508", "chains_tiny", "ee_tiny", "fspc_tiny")


# per-call arrays of the force fixtures (one per ALLFORCE / SLFORCES +
# FFORCES call of the reference harness)
PERCALL = ("SX", "SY", "SZ", "XCOM", "YCOM", "ZCOM", "FX", "FY", "FZ", "GX", "GY", "GZ", "HX", "HY", "HZ",
           "ENERGY", "POTE1", "POTL1", "PES141", "PSR141", "SPEE", "BR", "BE", "AR", "AE", "TR", "TE",
           "EEDIFF", "EEPAR", "EC")


def load_golden(name):
    """Fixture as a dump-like dict: static arrays as [a], per-call arrays as
    [call0, call1, ...]."""
    z = np.load(os.path.join(GOLDEN, name + ".npz"), allow_pickle=False)
    d = {}
    ncall = int(z["NCALL"])
    for k in z.files:
        if k[-2:] in ("_0", "_1") and k[:-2] in PERCALL:
            continue
        d[k] = [z[k]]
    for k in PERCALL:
        if "%s_0" % k in z.files:
            d[k] = [z["%s_%d" % (k, c)] for c in range(ncall)]
    d["lam"] = float(z["lam"])
    return d


def golden_positions(d, call=0):
    return np.stack([d["SX"][call], d["SY"][call], d["SZ"][call]], 1)


def golden_com(d, call=0):
    return np.stack([d["XCOM"][call], d["YCOM"][call], d["ZCOM"][call]], 1)


def golden_forces(d, key, call=0):
    return np.stack([d[key + "X"][call], d[key + "Y"][call], d[key + "Z"][call]], 1)


class DeviceArrays:
    """Device buffers from the HIP runtime the engine library links
    (libamdhip64.so.7, already loaded with it), for tests of entry points that
    take device pointers.  A second HIP runtime in the process (the one a
    PyTorch wheel bundles) would compete with it for the device."""

    def __init__(self):
        import ctypes as C
        from mdynamix_amd.engine import load_library
        load_library()
        self.C = C
        self.hip = C.CDLL("libamdhip64.so.7")
        self.hip.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        self.hip.hipFree.argtypes = [C.c_void_p]
        self.hip.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        self.ptrs = []

    def alloc(self, nbytes):
        p = self.C.c_void_p()
        assert self.hip.hipMalloc(self.C.byref(p), self.C.c_size_t(max(int(nbytes), 1))) == 0
        self.ptrs.append(p.value)
        return p.value

    def upload(self, a):
        a = np.ascontiguousarray(a)
        p = self.alloc(a.nbytes)
        assert self.hip.hipMemcpy(p, a.ctypes.data, a.nbytes, 1) == 0      # host to device
        return p

    def download(self, p, n, dtype=np.float64):
        out = np.zeros(n, dtype)
        assert self.hip.hipDeviceSynchronize() == 0
        assert self.hip.hipMemcpy(out.ctypes.data, p, out.nbytes, 2) == 0  # device to host
        return out

    def close(self):
        for p in self.ptrs:
            self.hip.hipFree(p)
        self.ptrs = []
