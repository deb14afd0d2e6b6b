"""Shared test helpers: golden fixture loading."""
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

GOLDEN = os.path.join(ROOT, "tests", "golden")
CASES = ("water_tiny", "water_508", "chains_tiny", "ee_tiny", "fspc_tiny")


def load_golden(name):
    """Fixture as a dump-like dict: static arrays as [a], per-call arrays as
    [call0, call1, ...]."""
    z = np.load(os.path.join(GOLDEN, name + ".npz"), allow_pickle=False)
    d = {}
    ncall = int(z["NCALL"])
    for k in z.files:
        if k[-2:] in ("_0", "_1") and k[:-2] in ("SX", "SY", "SZ", "XCOM", "YCOM", "ZCOM", "FX", "FY", "FZ",
                                                 "GX", "GY", "GZ", "HX", "HY", "HZ", "ENERGY", "POTE1",
                                                 "POTL1", "PES141", "PSR141", "SPEE"):
            continue
        d[k] = [z[k]]
    for k in ("SX", "SY", "SZ", "XCOM", "YCOM", "ZCOM", "FX", "FY", "FZ", "GX", "GY", "GZ", "HX", "HY",
              "HZ", "ENERGY", "POTE1", "POTL1", "PES141", "PSR141", "SPEE"):
        d[k] = [z["%s_%d" % (k, c)] for c in range(ncall)]
    d["lam"] = float(z["lam"])
    return d


def golden_positions(d, call=0):
    return np.stack([d["SX"][call], d["SY"][call], d["SZ"][call]], 1)


def golden_com(d, call=0):
    return np.stack([d["XCOM"][call], d["YCOM"][call], d["ZCOM"][call]], 1)


def golden_forces(d, key, call=0):
    return np.stack([d[key + "X"][call], d[key + "Y"][call], d[key + "Z"][call]], 1)
