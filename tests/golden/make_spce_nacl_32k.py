"""Fixture of the 32k SPC/E + 0.15 M NaCl box (96,174 atoms, 98.7 A,
R_cut 12 A, alpha R_cut 3.2, FEXP 11.5, 48,052 k vectors) from the FP64
oracle, which is pinned against the reference program on the smaller golden
fixtures (tests/test_oracle_golden.py).  The reference itself is not run
at this size here: its O(N^2) list build and 4.6e9-term reciprocal sum
take minutes per call, which the oracle also takes, and the oracle's
agreement with the reference is what the small fixtures establish.

Two calls:
  call 0  the generator's positions, reference lists (CHCNB semantics: no
          skin), ALLFORCE: list lengths and SHA-256 digests of the lists in
          the reference layout, a digest of the k set, F in FP64, G and H
          in FP32 (their check is at 1e-5), every energy/virial scalar;
  call 1  every atom displaced by U(-0.25, 0.25) A per axis (seed 2026),
          lists rebuilt at the displaced positions: F in FP64, the
          scalars and the pair count, for the engine's Verlet list (built
          at call 0's positions with a 1 A skin) evaluated at call 1's.

The positions are regenerated by mdynamix_amd.systems.cfg_large(2025) and
checked against the stored SHA-256 of their bytes.

    python tests/golden/make_spce_nacl_32k.py     # writes tests/golden/spce_nacl_32k.npz (~6 min, 2 processes)
"""
from __future__ import annotations

import hashlib
import multiprocessing as mp
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

OUT = os.path.join(HERE, "spce_nacl_32k.npz")
SEED = 2025
DISPLACE_SEED = 2026
DISPLACE = 0.25


def sha(a: np.ndarray) -> str:
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


def list_digest(a: np.ndarray, b: np.ndarray) -> str:
    return hashlib.sha256(np.ascontiguousarray(a, "<i4").tobytes() +
                          np.ascontiguousarray(b, "<i4").tobytes()).hexdigest()


def displaced(pos: np.ndarray) -> np.ndarray:
    return pos + np.random.default_rng(DISPLACE_SEED).uniform(-DISPLACE, DISPLACE, pos.shape)


def one_call(call: int):
    from mdynamix_amd import abi, systems, topology
    from oracle import orc
    s = systems.cfg_large(SEED)
    topo, prm = topology.from_system(s)
    pos = s.pos if call == 0 else displaced(s.pos)
    ts, ps = topo.struct(), topology.params_struct(dict(prm, list_mode=0, skin=0.0))
    t0 = time.time()
    lists = orc.chcnb(ts, ps, pos, cap_short=8_000_000, cap_long=60_000_000)
    kv = orc.inifur(ts, ps, topo.nmoltypes)
    f, g, h, e = orc.compute(topo, ts, ps, pos, lists, abi.MDX_FORCES_ALL, kvec=kv)
    print("call %d: %d + %d pairs, %d k vectors, %.0f s" % (call, len(lists[0]), len(lists[2]), len(kv[0]),
                                                        time.time() - t0), flush=True)
    out = {"ENERGY": np.array([e[k] for k in abi.MdxEnergies.SCALARS]),
           "NPAIRS": np.array([e["npairs_short"], e["npairs_long"]], np.int64), "F": f}
    if call == 0:
        out.update({"G": g.astype(np.float32), "H": h.astype(np.float32),
                    "LISTLEN": np.array([len(lists[0]), len(lists[2])], np.int64),
                    "DIGEST_SHORT": np.array(list_digest(lists[0], lists[1])),
                    "DIGEST_LONG": np.array(list_digest(lists[2], lists[3])),
                    "NKV": np.array(len(kv[0]), np.int64),
                    "KSET_SHA": np.array(sha(np.stack(kv[:3], 1).astype("<f8"))),
                    "POS_SHA": np.array(sha(s.pos.astype("<f8")))})
    else:
        out["POS_SHA"] = np.array(sha(pos.astype("<f8")))
    return out


def main() -> None:
    with mp.Pool(2) as pool:
        c0, c1 = pool.map(one_call, [0, 1])
    d = {"SEED": np.array(SEED), "DISPLACE_SEED": np.array(DISPLACE_SEED), "DISPLACE": np.array(DISPLACE)}
    for k, v in c0.items():
        d[k + "_0"] = v
    for k, v in c1.items():
        d[k + "_1"] = v
    np.savez_compressed(OUT, **d)
    print("wrote", OUT, os.path.getsize(OUT), "bytes")


if __name__ == "__main__":
    main()
