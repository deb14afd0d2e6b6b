"""Python client of the engine's serving loop (mdx_serve_* in the C ABI).

An `MDJob` holds everything one MD segment needs (topology tables,
parameters, MD parameters, start state, output arrays) and keeps those
arrays alive until its result has been polled, as the C ABI requires.
`Server` submits jobs to the native worker, which admits them by HBM
budget and advances all admitted jobs together with fused launches.
"""
from __future__ import annotations

import ctypes as C
from typing import Dict, List, Optional

import numpy as np

from . import abi, topology
from .abi import as_ptr
from .engine import MdxError, load_library

_pd = C.POINTER(C.c_double)


class MdxJobDesc(C.Structure):
    _fields_ = [
        ("topology", C.c_void_p), ("params", C.c_void_p), ("md", C.c_void_p),
        ("x", _pd), ("y", _pd), ("z", _pd), ("vx", _pd), ("vy", _pd), ("vz", _pd),
        ("nsteps", C.c_int32),
        ("x_out", _pd), ("y_out", _pd), ("z_out", _pd), ("vx_out", _pd), ("vy_out", _pd), ("vz_out", _pd),
    ]


class MdxJobResult(C.Structure):
    _fields_ = [
        ("id", C.c_int64), ("status", C.c_int32), ("nsteps", C.c_int32),
        ("t_submit", C.c_double), ("t_admit", C.c_double), ("t_done", C.c_double),
        ("epot", C.c_double), ("ekin", C.c_double), ("temperature", C.c_double),
        ("pair_interactions", C.c_int64), ("bytes", C.c_int64), ("message", C.c_char * 128),
    ]


def _lib():
    L = load_library()
    if not getattr(L, "_serve_bound", False):
        pv = C.c_void_p
        L.mdx_serve_create.argtypes = [C.c_int64, C.c_int32, C.c_int32, C.POINTER(pv)]
        L.mdx_serve_submit.argtypes = [pv, C.POINTER(MdxJobDesc), C.POINTER(C.c_int64)]
        L.mdx_serve_poll.argtypes = [pv, C.POINTER(MdxJobResult), C.c_int32, C.POINTER(C.c_int32),
                                     C.POINTER(C.c_int32), C.POINTER(C.c_int32)]
        L.mdx_serve_wait.argtypes = [pv, C.c_double]
        L.mdx_serve_stats.argtypes = [pv, C.POINTER(C.c_int64), _pd, _pd]
        L.mdx_serve_last_error.argtypes = [pv]
        L.mdx_serve_last_error.restype = C.c_char_p
        L.mdx_serve_destroy.argtypes = [pv]
        L.mdx_estimate_bytes.argtypes = [pv, pv, C.c_int32, C.POINTER(C.c_int64)]
        for n in ("mdx_serve_create", "mdx_serve_submit", "mdx_serve_poll", "mdx_serve_wait", "mdx_serve_stats",
                  "mdx_serve_destroy", "mdx_estimate_bytes"):
            getattr(L, n).restype = C.c_int
        L._serve_bound = True
    return L


class MDJob:
    """One MD segment of a generated system (kJ/mol, A, fs units)."""

    def __init__(self, system, nsteps: int, velocities: np.ndarray, lam: Optional[float] = None,
                 md_overrides: Optional[Dict] = None):
        self.system = system
        self.nsteps = int(nsteps)
        self.topo, self.prm = topology.from_system(system, lam)
        self.ts = self.topo.struct()
        self.ps = abi.make_params(**self.prm)
        md = topology.md_params(system)
        md.update(md_overrides or {})
        self.md = md
        self.mdp, self._keep = abi.make_md_params(md)
        self.pos = [np.ascontiguousarray(system.pos[:, k], dtype=np.float64) for k in range(3)]
        self.vel = [np.ascontiguousarray(velocities[:, k], dtype=np.float64) for k in range(3)]
        self.out = np.zeros((6, system.natoms))
        d = MdxJobDesc()
        d.topology = C.cast(C.pointer(self.ts), C.c_void_p)
        d.params = C.cast(C.pointer(self.ps), C.c_void_p)
        d.md = C.cast(C.pointer(self.mdp), C.c_void_p)
        d.x, d.y, d.z = [as_ptr(a) for a in self.pos]
        d.vx, d.vy, d.vz = [as_ptr(a) for a in self.vel]
        d.nsteps = self.nsteps
        d.x_out, d.y_out, d.z_out, d.vx_out, d.vy_out, d.vz_out = [as_ptr(self.out[k]) for k in range(6)]
        self.desc = d
        self.natoms = system.natoms

    def estimate_bytes(self) -> int:
        b = C.c_int64()
        st = _lib().mdx_estimate_bytes(C.byref(self.ts), C.byref(self.ps), 1, C.byref(b))
        if st != 0:
            raise MdxError(st, "estimate failed")
        return b.value

    def final_state(self):
        return self.out[0:3].T.copy(), self.out[3:6].T.copy()


class Server:
    def __init__(self, hbm_budget_bytes: int, max_slots: int = 16, quantum: int = 10):
        self.lib = _lib()
        s = C.c_void_p()
        st = self.lib.mdx_serve_create(int(hbm_budget_bytes), int(max_slots), int(quantum), C.byref(s))
        if st != 0:
            raise MdxError(st, "server creation failed")
        self._s = s
        self.jobs: Dict[int, MDJob] = {}

    def submit(self, job: MDJob) -> int:
        jid = C.c_int64()
        st = self.lib.mdx_serve_submit(self._s, C.byref(job.desc), C.byref(jid))
        if st != 0:
            raise MdxError(st, self.lib.mdx_serve_last_error(self._s).decode())
        self.jobs[jid.value] = job
        return jid.value

    def poll(self) -> List[Dict]:
        buf = (MdxJobResult * 64)()
        n, act, q = C.c_int32(), C.c_int32(), C.c_int32()
        out = []
        while True:
            self.lib.mdx_serve_poll(self._s, buf, 64, C.byref(n), C.byref(act), C.byref(q))
            for i in range(n.value):
                r = buf[i]
                out.append({f: (getattr(r, f).decode() if f == "message" else getattr(r, f))
                            for f, _ in MdxJobResult._fields_})
            if n.value < 64:
                break
        self.active, self.queued = act.value, q.value
        return out

    def wait(self, timeout_s: float = 600.0) -> bool:
        return self.lib.mdx_serve_wait(self._s, float(timeout_s)) == 0

    def stats(self) -> Dict:
        r, busy, clk = C.c_int64(), C.c_double(), C.c_double()
        self.lib.mdx_serve_stats(self._s, C.byref(r), C.byref(busy), C.byref(clk))
        return {"rounds": r.value, "busy_s": busy.value, "clock_s": clk.value}

    def close(self) -> None:
        if self._s and self._s.value:
            self.lib.mdx_serve_destroy(self._s)
            self._s = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
