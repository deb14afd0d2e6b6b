"""Reader for the binary dump written by the reference harness
(oracle/ref/mdx_dump.f) and driver that runs the reference program on a
generated system.  Test infrastructure: used only by tests/, by the
fixture generator under tests/golden/ and by the cpu_baseline leg of
bench.py (time_reference_stripe).

Record layout: 8-byte name, int32 kind (1 = int32, 2 = float64),
int32 count, count values.  Names that occur twice (per-call arrays of the
second force evaluation) are returned as lists.
"""
from __future__ import annotations

import os
import subprocess
import tempfile
from typing import Dict, List

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
REF_EXE = os.path.join(HERE, "_ref", "mdx_ref")


def read_dump(path: str) -> Dict[str, List[np.ndarray]]:
    raw = open(path, "rb").read()
    out: Dict[str, List[np.ndarray]] = {}
    off = 0
    while off < len(raw):
        name = raw[off:off + 8].decode().strip()
        kind, n = np.frombuffer(raw, dtype="<i4", count=2, offset=off + 8)
        off += 16
        dt = np.dtype("<i4") if kind == 1 else np.dtype("<f8")
        arr = np.frombuffer(raw, dtype=dt, count=int(n), offset=off).copy()
        off += int(n) * dt.itemsize
        out.setdefault(name, []).append(arr)
    return out


def run_reference(system, workdir: str | None = None, timeout: float = 600.0) -> Dict[str, List[np.ndarray]]:
    """Write the system's input files, run the reference harness in a
    scratch directory and return its dump."""
    if not os.path.exists(REF_EXE):
        raise FileNotFoundError("reference harness not built: run make -C oracle/ref")
    own = workdir is None
    d = tempfile.mkdtemp(prefix="mdxref_") if own else workdir
    system.write_reference_inputs(d)
    res = subprocess.run([REF_EXE, "md.in"], cwd=d, capture_output=True, text=True, timeout=timeout)
    dump = os.path.join(d, "mdx_dump.bin")
    if not os.path.exists(dump):
        raise RuntimeError("reference run failed:\n" + res.stdout[-4000:] + res.stderr[-2000:])
    out = read_dump(dump)
    out["_stdout"] = [np.array(res.stdout)]
    return out


TRAJ_EXE = os.path.join(HERE, "_ref", "mdx_traj")


def run_trajectory(system, nsteps: int, workdir: str | None = None, timeout: float = 600.0):
    """Run the reference integrator (oracle/ref/mdx_traj.f around the
    reference SULTAN) for nsteps leap-frog steps; returns the dump with the
    start state, the end state and the last energies."""
    if not os.path.exists(TRAJ_EXE):
        raise FileNotFoundError("trajectory harness not built: run make -C oracle/ref")
    own = workdir is None
    d = tempfile.mkdtemp(prefix="mdxtrj_") if own else workdir
    system.write_reference_inputs(d, nsteps=nsteps)
    res = subprocess.run([TRAJ_EXE, "md.in"], cwd=d, capture_output=True, text=True, timeout=timeout)
    dump = os.path.join(d, "mdx_traj.bin")
    if not os.path.exists(dump):
        raise RuntimeError("reference run failed:\n" + res.stdout[-4000:] + res.stderr[-2000:])
    out = read_dump(dump)
    out["_stdout"] = [np.array(res.stdout)]
    return out


PIMD_EXE = os.path.join(HERE, "_ref", "mdx_pimd")


def run_pimd(system, nsteps: int, workdir: str | None = None, timeout: float = 600.0) -> List[Dict]:
    """Run the reference path-integral integrator (oracle/ref/mdx_pimd.f
    around the reference PIMD) as system.pimd_beads MPI ranks, one bead per
    rank, for nsteps long steps; returns one dump per bead (start state,
    end state, thermostat chains, last energies)."""
    if not (os.path.exists(PIMD_EXE) and os.path.exists(MPIEXEC)):
        raise FileNotFoundError("PIMD harness not built (make -C oracle/ref pimd) or mpiexec missing")
    d = tempfile.mkdtemp(prefix="mdxpi_") if workdir is None else workdir
    system.write_reference_inputs(d, nsteps=nsteps)
    res = subprocess.run([MPIEXEC, "-n", str(int(system.pimd_beads)), PIMD_EXE, "md.in"], cwd=d,
                         capture_output=True, text=True, timeout=timeout)
    out = []
    for r in range(system.pimd_beads):
        path = os.path.join(d, "mdx_pimd_%d.bin" % r)
        if not os.path.exists(path):
            raise RuntimeError("reference PIMD run failed:\n" + res.stdout[-4000:] + res.stderr[-2000:])
        out.append(read_dump(path))
    out[0]["_stdout"] = [np.array(res.stdout)]
    return out


TIME_EXE = os.path.join(HERE, "_ref", "mdx_time")
TIME_MPI_EXE = os.path.join(HERE, "_ref", "mdx_time_mpi")
MPIEXEC = "/opt/conda/bin/mpiexec"


def _parse_time(res, t_total) -> Dict:
    for line in res.stdout.splitlines():
        f = line.split()
        if len(f) == 6 and f[0] == "MDXTIME":
            return {"ranks": int(f[1]), "mbsh": int(f[2]), "mbln": int(f[3]), "t_list": float(f[4]),
                    "t_force": float(f[5]), "t_total": t_total}
    raise RuntimeError("reference timing run failed:\n" + res.stdout[-2000:] + res.stderr[-1000:])


def time_reference_mpi(system, ranks: int, workdir: str | None = None, timeout: float = 300.0) -> Dict:
    """Run the reference's MPI build (timing harness linked to mpi_double.f
    and MPICH) as `ranks` processes on the host cores: its own CHCNB and
    one ALLFORCE with the replicated-data reductions, timed between
    barriers on rank 0.  Returns {"ranks", "t_list", "t_force", "mbsh",
    "mbln" (rank 0's list lengths), "t_total"}."""
    import time
    if not (os.path.exists(TIME_MPI_EXE) and os.path.exists(MPIEXEC)):
        raise FileNotFoundError("reference MPI timing binary or mpiexec missing")
    d = tempfile.mkdtemp(prefix="mdxmpi_") if workdir is None else workdir
    system.write_reference_inputs(d)
    t0 = time.perf_counter()
    res = subprocess.run([MPIEXEC, "-n", str(int(ranks)), TIME_MPI_EXE, "md.in"], cwd=d, capture_output=True,
                         text=True, timeout=timeout, env=dict(os.environ, MDX_TIME_STRIPES="0"))
    return _parse_time(res, time.perf_counter() - t0)


def time_reference_stripe(system, stripes: int, workdir: str | None = None, timeout: float = 300.0) -> Dict:
    """Run the reference (timing build of the dump harness, -O3) on rank 0's
    share of `stripes` replicated-data ranks: its own CHCNB and one
    ALLFORCE, timed inside the program with the wall clock.  Returns
    {"ranks", "t_list", "t_force", "mbsh", "mbln", "t_total"}; raises if the binary
    is missing or the run fails."""
    import time
    if not os.path.exists(TIME_EXE):
        raise FileNotFoundError("reference timing binary not built: run make -C oracle/ref")
    d = tempfile.mkdtemp(prefix="mdxtime_") if workdir is None else workdir
    system.write_reference_inputs(d)
    t0 = time.perf_counter()
    res = subprocess.run([TIME_EXE, "md.in"], cwd=d, capture_output=True, text=True, timeout=timeout,
                         env=dict(os.environ, MDX_TIME_STRIPES=str(int(stripes))))
    return _parse_time(res, time.perf_counter() - t0)

