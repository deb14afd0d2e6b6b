"""Benchmark of the MI355X force engine on the 32k SPC/E water + NaCl box.

Workload: 32,000 SPC/E waters with 87 Na+ and 87 Cl- (96,174 atoms) in a
98.7 A cubic box, R_cut 12 A, Verlet skin 1 A with a list rebuild every 10
steps, Ewald alpha*R_cut = 3.2 and FEXP = 11.5 (48,052 k-vectors).  One
step = one full force call (ALLFORCE: real-space pairs, reciprocal sum,
exclusion corrections, bonded terms, every energy and virial scalar copied
back) plus the list rebuild on every tenth step, on coordinates resident
in HBM.  Synthetic seeded coordinates (jittered lattice, random
orientations); nothing is loaded from disk.

Prints one JSON line (metric pair-interactions/s; ns/day, per-stage times,
roofline of the dominant kernel and the CPU baseline beside it: the
reference's own code, built from its sources by oracle/ref/Makefile, on
one host core and a 1/12 replicated-data stripe, with the FP64 port on the
same stripe beside it; the port alone when the reference build is
absent), plus the
batched form (`batched`: 32 path-integral beads x 12,000 atoms per fused
force call), the multiple-time-step MD segment (`respa`: 47,776-atom
chains-in-water box, 2 fs long steps of 10 short steps, ms per long step
and ns/day) and the served form (`served`: a Poisson stream of MD-segment
jobs of 1.5k-60k atoms with continuous batching, p50/p99 latency and
aggregate pair-interactions/s against one job at a time).

    python bench.py [--steps K] [--warmup W] [--no-cpu] [--no-batched] [--no-respa] [--no-served]
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

FP32_PEAK_TFLOPS = 157.3     # MI355X vector FP32 (published)
F16_MFMA_PEAK_TFLOPS = 2500.0  # MI355X dense f16 matrix (published)
HBM_PEAK_GBS = 8000.0        # MI355X HBM3E
PAIR_FLOP = 74               # per pair within R_cut (forces.f:1071-1198 operation count)
RECIP_FLOP = 21              # per (k, atom) term (forces.f:1288-1325 operation count)


def measured_traffic(kernel: str):
    """HBM bytes per launch of `kernel` from the committed rocprofv3 counter
    summary (lines 'traffic_bytes <kernel> <bytes>'), or None."""
    path = os.path.join(ROOT, "profiles", "traffic_spce_nacl_32k.md")
    if not os.path.exists(path):
        return None
    rows = [line.split() for line in open(path)]
    rows = [(" ".join(f[1:-1]), int(f[-1])) for f in rows if len(f) >= 3 and f[0] == "traffic_bytes"]
    for name, v in rows:
        if name == kernel:
            return v
    # the pair kernel at another occupancy (same work and bytes)
    if kernel.startswith("k_pairs<false, false"):
        for name, v in rows:
            if name.startswith("k_pairs<false, false"):
                return v
    return None


def measured_mfma(kernel: str):
    """Matrix instructions per launch of `kernel` (SQ_INSTS_MFMA) from the
    committed instruction-counter summary, or None."""
    path = os.path.join(ROOT, "profiles", "counters_spce_nacl_32k.md")
    if not os.path.exists(path):
        return None
    names = None
    for line in open(path):
        cells = [c.strip() for c in line.strip().strip("|").split("|")]
        if cells and cells[0] == "counter":
            names = cells[1:]
        elif names and cells and cells[0] == "SQ_INSTS_MFMA":
            for name, v in zip(names, cells[1:]):
                if name == kernel:
                    return float(v)
    return None


def build_engine(seed: int = 2025):
    from mdynamix_amd import systems, topology
    from mdynamix_amd.engine import ForceEngine
    s = systems.cfg_large(seed)
    topo, prm = topology.from_system(s)
    eng = ForceEngine(topo, prm)
    eng.set_positions(s.pos)
    return s, topo, prm, eng


def cpu_reference(s, numtask: int = 12):
    """The reference itself (its forces.f, main.f set-up and serial MPI
    stubs compiled from source by oracle/ref/Makefile with amdflang -O3,
    no fast-math) on one host core: CHCNB and one ALLFORCE of rank 0 of
    `numtask` replicated-data ranks (the reference's own striding:
    forces.f:1825-1830, 1273-1277, 1501-1502), wall-clocked inside the
    program and scaled by numtask to a full force call, with the list
    amortised over its 10-step rebuild interval."""
    from oracle import refdump
    r = refdump.time_reference_stripe(s, numtask, timeout=180.0)
    t_call = numtask * r["t_force"] + numtask * r["t_list"] / 10.0
    pairs_call = (r["mbsh"] + r["mbln"]) * numtask
    return {
        "value": pairs_call / t_call,
        "unit": "pair-interactions/s",
        "cores": 1,
        "kind": "reference",
        "ns_per_day": (1.0 / t_call) * 2.0 * 0.0864,
        "seconds_per_step": t_call,
        "sample": ("reference forces.f (amdflang -O3 serial build, oracle/_ref/mdx_time) on rank 0 of %d "
                   "replicated-data stripes of the same 96,174-atom box, one core: CHCNB %.2f s "
                   "(%d + %d listed pairs), ALLFORCE %.2f s (%.1f s sampled, %.1f s with the program's "
                   "start-up), scaled x%d" % (numtask, r["t_list"], r["mbsh"], r["mbln"], r["t_force"],
                                            r["t_list"] + r["t_force"], r["t_total"], numtask)),
    }


def cpu_reference_mpi(s, ranks: int):
    """The reference's MPI build (its forces.f with mpi_double.f, amdflang
    -O3 against MPICH, oracle/_ref/mdx_time_mpi) as `ranks` processes on the
    host cores: one full CHCNB and ALLFORCE of the same system with its
    replicated-data reductions, wall-clocked between barriers on rank 0;
    the list amortised over its 10-step rebuild interval."""
    from oracle import refdump
    r = refdump.time_reference_mpi(s, ranks, timeout=180.0)
    t_call = r["t_force"] + r["t_list"] / 10.0
    pairs_call = (r["mbsh"] + r["mbln"]) * r["ranks"]
    return {
        "value": pairs_call / t_call,
        "unit": "pair-interactions/s",
        "cores": r["ranks"],
        "kind": "reference",
        "ns_per_day": (1.0 / t_call) * 2.0 * 0.0864,
        "seconds_per_step": t_call,
        "sample": ("reference MDynaMix MPI build (forces.f + mpi_double.f, amdflang -O3, MPICH; "
                   "oracle/_ref/mdx_time_mpi) as %d ranks on %d host cores: one full CHCNB %.2f s and "
                   "ALLFORCE %.2f s of the same 96,174-atom box (%.1f s with the program's start-up); "
                   "pairs = rank 0's %d + %d listed pairs x %d" %
                   (r["ranks"], r["ranks"], r["t_list"], r["t_force"], r["t_total"], r["mbsh"], r["mbln"],
                    r["ranks"])),
    }


def cpu_baseline(s, topo, prm, numtask: int = 12):
    """The FP64 oracle (a scalar C port of the reference force path) on one
    host core, on the share of one of `numtask` replicated-data ranks
    (CHCNB outer-atom stripe, forces.f:1825-1830; FURIR k stripe,
    forces.f:1273-1277), scaled by numtask to one full force call; the
    list build is amortised over its 10-step rebuild interval."""
    from mdynamix_amd import topology
    from oracle import orc
    import ctypes as C
    from mdynamix_amd.abi import MdxEnergies, as_ptr
    ts = topo.struct()
    ps = topology.params_struct(dict(prm, list_mode=0, skin=0.0))
    pos = s.pos
    L = orc.lib()
    t0 = time.perf_counter()
    s1, s2, l1, l2 = orc.chcnb(ts, ps, pos, cap_short=8_000_000, cap_long=40_000_000, taskid=0, numtask=numtask)
    t_list = time.perf_counter() - t0
    x, y, z = [np.ascontiguousarray(pos[:, k]) for k in range(3)]
    f = [np.zeros(len(pos)) for _ in range(3)]
    e = MdxEnergies()
    pi = C.c_int32
    t0 = time.perf_counter()
    L.orc_localf(C.byref(ts), C.byref(ps), as_ptr(x), as_ptr(y), as_ptr(z), as_ptr(s1, pi), as_ptr(s2, pi),
                 len(s1), *[as_ptr(a) for a in f], C.byref(e))
    L.orc_forces(C.byref(ts), C.byref(ps), as_ptr(x), as_ptr(y), as_ptr(z), as_ptr(l1, pi), as_ptr(l2, pi),
                 len(l1), *[as_ptr(a) for a in f], C.byref(e))
    t_pairs = time.perf_counter() - t0
    rkx, rky, rkz, spee, spei = orc.inifur(ts, ps, topo.nmoltypes)
    t0 = time.perf_counter()
    L.orc_furir_task(C.byref(ts), C.byref(ps), as_ptr(x), as_ptr(y), as_ptr(z), as_ptr(rkx), as_ptr(rky),
                     as_ptr(rkz), len(rkx), 0, numtask, *[as_ptr(a) for a in f], C.byref(e))
    t_recip = time.perf_counter() - t0
    t_call = numtask * (t_pairs + t_recip) + numtask * t_list / 10.0
    pairs_call = (len(s1) + len(l1)) * numtask
    sampled = t_list + t_pairs + t_recip
    return {
        "value": pairs_call / t_call,
        "unit": "pair-interactions/s",
        "cores": 1,
        "kind": "port",
        "ns_per_day": (1.0 / t_call) * 2.0 * 0.0864,
        "seconds_per_step": t_call,
        "sample": ("FP64 oracle (scalar C port of forces.f, gcc -O2, no FMA) on rank 0 of %d replicated-data "
                   "stripes of the same 96,174-atom box: CHCNB stripe %.2f s, LOCALF+FORCES on its %d pairs "
                   "%.2f s, FURIR on %d of %d k-vectors %.2f s (%.1f s sampled), scaled x%d" %
                   (numtask, t_list, len(s1) + len(l1), t_pairs, len(range(numtask - 1, len(rkx), numtask)),
                    len(rkx), t_recip, sampled, numtask)),
    }


def bench_detail(eng, calls: int) -> dict:
    """The call the Fortran drop-in makes at every step (ALLFORCE with
    MDX_WANT_DETAIL, mdx_forces.f): forces F, G, H and the per-type energy
    breakdown copied back to host arrays, so the time includes the PCIe
    copies of 9 x 96,174 doubles."""
    import ctypes as C
    from mdynamix_amd import abi
    from mdynamix_amd.abi import MdxEnergies, as_ptr
    # the caller's arrays, allocated once as a Fortran program holds them
    out = np.zeros((9, eng.n))
    ptrs = [as_ptr(out[k]) for k in range(9)]
    e = MdxEnergies()

    def timed(want):
        def call():
            st = eng.lib.mdx_compute(eng._h, abi.MDX_FORCES_ALL, want, *ptrs, C.byref(e))
            if st != 0:
                raise RuntimeError("mdx_compute failed: %d" % st)
        call()
        t0 = time.perf_counter()
        for _ in range(calls):
            call()
        return 1e3 * (time.perf_counter() - t0) / calls
    return {"calls": calls, "ms_per_call": timed(abi.MDX_WANT_DETAIL), "ms_per_call_no_detail": timed(0),
            "what": "mdx_compute(MDX_FORCES_ALL, MDX_WANT_DETAIL) into host F, G, H arrays (PCIe included); "
                    "beside it the same host-array call without the breakdown"}


def bench_md(s, topo, prm, steps: int, warmup: int) -> dict:
    """MD on the device at the headline workload: the reference's
    single-time-step leap-frog with SHAKE (SULTAN, mdstep.f:29-260) for the
    rigid SPC/E waters, Nose-Hoover at 300 K, list rebuild every 10 steps,
    each step a full force call; between rebuilds a step is one captured
    hipGraph.  ns/day from the system's 2 fs step."""
    from mdynamix_amd import systems, topology
    from mdynamix_amd.engine import ForceEngine
    eng = ForceEngine(topo, prm)
    eng.set_positions(s.pos)
    eng.set_timing(False)
    md = topology.md_params(s)
    eng.md_init(md, systems.maxwell_velocities(s, 41))
    eng.md_run(warmup)
    t0 = time.perf_counter()
    tr = eng.md_run(steps)
    t = time.perf_counter() - t0
    fails = eng.md_shake_failures()
    eng.close()
    return {"workload": "spce_nacl_32k", "integrator": "SULTAN + SHAKE (single time step)",
            "dt_fs": s.dt_fs, "steps": steps, "ms_per_step": 1e3 * t / steps,
            "ns_per_day": (steps / t) * s.dt_fs * 0.0864, "final_temperature_K": float(tr[-1, 2]),
            "shake_failures": int(fails)}


def bench_fp64(s, topo, prm, calls: int, f_mixed) -> dict:
    """The same box in the FP64 precision mode (MDX_PRECISION_FP64): time per
    force call (forces left in HBM, energies read back, as the headline
    steps) and the mixed path's force RMS relative difference from it."""
    from mdynamix_amd import abi
    from mdynamix_amd.engine import ForceEngine
    eng = ForceEngine(topo, dict(prm, precision=abi.MDX_PRECISION_FP64))
    eng.set_positions(s.pos)
    eng.chcnb()
    f64 = eng.allforce()[0]
    eng.set_timing(True)
    for _ in range(2):
        eng.compute_device(0, 0, 0, energies=True)
    st = {}
    pairs = 0
    t0 = time.perf_counter()
    for _ in range(calls):
        e = eng.compute_device(0, 0, 0, energies=True)
        pairs += e["npairs_short"] + e["npairs_long"]
        for k, v in eng.stage_times().items():
            if k != "list":
                st[k] = st.get(k, 0.0) + v
    t = time.perf_counter() - t0
    eng.close()
    rms = float(np.sqrt(((f_mixed - f64) ** 2).sum() / (f64 ** 2).sum()))
    return {"calls": calls, "ms_per_call": 1e3 * t / calls, "pair_interactions_per_s": pairs / t,
            "stages_ms": {"pairs": round(st.get("pairs", 0.0) / calls, 4),
                          "recip_fp64": round(st.get("recip_force", 0.0) / calls, 4)},
            "mixed_vs_fp64_force_rms": rms,
            "what": "FP64 pair kernel (k_pairs64) and FP64 FURIR; the mixed path's forces against these"}


def call_roofline(pairs_per_s: float, kterms_per_s: float) -> dict:
    """Whole-call FLOP rate of a batched or served form against the FP32
    peak: 74 FLOP per pair within R_cut and 21 FLOP per (k vector, atom)
    term of the reciprocal sum, per second of wall time."""
    flops = PAIR_FLOP * pairs_per_s + RECIP_FLOP * kterms_per_s
    return {"bound": "valu+mfma", "achieved": round(flops / 1e12, 3), "peak": FP32_PEAK_TFLOPS,
            "unit": "TFLOP/s", "frac": round(flops / 1e12 / FP32_PEAK_TFLOPS, 4),
            "note": "whole force calls per wall second (pairs x 74 FLOP + k terms x 21 FLOP)"}


def bench_batched(steps: int, warmup: int, beads: int = 32, n_water: int = 4000) -> dict:
    """Path-integral batch: `beads` replicas of an SPC-FW box (4,000 waters,
    49.3 A) evaluated as one ring batch per force call (every kernel one
    fused launch over all beads, replayed from a hipGraph), list rebuild
    every 10 calls; beside it the same beads evaluated one at a time."""
    from mdynamix_amd import systems, topology
    from mdynamix_amd.engine import ForceEngine, ForceBatch
    bs = systems.pimd_beads(n_beads=beads, n_water=n_water, box=49.3, seed=2027)
    topo, prm = topology.from_system(bs[0])
    engs = []
    for b in bs:
        e = ForceEngine(topo, prm)
        e.set_positions(b.pos)
        e.set_timing(False)
        e.chcnb()
        engs.append(e)
    ring = ForceBatch(engs, ring_beadfac=1e-4)

    def run(k0, n, batched):
        pairs = 0
        for k in range(k0, k0 + n):
            if k % 10 == 0:
                for e in engs:
                    e.chcnb()
            if batched:
                es = ring.compute(energies=True)
            else:
                es = [e.compute_device(0, 0, 0, energies=True) for e in engs]
            pairs += sum(x["npairs_short"] + x["npairs_long"] for x in es)
        return pairs

    run(0, warmup, True)
    t0 = time.perf_counter()
    pairs = run(warmup, steps, True)
    tb = time.perf_counter() - t0
    run(0, warmup, False)
    t0 = time.perf_counter()
    pairs_a = run(warmup, steps, False)
    ta = time.perf_counter() - t0
    ring.close()
    pimd = bench_pimd_md(bs, engs, topo, max(steps // 2, 5), warmup)
    for e in engs:
        e.close()
    kterms = _count_k(prm) * topo.natoms * beads * steps
    return {"workload": "pimd_%dx%d_spcfw" % (beads, n_water), "beads": beads, "atoms_per_bead": topo.natoms,
            "steps": steps, "ms_per_call": 1e3 * tb / steps, "pair_interactions_per_s": pairs / tb,
            "one_at_a_time": {"ms_per_call": 1e3 * ta / steps, "pair_interactions_per_s": pairs_a / ta},
            "speedup": ta / tb, "roofline": call_roofline(pairs / tb, kterms / tb), "pimd_md": pimd}


def bench_pimd_md(bs, engs, topo, long_steps: int, warmup: int) -> dict:
    """The same beads advanced by the device path-integral loop (PIMD,
    pimd.f:18-314): 1 fs long steps of 4 short steps, bead masses m / P,
    ring springs of the physical BEADFAC at 300 K, BEADNOSE chains (5 fs),
    centroid GETCOM, list rebuild every 10 long steps; every short step is
    a fused force call over all beads."""
    from mdynamix_amd import abi, systems, topology
    from mdynamix_amd.engine import ForceBatch
    P = len(engs)
    md = topology.md_params(bs[0])
    md.update(integrator=abi.MDX_MD_PIMD, nfreq=4, dt=1.0, tau_t=0.0, rebuild_every=10)
    # Maxwell velocities of the bead mass m / P at 300 K (sqrt(P) x those of m)
    for k, (e, b) in enumerate(zip(engs, bs)):
        e.set_positions(b.pos)
        e.md_init(md, systems.maxwell_velocities(b, 100 + k) * np.sqrt(P))
    ring = ForceBatch(engs, ring_beadfac=topology.pimd_beadfac(300.0, P))
    ring.pimd_init(temperature=300.0, tau_bead=5.0, nchain=1)
    ring.md_run(warmup)
    t0 = time.perf_counter()
    tr = ring.md_run(long_steps)
    t = time.perf_counter() - t0
    ring.close()
    ms = 1e3 * t / long_steps
    return {"integrator": "PIMD (beads' multiple time step, BEADNOSE)", "dt_fs": 1.0, "nfreq": 4,
            "long_steps": long_steps, "ms_per_long_step": ms, "ns_per_day": 86400.0 / (ms * 1e-3) * 1e-6,
            "final_mean_bead_temperature_K": float(np.mean(tr[:, -1, 2])) if tr is not None else None}


def bench_respa(long_steps: int, warmup: int) -> dict:
    """Multiple-time-step MD segment on the device (DOUBLE, mdstep.f:281-530):
    256 flexible 46-site chains in 12,000 SPC/E waters (47,776 atoms, 80 A,
    R_cut 12 A, R_short 5 A, skin 1 A), 2 fs long steps of 10 short steps
    (9 fast force calls and one full call per long step), Nose-Hoover at
    300 K, list rebuild every 10 long steps; each long step is one captured
    hipGraph between rebuilds."""
    from mdynamix_amd import systems, topology
    from mdynamix_amd.engine import ForceEngine
    s = systems.cfg_chains()
    topo, prm = topology.from_system(s)
    eng = ForceEngine(topo, prm)
    eng.set_positions(s.pos)
    eng.set_timing(False)
    md = topology.md_params(s)
    eng.md_init(md, systems.maxwell_velocities(s, 31))
    eng.md_run(warmup)
    t0 = time.perf_counter()
    tr = eng.md_run(long_steps)
    t = time.perf_counter() - t0
    eng.close()
    return {"workload": "chains_256x46_spce_12000", "atoms": topo.natoms, "integrator": "DOUBLE (r-RESPA)",
            "dt_fs": s.dt_fs, "nfreq": md["nfreq"], "long_steps": long_steps,
            "ms_per_long_step": 1e3 * t / long_steps,
            "ns_per_day": (long_steps / t) * s.dt_fs * 0.0864,
            "final_temperature_K": float(tr[-1, 2])}


def bench_served(n_jobs: int, nsteps: int, rate: float, seed: int = 77) -> dict:
    """Served mixed-size stream: `n_jobs` MD segments of `nsteps` steps
    (SPC/E + 0.15 M NaCl boxes of 1.5k-60k atoms, log-uniform, one rigid
    10-60-site solute, at most n/50 sites, at lambda in {0, 0.05, ..., 1}),
    Poisson arrivals at
    `rate` jobs/s, served with continuous batching (up to 16 slots, 64 GiB
    HBM budget, 10-step rounds) and, on the same arrival times, one job at
    a time; latency = completion - submission."""
    import threading
    from mdynamix_amd import systems
    from mdynamix_amd.serve import MDJob, Server
    from mdynamix_amd import topology
    rng = np.random.default_rng(seed)
    jobs = []
    kterms = 0.0
    # sizes: log-spaced over 1.5k-60k atoms (the generator's atom count is
    # corrected once toward each target), in a shuffled arrival order
    targets = np.geomspace(1500, 60000, n_jobs)[rng.permutation(n_jobs)]
    for j in range(n_jobs):
        n = int(targets[j])
        # solute size capped at n/50 so it spans less than half the box
        ns = min(int(rng.integers(10, 61)), max(10, n // 50))
        lam = float(rng.integers(0, 21)) * 0.05
        s = systems.ee_job(n, ns, lam, seed + j, rigid_solute=True)
        if s.pos.shape[0] != n:
            s = systems.ee_job(int(round(n * n / s.pos.shape[0])), ns, lam, seed + j, rigid_solute=True)
        jobs.append(MDJob(s, nsteps, systems.maxwell_velocities(s, seed + j)))
        _, jprm = topology.from_system(s)
        kterms += float(_count_k(jprm)) * s.pos.shape[0] * nsteps
    arrivals = np.cumsum(rng.exponential(1.0 / rate, n_jobs))

    def serve(slots):
        srv = Server(hbm_budget_bytes=64 << 30, max_slots=slots, quantum=10)
        t0 = time.perf_counter()
        for j, job in enumerate(jobs):
            dt = arrivals[j] - (time.perf_counter() - t0)
            if dt > 0:
                time.sleep(dt)
            srv.submit(job)
        res = []
        deadline = time.perf_counter() + 1800.0
        while len(res) < n_jobs and time.perf_counter() < deadline:
            srv.wait(15.0)
            res += srv.poll()
            print("served %d/%d jobs (slots %d)" % (len(res), n_jobs, slots), file=sys.stderr, flush=True)
        st = srv.stats()
        srv.close()
        lat = np.array([r["t_done"] - r["t_submit"] for r in res])
        span = max(r["t_done"] for r in res) - min(r["t_submit"] for r in res)
        pairs = sum(r["pair_interactions"] for r in res)
        ok = all(r["status"] == 0 for r in res) and len(res) == n_jobs
        return {"p50_s": float(np.percentile(lat, 50)), "p99_s": float(np.percentile(lat, 99)),
                "mean_s": float(lat.mean()), "makespan_s": span, "pair_interactions_per_s": pairs / span,
                "rounds": st["rounds"], "all_ok": ok, "roofline": call_roofline(pairs / span, kterms / span)}

    served = serve(16)
    alone = serve(1)
    sizes = [j.natoms for j in jobs]
    return {"workload": "ee_stream_%djobs_%dsteps" % (n_jobs, nsteps), "jobs": n_jobs, "steps_per_job": nsteps,
            "arrival_rate_per_s": rate, "atoms_min": min(sizes), "atoms_max": max(sizes),
            "atoms_total": int(sum(sizes)), "served": served, "one_at_a_time": alone,
            "throughput_gain": served["pair_interactions_per_s"] / alone["pair_interactions_per_s"]}


def copy_calibration(mib: int, reps: int = 5) -> dict:
    """Device-to-device copies of a known size through the HIP runtime (its
    copy kernel), so hardware traffic counters read in a profiler can be
    scaled to bytes: each copy reads and writes `mib` MiB."""
    import ctypes as C
    hip = C.CDLL("libamdhip64.so")
    n = mib << 20
    a, b = C.c_void_p(), C.c_void_p()
    if hip.hipMalloc(C.byref(a), C.c_size_t(n)) or hip.hipMalloc(C.byref(b), C.c_size_t(n)):
        raise RuntimeError("hipMalloc failed")
    hip.hipMemset(a, 1, C.c_size_t(n))
    for _ in range(reps):
        hip.hipMemcpy(b, a, C.c_size_t(n), 3)          # hipMemcpyDeviceToDevice
    hip.hipDeviceSynchronize()
    hip.hipFree(a)
    hip.hipFree(b)
    return {"copy_bytes_read": n, "copy_bytes_written": n, "reps": reps}


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-tasks", type=int, default=12)
    ap.add_argument("--cpu-ranks", type=int, default=16,
                    help="MPI ranks of the reference CPU baseline (the host cores of one GPU's share)")
    ap.add_argument("--no-batched", action="store_true")
    ap.add_argument("--no-served", action="store_true")
    ap.add_argument("--no-respa", action="store_true")
    ap.add_argument("--serve-jobs", type=int, default=16)
    ap.add_argument("--serve-steps", type=int, default=1000)
    ap.add_argument("--serve-rate", type=float, default=4.0)
    ap.add_argument("--kspace-shape", default="", metavar="CHUNK,NSLICE[,GRID[,WAVES]]",
                    help="fixed k-space launch shape (and cluster grid scale, pair-kernel wavefronts "
                         "per SIMD) instead of autotuning (profiling runs)")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="only run device-to-device copies of MIB MiB (traffic counter calibration)")
    args = ap.parse_args()
    if args.copy_calibration:
        print(json.dumps(copy_calibration(args.copy_calibration)))
        return 0
    if args.gpus != 1:
        print("bench.py measures one GPU: this build has no multi-GPU path (one GPU per box)", file=sys.stderr)
        return 2

    from mdynamix_amd.engine import device_info
    s, topo, prm, eng = build_engine()
    dev = device_info()
    eng.chcnb()
    if args.kspace_shape:
        v = args.kspace_shape.split(",")
        c, n = int(v[0]), int(v[1])
        if len(v) > 2:
            eng.set_cluster_grid(float(v[2]))
        if len(v) > 3:
            eng.set_pair_waves(int(v[3]))
        eng.set_kspace_shape(c, n)
        tuned = {"chunk": c, "nslice": n, "column_width": eng.cluster_grid()["column_width"],
                 "pair_waves": eng.pair_waves(), "source": "--kspace-shape"}
    else:
        tuned = dict(eng.autotune(), source="autotune")
    print("k-space shape: %s" % tuned, file=sys.stderr, flush=True)
    rebuild = s.rebuild_every
    stage_sum = {}
    list_ms = []

    def step(k: int, record: bool):
        if k % rebuild == 0:
            eng.chcnb()
            if record:
                list_ms.append(eng.stage_times()["list"])
        e = eng.compute_device(eng_f[0], eng_f[1], eng_f[2], energies=True)
        if record:
            for kk, v in eng.stage_times().items():
                if kk != "list":
                    stage_sum[kk] = stage_sum.get(kk, 0.0) + v
        return e

    # forces stay in the engine's HBM buffer (no output pointers); every
    # step returns its energies, which synchronises the stream.  The timed
    # region is the production call: each force call forks into its three
    # stream chains (pairs, reciprocal sum, scattered terms) and is replayed
    # from the captured hipGraph; the list rebuild (every tenth step) runs
    # eagerly inside the region.
    eng_f = [0, 0, 0]
    eng.set_timing(False)
    for k in range(args.warmup):
        step(k, False)
    pairs = 0
    t0 = time.perf_counter()
    for k in range(args.steps):
        e = step(k, False)
        pairs += e["npairs_short"] + e["npairs_long"]
    t = time.perf_counter() - t0
    ms_step = 1e3 * t / args.steps
    value = pairs / t
    nsday = (args.steps / t) * s.dt_fs * 0.0864
    # the same steps launched one kernel after another on the engine's
    # stream with HIP events between the stages: each kernel's own duration
    # (the roofline below) and the stage breakdown
    eng.set_timing(True)
    for k in range(args.warmup):
        step(k, False)
    t0 = time.perf_counter()
    for k in range(args.steps):
        step(k, True)
    t_serial = time.perf_counter() - t0
    stages = {k: v / args.steps for k, v in stage_sum.items()}
    stages["list_per_rebuild"] = float(np.mean(list_ms)) if list_ms else 0.0
    # roofline of the dominant kernel
    e0 = e
    pairs_call = e0["npairs_short"] + e0["npairs_long"]
    kv_count = _count_k(prm)
    cand = {
        "k_pairs": (stages.get("pairs", 0.0), PAIR_FLOP * pairs_call),
        "k_sk": (stages.get("recip_sk", 0.0), RECIP_FLOP * kv_count * topo.natoms / 2.0),
        "k_recip_force": (stages.get("recip_force", 0.0), RECIP_FLOP * kv_count * topo.natoms / 2.0),
    }
    dom = max(cand, key=lambda k: cand[k][0])
    dms, dflop = cand[dom]
    achieved = dflop / (dms * 1e-3) / 1e12 if dms > 0 else 0.0
    if dom == "k_pairs":
        # the pair stage: k_pairs and the gather of its j-reaction records
        tp = measured_traffic("k_pairs<false, false, %d>" % tuned["pair_waves"])
        tg = measured_traffic("k_jgather")
        traffic = tp + tg if tp is not None and tg is not None else None
        dom = "k_pairs+k_jgather"
    else:
        # the instantiations this box runs (n_z up to 32 per row)
        traffic = measured_traffic({"k_sk": "k_sk<5>", "k_recip_force": "k_recip_force<32>"}[dom])
    entries, front = eng.list_entries()
    lanes = {"list_entries": entries, "front_entries": front, "pairs_in_cutoff": pairs_call,
             "lane_occupancy": round(pairs_call / (64.0 * entries), 4) if entries else None}
    # every candidate is compute-bound (a few hundred FLOP per HBM byte):
    # k_pairs on the vector ALU, the k-space passes on the matrix cores
    roof = {"kernel": dom, "bound": "valu" if dom.startswith("k_pairs") else "mfma",
            "achieved": round(achieved, 3), "peak": FP32_PEAK_TFLOPS,
            "unit": "TFLOP/s", "frac": round(achieved / FP32_PEAK_TFLOPS, 4), "traffic": traffic,
            "avg_ms": round(dms, 4), "algorithmic_flop": dflop,
            "lanes": lanes,
            "note": ("k_pairs+k_jgather: the pair stage (the half-list pair kernel and the per-atom sums of "
                     "its j-reaction records), 74 FLOP per pair within R_cut on the vector ALU (FP32, "
                     "unpacked: its own issue peak is 78.6 TFLOP/s; lanes.lane_occupancy: pairs within "
                     "R_cut per 64 lanes of the 8x8 cluster-pair entries); k_sk / k_recip_force: half of the 21 FLOP per (k, atom) "
                     "term of FURIR each, computed on the matrix cores as FP32 products split into f16 "
                     "hi/lo pairs (v_mfma_f32_16x16x32_f16, three products), FP64 sums; peak: "
                     "MI355X FP32 157.3 TFLOP/s (dense matrix or packed vector); traffic: HBM bytes per "
                     "launch from profiles/traffic_spce_nacl_32k.md (FETCH_SIZE/WRITE_SIZE, "
                     "copy-calibrated)")}
    # the k-space passes against the matrix cores' own peak, from the work
    # they execute: matrix instructions per launch (committed counters) x
    # 16,384 FLOP (v_mfma_f32_16x16x32_f16) over the stage's duration
    kspace = {}
    for stage, kern in (("recip_sk", "k_sk<5>"), ("recip_force", "k_recip_force<32>")):
        n_mfma, ms = measured_mfma(kern), stages.get(stage, 0.0)
        if n_mfma and ms > 0:
            tf = n_mfma * 16384.0 / (ms * 1e-3) / 1e12
            kspace[kern] = {"mfma_per_launch": n_mfma, "avg_ms": round(ms, 4), "achieved": round(tf, 1),
                            "peak": F16_MFMA_PEAK_TFLOPS, "unit": "TFLOP/s",
                            "frac": round(tf / F16_MFMA_PEAK_TFLOPS, 4),
                            "algorithmic_tflops": round(RECIP_FLOP * kv_count * topo.natoms / 2.0
                                                        / (ms * 1e-3) / 1e12, 1)}
    roof["kspace_matrix"] = kspace
    out = {
        "metric": "pair-interactions/s (32k SPC/E + NaCl, full force call incl. Ewald k-space)",
        "value": value,
        "unit": "pair-interactions/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_step,
        "ns_per_day": nsday,
        "timing": ("force calls forked into three stream chains and replayed from the captured hipGraph "
                   "(list rebuild every tenth step included); roofline and stages_ms from the serial pass"),
        "serial_stages": {"ms_per_step": 1e3 * t_serial / args.steps,
                          "pair_interactions_per_s": pairs / t_serial,
                          "what": "the same steps with every kernel in order on one stream and HIP events "
                                  "between the stages"},
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": ("fp32 pair and erfc arithmetic (FP32 sums over <= 16 list entries, then fp64 / 64-bit "
                  "fixed point); k space: fp32 phase products as f16 hi/lo pairs (22-bit) on the matrix "
                  "cores, each 32-atom stage's fp32 result summed in fp64 (structure factors), per-k-row "
                  "fp32 terms summed in fp64 (forces); fp64 tables, energies, virials"),
        "data": "synthetic (seeded jittered lattice, random orientations)",
        "config": {"workload": "spce_nacl_32k", "atoms": topo.natoms, "box_A": s.box[0], "rcut_A": s.rcut,
                   "skin_A": s.skin, "rebuild_every": rebuild, "kvectors": kv_count,
                   "pairs_per_call": pairs_call, "list": "verlet-cluster-8", "kspace_shape": tuned},
        "stages_ms": {k: round(v, 4) for k, v in stages.items()},
        "device": dev,
        "roofline": roof,
    }
    out["detail_call"] = bench_detail(eng, max(args.steps // 4, 3))
    f_mixed = eng.allforce()[0]
    eng.close()
    out["fp64_mode"] = bench_fp64(s, topo, prm, max(args.steps // 10, 3), f_mixed)
    print("fp64 mode: %.3f ms/call" % out["fp64_mode"]["ms_per_call"], file=sys.stderr, flush=True)
    print("plain: %.3f ms/step" % ms_step, file=sys.stderr, flush=True)
    out["md"] = bench_md(s, topo, prm, max(args.steps, 20), args.warmup)
    print("md: %.3f ms/step" % out["md"]["ms_per_step"], file=sys.stderr, flush=True)
    if not args.no_batched:
        out["batched"] = bench_batched(max(args.steps // 2, 5), args.warmup)
        print("batched: %.3f ms/call" % out["batched"]["ms_per_call"], file=sys.stderr, flush=True)
    if not args.no_respa:
        out["respa"] = bench_respa(max(args.steps, 20), args.warmup)
        print("respa: %.3f ms/long step" % out["respa"]["ms_per_long_step"], file=sys.stderr, flush=True)
    if not args.no_served:
        out["served"] = bench_served(args.serve_jobs, args.serve_steps, args.serve_rate)
    if not args.no_cpu:
        # the reference itself when its timing build is present (built from
        # the reference sources by oracle/ref/Makefile), beside the FP64
        # port (the oracle) on the same stripe
        port = cpu_baseline(s, topo, prm, args.cpu_tasks)
        brief = lambda r: {k: r[k] for k in ("value", "cores", "seconds_per_step", "sample")}
        single = mpi = None
        notes = []
        try:
            single = cpu_reference(s, args.cpu_tasks)
        except Exception as exc:   # binary absent or the run failed
            notes.append("single-core reference: " + str(exc).splitlines()[0][:160])
        try:
            mpi = cpu_reference_mpi(s, min(args.cpu_ranks, os.cpu_count() or 1))
        except Exception as exc:
            notes.append("MPI reference: " + str(exc).splitlines()[0][:160])
        cb = mpi or single or port
        if cb is not single and single is not None:
            cb["single_core"] = brief(single)
        if cb is not port:
            cb["port"] = brief(port)
        if notes:
            cb["unavailable"] = notes
        out["cpu_baseline"] = cb
        print("cpu baseline (%s): %.3g pair-interactions/s" % (out["cpu_baseline"]["kind"],
                                                               out["cpu_baseline"]["value"]),
              file=sys.stderr, flush=True)
        out["gpu_over_cpu"] = value / out["cpu_baseline"]["value"]
    print(json.dumps(out))
    return 0


def _count_k(prm) -> int:
    import math
    alpha, fexp = prm["alpha"], prm["fexp"]
    box = prm["box"]
    cutk2 = 4 * alpha * alpha * fexp
    cutk = math.sqrt(cutk2)
    tp = 2 * 3.1415926536
    km = [int(cutk * box[k] / tp + 1.0) for k in range(3)]
    n = 0
    for iz in range(-km[2], km[2] + 1):
        for ix in range(0, km[0] + 1):
            y1 = (0 if iz > 0 else 1) if ix == 0 else -km[1]
            for iy in range(y1, km[1] + 1):
                a, b, c = tp * ix / box[0], tp * iy / box[1], tp * iz / box[2]
                if (a * a + b * b) + c * c <= cutk2:
                    n += 1
    return n


if __name__ == "__main__":
    sys.exit(main())
