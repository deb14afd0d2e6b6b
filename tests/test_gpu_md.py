"""Device-resident MD segments against the reference integrators: the
leap-frog SULTAN with constraints (below) and the multiple-time-step DOUBLE
(test_respa_*, fixture traj_spcfw_respa).

The leap-frog fixture holds the reference program's own start state of the
508-water box and its state after 1 and 20 leap-frog steps of SULTAN
(mdstep.f:29-260: SHEJK constraints with TOLER = 1e-6, Nose-Hoover
SCALING(3) with QT = 30 fs, GETCOM wrapping, CHCNB every 10 steps),
written by oracle/ref/mdx_traj.f.  The device integrator starts from the
same state; its forces carry the engine's FP32 pair error (force RMS
~1e-6), so the states agree to what that error produces: after one step
positions within 1e-6 A, velocities within 1e-5 relative, temperature
within 1e-6 relative; after 20 steps of this far-from-equilibrium start
(T rises from 300 K to ~1400 K) positions within 1e-4 A (measured 2.4e-6).
SC is compared in units of its change per step for d(T/T0) = 1.
Constraints hold to the SHAKE tolerance (|r/RB - 1| <= 5e-13).
"""
import os

import numpy as np
import pytest

from tests.helpers import GOLDEN
from mdynamix_amd import abi, systems, topology
from mdynamix_amd.engine import ForceEngine, ForceBatch

pytestmark = pytest.mark.gpu


def load_golden(name):
    z = np.load(os.path.join(GOLDEN, name + ".npz"), allow_pickle=False)
    return {k: z[k] for k in z.files}


def _state(d, n):
    pos = np.stack([d["SX_%d" % n], d["SY_%d" % n], d["SZ_%d" % n]], 1)
    ts = d["TSCAL"]
    tstep, dt_fs = ts[0], ts[4] / 1e-15
    # SULTAN keeps momenta in internal units: S += TSTEP * V * MASSDI
    vel = np.stack([d["VX_%d" % n], d["VY_%d" % n], d["VZ_%d" % n]], 1) * d["MASSDI"][:, None] * tstep / dt_fs
    return pos, vel


def _md_engine(d):
    s = systems.cfg_small()
    topo, prm = topology.from_system(s)
    eng = ForceEngine(topo, prm)
    pos, vel = _state(d, 0)
    eng.set_positions(pos)
    ts = d["TSCAL"]
    md = topology.md_params(s)
    md.update(step0=1, temperature=ts[5], tau_t=ts[6], ndof=ts[8], shake_tol=ts[10], shake_bdtol=ts[11],
              dt=ts[4] / 1e-15)
    assert int(d["TSIZES"][2]) == md["rebuild_every"]
    eng.md_init(md, vel)
    return s, eng, md


def _bond_errors(pos, md):
    b = pos[md["cons_i"]] - pos[md["cons_j"]]
    return np.abs(np.sqrt((b * b).sum(1)) / md["cons_len"] - 1.0).max()


def test_md_segment_matches_reference_integrator():
    d = load_golden("traj_water_508")
    s, eng, md = _md_engine(d)
    print("case traj_water_508", eng.n, flush=True)
    ts = d["TSCAL"]
    dt_fs, qt = ts[4] / 1e-15, ts[6]
    tr1 = eng.md_run(1)
    pos1, vel1 = eng.md_state()
    rpos1, rvel1 = _state(d, 1)
    tr = eng.md_run(19)
    pos20, vel20 = eng.md_state()
    rpos20, rvel20 = _state(d, 20)
    # SC is kept per internal time unit (TSTEP = dt); TEMP from GETEMP
    dsc_step = dt_fs / qt ** 2         # d SC per step for d(T/T0) = 1
    res = {
        "pos1": np.abs(pos1 - rpos1).max(),
        "vel1": np.abs(vel1 - rvel1).max() / np.abs(rvel1).max(),
        "T1": abs(tr1[0, 2] / d["TEND_1"][1] - 1.0),
        "SC1": abs(tr1[0, 3] - d["TEND_1"][0] / dt_fs) / dsc_step,
        "pos20": np.abs(pos20 - rpos20).max(),
        "vel20": np.abs(vel20 - rvel20).max() / np.abs(rvel20).max(),
        "T20": abs(tr[-1, 2] / d["TEND_20"][1] - 1.0),
        "SC20": abs(tr[-1, 3] - d["TEND_20"][0] / dt_fs) / dsc_step,
        "bond1": _bond_errors(pos1, md),
        "bond20": _bond_errors(pos20, md),
    }
    print(res, flush=True)
    assert res["pos1"] <= 1e-6 and res["vel1"] <= 1e-5
    assert res["T1"] <= 1e-6 and res["SC1"] <= 1e-6
    assert res["pos20"] <= 1e-4 and res["vel20"] <= 1e-4
    assert res["T20"] <= 1e-6 and res["SC20"] <= 1e-5
    assert res["bond1"] <= 1e-11 and res["bond20"] <= 1e-11
    assert np.all(np.isfinite(tr))
    assert eng.md_shake_failures() == 0


def test_shake_failures_are_counted():
    """With MAXIT = 0 (one SHEJK sweep) every molecule stops before the
    tolerance is met (the limit is set from the host, every kernel exits
    normally): the run continues, as the reference's does after its
    NO CONVERGENCE message (mdstep.f:1264-1268), and the count says so."""
    s = systems.cfg_small()
    topo, prm = topology.from_system(s)
    eng = ForceEngine(topo, prm)
    try:
        eng.set_positions(s.pos)
        md = topology.md_params(s)
        md.update(shake_maxit=0)
        eng.md_init(md, systems.maxwell_velocities(s, 5))
        print("case shake maxit 0", flush=True)
        tr = eng.md_run(2)
        assert np.all(np.isfinite(tr))
        assert eng.md_shake_failures() > 0
    finally:
        eng.close()


def test_batched_md_matches_alone():
    """Two systems advanced by fused MD launches give the bits of each
    system advanced alone."""
    d = load_golden("traj_water_508")
    runs = []
    for batched in (False, True):
        s1, e1, _ = _md_engine(d)
        s2 = systems.water_nacl(300, 2, 21.5, 5, rcut=10.0, rshort=5.0, ewald_a=2.8, ewald_b=9.0, shake=True)
        topo2, prm2 = topology.from_system(s2)
        e2 = ForceEngine(topo2, prm2)
        e2.set_positions(s2.pos)
        rng = np.random.default_rng(4)
        e2.md_init(topology.md_params(s2), rng.normal(0.0, 0.005, (e2.n, 3)))
        if batched:
            b = ForceBatch([e1, e2])
            tr = b.md_run(12)
            b.close()
        else:
            tr = np.stack([e1.md_run(12), e2.md_run(12)])
        runs.append((tr, e1.md_state(), e2.md_state()))
    (ta, a1, a2), (tb, b1, b2) = runs
    assert np.array_equal(ta, tb)
    for x, y in zip(a1 + a2, b1 + b2):
        assert np.array_equal(x, y)


def _respa_engine(d):
    s = systems.cfg_flex_small()
    topo, prm = topology.from_system(s)
    eng = ForceEngine(topo, prm)
    pos, vel = _state(d, 0)
    eng.set_positions(pos)
    ts, sz = d["TSCAL"], d["TSIZES"]
    md = topology.md_params(s)
    # the reference ran DOUBLE (no constraints) with the same NFREQ and ICHNB
    assert md["integrator"] == abi.MDX_MD_RESPA and int(sz[5]) == 1 and int(sz[4]) == 0
    assert int(sz[6]) == md["nfreq"] and int(sz[2]) == md["rebuild_every"]
    md.update(step0=int(sz[7]), temperature=ts[5], tau_t=ts[6], ndof=ts[8], dt=ts[4] / 1e-15)
    eng.md_init(md, vel)
    return s, eng, md


def test_respa_segment_matches_reference_integrator():
    """The multiple-time-step integrator against the reference DOUBLE
    (mdstep.f:281-530) on 216 flexible SPC-FW waters + 4 ions: 2 fs long
    steps of 4 short steps, Nose-Hoover SCALING(1)/(2) with QT = 30 fs,
    GETCOM after every drift, the list rebuilt inside the 10th and 20th
    long steps.  Same start state; the device forces carry the engine's
    FP32 pair error, so after 1 long step positions agree within 1e-6 A
    and after 20 (a lattice start relaxing, T 300 -> 865 -> 237 K) within
    1e-4 A."""
    d = load_golden("traj_spcfw_respa")
    s, eng, md = _respa_engine(d)
    print("case traj_spcfw_respa", eng.n, flush=True)
    ts = d["TSCAL"]
    dt_fs, qt = ts[4] / 1e-15, ts[6]
    tr1 = eng.md_run(1)
    pos1, vel1 = eng.md_state()
    rpos1, rvel1 = _state(d, 1)
    tr = eng.md_run(19)
    pos20, vel20 = eng.md_state()
    rpos20, rvel20 = _state(d, 20)
    dsc_step = dt_fs / qt ** 2
    res = {
        "pos1": np.abs(pos1 - rpos1).max(),
        "vel1": np.abs(vel1 - rvel1).max() / np.abs(rvel1).max(),
        "T1": abs(tr1[0, 2] / d["TEND_1"][1] - 1.0),
        "SC1": abs(tr1[0, 3] - d["TEND_1"][0] / dt_fs) / dsc_step,
        "pos20": np.abs(pos20 - rpos20).max(),
        "vel20": np.abs(vel20 - rvel20).max() / np.abs(rvel20).max(),
        "T20": abs(tr[-1, 2] / d["TEND_20"][1] - 1.0),
        "SC20": abs(tr[-1, 3] - d["TEND_20"][0] / dt_fs) / dsc_step,
    }
    print(res, flush=True)
    assert res["pos1"] <= 1e-6 and res["vel1"] <= 1e-5
    assert res["T1"] <= 1e-6 and res["SC1"] <= 1e-6
    assert res["pos20"] <= 1e-4 and res["vel20"] <= 1e-4
    assert res["T20"] <= 1e-5 and res["SC20"] <= 1e-5
    assert np.all(np.isfinite(tr))


def test_batched_respa_matches_alone():
    """Two flexible systems advanced by fused multiple-time-step launches
    give the bits of each system advanced alone."""
    d = load_golden("traj_spcfw_respa")
    runs = []
    for batched in (False, True):
        s1, e1, _ = _respa_engine(d)
        s2 = systems.water_nacl(150, 1, 17.0, 6, name="f150", water=systems.spcfw_water(), rcut=8.0, rshort=5.0,
                                ewald_a=2.8, ewald_b=9.0, shake=False, nfreq=4, dt_fs=2.0)
        topo2, prm2 = topology.from_system(s2)
        e2 = ForceEngine(topo2, prm2)
        e2.set_positions(s2.pos)
        e2.md_init(topology.md_params(s2), systems.maxwell_velocities(s2, 8))
        if batched:
            b = ForceBatch([e1, e2])
            tr = b.md_run(12)
            b.close()
        else:
            tr = np.stack([e1.md_run(12), e2.md_run(12)])
        runs.append((tr, e1.md_state(), e2.md_state()))
    (ta, a1, a2), (tb, b1, b2) = runs
    assert np.array_equal(ta, tb)
    for x, y in zip(a1 + a2, b1 + b2):
        assert np.array_equal(x, y)
