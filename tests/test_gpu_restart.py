"""Job ingest and egress through the reference's restart files
(mdynamix_amd/restart.py; RESTRT, restart.f:23-417).

The 508-water box (SHAKE, Nose-Hoover, lists rebuilt every step) run by the
reference program for 3 steps and dumped; the device MD segment starts from
that dump (coordinates, momenta converted to A/fs, SC) and runs 3 more
steps; the result agrees with the reference's own 6-step run to the
engine's MD parity (positions within 1e-5 A, velocities within 1e-4 of
their largest value, SC within 1e-5 of its change per step).  The engine's
state is then written as a restart file, the reference program reads it
and runs 3 more steps, and the result agrees with a 9-step reference run
within 1e-4 A.  The reference program (oracle/_ref/mdx_md) is built from
its sources by oracle/ref/Makefile.
"""
import dataclasses
import os
import tempfile

import numpy as np
import pytest

from mdynamix_amd import restart, systems, topology
from mdynamix_amd.engine import ForceEngine
from tests.test_restart import MD, _run

pytestmark = [pytest.mark.gpu,
              pytest.mark.skipif(not os.path.exists(MD), reason="reference program not built (make -C oracle/ref)")]


def test_engine_continues_a_reference_restart_and_hands_back():
    s = dataclasses.replace(systems.cfg_small(), rebuild_every=1)
    topo, prm = topology.from_system(s)
    mass = topo.atom_mass
    straight6 = restart.read_dmp(_run(dataclasses.replace(s, dump_every=6), 6))
    straight9 = restart.read_dmp(_run(dataclasses.replace(s, dump_every=9), 9))
    half = restart.read_dmp(_run(dataclasses.replace(s, dump_every=3), 3))
    dt_fs = half.dt * 1e15
    eng = ForceEngine(topo, prm)
    try:
        print("case restart ingest 508 waters", flush=True)
        eng.set_positions(half.pos)
        eng.chcnb()
        md = topology.md_params(s)
        md.update(step0=3, sc0=half.nose_sc_per_fs)
        eng.md_init(md, half.velocities(mass))
        tr = eng.md_run(3)
        pos, vel = eng.md_state()
        v6 = straight6.velocities(mass)
        dsc_step = dt_fs / md["tau_t"] ** 2
        res = {"pos": float(np.abs(pos - straight6.pos).max()),
               "vel": float(np.abs(vel - v6).max() / np.abs(v6).max()),
               "sc": abs(tr[-1, 3] - straight6.nose_sc_per_fs) / dsc_step}
        print(res, flush=True)
        assert res["pos"] <= 1e-5 and res["vel"] <= 1e-4 and res["sc"] <= 1e-5
        out = restart.restart_from_state(pos, vel, mass, s.box, dt_fs, ntypes=len(s.types), nmols=topo.nmols,
                                         nose_sc=tr[-1, 3] * dt_fs, nsttot=6)
    finally:
        eng.close()
    d = tempfile.mkdtemp(prefix="mdxegr_")
    restart.write_dmp(os.path.join(d, s.name + ".dmp"), out)
    back = restart.read_dmp(_run(dataclasses.replace(s, read_restart=True, dump_every=9), 3, workdir=d))
    assert back.counters[1] == 9
    err = float(np.abs(back.pos - straight9.pos).max())
    print("egress: 3 reference steps from the engine's state, max |dr| %.2e A" % err, flush=True)
    assert err <= 1e-4
