"""Restart (.dmp) and start (.start) files (mdynamix_amd/restart.py)
against files the reference program itself wrote (RESTRT, restart.f:23-417;
main.f:551-584), built from the reference sources by oracle/ref/Makefile:

* every record read and written back byte for byte: unformatted and ASCII
  dumps of a water box (averages), the unformatted dump of an
  expanded-ensemble job (its subensemble records);
* the fields mean what the reference meant: in one run of the trajectory
  harness the dump's coordinates and momenta are those the harness records
  as the end state, and v = V M / (m DT) reproduces the velocities;
* a restart file built from a state (coordinates, velocities in A/fs, SC)
  lets the reference program continue: 3 + 3 steps through such a file
  equal 6 straight steps (lists rebuilt every step) to 1e-9 A;
* XMOL start files read back as written, box included.
"""
import dataclasses
import glob
import os
import subprocess
import tempfile

import numpy as np
import pytest

from mdynamix_amd import restart, systems, topology
from oracle import refdump

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
MD = os.path.join(ROOT, "oracle", "_ref", "mdx_md")
needs_ref = pytest.mark.skipif(not os.path.exists(MD), reason="reference program not built (make -C oracle/ref)")


def _run(system, nsteps, exe=MD, workdir=None):
    d = workdir or tempfile.mkdtemp(prefix="mdxrst_")
    system.write_reference_inputs(d, nsteps=nsteps)
    res = subprocess.run([exe, "md.in"], cwd=d, capture_output=True, text=True, timeout=300)
    f = glob.glob(os.path.join(d, "*.dmp"))
    assert f, res.stdout[-2000:] + res.stderr[-1000:]
    return f[0]


@needs_ref
@pytest.mark.parametrize("case", ["water_binary", "water_ascii", "lambda_job_binary"])
def test_dmp_round_trip_is_byte_exact(case):
    s = systems.ee_job(1500, 24, 0.6, 21) if case.startswith("lambda") else systems.cfg_small()
    s = dataclasses.replace(s, dump_every=3, dump_ascii=case.endswith("ascii"))
    path = _run(s, 3)
    r = restart.read_dmp(path)
    assert r.formatted == case.endswith("ascii")
    assert r.nstot == s.natoms and r.counters[1] == 3
    assert np.allclose(r.box, s.box)
    out = path + ".rt"
    restart.write_dmp(out, r)
    assert open(out, "rb").read() == open(path, "rb").read()


@needs_ref
def test_dmp_fields_are_the_reference_state():
    # the harness runs with NSTTOT = 1 (oracle/ref/mdx_traj.f), so the
    # dump of mod(NSTTOT + NSTEP, NDUMP) = 0 (mdstep.f:246) after its 4th
    # and last step needs NDUMP = 5
    s = dataclasses.replace(systems.cfg_small(), dump_every=5)
    d = tempfile.mkdtemp(prefix="mdxrst_")
    tr = refdump.run_trajectory(s, 4, workdir=d)
    r = restart.read_dmp(glob.glob(os.path.join(d, "*.dmp"))[0])
    end = np.stack([tr["SX"][1], tr["SY"][1], tr["SZ"][1]], 1)
    vend = np.stack([tr["VX"][1], tr["VY"][1], tr["VZ"][1]], 1)
    assert np.array_equal(r.pos, end) and np.array_equal(r.mom, vend)
    # v [A/fs] = V MASSDI TSTEP / DT[fs] (mdstep.f:109-111) with the harness' own MASSDI
    topo, _ = topology.from_system(s)
    ref_v = vend * tr["MASSDI"][0][:, None] * tr["TSCAL"][0][0] / (r.dt * 1e15)
    assert np.allclose(r.velocities(topo.atom_mass), ref_v, rtol=1e-13, atol=0.0)


@needs_ref
def test_reference_continues_from_a_built_restart_file():
    s = dataclasses.replace(systems.cfg_small(), rebuild_every=1, dump_every=6)
    topo, _ = topology.from_system(s)
    straight = restart.read_dmp(_run(s, 6))
    half = restart.read_dmp(_run(dataclasses.replace(s, dump_every=3), 3))
    built = restart.restart_from_state(half.pos, half.velocities(topo.atom_mass), topo.atom_mass, half.box,
                                       half.dt * 1e15, ntypes=len(s.types), nmols=topo.nmols,
                                       nose_sc=half.nose_sc, nsttot=3)
    d = tempfile.mkdtemp(prefix="mdxrst_")
    restart.write_dmp(os.path.join(d, s.name + ".dmp"), built)
    cont = restart.read_dmp(_run(dataclasses.replace(s, read_restart=True), 3, workdir=d))
    assert cont.counters[1] == 6
    assert np.abs(cont.pos - straight.pos).max() <= 1e-9
    assert np.abs(cont.mom - straight.mom).max() <= 1e-9 * np.abs(straight.mom).max()


def test_start_file_round_trip(tmp_path):
    s = systems.cfg_small()
    s.write_reference_inputs(str(tmp_path), nsteps=1)
    names, pos, box = restart.read_start(os.path.join(str(tmp_path), s.name + ".start"))
    assert np.array_equal(pos, s.pos) and np.allclose(box, s.box) and len(names) == s.natoms
    restart.write_start(os.path.join(str(tmp_path), "b.start"), names, pos, box)
    n2, p2, b2 = restart.read_start(os.path.join(str(tmp_path), "b.start"))
    assert n2 == names and np.array_equal(p2, pos) and np.array_equal(b2, box)
