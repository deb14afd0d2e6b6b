"""Path-integral MD on the device (PIMD, pimd.f:18-314) against the
reference's own PIMD run as four MPI ranks, one bead each
(tests/golden/pimd_spcfw_4.npz, written by make_golden.py pimd from
oracle/ref/mdx_pimd.f).

216 flexible SPC-FW waters + 2 Na+ + 2 Cl- (652 atoms) as 4 beads, 1 fs
long steps of 4 short steps, the ring springs of the reference's BEADFAC at
300 K, BEADNOSE chains of one link (BTAU 5 fs), GETCOM after every drift,
the list rebuilt inside the 5th and 10th long steps.  Both runs start from
the reference's start state of each bead (after its per-bead position
jitter, main.f:653-663).
"""
import numpy as np
import pytest

from mdynamix_amd import abi, systems, topology
from mdynamix_amd.engine import ForceEngine, ForceBatch
import os

pytestmark = pytest.mark.gpu

AVSNO = 6.02252e23
GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def load_golden(name):
    return np.load(os.path.join(GOLDEN, name + ".npz"), allow_pickle=False)


def _bead_state(d, b, n):
    pos = np.stack([d["SX_b%d_%d" % (b, n)], d["SY_b%d_%d" % (b, n)], d["SZ_b%d_%d" % (b, n)]], 1)
    ps = d["PSCAL"]
    tstep, dt_fs = ps[0], ps[4] / 1e-15
    # the reference keeps momenta in internal units: S += TSTEB * V * MASSDI
    vel = np.stack([d["VX_b%d_%d" % (b, n)], d["VY_b%d_%d" % (b, n)], d["VZ_b%d_%d" % (b, n)]], 1)
    return pos, vel * d["MASSDI"][:, None] * tstep / dt_fs


def _ring(d):
    import dataclasses
    s = dataclasses.replace(systems.cfg_flex_small(), dt_fs=1.0, nfreq=4, rebuild_every=5)
    topo, prm = topology.from_system(s)
    sz, ps = d["PSIZES"], d["PSCAL"]
    nbeads, nchain, nfreq = int(sz[4]), int(sz[5]), int(sz[6])
    assert int(sz[3]) == 1 and nfreq == s.nfreq and int(sz[2]) == s.rebuild_every
    md = topology.md_params(s)
    md.update(integrator=abi.MDX_MD_PIMD, tau_t=0.0, step0=int(sz[7]), dt=ps[4] / 1e-15, nfreq=nfreq)
    engs = []
    for b in range(nbeads):
        e = ForceEngine(topo, prm)
        pos, vel = _bead_state(d, b, 0)
        e.set_positions(pos)
        e.md_init(md, vel)
        engs.append(e)
    # BEADFAC from the reference's internal units to kJ/mol/A^2 per g/mol
    beadfac = ps[10] * ps[11] * AVSNO * 1e-3
    assert abs(beadfac / topology.pimd_beadfac(ps[5], nbeads) - 1.0) < 1e-12
    ring = ForceBatch(engs, ring_beadfac=beadfac)
    ring.pimd_init(temperature=ps[5], tau_bead=ps[6], nchain=nchain)
    return engs, ring


def test_pimd_beads_match_reference_run():
    """After 1 long step (4 short steps) every bead's positions agree within
    1e-7 A, velocities within 1e-6 (relative to the largest) and the
    BEADNOSE chain variables within 1e-6 of their range; after 10 long
    steps, with two list rebuilds, a lattice start relaxing (bead
    temperature ~380 K), within 1e-5 A, 1e-5 and 1e-5 (measured: 8e-9 A,
    1.4e-7, 2e-7; 3.4e-7 A, 7.9e-7, 4.7e-7).  The device forces carry the
    engine's FP32 pair error and its FP64 bead neighbours (the reference
    passes them as REAL*4, pimd.f:451-486)."""
    d = load_golden("pimd_spcfw_4")
    engs, ring = _ring(d)
    print("case pimd_spcfw_4 beads", len(engs), "atoms", engs[0].n, flush=True)
    try:
        res = {}
        tr = None
        for n0, n in ((0, 1), (1, 10)):
            tr = ring.md_run(n - n0)
            dpos = dvel = 0.0
            for b, e in enumerate(engs):
                pos, vel = e.md_state()
                rpos, rvel = _bead_state(d, b, n)
                dpos = max(dpos, float(np.abs(pos - rpos).max()))
                dvel = max(dvel, float(np.abs(vel - rvel).max() / np.abs(rvel).max()))
            xi = ring.pimd_chains()
            rxi = d["THERMO_b0_%d" % n]
            res["pos%d" % n] = dpos
            res["vel%d" % n] = dvel
            res["xi%d" % n] = float(np.abs(xi - rxi).max() / np.abs(rxi).max())
            # the chains are the same on every bead of the reference
            for b in range(1, len(engs)):
                assert np.array_equal(d["THERMO_b%d_%d" % (b, n)], rxi)
        print(res, flush=True)
        assert res["pos1"] <= 1e-7 and res["vel1"] <= 1e-6 and res["xi1"] <= 1e-6
        assert res["pos10"] <= 1e-5 and res["vel10"] <= 1e-5 and res["xi10"] <= 1e-5
        assert np.all(np.isfinite(tr))
    finally:
        ring.close()
        for e in engs:
            e.close()


def test_pimd_graph_replay_matches_eager():
    """The path-integral steps replayed from their captured hipGraph give
    the bits of the same steps launched eagerly."""
    d = load_golden("pimd_spcfw_4")
    states = []
    for graph in (False, True):
        engs, ring = _ring(d)
        try:
            ring.lib.mdx_batch_use_graph(ring._b, 1 if graph else 0)
            ring.md_run(3)
            states.append([e.md_state() for e in engs] + [ring.pimd_chains()])
        finally:
            ring.close()
            for e in engs:
                e.close()
    a, b = states
    for x, y in zip(a[:-1], b[:-1]):
        for u, v in zip(x, y):
            assert np.array_equal(u, v)
    assert np.array_equal(a[-1], b[-1])


def test_pimd_benchmark_ring_graph_matches_eager():
    """The benchmark's ring (32 beads x 4,000 flexible SPC-FW waters,
    12,000 atoms each, 49.3 A, 300 K): 3 long steps of the device PIMD loop
    (1 fs, 4 short steps, BEADNOSE 5 fs, list rebuild every 2 long steps)
    replayed from the captured hipGraph give the bits of the eager launches,
    and the bead temperatures stay finite and positive."""
    from mdynamix_amd import systems as sysm
    bs = sysm.pimd_beads(n_beads=32, n_water=4000, box=49.3, seed=2027)
    topo, prm = topology.from_system(bs[0])
    md = topology.md_params(bs[0])
    md.update(integrator=abi.MDX_MD_PIMD, nfreq=4, dt=1.0, tau_t=0.0, rebuild_every=2)
    out = []
    for graph in (False, True):
        engs = [ForceEngine(topo, prm) for _ in bs]
        ring = None
        try:
            for k, (e, b) in enumerate(zip(engs, bs)):
                e.set_positions(b.pos)
                e.md_init(md, sysm.maxwell_velocities(b, 100 + k))
            ring = ForceBatch(engs, ring_beadfac=topology.pimd_beadfac(300.0, 32))
            ring.pimd_init(temperature=300.0, tau_bead=5.0, nchain=1)
            ring.lib.mdx_batch_use_graph(ring._b, 1 if graph else 0)
            print("case pimd ring 32 x 12000 graph", graph, flush=True)
            tr = ring.md_run(3)
            out.append(([e.md_state() for e in engs], ring.pimd_chains(), tr))
        finally:
            if ring is not None:
                ring.close()
            for e in engs:
                e.close()
    (sa, xa, ta), (sb, xb, tb) = out
    for (pa, va), (pb, vb) in zip(sa, sb):
        assert np.array_equal(pa, pb) and np.array_equal(va, vb)
    assert np.array_equal(xa, xb) and np.array_equal(ta, tb)
    assert np.all(np.isfinite(ta)) and np.all(ta[:, :, 2] > 0.0)
