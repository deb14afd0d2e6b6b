"""The host restatement of the reference's path-integral MD
(oracle/pimd.py, on the FP64 oracle's forces) against the reference's own
PIMD run as four MPI ranks (tests/golden/pimd_spcfw_4.npz): the same start
state of every bead, one long step of four short steps."""
import os

import numpy as np

from mdynamix_amd import systems, topology
from oracle.pimd import PimdRef

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def _bead(d, b, n):
    pos = np.stack([d["SX_b%d_%d" % (b, n)], d["SY_b%d_%d" % (b, n)], d["SZ_b%d_%d" % (b, n)]], 1)
    ps = d["PSCAL"]
    vel = np.stack([d["VX_b%d_%d" % (b, n)], d["VY_b%d_%d" % (b, n)], d["VZ_b%d_%d" % (b, n)]], 1)
    return pos, vel * d["MASSDI"][:, None] * ps[0] / (ps[4] / 1e-15)


def test_oracle_pimd_step_matches_reference_run():
    """FP64 against FP64: positions, velocities and the BEADNOSE chain after
    one long step agree to rounding (1e-12 A, 1e-11 relative), and after
    ten (two list rebuilds, a lattice start relaxing) within 1e-12 A and
    1e-12 of their range."""
    import dataclasses
    d = np.load(os.path.join(GOLDEN, "pimd_spcfw_4.npz"), allow_pickle=False)
    sz, ps = d["PSIZES"], d["PSCAL"]
    s = dataclasses.replace(systems.cfg_flex_small(), dt_fs=1.0, nfreq=4, rebuild_every=5)
    topo, prm = topology.from_system(s)
    P = int(sz[4])
    states = [_bead(d, b, 0) for b in range(P)]
    r = PimdRef(topo, prm, s.box, [p for p, _ in states], [v for _, v in states], nfreq=int(sz[6]),
                dt=ps[4] / 1e-15, rebuild_every=int(sz[2]), beadfac=ps[10] * ps[11] * 6.02252e23 * 1e-3,
                temperature=ps[5], tau_bead=ps[6], nchain=int(sz[5]), step0=int(sz[7]),
                boltz=topology.BOLTZ_KJMOL)
    r.long_step()
    for b in range(P):
        rp, rv = _bead(d, b, 1)
        assert np.abs(r.X[b] - rp).max() <= 1e-12
        assert np.abs(r.V[b] - rv).max() <= 1e-11 * np.abs(rv).max()
    rxi = d["THERMO_b0_1"]
    assert np.abs(r.xi - rxi).max() <= 1e-11 * np.abs(rxi).max()
    for _ in range(9):
        r.long_step()
    for b in range(P):
        rp, rv = _bead(d, b, 10)
        assert np.abs(r.X[b] - rp).max() <= 1e-12
        assert np.abs(r.V[b] - rv).max() <= 1e-12 * np.abs(rv).max()
    rxi = d["THERMO_b0_10"]
    assert np.abs(r.xi - rxi).max() <= 1e-12 * np.abs(rxi).max()


def test_oracle_pimd_spring_constant_closed_form():
    """BEADFAC of the reference (setup.f:696-697, its internal units) equals
    P (k_B T / hbar)^2 / 2 per unit mass in kJ/mol/A^2 per g/mol."""
    d = np.load(os.path.join(GOLDEN, "pimd_spcfw_4.npz"), allow_pickle=False)
    ps = d["PSCAL"]
    got = ps[10] * ps[11] * 6.02252e23 * 1e-3
    assert abs(got / topology.pimd_beadfac(ps[5], int(d["PSIZES"][4])) - 1.0) < 1e-12
