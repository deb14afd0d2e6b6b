"""Expanded-ensemble energies on the device (mdx_ee_energies: the sums of
DIFEN, mdee.f:379-530, and DIFFUR, mdee.f:265-370, as functions of lambda).

* Against the reference program's own DIFFUR / DIFEN values toward every
  subensemble (tests/golden/ee_tiny.npz, two dumped calls, the second with
  the list of the first): within 1e-10 of each term's magnitude.
* Against the oracle (oracle/mdee.py, pinned on the same fixture) for 20
  values of lambda on a 1,500-atom lambda-coupled job, with reference lists
  and with a Verlet list at displaced coordinates (pairs within R_cut at
  the current coordinates): within 1e-10.
The residual differences are rounding (sincospi of reduced phases against
libm sin/cos, device exp, summation order).
"""
import numpy as np
import pytest

from tests.helpers import load_golden, golden_positions
from mdynamix_amd import abi, systems, topology
from mdynamix_amd.engine import ForceEngine
from oracle import mdee, orc

pytestmark = pytest.mark.gpu

TOL = 1e-10


def test_ee_energies_vs_reference():
    d = load_golden("ee_tiny")
    topo, prm = topology.from_dump(d, lam=d["lam"])
    eng = ForceEngine(topo, prm)
    try:
        print("case ee_tiny lambda energies", flush=True)
        eng.set_positions(golden_positions(d, 0))
        eng.chcnb()
        for call in range(len(d["SX"])):
            eng.set_positions(golden_positions(d, call))
            ed, ec, me, (pw1, pw2, bee) = mdee.from_fixture(d, call)
            ep, er = eng.ee_energies(ec, pw1, pw2, bee)
            for me1 in range(len(ec)):
                efn, efo, een, eeo = ed[me1]
                sf, se = max(abs(efn), abs(efo)), max(abs(een), abs(eeo))
                assert abs(er[me1] - efn) <= TOL * sf, (call, me1, er[me1], efn)
                assert abs(er[me - 1] - efo) <= TOL * sf, (call, me1, er[me - 1], efo)
                assert abs(ep[me1] - een) <= TOL * se, (call, me1, ep[me1], een)
                assert abs(ep[me - 1] - eeo) <= TOL * se, (call, me1, ep[me - 1], eeo)
            print("call %d: e_pair %s e_recip %s" % (call, ep, er), flush=True)
    finally:
        eng.close()


@pytest.mark.parametrize("mode", ["reference", "verlet"])
def test_ee_energies_vs_oracle(mode):
    s = systems.ee_job(1500, 24, 0.6, 21)
    topo, prm = topology.from_system(s)
    lams = np.linspace(0.05, 1.0, 20)
    ts = topo.struct()
    ps = topology.params_struct(dict(prm, list_mode=0, skin=0.0))
    if mode == "reference":
        prm = dict(prm, list_mode=abi.MDX_LIST_REFERENCE, skin=0.0)
        pos = s.pos
    else:
        prm = dict(prm, list_mode=abi.MDX_LIST_VERLET, skin=1.0)
        pos = s.pos + np.random.default_rng(4).uniform(-0.2, 0.2, s.pos.shape)
    eng = ForceEngine(topo, prm)
    try:
        print("case ee_job lambda energies", mode, topo.natoms, flush=True)
        eng.set_positions(s.pos)
        eng.chcnb()
        eng.set_positions(pos)
        ep, er = eng.ee_energies(lams, s.ee_pw1, s.ee_pw2, s.ee_bee)
        # the oracle's reference list at the coordinates the pairs are taken at
        lists = orc.chcnb(ts, ps, pos if mode == "verlet" else s.pos)
        op = mdee.difen(topo, prm, pos, lists, lams, s.ee_pw1, s.ee_pw2, s.ee_bee)
        orr = mdee.diffur(topo, prm, pos, lams, s.ee_pw1)
        assert np.all(np.abs(ep - op) <= TOL * np.abs(op).max()), (ep - op, op)
        assert np.all(np.abs(er - orr) <= TOL * np.abs(orr).max()), (er - orr, orr)
        assert np.ptp(op) > 1e-3 * np.abs(op).max()
        print("max rel diff: pair %.2e recip %.2e" % (np.abs(ep - op).max() / np.abs(op).max(),
                                                      np.abs(er - orr).max() / np.abs(orr).max()), flush=True)
    finally:
        eng.close()
