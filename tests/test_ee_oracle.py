"""The oracle's restatement of DIFEN / DIFFUR (oracle/mdee.py) against the
reference program's own values (tests/golden/ee_tiny.npz: the dump harness
calls the reference DIFFUR and DIFEN toward every subensemble at each
dumped call, with the reference's lists and coordinates).  Within 1e-12 of
each term's magnitude."""
import numpy as np

from tests.helpers import load_golden, golden_positions
from mdynamix_amd import topology
from oracle import mdee, orc


def test_difen_diffur_match_reference():
    d = load_golden("ee_tiny")
    topo, prm = topology.from_dump(d, lam=d["lam"])
    ts, ps = topo.struct(), topology.params_struct(dict(prm, list_mode=0, skin=0.0))
    lists = orc.chcnb(ts, ps, golden_positions(d, 0))       # built at the first call, kept
    checked = 0
    for call in range(len(d["SX"])):
        ed, ec, me, (pw1, pw2, bee) = mdee.from_fixture(d, call)
        pos = golden_positions(d, call)
        ep = mdee.difen(topo, prm, pos, lists, ec, pw1, pw2, bee)
        er = mdee.diffur(topo, prm, pos, ec, pw1)
        for me1 in range(len(ec)):
            efn, efo, een, eeo = ed[me1]
            scale_f, scale_e = max(abs(efn), abs(efo)), max(abs(een), abs(eeo))
            assert abs(er[me1] - efn) <= 1e-12 * scale_f, (call, me1, er[me1], efn)
            assert abs(er[me - 1] - efo) <= 1e-12 * scale_f, (call, me1, er[me - 1], efo)
            assert abs(ep[me1] - een) <= 1e-12 * scale_e, (call, me1, ep[me1], een)
            assert abs(ep[me - 1] - eeo) <= 1e-12 * scale_e, (call, me1, ep[me - 1], eeo)
            checked += 1
        # the two subensembles differ (the check would pass for a constant)
        assert abs(ep[0] - ep[1]) > 1e-3 * scale_e and abs(er[0] - er[1]) > 1e-3 * scale_f
    assert checked == 4
