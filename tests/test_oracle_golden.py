"""The FP64 oracle against the reference's own outputs (golden fixtures
written by tests/golden/make_golden.py from the reference program)."""
import hashlib

import numpy as np
import pytest

from tests.helpers import CASES, load_golden, golden_positions, golden_com, golden_forces
from mdynamix_amd import abi, topology
from oracle import orc


def _setup(name):
    d = load_golden(name)
    topo, prm = topology.from_dump(d, lam=d["lam"])
    return d, topo, topo.struct(), topology.params_struct(prm)


def _digest(a, b):
    return hashlib.sha256(np.ascontiguousarray(a, "<i4").tobytes() + np.ascontiguousarray(b, "<i4").tobytes()).hexdigest()


@pytest.mark.parametrize("name", CASES)
def test_lists_bit_exact(name):
    """CHCNB restatement: identical arrays, order and 1-4 signs."""
    d, topo, ts, ps = _setup(name)
    s1, s2, l1, l2 = orc.chcnb(ts, ps, golden_positions(d))
    assert (len(s1), len(l1)) == tuple(int(x) for x in d["LISTLEN"][0])
    assert _digest(s1, s2) == str(d["DIGEST_SHORT"][0])
    assert _digest(l1, l2) == str(d["DIGEST_LONG"][0])
    if "NBS1" in d:
        for a, k in zip((s1, s2, l1, l2), ("NBS1", "NBS2", "NBL1", "NBL2")):
            np.testing.assert_array_equal(a, d[k][0])


@pytest.mark.parametrize("name", CASES)
def test_kvectors_and_self_energy(name):
    d, topo, ts, ps = _setup(name)
    rkx, rky, rkz, spee, spei = orc.inifur(ts, ps, topo.nmoltypes)
    np.testing.assert_array_equal(rkx, d["RKX"][0])
    np.testing.assert_array_equal(rky, d["RKY"][0])
    np.testing.assert_array_equal(rkz, d["RKZ"][0])
    np.testing.assert_allclose(spee, d["SPEEI"][0], rtol=1e-14, atol=0)


@pytest.mark.parametrize("name", CASES)
def test_forces_and_energies(name):
    """ALLFORCE (or SLFORCES + FFORCES for the r-RESPA fixture) on the
    reference's lists, including the second call on displaced atoms with
    the stale list."""
    d, topo, ts, ps = _setup(name)
    lists = orc.chcnb(ts, ps, golden_positions(d))
    kv = orc.inifur(ts, ps, topo.nmoltypes)
    double = int(d["SIZES"][0][0]) == 2
    for call in range(len(d["SX"])):
        pos, com = golden_positions(d, call), golden_com(d, call)
        if double:
            _, g, _, es = orc.compute(topo, ts, ps, pos, lists, abi.MDX_FORCES_SLOW, com=com, kvec=kv)
            _, _, h, ef = orc.compute(topo, ts, ps, pos, lists, abi.MDX_FORCES_FAST, com=com, kvec=kv)
            e = {k: es[k] + ef[k] for k in abi.MdxEnergies.SCALARS}
            detail = ef   # FFORCES zeroes the type breakdowns last (service.f:421-428)
            spee = es["spee"]
        else:
            f, g, h, e = orc.compute(topo, ts, ps, pos, lists, com=com, kvec=kv)
            detail, spee = e, e["spee"]
            fr = golden_forces(d, "F", call)
            np.testing.assert_allclose(f, fr, rtol=0, atol=1e-12 * np.abs(fr).max())
        for key, arr in (("G", g), ("H", h)):
            ref = golden_forces(d, key, call)
            np.testing.assert_allclose(arr, ref, rtol=0, atol=1e-12 * max(np.abs(ref).max(), 1e-300))
        for k, ref in zip(abi.MdxEnergies.SCALARS, d["ENERGY"][call]):
            assert abs(e[k] - ref) <= 1e-12 * abs(ref) + 1e-300, (k, e[k], ref)
        nt, npairs = topo.nmoltypes, topo.ntypepairs
        np.testing.assert_allclose(detail["pote1"][:npairs], d["POTE1"][call][:npairs], rtol=1e-12, atol=1e-300)
        np.testing.assert_allclose(detail["potl1"][:npairs], d["POTL1"][call][:npairs], rtol=1e-12, atol=1e-300)
        np.testing.assert_allclose(detail["pes141"][:nt], d["PES141"][call][:nt], rtol=1e-12, atol=1e-300)
        np.testing.assert_allclose(detail["psr141"][:nt], d["PSR141"][call][:nt], rtol=1e-12, atol=1e-300)
        np.testing.assert_allclose(spee[:nt], d["SPEE"][call][:nt], rtol=1e-12, atol=1e-300)


@pytest.mark.parametrize("name", CASES)
def test_parameter_preparation(name):
    """Host parameter preparation (topology.from_system) reproduces the
    reference's tables: Coulomb factor, LJ A6/B12 per atom pair, exclusion
    lists with 1-4 marks, bonded term lists (units: EFACT converts kJ/mol
    to the reference's internal energy unit)."""
    import tests.golden.make_golden as mg
    d, topo, _, _ = _setup(name)
    s = mg.CASES[name][0]()
    mine, prm = topology.from_system(s)
    efact, coulf = d["SCALARS"][0][12], d["SCALARS"][0][9]
    assert abs(prm["coulomb_factor"] * efact / coulf - 1.0) < 1e-14
    idx = np.arange(0, topo.natoms, max(1, topo.natoms // 200))
    for a, b in (("a6", "a6"), ("b12", "b12"), ("a6_0", "a6_0"), ("a6_14", "a6_14")):
        ref = getattr(topo, a)[topo.atom_ljtype[idx][:, None], topo.atom_ljtype[idx][None, :]]
        me = getattr(mine, b)[mine.atom_ljtype[idx][:, None], mine.atom_ljtype[idx][None, :]] * efact
        np.testing.assert_allclose(me, ref, rtol=1e-13, atol=1e-300)
    np.testing.assert_array_equal(mine.excl_start, topo.excl_start)
    np.testing.assert_array_equal(mine.excl_atom, topo.excl_atom)
    np.testing.assert_array_equal(mine.excl_is14, topo.excl_is14)
    np.testing.assert_allclose(mine.atom_charge, topo.atom_charge, rtol=1e-14)
    np.testing.assert_array_equal(mine.bond_i, topo.bond_i)
    np.testing.assert_array_equal(mine.angle_j, topo.angle_j)
    np.testing.assert_array_equal(mine.tors_l, topo.tors_l)
    np.testing.assert_allclose(mine.bp_k * efact, topo.bp_k, rtol=1e-13)
    np.testing.assert_allclose(mine.ap_theta0, topo.ap_theta0, rtol=1e-13)
    np.testing.assert_allclose(mine.tp_k1 * efact, topo.tp_k1, rtol=1e-13)
