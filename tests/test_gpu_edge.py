"""Edge cases of the HIP engine against the FP64 oracle, through the C ABI:
systems smaller than one 8-atom cluster or with a ragged last cluster, a
single molecule (no intermolecular pair at all), and a box without charges
(Lennard-Jones only: zero structure factors, the force-pass operand scale
falls back to 2^0). Tolerances as in test_gpu_parity.py: force RMS
relative error <= 1e-5, total energy <= 1e-6, components <= 1e-5.
"""
import numpy as np
import pytest

from mdynamix_amd import abi, systems, topology
from mdynamix_amd.engine import ForceEngine
from oracle import orc

pytestmark = pytest.mark.gpu

FORCE_RMS_TOL = 1e-5
ENERGY_KEYS = ("pe1", "pe2", "pels1", "pels2", "spe", "pint")
GROUPS = (abi.MDX_FORCES_ALL, abi.MDX_FORCES_FAST, abi.MDX_FORCES_SLOW)


def _rms_rel(a, b):
    den = (b ** 2).sum()
    if den == 0.0:
        return float(np.abs(a).max())
    return float(np.sqrt(((a - b) ** 2).sum() / den))


def _check(e, ref):
    """Total within 1e-6 of the larger of |total| and the sum of component
    magnitudes: in a box of two waters the total (-0.53 kJ/mol) is what is
    left of +11.0 reciprocal and -11.5 self/exclusion energy, so 1e-6 of
    the total alone would ask for 5e-8 of the components."""
    scale = sum(abs(ref[k]) for k in ("pe1", "pe2", "pels1", "pels2", "spe")) + 1e-12
    tot, tot_ref = sum(e[k] for k in ENERGY_KEYS), sum(ref[k] for k in ENERGY_KEYS)
    assert abs(tot - tot_ref) <= 1e-6 * max(abs(tot_ref), scale), ("total", tot, tot_ref)
    for k in abi.MdxEnergies.SCALARS:
        assert abs(e[k] - ref[k]) <= 1e-5 * max(abs(ref[k]), 1e-3 * scale) + 1e-12, (k, e[k], ref[k])


def _run_vs_oracle(s, label):
    topo, prm = topology.from_system(s)
    prm["list_mode"] = abi.MDX_LIST_REFERENCE
    eng = ForceEngine(topo, prm)
    try:
        eng.set_positions(s.pos)
        eng.chcnb()
        ts, ps = topo.struct(), topology.params_struct(dict(prm, list_mode=0, skin=0.0))
        lists = orc.chcnb(ts, ps, s.pos)
        for group in GROUPS:
            print("case", label, topo.natoms, "atoms, group", group, flush=True)
            f, g, h, e = eng.compute(group)
            fo, go, ho, eo = orc.compute(topo, ts, ps, s.pos, lists, group)
            assert _rms_rel(f, fo) <= FORCE_RMS_TOL, (label, group)
            if group != abi.MDX_FORCES_FAST:
                assert _rms_rel(g, go) <= FORCE_RMS_TOL, (label, group)
            if group != abi.MDX_FORCES_SLOW:
                assert _rms_rel(h, ho) <= FORCE_RMS_TOL, (label, group)
            _check(e, eo)
            assert e["npairs_short"] == eo["npairs_short"] and e["npairs_long"] == eo["npairs_long"]
    finally:
        eng.close()


@pytest.mark.parametrize("n_water,n_ion", [(1, 0), (2, 0), (3, 1), (5, 2)])
def test_tiny_and_ragged_systems(n_water, n_ion):
    """3, 6, 11 and 19 atoms in a 20 A box: one partial cluster, a ragged
    last cluster, and (one water) no intermolecular pair."""
    s = systems.water_nacl(n_water, n_ion, 20.0, 100 + n_water, rcut=8.0, rshort=4.0, ewald_a=2.8,
                           ewald_b=9.0, shake=False)
    _run_vs_oracle(s, "w%d_i%d" % (n_water, n_ion))


def test_uncharged_lennard_jones_box():
    """125 neutral Lennard-Jones atoms (argon-like): no k-space work."""
    rng = np.random.default_rng(8)
    argon = systems.MolType("AR", [systems.Site("Ar", (0.0, 0.0, 0.0), 39.948, 0.0, 3.405, 0.996)])
    centres = systems.lattice_sites(125, (24.0, 24.0, 24.0), rng)
    s = systems.assemble("ar125", [argon], [125], centres, rng, box=(24.0, 24.0, 24.0), rcut=9.0, rshort=4.5,
                         ewald_a=2.8, ewald_b=9.0, shake=False)
    _run_vs_oracle(s, "ar125")
