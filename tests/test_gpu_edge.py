"""Edge cases of the HIP engine against the FP64 oracle, through the C ABI:
systems smaller than one 8-atom cluster or with a ragged last cluster, a
single molecule (no intermolecular pair at all), and a box without charges
(Lennard-Jones only: zero structure factors, the force-pass operand scale
falls back to 2^0). Tolerances as in test_gpu_parity.py: force RMS
relative error <= 1e-5, total energy <= 1e-6, components <= 1e-5.
"""
import dataclasses

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


def _check(e, ref, floor=1e-3, skip=()):
    """Total within 1e-6 of the larger of |total| and the sum of component
    magnitudes: in a box of two waters the total (-0.53 kJ/mol) is what is
    left of +11.0 reciprocal and -11.5 self/exclusion energy, so 1e-6 of
    the total alone would ask for 5e-8 of the components. Each component
    within 1e-5 of the larger of its magnitude and `floor` x that sum."""
    scale = sum(abs(ref[k]) for k in ("pe1", "pe2", "pels1", "pels2", "spe")) + 1e-12
    tot, tot_ref = sum(e[k] for k in ENERGY_KEYS), sum(ref[k] for k in ENERGY_KEYS)
    assert abs(tot - tot_ref) <= 1e-6 * max(abs(tot_ref), scale), ("total", tot, tot_ref)
    for k in abi.MdxEnergies.SCALARS:
        if k in skip:
            continue
        assert abs(e[k] - ref[k]) <= 1e-5 * max(abs(ref[k]), floor * scale) + 1e-12, (k, e[k], ref[k])


def _run_vs_oracle(s, label, expect_growth=False, groups=GROUPS, skip=()):
    topo, prm = topology.from_system(s)
    prm["list_mode"] = abi.MDX_LIST_REFERENCE
    eng = ForceEngine(topo, prm)
    try:
        eng.set_positions(s.pos)
        before = eng.memory_bytes()
        eng.chcnb()
        if expect_growth:
            # the overflowing build reallocated larger list buffers
            assert eng.memory_bytes() > before, (before, eng.memory_bytes())
        ts, ps = topo.struct(), topology.params_struct(dict(prm, list_mode=0, skin=0.0))
        lists = orc.chcnb(ts, ps, s.pos)
        for group in groups:
            print("case", label, topo.natoms, "atoms, group", group, flush=True)
            f, g, h, e = eng.compute(group)
            fo, go, ho, eo = orc.compute(topo, ts, ps, s.pos, lists, group)
            assert _rms_rel(f, fo) <= FORCE_RMS_TOL, (label, group)
            if group != abi.MDX_FORCES_FAST:
                assert _rms_rel(g, go) <= FORCE_RMS_TOL, (label, group)
            if group != abi.MDX_FORCES_SLOW:
                assert _rms_rel(h, ho) <= FORCE_RMS_TOL, (label, group)
            _check(e, eo, skip=skip)
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


def test_box_change_keeps_or_renews_the_k_set():
    """mdx_set_box: FURIR keeps the k set of the last INIFUR while every
    edge stays within 5 % of the box it was enumerated at (k vectors
    rescaled with the new box) and enumerates again beyond that
    (forces.f:1250-1265). Positions scale with the box; the oracle gets
    the k set of the matching box."""
    s = systems.water_nacl(216, 2, 19.0, 61, rcut=8.0, rshort=4.0, ewald_a=2.8, ewald_b=9.0, shake=False)
    topo, prm = topology.from_system(s)
    prm["list_mode"] = abi.MDX_LIST_REFERENCE
    box0 = np.array(s.box, float)
    ts = topo.struct()
    k0 = orc.inifur(ts, topology.params_struct(dict(prm, list_mode=0, skin=0.0)), topo.nmoltypes)
    eng = ForceEngine(topo, prm)
    try:
        eng.set_positions(s.pos)
        eng.chcnb()
        kset = k0
        # each factor is relative to box0; the k set follows the last
        # enumeration (box0 until an edge leaves the 5 % window)
        for factor, renew in (((1.02, 1.02, 1.02), False), ((1.0, 1.0, 1.04), False), ((1.07, 1.07, 1.07), True)):
            f3 = np.array(factor)
            box = box0 * f3
            pos = s.pos * f3
            prm2 = dict(prm, box=tuple(box))
            ps2 = topology.params_struct(dict(prm2, list_mode=0, skin=0.0))
            if renew:
                kset = orc.inifur(ts, ps2, topo.nmoltypes)
            print("case box factor", factor, "renew", renew, "nkv", len(kset[0]), flush=True)
            eng.set_box(box)
            eng.set_positions(pos)
            eng.chcnb()
            f, g, h, e = eng.allforce()
            lists = orc.chcnb(ts, ps2, pos)
            fo, go, ho, eo = orc.compute(topo, ts, ps2, pos, lists, abi.MDX_FORCES_ALL, kvec=kset)
            assert _rms_rel(f, fo) <= FORCE_RMS_TOL, factor
            # the reciprocal energy carries the k set: held to 1e-6
            assert abs(e["pels1"] - eo["pels1"]) <= 1e-6 * abs(eo["pels1"]), (e["pels1"], eo["pels1"])
            # at 1.02 the short-range electrostatic sum nearly cancels
            # (-8.8 kJ/mol from ~8,000 pair terms of both signs, 471 at
            # 1.00): its FP32 pair rounding (2.4e-4 kJ/mol, the same
            # absolute size as at 1.00) is held to 1e-5 of 1e-2 of the
            # component sum instead of 1e-3
            _check(e, eo, floor=1e-2)
    finally:
        eng.close()


def test_dense_droplet_in_large_box_regrows_lists():
    """652 molecules (liquid density) in a 27 A cube inside a 100 A box: the list
    capacities, sized from the mean density, are too small for the
    droplet, so the build overflows, grows its buffers and builds again
    (the reference stops instead, forces.f:2188-2194); lists and forces as
    the oracle. The molecular-virial corrections WIRS/WIRSS are left out:
    with no molecule split by the box they vanish to rounding (oracle
    -0.0015, engine -0.0035 kJ/mol for WIRSS, 1250 in the periodic 27 A
    box)."""
    s = systems.water_nacl(650, 1, 27.0, 73, rcut=8.0, rshort=4.0, ewald_a=2.8, ewald_b=6.0, shake=False)
    s = dataclasses.replace(s, box=(100.0, 100.0, 100.0))
    _run_vs_oracle(s, "droplet", expect_growth=True, groups=(abi.MDX_FORCES_ALL,), skip=("wirs", "wirss"))


@pytest.mark.parametrize("ipot", [2, 1])
def test_fspc_water_stretched_bond(ipot):
    """fSPC water (STRBND, forces.f:530-800) with one O-H bond pulled to
    1.45 A, beyond the 1.3 A switch of the Morse force (IPOT 2), and the
    harmonic variant (IPOT 1, records edited: no input sets it)."""
    s = systems.water_nacl(60, 1, 13.0, 17, name="fspc", water=systems.fspc_water(), rcut=6.0, rshort=3.5,
                           ewald_a=2.8, ewald_b=9.0, shake=False)
    pos = s.pos.copy()
    d = pos[1] - pos[0]
    pos[1] = pos[0] + d * (1.45 / np.linalg.norm(d))
    s = dataclasses.replace(s, pos=pos)
    topo, prm = topology.from_system(s)
    assert len(topo.spc_atom) == 60
    topo.spc_par[:, 10] = ipot
    prm["list_mode"] = abi.MDX_LIST_REFERENCE
    eng = ForceEngine(topo, prm)
    try:
        eng.set_positions(s.pos)
        eng.chcnb()
        ts, ps = topo.struct(), topology.params_struct(dict(prm, list_mode=0, skin=0.0))
        lists = orc.chcnb(ts, ps, s.pos)
        print("case fspc ipot", ipot, flush=True)
        f, g, h, e = eng.compute(abi.MDX_FORCES_ALL)
        fo, go, ho, eo = orc.compute(topo, ts, ps, s.pos, lists, abi.MDX_FORCES_ALL)
        assert _rms_rel(f, fo) <= FORCE_RMS_TOL and _rms_rel(h, ho) <= FORCE_RMS_TOL
        _check(e, eo)
    finally:
        eng.close()


def _slab(seed=41):
    """300 SPC/E waters and an ion pair on a jittered 10 x 10 x 3 grid
    (3.1 A spacing, liquid density) in a 31 x 31 x 9.3 A slab centred on
    z = 0 of a 31 x 31 x 100 A box: in wrapped coordinates the slab's two
    faces sit at z ~ 0-5 and z ~ 95-100, so the 8-atom clusters cut from an
    x-y column sorted by z can span almost the whole box height."""
    rng = np.random.default_rng(seed)
    types = [systems.spce_water(), systems.sodium(), systems.chloride()]
    g = (np.arange(10) + 0.5) * 3.1 - 15.5
    gz = (np.arange(3) - 1.0) * 3.1
    grid = np.stack(np.meshgrid(g, g, gz, indexing="ij"), -1).reshape(-1, 3)
    centres = grid[rng.permutation(len(grid))] + rng.uniform(-0.25, 0.25, grid.shape)
    centres = np.concatenate([centres, [[0.0, 0.0, 4.6], [0.0, 15.5, -4.6]]])
    return systems.assemble("slab", types, [300, 1, 1], centres, rng, box=(31.0, 31.0, 100.0), rcut=10.0,
                            rshort=5.0, ewald_a=2.8, ewald_b=9.0, shake=False)


def test_slab_in_tall_box_reference_lists():
    """Cluster image shifts come from the cluster anchors; a cluster that
    spans the periodic seam needs the minimum image per pair.  Reference
    lists (every listed pair evaluated, no cutoff test) against the oracle."""
    _run_vs_oracle(_slab(), "slab", groups=(abi.MDX_FORCES_ALL,), skip=("wirs", "wirss"))


def test_slab_in_tall_box_verlet_displaced():
    """The same slab with a Verlet list (skin 1 A) evaluated after every
    atom moved up to 0.2 A: forces and energies as the oracle with a list
    built at the displaced positions, and the same pair count."""
    s = _slab(43)
    topo, prm = topology.from_system(s)
    prm["list_mode"] = abi.MDX_LIST_VERLET
    prm["skin"] = 1.0
    eng = ForceEngine(topo, prm)
    try:
        eng.set_positions(s.pos)
        eng.chcnb()
        pos2 = s.pos + np.random.default_rng(44).uniform(-0.2, 0.2, s.pos.shape)
        eng.set_positions(pos2)
        print("case slab verlet", topo.natoms, flush=True)
        f, g, h, e = eng.allforce()
        ts, ps = topo.struct(), topology.params_struct(dict(prm, list_mode=0, skin=0.0))
        fo, go, ho, eo = orc.compute(topo, ts, ps, pos2, orc.chcnb(ts, ps, pos2), abi.MDX_FORCES_ALL)
        assert _rms_rel(f, fo) <= FORCE_RMS_TOL
        assert e["npairs_short"] + e["npairs_long"] == eo["npairs_short"] + eo["npairs_long"]
        tot = sum(e[k] for k in ENERGY_KEYS)
        tot_ref = sum(eo[k] for k in ENERGY_KEYS)
        assert abs(tot - tot_ref) <= 1e-6 * abs(tot_ref), (tot, tot_ref)
    finally:
        eng.close()
