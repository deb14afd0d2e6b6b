"""Parity of the HIP engine with the reference (golden fixtures from the
reference program) and with the FP64 oracle, through the C ABI.

Tolerances (stated by the engine's precision contract): pair arithmetic in
FP32 with FP64 accumulation; force RMS relative error <= 1e-5 against FP64,
energies <= 1e-6 relative; neighbour lists bit-exact (same pairs, order
and 1-4 signs after export in the reference layout).
"""
import dataclasses
import hashlib

import numpy as np
import pytest

from tests.helpers import CASES, load_golden, golden_positions, golden_com, golden_forces
from mdynamix_amd import abi, topology, systems
from mdynamix_amd.engine import ForceEngine
from oracle import orc

pytestmark = pytest.mark.gpu

FORCE_RMS_TOL = 1e-5
ENERGY_TOL = 1e-6      # total potential energy
COMPONENT_TOL = 1e-5   # each energy component and virial
ENERGY_KEYS = ("pe1", "pe2", "pels1", "pels2", "spe", "pint")


def rms_rel(a, b):
    return float(np.sqrt(((a - b) ** 2).sum() / max((b ** 2).sum(), 1e-300)))


def energy_scale(ref):
    keys = ("pe1", "pe2", "pels1", "pels2", "spe")
    return sum(abs(ref[k]) for k in keys)


def check_energies(e, ref, keys=abi.MdxEnergies.SCALARS):
    """Total potential energy to 1e-6 relative; every component and virial
    to 1e-5 relative (a component much smaller than the total is held to
    that fraction of the sum of component magnitudes)."""
    scale = energy_scale(ref)
    tot = sum(e[k] for k in ENERGY_KEYS)
    tot_ref = sum(ref[k] for k in ENERGY_KEYS)
    assert abs(tot - tot_ref) <= ENERGY_TOL * abs(tot_ref), ("total", tot, tot_ref)
    for k in keys:
        err = abs(e[k] - ref[k])
        assert err <= COMPONENT_TOL * max(abs(ref[k]), 1e-3 * scale) + 1e-300, (k, e[k], ref[k])


def _digest(a, b):
    return hashlib.sha256(np.ascontiguousarray(a, "<i4").tobytes() + np.ascontiguousarray(b, "<i4").tobytes()).hexdigest()


def _engine_from_golden(name):
    d = load_golden(name)
    topo, prm = topology.from_dump(d, lam=d["lam"])
    eng = ForceEngine(topo, prm)
    return d, topo, prm, eng


@pytest.mark.parametrize("name", CASES)
def test_list_bit_exact_vs_reference(name):
    d, topo, prm, eng = _engine_from_golden(name)
    print("case", name, topo.natoms, flush=True)
    eng.set_positions(golden_positions(d))
    eng.chcnb()
    assert eng.list_counts() == tuple(int(x) for x in d["LISTLEN"][0])
    s1, s2, l1, l2 = eng.get_list()
    assert _digest(s1, s2) == str(d["DIGEST_SHORT"][0])
    assert _digest(l1, l2) == str(d["DIGEST_LONG"][0])


@pytest.mark.parametrize("name", CASES)
def test_forces_energies_vs_reference(name):
    d, topo, prm, eng = _engine_from_golden(name)
    print("case", name, topo.natoms, flush=True)
    eng.set_positions(golden_positions(d))
    eng.chcnb()
    double = int(d["SIZES"][0][0]) == 2
    for call in range(len(d["SX"])):
        eng.set_positions(golden_positions(d, call))
        eng.set_molecule_com(golden_com(d, call))
        ref = dict(zip(abi.MdxEnergies.SCALARS, d["ENERGY"][call]))
        if double:
            _, g, _, es = eng.slforces()
            _, _, h, ef = eng.fforces()
            e = {k: es[k] + ef[k] for k in abi.MdxEnergies.SCALARS}
        else:
            f, g, h, e = eng.allforce()
            assert rms_rel(f, golden_forces(d, "F", call)) <= FORCE_RMS_TOL
        assert rms_rel(g, golden_forces(d, "G", call)) <= FORCE_RMS_TOL
        assert rms_rel(h, golden_forces(d, "H", call)) <= FORCE_RMS_TOL
        check_energies(e, ref)


@pytest.mark.parametrize("name", CASES)
def test_energy_breakdown_vs_reference(name):
    """MDX_WANT_DETAIL: bare Coulomb and LJ energy per molecule-type pair
    (POTE1, POTL1) and per molecule type inside molecules (PES141, PSR141)
    as LOCALF/FORCES accumulate them; each entry within 1e-5 relative (an
    entry much smaller than the rest is held to 1e-5 of 1e-3 x the sum of
    magnitudes)."""
    d, topo, prm, eng = _engine_from_golden(name)
    print("case", name, topo.natoms, flush=True)
    eng.set_positions(golden_positions(d))
    eng.chcnb()
    double = int(d["SIZES"][0][0]) == 2
    nt, npairs = topo.nmoltypes, topo.ntypepairs
    for call in range(len(d["SX"])):
        eng.set_positions(golden_positions(d, call))
        eng.set_molecule_com(golden_com(d, call))
        if double:
            eng.slforces(want=abi.MDX_WANT_DETAIL)
            _, _, _, e = eng.fforces(want=abi.MDX_WANT_DETAIL)   # FFORCES leaves the fast breakdown
        else:
            _, _, _, e = eng.allforce(want=abi.MDX_WANT_DETAIL)
        refs = {"pote1": d["POTE1"][call][:npairs], "potl1": d["POTL1"][call][:npairs],
                "pes141": d["PES141"][call][:nt], "psr141": d["PSR141"][call][:nt]}
        scale = sum(np.abs(v).sum() for v in refs.values())
        for key, ref in refs.items():
            err = np.abs(e[key] - ref)
            assert np.all(err <= COMPONENT_TOL * np.maximum(np.abs(ref), 1e-3 * scale)), (key, e[key], ref)


def _system_engine(s, list_mode=None):
    topo, prm = topology.from_system(s)
    if list_mode is not None:
        prm["list_mode"] = list_mode
    return topo, prm, ForceEngine(topo, prm)


def _oracle(topo, prm, pos, group=abi.MDX_FORCES_ALL):
    ts, ps = topo.struct(), topology.params_struct(dict(prm, list_mode=0, skin=0.0))
    lists = orc.chcnb(ts, ps, pos)
    return orc.compute(topo, ts, ps, pos, lists, group)


def test_verlet_mode_matches_fresh_reference_lists():
    """Verlet list (skin 1 A) evaluated at displaced positions equals the
    reference semantics with a list rebuilt at those positions."""
    s = systems.water_nacl(508, 4, 25.0, 77, rcut=10.0, rshort=5.0, ewald_a=2.8, ewald_b=9.95, skin=1.0)
    topo, prm, eng = _system_engine(s, abi.MDX_LIST_VERLET)
    eng.set_positions(s.pos)
    eng.chcnb()
    rng = np.random.default_rng(5)
    pos2 = s.pos + rng.uniform(-0.15, 0.15, s.pos.shape)
    eng.set_positions(pos2)
    f, g, h, e = eng.allforce()
    fo, go, ho, eo = _oracle(topo, prm, pos2)
    assert rms_rel(f, fo) <= FORCE_RMS_TOL
    # A Verlet list keeps the short/long class of every pair from its build
    # (the split of an r-RESPA step must not change between rebuilds), so
    # pairs that crossed R_short since then sit in the other class than in
    # a list rebuilt at pos2: compare the class sums.
    merged = lambda d: dict(d, pe1=d["pe1"] + d["pe2"], pe2=0.0, pels1=d["pels1"] + d["pels2"], pels2=0.0,
                            vir1=d["vir1"] + d["vir2"], vir2=0.0, virx=d["virx"] + d["virfx"],
                            viry=d["viry"] + d["virfy"], virz=d["virz"] + d["virfz"])
    check_energies(merged(e), merged(eo), keys=("pe1", "pels1", "spe", "pint", "vir1", "virb", "virx", "viry", "virz"))
    assert e["npairs_short"] + e["npairs_long"] == eo["npairs_short"] + eo["npairs_long"]


def test_chain_mixture_groups_vs_oracle():
    """Bonded terms, 1-4 pairs and the fast/slow split on a flexible-chain
    mixture (kJ/mol units from the host parameter preparation)."""
    s = systems.chain_mixture(6, 120, 20.0, 11, chain_sites=16, rcut=9.0, rshort=4.5, ewald_a=2.8, ewald_b=9.0,
                              shake=False)
    topo, prm, eng = _system_engine(s)
    eng.set_positions(s.pos)
    eng.chcnb()
    for group in (abi.MDX_FORCES_ALL, abi.MDX_FORCES_FAST, abi.MDX_FORCES_SLOW):
        f, g, h, e = eng.compute(group)
        fo, go, ho, eo = _oracle(topo, prm, s.pos, group)
        assert rms_rel(f, fo) <= FORCE_RMS_TOL, group
        check_energies(e, eo)


def test_ewald_lambda_coupled_solute_vs_oracle():
    s = systems.ee_job(1500, 24, 0.6, 21)
    topo, prm, eng = _system_engine(s, abi.MDX_LIST_REFERENCE)
    eng.set_positions(s.pos)
    eng.chcnb()
    f, g, h, e = eng.allforce()
    fo, go, ho, eo = _oracle(topo, prm, s.pos)
    assert rms_rel(f, fo) <= FORCE_RMS_TOL
    check_energies(e, eo)


def test_repeat_call_is_bit_identical():
    """Fixed-order sums: the same call twice gives identical bits."""
    d, topo, prm, eng = _engine_from_golden("water_508")
    eng.set_positions(golden_positions(d))
    eng.chcnb()
    f1, _, _, e1 = eng.allforce()
    f2, _, _, e2 = eng.allforce()
    assert np.array_equal(f1, f2)
    assert all(e1[k] == e2[k] for k in abi.MdxEnergies.SCALARS)


def test_autotuned_shape_keeps_parity():
    """The autotuner only changes launch shapes and the cluster grid: the
    reference-layout list it rebuilds is the same, and forces and energies
    stay within the tolerances."""
    d, topo, prm, eng = _engine_from_golden("water_508")
    eng.set_positions(golden_positions(d))
    eng.chcnb()
    lists0 = eng.get_list()
    shape = eng.autotune()
    print("case water_508 tuned", shape, eng.cluster_grid(), flush=True)
    assert shape["chunk"] > 0 and shape["nslice"] > 0 and shape["column_width"] > 0.0
    assert shape["pair_waves"] in (5, 6)
    for a, b in zip(lists0, eng.get_list()):
        assert np.array_equal(a, b)
    eng.set_molecule_com(golden_com(d, 0))
    f, g, h, e = eng.allforce()
    assert rms_rel(f, golden_forces(d, "F", 0)) <= FORCE_RMS_TOL
    check_energies(e, dict(zip(abi.MdxEnergies.SCALARS, d["ENERGY"][0])))


def test_pair_occupancy_choices_agree():
    """The plain pair kernel at five and at six wavefronts per SIMD (the
    autotuner's two choices) is the same arithmetic: identical forces and
    energies, through the eager call and the captured graph."""
    d, topo, prm, eng = _engine_from_golden("water_508")
    eng.set_positions(golden_positions(d))
    eng.chcnb()
    eng.set_molecule_com(golden_com(d, 0))
    out = {}
    for w in (6, 5):
        eng.set_pair_waves(w)
        assert eng.pair_waves() == w
        for timing in (True, False):
            eng.set_timing(timing)
            f, g, h, e = eng.allforce()
            print("case water_508 pair_waves", w, "timing", timing, flush=True)
            out[(w, timing)] = (f.copy(), g.copy(), h.copy(), dict(e))
    ref = out[(6, True)]
    for k, v in out.items():
        for a, b in zip(v[:3], ref[:3]):
            assert np.array_equal(a, b), k
        for name in ref[3]:
            assert np.array_equal(np.asarray(v[3][name]), np.asarray(ref[3][name])), (k, name)
    eng.set_timing(True)
    with pytest.raises(Exception):
        eng.set_pair_waves(4)


def test_all_bonded_kinds_vs_oracle():
    """Morse bonds and cosine, Ryckaert-Bellemans and improper torsions
    (the oracle's forms are pinned by tests/test_oracle_math.py) on a box of
    charged 6-site chains, all three force groups."""
    rng = np.random.default_rng(8)
    geo = systems._chain_geometry(6, rng, 1.53, 112.0)
    q = [0.2, -0.1, 0.0, 0.1, -0.3, 0.1]
    sites = [systems.Site("C%d" % (i + 1), tuple(geo[i]), 14.0, q[i], 3.6, 0.4) for i in range(6)]
    bonds = [(1 if i == 2 else 0, i + 1, i + 2, 1.53 + 0.01 * i, 1200.0 + 10 * i, 300.0, 1.9) for i in range(5)]
    angles = [(i + 1, i + 2, i + 3, 112.0 + i, 250.0 + 5 * i) for i in range(4)]
    tors = [(i + 1, i + 2, i + 3, i + 4, 15.0 * (i + 1), 2.0 + i, 3) for i in range(3)]
    mt = systems.MolType("CH6", sites, bonds, angles, tors, c14el=0.5, c14lj=0.5)
    centres = systems.lattice_sites(27, (24.0, 24.0, 24.0), rng)
    s = systems.assemble("kinds", [mt], [27], centres, rng, box=(24.0, 24.0, 24.0), rcut=9.0, rshort=4.5,
                         ewald_a=2.8, ewald_b=9.0, shake=False)
    # off the built geometry's exact trans dihedrals (sin phi = 0, where the
    # reference's force formula degenerates; see the test below)
    s = dataclasses.replace(s, pos=s.pos + rng.normal(0.0, 0.03, s.pos.shape))
    topo, prm = topology.from_system(s)
    assert len(topo.tp_kind) >= 3 and len(topo.bp_morse) >= 3
    topo.bp_morse[:] = 0
    topo.bp_morse[2] = 1
    topo.tp_kind[:3] = [0, 5, -1]
    topo.tp_k1[1], topo.tp_k2[1], topo.tp_k3[1], topo.tp_k4[1], topo.tp_k5[1] = 3.0, -2.0, 1.5, -0.5, 0.25
    eng = ForceEngine(topo, prm)
    eng.set_positions(s.pos)
    eng.chcnb()
    print("case bonded kinds", topo.natoms, flush=True)
    for group in (abi.MDX_FORCES_ALL, abi.MDX_FORCES_FAST, abi.MDX_FORCES_SLOW):
        f, g, h, e = eng.compute(group)
        fo, go, ho, eo = _oracle(topo, prm, s.pos, group)
        assert rms_rel(f, fo) <= FORCE_RMS_TOL, group
        check_energies(e, eo)
        if group != abi.MDX_FORCES_SLOW:
            assert abs(e["pint"] - eo["pint"]) <= 1e-9 * max(abs(eo["pint"]), 1.0)


@pytest.mark.parametrize("kind", [0, 5])
def test_dihedral_at_exact_trans_is_the_energy_gradient(kind):
    """A 4-site chain at exactly 180 degrees (sin phi = 0) with a periodic
    dihedral form whose dE/dphi is not zero there (cosine with delta = 15
    deg, the cos^n series with delta = 30 deg; the improper form K (phi -
    phi0)^2 has its branch cut at 180 degrees).  The
    reference's GETTRS force divides d(cos phi)/dr, which vanishes at 180
    degrees, by sin phi clamped to 1e-10 (forces.f:383-388) and loses the
    torque there; the engine's gradient of phi has no 1/sin phi, so its
    force is the energy gradient.  The chain lies in the z = 0 plane with
    its bonds and angles at their equilibria: the torque is along z and
    checked against central differences of the oracle's bonded energy (no
    charges, no Lennard-Jones, a single molecule) to 1e-6 relative, and the
    in-plane forces vanish by symmetry (in-plane differences of the
    oracle's energy are not used: its phi comes from acos near -1, whose
    rounding moves phi by ~1e-8 rad under in-plane steps)."""
    sites = [systems.Site("C%d" % (i + 1), (0.0, 0.0, 0.0), 14.0, 0.0, 0.0, 0.0) for i in range(4)]
    bonds = [(0, i + 1, i + 2, 1.53, 1200.0, 0.0, 0.0) for i in range(3)]
    angles = [(i + 1, i + 2, i + 3, 112.0, 250.0) for i in range(2)]
    delta = {0: 15.0, 5: 30.0}[kind]
    tors = [(1, 2, 3, 4, delta, 2.0, 3)]
    mt = systems.MolType("C4T", sites, bonds, angles, tors, c14el=0.0, c14lj=0.0)
    th = np.radians(112.0)
    pos = np.array([[0.0, 0.0, 0.0], [1.53, 0.0, 0.0],
                    [1.53 - 1.53 * np.cos(th), 1.53 * np.sin(th), 0.0],
                    [3.06 - 1.53 * np.cos(th), 1.53 * np.sin(th), 0.0]])
    pos[3] = pos[2] + (pos[1] - pos[0])          # anti (trans) in the plane
    pos += 10.0
    s = systems.System("trans", [mt], [1], (20.0, 20.0, 20.0), pos, rcut=8.0, rshort=4.0, ewald_a=2.8,
                       ewald_b=9.0, shake=False)
    topo, prm = topology.from_system(s)
    topo.tp_kind[0] = kind
    if kind == 5:
        topo.tp_k1[0], topo.tp_k2[0], topo.tp_k3[0], topo.tp_k4[0], topo.tp_k5[0] = 3.0, -2.0, 1.5, -0.5, 0.25
    eng = ForceEngine(topo, prm)
    try:
        eng.set_positions(pos)
        eng.chcnb()
        f, _, _, e = eng.fforces()
    finally:
        eng.close()
    ts, ps = topo.struct(), topology.params_struct(dict(prm, list_mode=0, skin=0.0))

    def energy(p):
        return orc.compute(topo, ts, ps, p, orc.chcnb(ts, ps, p), abi.MDX_FORCES_FAST)[3]["pint"]
    h = 1e-5
    gz = np.zeros(4)
    for a in range(4):
        pp, pm = pos.copy(), pos.copy()
        pp[a, 2] += h
        pm[a, 2] -= h
        gz[a] = (energy(pp) - energy(pm)) / (2 * h)
    scale = np.abs(gz).max()
    print("case trans kind", kind, "max |F_z + dE/dz|", np.abs(f[:, 2] + gz).max(), "of", scale, flush=True)
    assert scale > 1e-3
    assert np.abs(f[:, 2] + gz).max() <= 1e-6 * scale
    assert np.abs(f[:, :2]).max() <= 1e-6 * scale
