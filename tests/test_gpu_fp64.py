"""The engine's FP64 precision mode (MDX_PRECISION_FP64): the lists,
classes and outputs of the default mixed-precision path with every pair
term (k_pairs64) and the reciprocal sum (the FP64 FURIR kernels) in FP64.

Against the reference program's own fixtures (the five golden systems),
the FP64 oracle (a flexible-chain mixture per force group, a
lambda-coupled solute with soft-core pairs) and the oracle's fixture of the
32k SPC/E + NaCl box at its full size: force RMS relative error <= 1e-10
(F, G, H), every energy component and virial <= 1e-10 relative (a
component much smaller than the rest is held to 1e-10 of 1e-3 x the sum of
component magnitudes), the per-type breakdown likewise, pair counts equal.
At the full-size box the mixed path is compared with the FP64 GPU forces
as well (held to its own 1e-5 contract).  The residual differences are
rounding: libm against the device's exp and sincospi, summation order.
"""
import numpy as np
import pytest

from tests.helpers import CASES, load_golden, golden_positions, golden_com, golden_forces
from tests.test_gpu_parity import rms_rel, energy_scale, FORCE_RMS_TOL
from mdynamix_amd import abi, topology, systems
from mdynamix_amd.engine import ForceEngine, ForceBatch
from oracle import orc

pytestmark = pytest.mark.gpu

TOL = 1e-10
FP64 = abi.MDX_PRECISION_FP64


def check_close(e, ref, keys=abi.MdxEnergies.SCALARS, tol=TOL):
    scale = energy_scale(ref)
    for k in keys:
        err = abs(e[k] - ref[k])
        assert err <= tol * max(abs(ref[k]), 1e-3 * scale) + 1e-300, (k, e[k], ref[k])


def _golden_engine(name, precision=FP64):
    d = load_golden(name)
    topo, prm = topology.from_dump(d, lam=d["lam"])
    return d, topo, prm, ForceEngine(topo, dict(prm, precision=precision))


@pytest.mark.parametrize("name", CASES)
def test_fp64_forces_energies_vs_reference(name):
    d, topo, prm, eng = _golden_engine(name)
    try:
        print("case fp64", name, topo.natoms, flush=True)
        eng.set_positions(golden_positions(d))
        eng.chcnb()
        double = int(d["SIZES"][0][0]) == 2
        worst = 0.0
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
                worst = max(worst, rms_rel(f, golden_forces(d, "F", call)))
                assert rms_rel(f, golden_forces(d, "F", call)) <= TOL
            worst = max(worst, rms_rel(g, golden_forces(d, "G", call)), rms_rel(h, golden_forces(d, "H", call)))
            assert rms_rel(g, golden_forces(d, "G", call)) <= TOL
            assert rms_rel(h, golden_forces(d, "H", call)) <= TOL
            check_close(e, ref)
        print("fp64 %s: worst force rms rel %.2e" % (name, worst), flush=True)
    finally:
        eng.close()


@pytest.mark.parametrize("name", CASES)
def test_fp64_energy_breakdown_vs_reference(name):
    d, topo, prm, eng = _golden_engine(name)
    try:
        print("case fp64 detail", name, flush=True)
        eng.set_positions(golden_positions(d))
        eng.chcnb()
        double = int(d["SIZES"][0][0]) == 2
        nt, npairs = topo.nmoltypes, topo.ntypepairs
        for call in range(len(d["SX"])):
            eng.set_positions(golden_positions(d, call))
            eng.set_molecule_com(golden_com(d, call))
            if double:
                eng.slforces(want=abi.MDX_WANT_DETAIL)
                _, _, _, e = eng.fforces(want=abi.MDX_WANT_DETAIL)
            else:
                _, _, _, e = eng.allforce(want=abi.MDX_WANT_DETAIL)
            refs = {"pote1": d["POTE1"][call][:npairs], "potl1": d["POTL1"][call][:npairs],
                    "pes141": d["PES141"][call][:nt], "psr141": d["PSR141"][call][:nt]}
            scale = sum(np.abs(v).sum() for v in refs.values())
            for key, ref in refs.items():
                err = np.abs(e[key] - ref)
                assert np.all(err <= TOL * np.maximum(np.abs(ref), 1e-3 * scale)), (key, e[key], ref)
    finally:
        eng.close()


def _oracle(topo, prm, pos, group=abi.MDX_FORCES_ALL):
    ts, ps = topo.struct(), topology.params_struct(dict(prm, list_mode=0, skin=0.0))
    lists = orc.chcnb(ts, ps, pos)
    return orc.compute(topo, ts, ps, pos, lists, group)


def test_fp64_groups_and_soft_core_vs_oracle():
    """FFORCES / SLFORCES / ALLFORCE on a flexible-chain mixture (bonded
    terms, 1-4 pairs) and a lambda-coupled solute at lambda = 0.6 (soft-core
    shift RAE in the pairs touching it), against the FP64 oracle."""
    s = systems.chain_mixture(6, 120, 20.0, 11, chain_sites=16, rcut=9.0, rshort=4.5, ewald_a=2.8, ewald_b=9.0,
                              shake=False)
    topo, prm = topology.from_system(s)
    eng = ForceEngine(topo, dict(prm, precision=FP64))
    try:
        print("case fp64 chain mixture", topo.natoms, flush=True)
        eng.set_positions(s.pos)
        eng.chcnb()
        for group in (abi.MDX_FORCES_ALL, abi.MDX_FORCES_FAST, abi.MDX_FORCES_SLOW):
            f, g, h, e = eng.compute(group)
            fo, go, ho, eo = _oracle(topo, prm, s.pos, group)
            assert rms_rel(f, fo) <= TOL, group
            check_close(e, eo)
    finally:
        eng.close()
    s = systems.ee_job(1500, 24, 0.6, 21)
    topo, prm = topology.from_system(s)
    prm = dict(prm, list_mode=abi.MDX_LIST_REFERENCE, skin=0.0)
    eng = ForceEngine(topo, dict(prm, precision=FP64))
    try:
        print("case fp64 lambda-coupled solute", topo.natoms, flush=True)
        eng.set_positions(s.pos)
        eng.chcnb()
        f, g, h, e = eng.allforce()
        fo, go, ho, eo = _oracle(topo, prm, s.pos)
        assert rms_rel(f, fo) <= TOL
        assert rms_rel(g, go) <= TOL and rms_rel(h, ho) <= TOL
        check_close(e, eo)
    finally:
        eng.close()


def test_fp64_verlet_lists_at_displaced_positions_vs_oracle():
    """Verlet list (1 A skin) evaluated after every atom moved: the FP64
    cutoff test gives the pair set of a list rebuilt at the new positions."""
    s = systems.water_nacl(508, 4, 25.0, 77, rcut=10.0, rshort=5.0, ewald_a=2.8, ewald_b=9.95, skin=1.0)
    topo, prm = topology.from_system(s)
    prm = dict(prm, list_mode=abi.MDX_LIST_VERLET)
    eng = ForceEngine(topo, dict(prm, precision=FP64))
    try:
        eng.set_positions(s.pos)
        eng.chcnb()
        pos2 = s.pos + np.random.default_rng(5).uniform(-0.15, 0.15, s.pos.shape)
        eng.set_positions(pos2)
        f, g, h, e = eng.allforce()
        fo, go, ho, eo = _oracle(topo, prm, pos2)
        assert rms_rel(f, fo) <= TOL
        assert e["npairs_short"] + e["npairs_long"] == eo["npairs_short"] + eo["npairs_long"]
        merged = lambda d: dict(d, pe1=d["pe1"] + d["pe2"], pe2=0.0, pels1=d["pels1"] + d["pels2"], pels2=0.0,
                                vir1=d["vir1"] + d["vir2"], vir2=0.0, virx=d["virx"] + d["virfx"],
                                viry=d["viry"] + d["virfy"], virz=d["virz"] + d["virfz"])
        check_close(merged(e), merged(eo), keys=("pe1", "pels1", "spe", "pint", "vir1", "virb", "virx", "viry", "virz"))
    finally:
        eng.close()


def test_fp64_full_size_box_vs_oracle_and_mixed_path():
    """The 32k SPC/E + NaCl box (96,174 atoms, 48,052 k vectors) in both
    modes with the reference lists: the FP64 GPU forces and energies
    against the oracle's FP64 fixture; the mixed path's forces against the
    FP64 GPU ones."""
    from tests.test_gpu_spce_nacl_32k import _fixture, _system
    fx = _fixture()
    s = _system(fx)
    topo, prm = topology.from_system(s)
    prm = dict(prm, list_mode=abi.MDX_LIST_REFERENCE, skin=0.0)
    ref = dict(zip(abi.MdxEnergies.SCALARS, fx["ENERGY_0"]))
    out = {}
    for precision in (FP64, abi.MDX_PRECISION_MIXED):
        eng = ForceEngine(topo, dict(prm, precision=precision))
        try:
            print("case full-size box precision", precision, flush=True)
            eng.set_positions(s.pos)
            eng.chcnb()
            out[precision] = eng.allforce()
        finally:
            eng.close()
    f64, g64, h64, e64 = out[FP64]
    fmx, gmx, hmx, emx = out[abi.MDX_PRECISION_MIXED]
    assert (e64["npairs_short"], e64["npairs_long"]) == tuple(int(v) for v in fx["NPAIRS_0"])
    err64 = rms_rel(f64, fx["F_0"])
    errmx = rms_rel(fmx, f64)
    print("full-size box: FP64 GPU vs oracle F rms %.2e; mixed vs FP64 GPU F %.2e, G %.2e, H %.2e" % (
        err64, errmx, rms_rel(gmx, g64), rms_rel(hmx, h64)), flush=True)
    assert err64 <= TOL
    check_close(e64, ref)
    assert errmx <= FORCE_RMS_TOL
    assert rms_rel(gmx, g64) <= FORCE_RMS_TOL and rms_rel(hmx, h64) <= FORCE_RMS_TOL


def test_fp64_batch_with_mixed_system_matches_alone_and_graph():
    """A batch holding an FP64-mode system beside a mixed-mode one: each
    system's forces and energies carry the bits of its call alone, eager
    and from the captured graph, call after call; repeated calls are
    bit-identical (integer j-reaction sums)."""
    d1, _, _, e64 = _golden_engine("ee_tiny")
    d2, _, _, emx = _golden_engine("water_508", precision=abi.MDX_PRECISION_MIXED)
    try:
        for d, e in ((d1, e64), (d2, emx)):
            e.set_positions(golden_positions(d))
            e.chcnb()
        alone = []
        for e in (e64, emx):
            e.set_timing(True)
            f, g, h, en = e.compute(abi.MDX_FORCES_ALL)
            f2, g2, h2, en2 = e.compute(abi.MDX_FORCES_ALL)
            assert np.array_equal(f, f2) and np.array_equal(g, g2) and np.array_equal(h, h2)
            alone.append(((f, g, h), en))
        for graph in (False, True):
            batch = ForceBatch([e64, emx], graph=graph)
            try:
                for rep in range(3):
                    eb = batch.compute(abi.MDX_FORCES_ALL)
                    for s, e in enumerate((e64, emx)):
                        for x, y in zip(e.get_forces(), alone[s][0]):
                            assert np.array_equal(x, y), (graph, rep, s)
                        for k in abi.MdxEnergies.SCALARS + ("npairs_short", "npairs_long"):
                            assert eb[s][k] == alone[s][1][k], (graph, rep, s, k)
            finally:
                batch.close()
    finally:
        e64.close()
        emx.close()
