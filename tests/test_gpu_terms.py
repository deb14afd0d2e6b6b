"""Per-template bonded sums BR/BE, AR/AE, TR/TE (MDX_WANT_TERMS,
mdx_get_terms) against the values the reference program's GETBND, STRBND,
GETANG and GETTRS left in COMMON at each dumped force call
(forces.f:119-182, 655-663, 232-273, 336-447; the golden fixtures hold
them per call).

Tolerance: each entry within 1e-10 relative, plus 3e-8 rad for angles and
dihedral angles and 1e-12 of the fixture's largest bonded energy for
energies: the reference takes the angles as acos of a cosine
(forces.f:265, 385), which near cos = +-1 (trans dihedrals, 180 deg)
resolves an angle only to sqrt(2 eps) ~ 2e-8 rad where the engine's atan2
of the dihedral stays exact, and the energy of such a dihedral (zero at
trans for these forms) then keeps a residue of K O(2e-8^2) ~ 1e-15 of its
scale.  Otherwise the differences are rounding (libm against the device's
sqrt, acos and atan2, summation order).
"""
import numpy as np
import pytest

from tests.helpers import CASES, load_golden, golden_positions, golden_com
from mdynamix_amd import abi, topology
from mdynamix_amd.engine import ForceEngine

pytestmark = pytest.mark.gpu

TOL = 1e-10
ACOS_FLOOR = 3e-8
KEYS = (("br", "BR"), ("be", "BE"), ("ar", "AR"), ("ae", "AE"), ("tr", "TR"), ("te", "TE"))


def _terms_call(eng, double):
    if double:
        eng.slforces()
        eng.fforces(want=abi.MDX_WANT_TERMS)
    else:
        eng.allforce(want=abi.MDX_WANT_TERMS)
    return eng.terms()


@pytest.mark.parametrize("name", CASES)
def test_terms_vs_reference(name):
    d = load_golden(name)
    topo, prm = topology.from_dump(d, lam=d["lam"])
    eng = ForceEngine(topo, prm)
    try:
        print("case terms", name, flush=True)
        eng.set_positions(golden_positions(d))
        eng.chcnb()
        double = int(d["SIZES"][0][0]) == 2
        escale = max([np.abs(d[K][c]).max(initial=0.0) for c in range(len(d["SX"])) for K in ("BE", "AE", "TE")])
        nonzero = 0
        for call in range(len(d["SX"])):
            eng.set_positions(golden_positions(d, call))
            eng.set_molecule_com(golden_com(d, call))
            t = _terms_call(eng, double)
            for k, K in KEYS:
                ref = d[K][call]
                assert t[k].shape == ref.shape, (k, t[k].shape, ref.shape)
                err = np.abs(t[k] - ref)
                tol = TOL * np.abs(ref) + (ACOS_FLOOR if k in ("ar", "tr") else 0.0) + \
                    (1e-12 * escale if k in ("be", "ae", "te") else 0.0)
                assert np.all(err <= tol), (name, call, k, t[k] - ref, ref)
                nonzero += int(np.count_nonzero(ref))
        assert nonzero > 0
    finally:
        eng.close()


def test_terms_update_only_on_fast_calls_with_the_flag():
    """mdx_get_terms keeps the sums of the last fast-group call made with
    MDX_WANT_TERMS: a slow-group call or a call without the flag leaves
    them, as SLFORCES leaves BR..TE (only GETBND/GETANG/GETTRS set them)."""
    d = load_golden("chains_tiny")
    topo, prm = topology.from_dump(d, lam=d["lam"])
    eng = ForceEngine(topo, prm)
    try:
        eng.set_positions(golden_positions(d))
        eng.chcnb()
        t0 = eng.terms()
        assert all(not np.any(v) for v in t0.values())
        eng.allforce(want=abi.MDX_WANT_TERMS)
        t1 = eng.terms()
        eng.set_positions(golden_positions(d) + 0.01)
        eng.allforce()
        eng.slforces(want=abi.MDX_WANT_TERMS)
        t2 = eng.terms()
        for k in t1:
            assert np.array_equal(t1[k], t2[k]), k
        eng.set_positions(golden_positions(d) + np.random.default_rng(2).normal(0.0, 0.02, (eng.n, 3)))
        eng.fforces(want=abi.MDX_WANT_TERMS)
        t3 = eng.terms()
        assert not np.array_equal(t1["br"], t3["br"])
        eng.fforces(want=abi.MDX_WANT_TERMS)
        t4 = eng.terms()
        for k in t3:
            assert np.array_equal(t3[k], t4[k]), k
    finally:
        eng.close()
