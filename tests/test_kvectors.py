"""The engine's k set (INIFUR, forces.f:1356-1427) is index work and must be
bit-exact: its host-side enumeration, called through the C ABI without a
device, against the RKX/RKY/RKZ arrays the reference program wrote for every
golden fixture, and against the oracle's INIFUR at the parameters of the
96,174-atom box (R_cut 12, alpha R_cut 3.2, FEXP 11.5, 98.7 A), whose
count 48,052 the sphere n^2 (2 pi / L)^2 <= 4 alpha^2 FEXP fixes.  A
missing or extra vector on the sphere's edge fails these tests even where
it would move the energy by less than the force tolerance."""
import numpy as np
import pytest

from tests.helpers import CASES, load_golden
from mdynamix_amd import systems, topology
from mdynamix_amd.engine import enumerate_kvectors
from oracle import orc

TWOPI = 2.0 * 3.1415926536       # the reference's PI (prcm.h)


def _as_rk(n):
    return TWOPI * n[:, 0].astype(np.float64), TWOPI * n[:, 1].astype(np.float64), TWOPI * n[:, 2].astype(np.float64)


@pytest.mark.parametrize("name", CASES)
def test_engine_kset_equals_reference_fixture(name):
    d = load_golden(name)
    topo, prm = topology.from_dump(d, lam=d["lam"])
    n = enumerate_kvectors(prm)
    rkx, rky, rkz = _as_rk(n)
    np.testing.assert_array_equal(rkx, d["RKX"][0])
    np.testing.assert_array_equal(rky, d["RKY"][0])
    np.testing.assert_array_equal(rkz, d["RKZ"][0])


def test_engine_kset_at_the_32k_box():
    s = systems.cfg_large()
    topo, prm = topology.from_system(s)
    n = enumerate_kvectors(prm)
    assert len(n) == 48052
    ts, ps = topo.struct(), topology.params_struct(dict(prm, list_mode=0, skin=0.0))
    rkx, rky, rkz, _, _ = orc.inifur(ts, ps, topo.nmoltypes)
    ex, ey, ez = _as_rk(n)
    np.testing.assert_array_equal(ex, rkx)
    np.testing.assert_array_equal(ey, rky)
    np.testing.assert_array_equal(ez, rkz)
    # every vector inside the sphere, every integer vector of the half
    # space inside the sphere present (brute force over the cube)
    alpha, fexp, L = prm["alpha"], prm["fexp"], np.asarray(prm["box"])
    cut2 = 4.0 * alpha * alpha * fexp
    k2 = ((TWOPI * n / L) ** 2).sum(1)
    assert np.all(k2 <= cut2)
    km = int(np.sqrt(cut2) * L[0] / TWOPI + 1.0)
    g = np.arange(-km, km + 1)
    cube = np.stack(np.meshgrid(g, g, g, indexing="ij"), -1).reshape(-1, 3)
    half = (cube[:, 0] > 0) | ((cube[:, 0] == 0) & ((cube[:, 1] > 0) | ((cube[:, 1] == 0) & (cube[:, 2] > 0))))
    inside = (((TWOPI * cube / L) ** 2).sum(1) <= cut2) & half
    assert inside.sum() == len(n)
    assert set(map(tuple, cube[inside])) == set(map(tuple, n))


def test_single_molecule_has_no_kvectors_but_parameters_do():
    """FURIR returns at once for one molecule (forces.f:1266); the
    parameter-only enumeration does not know the topology."""
    s = systems.water_nacl(1, 0, 20.0, 3, rcut=8.0, rshort=4.0, ewald_a=2.8, ewald_b=9.0, shake=False)
    topo, prm = topology.from_system(s)
    assert len(enumerate_kvectors(prm)) > 0
