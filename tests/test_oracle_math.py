"""Mathematical checks of the oracle's bonded terms that the reference runs
(golden fixtures) do not exercise: Morse bonds with the harmonic fallback
(GETBND, forces.f:159-177), cosine (type 0), Ryckaert-Bellemans (type 5)
and improper (type -1) torsions (GETTRS, forces.f:393-445), next to the
harmonic bond/angle and tors1 terms the fixtures cover.

* Forces are the exact negative gradient of the energy: central finite
  differences of the oracle's PINT against its fast forces H.
* Conventions fixed by the mathematics: the IUPAC dihedral (cis = 0,
  trans = 180 degrees, sign from the handedness), bond and angle minima,
  the Morse well depth, values at special angles.

Each geometry is one molecule in a large box with zero charges and LJ, so
PINT and H are the bonded terms alone.
"""
import math

import numpy as np
import pytest

from mdynamix_amd import abi, systems, topology
from oracle import orc


def _molecule(coords, bonds=(), angles=(), torsions=()):
    sites = [systems.Site("A%d" % k, tuple(c), 12.0 + k, 0.0, 3.0, 0.0) for k, c in enumerate(coords)]
    mt = systems.MolType("TST", sites, list(bonds), list(angles), list(torsions))
    pos = np.array(coords, dtype=np.float64)
    s = systems.System(name="tst", types=[mt], counts=[1], box=(30.0, 30.0, 30.0), pos=pos, rcut=8.0,
                       rshort=4.0, ewald_a=2.8, ewald_b=9.0, shake=False)
    topo, prm = topology.from_system(s)
    return s, topo, prm


def _eval(topo, prm, pos):
    ts = topo.struct()
    ps = topology.params_struct(dict(prm, list_mode=0, skin=0.0))
    lists = orc.chcnb(ts, ps, pos)
    f, g, h, e = orc.compute(topo, ts, ps, pos, lists, abi.MDX_FORCES_FAST)
    return e["pint"], h


def _fd_check(topo, prm, pos, step=1e-5, rtol=1e-6):
    e0, h = _eval(topo, prm, pos)
    fd = np.zeros_like(pos)
    for a in range(pos.shape[0]):
        for k in range(3):
            p = pos.copy(); p[a, k] += step
            m = pos.copy(); m[a, k] -= step
            fd[a, k] = -(_eval(topo, prm, p)[0] - _eval(topo, prm, m)[0]) / (2 * step)
    scale = max(np.abs(h).max(), 1e-12)
    # absolute floor: central-difference truncation h^2 E'''/6 where the force vanishes
    assert np.abs(fd - h).max() <= rtol * scale + 1e-6, (fd, h)
    assert np.abs(h.sum(0)).max() <= 1e-9 * scale + 1e-12        # Newton's third law
    return e0


def _dihedral_coords(phi_deg, b=1.5, theta=110.0):
    """i-j-k-l with bond length b, bond angles theta and dihedral phi:
    j at the origin, k on +x, i in the xy plane on the +y side, l rotated by
    phi about the j-k axis (phi = 0: cis)."""
    th, ph = math.radians(theta), math.radians(phi_deg)
    j = np.zeros(3)
    k = np.array([b, 0.0, 0.0])
    i = j + b * np.array([math.cos(th), math.sin(th), 0.0])
    l = k + b * np.array([-math.cos(th), math.sin(th) * math.cos(ph), math.sin(th) * math.sin(ph)])
    return [tuple(i + 10.0), tuple(j + 10.0), tuple(k + 10.0), tuple(l + 10.0)]


def _torsion_topo(phi_deg, kind, **par):
    coords = _dihedral_coords(phi_deg)
    s, topo, prm = _molecule(coords, torsions=[(1, 2, 3, 4, 0.0, 1.0, 1)])
    topo.tp_kind[:] = kind
    for name, v in par.items():
        getattr(topo, name)[:] = v
    return s, topo, prm


def test_dihedral_convention_iupac():
    """cis = 0, trans = 180 degrees: E = K (1 + cos(phi)) is 2K at cis and 0
    at trans; 90 and -90 differ only through delta."""
    for phi, want in ((0.0, 2.0), (180.0, 0.0), (90.0, 1.0), (-90.0, 1.0)):
        s, topo, prm = _torsion_topo(phi, 0, tp_k=1.0, tp_mult=1, tp_delta=0.0)
        e, _ = _eval(topo, prm, s.pos)
        assert e == pytest.approx(want, abs=1e-9), phi
    # the sign of phi: delta = 90 degrees gives 1 + cos(phi - 90) = 1 + sin(phi)
    for phi, want in ((90.0, 2.0), (-90.0, 0.0)):
        s, topo, prm = _torsion_topo(phi, 0, tp_k=1.0, tp_mult=1, tp_delta=math.pi / 2)
        e, _ = _eval(topo, prm, s.pos)
        assert e == pytest.approx(want, abs=1e-9), phi


@pytest.mark.parametrize("phi", [35.0, 100.0, -150.0])
def test_cosine_torsion_type0(phi):
    s, topo, prm = _torsion_topo(phi, 0, tp_k=2.5, tp_mult=3, tp_delta=math.radians(20.0))
    e = _fd_check(topo, prm, s.pos)
    assert e == pytest.approx(2.5 * (1 + math.cos(3 * math.radians(phi) - math.radians(20.0))), rel=1e-9)


@pytest.mark.parametrize("phi", [40.0, 170.0, -75.0])
def test_ryckaert_bellemans_type5(phi):
    c = [9.28, 12.16, -13.12, -3.06, 26.24, -31.5]
    s, topo, prm = _torsion_topo(phi, 5, tp_k=c[0], tp_k1=c[1], tp_k2=c[2], tp_k3=c[3], tp_k4=c[4],
                                 tp_k5=c[5], tp_delta=math.pi)
    e = _fd_check(topo, prm, s.pos)
    # psi = phi - 180 degrees: at trans all cosines are 1
    if phi == 170.0:
        psi = math.radians(phi - 180.0)
        assert e == pytest.approx(sum(ck * math.cos(psi) ** n for n, ck in enumerate(c)), rel=1e-9)
    s, topo, prm = _torsion_topo(180.0, 5, tp_k=c[0], tp_k1=c[1], tp_k2=c[2], tp_k3=c[3], tp_k4=c[4],
                                 tp_k5=c[5], tp_delta=math.pi)
    assert _eval(topo, prm, s.pos)[0] == pytest.approx(sum(c), rel=1e-9)


@pytest.mark.parametrize("phi", [20.0, 95.0, -60.0])
def test_improper_type_minus1(phi):
    s, topo, prm = _torsion_topo(phi, -1, tp_k=0.8, tp_delta=math.radians(10.0))
    e = _fd_check(topo, prm, s.pos)
    assert e == pytest.approx(0.8 * (math.radians(phi) - math.radians(10.0)) ** 2, rel=1e-9)


@pytest.mark.parametrize("phi", [60.0, 180.0, -110.0])
def test_tors1_type1(phi):
    s, topo, prm = _torsion_topo(phi, 1, tp_k1=1.3, tp_k2=-0.7, tp_k3=2.1)
    e = _fd_check(topo, prm, s.pos)
    p = math.radians(phi)
    assert e == pytest.approx(1.3 * (1 + math.cos(p)) - 0.7 * (1 - math.cos(2 * p)) + 2.1 * (1 + math.cos(3 * p)),
                              rel=1e-9)


@pytest.mark.parametrize("r", [1.35, 1.53, 1.9, 3.5])
def test_morse_bond_and_fallback(r):
    """D (1 - exp(-rho (r - r0)))^2 with the harmonic K (r - r0)^2 beyond
    r - r0 > r0."""
    s, topo, prm = _molecule([(10.0, 10.0, 10.0), (10.0 + r, 10.0, 10.0)],
                             bonds=[(1, 1, 2, 1.53, 300.0, 400.0, 2.0)])
    topo.bp_morse[:] = 1
    e = _fd_check(topo, prm, s.pos)
    k, d, rho, r0 = topo.bp_k[0], topo.bp_d[0], topo.bp_rho[0], topo.bp_r0[0]
    want = k * (r - r0) ** 2 if r - r0 > r0 else d * (1 - math.exp(-rho * (r - r0))) ** 2
    assert e == pytest.approx(want, rel=1e-9, abs=1e-12)


@pytest.mark.parametrize("r", [1.2, 1.53, 1.8])
def test_harmonic_bond(r):
    s, topo, prm = _molecule([(10.0, 10.0, 10.0), (10.0 + r * 0.6, 10.0 + r * 0.8, 10.0)],
                             bonds=[(0, 1, 2, 1.53, 300.0, 0.0, 0.0)])
    e = _fd_check(topo, prm, s.pos)
    assert e == pytest.approx(topo.bp_k[0] * (r - 1.53) ** 2, rel=1e-9, abs=1e-12)


@pytest.mark.parametrize("theta", [95.0, 112.0, 150.0])
def test_harmonic_angle(theta):
    t = math.radians(theta)
    coords = [(10.0 + 1.5, 10.0, 10.0), (10.0, 10.0, 10.0), (10.0 + 1.4 * math.cos(t), 10.0 + 1.4 * math.sin(t), 10.0)]
    s, topo, prm = _molecule(coords, angles=[(1, 2, 3, 112.0, 50.0)])
    e = _fd_check(topo, prm, s.pos)
    assert e == pytest.approx(topo.ap_k[0] * (t - math.radians(112.0)) ** 2, rel=1e-9, abs=1e-12)


@pytest.mark.parametrize("ipot", [2, 1])
def test_strbnd_forces_are_energy_gradients(ipot):
    """STRBND (forces.f:530-800) alone (charges and LJ off): forces equal
    minus the central difference of PINT for both variants, away from the
    1.3 A switch of the Morse force."""
    import dataclasses
    from mdynamix_amd import systems, topology
    s = systems.water_nacl(8, 0, 12.0, 5, name="fspc", water=systems.fspc_water(), rcut=5.0, rshort=3.0,
                           ewald_a=2.8, ewald_b=4.0, shake=False)
    rng = np.random.default_rng(3)
    pos = s.pos + rng.normal(0.0, 0.05, s.pos.shape)
    topo, prm = topology.from_system(dataclasses.replace(s, pos=pos))
    topo.spc_par[:, 10] = ipot
    topo.atom_charge[:] = 0.0
    topo.atom_charge0[:] = 0.0
    topo.a6[:] = 0.0
    topo.b12[:] = 0.0
    ts, ps = topo.struct(), topology.params_struct(dict(prm, list_mode=0, skin=0.0))
    lists = orc.chcnb(ts, ps, pos)
    _, _, h, e = orc.compute(topo, ts, ps, pos, lists, abi.MDX_FORCES_FAST)
    assert e["pint"] != 0.0
    step = 1e-6
    for atom in (0, 1, 2, 7):
        for c in range(3):
            pp, pm = pos.copy(), pos.copy()
            pp[atom, c] += step
            pm[atom, c] -= step
            ep = orc.compute(topo, ts, ps, pp, lists, abi.MDX_FORCES_FAST)[3]["pint"]
            em = orc.compute(topo, ts, ps, pm, lists, abi.MDX_FORCES_FAST)[3]["pint"]
            fd = -(ep - em) / (2 * step)
            assert abs(h[atom, c] - fd) <= 1e-6 * max(1.0, abs(fd)), (atom, c, h[atom, c], fd)
