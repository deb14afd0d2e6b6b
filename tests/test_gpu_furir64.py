"""FURIR in FP64 on device arrays (mdx_furir_device) against the oracle's
FURIR (forces.f:1217-1339, pinned on the reference fixtures): forces and
the energy QPE with its virial projections to 1e-10 relative (both FP64;
the device takes sines and cosines from sincospi of phase turns reduced to
[-1/2, 1/2], the oracle from libm of the full phase, up to ~40 rad here:
measured 1.2e-11 in force and 3e-12 in energy), on the 508-water box
and a 3,400-atom water + NaCl box; the same call twice gives the same bits.
"""
import ctypes as C

import numpy as np
import pytest

from mdynamix_amd import abi, systems, topology
from mdynamix_amd.abi import MdxEnergies, as_ptr
from mdynamix_amd.engine import furir_device, furir_workspace_bytes
from oracle import orc
from tests.helpers import DeviceArrays

pytestmark = pytest.mark.gpu


def _oracle_furir(topo, prm, pos):
    ts, ps = topo.struct(), topology.params_struct(dict(prm, list_mode=0, skin=0.0))
    rkx, rky, rkz, _, _ = orc.inifur(ts, ps, topo.nmoltypes)
    n = len(pos)
    f = [np.zeros(n) for _ in range(3)]
    e = MdxEnergies()
    x, y, z = [np.ascontiguousarray(pos[:, k]) for k in range(3)]
    orc.lib().orc_furir_task(C.byref(ts), C.byref(ps), as_ptr(x), as_ptr(y), as_ptr(z), as_ptr(rkx), as_ptr(rky),
                             as_ptr(rkz), len(rkx), 0, 1, *[as_ptr(a) for a in f], C.byref(e))
    return np.stack(f, 1), np.array([e.pels1, e.virx, e.viry, e.virz])


def run_device(prm, pos, q, call=None):
    """One mdx_furir_device call (or `call` with its arguments) on device
    copies of pos and q; returns forces (n, 3) and the four energy terms."""
    dev = DeviceArrays()
    try:
        n = len(pos)
        xyzq = [dev.upload(np.ascontiguousarray(a, np.float64)) for a in (pos[:, 0], pos[:, 1], pos[:, 2], q)]
        f = [dev.alloc(8 * n) for _ in range(3)]
        e4 = dev.alloc(32)
        wsb = furir_workspace_bytes(prm, n)
        ws = dev.alloc(wsb)
        (call or furir_device)(prm, n, xyzq, f, e4, ws, wsb)
        forces = np.stack([dev.download(p, n) for p in f], 1)
        return forces, dev.download(e4, 4)
    finally:
        dev.close()


@pytest.mark.parametrize("which", ["water_508", "wnacl_3400"])
def test_furir64_matches_oracle(which):
    if which == "water_508":
        s = systems.cfg_small()
    else:
        s = systems.water_nacl(1120, 20, 33.0, 19, rcut=10.0, rshort=5.0, ewald_a=3.0, ewald_b=10.0, shake=False)
    topo, prm = topology.from_system(s)
    print("case furir64", which, topo.natoms, flush=True)
    f, e4 = run_device(prm, s.pos, topo.atom_charge)
    fo, eo = _oracle_furir(topo, prm, s.pos)
    err = np.sqrt(((f - fo) ** 2).sum() / (fo ** 2).sum())
    print("furir64 %s: force rms rel %.2e, energy rel %.2e" % (which, err, abs(e4[0] - eo[0]) / abs(eo[0])), flush=True)
    assert err <= 1e-10
    assert np.all(np.abs(e4 - eo) <= 1e-10 * np.abs(eo).max())
    f2, e42 = run_device(prm, s.pos, topo.atom_charge)
    assert np.array_equal(f, f2) and np.array_equal(e4, e42)
