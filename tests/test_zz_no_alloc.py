"""mdx_furir_device_noalloc, the variant of mdx_furir_device that allocates
nothing (its k set, structure factors and partials live in the library's
static per-GPU scratch of MDX_NOALLOC_SCRATCH_BYTES): on two calls in a row,
the second with a smaller input, it returns the same bits as
mdx_furir_device with a caller workspace (the same kernels in the same
order); beyond the scratch it returns MDX_ERR_LIMIT before any device work
(checked on the host, no GPU needed)."""
import ctypes as C

import numpy as np
import pytest

from mdynamix_amd import abi, systems, topology
from mdynamix_amd.engine import MdxError, furir_device_noalloc, furir_workspace_bytes, load_library


def test_noalloc_rejects_beyond_scratch():
    s = systems.water_nacl(20000, 2, 80.0, 3, rcut=9.0, rshort=4.5, ewald_a=2.8, ewald_b=40.0, shake=False)
    topo, prm = topology.from_system(s)
    need = furir_workspace_bytes(prm, topo.natoms)
    assert need > abi.MDX_NOALLOC_SCRATCH_BYTES
    dummy = np.zeros(8)
    p = dummy.ctypes.data
    with pytest.raises(MdxError) as ex:
        furir_device_noalloc(prm, topo.natoms, [p] * 4, [p] * 3, p)
    assert ex.value.status == abi.MDX_ERR_LIMIT


@pytest.mark.gpu
def test_noalloc_matches_workspace_call_twice():
    from tests.test_gpu_furir64 import run_device
    big = systems.water_nacl(1120, 20, 33.0, 19, rcut=10.0, rshort=5.0, ewald_a=3.0, ewald_b=10.0, shake=False)
    small = systems.water_nacl(200, 3, 19.0, 23, rcut=8.0, rshort=4.0, ewald_a=2.8, ewald_b=9.0, shake=False)
    for s in (big, small):
        topo, prm = topology.from_system(s)
        assert furir_workspace_bytes(prm, topo.natoms) <= abi.MDX_NOALLOC_SCRATCH_BYTES
        print("case noalloc", topo.natoms, flush=True)
        f_ws, e_ws = run_device(prm, s.pos, topo.atom_charge)

        def noalloc(prm_, n, xyzq, forces, e4, ws, wsb):
            furir_device_noalloc(prm_, n, xyzq, forces, e4)
        f_na, e_na = run_device(prm, s.pos, topo.atom_charge, call=noalloc)
        assert np.array_equal(f_na, f_ws), np.abs(f_na - f_ws).max()
        assert np.array_equal(e_na, e_ws)
