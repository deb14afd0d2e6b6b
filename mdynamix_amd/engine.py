"""Python host side of the MI355X force engine (ctypes over libmdx.so).

`ForceEngine` mirrors the reference force-routine interface (forces.f):

    chcnb()        CHCNB: rebuild the neighbour lists
    allforce()     ALLFORCE: F = G + H with every energy/virial scalar
    fforces()      FFORCES: fast forces H (short pairs + bonded)
    slforces()     SLFORCES: slow forces G (k-space, exclusions, long pairs)

Errors of the C ABI raise `MdxError` carrying the status name and the
engine's message.  There is no CPU fallback: without the HIP library or a
device the constructor raises.
"""
from __future__ import annotations

import ctypes as C
import os
from typing import Dict, Optional, Tuple

import numpy as np

from . import abi
from .abi import MdxEnergies, as_ptr

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libmdx.so")

_lib = None


class MdxError(RuntimeError):
    def __init__(self, status: int, msg: str):
        super().__init__("%s: %s" % (abi.STATUS_NAMES.get(status, str(status)), msg))
        self.status = status


def load_library() -> C.CDLL:
    """Load the in-tree HIP library; raise if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise FileNotFoundError("libmdx.so not built: run python -m mdynamix_amd.build")
    L = C.CDLL(LIB_PATH)
    pd, pv = C.POINTER(C.c_double), C.c_void_p
    pi32 = C.POINTER(C.c_int32)
    L.mdx_version.restype = C.c_char_p
    L.mdx_last_error.restype = C.c_char_p
    L.mdx_last_error.argtypes = [pv]
    L.mdx_device_info.argtypes = [pi32, pi32, C.c_char_p, C.c_int32]
    L.mdx_create.argtypes = [pv, pv, C.POINTER(pv)]
    L.mdx_destroy.argtypes = [pv]
    L.mdx_set_positions.argtypes = [pv, pd, pd, pd]
    L.mdx_set_positions_device.argtypes = [pv, pv, pv, pv]
    L.mdx_set_box.argtypes = [pv, pd]
    L.mdx_set_molecule_com.argtypes = [pv, pd, pd, pd]
    L.mdx_build_list.argtypes = [pv]
    L.mdx_list_counts.argtypes = [pv, C.POINTER(C.c_int64), C.POINTER(C.c_int64)]
    L.mdx_list_entries.argtypes = [pv, C.POINTER(C.c_int64), C.POINTER(C.c_int64)]
    L.mdx_get_list.argtypes = [pv, pi32, pi32, C.c_int64, pi32, pi32, C.c_int64]
    L.mdx_compute.argtypes = [pv, C.c_int32, C.c_int32] + [pd] * 9 + [pv]
    L.mdx_compute_device.argtypes = [pv, C.c_int32, C.c_int32, pv, pv, pv, pv]
    L.mdx_stage_times.argtypes = [pv, pd, C.c_int32]
    L.mdx_get_forces.argtypes = [pv] + [pd] * 9
    L.mdx_get_terms.argtypes = [pv] + [pd] * 6
    L.mdx_ee_energies.argtypes = [pv, C.POINTER(abi.MdxEeParams), C.c_int32, pd, pd, pd]
    L.mdx_set_timing.argtypes = [pv, C.c_int32]
    L.mdx_memory_bytes.argtypes = [pv, C.POINTER(C.c_int64)]
    L.mdx_batch_create.argtypes = [C.POINTER(pv), C.c_int32, C.POINTER(pv)]
    L.mdx_batch_set_ring.argtypes = [pv, C.c_double]
    L.mdx_batch_use_graph.argtypes = [pv, C.c_int32]
    L.mdx_batch_compute.argtypes = [pv, C.c_int32, C.c_int32, pv]
    L.mdx_batch_last_error.argtypes = [pv]
    L.mdx_batch_last_error.restype = C.c_char_p
    L.mdx_batch_destroy.argtypes = [pv]
    L.mdx_autotune.argtypes = [pv, pi32, pi32]
    L.mdx_cluster_grid.argtypes = [pv, C.POINTER(C.c_double), pi32, pi32]
    L.mdx_set_cluster_grid.argtypes = [pv, C.c_double]
    L.mdx_set_pair_waves.argtypes = [pv, C.c_int32]
    L.mdx_pair_waves.argtypes = [pv, pi32]
    L.mdx_set_kspace_shape.argtypes = [pv, C.c_int32, C.c_int32]
    L.mdx_get_kvectors.argtypes = [pv, pi32, pi32, pi32, C.c_int64, C.POINTER(C.c_int64)]
    L.mdx_enumerate_kvectors.argtypes = [pv, pi32, pi32, pi32, C.c_int64, C.POINTER(C.c_int64)]
    L.mdx_md_init.argtypes = [pv, pv, pd, pd, pd]
    L.mdx_md_run.argtypes = [pv, C.c_int32, pv]
    L.mdx_md_get_state.argtypes = [pv] + [pd] * 6
    L.mdx_batch_md_run.argtypes = [pv, C.c_int32, pv]
    L.mdx_md_shake_failures.argtypes = [pv, C.POINTER(C.c_int64)]
    L.mdx_furir_workspace_bytes.argtypes = [pv, C.c_int32, C.POINTER(C.c_int64)]
    # p, natoms, x, y, z, q, fx, fy, fz, energy4, [workspace, bytes,] stream
    L.mdx_furir_device.argtypes = [pv, C.c_int32] + [pv] * 8 + [pv, C.c_int64, pv]
    L.mdx_furir_device_noalloc.argtypes = [pv, C.c_int32] + [pv] * 8 + [pv]
    L.mdx_batch_pimd_init.argtypes = [pv, C.POINTER(abi.MdxPimdParams)]
    L.mdx_batch_pimd_chains.argtypes = [pv, pd]
    for name in ("mdx_create", "mdx_destroy", "mdx_set_positions", "mdx_set_positions_device", "mdx_set_box",
                 "mdx_set_molecule_com", "mdx_build_list", "mdx_list_counts", "mdx_list_entries", "mdx_get_list", "mdx_compute", "mdx_get_terms", "mdx_ee_energies",
                 "mdx_compute_device", "mdx_stage_times", "mdx_device_info", "mdx_get_forces", "mdx_set_timing",
                 "mdx_memory_bytes", "mdx_batch_create", "mdx_batch_set_ring", "mdx_batch_use_graph",
                 "mdx_batch_compute", "mdx_batch_destroy", "mdx_md_init", "mdx_md_run", "mdx_md_get_state",
                 "mdx_batch_md_run", "mdx_autotune", "mdx_cluster_grid", "mdx_set_cluster_grid",
                 "mdx_set_pair_waves", "mdx_pair_waves", "mdx_set_kspace_shape", "mdx_get_kvectors",
                 "mdx_enumerate_kvectors", "mdx_md_shake_failures", "mdx_furir_workspace_bytes",
                 "mdx_furir_device", "mdx_furir_device_noalloc", "mdx_batch_pimd_init", "mdx_batch_pimd_chains"):
        getattr(L, name).restype = C.c_int
    _lib = L
    return L


STAGES = ("list", "prep", "pairs", "special", "recip_prep", "recip_sk", "recip_sk_reduce", "recip_force",
          "bonded", "reduce")


class ForceEngine:
    """One system on the GPU.  `topo` is a TopologyArrays, `params` a dict
    of mdx_params fields (see topology.from_system / from_dump)."""

    def __init__(self, topo: abi.TopologyArrays, params: Dict):
        self.lib = load_library()
        self.topo = topo
        self.params = dict(params)
        self._ts = topo.struct()
        self._ps = abi.make_params(**params)
        h = C.c_void_p()
        st = self.lib.mdx_create(C.byref(self._ts), C.byref(self._ps), C.byref(h))
        self._h = h
        if st != 0:
            msg = self.lib.mdx_last_error(h).decode() if h.value else ""
            if h.value:
                self.lib.mdx_destroy(h)
                self._h = C.c_void_p()
            raise MdxError(st, msg)
        self.n = topo.natoms

    def _check(self, st: int) -> None:
        if st != 0:
            raise MdxError(st, self.lib.mdx_last_error(self._h).decode())

    def close(self) -> None:
        if self._h and self._h.value:
            self.lib.mdx_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # -- state ---------------------------------------------------------
    def set_positions(self, pos: np.ndarray) -> None:
        pos = np.asarray(pos, dtype=np.float64)
        self._xyz = [np.ascontiguousarray(pos[:, k]) for k in range(3)]
        self._check(self.lib.mdx_set_positions(self._h, *[as_ptr(a) for a in self._xyz]))

    def set_positions_device(self, x_ptr: int, y_ptr: int, z_ptr: int) -> None:
        self._check(self.lib.mdx_set_positions_device(self._h, C.c_void_p(x_ptr), C.c_void_p(y_ptr),
                                                      C.c_void_p(z_ptr)))

    def set_box(self, box) -> None:
        b = np.ascontiguousarray(box, dtype=np.float64)
        self._check(self.lib.mdx_set_box(self._h, as_ptr(b)))

    def set_molecule_com(self, com: Optional[np.ndarray]) -> None:
        if com is None:
            self._check(self.lib.mdx_set_molecule_com(self._h, None, None, None))
            return
        self._com = [np.ascontiguousarray(com[:, k], dtype=np.float64) for k in range(3)]
        self._check(self.lib.mdx_set_molecule_com(self._h, *[as_ptr(a) for a in self._com]))

    # -- reference-named entry points ------------------------------------
    def chcnb(self) -> None:
        self._check(self.lib.mdx_build_list(self._h))

    build_list = chcnb

    def list_counts(self) -> Tuple[int, int]:
        s, l = C.c_int64(), C.c_int64()
        self._check(self.lib.mdx_list_counts(self._h, C.byref(s), C.byref(l)))
        return s.value, l.value

    def list_entries(self) -> Tuple[int, int]:
        """Cluster-pair entries of the last build: (all, with a short-class pair)."""
        e, f = C.c_int64(), C.c_int64()
        self._check(self.lib.mdx_list_entries(self._h, C.byref(e), C.byref(f)))
        return e.value, f.value

    def get_list(self):
        """NBS1, NBS2, NBL1, NBL2 in the reference layout and order."""
        ns, nl = self.list_counts()
        s1 = np.zeros(max(ns, 1), np.int32); s2 = np.zeros(max(ns, 1), np.int32)
        l1 = np.zeros(max(nl, 1), np.int32); l2 = np.zeros(max(nl, 1), np.int32)
        pi = C.c_int32
        self._check(self.lib.mdx_get_list(self._h, as_ptr(s1, pi), as_ptr(s2, pi), ns, as_ptr(l1, pi),
                                          as_ptr(l2, pi), nl))
        return s1[:ns], s2[:ns], l1[:nl], l2[:nl]

    def compute(self, group: int = abi.MDX_FORCES_ALL, want: int = 0, split: bool = True):
        """(F, G, H, energies); G and H are None unless split."""
        n = self.n
        f = np.zeros((3, n))
        g = np.zeros((3, n)) if split else None
        h = np.zeros((3, n)) if split else None
        e = MdxEnergies()
        ptrs = [as_ptr(f[k]) for k in range(3)]
        ptrs += [as_ptr(g[k]) for k in range(3)] if split else [None] * 3
        ptrs += [as_ptr(h[k]) for k in range(3)] if split else [None] * 3
        self._check(self.lib.mdx_compute(self._h, group, want, *ptrs, C.byref(e)))
        ed = e.as_dict(self.topo.nmoltypes, self.topo.ntypepairs)
        return f.T.copy(), (g.T.copy() if split else None), (h.T.copy() if split else None), ed

    def allforce(self, **kw):
        return self.compute(abi.MDX_FORCES_ALL, **kw)

    def fforces(self, **kw):
        return self.compute(abi.MDX_FORCES_FAST, **kw)

    def slforces(self, **kw):
        return self.compute(abi.MDX_FORCES_SLOW, **kw)

    def compute_device(self, fx_ptr: int, fy_ptr: int, fz_ptr: int, group: int = abi.MDX_FORCES_ALL,
                       energies: bool = False):
        e = MdxEnergies() if energies else None
        ptr = lambda p: C.c_void_p(p) if p else None
        self._check(self.lib.mdx_compute_device(self._h, group, 0, ptr(fx_ptr), ptr(fy_ptr), ptr(fz_ptr),
                                                C.byref(e) if e is not None else None))
        return e.as_dict(self.topo.nmoltypes, self.topo.ntypepairs) if e is not None else None

    def stage_times(self) -> Dict[str, float]:
        ms = np.zeros(len(STAGES))
        self._check(self.lib.mdx_stage_times(self._h, as_ptr(ms), len(STAGES)))
        return dict(zip(STAGES, ms.tolist()))

    def set_timing(self, on: bool) -> None:
        """Per-stage event timing (eager launches) or, when off, replay of
        the force call from a captured hipGraph."""
        self._check(self.lib.mdx_set_timing(self._h, 1 if on else 0))

    def autotune(self) -> Dict[str, float]:
        """Pick the k-space launch shape, the cluster grid and the pair
        kernel's wavefronts per SIMD of this size class by timing (see
        mdx_autotune; the list is rebuilt at the current coordinates);
        returns the chosen shape, column width and wavefronts."""
        c, sl = C.c_int32(), C.c_int32()
        self._check(self.lib.mdx_autotune(self._h, C.byref(c), C.byref(sl)))
        return {"chunk": c.value, "nslice": sl.value, "column_width": self.cluster_grid()["column_width"],
                "pair_waves": self.pair_waves()}

    def set_pair_waves(self, waves: int) -> None:
        """A fixed occupancy of the plain pair kernel, 5 or 6 wavefronts per
        SIMD (mdx_set_pair_waves)."""
        self._check(self.lib.mdx_set_pair_waves(self._h, int(waves)))

    def pair_waves(self) -> int:
        """Wavefronts per SIMD of the plain pair kernel (mdx_pair_waves)."""
        w = C.c_int32()
        self._check(self.lib.mdx_pair_waves(self._h, C.byref(w)))
        return w.value

    def set_cluster_grid(self, scale: float) -> None:
        """A fixed cluster grid (mdx_set_cluster_grid)."""
        self._check(self.lib.mdx_set_cluster_grid(self._h, float(scale)))

    def cluster_grid(self) -> Dict[str, float]:
        """The cluster grid in use (mdx_cluster_grid)."""
        w, nx, ny = C.c_double(), C.c_int32(), C.c_int32()
        self._check(self.lib.mdx_cluster_grid(self._h, C.byref(w), C.byref(nx), C.byref(ny)))
        return {"column_width": w.value, "ncx": nx.value, "ncy": ny.value}

    def kvectors(self) -> np.ndarray:
        """The engine's k set as an (nkv, 3) array of integer triples n in
        INIFUR order (RKX = 2 pi n_x, forces.f:1375-1427)."""
        return _kvectors(lambda *a: self.lib.mdx_get_kvectors(self._h, *a), self._check)

    def set_kspace_shape(self, chunk: int, nslice: int) -> None:
        self._check(self.lib.mdx_set_kspace_shape(self._h, int(chunk), int(nslice)))

    def terms(self) -> Dict[str, np.ndarray]:
        """Per-template bonded sums of the last call made with
        MDX_WANT_TERMS: br/be (bond templates), ar/ae (angle), tr/te
        (dihedral), as GETBND/STRBND/GETANG/GETTRS leave BR..TE."""
        t = self.topo
        nb, na, nt = len(t.bp_r0), len(t.ap_k), len(t.tp_kind)
        out = {k: np.zeros(max(m, 1)) for k, m in (("br", nb), ("be", nb), ("ar", na), ("ae", na),
                                                    ("tr", nt), ("te", nt))}
        self._check(self.lib.mdx_get_terms(self._h, *[as_ptr(out[k]) for k in ("br", "be", "ar", "ae", "tr", "te")]))
        return {k: v[:m] for (k, v), m in zip(out.items(), (nb, nb, na, na, nt, nt))}

    def ee_energies(self, lambdas, pw1: float, pw2: float, bee: float):
        """(e_pair, e_recip) at each lambda: the sums DIFEN and DIFFUR form
        for an expanded-ensemble move (mdee.f:265-534) as functions of the
        coupling parameter, at the current coordinates and list."""
        lam = np.ascontiguousarray(lambdas, dtype=np.float64)
        ep, er = np.zeros(len(lam)), np.zeros(len(lam))
        p = abi.MdxEeParams(pw1, pw2, bee)
        self._check(self.lib.mdx_ee_energies(self._h, C.byref(p), len(lam), as_ptr(lam), as_ptr(ep), as_ptr(er)))
        return ep, er

    def get_forces(self):
        """(F, G, H) of the last force call, e.g. after a batched call."""
        n = self.n
        out = np.zeros((9, n))
        self._check(self.lib.mdx_get_forces(self._h, *[as_ptr(out[k]) for k in range(9)]))
        return out[0:3].T.copy(), out[3:6].T.copy(), out[6:9].T.copy()

    # -- device-resident MD segments ------------------------------------
    def md_init(self, md: Dict, velocities: np.ndarray) -> None:
        """Start an MD segment (leap-frog + constraints + thermostat on the
        device) from the current positions; `md` as topology.md_params."""
        p, keep = abi.make_md_params(md)
        v = [np.ascontiguousarray(velocities[:, k], dtype=np.float64) for k in range(3)]
        self._check(self.lib.mdx_md_init(self._h, C.byref(p), *[as_ptr(a) for a in v]))

    def md_run(self, nsteps: int, trace: bool = True):
        """nsteps steps; returns the per-step (Epot, Ekin, T, SC, pairs) array."""
        tr = np.zeros((max(nsteps, 1), 5)) if trace else None
        self._check(self.lib.mdx_md_run(self._h, nsteps, as_ptr(tr) if trace else None))
        return tr[:nsteps] if trace else None

    def md_state(self):
        """(positions, velocities) as (n, 3) arrays."""
        out = np.zeros((6, self.n))
        self._check(self.lib.mdx_md_get_state(self._h, *[as_ptr(out[k]) for k in range(6)]))
        return out[0:3].T.copy(), out[3:6].T.copy()

    def md_shake_failures(self) -> int:
        """Constraint failures counted by the device SHAKE since md_init."""
        c = C.c_int64()
        self._check(self.lib.mdx_md_shake_failures(self._h, C.byref(c)))
        return c.value

    def memory_bytes(self) -> int:
        b = C.c_int64()
        self._check(self.lib.mdx_memory_bytes(self._h, C.byref(b)))
        return b.value


class ForceBatch:
    """K engines evaluated by one set of fused launches per force call
    (path-integral beads with `ring`, lambda states or independent jobs).
    Each engine keeps its own positions and lists."""

    def __init__(self, engines, ring_beadfac: Optional[float] = None, graph: bool = True):
        self.lib = load_library()
        self.engines = list(engines)
        arr = (C.c_void_p * len(self.engines))(*[e._h.value for e in self.engines])
        b = C.c_void_p()
        st = self.lib.mdx_batch_create(arr, len(self.engines), C.byref(b))
        self._b = b
        if st != 0:
            if b.value:
                self.lib.mdx_batch_destroy(b)
            self._b = C.c_void_p()
            raise MdxError(st, "batch creation failed")
        if ring_beadfac is not None:
            self._check(self.lib.mdx_batch_set_ring(self._b, ring_beadfac))
        self._check(self.lib.mdx_batch_use_graph(self._b, 1 if graph else 0))

    def _check(self, st: int) -> None:
        if st != 0:
            raise MdxError(st, self.lib.mdx_batch_last_error(self._b).decode())

    def compute(self, group: int = abi.MDX_FORCES_ALL, energies: bool = True):
        k = len(self.engines)
        e = (MdxEnergies * k)()
        self._check(self.lib.mdx_batch_compute(self._b, group, 0, C.cast(e, C.c_void_p) if energies else None))
        if not energies:
            return None
        return [e[s].as_dict(self.engines[s].topo.nmoltypes, self.engines[s].topo.ntypepairs) for s in range(k)]

    def md_run(self, nsteps: int, trace: bool = True):
        """nsteps MD steps of every engine (each set up by md_init); returns
        a (K, nsteps, 5) trace."""
        k = len(self.engines)
        tr = np.zeros((k, max(nsteps, 1), 5)) if trace else None
        self._check(self.lib.mdx_batch_md_run(self._b, nsteps, as_ptr(tr) if trace else None))
        return tr[:, :nsteps] if trace else None

    def pimd_init(self, temperature: float, tau_bead: float, nchain: int = 1) -> None:
        """Path-integral MD of a ring batch (every engine md_init'ed with
        integrator MDX_MD_PIMD): bead thermostat chains (PI_thermo T BTAU
        JNH; tau_bead 0: none) and the ring forces at the start."""
        p = abi.MdxPimdParams(float(temperature), float(tau_bead), int(nchain))
        self._nchain = max(1, int(nchain))
        self._check(self.lib.mdx_batch_pimd_init(self._b, C.byref(p)))

    def pimd_chains(self) -> np.ndarray:
        """Bead thermostat chain variables, (nchain, n)."""
        xi = np.zeros((self._nchain, self.engines[0].n))
        self._check(self.lib.mdx_batch_pimd_chains(self._b, as_ptr(xi)))
        return xi

    def close(self) -> None:
        if self._b and self._b.value:
            self.lib.mdx_batch_destroy(self._b)
            self._b = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def _kvectors(call, check) -> np.ndarray:
    n = C.c_int64()
    check(call(None, None, None, 0, C.byref(n)))
    out = np.zeros((3, max(n.value, 1)), np.int32)
    pi = C.c_int32
    check(call(as_ptr(out[0], pi), as_ptr(out[1], pi), as_ptr(out[2], pi), out.shape[1], C.byref(n)))
    return out[:, :n.value].T.copy()


def enumerate_kvectors(params: Dict) -> np.ndarray:
    """INIFUR's k set for `params` (a dict of mdx_params fields) from the
    engine's host-side enumeration; needs the library but no device."""
    L = load_library()
    ps = abi.make_params(**params)

    def check(st):
        if st != 0:
            raise MdxError(st, "mdx_enumerate_kvectors")
    return _kvectors(lambda *a: L.mdx_enumerate_kvectors(C.byref(ps), *a), check)


def furir_workspace_bytes(params: Dict, natoms: int) -> int:
    """Device workspace mdx_furir_device needs for natoms atoms."""
    L = load_library()
    ps = abi.make_params(**params)
    b = C.c_int64()
    st = L.mdx_furir_workspace_bytes(C.byref(ps), natoms, C.byref(b))
    if st != 0:
        raise MdxError(st, "mdx_furir_workspace_bytes")
    return b.value


def furir_device(params: Dict, natoms: int, xyzq, forces, energy4: int, workspace: int, workspace_bytes: int,
                 stream: int = 0) -> None:
    """FURIR in FP64 on device arrays (addresses as ints): xyzq = (x, y, z,
    q), forces = (fx, fy, fz), energy4 = QPE and the three virial
    projections; queued on `stream` (0: default), not waited for."""
    L = load_library()
    ps = abi.make_params(**params)
    v = lambda p: C.c_void_p(p) if p else None
    st = L.mdx_furir_device(C.byref(ps), natoms, *[v(p) for p in xyzq], *[v(p) for p in forces], v(energy4),
                            v(workspace), workspace_bytes, v(stream))
    if st != 0:
        raise MdxError(st, "mdx_furir_device")


def furir_device_noalloc(params: Dict, natoms: int, xyzq, forces, energy4: int, stream: int = 0) -> None:
    """mdx_furir_device on the library's static per-GPU scratch (no device
    allocation); raises MdxError(MDX_ERR_LIMIT) beyond its size."""
    L = load_library()
    ps = abi.make_params(**params)
    v = lambda p: C.c_void_p(p) if p else None
    st = L.mdx_furir_device_noalloc(C.byref(ps), natoms, *[v(p) for p in xyzq], *[v(p) for p in forces],
                                    v(energy4), v(stream))
    if st != 0:
        raise MdxError(st, "mdx_furir_device_noalloc")


def device_info() -> Dict:
    L = load_library()
    cu, clk = C.c_int32(), C.c_int32()
    name = C.create_string_buffer(256)
    st = L.mdx_device_info(C.byref(cu), C.byref(clk), name, 256)
    if st != 0:
        raise MdxError(st, "no HIP device")
    return {"compute_units": cu.value, "clock_mhz": clk.value, "name": name.value.decode()}
