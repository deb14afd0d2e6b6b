"""ctypes mirror of include/mdx.h.

`TopologyArrays` holds the numpy arrays behind one `mdx_topology` and
builds the struct that points at them; `make_params` fills `mdx_params`.
Both the HIP engine (mdynamix_amd/libmdx.so) and the CPU oracle take these
structs, so every test feeds both sides from the same bytes.
"""
from __future__ import annotations

import ctypes as C
import dataclasses
from typing import Dict, Optional

import numpy as np

MDX_OK = 0
MDX_ERR_ARG = 1
MDX_ERR_LIST_OVERFLOW = 2
MDX_ERR_NKV = 3
MDX_ERR_HIP = 4
MDX_ERR_NOMEM = 5
MDX_ERR_STATE = 6
MDX_ERR_LIMIT = 7
STATUS_NAMES = {0: "MDX_OK", 1: "MDX_ERR_ARG", 2: "MDX_ERR_LIST_OVERFLOW", 3: "MDX_ERR_NKV",
                4: "MDX_ERR_HIP", 5: "MDX_ERR_NOMEM", 6: "MDX_ERR_STATE", 7: "MDX_ERR_LIMIT"}
MDX_NOALLOC_SCRATCH_BYTES = 64 << 20

MDX_LIST_REFERENCE = 0
MDX_LIST_VERLET = 1
MDX_PRECISION_MIXED = 0
MDX_PRECISION_FP64 = 1
MDX_FORCES_ALL = 0
MDX_MD_LEAPFROG = 0                  # SULTAN: leap-frog with SHEJK constraints
MDX_MD_RESPA = 1                     # DOUBLE: multiple time step (NFREQ short steps)
MDX_MD_PIMD = 2                      # PIMD: path-integral beads of a ring batch
MDX_FORCES_FAST = 1
MDX_FORCES_SLOW = 2
MDX_WANT_DETAIL = 1
MDX_WANT_TERMS = 2
MDX_MAX_MOLTYPES = 16
MDX_MAX_TYPEPAIRS = MDX_MAX_MOLTYPES * (MDX_MAX_MOLTYPES + 1) // 2 + 1

_pi = C.POINTER(C.c_int32)
_pd = C.POINTER(C.c_double)
_pb = C.POINTER(C.c_int8)


class MdxTopology(C.Structure):
    _fields_ = [
        ("natoms", C.c_int32), ("nmols", C.c_int32), ("nmoltypes", C.c_int32),
        ("nljtypes", C.c_int32), ("ntypepairs", C.c_int32),
        ("atom_moltype", _pi), ("atom_mol", _pi), ("atom_ljtype", _pi),
        ("atom_charge", _pd), ("atom_charge0", _pd), ("atom_mass", _pd),
        ("type_ee", _pi), ("type_move", _pi), ("type_c14el", _pd), ("type_c14lj", _pd),
        ("type_pair", _pi),
        ("a6", _pd), ("b12", _pd), ("a6_0", _pd), ("b12_0", _pd), ("a6_14", _pd), ("b12_14", _pd),
        ("excl_start", _pi), ("excl_atom", _pi), ("excl_is14", _pb),
        ("nbonds", C.c_int32), ("bond_i", _pi), ("bond_j", _pi), ("bond_param", _pi),
        ("nbondparams", C.c_int32), ("bp_morse", _pi), ("bp_r0", _pd), ("bp_k", _pd),
        ("bp_d", _pd), ("bp_rho", _pd),
        ("nangles", C.c_int32), ("angle_i", _pi), ("angle_j", _pi), ("angle_k", _pi),
        ("angle_param", _pi), ("nangleparams", C.c_int32), ("ap_theta0", _pd), ("ap_k", _pd),
        ("ntors", C.c_int32), ("tors_i", _pi), ("tors_j", _pi), ("tors_k", _pi), ("tors_l", _pi),
        ("tors_param", _pi), ("ntorsparams", C.c_int32), ("tp_kind", _pi), ("tp_mult", _pi),
        ("tp_k", _pd), ("tp_delta", _pd), ("tp_k1", _pd), ("tp_k2", _pd), ("tp_k3", _pd),
        ("tp_k4", _pd), ("tp_k5", _pd),
        ("nspc", C.c_int32), ("spc_atom", _pi), ("spc_par", _pd), ("spc_bond", _pi),
    ]


class MdxParams(C.Structure):
    _fields_ = [
        ("box", C.c_double * 3), ("rcut", C.c_double), ("rshort", C.c_double),
        ("skin", C.c_double), ("alpha", C.c_double), ("fexp", C.c_double),
        ("coulomb_factor", C.c_double), ("list_mode", C.c_int32), ("ee_enabled", C.c_int32),
        ("ee_raex", C.c_double), ("ee_self_scale", C.c_double), ("precision", C.c_int32),
    ]


class MdxEeParams(C.Structure):
    _fields_ = [("pw1", C.c_double), ("pw2", C.c_double), ("bee", C.c_double)]


class MdxMdParams(C.Structure):
    _fields_ = [
        ("dt", C.c_double), ("accel_factor", C.c_double), ("ekin_factor", C.c_double), ("boltz", C.c_double),
        ("temperature", C.c_double), ("tau_t", C.c_double), ("sc0", C.c_double), ("ndof", C.c_double),
        ("shake_tol", C.c_double), ("shake_bdtol", C.c_double), ("shake_maxit", C.c_int32),
        ("rebuild_every", C.c_int32), ("step0", C.c_int32), ("ncons", C.c_int32),
        ("cons_i", C.POINTER(C.c_int32)), ("cons_j", C.POINTER(C.c_int32)), ("cons_len", C.POINTER(C.c_double)),
        ("integrator", C.c_int32), ("nfreq", C.c_int32),
    ]


class MdxPimdParams(C.Structure):
    _fields_ = [("temperature", C.c_double), ("tau_bead", C.c_double), ("nchain", C.c_int32)]


def make_md_params(d: Dict):
    """MdxMdParams from a dict (topology.md_params); returns (struct, arrays
    kept alive by the caller)."""
    p = MdxMdParams()
    keep = []
    for name, ctype in MdxMdParams._fields_:
        if name in ("cons_i", "cons_j", "cons_len"):
            continue
        if name == "ncons":
            continue
        if name in ("integrator", "nfreq"):
            setattr(p, name, int(d.get(name, MDX_MD_LEAPFROG if name == "integrator" else 1)))
            continue
        setattr(p, name, d[name])
    ci = np.ascontiguousarray(d["cons_i"], np.int32)
    cj = np.ascontiguousarray(d["cons_j"], np.int32)
    ln = np.ascontiguousarray(d["cons_len"], np.float64)
    p.ncons = len(ci)
    p.cons_i = ci.ctypes.data_as(C.POINTER(C.c_int32))
    p.cons_j = cj.ctypes.data_as(C.POINTER(C.c_int32))
    p.cons_len = ln.ctypes.data_as(C.POINTER(C.c_double))
    keep += [ci, cj, ln]
    return p, keep


class MdxEnergies(C.Structure):
    _fields_ = [
        ("pe1", C.c_double), ("pe2", C.c_double), ("pels1", C.c_double), ("pels2", C.c_double),
        ("spe", C.c_double), ("pint", C.c_double), ("vir1", C.c_double), ("vir2", C.c_double),
        ("virb", C.c_double), ("virx", C.c_double), ("viry", C.c_double), ("virz", C.c_double),
        ("virfx", C.c_double), ("virfy", C.c_double), ("virfz", C.c_double),
        ("wirs", C.c_double), ("wirss", C.c_double),
        ("npairs_short", C.c_int64), ("npairs_long", C.c_int64),
        ("pote1", C.c_double * MDX_MAX_TYPEPAIRS), ("potl1", C.c_double * MDX_MAX_TYPEPAIRS),
        ("pes141", C.c_double * MDX_MAX_MOLTYPES), ("psr141", C.c_double * MDX_MAX_MOLTYPES),
        ("spee", C.c_double * MDX_MAX_MOLTYPES),
        ("qek", C.c_double), ("virpi", C.c_double),
    ]

    SCALARS = ("pe1", "pe2", "pels1", "pels2", "spe", "pint", "vir1", "vir2", "virb", "virx",
               "viry", "virz", "virfx", "virfy", "virfz", "wirs", "wirss")

    def as_dict(self, nmoltypes: int = MDX_MAX_MOLTYPES, ntypepairs: int = MDX_MAX_TYPEPAIRS) -> Dict:
        # read through the struct's buffer (one view per field group; the
        # scalars are the leading doubles, in SCALARS order)
        c = type(self)
        sc = np.frombuffer(self, np.float64, len(self.SCALARS), 0).tolist()
        d = dict(zip(self.SCALARS, sc))
        d["qek"], d["virpi"] = np.frombuffer(self, np.float64, 2, c.qek.offset).tolist()
        d["npairs_short"], d["npairs_long"] = np.frombuffer(self, np.int64, 2, c.npairs_short.offset).tolist()
        d["pote1"] = np.frombuffer(self, np.float64, ntypepairs, c.pote1.offset).copy()
        d["potl1"] = np.frombuffer(self, np.float64, ntypepairs, c.potl1.offset).copy()
        d["pes141"] = np.frombuffer(self, np.float64, nmoltypes, c.pes141.offset).copy()
        d["psr141"] = np.frombuffer(self, np.float64, nmoltypes, c.psr141.offset).copy()
        d["spee"] = np.frombuffer(self, np.float64, nmoltypes, c.spee.offset).copy()
        return d


def _i32(a) -> np.ndarray:
    return np.ascontiguousarray(np.asarray(a, dtype=np.int32).ravel())


def _f64(a) -> np.ndarray:
    return np.ascontiguousarray(np.asarray(a, dtype=np.float64).ravel())


@dataclasses.dataclass
class TopologyArrays:
    """Flat 0-based topology in the layout of mdx_topology."""
    atom_moltype: np.ndarray
    atom_mol: np.ndarray
    atom_ljtype: np.ndarray
    atom_charge: np.ndarray
    atom_charge0: np.ndarray
    atom_mass: np.ndarray
    type_ee: np.ndarray
    type_move: np.ndarray
    type_c14el: np.ndarray
    type_c14lj: np.ndarray
    type_pair: np.ndarray          # (nmoltypes, nmoltypes)
    ntypepairs: int
    a6: np.ndarray                 # (nlj, nlj)
    b12: np.ndarray
    a6_0: np.ndarray
    b12_0: np.ndarray
    a6_14: np.ndarray
    b12_14: np.ndarray
    excl_start: np.ndarray
    excl_atom: np.ndarray
    excl_is14: np.ndarray
    bond_i: np.ndarray
    bond_j: np.ndarray
    bond_param: np.ndarray
    bp_morse: np.ndarray
    bp_r0: np.ndarray
    bp_k: np.ndarray
    bp_d: np.ndarray
    bp_rho: np.ndarray
    angle_i: np.ndarray
    angle_j: np.ndarray
    angle_k: np.ndarray
    angle_param: np.ndarray
    ap_theta0: np.ndarray
    ap_k: np.ndarray
    tors_i: np.ndarray
    tors_j: np.ndarray
    tors_k: np.ndarray
    tors_l: np.ndarray
    tors_param: np.ndarray
    tp_kind: np.ndarray
    tp_mult: np.ndarray
    tp_k: np.ndarray
    tp_delta: np.ndarray
    tp_k1: np.ndarray
    tp_k2: np.ndarray
    tp_k3: np.ndarray
    tp_k4: np.ndarray
    tp_k5: np.ndarray
    # fSPC water (STRBND): per molecule atoms O, H1, H2 and 12 parameters
    spc_atom: np.ndarray = dataclasses.field(default_factory=lambda: np.zeros((0, 3), np.int32))
    spc_par: np.ndarray = dataclasses.field(default_factory=lambda: np.zeros((0, 12)))
    # bond template (parameter index) of each record's R_OH1 bond, or None
    spc_bond: Optional[np.ndarray] = None
    _keep: Optional[list] = None

    @property
    def natoms(self) -> int:
        return int(len(self.atom_mol))

    @property
    def nmols(self) -> int:
        return int(self.atom_mol.max()) + 1 if len(self.atom_mol) else 0

    @property
    def nmoltypes(self) -> int:
        return int(len(self.type_ee))

    @property
    def nljtypes(self) -> int:
        return int(self.a6.shape[0])

    def struct(self) -> MdxTopology:
        """Build the ctypes struct; the arrays stay referenced by self."""
        keep = []

        def ptr(a, kind):
            a = {"i": _i32, "d": _f64}[kind](a) if kind != "b" else np.ascontiguousarray(a, dtype=np.int8).ravel()
            if a.size == 0:
                a = np.zeros(1, dtype=a.dtype)
            keep.append(a)
            ctype = {"i": C.c_int32, "d": C.c_double, "b": C.c_int8}[kind]
            return a.ctypes.data_as(C.POINTER(ctype))

        s = MdxTopology()
        s.natoms = self.natoms
        s.nmols = self.nmols
        s.nmoltypes = self.nmoltypes
        s.nljtypes = self.nljtypes
        s.ntypepairs = int(self.ntypepairs)
        for name in ("atom_moltype", "atom_mol", "atom_ljtype", "type_ee", "type_move", "type_pair",
                     "excl_start", "excl_atom", "bond_i", "bond_j", "bond_param", "bp_morse",
                     "angle_i", "angle_j", "angle_k", "angle_param", "tors_i", "tors_j", "tors_k",
                     "tors_l", "tors_param", "tp_kind", "tp_mult"):
            setattr(s, name, ptr(getattr(self, name), "i"))
        for name in ("atom_charge", "atom_charge0", "atom_mass", "type_c14el", "type_c14lj", "a6",
                     "b12", "a6_0", "b12_0", "a6_14", "b12_14", "bp_r0", "bp_k", "bp_d", "bp_rho",
                     "ap_theta0", "ap_k", "tp_k", "tp_delta", "tp_k1", "tp_k2", "tp_k3", "tp_k4",
                     "tp_k5"):
            setattr(s, name, ptr(getattr(self, name), "d"))
        s.excl_is14 = ptr(self.excl_is14, "b")
        s.nbonds = len(self.bond_i)
        s.nbondparams = len(self.bp_r0)
        s.nangles = len(self.angle_i)
        s.nangleparams = len(self.ap_k)
        s.ntors = len(self.tors_i)
        s.ntorsparams = len(self.tp_kind)
        s.nspc = int(len(self.spc_atom))
        s.spc_atom = ptr(np.asarray(self.spc_atom).ravel(), "i")
        s.spc_par = ptr(np.asarray(self.spc_par).ravel(), "d")
        s.spc_bond = ptr(self.spc_bond, "i") if self.spc_bond is not None else None
        self._keep = keep
        return s


def make_params(box, rcut, rshort, alpha, fexp, coulomb_factor, list_mode=MDX_LIST_REFERENCE,
                skin=0.0, ee_enabled=False, ee_raex=0.0, ee_self_scale=1.0,
                precision=MDX_PRECISION_MIXED) -> MdxParams:
    p = MdxParams()
    for k in range(3):
        p.box[k] = float(box[k])
    p.rcut = float(rcut)
    p.rshort = float(rshort)
    p.skin = float(skin)
    p.alpha = float(alpha)
    p.fexp = float(fexp)
    p.coulomb_factor = float(coulomb_factor)
    p.list_mode = int(list_mode)
    p.ee_enabled = 1 if ee_enabled else 0
    p.ee_raex = float(ee_raex)
    p.ee_self_scale = float(ee_self_scale)
    p.precision = int(precision)
    return p


def as_ptr(a: np.ndarray, ctype=C.c_double):
    return a.ctypes.data_as(C.POINTER(ctype))
