"""Host-side parameter preparation for the force engine.

`from_system` derives, from a `System`, the tables the reference builds in
IUNITS/BNDLST (setup.f:27-702, 753-939): non-bonded type merging, Lennard-
Jones tables by combination rule, type-pair energy indices MDX, bound-atom
lists with 1-4 marks, and per-molecule bonded term lists.  Energies are in
kJ/mol, lengths in Angstrom, charges in e and the Coulomb factor is the
reference's 1389.165 kJ/mol A e^-2 (prcm.h constants).

`from_dump` takes the same tables from a dump of the reference program
(oracle/ref), in the reference's internal units, so the engine and the
oracle can be run on exactly the parameters the reference used.
"""
from __future__ import annotations

import math
from typing import Dict, List, Tuple

import numpy as np

from .abi import TopologyArrays, make_params, MDX_LIST_REFERENCE, MDX_LIST_VERLET
from .systems import COULF_KJ, System

TODGR = 180.0 / 3.1415926536          # prcm.h:31
FSPC_EEP, FSPC_GGP = -884.63, 467.31  # STRBND cross constants, kJ/mol/A^2 (forces.f:537)


def _lb(s1, s2, e1, e2) -> Tuple[float, float]:
    """Lorentz-Berthelot rule, the reference default (setup.f:743-745)."""
    s = 0.5 * (s1 + s2)
    ee = math.sqrt(e1 * e2)
    return -4.0 * ee * s ** 6, 4.0 * ee * s ** 12


def type_pair_index(ntypes: int) -> Tuple[np.ndarray, int]:
    """MDX(ITYP,JTYP) - 1 and MOLINT (setup.f:237-252)."""
    mdx = np.zeros((ntypes, ntypes), np.int32)
    mt = 0
    for i in range(ntypes):
        for j in range(i, ntypes):
            mdx[i, j] = mdx[j, i] = mt
            mt += 1
    return mdx, mt + 1


def bound_lists(system: System) -> Tuple[List[List[int]], List[List[bool]]]:
    """Per-atom bound neighbours with 1-4 flags, registered in the order of
    BNDLST (setup.f:781-849): bonds (1-2), then bond pairs sharing an atom
    (1-3), then bond triples (1-4).  A pair already registered keeps its
    first mark (BNDREG, setup.f:893-939).  Bonds with ID = 2 are skipped."""
    nat = system.natoms
    nb: List[List[int]] = [[] for _ in range(nat)]
    is14: List[List[bool]] = [[] for _ in range(nat)]
    base = 0
    for t, cnt in zip(system.types, system.counts):
        ns = t.nsites
        bonds = [(b[1] - 1, b[2] - 1) for b in t.bonds if b[0] != 2]
        reg: List[Tuple[int, int, bool]] = []

        def add(a, b, f14):
            if a != b:
                reg.append((a, b, f14))
        for a, b in bonds:
            add(a, b, False)
        for s in range(ns):
            for jj, (ib, jb) in enumerate(bonds):
                if s == ib:
                    for ik, jk in bonds[jj + 1:]:
                        if s == ik:
                            add(jb, jk, False)
                        elif s == jk:
                            add(jb, ik, False)
                elif s == jb:
                    for ik, jk in bonds[jj + 1:]:
                        if s == ik:
                            add(ib, jk, False)
                        elif s == jk:
                            add(ib, ik, False)
        for (is_, js) in bonds:
            for (ij, jj_) in bonds:
                if is_ == jj_:
                    for (ik, jk) in bonds:
                        if js == ik:
                            add(ij, jk, True)
                        elif js == jk:
                            add(ij, ik, True)
                elif is_ == ij:
                    for (ik, jk) in bonds:
                        if js == ik:
                            add(jj_, jk, True)
                        elif js == jk:
                            add(jj_, ik, True)
        # one molecule's lists, in BNDREG order
        loc_nb: List[List[int]] = [[] for _ in range(ns)]
        loc_14: List[List[bool]] = [[] for _ in range(ns)]
        for a, b, f14 in reg:
            if b not in loc_nb[a]:
                loc_nb[a].append(b)
                loc_14[a].append(f14)
            if a not in loc_nb[b]:
                loc_nb[b].append(a)
                loc_14[b].append(f14)
        for m in range(cnt):
            off = base + m * ns
            for s in range(ns):
                nb[off + s] = [off + x for x in loc_nb[s]]
                is14[off + s] = list(loc_14[s])
        base += cnt * ns
    return nb, is14


def from_system(system: System, lam: float | None = None) -> Tuple[TopologyArrays, Dict]:
    """Engine topology and parameters for a generated system (kJ/mol units).

    With an expanded-ensemble type (system.ee_type) the type's charges and
    every LJ pair touching it are scaled by lambda^PW1 and the soft-core
    shift is RAEX = BEE (1 - lambda^PW2) (RECINT, mdee.f:86-137)."""
    types, counts = system.types, system.counts
    ntypes = len(types)
    lam = system.ee_lambda if lam is None else lam
    ee = np.ones(ntypes, np.int32)
    if system.ee_type is not None:
        ee[system.ee_type] = 0
    # non-bonded types: one per distinct (sigma, eps, ee flag)
    ljkeys: List[Tuple[float, float, int]] = []
    site_lj: List[List[int]] = []
    for it, t in enumerate(types):
        row = []
        for s in t.sites:
            key = (s.sigma, s.eps, int(ee[it]))
            if key not in ljkeys:
                ljkeys.append(key)
            row.append(ljkeys.index(key))
        site_lj.append(row)
    nlj = len(ljkeys)
    a6 = np.zeros((nlj, nlj))
    b12 = np.zeros((nlj, nlj))
    for i, (si, ei, _) in enumerate(ljkeys):
        for j, (sj, ej, _) in enumerate(ljkeys):
            a6[i, j], b12[i, j] = _lb(si, sj, ei, ej)
    a6_0, b12_0 = a6.copy(), b12.copy()
    scale = lam ** system.ee_pw1
    if system.ee_type is not None:
        for i, (_, _, fi) in enumerate(ljkeys):
            for j, (_, _, fj) in enumerate(ljkeys):
                if fi == 0 or fj == 0:
                    a6[i, j] *= scale
                    b12[i, j] *= scale
    mdx, molint = type_pair_index(ntypes)
    # per-atom arrays
    moltype, mol, ljt, q, q0, mass = [], [], [], [], [], []
    im = 0
    for it, (t, c) in enumerate(zip(types, counts)):
        for _ in range(c):
            for k, s in enumerate(t.sites):
                moltype.append(it)
                mol.append(im)
                ljt.append(site_lj[it][k])
                q0.append(s.charge)
                q.append(s.charge * (scale if ee[it] == 0 else 1.0))
                mass.append(s.mass)
            im += 1
    nb, is14 = bound_lists(system)
    excl_start = np.zeros(system.natoms + 1, np.int32)
    excl_start[1:] = np.cumsum([len(x) for x in nb])
    excl_atom = np.array([a for x in nb for a in x], np.int32)
    excl_is14 = np.array([f for x in is14 for f in x], np.int8)
    # bonded terms
    bi, bj, bpar = [], [], []
    bpm, br0, bk, bd, brho = [], [], [], [], []
    ai, aj, ak, apar, ath, akk = [], [], [], [], [], []
    ti, tj, tk, tl, tpar = [], [], [], [], []
    tkind, tmult, tkk, tdel, tk1, tk2, tk3, tk4, tk5 = [], [], [], [], [], [], [], [], []
    spc_at, spc_pr, spc_bd = [], [], []
    base = 0
    for t, c in zip(types, counts):
        b0, a0, t0 = len(br0), len(ath), len(tkind)
        for b in t.bonds:
            bpm.append(1 if b[0] == 1 else 0)
            br0.append(b[3]); bk.append(b[4]); bd.append(b[5]); brho.append(b[6])
        for a in t.angles:
            ath.append(a[3] / TODGR); akk.append(a[4])
        for tt in t.torsions:
            tkind.append(0); tmult.append(tt[6]); tkk.append(tt[5]); tdel.append(tt[4] / TODGR)
            tk1.append(0.0); tk2.append(0.0); tk3.append(0.0); tk4.append(0.0); tk5.append(0.0)
        for tt in t.tors1:
            tkind.append(1); tmult.append(0); tkk.append(0.0); tdel.append(0.0)
            tk1.append(tt[4]); tk2.append(tt[5]); tk3.append(tt[6]); tk4.append(0.0); tk5.append(0.0)
        ns = t.nsites
        for m in range(c):
            off = base + m * ns
            if t.fspc:
                # STRBND replaces the type's listed bonds, angles and
                # torsions (setup.f:297): one record per molecule
                b1, b2, b3 = t.bonds[0], t.bonds[1], t.bonds[2]
                spc_at.append((off, off + 1, off + 2))
                spc_pr.append((b1[3], b2[3], b3[3], b1[5], b2[5], b3[4], b1[6], b2[6],
                               FSPC_EEP, FSPC_GGP, 2.0, 1.0))
                spc_bd.append(b0)
                continue
            for kb, b in enumerate(t.bonds):
                bi.append(off + b[1] - 1); bj.append(off + b[2] - 1); bpar.append(b0 + kb)
            for ka, a in enumerate(t.angles):
                ai.append(off + a[0] - 1); aj.append(off + a[1] - 1); ak.append(off + a[2] - 1)
                apar.append(a0 + ka)
            for kt, tt in enumerate(list(t.torsions) + list(t.tors1)):
                ti.append(off + tt[0] - 1); tj.append(off + tt[1] - 1); tk.append(off + tt[2] - 1)
                tl.append(off + tt[3] - 1); tpar.append(t0 + kt)
        base += c * ns
    topo = TopologyArrays(
        atom_moltype=np.array(moltype, np.int32), atom_mol=np.array(mol, np.int32),
        atom_ljtype=np.array(ljt, np.int32), atom_charge=np.array(q), atom_charge0=np.array(q0),
        atom_mass=np.array(mass), type_ee=ee, type_move=np.ones(ntypes, np.int32),
        type_c14el=np.array([t.c14el if t.c14el is not None else 1.0 for t in types]),
        type_c14lj=np.array([t.c14lj if t.c14lj is not None else 1.0 for t in types]),
        type_pair=mdx, ntypepairs=molint, a6=a6, b12=b12, a6_0=a6_0, b12_0=b12_0,
        a6_14=a6_0.copy(), b12_14=b12_0.copy(),
        excl_start=excl_start, excl_atom=excl_atom, excl_is14=excl_is14,
        bond_i=np.array(bi, np.int32), bond_j=np.array(bj, np.int32), bond_param=np.array(bpar, np.int32),
        bp_morse=np.array(bpm, np.int32), bp_r0=np.array(br0), bp_k=np.array(bk), bp_d=np.array(bd),
        bp_rho=np.array(brho),
        angle_i=np.array(ai, np.int32), angle_j=np.array(aj, np.int32), angle_k=np.array(ak, np.int32),
        angle_param=np.array(apar, np.int32), ap_theta0=np.array(ath), ap_k=np.array(akk),
        tors_i=np.array(ti, np.int32), tors_j=np.array(tj, np.int32), tors_k=np.array(tk, np.int32),
        tors_l=np.array(tl, np.int32), tors_param=np.array(tpar, np.int32),
        tp_kind=np.array(tkind, np.int32), tp_mult=np.array(tmult, np.int32), tp_k=np.array(tkk),
        tp_delta=np.array(tdel), tp_k1=np.array(tk1), tp_k2=np.array(tk2), tp_k3=np.array(tk3),
        tp_k4=np.array(tk4), tp_k5=np.array(tk5),
        spc_atom=np.array(spc_at, np.int32).reshape(-1, 3), spc_par=np.array(spc_pr).reshape(-1, 12),
        spc_bond=np.array(spc_bd, np.int32))
    raex = system.ee_bee * (1.0 - lam ** system.ee_pw2) if system.ee_type is not None else 0.0
    params = dict(box=system.box, rcut=system.rcut, rshort=system.rshort, alpha=system.alpha,
                  fexp=system.ewald_b, coulomb_factor=COULF_KJ,
                  list_mode=MDX_LIST_VERLET if system.skin > 0 else MDX_LIST_REFERENCE,
                  skin=system.skin, ee_enabled=system.ee_type is not None, ee_raex=raex,
                  ee_self_scale=lam)
    return topo, params


def _spc_records(d: Dict[str, List[np.ndarray]]) -> Tuple[np.ndarray, np.ndarray, np.ndarray]:
    """STRBND records (forces.f:555-590) of every type with IPOT 1 or 2:
    atoms ISADDR+3(n-1)+1..3 of molecule n, parameters from the type's
    first three bond records (IADB), cross constants EEP, GGP times
    FACTOR = 1e3/AVSNO/UNITE; and each record's first bond template
    (IADB - 1, where STRBND adds its BR/BE terms, forces.f:655-663)."""
    g = lambda k: d[k][0]
    ipot, iadb, nspec, nsits = g("IPOT"), g("IADB"), g("NSPEC"), g("NSITS")
    rb, fb, db, ro, lmove = g("RB"), g("FB"), g("DB"), g("RO"), g("LMOVE")
    factor = 1.0e3 / 6.02252e23 / g("SCALARS")[11]
    at, pr, bd = [], [], []
    base = 0
    for it in range(len(ipot)):
        if ipot[it] in (1, 2):
            i1 = iadb[it] - 1
            i2, i3 = i1 + 1, i1 + 2
            a_, b_ = (db[i1], db[i2]) if ipot[it] == 2 else (fb[i1], fb[i2])
            for m in range(nspec[it]):
                off = base + m * nsits[it]
                at.append((off, off + 1, off + 2))
                pr.append((rb[i1], rb[i2], rb[i3], a_, b_, fb[i3], ro[i1], ro[i2], FSPC_EEP * factor,
                           FSPC_GGP * factor, float(ipot[it]), float(lmove[it] != 0)))
                bd.append(i1)
        base += nspec[it] * nsits[it]
    return np.array(at, np.int32).reshape(-1, 3), np.array(pr, float).reshape(-1, 12), np.array(bd, np.int32)


def from_dump(d: Dict[str, List[np.ndarray]], lam: float = 1.0) -> Tuple[TopologyArrays, Dict]:
    """Topology and parameters exactly as the reference held them (internal
    units), from a dump of oracle/ref.  `lam` is EC(ME) of an expanded-
    ensemble run (charge scale of the lambda-coupled type in the Ewald
    self term, forces.f:1434-1435)."""
    g = lambda k: d[k][0]
    sizes = g("SIZES")
    nstot, ntypes, nop, nnbt = (int(x) for x in sizes[1:5])
    molint = int(sizes[9])
    sc = g("SCALARS")
    nnbb = g("NNBB")
    inbb = g("INBB")
    excl_start = np.zeros(nstot + 1, np.int32)
    excl_start[1:] = np.cumsum(nnbb)
    ex = inbb.astype(np.int64)
    nbond = len(g("IBI"))
    tp = lambda k: g(k).reshape(nnbt, nnbt)
    topo = TopologyArrays(
        atom_moltype=g("ITYPE") - 1, atom_mol=g("NNUM") - 1, atom_ljtype=g("INBTA") - 1,
        atom_charge=g("Q"), atom_charge0=g("CHARGEA"), atom_mass=g("MASSA"),
        type_ee=g("IEE"), type_move=g("LMOVE"), type_c14el=g("C14EL"), type_c14lj=g("C14LJ"),
        type_pair=g("MDX").reshape(ntypes, ntypes) - 1, ntypepairs=molint,
        a6=tp("A6LJ"), b12=tp("B12LJ"), a6_0=tp("A6LJ0"), b12_0=tp("B12LJ0"),
        a6_14=tp("A6_14"), b12_14=tp("B12_14"),
        excl_start=excl_start, excl_atom=(np.abs(ex) - 1).astype(np.int32),
        excl_is14=(ex < 0).astype(np.int8),
        bond_i=g("IBI") - 1, bond_j=g("JBJ") - 1, bond_param=g("IBUK") - 1,
        bp_morse=(g("ID") == 1).astype(np.int32), bp_r0=g("RB"), bp_k=g("FB"), bp_d=g("DB"),
        bp_rho=g("RO"),
        angle_i=g("IAI") - 1, angle_j=g("JAJ") - 1, angle_k=g("KAK") - 1, angle_param=g("IAUK") - 1,
        ap_theta0=g("RA"), ap_k=g("FA"),
        tors_i=g("ITI") - 1, tors_j=g("JTJ") - 1, tors_k=g("KTK") - 1, tors_l=g("LTL") - 1,
        tors_param=g("ITUK") - 1, tp_kind=g("ITORS"), tp_mult=g("NMUL"), tp_k=g("FT"),
        tp_delta=g("RT"), tp_k1=g("FT1"), tp_k2=g("FT2"), tp_k3=g("FT3"), tp_k4=g("FT4"),
        tp_k5=g("FT5"))
    assert nbond == len(topo.bond_i)
    if "IPOT" in d:
        topo.spc_atom, topo.spc_par, topo.spc_bond = _spc_records(d)
    ee_enabled = bool(sizes[11])
    params = dict(box=(sc[0], sc[1], sc[2]), rcut=sc[3], rshort=sc[5], alpha=sc[7], fexp=sc[8],
                  coulomb_factor=sc[9], list_mode=MDX_LIST_REFERENCE, skin=0.0,
                  ee_enabled=ee_enabled, ee_raex=sc[14], ee_self_scale=lam)
    return topo, params


def params_struct(params: Dict):
    return make_params(**params)


# prcm.h:22 (reference physical constants)
AVSNO = 6.02252e23
BOLTZ_J = 1.38054e-23
BOLTZ_KJMOL = AVSNO * BOLTZ_J * 1e-3     # kJ/mol/K


def pimd_beadfac(temperature: float, nbeads: int) -> float:
    """Spring constant per unit mass of the path-integral ring in kJ/mol/A^2
    per g/mol (BEADFAC, setup.f:696-697, converted from the reference's
    internal units): E = BEADFAC m |r_k - r_k+1|^2 per link with
    BEADFAC = P (k_B T / hbar)^2 / 2, the reference's k_B = 1.38054e-23 J/K
    and hbar = 1.0545e-34 J s (prcm.h:22-26)."""
    kt_hbar = 1.38054e-23 * temperature / 1.0545e-34
    return 0.5e-26 * nbeads * kt_hbar * kt_hbar


def md_params(system, tolerance: float = 1e-6) -> Dict:
    """MD-segment parameters of a generated system in the host units
    (A, fs, kJ/mol, g/mol): the constrained bonds of every molecule when
    the system runs with SHAKE (all bonds with ID != 2 of the constrained
    molecule types -- ISHEJK, input.f:292-296 -- in .mmol order, as SHEJK
    walks them, mdstep.f:1115-1140),
    the degrees of freedom FNST (main.f:297-348: 3 per atom, minus one per
    constraint, minus 3) and the Nose-Hoover time constant.  Without SHAKE
    the integrator is the multiple-time-step DOUBLE with the system's NFREQ
    (the reference chooses it the same way, main.f:918-926)."""
    ci, cj, ln = [], [], []
    off = 0
    ndof = 0.0
    for k, (t, cnt) in enumerate(zip(system.types, system.counts)):
        con = system.shake and (system.shake_types is None or k in system.shake_types)
        for m in range(cnt):
            if con:
                for b in t.bonds:
                    if b[0] != 2:
                        ci.append(off + b[1] - 1)
                        cj.append(off + b[2] - 1)
                        ln.append(b[3])
            off += t.nsites
        ndof += 3.0 * t.nsites * cnt
        if con:
            ndof -= sum(1 for b in t.bonds if b[0] != 2) * cnt
    ndof -= 3.0
    return dict(dt=system.dt_fs, accel_factor=1e-4, ekin_factor=1e4, boltz=BOLTZ_KJMOL,
                temperature=system.temperature, tau_t=30.0, sc0=0.0, ndof=ndof, shake_tol=tolerance,
                shake_bdtol=float(np.float32(0.0001)), shake_maxit=2000, rebuild_every=system.rebuild_every,
                step0=0, cons_i=np.array(ci, np.int32), cons_j=np.array(cj, np.int32),
                cons_len=np.array(ln, np.float64),
                integrator=0 if system.shake else 1, nfreq=1 if system.shake else int(system.nfreq))
