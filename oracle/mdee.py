"""Host restatement of the reference's expanded-ensemble energy differences
(test infrastructure only: the checker of the device mdx_ee_energies,
never part of the product path).

Both routines are evaluated as functions of the coupling parameter lambda
of the lambda-coupled molecule types (IEE = 0): the reference's "old"
energies (EEO, EFO, with the current Q, A6LJ, B12LJ and RAEX that RECINT
set from EC(ME), mdee.f:86-137) are these functions at EC(ME), its "new"
ones (EEN, EFN) at EC(ME1).

difen (DIFEN, mdee.f:379-530): over every listed pair (long list NBL1/NBL2,
then short list NBS1/NBS2; 1-based, 1-4 pairs with a negative first index)
with at least one atom of a lambda-coupled type and the two atoms in
different molecules, no cutoff test:
    s = lambda^PW1, shift = BEE (1 - lambda^PW2)
    charge scale: s per lambda-coupled atom of the pair (ESCL)
    LJ scale: s when either atom is lambda-coupled (ESCI)
    E = qscale q0_i q0_j COULF erfc(alpha r) / (r + shift)
        + (ljscale A6_0 + ljscale B12_0 (r + shift)^-6) (r + shift)^-6
with the A&S erfc polynomial (prcm.h:27-29) and the one-shift minimum
image (service.f:953-956).

diffur (DIFFUR, mdee.f:265-370): sum over the k set of AKC |S(k)|^2 with
S(k) = sum_i q_i e^{i k.r_i}, q_i = s CHARGE_i for lambda-coupled atoms and
the current Q_i for the others; minus the self terms of the lambda-coupled
types COULF (s CHARGE)^2 alpha / sqrt(pi) NSPEC per site; plus, over every
pair of sites of each lambda-coupled molecule, s^2 COULF CHARGE_i CHARGE_j
(erfc(alpha r) - 1) / r.
"""
from __future__ import annotations

from typing import Sequence, Tuple

import numpy as np

from oracle import orc

# prcm.h:27-32
A1, A2, A3, A4, A5, B1 = 0.254829592, -0.284496736, 1.421413741, -1.453152027, 1.061405429, 0.3275911
PI = 3.1415926536


def _pbc(d: np.ndarray, box: np.ndarray) -> np.ndarray:
    return d - np.where(np.abs(d) > 0.5 * box, np.sign(d) * box, 0.0)


def _erfc(ar: np.ndarray) -> np.ndarray:
    t = 1.0 / (1.0 + B1 * ar)
    return ((((A5 * t + A4) * t + A3) * t + A2) * t + A1) * t * np.exp(-ar ** 2)


def difen(topo, prm, pos: np.ndarray, lists, lams: Sequence[float], pw1: float, pw2: float,
          bee: float, rcut2: float = np.inf) -> np.ndarray:
    """Pair part per lambda.  rcut2: a cutoff test (Verlet lists); the
    reference lists take none."""
    s1, s2, l1, l2 = lists
    i = np.abs(np.concatenate([l1, s1]).astype(np.int64)) - 1
    j = np.concatenate([l2, s2]).astype(np.int64) - 1
    ee = np.asarray(topo.type_ee)[np.asarray(topo.atom_moltype)] == 0
    mol = np.asarray(topo.atom_mol)
    sel = (ee[i] | ee[j]) & (mol[i] != mol[j])
    i, j = i[sel], j[sel]
    box = np.asarray(prm["box"], float)
    d = _pbc(pos[i] - pos[j], box)
    rr = (d[:, 0] ** 2 + d[:, 1] ** 2) + d[:, 2] ** 2
    keep = rr <= rcut2
    i, j, rr = i[keep], j[keep], rr[keep]
    r = np.sqrt(rr)
    erfcr = _erfc(prm["alpha"] * r)
    q0 = np.asarray(topo.atom_charge0)
    lj = np.asarray(topo.atom_ljtype)
    a6 = np.asarray(topo.a6_0)[lj[i], lj[j]]
    b12 = np.asarray(topo.b12_0)[lj[i], lj[j]]
    nee = ee[i].astype(int) + ee[j].astype(int)
    out = []
    for lam in lams:
        s = lam ** pw1
        shift = bee * (1.0 - lam ** pw2)
        qs = np.where(nee == 2, s * s, s)
        r1ni = 1.0 / (r + shift)
        ees = qs * q0[i] * q0[j] * prm["coulomb_factor"] * erfcr * r1ni
        r6 = r1ni ** 6
        esr = (s * a6 + s * b12 * r6) * r6
        out.append(float(np.sum(ees + esr)))
    return np.array(out)


def diffur(topo, prm, pos: np.ndarray, lams: Sequence[float], pw1: float) -> np.ndarray:
    """Reciprocal part with the self and intramolecular corrections, per
    lambda, over the INIFUR k set of the box in prm."""
    from mdynamix_amd import topology
    ts = topo.struct()
    ps = topology.params_struct(dict(prm, list_mode=0, skin=0.0))
    rkx, rky, rkz, _, _ = orc.inifur(ts, ps, topo.nmoltypes)
    box = np.asarray(prm["box"], float)
    k = np.stack([rkx / box[0], rky / box[1], rkz / box[2]], 1)
    rk2 = (k[:, 0] ** 2 + k[:, 1] ** 2) + k[:, 2] ** 2
    alp2 = prm["alpha"] ** 2
    coulf = prm["coulomb_factor"]
    cfur = 4.0 * PI * coulf / (box[0] * box[1] * box[2])
    akc = cfur * np.exp(-0.25 * rk2 / alp2) / rk2
    ee = np.asarray(topo.type_ee)[np.asarray(topo.atom_moltype)] == 0
    q = np.asarray(topo.atom_charge)
    q0 = np.asarray(topo.atom_charge0)
    cs_a = np.zeros(len(k)); sn_a = np.zeros(len(k)); cs_b = np.zeros(len(k)); sn_b = np.zeros(len(k))
    for c0 in range(0, len(pos), 512):
        ph = k @ pos[c0:c0 + 512].T
        c, s_ = np.cos(ph), np.sin(ph)
        qa = np.where(ee[c0:c0 + 512], 0.0, q[c0:c0 + 512])
        qb = np.where(ee[c0:c0 + 512], q0[c0:c0 + 512], 0.0)
        cs_a += c @ qa; sn_a += s_ @ qa; cs_b += c @ qb; sn_b += s_ @ qb
    # self terms and intramolecular pairs of the lambda-coupled molecules
    abc = prm["alpha"] / np.sqrt(PI)
    mol = np.asarray(topo.atom_mol)
    self0 = float(np.sum(coulf * q0[ee] ** 2 * abc))          # NSPEC x sites = every such atom
    intra0 = 0.0
    for m in np.unique(mol[ee]):
        at = np.nonzero(mol == m)[0]
        for a in range(len(at)):
            for b in range(a + 1, len(at)):
                ia, ib = at[a], at[b]
                d = _pbc(pos[ia] - pos[ib], box)
                r2 = (d[0] ** 2 + d[1] ** 2) + d[2] ** 2
                r1 = np.sqrt(r2)
                intra0 += coulf * q0[ia] * q0[ib] * (_erfc(prm["alpha"] * r1) - 1.0) * (r1 / r2)
    out = []
    for lam in lams:
        s = lam ** pw1
        cs, sn = cs_a + s * cs_b, sn_a + s * sn_b
        out.append(float(np.sum((sn ** 2 + cs ** 2) * akc)) - s * s * self0 + s * s * intra0)
    return np.array(out)


def from_fixture(d, call: int) -> Tuple[np.ndarray, np.ndarray, np.ndarray, Tuple[float, float, float]]:
    """The reference's own (EFN, EFO, EEN, EEO) per subensemble ME1 at a
    dumped call, the subensembles' EC and (PW1, PW2, BEE)."""
    ed = np.asarray(d["EEDIFF"][call]).reshape(-1, 4)
    par = np.asarray(d["EEPAR"][call])
    return ed, np.asarray(d["EC"][call]), int(par[0]), (float(par[1]), float(par[2]), float(par[3]))
