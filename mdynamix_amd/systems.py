"""Seeded molecular systems for the MDynaMix force engine.

A `System` holds what the reference program reads from its input files:
molecule types (the content of `.mmol` files, format read by READMOL,
input.f:936-1340), molecule counts, the periodic box, run parameters of the
keyword file (INPUT, input.f:166-674) and atom coordinates (XMOL start file,
main.f:551-584).  The same object is written out as those files so the
reference program and this engine see identical inputs.

Builders cover the workloads the engine is measured on:

* SPC/E water + Na+/Cl- boxes (1,532 atoms at 25 A; 96,174 atoms at 98.7 A),
* flexible united-atom chains solvated in SPC/E water (bonds, angles,
  MM3-type dihedrals, 1-4 pairs),
* SPC-FW flexible water for path-integral beads,
* water + NaCl boxes with one lambda-coupled solute (expanded-ensemble jobs).

Molecules are placed on a jittered lattice with random orientations drawn
from one numpy Generator seed, so a (builder, seed) pair fixes every
coordinate bit.
"""
from __future__ import annotations

import dataclasses
import math
import os
from typing import List, Optional, Sequence, Tuple

import numpy as np

# Reference physical constants (prcm.h:20-33).  PI and 2/sqrt(pi) are the
# truncated literals the reference computes with, not the exact values.
AVSNO = 6.02252e23
BOLTZ = 1.38054e-23
ELECHG = 1.602e-19
EPS0 = 8.854e-12
PI_REF = 3.1415926536
TOTPI = 1.128379169
# Coulomb constant in kJ/mol * A / e^2 with the reference constants
# (setup.f:109 divided by the internal energy unit, times ENERF*1e-3).
COULF_KJ = (ELECHG / EPS0) * ELECHG / (4.0 * PI_REF * 1.0e-10) * AVSNO * 1.0e-3


@dataclasses.dataclass
class Site:
    name: str
    xyz: Tuple[float, float, float]
    mass: float
    charge: float
    sigma: float
    eps: float


@dataclasses.dataclass
class MolType:
    """One `.mmol` molecule type (READMOL, input.f:936-1340)."""
    name: str
    sites: List[Site]
    # (ID, i, j, r0, K, D, rho) with 1-based site indices; E = K (r-r0)^2
    bonds: List[Tuple[int, int, int, float, float, float, float]] = dataclasses.field(default_factory=list)
    # (i, j, k, theta0_deg, K); E = K (theta-theta0)^2
    angles: List[Tuple[int, int, int, float, float]] = dataclasses.field(default_factory=list)
    # AMBER-type (i, j, k, l, delta_deg, K, mult); E = K (1 + cos(m phi - delta))
    torsions: List[Tuple[int, int, int, int, float, float, int]] = dataclasses.field(default_factory=list)
    # MM3-type "Tors1" (i, j, k, l, K1, K2, K3)
    tors1: List[Tuple[int, int, int, int, float, float, float]] = dataclasses.field(default_factory=list)
    c14el: Optional[float] = None
    c14lj: Optional[float] = None
    # "fSPC" keyword (input.f:1304-1313): flexible SPC water, whose
    # intramolecular terms STRBND computes from the first three bonds
    fspc: bool = False

    @property
    def nsites(self) -> int:
        return len(self.sites)

    def mass(self) -> float:
        return sum(s.mass for s in self.sites)

    def local_coords(self) -> np.ndarray:
        """Site coordinates with the molecular centre of mass at zero
        (input.f:1000-1028)."""
        r = np.array([s.xyz for s in self.sites], dtype=np.float64)
        m = np.array([s.mass for s in self.sites], dtype=np.float64)
        return r - (m[:, None] * r).sum(0) / m.sum()

    def to_mmol(self) -> str:
        out = ["# molecule %s (written by mdynamix_amd.systems)" % self.name,
               "# number of sites", " %d" % self.nsites,
               "# name  X  Y  Z  mass  charge  sigma  epsilon"]
        for s in self.sites:
            out.append("%-4s %.10f %.10f %.10f %.6f %.6f %.6f %.6f" % (
                s.name, s.xyz[0], s.xyz[1], s.xyz[2], s.mass, s.charge, s.sigma, s.eps))
        out += ["# reference lines", "1", "   generated topology", "# bonds", " %d" % len(self.bonds)]
        for b in self.bonds:
            out.append(" %d %d %d %.8f %.8f %.6f %.6f" % b)
        out += ["# angles", " %d" % len(self.angles)]
        for a in self.angles:
            out.append(" %d %d %d %.8f %.8f" % a)
        out += ["# dihedrals", " %d" % len(self.torsions)]
        for t in self.torsions:
            out.append(" %d %d %d %d %.8f %.8f %d" % t)
        if self.tors1:
            out += ["Tors1", " %d" % len(self.tors1)]
            for t in self.tors1:
                out.append(" %d %d %d %d %.8f %.8f %.8f" % t)
        if self.c14el is not None:
            out.append("sel14 %.8f" % self.c14el)
        if self.c14lj is not None:
            out.append("slj14 %.8f" % self.c14lj)
        if self.fspc:
            out.append("fSPC")
        out.append("End")
        return "\n".join(out) + "\n"


@dataclasses.dataclass
class System:
    """A periodic system plus the run parameters of the force path."""
    name: str
    types: List[MolType]
    counts: List[int]
    box: Tuple[float, float, float]
    pos: np.ndarray                  # (N, 3) float64, atoms ordered type by type
    rcut: float = 10.0
    rshort: float = 5.0
    ewald_a: float = 2.8             # alpha * R_cut ("Electrostatics Ewald A B")
    ewald_b: float = 9.0             # FEXP
    dt_fs: float = 2.0
    shake: bool = True
    # molecule types constrained under SHAKE (None: every type with bonds)
    shake_types: Optional[Tuple[int, ...]] = None
    nfreq: int = 10                  # double time step ratio when not SHAKE
    temperature: float = 300.0
    skin: float = 0.0
    rebuild_every: int = 10
    ee_type: Optional[int] = None    # 0-based index of a lambda-coupled type
    ee_lambda: float = 1.0
    ee_pw1: float = 1.0
    ee_pw2: float = 1.0 / 3.0
    ee_bee: float = 3.0
    seed: int = 0
    # path-integral run of the reference ("PIMD P", input.f:642-659) and its
    # per-atom bead thermostat chains ("PI_thermo T BTAU JNH",
    # input.f:660-666; it switches the Nose thermostat off)
    pimd_beads: int = 0
    pi_thermo: Optional[Tuple[float, float, int]] = None
    # reference input "Average_int yes": per-template bond, angle and
    # dihedral averages (ISTAV, input.f:372-374; GETAVR, aver.f)
    average_int: bool = False
    # restart files (RESTRT, restart.f): "Dump_restart N [ASCII]" and
    # "Read_restart yes [ASCII]" of the reference input (input.f:191-205)
    dump_every: int = 1000000
    dump_ascii: bool = False
    read_restart: bool = False
    read_ascii: bool = False

    # ------------------------------------------------------------------
    @property
    def natoms(self) -> int:
        return int(sum(t.nsites * c for t, c in zip(self.types, self.counts)))

    @property
    def nmol(self) -> int:
        return int(sum(self.counts))

    @property
    def alpha(self) -> float:
        """Ewald alpha (main.f:804-806: ALPHAD = ALPHA / RCUTT when the
        cutoff is at most half the box)."""
        cl = self.box[0]
        if self.rcut <= 0.5 * cl:
            return self.ewald_a / self.rcut
        return 2.0 * self.ewald_a / cl

    def write_reference_inputs(self, directory: str, verbose: int = 5, nsteps: int = 1) -> None:
        """Write md.in, the .mmol files and the XMOL start file."""
        os.makedirs(directory, exist_ok=True)
        for t in self.types:
            with open(os.path.join(directory, t.name + ".mmol"), "w") as f:
                f.write(t.to_mmol())
        lines = ["Main_filename %s" % self.name, "Verbose_level %d" % verbose,
                 "Path_DB .",
                 "Read_restart %s" % (("yes ASCII" if self.read_ascii else "yes") if self.read_restart else "no"),
                 "Dump_restart %d%s" % (self.dump_every, " ASCII" if self.dump_ascii else ""), "Check_only no",
                 "Molecule_types %d" % len(self.types)]
        for t, c in zip(self.types, self.counts):
            lines.append("%s %d" % (t.name, c))
        lines += ["Box %.12f %.12f %.12f" % self.box, "PBC rect"]
        if self.pi_thermo is None:
            lines.append("Nose_thermostat %.3f 30." % self.temperature)
        lines += ["Time_step %.6f" % self.dt_fs, "Number_steps %d" % nsteps]
        if self.shake:
            if self.shake_types is None:
                lines.append("Constraints 1.d-6")
            else:
                flags = " ".join("1" if k in self.shake_types else "0" for k in range(len(self.types)))
                lines.append("Constraints 1.d-6 %s" % flags)
        else:
            lines.append("Double_timestep %d" % self.nfreq)
        lines += ["R_cutoff %.10f" % self.rcut, "R_short %.10f" % self.rshort,
                  "Neighbour_list %d" % self.rebuild_every,
                  "Electrostatics Ewald %.12f %.12f" % (self.ewald_a, self.ewald_b),
                  "Startup XMOL", "Output 1", "COM_check yes 1000000"]
        if self.average_int:
            # one averaging series over the run (GETAVR(2), mdstep.f:243)
            lines += ["Average_int yes", "Serie_av %d" % nsteps]
        if self.ee_type is not None:
            # two sub-ensembles: the job's lambda (started in) and lambda=1
            lines += ["MDEE 2 list", "%.12f 0.0" % self.ee_lambda, "1.0 0.0",
                      "EE_type %d" % (self.ee_type + 1), "EE_start 1",
                      "EE_offset %.12f" % self.ee_bee,
                      "EE_power %.12f %.12f" % (self.ee_pw1, self.ee_pw2)]
        if self.pimd_beads:
            lines.append("PIMD %d" % self.pimd_beads)
            if self.pi_thermo is not None:
                lines.append("PI_thermo %.6f %.6f %d" % self.pi_thermo)
        lines.append("End")
        with open(os.path.join(directory, "md.in"), "w") as f:
            f.write("\n".join(lines) + "\n")
        from .restart import write_start
        names = [s.name[0] for t, c in zip(self.types, self.counts) for _ in range(c) for s in t.sites]
        write_start(os.path.join(directory, self.name + ".start"), names, self.pos, self.box)


# ----------------------------------------------------------------------
# Molecule types
# ----------------------------------------------------------------------
def _water_geometry(r_oh: float, theta_deg: float) -> List[Tuple[float, float, float]]:
    half = math.radians(theta_deg) / 2.0
    return [(0.0, 0.0, 0.0),
            (0.0, -r_oh * math.sin(half), r_oh * math.cos(half)),
            (0.0, r_oh * math.sin(half), r_oh * math.cos(half))]


def spce_water() -> MolType:
    """Rigid SPC/E: q_O=-0.8476, q_H=+0.4238, r_OH=1.0 A, HOH=109.47 deg,
    sigma_OO=3.166 A, eps_OO=0.650 kJ/mol.  The H-H distance is a third
    harmonic bond so SHAKE keeps the triangle rigid."""
    g = _water_geometry(1.0, 109.47)
    hh = round(math.dist(g[1], g[2]), 8)     # as written to the .mmol file
    sites = [Site("OW", g[0], 15.9994, -0.8476, 3.166, 0.650),
             Site("HW1", g[1], 1.008, 0.4238, 0.0, 0.0),
             Site("HW2", g[2], 1.008, 0.4238, 0.0, 0.0)]
    bonds = [(0, 1, 2, 1.0, 2811.0, 0.0, 0.0), (0, 1, 3, 1.0, 2811.0, 0.0, 0.0),
             (0, 2, 3, hh, 687.0, 0.0, 0.0)]
    return MolType("SPCE", sites, bonds)


def fspc_water() -> MolType:
    """Flexible SPC water of the "fSPC" kind (STRBND, forces.f:530-800):
    the molecule must be named H2O*; Morse O-H bonds (depth D, range rho)
    and a harmonic H-H term from the first three bonds, plus STRBND's fixed
    cross terms.  Parameter values in the spirit of Toukan and Rahman
    (kJ/mol, A)."""
    g = _water_geometry(1.0, 109.47)
    sites = [Site("OW", g[0], 15.9994, -0.82, 3.165492, 0.6502),
             Site("HW1", g[1], 1.008, 0.41, 0.0, 0.0),
             Site("HW2", g[2], 1.008, 0.41, 0.0, 0.0)]
    rhh = float(np.linalg.norm(np.asarray(g[1]) - np.asarray(g[2])))
    bonds = [(1, 1, 2, 1.0, 2000.0, 432.58, 2.566), (1, 1, 3, 1.0, 2000.0, 432.58, 2.566),
             (0, 2, 3, rhh, 228.3, 0.0, 0.0)]
    return MolType("H2O", sites, bonds, [], fspc=True)


def spcfw_water() -> MolType:
    """Flexible SPC-FW: O-H k=4431.53 kJ/mol/A^2 r0=1.012 A, H-O-H
    k=317.57 kJ/mol/rad^2 theta0=113.24 deg.  The reference energy has no
    1/2 (E = K dr^2), so the file stores K = k/2."""
    g = _water_geometry(1.012, 113.24)
    sites = [Site("OW", g[0], 15.9994, -0.82, 3.165492, 0.6502),
             Site("HW1", g[1], 1.008, 0.41, 0.0, 0.0),
             Site("HW2", g[2], 1.008, 0.41, 0.0, 0.0)]
    bonds = [(0, 1, 2, 1.012, 4431.53 / 2, 0.0, 0.0), (0, 1, 3, 1.012, 4431.53 / 2, 0.0, 0.0)]
    angles = [(2, 1, 3, 113.24, 317.57 / 2)]
    return MolType("SPCFW", sites, bonds, angles)


def sodium() -> MolType:
    return MolType("NAP", [Site("Na", (0.0, 0.0, 0.0), 22.9898, 1.0, 2.16, 1.47)])


def chloride() -> MolType:
    return MolType("CLM", [Site("Cl", (0.0, 0.0, 0.0), 35.453, -1.0, 4.83, 0.054)])


def _chain_geometry(n: int, rng: np.random.Generator, bond: float, theta_deg: float,
                    compact: Optional[float] = None) -> np.ndarray:
    """A chain conformation: fixed bond and angle, trans-rich dihedrals.
    With `compact` (A) a globule instead: each new site takes the first of
    the three staggered dihedrals (in random order) that keeps it within
    `compact` of the centroid so far and 3 A from every site four or more
    bonds back, else the one farthest from those sites."""
    theta = math.radians(theta_deg)
    x = np.zeros((n, 3))
    x[1] = [bond, 0.0, 0.0]
    x[2] = x[1] + bond * np.array([-math.cos(theta), math.sin(theta), 0.0])

    def place(i, phi):
        b = x[i - 1] - x[i - 2]
        b /= np.linalg.norm(b)
        nrm = np.cross(x[i - 2] - x[i - 3], b)
        nrm /= np.linalg.norm(nrm)
        m = np.cross(nrm, b)
        d = np.array([-bond * math.cos(theta), bond * math.sin(theta) * math.cos(phi),
                      bond * math.sin(theta) * math.sin(phi)])
        return x[i - 1] + d[0] * b + d[1] * m + d[2] * nrm

    for i in range(3, n):
        if compact is None:
            phi = math.pi if rng.random() < 0.75 else rng.choice([math.pi / 3, -math.pi / 3])
            x[i] = place(i, phi)
            continue
        best, best_d = None, -1.0
        for phi in rng.permutation([math.pi, math.pi / 3, -math.pi / 3]):
            p = place(i, phi)
            dmin = np.sqrt(((x[:i - 3] - p) ** 2).sum(1)).min() if i > 3 else 9.0
            if dmin >= 3.0 and np.linalg.norm(p - x[:i].mean(0)) <= compact:
                best = p
                break
            if dmin > best_d:
                best, best_d = p, dmin
        x[i] = best
    return x


def ua_chain(n_sites: int = 46, seed: int = 7, n_charged: int = 4, compact: Optional[float] = None) -> MolType:
    """Flexible united-atom chain: bonds k=2500 kJ/mol/A^2 r0=1.53 A,
    angles k=500 kJ/mol/rad^2 theta0=112 deg, MM3-type dihedrals with
    seeded coefficients in [-5, 5] kJ/mol, LJ sigma=3.9 A eps=0.5 kJ/mol,
    charges +-0.3 on `n_charged` sites (net zero)."""
    rng = np.random.default_rng(seed)
    geo = _chain_geometry(n_sites, rng, 1.53, 112.0, compact)
    charged = np.linspace(0, n_sites - 1, n_charged).round().astype(int)
    q = np.zeros(n_sites)
    for k, i in enumerate(charged):
        q[i] = 0.3 if k % 2 == 0 else -0.3
    sites = [Site("C%d" % (i + 1), tuple(geo[i]), 14.027, float(q[i]), 3.9, 0.5) for i in range(n_sites)]
    bonds = [(0, i + 1, i + 2, 1.53, 2500.0, 0.0, 0.0) for i in range(n_sites - 1)]
    angles = [(i + 1, i + 2, i + 3, 112.0, 500.0) for i in range(n_sites - 2)]
    coef = np.round(rng.uniform(-5.0, 5.0, size=(3,)), 6)   # as written to the .mmol file
    tors1 = [(i + 1, i + 2, i + 3, i + 4, float(coef[0]), float(coef[1]), float(coef[2]))
             for i in range(n_sites - 3)]
    return MolType("UACH", sites, bonds, angles, [], tors1, c14el=0.5, c14lj=0.5)


def lj_solute(n_sites: int, seed: int, bonded: bool = False) -> MolType:
    """A charged LJ solute of `n_sites` sites (lambda-coupled in
    expanded-ensemble jobs), alternating charges summing to zero.

    Default: a compact random cluster, sites 1.5 A apart, no bonded terms
    (force-evaluation fixtures).  With `bonded`: a branched tree for MD
    segments -- bonds of 1.5 A (constrained under SHAKE: n - 1 constraints
    on a tree, so the iteration converges in a few sweeps), a harmonic angle
    on every pair of bonds sharing a site (which also excludes the 1-3
    pairs), every other site pair at least 3.0 A apart."""
    rng = np.random.default_rng(seed)
    q = np.array([0.25 if i % 2 == 0 else -0.25 for i in range(n_sites)])
    if n_sites % 2 == 1:
        q[-1] = 0.0
    if not bonded:
        pts = [np.zeros(3)]
        while len(pts) < n_sites:
            base = pts[rng.integers(len(pts))]
            v = rng.normal(size=3)
            cand = base + 1.5 * v / np.linalg.norm(v)
            if min(np.linalg.norm(cand - p) for p in pts) > 1.4:
                pts.append(cand)
        sites = [Site("S%d" % (i + 1), tuple(pts[i]), 12.011, float(q[i]), 3.4, 0.4) for i in range(n_sites)]
        return MolType("SOLU", sites)
    pts = [np.zeros(3)]
    nbr = [[]]
    while len(pts) < n_sites:
        # grow from sites near the centre (compact tree)
        c = np.mean(pts, axis=0)
        pick = rng.integers(len(pts), size=4)
        p = int(min(pick, key=lambda k: np.linalg.norm(pts[k] - c)))
        if len(nbr[p]) >= 3:
            continue
        v = rng.normal(size=3)
        cand = pts[p] + 1.5 * v / np.linalg.norm(v)
        near = set(nbr[p]) | {p}
        ok = all(np.linalg.norm(cand - pts[k]) >= 2.3 for k in nbr[p])
        ok = ok and all(np.linalg.norm(cand - pts[k]) >= 3.0 for k in range(len(pts)) if k not in near)
        if ok:
            nbr[p].append(len(pts))
            nbr.append([p])
            pts.append(cand)
    sites = [Site("S%d" % (i + 1), tuple(pts[i]), 12.011, float(q[i]), 3.4, 0.4) for i in range(n_sites)]
    bonds = [(0, i + 1, j + 1, 1.5, 2000.0, 0.0, 0.0) for i in range(n_sites) for j in nbr[i] if i < j]
    angles = []
    for c in range(n_sites):
        for a in range(len(nbr[c])):
            for b in range(a + 1, len(nbr[c])):
                i, k = nbr[c][a], nbr[c][b]
                u, w = pts[i] - pts[c], pts[k] - pts[c]
                th = math.degrees(math.acos(float(np.dot(u, w) / (np.linalg.norm(u) * np.linalg.norm(w)))))
                angles.append((i + 1, c + 1, k + 1, round(th, 4), 200.0))
    return MolType("SOLU", sites, bonds, angles)


# ----------------------------------------------------------------------
# Placement
# ----------------------------------------------------------------------
def random_rotation(rng: np.random.Generator) -> np.ndarray:
    """Uniform random rotation from a unit quaternion."""
    q = rng.normal(size=4)
    q /= np.linalg.norm(q)
    a, b, c, d = q
    return np.array([[a*a+b*b-c*c-d*d, 2*(b*c-a*d), 2*(b*d+a*c)],
                     [2*(b*c+a*d), a*a-b*b+c*c-d*d, 2*(c*d-a*b)],
                     [2*(b*d-a*c), 2*(c*d+a*b), a*a-b*b-c*c+d*d]])


def lattice_sites(nmol: int, box: Sequence[float], rng: np.random.Generator,
                  jitter: float = 0.1) -> np.ndarray:
    """Molecule centres on a jittered simple-cubic lattice inside
    [-L/2, L/2)^3; `jitter` is a fraction of the lattice spacing."""
    n = int(math.ceil(nmol ** (1.0 / 3.0) - 1e-9))
    g = (np.arange(n) + 0.5) / n - 0.5
    grid = np.stack(np.meshgrid(g, g, g, indexing="ij"), -1).reshape(-1, 3)
    sel = rng.permutation(len(grid))[:nmol]
    c = grid[sel] * np.asarray(box)
    c += rng.uniform(-jitter, jitter, size=c.shape) * (np.asarray(box) / n)
    return c


def assemble(name: str, types: List[MolType], counts: List[int], centres: np.ndarray,
             rng: np.random.Generator, **params) -> System:
    """Atom coordinates from molecule centres (one centre per molecule, in
    type order) with a random orientation per molecule."""
    pos = []
    k = 0
    for t, c in zip(types, counts):
        loc = t.local_coords()
        for _ in range(c):
            r = random_rotation(rng) if t.nsites > 1 else np.eye(3)
            pos.append(loc @ r.T + centres[k])
            k += 1
    return System(name, types, counts, tuple(float(b) for b in params.pop("box")),
                  np.concatenate(pos).astype(np.float64), **params)


def water_nacl(n_water: int, n_ion: int, box: float, seed: int, name: str = "wnacl",
               water: Optional[MolType] = None, **params) -> System:
    """`n_water` waters plus `n_ion` Na+ and `n_ion` Cl- on a jittered
    lattice; the ions take the places of randomly chosen waters."""
    rng = np.random.default_rng(seed)
    water = water or spce_water()
    nmol = n_water + 2 * n_ion
    centres = lattice_sites(nmol, (box, box, box), rng)
    order = rng.permutation(nmol)
    centres = centres[order]      # ions are the last 2*n_ion random sites
    types = [water, sodium(), chloride()]
    counts = [n_water, n_ion, n_ion]
    if n_ion == 0:
        types, counts = [water], [n_water]
    return assemble(name, types, counts, centres, rng, box=(box, box, box), seed=seed, **params)


def cfg_small(seed: int = 2024) -> System:
    """508 SPC/E + 4 Na+ + 4 Cl- in a 25 A box, R_cut 10 A, alpha 0.28,
    |k| cut for the n^2 <= 49 sphere (B=9.95 -> 709 vectors)."""
    return water_nacl(508, 4, 25.0, seed, name="w508", rcut=10.0, rshort=5.0,
                      ewald_a=2.8, ewald_b=9.95, shake=True)


def cfg_flex_small(seed: int = 2030) -> System:
    """216 flexible SPC-FW waters + 2 Na+ + 2 Cl- in a 19 A box, R_cut 9 A,
    R_short 5 A, no constraints: the multiple-time-step integrator with a
    2 fs long step split into 4 short steps."""
    return water_nacl(216, 2, 19.0, seed, name="f216", water=spcfw_water(), rcut=9.0, rshort=5.0,
                      ewald_a=2.8, ewald_b=9.0, shake=False, nfreq=4, dt_fs=2.0)


def cfg_large(seed: int = 2025) -> System:
    """32,000 SPC/E + 87 Na+ + 87 Cl- (96,174 atoms) in a 98.7 A box,
    R_cut 12 A, skin 1 A, alpha*R_cut=3.2 and B=11.5 for ~1e-5 relative
    force error."""
    return water_nacl(32000, 87, 98.7, seed, name="w32k", rcut=12.0, rshort=5.0,
                      ewald_a=3.2, ewald_b=11.5, shake=True, skin=1.0)


def chain_mixture(n_chain: int, n_water: int, box: float, seed: int, name: str = "chains",
                  chain_sites: int = 46, water_factor: float = 1.6, chain_compact: Optional[float] = None,
                  **params) -> System:
    """Flexible chains solvated in SPC/E water.  Chains are placed first on a
    coarse lattice; waters fill lattice sites (water_factor x n_water
    candidates) farther than 2.6 A from any chain atom."""
    from scipy.spatial import cKDTree
    rng = np.random.default_rng(seed)
    chain = ua_chain(chain_sites, seed=seed + 1, compact=chain_compact)
    water = spce_water()
    L = np.array([box, box, box])
    cc = lattice_sites(n_chain, L, rng, jitter=0.05)
    loc = chain.local_coords()
    # chains farther apart than twice the chain radius plus the 3.2 A
    # clearance cannot change a placement test, so only nearer ones are
    # compared (same result as comparing against every placed chain)
    reach = 2.0 * np.sqrt((loc ** 2).sum(1)).max() + 3.3
    chain_atoms = []
    for k, c in enumerate(cc):
        # orientation with no chain atom closer than 3.2 A to those placed
        best, best_d = None, -1.0
        dc = cc[:k] - c
        dc -= L * np.round(dc / L)
        near = [chain_atoms[q] for q in np.nonzero((dc ** 2).sum(1) < reach ** 2)[0]]
        tree = cKDTree(np.mod(np.concatenate(near), box), boxsize=box) if near else None
        for _ in range(60):
            cand_atoms = loc @ random_rotation(rng).T + c
            if tree is None:
                best = cand_atoms
                break
            dmin = tree.query(np.mod(cand_atoms, box))[0].min()
            if dmin > best_d:
                best, best_d = cand_atoms, dmin
            if dmin > 3.2:
                break
        chain_atoms.append(best)
    ca = np.concatenate(chain_atoms) if chain_atoms else np.zeros((0, 3))
    # candidate water sites: dense lattice minus those near chains
    cand = lattice_sites(int(n_water * water_factor) + 64, L, rng, jitter=0.08)
    if len(ca):
        d = cKDTree(np.mod(ca, box), boxsize=box).query(np.mod(cand, box))[0]
        cand = cand[d > 2.6]
    if len(cand) < n_water:
        raise ValueError("box too small for %d waters next to %d chains" % (n_water, n_chain))
    wc = cand[rng.permutation(len(cand))[:n_water]]
    pos = [ca]
    wl = water.local_coords()
    for c in wc:
        pos.append(wl @ random_rotation(rng).T + c)
    return System(name, [chain, water], [n_chain, n_water], (box, box, box),
                  np.concatenate(pos), seed=seed, **params)


def cfg_chains(seed: int = 2026) -> System:
    """256 chains of 46 united atoms in 12,000 SPC/E waters, double time
    step 0.2/2 fs.  The chains are compact globules (radius ~7 A) and the
    box is 92 A (0.81 g/cm^3): at 80-86 A (0.9-1.0 g/cm^3) rigid random
    placement leaves chain sites overlapping, whose forces blow up the
    first steps."""
    return chain_mixture(256, 12000, 92.0, seed, name="ch256", water_factor=2.0, chain_compact=7.0,
                         rcut=12.0, rshort=5.0, ewald_a=3.2, ewald_b=11.5, shake=False, nfreq=10, dt_fs=2.0,
                         skin=1.0)


def pimd_beads(n_beads: int = 32, n_water: int = 4000, box: float = 49.3, seed: int = 2027,
               temperature: float = 300.0) -> List[System]:
    """`n_beads` replicas of one SPC-FW water box; bead atoms are spread
    around the centroid by a Gaussian of the thermal de Broglie width
    lambda = hbar / sqrt(m kT) (per atom mass)."""
    base = water_nacl(n_water, 0, box, seed, name="pimd", water=spcfw_water(), rcut=10.0,
                      rshort=5.0, ewald_a=3.0, ewald_b=9.0, shake=False, nfreq=10,
                      temperature=temperature, skin=1.0)
    rng = np.random.default_rng(seed + 1)
    hbar = 1.0545e-34
    masses = np.concatenate([[s.mass for s in t.sites] * c for t, c in zip(base.types, base.counts)])
    width = hbar / np.sqrt(masses * 1e-3 / AVSNO * BOLTZ * temperature) * 1e10   # A
    beads = []
    for b in range(n_beads):
        s = dataclasses.replace(base, name="bead%02d" % b,
                                pos=base.pos + rng.normal(size=base.pos.shape) * width[:, None] / math.sqrt(3))
        beads.append(s)
    return beads


def ee_job(n_atoms_target: int, n_solute: int, lam: float, seed: int, rigid_solute: bool = False) -> System:
    """Expanded-ensemble style job: SPC/E water + 0.15 M NaCl sized to about
    `n_atoms_target` atoms, plus one solute of `n_solute` sites coupled at
    `lam` (LJ and charges scaled by lam^PW1, soft-core shift BEE(1-lam^PW2)).
    `rigid_solute` gives the solute its bond network (for MD segments); the
    bonds are then flexible (harmonic) and only the waters are constrained."""
    rng = np.random.default_rng(seed)
    n_w = max(64, int((n_atoms_target - n_solute) / 3.0))
    n_ion = max(1, int(round(n_w * 0.15 / 55.5)))
    n_w -= 2 * n_ion
    nmol = n_w + 2 * n_ion + 1
    box = (nmol / 0.0334) ** (1.0 / 3.0)
    box = max(box, 2.0 * 10.0 + 1.0)
    sol = lj_solute(n_solute, seed + 11, bonded=rigid_solute)
    centres = lattice_sites(nmol, (box, box, box), rng)
    # the solute takes the site nearest the box centre; water and ion sites
    # closer than 3.4 A to any solute atom are dropped
    k0 = int(np.argmin((centres ** 2).sum(1)))
    sc = centres[k0]
    rot = random_rotation(rng)
    sol_atoms = sol.local_coords() @ rot.T + sc
    others = np.delete(centres, k0, axis=0)
    d = others[:, None, :] - sol_atoms[None, :, :]
    d -= box * np.round(d / box)
    far = np.sqrt((d ** 2).sum(-1)).min(1) > 3.4
    others = others[far]
    n_w = min(n_w, len(others) - 2 * n_ion)
    others = others[rng.permutation(len(others))[:n_w + 2 * n_ion]]
    centres = np.concatenate([others[:n_w], others[n_w:n_w + n_ion], others[n_w + n_ion:], sc[None]])
    types = [spce_water(), sodium(), chloride(), sol]
    counts = [n_w, n_ion, n_ion, 1]
    sysm = assemble("ee%d" % seed, types[:3], counts[:3], centres[:-1], rng, box=(box, box, box), seed=seed,
                    rcut=10.0, rshort=5.0, ewald_a=2.8, ewald_b=9.0, shake=True, skin=1.0)
    return dataclasses.replace(sysm, types=types, counts=counts, pos=np.concatenate([sysm.pos, sol_atoms]),
                               ee_type=3, ee_lambda=lam, shake_types=(0,) if rigid_solute else None)


def maxwell_velocities(system: System, seed: int, temperature: float | None = None) -> np.ndarray:
    """Maxwell-Boltzmann velocities (A/fs) at the system temperature with
    the total momentum removed."""
    t = system.temperature if temperature is None else temperature
    masses = np.concatenate([[s.mass for s in ty.sites] * c for ty, c in zip(system.types, system.counts)])
    rng = np.random.default_rng(seed)
    kt = AVSNO * BOLTZ * t * 1e-3                      # kJ/mol
    v = rng.normal(size=(len(masses), 3)) * np.sqrt(kt / masses * 1e-4)[:, None]
    v -= (masses[:, None] * v).sum(0) / masses.sum()
    return v
