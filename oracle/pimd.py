"""Host restatement of the reference's path-integral MD (test infrastructure
only: used by tests/ as the checker of the device PIMD loop, never by the
product path).

It follows pimd.f's PIMD (pimd.f:18-314) with the forces of the FP64 oracle
(oracle/mdx_oracle.c through orc.compute), bead by bead:
  - bead masses m / P for the dynamics and kinetic energies (MASSD,
    setup.f:149), physical forces / P (pimd.f:61-69, 155-166),
  - PIFORCES springs BEADFAC m |r_k - r_k+-1|^2 to both neighbours, whose
    coordinates arrive as REAL*4 (pimd.f:321-394, 443-486),
  - per long step: half kick with G, NFREQ short steps (half kick with H,
    BEADNOSE, drift, GETCOM, CHCNB in short step NFREQ/2+1 of every
    ICHNB-th long step, fast forces -- slow and fast at the last --, half
    kick with H), half kick with G (pimd.f:92-204); PI_SCALING is inactive
    under PI_thermo (input.f:660-664),
  - BEADNOSE: per atom the kinetic energy summed over the beads, the
    temperature FEK CONVEQ and one Nose-Hoover chain shared by the beads
    (pimd.f:402-441),
  - GETCOM: each bead's molecules shifted by the image of the centroid's
    centre of mass (service.f:463-563 with PI_COM, pimd.f:498-520).
Units: A, fs, kJ/mol, g/mol (velocities in A/fs).
"""
from __future__ import annotations

from typing import Dict, List

import numpy as np

from oracle import orc

ACCEL = 1e-4        # kJ/mol/A / (g/mol) -> A/fs^2
EKIN = 1e4          # (g/mol) (A/fs)^2 -> kJ/mol


def _pbc_one_shift(d: np.ndarray, box: np.ndarray) -> np.ndarray:
    """if |d| > L/2: d -= sign(L, d), one shift per axis (PBC, service.f:953-956)."""
    return d - np.where(np.abs(d) > 0.5 * box, np.sign(d) * box, 0.0)


class PimdRef:
    """P beads of one system advanced by the reference PIMD scheme."""

    def __init__(self, topo, params: Dict, box, pos: List[np.ndarray], vel: List[np.ndarray], nfreq: int,
                 dt: float, rebuild_every: int, beadfac: float, temperature: float, tau_bead: float,
                 nchain: int = 1, step0: int = 0, boltz: float = 0.008314471):
        from mdynamix_amd import topology
        self.topo = topo
        self.ts = topo.struct()
        self.ps = topology.params_struct(dict(params, list_mode=0, skin=0.0))
        self.kv = orc.inifur(self.ts, self.ps, topo.nmoltypes)
        self.box = np.asarray(box, dtype=np.float64)
        self.P = len(pos)
        self.X = [np.array(p, dtype=np.float64) for p in pos]
        self.V = [np.array(v, dtype=np.float64) for v in vel]
        self.m = np.asarray(topo.atom_mass, dtype=np.float64)
        self.mol = np.asarray(topo.atom_mol)
        self.nfreq, self.dt, self.ichnb = nfreq, dt, rebuild_every
        self.beadfac, self.t0, self.tau, self.nchain = beadfac, temperature, tau_bead, max(1, nchain)
        self.boltz = boltz
        self.mstep = step0
        self.xi = np.zeros((self.nchain, len(self.m)))
        self.lists = [orc.chcnb(self.ts, self.ps, x) for x in self.X]
        self.G = [None] * self.P
        self.H = [None] * self.P
        self._forces(0)

    # ------------------------------------------------------------------
    def _springs(self, b: int) -> np.ndarray:
        f = np.zeros_like(self.X[b])
        for nb in ((b + 1) % self.P, (b - 1) % self.P):
            o = self.X[nb].astype(np.float32).astype(np.float64)     # REAL*4 transfer
            d = _pbc_one_shift(self.X[b] - o, self.box)
            bb = np.sqrt((d * d).sum(1))
            dbnd = bb * self.beadfac * self.m
            f += d * (-2.0 * dbnd / bb)[:, None]
        return f

    def _forces(self, group: int) -> None:
        """FFORCES (group 1) or SLFORCES + FFORCES (group 0), scaled by 1 / P,
        then the springs added to the fast forces."""
        for b in range(self.P):
            _, g, h, _ = orc.compute(self.topo, self.ts, self.ps, self.X[b], self.lists[b], group=group,
                                     kvec=self.kv)
            if group != 1:
                self.G[b] = g / self.P
            self.H[b] = h / self.P
        for b in range(self.P):
            self.H[b] = self.H[b] + self._springs(b)

    def _kick(self, F: List[np.ndarray], dt: float) -> None:
        # bead mass m / P
        for b in range(self.P):
            self.V[b] = self.V[b] + 0.5 * dt * ACCEL * F[b] * self.P / self.m[:, None]

    def _beadnose(self, dti: float) -> None:
        s = sum((v * v).sum(1) for v in self.V)
        # FEK CONVEQ with bead mass m / P: (2/3) sum_b 1/2 (m/P) v^2 / (k_B P)
        atemp = EKIN * self.m * s / (3.0 * self.boltz * self.P * self.P)
        inv_tau2 = 1.0 / (self.tau * self.tau)
        fnh = inv_tau2 / (3.0 * self.P)
        xi = self.xi
        xi[0] = xi[0] + (atemp / self.t0 - 1.0) * inv_tau2 * dti
        for j in range(1, self.nchain):
            xi[j] = xi[j] + (xi[j - 1] ** 2 - fnh) * dti
            xi[j - 1] = xi[j - 1] * (1.0 - xi[j] * dti)
        scv = 1.0 - xi[0] * dti
        for b in range(self.P):
            self.V[b] = self.V[b] * scv[:, None]

    def _getcom(self) -> None:
        nm = int(self.mol.max()) + 1
        w = np.zeros(nm)
        np.add.at(w, self.mol, self.m)
        coms = []
        for b in range(self.P):
            c = np.zeros((nm, 3))
            for k in range(3):
                np.add.at(c[:, k], self.mol, self.m * self.X[b][:, k])
            coms.append(c / w[:, None])
        cen = sum(coms) / self.P
        sh = _pbc_one_shift(cen, self.box) - cen
        for b in range(self.P):
            c = coms[b]
            self.X[b] = (self.X[b] - c[self.mol]) + (c + sh)[self.mol]

    def long_step(self) -> None:
        nf2 = self.nfreq // 2 + 1
        dti = self.dt / self.nfreq
        self.mstep += 1
        self._kick(self.G, self.dt)
        for k in range(1, self.nfreq + 1):
            self._kick(self.H, dti)
            if self.tau > 0.0:
                self._beadnose(dti)
            for b in range(self.P):
                self.X[b] = self.V[b] * dti + self.X[b]
            self._getcom()
            if k == nf2 and self.ichnb > 0 and self.mstep % self.ichnb == 0:
                self.lists = [orc.chcnb(self.ts, self.ps, x) for x in self.X]
            self._forces(1 if k < self.nfreq else 0)
            self._kick(self.H, dti)
        self._kick(self.G, self.dt)
