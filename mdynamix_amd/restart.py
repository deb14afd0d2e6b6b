"""Restart (.dmp) and start (.start, XMOL) files of the reference program,
for job ingest and egress: a job can start on the engine from the state a
reference run dumped, and hand its state back as a restart file the
reference program continues from.

.dmp (RESTRT, restart.f:23-417): the reference writes it unformatted
(Fortran sequential records, each framed by its 4-byte length) or, with
"Dump_restart N ASCII", formatted; both hold, in this order,
  LABEL                                     (character*32)
  NOP, NTYPES, NSTOT, NNAV, DT               (DT in s)
  NSTEP, NSTTOT, NSTEPA, NAVT, NHIST, TIM, TIMA, NTRKF, NTRKZ, ME, NEOLD, I3, I4
  SC, SCL, BOXL, VOL, RDFCUT, BOYL, BOZL, SCLX, SCLY, SCLZ, SCM(NTYPES)
                                            (+ 10 reserved values, unformatted)
  11 accumulated CPU times
  (SX, SY, SZ) per atom                     (A)
  (VX, VY, VZ) per atom                     (internal momenta, see below)
then, for a normal run, (AV, AW) pairs of NNAV averages and NHIST records
of NNAV window averages; for an expanded-ensemble run (NEOLD = NE > 0) the
subensemble parameters, per-subensemble averages, history, the transition
matrix and the walk counters (restart.f:387-395, 343-353; kept as read,
unformatted only: the reference's own formatted writer stops there, as it
formats the REAL counters of restart.f:351-352 with I10); for a
path-integral bead (unformatted) the bead-thermostat chains
(restart.f:396-400).  read_dmp parses every record and write_dmp writes
them back in the reference's layout and formats (unformatted records
byte for byte; formatted lines with the reference's edit descriptors).

Velocities: the integrators advance S += TSTEP V MASSDI per time step DT
(mdstep.f:109-111), TSTEP = 1 in the reference's internal units
(setup.f:102), MASSDI = NBEADS M / m_i with M the total mass
(setup.f:149), so v [A/fs] = V M NBEADS / (m_i DT[fs]) -- velocities() and
momenta() convert.

.start (main.f:551-584): XMOL -- the atom count, a comment line that may
hold "BOX lx ly lz" among its first 80 characters, then one "name x y z"
line per atom.
"""
from __future__ import annotations

import dataclasses
import re
import struct
from typing import List, Optional, Sequence, Tuple

import numpy as np


@dataclasses.dataclass
class Restart:
    label: bytes
    nop: int
    ntypes: int
    nstot: int
    nnav: int
    dt: float                      # s
    counters: List[int]            # NSTEP, NSTTOT, NSTEPA, NAVT, NHIST
    tim: float
    tima: float
    tail: List[int]                # NTRKF, NTRKZ, ME, NEOLD, I3, I4
    box_rec: np.ndarray            # SC, SCL, BOXL, VOL, RDFCUT, BOYL, BOZL, SCLX, SCLY, SCLZ, SCM(NTYPES)
    reserved: np.ndarray           # 10 values after SCM (unformatted files only)
    times: np.ndarray              # 11 CPU times
    pos: np.ndarray                # (NSTOT, 3) A
    mom: np.ndarray                # (NSTOT, 3) internal momenta VX, VY, VZ
    av: np.ndarray = dataclasses.field(default_factory=lambda: np.zeros(0))
    aw: np.ndarray = dataclasses.field(default_factory=lambda: np.zeros(0))
    hist: np.ndarray = dataclasses.field(default_factory=lambda: np.zeros((0, 0)))
    # expanded ensemble (NEOLD = NE > 0): raw records, kept as read
    ee_records: Optional[list] = None
    # path-integral bead: JNH and the chains (JNH, NSTOT), + the 10 reserved
    thermo: Optional[np.ndarray] = None
    thermo_reserved: Optional[np.ndarray] = None
    formatted: bool = False

    @property
    def nhist(self) -> int:
        return int(self.counters[4])

    @property
    def box(self) -> np.ndarray:
        b = self.box_rec
        return np.array([b[2], b[5], b[6]])

    @property
    def nose_sc(self) -> float:
        """SC, the Nose-Hoover thermostat's variable, per time step DT
        (SCALING, mdstep.f:534-1093, works in steps)."""
        return float(self.box_rec[0])

    @property
    def nose_sc_per_fs(self) -> float:
        """SC per fs, as the engine's MD segments keep it (mdx_md_params.sc0)."""
        return self.nose_sc / (self.dt * 1e15)

    def velocities(self, masses: Sequence[float], nbeads: int = 1) -> np.ndarray:
        """(NSTOT, 3) velocities in A/fs from the internal momenta."""
        m = np.asarray(masses, dtype=np.float64)
        return self.mom * (nbeads * m.sum() / m)[:, None] / (self.dt * 1e15)


def momenta(vel: np.ndarray, masses: Sequence[float], dt_fs: float, nbeads: int = 1) -> np.ndarray:
    """Internal momenta V of the .dmp from velocities in A/fs."""
    m = np.asarray(masses, dtype=np.float64)
    return np.asarray(vel) * (m / (nbeads * m.sum()))[:, None] * dt_fs


# -- unformatted -----------------------------------------------------------
def _records(b: bytes) -> List[bytes]:
    out, o = [], 0
    while o < len(b):
        n = struct.unpack_from("<i", b, o)[0]
        if n < 0 or o + 8 + n > len(b) or struct.unpack_from("<i", b, o + 4 + n)[0] != n:
            raise ValueError("not a sequential unformatted file (record at byte %d)" % o)
        out.append(b[o + 4:o + 4 + n])
        o += 8 + n
    return out


def _frame(rec: bytes) -> bytes:
    return struct.pack("<i", len(rec)) + rec + struct.pack("<i", len(rec))


def _read_unformatted(b: bytes) -> Restart:
    recs = _records(b)
    label = recs[0]
    nop, ntypes, nstot, nnav = struct.unpack_from("<4i", recs[1])
    (dt,) = struct.unpack_from("<d", recs[1], 16)
    c = list(struct.unpack_from("<5i", recs[2]))
    tim, tima = struct.unpack_from("<2d", recs[2], 20)
    tail = list(struct.unpack_from("<6i", recs[2], 36))
    r3 = np.frombuffer(recs[3], "<f8")
    box_rec, reserved = r3[:10 + ntypes].copy(), r3[10 + ntypes:].copy()
    times = np.frombuffer(recs[4], "<f8").copy()
    pos = np.frombuffer(recs[5], "<f8").reshape(nstot, 3).copy()
    mom = np.frombuffer(recs[6], "<f8").reshape(nstot, 3).copy()
    r = Restart(label, nop, ntypes, nstot, nnav, dt, c, tim, tima, tail, box_rec, reserved, times, pos, mom)
    k = 7
    if tail[3] != 0:
        # expanded ensemble: every remaining record up to a bead block
        r.ee_records = recs[k:]
        return r
    avw = np.frombuffer(recs[k], "<f8").reshape(nnav, 2)
    r.av, r.aw = avw[:, 0].copy(), avw[:, 1].copy()
    k += 1
    r.hist = np.array([np.frombuffer(recs[k + j], "<f8") for j in range(c[4])]).reshape(c[4], nnav)
    k += c[4]
    if k < len(recs):
        (jnh,) = struct.unpack_from("<i", recs[k])
        th = np.frombuffer(recs[k + 1], "<f8")
        r.thermo = th[:jnh * nstot].reshape(jnh, nstot).copy()
        r.thermo_reserved = th[jnh * nstot:].copy()
        if recs[k + 2] != b"EOF":
            raise ValueError("bead block does not end with EOF")
    return r


def _write_unformatted(r: Restart) -> bytes:
    out = [_frame(r.label),
           _frame(struct.pack("<4id", r.nop, r.ntypes, r.nstot, r.nnav, r.dt)),
           _frame(struct.pack("<5i2d6i", *r.counters, r.tim, r.tima, *r.tail)),
           _frame(np.concatenate([r.box_rec, r.reserved]).astype("<f8").tobytes()),
           _frame(np.asarray(r.times, "<f8").tobytes()),
           _frame(np.asarray(r.pos, "<f8").tobytes()),
           _frame(np.asarray(r.mom, "<f8").tobytes())]
    if r.ee_records is not None:
        out += [_frame(x) for x in r.ee_records]
        return b"".join(out)
    out.append(_frame(np.stack([r.av, r.aw], 1).astype("<f8").tobytes()))
    out += [_frame(np.asarray(h, "<f8").tobytes()) for h in r.hist]
    if r.thermo is not None:
        out.append(_frame(struct.pack("<i", r.thermo.shape[0])))
        out.append(_frame(np.concatenate([r.thermo.ravel(), r.thermo_reserved]).astype("<f8").tobytes()))
        out.append(_frame(b"EOF"))
    return b"".join(out)


# -- formatted -------------------------------------------------------------
def fmt_d(v: float, w: int = 16, d: int = 9) -> str:
    """Fortran Dw.d output: 0.ddddddddd, exponent D+ee (or +eee without D
    when it needs three digits), right-justified in w characters."""
    v = float(v)
    if v == 0.0:
        s = "0." + "0" * d + "D+00"
    else:
        mant, ex = ("%.*e" % (d - 1, abs(v))).split("e")
        digits = mant.replace(".", "")
        e = int(ex) + 1
        s = "0." + digits + ("D%+03d" % e if abs(e) < 100 else "%+04d" % e)
        if v < 0:
            s = "-" + s
    return s.rjust(w)


def _lines_d(vals, per: int = 5) -> List[str]:
    vals = list(vals)
    if not vals:
        return [""]
    return ["".join(fmt_d(v) for v in vals[k:k + per]) for k in range(0, len(vals), per)]


_NUM = re.compile(r"[-+]?\d*\.?\d+(?:[DEde][-+]?\d+|[-+]\d{3})?")


def _parse_d(tok: str) -> float:
    t = tok.replace("D", "E").replace("d", "E")
    m = re.match(r"^([-+]?\d*\.\d+)([-+]\d{3})$", t)
    if m:
        t = m.group(1) + "E" + m.group(2)
    return float(t)


def _read_formatted(text: str) -> Restart:
    lines = text.split("\n")
    label = lines[0].encode()
    p = 1

    def ints(line, widths):
        out, o = [], 0
        for w in widths:
            out.append(int(line[o:o + w]))
            o += w
        return out, o

    (nop, ntypes, nstot, nnav), o = ints(lines[p], [9] * 4)
    dt = _parse_d(lines[p][o + 1:o + 17])
    p += 1
    c, o = ints(lines[p], [9] * 5)
    tim, tima = _parse_d(lines[p][o:o + 16]), _parse_d(lines[p][o + 16:o + 32])
    tail, _ = ints(lines[p][o + 32:], [9] * 6)
    p += 1

    def reals(count):
        nonlocal p
        vals = []
        nl = max(1, -(-count // 5))
        for k in range(nl):
            ln = lines[p + k]
            vals += [_parse_d(ln[q:q + 16]) for q in range(0, len(ln.rstrip()), 16)]
        p += nl
        return np.array(vals[:count])

    box_rec = reals(10 + ntypes)
    times = reals(11)
    pos = reals(3 * nstot).reshape(nstot, 3)
    mom = reals(3 * nstot).reshape(nstot, 3)
    r = Restart(label, nop, ntypes, nstot, nnav, dt, c, tim, tima, tail, box_rec, np.zeros(10), times, pos, mom,
                formatted=True)
    if tail[3] != 0:
        r.ee_records = lines[p:]
        return r
    avw = reals(2 * nnav).reshape(nnav, 2)
    r.av, r.aw = avw[:, 0].copy(), avw[:, 1].copy()
    r.hist = np.array([reals(nnav) for _ in range(c[4])]).reshape(c[4], nnav)
    rest = [ln for ln in lines[p:] if ln.strip()]
    if rest:
        jnh = int(lines[p])
        p += 1
        r.thermo = reals(jnh * nstot).reshape(jnh, nstot)
    return r


def _write_formatted(r: Restart) -> str:
    out = [r.label.decode(errors="replace"),
           "".join("%9d" % v for v in (r.nop, r.ntypes, r.nstot, r.nnav)) + " " + fmt_d(r.dt),
           "".join("%9d" % v for v in r.counters) + fmt_d(r.tim) + fmt_d(r.tima) +
           "".join("%9d" % v for v in r.tail)]
    out += _lines_d(r.box_rec)
    out += _lines_d(r.times)
    out += _lines_d(np.asarray(r.pos).ravel())
    out += _lines_d(np.asarray(r.mom).ravel())
    if r.ee_records is not None:
        out += list(r.ee_records)
        return "\n".join(out)
    out += _lines_d(np.stack([r.av, r.aw], 1).ravel())
    for h in r.hist:
        out += _lines_d(h)
    if r.thermo is not None:
        out.append("%6d" % r.thermo.shape[0])
        out += _lines_d(r.thermo.ravel())
    return "\n".join(out) + "\n"


def read_dmp(path: str) -> Restart:
    """A restart file of either form (detected from its first bytes)."""
    with open(path, "rb") as f:
        b = f.read()
    if len(b) >= 4 and struct.unpack_from("<i", b)[0] == 32 and b[4:6] != b"  ":
        try:
            return _read_unformatted(b)
        except (ValueError, struct.error):
            pass
    try:
        return _read_formatted(b.decode("latin-1"))
    except (IndexError, ValueError) as exc:
        raise ValueError("%s: not a complete restart file (%s)" % (path, exc)) from None


def write_dmp(path: str, r: Restart, formatted: Optional[bool] = None) -> None:
    fmt = r.formatted if formatted is None else formatted
    if fmt:
        with open(path, "w", encoding="latin-1", newline="\n") as f:
            f.write(_write_formatted(r))
    else:
        with open(path, "wb") as f:
            f.write(_write_unformatted(r))


def restart_from_state(pos: np.ndarray, vel: np.ndarray, masses: Sequence[float], box: Sequence[float],
                       dt_fs: float, ntypes: int, nmols: int, nose_sc: float = 0.0, nsttot: int = 0,
                       label: bytes = b"MDYN_v5.3.0".ljust(32)) -> Restart:
    """A restart record of a plain MD state (no accumulated averages):
    coordinates (A), velocities (A/fs), the thermostat's SC per step (the
    engine's per-fs value times DT[fs]) and the steps made so far; the
    reference reads it with Read_restart yes and continues from it."""
    n = len(pos)
    box = np.asarray(box, float)
    vol = float(box[0] * box[1] * box[2])
    box_rec = np.zeros(10 + ntypes)
    box_rec[0], box_rec[2], box_rec[3], box_rec[5], box_rec[6] = nose_sc, box[0], vol, box[1], box[2]
    dt = dt_fs * 1e-15
    return Restart(label=label, nop=nmols, ntypes=ntypes, nstot=n, nnav=0, dt=dt,
                   counters=[0, nsttot, nsttot, 0, 0], tim=nsttot * dt, tima=nsttot * dt,
                   tail=[0, 0, 1, 0, 0, 0], box_rec=box_rec, reserved=np.zeros(10), times=np.zeros(11),
                   pos=np.asarray(pos, float), mom=momenta(vel, masses, dt_fs))


# -- XMOL .start -----------------------------------------------------------
def read_start(path: str) -> Tuple[List[str], np.ndarray, Optional[np.ndarray]]:
    """(names, positions (n, 3) A, box or None) of an XMOL start file."""
    with open(path) as f:
        n = int(f.readline().split()[0])
        comment = f.readline()
        box = None
        k = comment[:80].find("BOX")
        if k >= 0:
            try:
                box = np.array([float(v) for v in comment[k + 4:128].split()[:3]])
            except ValueError:
                box = None
        names, pos = [], np.zeros((n, 3))
        for i in range(n):
            parts = f.readline().split()
            names.append(parts[0])
            pos[i] = [float(v) for v in parts[1:4]]
    return names, pos, box


def write_start(path: str, names: Sequence[str], pos: np.ndarray, box: Optional[Sequence[float]] = None,
                comment: str = "generated") -> None:
    with open(path, "w") as f:
        f.write("%d\n" % len(pos))
        f.write("%s  BOX %.12f %.12f %.12f\n" % ((comment,) + tuple(box)) if box is not None else comment + "\n")
        for nm, (x, y, z) in zip(names, pos):
            f.write("%s %.17g %.17g %.17g\n" % (nm, x, y, z))
