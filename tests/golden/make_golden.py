"""Generate the golden fixtures of the force path from the reference program.

Runs the reference MDynaMix (built by `make -C oracle/ref` from the sources
under /root/reference into oracle/_ref/mdx_ref) on seeded systems and
stores, per system, the inputs exactly as the reference held them
(parameters in its internal units, positions after its start-up) and its
outputs (neighbour lists or their digest, forces F/G/H, every energy and
virial scalar, per-type breakdowns).  Fixtures are data only.

    python tests/golden/make_golden.py        # rewrites tests/golden/*.npz
    python tests/golden/make_golden.py traj   # trajectory fixtures only
    python tests/golden/make_golden.py traj traj_spcfw_respa   # one of them
    python tests/golden/make_golden.py force fspc_tiny         # one force fixture
    python tests/golden/make_golden.py pimd   # path-integral fixtures (MPI reference)
"""
from __future__ import annotations

import dataclasses

import hashlib
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from mdynamix_amd import systems  # noqa: E402
from oracle import refdump  # noqa: E402

# name -> (builder, lambda, keep full lists)
CASES = {
    "water_tiny": (lambda: systems.water_nacl(20, 1, 12.0, 1, name="wtiny", rcut=5.5, rshort=3.0,
                                              ewald_a=2.8, ewald_b=9.0), 1.0, True),
    "water_508": (systems.cfg_small, 1.0, False),
    "chains_tiny": (lambda: systems.chain_mixture(4, 60, 17.0, 5, name="chtiny", chain_sites=12,
                                                  rcut=8.0, rshort=4.0, ewald_a=2.8, ewald_b=9.0,
                                                  shake=False, nfreq=10), 1.0, True),
    "ee_tiny": (lambda: systems.ee_job(400, 10, 0.35, 9), 0.35, True),
    "fspc_tiny": (lambda: _jitter(systems.water_nacl(60, 1, 13.0, 17, name="fspc", water=systems.fspc_water(),
                                                     rcut=6.0, rshort=3.5, ewald_a=2.8, ewald_b=9.0,
                                                     shake=False, nfreq=4), 0.04, 23), 1.0, True),
}

def _jitter(s, sigma, seed):
    """Atoms displaced by N(0, sigma) A: bonds away from equilibrium, so
    the intramolecular terms are not zero."""
    rng = np.random.default_rng(seed)
    return dataclasses.replace(s, pos=s.pos + rng.normal(0.0, sigma, s.pos.shape))


STATIC = ["SIZES", "SCALARS", "ITYPE", "NNUM", "NSITE", "INBTA", "CHARGEA", "Q", "MASSDA", "MASSA",
          "IEE", "LMOVE", "C14EL", "C14LJ", "IPOT", "IADB", "NSPEC", "NSITS", "SPEEI", "MDX", "A6LJ", "B12LJ",
          "A6LJ0", "B12LJ0", "A6_14", "B12_14", "NNBB", "INBB", "RKX", "RKY", "RKZ", "IBI", "JBJ",
          "IBUK", "ID", "RB", "FB", "DB", "RO", "IAI", "JAJ", "KAK", "IAUK", "RA", "FA", "ITI", "JTJ",
          "KTK", "LTL", "ITUK", "ITORS", "NMUL", "FT", "RT", "FT1", "FT2", "FT3", "FT4", "FT5",
          "LISTLEN"]
PERCALL = ["SX", "SY", "SZ", "XCOM", "YCOM", "ZCOM", "FX", "FY", "FZ", "GX", "GY", "GZ", "HX", "HY",
           "HZ", "ENERGY", "POTE1", "POTL1", "PES141", "PSR141", "SPEE", "BR", "BE", "AR", "AE", "TR", "TE",
           "EEDIFF", "EEPAR", "EC"]


def list_digest(a: np.ndarray, b: np.ndarray) -> str:
    return hashlib.sha256(np.ascontiguousarray(a, "<i4").tobytes() +
                          np.ascontiguousarray(b, "<i4").tobytes()).hexdigest()


# trajectory fixtures: name -> (builder, step counts); the reference
# integrator (oracle/ref/mdx_traj.f around SULTAN with constraints, DOUBLE
# without) from the reference's own start state
TRAJ_CASES = {
    "traj_water_508": (systems.cfg_small, (1, 20)),
    "traj_spcfw_respa": (systems.cfg_flex_small, (1, 20)),
}


# path-integral fixtures: name -> (builder, long-step counts); the
# reference PIMD (oracle/ref/mdx_pimd.f, one MPI rank per bead) from the
# reference's own start state (its per-bead position jitter included).
# PI_thermo chains of one link: the reference's THERMO array has one
# column (INHM = 1, dimpar.h:61), longer chains overrun it.
def _pimd_case():
    s = systems.cfg_flex_small()
    return dataclasses.replace(s, dt_fs=1.0, nfreq=4, rebuild_every=5, pimd_beads=4,
                               pi_thermo=(300.0, 5.0, 1))


PIMD_CASES = {
    "pimd_spcfw_4": (_pimd_case, (1, 10)),
}


def main_pimd(names=None) -> None:
    for name, (builder, counts) in PIMD_CASES.items():
        if names and name not in names:
            continue
        s = builder()
        out = {}
        for n in counts:
            beads = refdump.run_pimd(s, n)
            if "PSIZES" not in out:
                out["PSIZES"] = beads[0]["PSIZES"][0]
                out["PSCAL"] = beads[0]["PSCAL"][0]
                out["MASSDI"] = beads[0]["MASSDI"][0]
            for b, d in enumerate(beads):
                for k in ("SX", "SY", "SZ", "VX", "VY", "VZ"):
                    key0 = "%s_b%d_0" % (k, b)
                    if key0 not in out:
                        out[key0] = d[k][0]
                    else:
                        assert np.array_equal(out[key0], d[k][0]), "start state differs between runs"
                    out["%s_b%d_%d" % (k, b, n)] = d[k][1]
                out["THERMO_b%d_%d" % (b, n)] = np.stack(d["THERMO"])
                out["PEND_b%d_%d" % (b, n)] = d["PEND"][0]
        out["COUNTS"] = np.array(counts)
        path = os.path.join(HERE, name + ".npz")
        np.savez_compressed(path, **out)
        print("%-16s atoms=%d beads=%d steps=%s -> %s (%d KB)" % (name, s.natoms, s.pimd_beads, counts, path,
                                                               os.path.getsize(path) // 1024))


def parse_steps(stdout: str) -> np.ndarray:
    """Per-step lines of SULTAN (mdstep.f:187-190) or DOUBLE
    (mdstep.f:489-490): MSTEP, POTEN, PENS, PEL, ETOT, TEMP, TRYCK(M),
    DENS."""
    rows = []
    for line in stdout.splitlines():
        f = line.split()
        if len(f) == 8:
            try:
                rows.append([float(x) for x in f])
            except ValueError:
                pass
    return np.array(rows)


def main_traj(names=None) -> None:
    for name, (builder, counts) in TRAJ_CASES.items():
        if names and name not in names:
            continue
        s = builder()
        out = {}
        for n in counts:
            d = refdump.run_trajectory(s, n)
            if "TSIZES" not in out:
                out["TSIZES"] = d["TSIZES"][0]
                out["TSCAL"] = d["TSCAL"][0]
                out["MASSDI"] = d["MASSDI"][0]
                for k in ("SX", "SY", "SZ", "VX", "VY", "VZ"):
                    out[k + "_0"] = d[k][0]
            else:
                for k in ("SX", "VX"):
                    assert np.array_equal(out[k + "_0"], d[k][0]), "start state differs between runs"
            for k in ("SX", "SY", "SZ", "VX", "VY", "VZ"):
                out["%s_%d" % (k, n)] = d[k][1]
            out["TEND_%d" % n] = d["TEND"][0]
            out["STEPS_%d" % n] = parse_steps(str(d["_stdout"][0]))
        out["COUNTS"] = np.array(counts)
        path = os.path.join(HERE, name + ".npz")
        np.savez_compressed(path, **out)
        print("%-16s atoms=%d steps=%s -> %s (%d KB)" % (name, s.natoms, counts, path, os.path.getsize(path) // 1024))


def main() -> None:
    if sys.argv[1:2] == ["traj"]:
        main_traj(sys.argv[2:])
        return
    if sys.argv[1:2] == ["pimd"]:
        main_pimd(sys.argv[2:])
        return
    only = None
    if sys.argv[1:2] == ["force"]:
        only = set(sys.argv[2:])
    else:
        main_traj()
    for name, (builder, lam, full_lists) in CASES.items():
        if only and name not in only:
            continue
        s = builder()
        d = refdump.run_reference(s)
        out = {"lam": np.array(lam)}
        for k in STATIC:
            if k in d:
                out[k] = d[k][0]
        ncall = len(d["SX"])
        out["NCALL"] = np.array(ncall)
        for c in range(ncall):
            for k in PERCALL:
                if k in d:
                    out["%s_%d" % (k, c)] = d[k][c]
        if full_lists:
            for k in ("NBS1", "NBS2", "NBL1", "NBL2"):
                out[k] = d[k][0]
        out["DIGEST_SHORT"] = np.array(list_digest(d["NBS1"][0], d["NBS2"][0]))
        out["DIGEST_LONG"] = np.array(list_digest(d["NBL1"][0], d["NBL2"][0]))
        path = os.path.join(HERE, name + ".npz")
        np.savez_compressed(path, **out)
        print("%-12s atoms=%d calls=%d lists=%s -> %s (%d KB)" % (
            name, s.natoms, ncall, tuple(d["LISTLEN"][0]), path, os.path.getsize(path) // 1024))


if __name__ == "__main__":
    main()
