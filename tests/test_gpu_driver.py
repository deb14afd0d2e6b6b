"""The unchanged reference driver (main.f, mdstep.f SULTAN with SHAKE) run
with the ISO_C_BINDING shim in place of forces.f, i.e. on the HIP engine,
against the same program built with the reference forces.f, on the
1,532-atom water + NaCl box.  Both binaries are built from the reference
sources by oracle/ref/Makefile (targets mdx_md and drv)."""
import os
import subprocess
import tempfile

import numpy as np
import pytest

from mdynamix_amd import systems

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
MD = os.path.join(ROOT, "oracle", "_ref", "mdx_md")
DRV = os.path.join(ROOT, "oracle", "_ref", "mdx_drv")


def run_md(exe, system, nsteps):
    d = tempfile.mkdtemp(prefix="mdxdrv_")
    system.write_reference_inputs(d, nsteps=nsteps)
    res = subprocess.run([exe, "md.in"], cwd=d, capture_output=True, text=True, timeout=300)
    lines = res.stdout.splitlines()
    rows = []
    for k, ln in enumerate(lines):
        if ln.strip().startswith("STEP") and "POTE" in ln:
            for l2 in lines[k + 1:]:
                parts = l2.split()
                if len(parts) == 8 and parts[0].isdigit():
                    rows.append([float(p) for p in parts])
                elif parts and parts[0] == "main":
                    break
            break
    assert rows, res.stdout[-3000:] + res.stderr[-2000:]
    return np.array(rows)


@pytest.mark.skipif(not (os.path.exists(MD) and os.path.exists(DRV)),
                    reason="reference binaries not built (make -C oracle/ref all drv needs /root/reference)")
def test_reference_driver_on_gpu_engine_matches_cpu_reference():
    s = systems.cfg_small()
    print("driver: 1532 atoms, 10 SHAKE steps", flush=True)
    ref = run_md(MD, s, 10)
    gpu = run_md(DRV, s, 10)
    assert ref.shape == gpu.shape == (10, 8)
    # columns: step, POTE, PE(int), E(el), Etot, TEMP, PRES, Dens (printed
    # with 3 decimals per molecule; the first steps agree to print precision)
    for k in range(3):
        for col in (1, 3, 4, 5):
            assert abs(gpu[k, col] - ref[k, col]) <= 2e-3 + 2e-5 * abs(ref[k, col]), (k, col, gpu[k], ref[k])


@pytest.mark.skipif(not (os.path.exists(MD) and os.path.exists(DRV)),
                    reason="reference binaries not built (make -C oracle/ref all drv needs /root/reference)")
def test_reference_driver_on_gpu_engine_fspc_water():
    """Flexible fSPC water ("fSPC" keyword, IPOT 2): the shim hands the
    engine STRBND records built from IPOT/IADB (forces.f:555-605) and
    lists none of the type's bonds (setup.f:297); 182 atoms, no SHAKE."""
    s = systems.water_nacl(60, 1, 13.0, 17, name="fspc", water=systems.fspc_water(), rcut=6.0, rshort=3.5,
                           ewald_a=2.8, ewald_b=9.0, shake=False, nfreq=4)
    print("driver: 182 atoms fSPC water, 5 steps", flush=True)
    ref = run_md(MD, s, 5)
    gpu = run_md(DRV, s, 5)
    assert ref.shape == gpu.shape == (5, 8)
    assert abs(ref[0, 2]) > 1e-3, ref[0]      # intramolecular energy is not zero
    for k in range(3):
        for col in (1, 2, 3, 4, 5):
            assert abs(gpu[k, col] - ref[k, col]) <= 2e-3 + 2e-5 * abs(ref[k, col]), (k, col, gpu[k], ref[k])


def run_averages(exe, system, nsteps):
    """Per-template lines GETAVR prints with Average_int yes (aver.f:300-440):
    'name - name  <length> +/- <fluct>  <energy> +/- <fluct>' for bonds, the
    same with three or four names for angles (degrees) and dihedrals."""
    import dataclasses
    d = tempfile.mkdtemp(prefix="mdxavg_")
    dataclasses.replace(system, average_int=True).write_reference_inputs(d, nsteps=nsteps)
    res = subprocess.run([exe, "md.in"], cwd=d, capture_output=True, text=True, timeout=300)
    rows = []
    for ln in res.stdout.splitlines():
        if "+/-" in ln and ":" not in ln:
            head, _, _ = ln.partition("+/-")
            nums = ln.replace("+/-", " ").split()
            vals = []
            for tok in reversed(nums):
                try:
                    vals.append(float(tok))
                except ValueError:
                    break
            if len(vals) == 4:       # template lines (not the per-type temperatures)
                rows.append((head.split()[0], vals[::-1]))
    assert rows, res.stdout[-3000:] + res.stderr[-2000:]
    return rows


@pytest.mark.skipif(not (os.path.exists(MD) and os.path.exists(DRV)),
                    reason="reference binaries not built (make -C oracle/ref all drv needs /root/reference)")
@pytest.mark.parametrize("case", ["chains_respa", "water_shake"])
def test_reference_driver_bonded_averages(case):
    """Average_int yes (ISTAV, input.f:372-374): the shim fills BR/BE, AR/AE,
    TR/TE from mdx_get_terms on every fast-group call (FFORCES under the
    multiple time step, ALLFORCE under SHAKE), and GETAVR's per-template
    averages of bond lengths, angles, dihedrals and their energies printed
    by the driver on the engine equal the reference program's to print
    precision (4 decimals; one unit of rounding allowed) over 10 steps."""
    if case == "chains_respa":
        s = systems.chain_mixture(4, 60, 17.0, 5, name="chavg", chain_sites=12, rcut=7.0, rshort=4.0,
                                  ewald_a=2.8, ewald_b=9.0, shake=False)
    else:
        s = systems.cfg_small()
    print("driver averages:", case, s.natoms, flush=True)
    ref = run_averages(MD, s, 10)
    gpu = run_averages(DRV, s, 10)
    assert len(ref) == len(gpu) and len(ref) >= 6
    nonzero = 0
    for (nr, vr), (ng, vg) in zip(ref, gpu):
        assert nr == ng and len(vr) == len(vg) == 4, (nr, ng, vr, vg)
        for a, b in zip(vg, vr):
            assert abs(a - b) <= 2e-4 + 1e-6 * abs(b), (nr, vg, vr)
        nonzero += sum(1 for v in vr if abs(v) > 1e-3)
    assert nonzero > 0
