"""CPU checks of the drop-in boundary: the HIP library loads and exports
every entry point include/mdx.h declares; the ctypes structs match the
header's layout."""
import ctypes as C
import os
import re
import shutil
import subprocess

import pytest

from mdynamix_amd import abi, engine

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    text = open(os.path.join(ROOT, "include", "mdx.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(mdx_[a-z_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    lib = engine.load_library()
    names = declared_functions()
    assert "mdx_create" in names and "mdx_compute" in names
    for name in names:
        assert hasattr(lib, name), name


def test_version_string():
    assert b"gfx950" in engine.load_library().mdx_version()


def test_struct_layout():
    # mdx_params: 3 box + rcut + rshort + skin + alpha + fexp + coulf (9 doubles), 2 ints, 2 doubles,
    # precision (int, padded to 8 bytes)
    assert C.sizeof(abi.MdxParams) == 9 * 8 + 2 * 4 + 2 * 8 + 8
    # energies: 17 doubles, 2 int64, 2*MAX_TYPEPAIRS + 3*MAX_MOLTYPES doubles
    assert C.sizeof(abi.MdxEnergies) == 8 * (17 + 2 + 2 * abi.MDX_MAX_TYPEPAIRS + 3 * abi.MDX_MAX_MOLTYPES + 2)


def _c_layout(struct, fields, tmp_path):
    """sizeof and field offsets of a header struct, from a host C compiler."""
    src = tmp_path / "layout.c"
    lines = ['#include <stdio.h>', '#include <stddef.h>', '#include "mdx.h"', "int main(void) {",
             '    printf("%%zu\\n", sizeof(%s));' % struct]
    lines += ['    printf("%%zu\\n", offsetof(%s, %s));' % (struct, f) for f in fields]
    lines += ["    return 0;", "}"]
    src.write_text("\n".join(lines) + "\n")
    exe = tmp_path / "layout"
    res = subprocess.run(["gcc", "-std=c99", "-I" + os.path.join(ROOT, "include"), str(src), "-o", str(exe)],
                         capture_output=True, text=True)
    assert res.returncode == 0, res.stderr
    out = subprocess.run([str(exe)], capture_output=True, text=True, check=True).stdout.split()
    return int(out[0]), [int(v) for v in out[1:]]


@pytest.mark.skipif(shutil.which("gcc") is None, reason="no host C compiler")
@pytest.mark.parametrize("ctype, cname", [(abi.MdxParams, "mdx_params"), (abi.MdxMdParams, "mdx_md_params")])
def test_ctypes_struct_matches_header(ctype, cname, tmp_path):
    """Every ctypes field sits at the header's offset (host compiler)."""
    fields = [f[0] for f in ctype._fields_]
    size, offs = _c_layout(cname, fields, tmp_path)
    assert size == C.sizeof(ctype)
    for f, off in zip(fields, offs):
        assert getattr(ctype, f).offset == off, f


def test_create_rejects_bad_input_without_device_work():
    """Argument validation happens before any device call."""
    lib = engine.load_library()
    h = C.c_void_p()
    assert lib.mdx_create(None, None, C.byref(h)) == abi.MDX_ERR_ARG
