"""The C-ABI library loads and exports every symbol declared in include/pyib_hip.h (no GPU needed)."""
import ctypes
import os
import re
import shutil
import subprocess

import pytest

import pyib_hip  # noqa: F401
from pyib_hip import _native as N

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared():
    text = open(os.path.join(ROOT, "include", "pyib_hip.h")).read()
    return sorted(set(re.findall(r"\b(pyib_[a-z0-9_]+)\s*\(", text)))


def test_header_symbols_exported():
    lib = ctypes.CDLL(N.LIB_PATH)
    names = declared()
    assert len(names) >= 25
    for name in names:
        assert hasattr(lib, name), name
    assert set(names) == set(N.EXPORTED)


def test_struct_layout_matches_header():
    text = open(os.path.join(ROOT, "include", "pyib_hip.h")).read()
    body = text[text.index("typedef struct pyib_mesh {") + len("typedef struct pyib_mesh {"):text.index("} pyib_mesh;")]
    body = re.sub(r"/\*.*?\*/", "", body, flags=re.S)
    fields = re.findall(r"\*?\s*([a-z_0-9]+)(?:\s*,\s*\*?\s*([a-z_0-9]+))?(?:\s*,\s*([a-z_0-9]+))?(?:\s*,\s*([a-z_0-9]+))?;", body)
    names = [n for grp in fields for n in grp if n]
    assert names == [f[0] for f in N.Mesh._fields_]


def test_load_fails_loudly_when_missing(tmp_path):
    saved = N._lib
    try:
        N._lib = None
        try:
            N.load(str(tmp_path / "missing.so"))
            raised = False
        except RuntimeError:
            raised = True
        assert raised
    finally:
        N._lib = saved


@pytest.mark.skipif(shutil.which("gcc") is None, reason="needs gcc")
def test_struct_offsets_match_the_c_compiler(tmp_path):
    """Every ctypes mirror of a C-ABI struct has the C compiler's size and field offsets."""
    structs = {"pyib_mesh": N.Mesh, "pyib_work": N.Work, "pyib_cg": N.CG, "pyib_level": N.Level,
               "pyib_grid": N.Grid, "pyib_pgraph": N.PGraph}
    lines = ["#include <stdio.h>", "#include <stddef.h>", '#include "pyib_hip.h"', "int main(void) {"]
    for cname, cls in structs.items():
        lines.append(f'  printf("{cname} sizeof %zu\\n", sizeof({cname}));')
        for f in cls._fields_:
            lines.append(f'  printf("{cname} {f[0]} %zu\\n", offsetof({cname}, {f[0]}));')
    lines += ["  return 0;", "}"]
    src = tmp_path / "offsets.c"
    src.write_text("\n".join(lines) + "\n")
    exe = tmp_path / "offsets"
    subprocess.run(["gcc", "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe)], check=True)
    got = {}
    for line in subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout.splitlines():
        struct, field, value = line.split()
        got[(struct, field)] = int(value)
    for cname, cls in structs.items():
        assert got[(cname, "sizeof")] == ctypes.sizeof(cls), cname
        for f in cls._fields_:
            assert got[(cname, f[0])] == getattr(cls, f[0]).offset, (cname, f[0])
