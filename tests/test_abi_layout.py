"""The ctypes mirrors of the header's structs (gsofa_amd/_lib.py) have the C
layout: a C program prints sizeof and every field's offsetof, and each must
equal the ctypes Structure's.  A drift here would corrupt calls silently."""
import os
import shutil
import subprocess

import pytest

from conftest import ROOT

PAIRS = [("Matrix", "gsofa_matrix"), ("Csr", "gsofa_csr"), ("Options", "gsofa_options"),
         ("Result", "gsofa_result"), ("Supernodal", "gsofa_supernodal"),
         ("ServerOptions", "gsofa_server_options"), ("RequestInfo", "gsofa_request_info")]


@pytest.mark.skipif(shutil.which("gcc") is None, reason="gcc missing")
def test_struct_layouts_match_header(tmp_path):
    from gsofa_amd import _lib
    lines = ["#include <stdio.h>", "#include <stddef.h>", '#include "gsofa.h"', "int main(void) {"]
    for py, c in PAIRS:
        lines.append('    printf("%s sizeof %%zu\\n", sizeof(%s));' % (c, c))
        for name, _ in getattr(_lib, py)._fields_:
            lines.append('    printf("%s %s %%zu\\n", offsetof(%s, %s));' % (c, name, c, name))
    lines += ["    return 0;", "}"]
    src = tmp_path / "layout.c"
    src.write_text("\n".join(lines) + "\n")
    exe = tmp_path / "layout"
    cc = subprocess.run(["gcc", "-std=c99", "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe)],
                        capture_output=True, text=True, timeout=120)
    assert cc.returncode == 0, cc.stderr  # a ctypes field the header lacks fails here
    got = {}
    for line in subprocess.run([str(exe)], capture_output=True, text=True, timeout=60).stdout.splitlines():
        struct, field, value = line.split()
        got[(struct, field)] = int(value)
    for py, c in PAIRS:
        cls = getattr(_lib, py)
        import ctypes
        assert got[(c, "sizeof")] == ctypes.sizeof(cls), c
        for name, _ in cls._fields_:
            assert got[(c, name)] == getattr(cls, name).offset, (c, name)


def test_every_header_function_has_ctypes_argtypes():
    """Without argtypes, ctypes passes 64-bit pointers as C ints."""
    import re
    from gsofa_amd import _lib
    L = _lib.lib()
    hdr = open(os.path.join(ROOT, "include", "gsofa.h")).read()
    names = sorted(set(re.findall(r"\b(gsofa_[a-z0-9_]+)\s*\(", hdr)))
    assert names
    assert [n for n in names if getattr(L, n).argtypes is None] == []
