"""The C ABI from plain C (what a cgo, JNI or ctypes binding sees): the
header compiles as strict C99, a C program links against libgsofa.so and
runs its host-side calls; without a GPU the device call reports
GSOFA_ERR_NO_DEVICE instead of failing later.  The program is
tests/dropin/c99_check.c (built by __graft_entry__.build(); its GPU run is
in test_dropin.py)."""
import os
import shutil
import subprocess

import pytest

from conftest import ROOT

LIB = os.path.join(ROOT, "gsofa_amd", "libgsofa.so")
SRC = os.path.join(ROOT, "tests", "dropin", "c99_check.c")


def _has_gpu():
    try:
        import torch
        return torch.cuda.is_available()
    except ImportError:
        return False


@pytest.mark.skipif(not os.path.exists(LIB) or shutil.which("gcc") is None, reason="library or gcc missing")
def test_c99_program_links_and_runs(tmp_path):
    exe = tmp_path / "t"
    cc = subprocess.run(["gcc", "-std=c99", "-Wall", "-Wextra", "-pedantic", "-Werror",
                         "-I", os.path.join(ROOT, "include"), SRC, "-o", str(exe),
                         "-L", os.path.dirname(LIB), "-lgsofa", "-Wl,-rpath," + os.path.dirname(LIB)],
                        capture_output=True, text=True, timeout=120)
    assert cc.returncode == 0, cc.stderr
    out = subprocess.run([str(exe)], capture_output=True, text=True, timeout=120)
    assert out.returncode == 0, out.stderr
    nvg, vid, rc, rp, rs, fills = map(int, out.stdout.split())
    assert (nvg, vid, rc, rp) == (1, -1, 0, 0)
    assert rs == (0 if _has_gpu() else -6)
    assert fills == 0  # the 3x3 tridiagonal pattern has no fill
