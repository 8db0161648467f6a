"""The C ABI from plain C (what a cgo, JNI or ctypes binding sees): the
header compiles as strict C99, a C program links against libgsofa.so and
runs its host-side calls; without a GPU the device call reports
GSOFA_ERR_NO_DEVICE instead of failing later."""
import os
import shutil
import subprocess

import pytest

from conftest import ROOT

LIB = os.path.join(ROOT, "gsofa_amd", "libgsofa.so")

PROG = r"""
#include <stdio.h>
#include "gsofa.h"
int main(void) {
    gsofa_options o;
    gsofa_options_default(&o);
    int32_t colbeg[3] = {0, 2, 5}, colend[3] = {2, 5, 7}, csr[7] = {0, 1, 0, 1, 2, 1, 2};
    gsofa_matrix A = {3, 7, colbeg, colend, csr};
    int32_t perm[3];
    int rc = gsofa_order_nd(&A, 64, perm);
    gsofa_csr B;
    int rp = gsofa_permute(&A, perm, &B);
    if (rp == GSOFA_OK) gsofa_csr_free(&B);
    gsofa_result r;
    int rs = gsofa_symbfact(&A, &o, &r);
    if (rs == GSOFA_OK) gsofa_result_free(&r);
    printf("%d %d %d %d %d\n", o.n_virtual_gpu, o.virtual_gpu_id, rc, rp, rs);
    return 0;
}
"""


@pytest.mark.skipif(not os.path.exists(LIB) or shutil.which("gcc") is None, reason="library or gcc missing")
def test_c99_program_links_and_runs(tmp_path):
    src = tmp_path / "t.c"
    src.write_text(PROG)
    exe = tmp_path / "t"
    cc = subprocess.run(["gcc", "-std=c99", "-Wall", "-Wextra", "-pedantic", "-Werror",
                         "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe),
                         "-L", os.path.dirname(LIB), "-lgsofa", "-Wl,-rpath," + os.path.dirname(LIB)],
                        capture_output=True, text=True, timeout=120)
    assert cc.returncode == 0, cc.stderr
    out = subprocess.run([str(exe)], capture_output=True, text=True, timeout=120)
    assert out.returncode == 0, out.stderr
    nvg, vid, rc, rp, rs = map(int, out.stdout.split())
    assert (nvg, vid, rc, rp) == (1, -1, 0, 0)
    try:
        import torch
        has_gpu = torch.cuda.is_available()
    except ImportError:
        has_gpu = False
    assert rs == (0 if has_gpu else -6)
