"""The drop-in surface end to end: the reference-side binding of
symbfact_min_id (tests/dropin/symbfact_min_id.cpp, the code INTEGRATION.md
shows) with a caller shaped like the reference's main.cu, the strict-C99
program, and the command line, on BCSSTK18 written in the reference's ASCII
format.  The reference logs 980,136 fills for it
(with_space_optimization/space_optimization_log_1_264_GPU_warp_centric.dat:2,
printed by with_space_optimization/symbfact.cuh:2526)."""
import os
import re
import subprocess

import pytest

from conftest import ROOT

DROPIN = os.path.join(ROOT, "tests", "dropin", "dropin")
C99 = os.path.join(ROOT, "tests", "dropin", "c99_check")
CLI = os.path.join(ROOT, "gsofa_amd", "bin", "gsofa")
SHIM = os.path.join(ROOT, "tests", "dropin", "symbfact_min_id.cpp")
BCSSTK18_FILLS = 980136


def _has_gpu():
    try:
        import torch
        return torch.cuda.is_available()
    except ImportError:
        return False


def _ascii(A, d):
    import gsofa_amd as g
    p = [str(d / s) for s in ("colbeg.dat", "colend.dat", "csr.dat")]
    g.write_ascii(A, *p)
    return p


def _run(*args):
    print("run:", " ".join(map(str, args)), flush=True)
    return subprocess.run([*args], capture_output=True, text=True, timeout=300)


def test_integration_doc_shows_the_compiled_binding():
    """INTEGRATION.md's drop-in code block is the file that is compiled and run."""
    doc = open(os.path.join(ROOT, "INTEGRATION.md")).read()
    blocks = re.findall(r"```c\+\+\n(.*?)```", doc, re.S)
    shim = open(SHIM).read()
    assert any(b.strip() == shim.strip() for b in blocks)


@pytest.mark.skipif(not os.path.exists(DROPIN), reason="drop-in binding not built")
@pytest.mark.skipif(_has_gpu(), reason="checks the no-device path")
def test_dropin_without_device(bcsstk18, tmp_path):
    r = _run(DROPIN, *_ascii(bcsstk18, tmp_path), "1", "0")
    assert r.returncode == 3, (r.stdout, r.stderr)
    assert "Number of GPUs used: 1" in r.stdout
    assert "caller: fill_count=0 time_ms=0.000" in r.stdout
    assert "gsofa: no usable GPU" in r.stderr


@pytest.mark.gpu
@pytest.mark.parametrize("n_gpu", ["1", "4"])
def test_dropin_binding_bcsstk18(bcsstk18, tmp_path, n_gpu):
    r = _run(DROPIN, *_ascii(bcsstk18, tmp_path), n_gpu, "0")
    assert r.returncode == 0, (r.stdout, r.stderr)
    assert "Total Number of fill-ins detected: %d" % BCSSTK18_FILLS in r.stdout
    m = re.search(r"caller: fill_count=(\d+) time_ms=([0-9.]+)", r.stdout)
    assert m and int(m.group(1)) == BCSSTK18_FILLS
    assert 0.0 < float(m.group(2)) < 60000.0  # milliseconds, as the reference's out-parameter


@pytest.mark.gpu
def test_c99_program_on_device(bcsstk18, tmp_path):
    r = _run(C99, *_ascii(bcsstk18, tmp_path))
    assert r.returncode == 0, (r.stdout, r.stderr)
    nvg, vid, rc, rp, rs, fills = map(int, r.stdout.split())
    assert (nvg, vid, rc, rp, rs) == (1, -1, 0, 0, 0)
    assert fills == BCSSTK18_FILLS


@pytest.mark.gpu
@pytest.mark.parametrize("extra", [[], ["--spill"], ["--counts-only"], ["--chunk", "64"]])
def test_cli_on_device(bcsstk18, tmp_path, extra):
    r = _run(CLI, *_ascii(bcsstk18, tmp_path), "6", "1", "32", *extra)
    assert r.returncode == 0, (r.stdout, r.stderr)
    assert "Number of GPUs used: 6" in r.stdout
    assert "Total Number of fill-ins detected: %d" % BCSSTK18_FILLS in r.stdout
    slow = float(re.search(r"Slowest virtual GPU takes: ([0-9.e+-]+) ms", r.stdout).group(1))
    fast = float(re.search(r"Fastest virtual GPU takes: ([0-9.e+-]+) ms", r.stdout).group(1))
    assert 0.0 <= fast <= slow
    if "--counts-only" not in extra:
        assert re.search(r"Supernodes: \d+", r.stdout)


@pytest.mark.gpu
def test_cli_share_and_order_on_device(bcsstk18, tmp_path):
    """One share of six (--share) counts that share's rows; --order-nd factors
    P A P^T, whose count the oracle gives for the same permutation."""
    import gsofa_amd as g
    from oracle import oracle
    p = _ascii(bcsstk18, tmp_path)
    shares = 0
    for sh in range(6):
        r = _run(CLI, *p, "6", "1", "--share", str(sh))
        assert r.returncode == 0, (r.stdout, r.stderr)
        shares += int(re.search(r"Total Number of fill-ins detected: (\d+)", r.stdout).group(1))
    assert shares == BCSSTK18_FILLS
    r = _run(CLI, *p, "1", "1", "--order-nd", "64")
    assert r.returncode == 0, (r.stdout, r.stderr)
    B = g.permute(bcsstk18, g.order_nd(bcsstk18, 64))
    want = oracle.symbfact(B.colbeg, B.colend, B.csr, B.n).fill_count
    assert "Total Number of fill-ins detected: %d" % want in r.stdout
