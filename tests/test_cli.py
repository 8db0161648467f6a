"""The drop-in command line (gsofa_amd/bin/gsofa), on the CPU: the reference's
positional argv, usage errors, unreadable inputs, and on a machine without a
GPU a clean status after the host-side steps (reading, --order-nd)."""
import os
import subprocess

import pytest

from conftest import ROOT

CLI = os.path.join(ROOT, "gsofa_amd", "bin", "gsofa")


def _run(*args):
    return subprocess.run([CLI, *args], capture_output=True, text=True, timeout=120)


def _has_gpu():
    try:
        import torch
        return torch.cuda.is_available()
    except ImportError:
        return False


@pytest.mark.skipif(not os.path.exists(CLI), reason="CLI not built")
def test_cli_usage_and_bad_inputs(tmp_path):
    r = _run()
    assert r.returncode == 1 and "usage:" in r.stderr
    r = _run("a", "b", "c", "0", "1")  # N_GPU must be >= 1
    assert r.returncode == 1 and "usage:" in r.stderr
    r = _run(str(tmp_path / "x"), str(tmp_path / "y"), str(tmp_path / "z"), "1", "1")
    assert r.returncode == 2 and "reading input failed" in r.stderr


@pytest.mark.skipif(not os.path.exists(CLI), reason="CLI not built")
@pytest.mark.skipif(_has_gpu(), reason="checks the no-device path")
def test_cli_host_steps_without_device(bcsstk18, tmp_path):
    import gsofa_amd as g
    p = [str(tmp_path / s) for s in ("b.dat", "e.dat", "c.dat")]
    g.write_ascii(bcsstk18, *p)
    r = _run(*p, "2", "0.5", "--order-nd", "64")
    assert r.returncode == 3, (r.stdout, r.stderr)
    assert "Number of GPUs used: 2" in r.stdout
    assert "Reordered by nested dissection (leaf 64)" in r.stdout
    assert "gsofa:" in r.stderr
