"""Host checks of device math that replaces a reference expression by an
exact equivalent: the division-free box test of the neighbor search
(mega::quotient_below_one, BIN_SEARCH.F90:268-289) against the division
itself, compiled for the host from the same header the kernels use."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _hipcc():
    for cand in ("/opt/rocm/bin/hipcc", shutil.which("hipcc") or ""):
        if cand and os.path.exists(cand):
            return cand
    return None


@pytest.mark.skipif(_hipcc() is None, reason="hipcc not available")
def test_quotient_below_one_equals_division(tmp_path):
    exe = tmp_path / "kat_quotient"
    subprocess.run([_hipcc(), "-O2", "-std=c++17", "-I", os.path.join(ROOT, "taylor_amd", "csrc"),
                    os.path.join(ROOT, "tests", "kat_quotient.cpp"), "-o", str(exe)],
                   check=True, capture_output=True)
    out = subprocess.run([str(exe)], capture_output=True, text=True)
    assert out.returncode == 0, out.stdout + out.stderr
    assert "mismatches 0" in out.stdout
