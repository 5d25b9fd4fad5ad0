"""Host checks of device math, compiled for the host from the same headers
the kernels use:
- the division-free box test of the neighbor search (mega::quotient_below_one,
  BIN_SEARCH.F90:268-289) against the division it replaces;
- the search's FP32 prefilter of that test (mega::box_rel, box_limits,
  box_axis): wherever it decides, it agrees with the exact test;
- the double-double power of the geometry preprocessing (dd::pow_dd, for
  VOL**THRD in PRE_MODEL_SMOOTHING_LENGTHS.F90:61): correctly rounded
  (against sqrt, x*x, 1/x) and within one ulp of the C library's pow."""
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


def _run_kat(tmp_path, name):
    exe = tmp_path / name
    subprocess.run([_hipcc(), "-O2", "-std=c++17", "-I", os.path.join(ROOT, "taylor_amd", "csrc"),
                    os.path.join(ROOT, "tests", name + ".cpp"), "-o", str(exe)],
                   check=True, capture_output=True)
    out = subprocess.run([str(exe)], capture_output=True, text=True)
    assert out.returncode == 0, out.stdout + out.stderr
    return out.stdout


@pytest.mark.skipif(_hipcc() is None, reason="hipcc not available")
def test_quotient_below_one_equals_division(tmp_path):
    assert "mismatches 0" in _run_kat(tmp_path, "kat_quotient")


@pytest.mark.skipif(_hipcc() is None, reason="hipcc not available")
def test_pow_dd_correctly_rounded(tmp_path):
    out = _run_kat(tmp_path, "kat_pow")
    assert "exact_mismatches 0" in out and "libm_beyond_one_ulp 0" in out


@pytest.mark.skipif(_hipcc() is None, reason="hipcc not available")
def test_box_prefilter_agrees_with_the_exact_test(tmp_path):
    out = _run_kat(tmp_path, "kat_boxfilter")
    print(out)
    assert "mismatches 0" in out
    words = out.split()
    random_checked, random_decided = int(words[5]), int(words[7])
    assert random_decided > 0.99 * random_checked   # it decides almost every pair away from the edge

