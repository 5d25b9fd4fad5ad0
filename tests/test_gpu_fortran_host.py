"""The Fortran host (taylor_amd/host/mega_drive, bind(C) interfaces of
include/mega.h) drives the same device step as the Python host: identical
final positions and plastic strain on the same model."""
import os
import subprocess

import numpy as np
import pytest

from taylor_amd import deck, mega

pytestmark = pytest.mark.gpu
DRIVE = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "taylor_amd", "host", "mega_drive")


@pytest.mark.skipif(not os.path.exists(DRIVE), reason="Fortran host not built")
@pytest.mark.parametrize("precision", [mega.MEGA_FP64, mega.MEGA_MIXED])
def test_fortran_host_equals_python_host(tmp_path, precision):
    prob = deck.taylor_problem(0.7e-3, length=6e-3, seed=0)
    model, result = str(tmp_path / "model.bin"), str(tmp_path / "result.bin")
    mega.write_model(model, prob, nsteps=30, chunk=10, precision=precision)
    out = subprocess.run([DRIVE, model, result], capture_output=True, text=True, timeout=120)
    print(out.stdout)
    assert out.returncode == 0, out.stderr
    r = np.fromfile(result, dtype=np.float64)
    x_f, eps_f = r[:3 * prob.n].reshape(prob.n, 3), r[3 * prob.n:]
    g = mega.Mega(prob, precision=precision)
    for _ in range(3):
        g.step(10)
    x, eps = g.download("x"), g.download("eps")
    g.close()
    assert np.array_equal(x, x_f) and np.array_equal(eps, eps_f)
