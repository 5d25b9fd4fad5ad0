"""The whole CPU-runnable Taylor bar (spacing 0.5 mm, seed 0, 190 m/s,
80 us, 2446 steps) on the GPU against the reference program's own final
state (tests/golden/taylor_h05_final.npz, written by make_golden.py --final
with MEGA -j 1): final bar length and mushroom radius within 0.5 %, for
FP64, mixed precision, and mixed precision with the learned return map."""
import os

import numpy as np
import pytest

from taylor_amd import deck, mega

pytestmark = pytest.mark.gpu
GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "taylor_h05_final.npz")
WEIGHTS = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "taylor_amd", "data",
                       "j2_copper.npz")


def _measures(prob, x, eps, z):
    wall = prob.sets["WALL"] - 1
    L = x[:, 2].max()
    R = np.sqrt(x[wall, 0] ** 2 + x[wall, 1] ** 2).max()
    vol = prob.vol
    l2 = np.sqrt((vol * (eps - z["eps"]) ** 2).sum() / (vol * z["eps"] ** 2).sum())
    return L, R, l2


@pytest.mark.parametrize("mode", ["fp64", "mixed", "surrogate"])
def test_final_length_and_radius_vs_reference(mode):
    z = np.load(GOLD)
    prob = deck.taylor_problem(0.5e-3, seed=0)
    nsteps = int(z["nsteps"][0])
    assert nsteps == int(np.floor(80e-6 / prob.dt)) + 1
    if mode == "surrogate" and not os.path.exists(WEIGHTS):
        pytest.skip("no trained weights")
    g = mega.Mega(prob, precision=mega.MEGA_FP64 if mode == "fp64" else mega.MEGA_MIXED)
    if mode == "surrogate":
        from taylor_amd import surrogate
        surrogate.attach(g, WEIGHTS)
    g.step(nsteps)
    x, eps = g.download("x"), g.download("eps")
    st = g.surrogate_stats() if mode == "surrogate" else None
    g.close()
    wall = prob.sets["WALL"] - 1
    L0 = z["x"][:, 2].max()
    R0 = np.sqrt(z["x"][wall, 0] ** 2 + z["x"][wall, 1] ** 2).max()
    L, R, l2 = _measures(prob, x, eps, z)
    print(f"{mode}: length {L * 1e3:.4f} mm (reference {L0 * 1e3:.4f}, rel {abs(L - L0) / L0:.2e}), radius "
          f"{R * 1e3:.4f} mm (reference {R0 * 1e3:.4f}, rel {abs(R - R0) / R0:.2e}), eqps L2 rel {l2:.2e}"
          + (f", surrogate {st}" if st else ""))
    assert np.isfinite(x).all()
    assert abs(L - L0) <= 0.005 * L0
    assert abs(R - R0) <= 0.005 * R0
