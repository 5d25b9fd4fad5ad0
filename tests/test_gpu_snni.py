"""Smoothed gradients (ITYPE_INT = 1, Lagrangian RK): the stabilized
non-conforming nodal integration whose smoothing the reference keeps
commented out (FINT.F90:441-496; box of PRE_MODEL_SMOOTHING_LENGTHS.F90:49-90).
An extension with no reference run behind it: the oracle's or_snni_grad is
pinned by the linear reproducing conditions (tests/test_oracle_kat.py), the
device follows the oracle, and a run past the step where direct nodal
integration fails shows what the smoothing is for."""
import numpy as np
import pytest

from oracle_bind import OracleRun, oracle_search
from test_oracle_kat import snni_grads
from taylor_amd import deck, mega

pytestmark = pytest.mark.gpu
H, LEN = 0.7e-3, 6e-3


def rel_err(a, b):
    a, b = np.asarray(a), np.asarray(b)
    return float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-300))


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_snni_gradients_match_oracle(oracle, precision):
    prob = deck.taylor_problem(H, length=LEN, seed=0, itype_int=1)
    g = mega.Mega(prob, precision=mega.MEGA_FP64 if precision == "fp64" else mega.MEGA_MIXED)
    g.search()
    g.shapes()
    _, dshp = g.get_shapes()
    g.close()
    rp, col, _ = oracle_search(oracle, prob)
    ref = np.concatenate(snni_grads(oracle, prob, rp, col))
    e = rel_err(dshp, ref)
    print(f"{precision}: smoothed gradients vs oracle {e:.2e}", flush=True)
    assert e < (1e-12 if precision == "fp64" else 2e-7)


@pytest.mark.parametrize("lagrangian,precision", [(1, "fp64"), (1, "mixed"), (0, "fp64"), (0, "mixed")])
def test_snni_steps_match_oracle(oracle, lagrangian, precision):
    """40 steps with the smoothed gradients against the oracle's run, Lagrangian
    and semi-Lagrangian (bins every 3 steps)."""
    prob = deck.taylor_problem(H, length=LEN, seed=0, itype_int=1, lagrangian=lagrangian, pdsearch_steps=3)
    g = mega.Mega(prob, precision=mega.MEGA_FP64 if precision == "fp64" else mega.MEGA_MIXED)
    g.step(40)
    run = OracleRun(oracle, prob).step(40)
    A = run.arr
    errs = {"x": rel_err(g.download("x"), A["x"]), "v": rel_err(g.download("v"), A["v"]),
            "fint": rel_err(g.download("fint"), A["fint"]), "stress": rel_err(g.download("stress"), A["stress"]),
            "eps": rel_err(g.download("eps"), A["state"][:, 2])}
    g.close()
    print(f"L{lagrangian} {precision}, 40 steps with smoothed gradients vs oracle: {errs}", flush=True)
    if precision == "fp64":
        assert errs["x"] < 1e-12 and errs["v"] < 1e-9 and errs["fint"] < 1e-9 and errs["stress"] < 1e-9
    else:
        assert errs["x"] < 1e-7 and errs["v"] < 1e-4 and errs["fint"] < 1e-4 and errs["stress"] < 1e-4


def test_snni_runs_past_the_direct_integration_failure():
    """The 181k-node bar (spacing 0.19 mm, jitter seed 1, 190 m/s), Lagrangian
    FP64: with direct nodal integration the reference program and the device
    both fail the return map in step 1,338; with the smoothed gradients the
    run stays finite and every node slower than 5 v0 through 3,000 steps."""
    prob = deck.taylor_problem(0.19e-3, seed=1, v0=190.0, itype_int=1)
    g = mega.Mega(prob, precision=mega.MEGA_FP64)
    wall = prob.sets["WALL"] - 1
    done = 0
    for chunk in (500,) * 6:
        st = g.step(chunk)
        done += chunk
        x, v = g.download("x"), g.download("vphy_step")
        vmax = float(np.sqrt((v * v).sum(1)).max())
        L = float(x[:, 2].max())
        R = float(np.sqrt(x[wall, 0] ** 2 + x[wall, 1] ** 2).max())
        print(f"step {done}: status {st}, length {L * 1e3:.4f} mm, radius {R * 1e3:.4f} mm, max speed {vmax:.1f} m/s",
              flush=True)
        assert st == 0 and np.isfinite(x).all() and vmax < 5 * 190.0
    g.close()
