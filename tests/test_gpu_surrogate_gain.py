"""What the trained networks gain over the method without them, at an impact
speed they were not trained on (220 m/s; training runs at 150, 190 and
250 m/s): share of plastic return maps answered, node-kernel time with and
without the network, and the error of the run against the FP64 run of the
exact method (bar length within 0.5 %).  Two materials: the benchmark's
copper with linear hardening (taylor_amd/data/j2_copper.npz), where the exact
return map is one closed-form Newton step, and the reference deck's own
aluminium with power-law hardening (j2_powerlaw.npz), where VON_MISES iterates
Newton with a power per iteration (CONSTITUTE_VONMISES.F90:94-119)."""
import os

import numpy as np
import pytest

from taylor_amd import deck, mega, surrogate

pytestmark = pytest.mark.gpu
WEIGHTS = surrogate.DEFAULT_WEIGHTS


def _run(prob, mode, steps, weights=WEIGHTS):
    g = mega.Mega(prob, precision=mega.MEGA_FP64 if mode == "fp64" else mega.MEGA_MIXED)
    if mode == "surrogate":
        surrogate.attach(g, weights)
    g.step(1)
    g.profile(True)
    g.step(steps - 1)
    kt = g.kernel_times()
    x, eps = g.download("x"), g.download("eps")
    st = g.surrogate_stats() if mode == "surrogate" else None
    g.close()
    wall = prob.sets["WALL"] - 1
    L = x[:, 2].max()
    R = np.sqrt(x[wall, 0] ** 2 + x[wall, 1] ** 2).max()
    return {"L": L, "R": R, "eps": eps, "k_node_ms": kt["k_node"][0], "stats": st,
            "step_ms": sum(v[0] * v[1] for v in kt.values()) / (steps - 1)}


@pytest.mark.parametrize("material", ["copper", "powerlaw"])
def test_surrogate_gain_at_220_mps(material):
    weights = WEIGHTS if material == "copper" else surrogate.POWERLAW_WEIGHTS
    if not os.path.exists(weights):
        pytest.skip("no trained weights")
    # 181k-node bar, 600 steps (inside the window the formulation stays finite)
    prob = deck.taylor_problem(0.19e-3, v0=220.0, seed=1, material=surrogate.MATERIALS[material]())
    res = {m: _run(prob, m, 600, weights) for m in ("fp64", "mixed", "surrogate")}
    ref = res["fp64"]
    vol = prob.vol
    for m in ("mixed", "surrogate"):
        r = res[m]
        r["L_err"] = abs(r["L"] - ref["L"]) / ref["L"]
        r["R_err"] = abs(r["R"] - ref["R"]) / ref["R"]
        r["eps_l2"] = float(np.sqrt((vol * (r["eps"] - ref["eps"]) ** 2).sum() / (vol * ref["eps"] ** 2).sum()))
    st = res["surrogate"]["stats"]
    answered = (st["evaluated"] - st["rejected"]) / max(1, st["plastic"])
    gain = res["mixed"]["k_node_ms"] / res["surrogate"]["k_node_ms"]
    for m in ("fp64", "mixed", "surrogate"):
        r = res[m]
        print(f"{m:9s} L {r['L'] * 1e3:.4f} mm R {r['R'] * 1e3:.4f} mm k_node {r['k_node_ms']:.4f} ms "
              f"step {r['step_ms']:.4f} ms" + (f" L err {r['L_err']:.2e} R err {r['R_err']:.2e} eqps L2 {r['eps_l2']:.2e}"
                                               if m != "fp64" else ""))
    print(f"{material}: surrogate stats {st}: network answered {answered:.3f} of plastic return maps; "
          f"k_node time ratio (exact / surrogate) {gain:.3f}")
    assert answered > 0.3
    assert res["surrogate"]["L_err"] <= 0.005 and res["surrogate"]["R_err"] <= 0.005
