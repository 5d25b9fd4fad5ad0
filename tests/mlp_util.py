"""Shared helpers of the learned-return-map tests: exact J2 tuples from the
oracle's VON_MISES restatement (pinned to the reference's golden vectors)."""
from __future__ import annotations

import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, os.path.join(ROOT, "oracle"))
import mlp_oracle  # noqa: E402,F401


def j2_tuples(oracle, props, n, rng, plastic_only=True):
    """n tuples (de[6], rotated stress[6], eqps_n, stress[6], eqps) as float32
    rows of 20, from random rotated stresses near the yield surface and random
    strain increments, returned through or_von_mises; only plastic trial
    states when plastic_only."""
    lam, mu, l2m, sig0, B = props[27], props[28], props[29], props[3], props[4]
    out = []
    while len(out) < n:
        s = rng.normal(size=6) * sig0 * 0.6
        de = rng.normal(size=6) * 2e-3 * rng.uniform(0.01, 1.0)
        e0 = abs(rng.normal()) * 0.3
        s32, de32, e32 = s.astype(np.float32), de.astype(np.float32), np.float32(e0)
        sd, ded = s32.astype(np.float64), de32.astype(np.float64)
        pred = sd.copy()
        pred[0] += l2m * ded[0] + lam * ded[1] + lam * ded[2]
        pred[1] += lam * ded[0] + l2m * ded[1] + lam * ded[2]
        pred[2] += lam * ded[0] + lam * ded[1] + l2m * ded[2]
        pred[3:] += mu * ded[3:]
        dev = pred.copy()
        dev[:3] -= pred[:3].sum() / 3
        f = np.sqrt((dev[:3] ** 2).sum() + 2 * (dev[3:] ** 2).sum()) - np.sqrt(2 / 3) * sig0 * (1 + B * float(e32))
        if plastic_only and f <= 0:
            continue
        so, st = np.zeros(6), np.zeros(20)
        st[2] = float(e32)
        assert oracle.or_von_mises(so, pred, st, props) >= 0
        out.append(np.concatenate([de32, s32, [e32], so, [st[2]]]).astype(np.float32))
    return np.array(out, dtype=np.float32)
