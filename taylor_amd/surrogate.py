"""Learned J2 return map: harvest training tuples from the exact method's own
runs, train the network on the GPU with the library's own kernels, and save,
load and attach the result.

The network (13-64-64-64-7, SiLU, bf16 operands on the matrix cores) stands
in for VON_MISES (CONSTITUTE_VONMISES.F90:21-158); the step falls back to the
exact return map outside the training envelope and wherever the network's
answer misses the yield surface by more than tol * sigma_y.

    python -m taylor_amd.surrogate                       # copper, linear hardening
    python -m taylor_amd.surrogate --material powerlaw   # the reference deck's aluminium, power law

Training runs: the 181k-node bar (spacing 0.19 mm, jitter seed 1) in the
semi-Lagrangian formulation with bins rebuilt every 20 steps, at 150, 190 and
250 m/s, every 4th step of the first 2,000 (or until the run leaves the range
in which the formulation stays finite).
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

if __package__ in (None, ""):
    sys.path.insert(0, os.path.dirname(os.path.dirname(os.path.abspath(__file__))))

from taylor_amd import deck, mega  # noqa: E402

HERE = os.path.dirname(os.path.abspath(__file__))
DEFAULT_WEIGHTS = os.path.join(HERE, "data", "j2_copper.npz")
POWERLAW_WEIGHTS = os.path.join(HERE, "data", "j2_powerlaw.npz")
MATERIALS = {"copper": deck.copper_linear, "powerlaw": deck.aluminium_powerlaw}


def harvest(prob, steps: int, every: int, cap: int, precision: int = mega.MEGA_FP64, log=None, chunk: int = 100):
    """Run `steps` explicit steps and return the plastic return-map tuples of
    every `every`-th step: float32 rows (de[6], rotated stress[6], eqps_n,
    stress[6], eqps).  The formulation loses stability at fine spacing: a
    chunk in which a return map fails, the state stops being finite or a node
    moves faster than 5 times the impact speed is dropped and the run ends."""
    g = mega.Mega(prob, precision=precision)
    g.harvest(cap, every=every)
    v0 = float(np.abs(prob.vinit).max())
    out, done = [], 0
    while done < steps:
        k = min(chunk, steps - done)
        try:
            st = g.step(k)
        except mega.MegaError as e:   # e.g. nodes piling into the edge bins
            if log:
                log(f"    stopped at step {done + k}: {e}")
            break
        done += k
        tup, count = g.harvest_take(cap)
        if count > cap:
            raise RuntimeError(f"harvest buffer overflow: {count} tuples > {cap}")
        v = g.download("v")
        vmax = float(np.sqrt((v * v).sum(1)).max())
        if st != 0 or not (np.isfinite(g.download("x")).all() and np.isfinite(tup).all()) or not vmax <= 5.0 * v0:
            # the formulation's own instability has set in: keep what came before
            if log:
                log(f"    run left the finite range by step {done} (status {st}, max speed {vmax:.0f}); chunk dropped")
            break
        out.append(tup)
        if log and (done % 1000 == 0 or done == steps):
            log(f"    step {done}/{steps}: {sum(t.shape[0] for t in out)} tuples")
    g.close()
    return np.concatenate(out) if out else np.zeros((0, 20), np.float32)


def train(tuples, props, iters: int, batch: int, lr0: float, lr1: float, seed: int = 0, chunk: int = 2000,
          log=None):
    """Adam on the GPU; the learning rate follows one cosine from lr0 to lr1
    over all iterations, run in chunks so that progress can be reported."""
    m = mega.Mlp(seed)
    m.set_data(tuples, props)
    t0 = time.perf_counter()
    sched = lambda k: lr1 + (lr0 - lr1) * 0.5 * (1.0 + np.cos(np.pi * k / max(1, iters)))  # noqa: E731
    parts, done = [], 0
    while done < iters:
        k = min(chunk, iters - done)
        parts.append(m.train(k, batch=batch, lr0=sched(done), lr1=sched(done + k)))
        done += k
        if log:
            log(f"    iteration {done}/{iters}: loss {parts[-1][-min(200, k):].mean():.4e}")
    hist = np.concatenate(parts) if parts else np.zeros(0, np.float32)
    dt = time.perf_counter() - t0
    out = {"params": m.params, "norm": m.norm, "loss": hist, "train_seconds": dt}
    m.close()
    return out


def return_residual(pred, s, eps_n, eps, props):
    """Largest residual of the radial-return equations (yield condition, flow
    direction, pressure, plastic multiplier) in units of sigma_y(eps), as the
    step's acceptance test computes it."""
    def dev(v):
        d = v.copy()
        d[:, :3] -= v[:, :3].sum(1, keepdims=True) / 3
        return d

    def tn(v):
        return np.sqrt((v[:, :3] ** 2).sum(1) + 2 * (v[:, 3:] ** 2).sum(1))

    a, b = dev(pred), dev(s)
    na, nb = tn(a), tn(b)
    sy = props[3] * (1.0 + props[4] * eps) ** props[5]
    r = np.maximum.reduce([np.abs(np.sqrt(1.5) * nb - sy), tn(b - (nb / na)[:, None] * a),
                           np.abs(s[:, :3].sum(1) - pred[:, :3].sum(1)) / 3,
                           np.abs(na - nb - 2 * props[28] * np.sqrt(1.5) * (eps - eps_n))])
    return r / sy


def evaluate(w: dict, tuples, tol: float = 1e-3, seed: int = 0) -> dict:
    """Offline check of trained weights on tuples of the exact method: share
    inside the envelope, share the step would accept (return-map residual
    within tol, eqps not decreasing) and the accepted stresses' error."""
    u = np.asarray(tuples, np.float32)
    p = np.asarray(w["props"], np.float64)
    nm = np.asarray(w["norm"], np.float32)
    x = u[:, :13]
    inside = np.all((x >= nm[33:46]) & (x <= nm[46:59]), axis=1)
    xh = ((x - nm[0:13]) * nm[13:26]).astype(np.float32)
    m = mega.Mlp(seed)
    m.params = w["params"]
    y = m.forward(xh).astype(np.float64)
    m.close()
    lam, mu, l2m = p[27], p[28], p[29]
    de, s_rot = u[:, 0:6].astype(np.float64), u[:, 6:12].astype(np.float64)
    tr = s_rot.copy()
    tr[:, 0] += l2m * de[:, 0] + lam * de[:, 1] + lam * de[:, 2]
    tr[:, 1] += lam * de[:, 0] + l2m * de[:, 1] + lam * de[:, 2]
    tr[:, 2] += lam * de[:, 0] + lam * de[:, 1] + l2m * de[:, 2]
    tr[:, 3:] += mu * de[:, 3:]
    os_ = nm[26:33].astype(np.float64)
    s = tr - os_[:6] * y[:, :6]
    e = u[:, 12] + os_[6] * y[:, 6]
    ok = inside & (return_residual(tr, s, u[:, 12], e, p) <= tol) & (e >= u[:, 12])
    err = np.abs(s - u[:, 13:19]).max(1) / p[3]
    res = np.where(inside, return_residual(tr, s, u[:, 12], e, p), np.inf)
    return {"tuples": int(u.shape[0]), "inside": float(inside.mean()), "accepted": float(ok.mean()),
            "accepted_at_tol": {f"{t:g}": float(((res <= t) & (e >= u[:, 12])).mean()) for t in (3e-4, 1e-3, 3e-3, 1e-2)},
            "stress_err_accepted_max": float(err[ok].max()) if ok.any() else None,
            "stress_err_accepted_rms": float(np.sqrt((err[ok] ** 2).mean())) if ok.any() else None,
            "stress_err_all_rms": float(np.sqrt((err[inside] ** 2).mean())) if inside.any() else None}


def save(path, params, norm, props, meta: dict):
    os.makedirs(os.path.dirname(os.path.abspath(path)), exist_ok=True)
    np.savez_compressed(path, params=np.asarray(params, np.float32), norm=np.asarray(norm, np.float32),
                        props=np.asarray(props, np.float64), meta=np.array(json.dumps(meta)))


def load(path=DEFAULT_WEIGHTS):
    z = np.load(path)
    return {"params": z["params"], "norm": z["norm"], "props": z["props"], "meta": json.loads(str(z["meta"]))}


def attach(g: "mega.Mega", weights=None, tol: float = 1e-3):
    w = load(weights or DEFAULT_WEIGHTS) if not isinstance(weights, dict) else weights
    g.set_surrogate(w["params"], w["norm"], w["props"], tol=tol)
    return w


def main():
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--material", choices=sorted(MATERIALS), default="copper")
    ap.add_argument("--out", default=None)
    ap.add_argument("--velocities", type=float, nargs="+", default=[150.0, 190.0, 250.0])
    ap.add_argument("--h", type=float, default=0.19e-3, help="lattice spacing of the harvest runs")
    ap.add_argument("--seed", type=int, default=1)
    ap.add_argument("--lagrangian", type=int, default=0, help="1: Lagrangian harvest runs (default semi-Lagrangian)")
    ap.add_argument("--pdsearch", type=int, default=20, help="bin rebuild period of the semi-Lagrangian runs")
    ap.add_argument("--steps", type=int, default=2000)
    ap.add_argument("--every", type=int, default=4)
    ap.add_argument("--h2", type=float, default=0.0, help="spacing of optional full-length (80 us) runs (0: none)")
    ap.add_argument("--every2", type=int, default=2)
    ap.add_argument("--cap", type=int, default=20_000_000)
    ap.add_argument("--iters", type=int, default=300000)
    ap.add_argument("--batch", type=int, default=16384)
    ap.add_argument("--lr0", type=float, default=2e-3)
    ap.add_argument("--lr1", type=float, default=1e-5)
    args = ap.parse_args()

    def log(*a):
        print(*a, flush=True)

    out_path = args.out or (DEFAULT_WEIGHTS if args.material == "copper" else POWERLAW_WEIGHTS)
    mat = MATERIALS[args.material]()
    sets = []
    runs = [(args.h, args.seed, args.steps, args.every, args.lagrangian)]
    if args.h2 > 0:
        runs.append((args.h2, 0, None, args.every2, 1))
    for v in args.velocities:
        for h, seed, steps, every, lag in runs:
            prob = deck.taylor_problem(h, v0=v, seed=seed, material=mat, lagrangian=lag, pdsearch_steps=args.pdsearch)
            n_steps = steps if steps else int(np.floor(80e-6 / prob.dt)) + 1
            log(f"harvest: h {h * 1e3:.3f} mm, {prob.n} nodes, {v} m/s, {n_steps} steps, every {every}")
            t0 = time.perf_counter()
            sets.append(harvest(prob, n_steps, every, args.cap, log=log))
            log(f"  {sets[-1].shape[0]} tuples in {time.perf_counter() - t0:.1f} s")
    tup = np.concatenate(sets)
    log(f"training set: {tup.shape[0]} tuples")
    res = train(tup, mat.props, args.iters, args.batch, args.lr0, args.lr1, log=log)
    hist = res["loss"]
    log(f"trained {args.iters} iterations in {res['train_seconds']:.1f} s: loss {hist[:100].mean():.4e} -> "
        f"{hist[-500:].mean():.4e}")
    meta = {"material": args.material, "lagrangian": args.lagrangian, "pdsearch": args.pdsearch,
            "velocities": args.velocities, "h": args.h, "seed": args.seed, "steps": args.steps, "every": args.every,
            "h2": args.h2, "every2": args.every2, "tuples": int(tup.shape[0]), "iters": args.iters,
            "batch": args.batch, "lr0": args.lr0, "lr1": args.lr1, "final_loss": float(hist[-500:].mean()),
            "train_seconds": res["train_seconds"]}
    w = {"params": res["params"], "norm": res["norm"], "props": mat.props}
    rng = np.random.default_rng(0)
    sample = tup[rng.choice(tup.shape[0], size=min(tup.shape[0], 2_000_000), replace=False)]
    meta["train_eval"] = evaluate(w, sample)
    log("training-set check:", meta["train_eval"])
    prob = deck.taylor_problem(args.h, v0=220.0, seed=args.seed, material=mat, lagrangian=args.lagrangian,
                               pdsearch_steps=args.pdsearch)
    held = harvest(prob, min(args.steps, 600), 5, args.cap)
    meta["heldout_220"] = evaluate(w, held)
    log("held-out 220 m/s check:", meta["heldout_220"])
    save(out_path, res["params"], res["norm"], mat.props, meta)
    np.save(os.path.splitext(out_path)[0] + "_loss.npy", hist[:: max(1, len(hist) // 2000)])
    log(f"saved {out_path}")


if __name__ == "__main__":
    main()
