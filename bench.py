"""Benchmark of the explicit meshfree step on one MI355X.

Workload: the Taylor bar of 25.4 mm x 3.8 mm radius on a cubic lattice of
spacing 0.088 mm (jitter 5 % of the spacing, seed 2; ~1.75M nodes, ~58
neighbors per node), copper J2 with linear hardening, 220 m/s onto a fixed
plane (the learned return map's test speed, outside its training runs),
dt = 0.3 h / c_p, Lagrangian finite strain with direct nodal integration.
Synthetic deck.

A "step" is one explicit step of the whole model (predictor, interpolation,
internal force with the constitution, corrector).  The timed region is K
steps after an untimed prelude of up to 600 steps (into the plastic regime,
where about a quarter of the nodes yield; cut so that the timed steps end
inside the window in which the run stays finite) and W warm-up steps, with
all inputs resident in HBM, run in three modes: fp64 (plain), mixed (FP32 shape data, stress, return map and force;
FP64 positions, F, F^-1 and rotation) and surrogate (mixed, with the learned
J2 return map of taylor_amd/data/j2_copper.npz and the exact return map as
fallback).  value = surrogate node-steps/s; the other rates are beside it.
During the timed region HIP events bracket the force kernel only (the
roofline's kernel); kernel_ms, the per-kernel breakdown, comes from up to 50
further steps with events on every kernel, which add a few percent.

impact_start_window: the same three modes at 190 m/s over steps W+1..W+K
from the impact, beside the headline.

Unless --no-validate: (1) the three modes run the benchmark bar for
--stable-steps steps (the window in which the reference formulation stays
finite at this resolution; it loses stability later in FP64 too, as the
reference program does at 0.19 mm) and their bar length, mushroom radius
and eqps field are compared with FP64, and again over 3,500 steps with the
smoothed gradients (ITYPE_INT = 1, an extension that keeps the run finite
about four times longer); (2) the
whole 80 us CPU-runnable bar (spacing 0.5 mm) runs in the three modes and is
compared with FP64 and with the reference program's final state
(tests/golden/taylor_h05_final.npz).  Gate: length within 0.5 %.

semi_lagrangian (unless --no-semi): the same bar in the semi-Lagrangian
formulation, FP64 and mixed rates beside the headline (not part of value);
pd_plate (unless --no-pd): correspondence peridynamics on the 6.53M-particle
plate (100 x 100 x 5 mm, spacing 0.2 mm), mixed and FP64 rates and the pair
kernels' roofline fractions;
preprocessing: the one-time geometry preprocessing of the mesh on the GPU
(mega_preprocess) beside its host restatement.

roofline: the dominant kernel's algorithmic bytes per launch (compulsory
HBM traffic: every pair and node array it must touch read or written once,
neighbor re-reads excluded) over its mean launch time from HIP events on the
context's stream during the timed region; peak 8 TB/s.  traffic: HBM bytes per launch measured with
rocprofv3 counters, read from profiles/ when present.

cpu_baseline: the reference program (oracle/_ref/MEGA, OpenMP) on the host
cores, on a bounded sample of the same workload (same spacing and material,
a 2.54 mm slice of the bar, ~181k nodes): node-steps/s from the difference of
a 12-step and a 2-step run (setup cancels).  Falls back to the oracle
restatement when the reference build is absent.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--prelude P] [--modes fp64,mixed,surrogate]
                    [--stable-steps S] [--no-validate] [--no-cpu] [--no-semi] [--no-pd] [--pd-only]
"""
from __future__ import annotations

import argparse
import json
import os
import subprocess
import sys
import tempfile
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

from taylor_amd import deck  # noqa: E402
from taylor_amd import mega  # noqa: E402
from taylor_amd import surrogate  # noqa: E402

H = 0.088e-3
SEED = 2
V0 = 220.0            # the learned return map's test speed (its training runs were 150, 190, 250 m/s)
V0_IMPACT = 190.0     # the speed of the impact-start window beside the headline
T_END = 80e-6
PEAK_HBM = 8.0e12
PEAK_FP64 = 78.6e12   # MI355X vector FP64 (flop/s)
# steps over which the 1.75M-node bar stays finite at V0 in every mode (direct
# nodal integration has no live stabilisation; at 220 m/s the FP64 step moves
# a node faster than 5 v0 by step 950): the prelude is cut so that the timed
# steps end inside it
FINITE_WINDOW = 850
# with the smoothed gradients (ITYPE_INT = 1, an extension) the same bar stays
# finite in every mode through ~3,750 steps at V0 (21.6 us)
SNNI_WINDOW = 3500


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def kernel_bytes(name: str, n: int, nnz: int, mixed: bool) -> float:
    """Algorithmic (compulsory) HBM bytes of one launch of the step kernels.
    r: bytes of the shape/stress type; node masters are FP64."""
    r = 4 if mixed else 8
    if name == "k_gather":
        # pairs: col + phi + grad phi; nodes: own d (FP64) read, uphy r/w, the gather
        # sums written as the steps carry F (L0 and, mixed, S: 12 floats; FP64 L0
        # padded to 10 doubles).  x = X0 + u is stored only on a call's last step
        # (nothing in the step reads it)
        return nnz * (4 + r + 3 * r) + n * (24 + 48 + (12 if mixed else 10) * r)
    if name == "k_node":
        # the gather sums read (12 floats / 10 doubles), stress/strain r/w, eps r/w, vol,
        # mat id, the force record P written whole (12 floats / 10 doubles), own d
        # (mixed: S x d_i), the carried deformation gradient r/w (10 doubles)
        return n * ((12 if mixed else 10) * r + 4 * 6 * r + 2 * r + 8 + 4 + (12 if mixed else 10) * r +
                    (24 if mixed else 0) + 2 * 80)
    if name == "k_force":
        # pairs: source + grad phi (transposed); nodes: P read once, mass (one per node),
        # ebc, mat id, v r/w, d write, dtot r/w (+ the FP32 copy of d in mixed).  fint,
        # fext and a are stored only on a call's last steps (the next predictor is
        # fused, so nothing in the step reads them)
        return nnz * (4 + 3 * r) + n * (10 * r + 8 + 12 + 4 + 48 + 24 + 48 + (12 if mixed else 0))
    if name == "k_predict":
        # v and a read, d written, dtot r/w, v written (+ the FP32 copy of d in mixed)
        return n * (24 + 24 + 24 + 48 + 24 + (12 if mixed else 0))
    if name == "k_search_count":
        # one wave per bin tests the 27 bins' candidates; per node: own x and window
        # (a candidate's are re-reads), its slot in the bin list, the count, the kept
        # ballot masks (8 x 8 B) and the 27 per-source-bin counts of its transposed row
        return n * (24 + 24 + 4 + 4 + 64 + 27 * 4)
    if name == "k_search_fill":
        # the same walk over the kept masks (tests again beyond them): per node the masks,
        # row start, the 27 transposed group cursors read and written; per pair the
        # neighbor index and its transposed slot written
        return n * (64 + 4 + 24 + 24 + 2 * 27 * 4) + nnz * (4 + 4)
    if name == "k_pd_dt":
        # the minimum over particles of the shortest bond (formed by the gather) over
        # the wave speed: both read once
        return n * (8 + 8)
    if name == "k_pd_gather":
        # pairs: bond index + influence weight; nodes: own X0, x, d, volume, row start,
        # the 18 gather sums written
        return nnz * (4 + r) + n * (24 + 24 + 24 + 8 + 4 + 18 * r)
    if name == "k_pd_node":
        # sums read, K^-1 read, stress/strain r/w, eps r/w, volume, mat id, T record written
        return n * (18 * r + 72 + 24 * r + 2 * r + 8 + 4 + 10 * r)
    if name == "k_pd_force":
        # pairs: source + influence weight (transposed); nodes: row start, own X0, own T,
        # g vector, then the corrector's node arrays as k_force (mass, ebc, mat id, v r/w,
        # d write, dtot r/w, the FP32 copy of d in mixed) and, since the particles move
        # with d, uphy r/w and x written
        return nnz * (4 + r) + n * (4 + 24 + 10 * r + 24 + 8 + 12 + 4 + 48 + 24 + 48 + 48 + 24 +
                                    (12 if mixed else 0))
    if name == "k_rk":
        # the semi-Lagrangian launch, gather sums fused: pairs: col + tpos read,
        # phi written in list order and {grad phi, source} in transposed order;
        # nodes: own x, window, EBC flag, row, own d (FP32 copy in mixed), sums written
        return nnz * (4 + 4 + r + 3 * r + 4) + n * (24 + 24 + 4 + 8 + (16 if mixed else 24) + (48 if mixed else 80))
    return 0.0


def node_step_bytes(n: int, nnz: int, mixed: bool) -> float:
    """Algorithmic bytes per node-step of the whole step (fused-pass model: one
    pass over the pairs, index + phi + grad phi, and each node array once)."""
    kbar = nnz / n
    return kbar * (20 if mixed else 36) + (368 if mixed else 468)


def semi_step_bytes(n: int, nnz: int, mixed: bool) -> float:
    """The same for the semi-Lagrangian step: the search writes the indices,
    the force pass reads them and writes phi, the next interpolation reads
    index + phi (RK in registers, bins amortised over 20 steps)."""
    kbar = nnz / n
    return kbar * (20 if mixed else 28) + (400 if mixed else 500)


PD_DLT_FAC = 0.5


def pd_plate(steps: int, warmup: int, modes=("mixed", "fp64")) -> dict:
    """Correspondence peridynamics (CONSTRUCT_FINT_PD) on the plate:
    100 x 100 x 5 mm at spacing 0.2 mm (jitter seed 3, 6.53M particles),
    elastic aluminium, horizon ~3 spacings, clamped lateral faces, a 5 mm
    plug at 300 m/s; mixed (FP32 bond forces, FP64 integration) and FP64
    node-steps/s, per-kernel times and each pair kernel's fraction of the
    HBM roofline."""
    t = time.time()
    prob = deck.pd_plate_problem(preprocess_fn=mega.geometry)
    log(f"PD plate: {prob.n} particles, deck {time.time() - t:.1f} s")
    out = {"workload": "PD plate 100 x 100 x 5 mm, h 0.2 mm, jitter 5% seed 3, elastic aluminium, horizon 3.0 h, "
                       "clamped edges, 5 mm plug at 300 m/s, correspondence PD (FINT_PD)",
           "particles": prob.n, "steps": steps}
    for name in modes:
        wall, kt_timed, kt, nnz, mx, _ = timed_run(prob, name, steps, warmup)
        rate = prob.n * steps / wall
        mixed = name != "fp64"
        kern = {}
        for k in ("k_pd_gather", "k_pd_node", "k_pd_force"):
            if k in kt:
                ms = kt_timed[k][0] if k in kt_timed else kt[k][0]
                b = kernel_bytes(k, prob.n, nnz, mixed)
                kern[k] = {"ms": ms, "algorithmic_bytes": b, "achieved_GB_s": b / (ms * 1e-3) / 1e9,
                           "frac": b / (ms * 1e-3) / PEAK_HBM}
        out[name] = {"rate": rate, "ms_per_step": 1e3 * wall / steps, "bonds": nnz, "bonds_per_particle": nnz / prob.n,
                     "max_bonds": mx, "kernels": kern}
        log(f"PD plate {name}: {1e3 * wall / steps:.3f} ms/step, {rate:.4e} node-steps/s, "
            + ", ".join(f"{k} {v['ms']:.3f} ms ({v['frac']:.2f} of HBM)" for k, v in kern.items()))
    if "mixed" in modes:
        # the automatic time step (IAUTO_TS = 1, DLT_FAC 0.5): the same step plus k_pd_dt
        wall, _, kt, nnz, _, _ = timed_run(prob, "mixed", steps, warmup, dlt_fac=PD_DLT_FAC)
        b = kernel_bytes("k_pd_dt", prob.n, nnz, True)
        ms = kt["k_pd_dt"][0]
        out["mixed_automatic_step"] = {"dlt_fac": PD_DLT_FAC, "rate": prob.n * steps / wall,
                                       "ms_per_step": 1e3 * wall / steps, "k_pd_dt_ms": ms,
                                       "k_pd_dt_frac": b / (ms * 1e-3) / PEAK_HBM,
                                       "kernel_ms": {k: v[0] for k, v in kt.items()}}
        log(f"PD plate mixed, automatic step: {1e3 * wall / steps:.3f} ms/step, k_pd_dt {ms:.3f} ms "
            f"({out['mixed_automatic_step']['k_pd_dt_frac']:.2f} of HBM), k_pd_gather {kt['k_pd_gather'][0]:.3f} ms")
    return out


def preprocessing() -> dict:
    """The one-time geometry preprocessing (GET_ADIAL, GET_VOLUMES,
    GET_SMOOTHING_LENGTHS) of the benchmark mesh: mega_preprocess on the GPU,
    host arrays in and out (PCIe included), beside the host restatement."""
    mesh = deck.taylor_bar_mesh(H, seed=SEED)
    mega.preprocess(mesh.coo[:1000], mesh.elcon[:10] * 0 + 1)     # load the library, warm the runtime
    t0 = time.perf_counter()
    win, vol, _, dmax = mega.preprocess(mesh.coo, mesh.elcon)
    t1 = time.perf_counter()
    ref = deck.preprocess(mesh)
    t2 = time.perf_counter()
    return {"nodes": int(mesh.coo.shape[0]), "elements": int(mesh.elcon.shape[0]),
            "gpu_ms_with_transfers": 1e3 * (t1 - t0), "host_numpy_ms": 1e3 * (t2 - t1),
            "windows_volumes_equal": bool(np.array_equal(win, ref.win0) and np.array_equal(vol, ref.vol))}


def rk_flops(n: int, nnz: int) -> float:
    """Minimal FP64 flops of RK1 for all nodes (symmetric moment matrix M and
    its three derivatives, the 4x4 inverse, the shape functions and their
    gradients): about 250 per neighbor and 1,500 per node."""
    return 250.0 * nnz + 1500.0 * n


def semi_lagrangian(steps: int, warmup: int) -> dict:
    """The same bar in the semi-Lagrangian formulation (SOFT_SEARCH and RK1
    every step, GET_BINS every 20 steps): FP64 and mixed rates, and the
    shape-function kernel against both its rooflines (HBM bytes and FP64
    flops; its moment matrix and inverse are FP64 in both modes)."""
    prob = deck.taylor_problem(H, v0=V0_IMPACT, seed=SEED, lagrangian=0, pdsearch_steps=20)
    out = {"workload": "same bar, semi-Lagrangian, bins every 20 steps", "steps": steps}
    for name in ("fp64", "mixed"):
        wall, _, kt, nnz, mx, _ = timed_run(prob, name, steps, warmup)
        rate = prob.n * steps / wall
        out[name] = {"rate": rate, "ms_per_step": 1e3 * wall / steps,
                     "kernel_ms": {k: v[0] for k, v in kt.items()},
                     "step_roofline_frac": rate * semi_step_bytes(prob.n, nnz, name != "fp64") / PEAK_HBM}
        if "k_rk" in kt:
            s = kt["k_rk"][0] * 1e-3
            out[name]["k_rk"] = {"ms": kt["k_rk"][0],
                                 "hbm_frac": kernel_bytes("k_rk", prob.n, nnz, name != "fp64") / s / PEAK_HBM,
                                 "fp64_flops": rk_flops(prob.n, nnz),
                                 "fp64_frac": rk_flops(prob.n, nnz) / s / PEAK_FP64}
        log(f"semi-Lagrangian {name}: {1e3 * wall / steps:.3f} ms/step, {rate:.4e} node-steps/s"
            + (f"; k_rk {out[name]['k_rk']['hbm_frac']:.3f} of HBM, {out[name]['k_rk']['fp64_frac']:.3f} of FP64"
               if "k_rk" in out[name] else ""))
    return out


def read_traffic(kernel: str, mixed: bool):
    """HBM bytes per dispatch of a kernel from the committed rocprofv3 counter
    summary (profiles/pmc_<kernel>_taylor1p75M_<mode>.txt), or None."""
    path = os.path.join(ROOT, "profiles", f"pmc_{kernel}_taylor1p75M_{'mixed' if mixed else 'fp64'}.txt")
    if not os.path.exists(path):
        return None
    for line in open(path):
        if line.startswith("traffic_bytes_per_dispatch"):
            return float(line.split()[1])
    return None


MODES = {"fp64": mega.MEGA_FP64, "mixed": mega.MEGA_MIXED, "surrogate": mega.MEGA_MIXED}


def make(prob, mode: str):
    g = mega.Mega(prob, precision=MODES[mode])
    if mode == "surrogate":
        surrogate.attach(g, surrogate.DEFAULT_WEIGHTS)
    return g


def timed_run(prob, mode: str, steps: int, warmup: int, prelude: int = 0, dlt_fac: float | None = None):
    """Wall time of `steps` steps after `prelude` + `warmup` untimed steps,
    with HIP events on the force kernel only (the roofline's kernel; events on
    every kernel add a few percent), then the per-kernel breakdown from a
    further profiled run of up to 50 steps.  Returns (wall, force events of
    the timed region, breakdown, nnz, max neighbors, surrogate counters of the
    timed region).  dlt_fac: the automatic time step (peridynamics) with that
    DLT_FAC instead of the fixed step."""
    g = make(prob, mode)
    adv = g.step if dlt_fac is None else (lambda k: g.step_auto(dlt_fac, k))
    adv(max(prelude + warmup, 1))        # includes the first step (search, shapes)
    if mode == "surrogate":
        g.surrogate_stats(reset=True)
    g.profile(2)
    t0 = time.perf_counter()
    adv(steps)                           # returns after the stream is synchronised
    t1 = time.perf_counter()
    kt_timed = g.kernel_times()
    nnz, mx, _ = g.stats()
    st = g.surrogate_stats() if mode == "surrogate" else None
    g.profile(1)
    adv(min(steps, 50))
    kt = g.kernel_times()
    g.close()
    return (t1 - t0), kt_timed, kt, nnz, mx, st


def run_to(prob, mode: str, nsteps: int):
    """Final length (max z above the wall plane), mushroom radius (max radius
    over the impact-face nodes) and eqps field after nsteps."""
    g = make(prob, mode)
    done = 0
    while done < nsteps:
        k = min(2000, nsteps - done)
        g.step(k)
        done += k
    x = g.download("x")
    eps = g.download("eps")
    st = g.surrogate_stats() if mode == "surrogate" else None
    g.close()
    wall = prob.sets["WALL"] - 1
    length = float(x[:, 2].max())
    radius = float(np.sqrt(x[wall, 0] ** 2 + x[wall, 1] ** 2).max())
    return {"length_m": length, "radius_m": radius, "eps": eps, "surrogate_stats": st,
            "finite": bool(np.isfinite(x).all())}


def compare(prob, runs: dict, ref_key: str = "fp64") -> dict:
    ref = runs[ref_key]
    vol = prob.vol
    out = {}
    for k, r in runs.items():
        d = {"length_m": r["length_m"], "radius_m": r["radius_m"], "finite": r["finite"]}
        if k != ref_key:
            d["length_rel_err"] = abs(r["length_m"] - ref["length_m"]) / ref["length_m"]
            d["radius_rel_err"] = abs(r["radius_m"] - ref["radius_m"]) / ref["radius_m"]
            d["eqps_l2_rel_err"] = float(np.sqrt((vol * (r["eps"] - ref["eps"]) ** 2).sum() /
                                                 (vol * ref["eps"] ** 2).sum()))
            d["length_gate_0p5pct"] = bool(d["length_rel_err"] <= 0.005)
        if r.get("surrogate_stats"):
            d["surrogate_stats"] = r["surrogate_stats"]
        out[k] = d
    return out


def validate(prob, stable_steps: int):
    """(1) the benchmark bar at the headline speed over the window in which
    the formulation stays finite at this resolution (its last finite state is
    the attainable "final" one), and over the longer window the smoothed
    gradients (ITYPE_INT = 1) allow; (2) the whole CPU-runnable bar (spacing
    0.5 mm, 190 m/s, 80 us), also against the reference program's final state
    when the fixture is present."""
    res = {}
    runs = {m: run_to(prob, m, stable_steps) for m in MODES}
    res["taylor_1p75M_window"] = {"v0_m_s": V0, "steps": stable_steps, "t_end_s": stable_steps * prob.dt,
                                  **compare(prob, runs)}
    ps = deck.taylor_problem(H, v0=V0, seed=SEED, itype_int=1)
    runs = {m: run_to(ps, m, SNNI_WINDOW) for m in MODES}
    res["taylor_1p75M_smoothed_gradients_window"] = {"v0_m_s": V0, "steps": SNNI_WINDOW,
                                                     "t_end_s": SNNI_WINDOW * ps.dt, **compare(ps, runs)}
    p5 = deck.taylor_problem(0.5e-3, v0=V0_IMPACT, seed=0)
    n5 = int(np.floor(T_END / p5.dt)) + 1
    runs = {m: run_to(p5, m, n5) for m in MODES}
    out = {"steps": n5, "t_end_s": T_END, "nodes": p5.n, **compare(p5, runs)}
    fix = os.path.join(ROOT, "tests", "golden", "taylor_h05_final.npz")
    if os.path.exists(fix):
        z = np.load(fix)
        wall = p5.sets["WALL"] - 1
        L0 = float(z["x"][:, 2].max())
        R0 = float(np.sqrt(z["x"][wall, 0] ** 2 + z["x"][wall, 1] ** 2).max())
        out["reference_program"] = {"length_m": L0, "radius_m": R0}
        for m in MODES:
            out[m]["length_rel_err_vs_reference"] = abs(runs[m]["length_m"] - L0) / L0
            out[m]["radius_rel_err_vs_reference"] = abs(runs[m]["radius_m"] - R0) / R0
    res["taylor_h0p5_80us"] = out
    return res


def host_cpu() -> str:
    try:
        with open("/proc/cpuinfo") as f:
            for ln in f:
                if ln.startswith("model name"):
                    return ln.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline_reference(sample_len: float, cores: int):
    from oracle_bind import REF_MEGA  # noqa: E402
    if not os.path.exists(REF_MEGA):
        return None
    mesh = deck.taylor_bar_mesh(H, length=sample_len, seed=SEED)
    prob = deck.taylor_problem(H, v0=V0, seed=SEED, length=sample_len)
    times = {}
    with tempfile.TemporaryDirectory() as wd:
        kfile = os.path.join(wd, "bar.k")
        deck.write_k(kfile, mesh)
        for nsteps in (2, 12):
            prob.time_end = prob.dt * (nsteps - 0.5)
            deck.write_control(os.path.join(wd, "Control.dat"), kfile, prob, ["BAR", "WALL"],
                               time_output=1.0, out_vars="node_id")
            t0 = time.perf_counter()
            with open(os.path.join(wd, "mega.out"), "w") as fo:
                subprocess.run([REF_MEGA, "-j", str(cores)], cwd=wd, stdin=subprocess.DEVNULL, stdout=fo,
                               stderr=fo, check=True, timeout=600)
            times[nsteps] = time.perf_counter() - t0
    rate = prob.n * 10 / (times[12] - times[2])
    return {"value": rate, "unit": "node-steps/s", "cores": cores, "kind": "reference",
            "cpu_model": host_cpu(), "host_threads": os.cpu_count(),
            "sample": f"reference program MEGA -j {cores} ({cores} of {os.cpu_count()} host threads, "
                      f"{host_cpu()}), Taylor bar slice 2.54 mm at h=0.088 mm ({prob.n} nodes), {V0:.0f} m/s, "
                      f"FP64 Lagrangian; 12-step minus 2-step wall time ({times[12]:.1f} s - {times[2]:.1f} s). "
                      f"A slice, because the reference sizes its neighbor storage as GNUMP*1000 entries "
                      f"(HANDELER.F90:147): about 157 GiB at the whole bar's 1.75M nodes"}


def cpu_baseline_port(sample_len: float, cores: int):
    from oracle_bind import OracleRun, load_oracle  # noqa: E402
    prob = deck.taylor_problem(H, v0=V0, seed=SEED, length=sample_len)
    run = OracleRun(load_oracle(), prob, nthreads=cores)
    run.step(1)
    t0 = time.perf_counter()
    run.step(10)
    dt = time.perf_counter() - t0
    return {"value": prob.n * 10 / dt, "unit": "node-steps/s", "cores": cores, "kind": "port",
            "cpu_model": host_cpu(), "host_threads": os.cpu_count(),
            "sample": f"oracle restatement, {cores} of {os.cpu_count()} host threads ({host_cpu()}), Taylor bar "
                      f"slice 2.54 mm at h=0.088 mm ({prob.n} nodes), FP64 Lagrangian, 10 steps after the first"}


def copy_calibration():
    """Reference kernels for the HBM byte counters (rocprofv3 --pmc FETCH_SIZE /
    WRITE_SIZE): five device-to-device copies of 1 GiB with torch's copy kernel
    (1 GiB read and 1 GiB written per dispatch), then, for windows of 128 B,
    4 KiB and 64 KiB, five gathers of 4 B elements (torch.take) from a 256 MiB
    array through an index that visits every element once, shuffled within
    each window: 256 MiB of elements and 512 MiB of int64 indices read, 256 MiB
    written per dispatch, whatever the window; the counters then show how the
    reading for scattered loads departs from the copy's."""
    import torch
    a = torch.ones(1 << 28, dtype=torch.float32, device="cuda")
    b = torch.empty_like(a)
    for _ in range(5):
        b.copy_(a)
    g = torch.ones(1 << 26, dtype=torch.float32, device="cuda")
    for w in (1 << 5, 1 << 10, 1 << 14):   # elements per window: 128 B, 4 KiB, 64 KiB
        perm = torch.randperm(w, generator=torch.Generator().manual_seed(1)).to("cuda")
        idx = (torch.arange(1 << 26, device="cuda").view(-1, w)[:, perm]).reshape(-1).contiguous()
        for _ in range(5):
            torch.take(g, idx)
        del idx
    torch.cuda.synchronize()
    print(json.dumps({"copy_bytes_per_dispatch": a.numel() * 4, "dispatches": 5,
                      "gather_read_bytes_per_dispatch": g.numel() * (4 + 8),
                      "gather_write_bytes_per_dispatch": g.numel() * 4}))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--modes", default="fp64,mixed,surrogate")
    ap.add_argument("--prelude", type=int, default=600,
                    help="untimed steps before the warm-up, into the plastic regime (cut to the finite window)")
    ap.add_argument("--stable-steps", type=int, default=FINITE_WINDOW)
    ap.add_argument("--no-validate", action="store_true")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-semi", action="store_true",
                    help="skip the side measurements (semi-Lagrangian step, preprocessing)")
    ap.add_argument("--no-pd", action="store_true", help="skip the peridynamic plate side measurement")
    ap.add_argument("--no-impact", action="store_true",
                    help="skip the impact-start window beside the headline (profiling runs)")
    ap.add_argument("--pd-only", action="store_true",
                    help="only the peridynamic plate side measurement (for profiling its kernels)")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only run 5 device copies of 1 GiB (reference kernel for the HBM byte counters)")
    args = ap.parse_args()
    if args.copy_calibration:
        copy_calibration()
        return
    if args.gpus != 1:
        print("bench.py measures one MI355X: this build has no multi-GPU path (run with --gpus 1)", file=sys.stderr)
        sys.exit(2)
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    modes = [m for m in args.modes.split(",") if m]
    if args.pd_only:
        print(json.dumps({"pd_plate": pd_plate(min(args.steps, 20), 3)}))
        return

    t = time.time()
    prob = deck.taylor_problem(H, v0=V0, seed=SEED)
    n = prob.n
    log(f"workload: {n} nodes, dt {prob.dt:.4e} s, deck {time.time() - t:.1f} s")
    prelude = max(0, min(args.prelude, FINITE_WINDOW - args.warmup - args.steps))
    first = prelude + args.warmup + 1

    res = {}
    for name in modes:
        wall, kt_timed, kt, nnz, mx, st = timed_run(prob, name, args.steps, args.warmup, prelude)
        res[name] = dict(wall=wall, kt=kt, kt_timed=kt_timed, nnz=nnz, max_neighbors=mx, rate=n * args.steps / wall,
                         stats=st)
        log(f"{name} (steps {first}-{first + args.steps - 1}, {V0:.0f} m/s): {1e3 * wall / args.steps:.3f} ms/step, "
            f"{res[name]['rate']:.4e} node-steps/s, "
            + ", ".join(f"{k} {v[0]:.3f} ms" for k, v in kt.items()) + (f", surrogate {st}" if st else ""))

    # beside the headline: the impact-start window (steps W+1 .. W+K) at 190 m/s
    impact = {}
    v_unit = prob.vinit / V0
    prob.vinit = v_unit * V0_IMPACT
    for name in ([] if args.no_impact else modes):
        wall, _, _, _, _, st = timed_run(prob, name, args.steps, args.warmup, 0)
        impact[name] = {"rate": n * args.steps / wall, "ms_per_step": 1e3 * wall / args.steps,
                        "surrogate_stats": st}
        log(f"{name} (impact start, steps {args.warmup + 1}-{args.warmup + args.steps}, {V0_IMPACT:.0f} m/s): "
            f"{impact[name]['ms_per_step']:.3f} ms/step")
    prob.vinit = v_unit * V0

    head = "surrogate" if "surrogate" in res else modes[-1]
    top = res[head]
    mixed_prec = head != "fp64"
    nnz = top["nnz"]
    # dominant kernel of the reported step; its mean duration from the events
    # of the timed region
    kname, (kms, _) = max(top["kt"].items(), key=lambda kv: kv[1][0] * kv[1][1])
    if kname in top["kt_timed"]:
        kms = top["kt_timed"][kname][0]
    kb = kernel_bytes(kname, n, nnz, mixed_prec)
    achieved = kb / (kms * 1e-3)
    traffic = read_traffic(kname, mixed_prec)

    semi = None if args.no_semi else semi_lagrangian(min(args.steps, 50), 3)
    pd = None if args.no_pd else pd_plate(min(args.steps, 20), 3)
    pre = None if args.no_semi else preprocessing()
    if pre:
        log("preprocessing:", pre)

    validation = None
    if not args.no_validate:
        validation = validate(prob, args.stable_steps)
        log("validation:", json.dumps(validation))

    cpu = None
    if not args.no_cpu:
        cores = max(1, min(16, os.cpu_count() or 1))
        try:
            cpu = cpu_baseline_reference(2.54e-3, cores)
        except Exception as e:   # the reference program is only a reported baseline
            log("reference CPU baseline failed:", e)
        if cpu is None:
            cpu = cpu_baseline_port(2.54e-3, cores)
        log("cpu_baseline:", cpu)

    st = top["stats"]
    out = {
        "metric": "node-steps/sec, surrogate+FP32 vs plain FP64, final length err <=0.5%",
        "value": top["rate"],
        "unit": "node-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1e3 * top["wall"] / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32+f64, bf16 network" if head == "surrogate" else ("f32+f64" if mixed_prec else "f64"),
        "data": "synthetic",
        "config": {"workload": f"Taylor bar 25.4 mm x r 3.8 mm, h 0.088 mm, jitter 5% seed 2, copper J2 linear "
                               f"hardening, {V0:.0f} m/s, Lagrangian DNI, {head}",
                   "nodes": n, "pairs": nnz, "mean_neighbors": nnz / n, "max_neighbors": top["max_neighbors"],
                   "dt_s": prob.dt, "surrogate": head == "surrogate", "v0_m_s": V0,
                   "timed_window": f"steps {first}-{first + args.steps - 1} after {prelude} untimed steps into the "
                                   f"plastic regime and {args.warmup} warm-up steps"},
        "rates": {m: r["rate"] for m, r in res.items()},
        "ms_per_step_by_mode": {m: 1e3 * r["wall"] / args.steps for m, r in res.items()},
        "kernel_ms": {m: {k: v[0] for k, v in r["kt"].items()} for m, r in res.items()},
        "surrogate_stats": st,
        "surrogate_answered_frac": ((st["evaluated"] - st["rejected"]) / max(1, st["plastic"])) if st else None,
        "step_roofline_frac": {m: r["rate"] * node_step_bytes(n, nnz, m != "fp64") / PEAK_HBM
                               for m, r in res.items()},
        "roofline": {"bound": "hbm", "kernel": kname, "achieved": achieved / 1e9, "peak": PEAK_HBM / 1e9,
                     "unit": "GB/s", "frac": achieved / PEAK_HBM, "traffic": traffic,
                     "algorithmic_bytes": kb},
        "impact_start_window": {"v0_m_s": V0_IMPACT, "steps": f"{args.warmup + 1}-{args.warmup + args.steps}",
                                "modes": impact},
        "semi_lagrangian": semi,
        "pd_plate": pd,
        "preprocessing": pre,
        "validation": validation,
        "cpu_baseline": cpu,
    }
    if "fp64" in res:
        out["value_over_fp64"] = top["rate"] / res["fp64"]["rate"]
    print(json.dumps(out))


if __name__ == "__main__":
    main()
