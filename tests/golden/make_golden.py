"""Write the golden fixtures under tests/golden/ from the reference program.

Runs only where the reference was compiled (oracle/_ref, built from the
reference's own sources by oracle/Makefile): calls the reference subroutines
through the bind(C) harness and runs the reference program MEGA on synthetic
decks written by taylor_amd.deck, then stores inputs and outputs as small
compressed .npz files.  The fixtures are data only.

    python tests/golden/make_golden.py            # dense, geometry, search/RK/FINT, runs
    python tests/golden/make_golden.py --pd       # peridynamics vectors and run
    python tests/golden/make_golden.py --pd-auto  # peridynamics with the automatic time step
    python tests/golden/make_golden.py --final    # the whole 80 us bar at 0.5 mm
    python tests/golden/make_golden.py --damage   # material types 6 and 3
    python tests/golden/make_golden.py --pd-damage   # peridynamics with material types 6 and 3
    python tests/golden/make_golden.py --vtk      # the reference's VTK frames of a small deck
    python tests/golden/make_golden.py --trajectory [--semi]   # 181k-node bar, 1,400 / 500 steps, 8 threads
"""
from __future__ import annotations

import ctypes as C
import glob
import os
import subprocess
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "tests"))

from damage_util import damage_material  # noqa: E402
from oracle_bind import REF_MEGA, load_ref, ref_set_control  # noqa: E402
from taylor_amd import deck  # noqa: E402
from vtk_read import read_unformatted_vtk  # noqa: E402

H, LEN = 0.7e-3, 6e-3


def small(lagrangian=1, **kw):
    return deck.taylor_problem(H, length=LEN, seed=0, lagrangian=lagrangian, pdsearch_steps=3, **kw)


def dense_vectors(ref, rng):
    out = {}
    n = 400
    xs = rng.uniform(0, 1.2, n)
    k = np.zeros((n, 2))
    for i, x in enumerate(xs):
        a, b = C.c_double(), C.c_double()
        ref.ref_mls_kernel0(float(x), 1.0, 3, C.byref(a), C.byref(b))
        k[i] = a.value, b.value
    out["kernel_x"], out["kernel_out"] = xs, k
    A4 = rng.normal(size=(n, 16)) + 4 * np.eye(4).ravel()
    out["m44_in"] = A4
    out["m44_out"] = np.array([_call(ref.ref_m44inv, a, 16) for a in A4])
    A3 = rng.normal(size=(n, 9))
    out["inv3_in"] = A3
    out["inv3_out"] = np.array([_call(ref.ref_inv3, a, 9) for a in A3])
    out["inverse_out"] = np.array([_call(lambda a, b: ref.ref_inverse(a, 3, b), a, 9) for a in A3])
    det = []
    for a in A3:
        d = C.c_double()
        ref.ref_determinant(np.ascontiguousarray(a), C.byref(d))
        det.append(d.value)
    out["det_out"] = np.array(det)
    L = rng.normal(size=(n, 9)) * 0.02
    out["hw_in"] = L
    hw = []
    for l in L:
        R, S, D = np.zeros(9), np.zeros(6), np.zeros(6)
        ref.ref_hughes_winget(np.ascontiguousarray(l), R, S, D)
        hw.append(np.concatenate([R, S, D]))
    out["hw_out"] = np.array(hw)
    V = rng.normal(size=(n, 6))
    rot = []
    for r, v in zip(out["hw_out"][:, :9], V):
        w = v.copy()
        ref.ref_rotate_tensor(np.ascontiguousarray(r), w)
        rot.append(w)
    out["rot_in"], out["rot_out"] = V, np.array(rot)
    # J2 return maps for the linear law (config materials) and the shipped power law
    for tag, mat in (("lin", deck.copper_linear()), ("pow", deck.Material.j2(78.2e9, 0.3, 2700.0, 0.29e9, 125.0, 0.1))):
        pred = rng.normal(size=(n, 6)) * mat.props[3] * rng.uniform(0.1, 3.0, size=(n, 1))
        eps0 = np.abs(rng.normal(size=n)) * 0.1
        res = []
        for p, e in zip(pred, eps0):
            s = np.zeros(6)
            st = np.zeros(20)
            st[2] = e
            ref.ref_von_mises(s, np.zeros(6), np.ascontiguousarray(p), st, mat.props)
            res.append(np.concatenate([s, [st[2]]]))
        out[f"vm_{tag}_props"], out[f"vm_{tag}_pred"], out[f"vm_{tag}_eps0"] = mat.props, pred, eps0
        out[f"vm_{tag}_out"] = np.array(res)
    return out


def _call(f, a, m):
    o = np.zeros(m)
    f(np.ascontiguousarray(a), o)
    return o


def geometry_vectors(ref):
    mesh = deck.taylor_bar_mesh(H, length=LEN, seed=0)
    n, ne = mesh.coo.shape[0], mesh.elcon.shape[0]
    ad, vol, sm, dm = np.zeros((n, 3)), np.zeros(n), np.zeros((n, 6)), np.zeros(3)
    ref.ref_preprocess(n, ne, np.ascontiguousarray(mesh.elcon), np.ascontiguousarray(mesh.coo), ad, vol, sm, dm)
    return {"coo": mesh.coo, "elcon": mesh.elcon, "adial": ad, "vol": vol, "sm_len": sm, "dist_max": dm}


def search_rk_fint_vectors(ref, rng):
    out = {}
    prob = small()
    ref_set_control(ref, prob)
    n = prob.n
    coo, win = np.ascontiguousarray(prob.coo), np.ascontiguousarray(prob.win)
    d = prob.dist_max
    geo, nb = np.zeros(6), np.zeros(3, np.int32)
    ref.ref_bin_geometry(n, coo, win, float(d[0]), float(d[1]), float(d[2]), geo, nb)
    ijk = np.zeros((n, 3), np.int32)
    ref.ref_get_bins(n, coo, ijk)
    sm = np.zeros((n, 6))

    def soft(x):
        dim = n * 200
        gn, gs, st = np.zeros(n, np.int32), np.zeros(n, np.int32), np.zeros(dim, np.int32)
        mx, nz = C.c_int(), C.c_int()
        ref.ref_soft_search(n, np.ascontiguousarray(x), win, sm, float(d[0]), float(d[1]), float(d[2]), gn, gs, dim, st,
                            C.byref(mx), C.byref(nz))
        return gn, gs, st[:nz.value]

    gn, gs, st = soft(coo)
    out.update(bin_geom=geo, bin_nb=nb, bin_ijk=ijk, gn=gn, gstart=gs, gstack=st)
    # stale bins: nodes moved, bins kept (HANDELER between PDSEARCH rebuilds)
    xm = coo + rng.normal(size=coo.shape) * 0.3 * H
    gn2, gs2, st2 = soft(xm)
    out.update(x_moved=xm, gn_stale=gn2, gstart_stale=gs2, gstack_stale=st2)
    # RK1 at every node (QL on, singular kernel at the wall)
    ebc = np.ascontiguousarray(prob.ebc_node, dtype=np.int32)
    shp, dshp = [], []
    for i in range(n):
        a, b = gs[i] - 1, gs[i] - 1 + gn[i]
        ls = np.ascontiguousarray(st[a:b])
        s1, d1 = np.zeros(gn[i]), np.zeros(3 * gn[i])
        ref.ref_rk1(coo[i].copy(), n, coo, win, ls, int(gn[i]), ebc, int(ebc[i]), s1, d1)
        shp.append(s1)
        dshp.append(d1.reshape(-1, 3))
    shp, dshp = np.concatenate(shp), np.concatenate(dshp)
    out.update(rk_shp=shp, rk_dshp=dshp)
    # one CONSTRUCT_FINT call per formulation on a yielding random state
    nnz = int(gn.sum())
    for lag in (1, 0):
        p2 = small(lagrangian=lag)
        ref_set_control(ref, p2)
        dtot = rng.normal(size=(n, 3)) * 2e-6
        dinc = rng.normal(size=(n, 3)) * 4e-7 + np.array([0, 0, -1e-6])
        stress = rng.normal(size=(n, 6)) * 3e8
        strain = rng.normal(size=(n, 6)) * 1e-3
        state = np.zeros((n, 20))
        state[:, 2] = np.abs(rng.normal(size=n)) * 0.02
        x_cur = coo + dtot
        s, e, t = stress.copy(), strain.copy(), state.copy()
        f, fx, wv, seq = np.zeros((n, 3)), np.zeros((n, 3)), np.zeros(n), np.zeros(n)
        w = C.c_double(0)
        props = np.ascontiguousarray(p2.props_table()[p2.mat_id])
        ref.ref_construct_fint(n, win, np.ascontiguousarray(p2.vol), coo, np.ascontiguousarray(x_cur), props, t, s, e,
                               np.ascontiguousarray(dinc), np.ascontiguousarray(dtot), np.full(n, 2, np.int32), ebc,
                               gn, gs, nnz, st, shp.copy(), np.ascontiguousarray(dshp).copy(), int(gn.max()), 0,
                               f, fx, wv, np.zeros((n, 3)), np.ones(n, np.int32), seq, C.byref(w), p2.dt)
        out.update({f"fint{lag}_dtot": dtot, f"fint{lag}_dinc": dinc, f"fint{lag}_stress0": stress,
                    f"fint{lag}_strain0": strain, f"fint{lag}_eps0": state[:, 2].copy(), f"fint{lag}_f": f,
                    f"fint{lag}_stress": s, f"fint{lag}_strain": e, f"fint{lag}_eps": t[:, 2].copy(),
                    f"fint{lag}_work": np.array([w.value]), f"fint{lag}_wave": wv, f"fint{lag}_strain_eq": seq})
    return out


def pd_vectors(ref, rng):
    """Peridynamics (CONSTRUCT_FINT_PD) on the small bar: UDFM_SHAPE_TENSOR at
    every node in SI units (the shape tensor is ~1e-16 there, so INVERSE stops
    early and returns it unreduced) and in mm units, and one CONSTRUCT_FINT_PD
    call per LINIT value on a yielding random state in mm units."""
    out = {}
    for tag, mm in (("si", False), ("mm", True)):
        prob = small(peridynamics=1, mm_units=mm)
        ref_set_control(ref, prob)
        n = prob.n
        coo, win, vol = (np.ascontiguousarray(prob.coo), np.ascontiguousarray(prob.win),
                         np.ascontiguousarray(prob.vol))
        d = prob.dist_max
        geo, nb = np.zeros(6), np.zeros(3, np.int32)
        ref.ref_bin_geometry(n, coo, win, float(d[0]), float(d[1]), float(d[2]), geo, nb)
        ijk = np.zeros((n, 3), np.int32)
        ref.ref_get_bins(n, coo, ijk)
        dim = n * 200
        gn, gs, st = np.zeros(n, np.int32), np.zeros(n, np.int32), np.zeros(dim, np.int32)
        mx, nz = C.c_int(), C.c_int()
        ref.ref_soft_search(n, coo, win, np.zeros((n, 6)), float(d[0]), float(d[1]), float(d[2]), gn, gs, dim, st,
                            C.byref(mx), C.byref(nz))
        st = st[:nz.value]
        ebc = np.ascontiguousarray(prob.ebc_node, dtype=np.int32)
        w, invk = [], []
        for i in range(n):
            a, b = gs[i] - 1, gs[i] - 1 + gn[i]
            s1, k1 = np.zeros(gn[i]), np.zeros(9)
            ref.ref_udfm(coo[i].copy(), n, coo, vol, win, np.ascontiguousarray(st[a:b]), int(gn[i]), ebc, int(ebc[i]),
                         s1, k1)
            w.append(s1)
            invk.append(k1)
        w, invk = np.concatenate(w), np.array(invk)
        out.update({f"{tag}_gn": gn, f"{tag}_gstack": st, f"{tag}_omega": w, f"{tag}_invk": invk})
        if not mm:
            continue
        nnz = int(gn.sum())
        props = np.ascontiguousarray(prob.props_table()[prob.mat_id])
        L = 0.7   # lattice spacing in the deck's units
        sig0 = prob.materials[0].props[3]
        for linit in (1, 0):
            dtot = rng.normal(size=(n, 3)) * 2e-3 * L
            dinc = rng.normal(size=(n, 3)) * 4e-4 * L + np.array([0, 0, -1e-3 * L])
            stress = rng.normal(size=(n, 6)) * 0.75 * sig0
            strain = rng.normal(size=(n, 6)) * 1e-3
            state = np.zeros((n, 20))
            state[:, 2] = np.abs(rng.normal(size=n)) * 0.02
            s, e, t = stress.copy(), strain.copy(), state.copy()
            ww = np.zeros(nnz) if linit else w.copy()
            kk = np.zeros((n, 9)) if linit else invk.copy()
            f, wv = np.zeros((n, 3)), np.zeros(n)
            ref.ref_construct_fint_pd(n, win, vol, coo, np.ascontiguousarray(coo + dtot), props, t, s, e,
                                      np.ascontiguousarray(dinc), np.ascontiguousarray(dtot), np.full(n, 2, np.int32),
                                      ebc, gn, gs, nnz, st, ww, kk, int(gn.max()), linit, f, wv, prob.dt)
            out.update({f"pd{linit}_dtot": dtot, f"pd{linit}_dinc": dinc, f"pd{linit}_stress0": stress,
                        f"pd{linit}_strain0": strain, f"pd{linit}_eps0": state[:, 2].copy(), f"pd{linit}_f": f,
                        f"pd{linit}_stress": s, f"pd{linit}_strain": e, f"pd{linit}_eps": t[:, 2].copy(),
                        f"pd{linit}_wave": wv})
    return out


def mega_run(prob, mesh, nsteps, frames, workdir, extra=()):
    """Run the reference program on the deck; return the chosen frames."""
    kfile = os.path.join(workdir, "bar.k")
    deck.write_k(kfile, mesh)
    prob.time_end = prob.dt * (nsteps - 0.5)
    deck.write_control(os.path.join(workdir, "Control.dat"), kfile, prob, ["BAR", "WALL"], time_output=-1.0,
                       out_vars="unformatted\ndisplacement\nvelocity\nfint\nstress\nstrain\neps" +
                       "".join("\n" + v for v in extra))
    with open(os.path.join(workdir, "mega.out"), "w") as fo:
        subprocess.run([REF_MEGA, "-j", "1"], cwd=workdir, stdin=subprocess.DEVNULL, stdout=fo, stderr=fo,
                       check=True, timeout=3600)
    out = {}
    for k in frames:
        fr = read_unformatted_vtk(os.path.join(workdir, "MEGAOUT_%04d.vtk" % k))
        out[f"step{k}_x"] = fr["points"]
        out[f"step{k}_u"] = np.stack([fr["displacement_x"], fr["displacement_y"], fr["displacement_z"]], 1)
        out[f"step{k}_vphy"] = np.stack([fr["velocity_x"], fr["velocity_y"], fr["velocity_z"]], 1)
        out[f"step{k}_fint"] = np.stack([fr["fint_x"], fr["fint_y"], fr["fint_z"]], 1)
        out[f"step{k}_stress"] = np.stack([fr[f"stress_{c}"] for c in ("xx", "yy", "zz", "yz", "xz", "xy")], 1)
        out[f"step{k}_eps"] = fr["eps"]
        for v in extra:
            out[f"step{k}_{v}"] = fr[v]
    return out


def damage_vectors(ref, rng):
    """VON_MISES_DAM and DRUCK_PRAG of the compiled reference on random
    states spanning their branches (elastic / plastic or sheared, damage
    below threshold / growing / capped, CE = 1 and a power law, Q_PSI 0 and
    not)."""
    out = {}
    n = 400
    for tag, fn in (("dam6", ref.ref_von_mises_dam), ("dp3", ref.ref_druck_prag)):
        P, S0, T, I, ST0, S1, ST1 = [], [], [], [], [], [], []
        for t in range(n):
            if tag == "dam6":
                p = deck.Material.j2_damage(E=117e9, nu=0.35, rho=8930.0, sig0=400e6, B=rng.uniform(0, 3),
                                            CE=(1.0, 0.3)[t % 2], k_i=rng.uniform(0.01, 0.1),
                                            k_c=rng.uniform(0.2, 0.6), dam_max=rng.uniform(0.5, 0.99)).props
                scale = 4e8
            else:
                p = deck.Material.drucker_prager(E=30e9, nu=0.2, rho=2400.0, dp_type=1 + t % 3,
                                                 phi_deg=rng.uniform(10, 45), c=rng.uniform(1e6, 2e7),
                                                 psi_deg=0.0, t_cut=3e6, k_i=rng.uniform(0.001, 0.01),
                                                 k_c=rng.uniform(0.02, 0.1), dam_max=rng.uniform(0.5, 0.99)).props
                if t % 4 == 3:
                    p[26] = rng.uniform(0.05, 0.3)   # a dilatant cone, beyond what the reader prepares
                scale = 3e7
            s = rng.normal(size=6) * scale * rng.uniform(0.1, 2)
            tot = rng.normal(size=6) * rng.uniform(0, 0.4 if tag == "dam6" else 0.08)
            inc = rng.normal(size=6) * 1e-3 * rng.uniform(0, 3)
            st = np.zeros(20)
            st[0], st[1], st[2] = rng.uniform(0, 0.1), rng.uniform(0, 0.1), rng.uniform(0, 0.5)
            st[3] = rng.uniform(0, 0.5) * (t % 2)
            st[4] = rng.uniform(0, 0.2) * (t % 3 == 0)
            s1, st1 = s.copy(), st.copy()
            fn(s1, np.ascontiguousarray(tot), np.ascontiguousarray(inc), np.zeros(6), st1, np.ascontiguousarray(p))
            P.append(p); S0.append(s); T.append(tot); I.append(inc); ST0.append(st); S1.append(s1); ST1.append(st1)
        for k, v in (("props", P), ("stress0", S0), ("total", T), ("inc", I), ("state0", ST0), ("stress", S1),
                     ("state", ST1)):
            out[f"{tag}_{k}"] = np.array(v)
    return out


# the deck of the VTK frame fixture: a small type-6 bar, frames every 5 steps
VTK_VARS = ("unformatted", "node_id", "displacement", "velocity", "acceleration", "fint", "bodyid", "material_type",
            "initial_coordinate", "init_velocity", "normalized_window", "nodal_volume", "nodal_mass",
            "Poissons_ratio", "Youngs_modulus", "density", "physical_window", "max_wave_vel", "stress", "strain",
            "eps", "eps_shear", "eps_vol", "dam")


def vtk_deck(workdir):
    """Write the frame fixture's deck into workdir; returns the problem."""
    h, radius, length = 1.0e-3, 2.0e-3, 3.0e-3
    mesh = deck.taylor_bar_mesh(h, radius=radius, length=length, seed=0)
    prob = deck.taylor_problem(h, radius=radius, length=length, seed=0, material=damage_material(6))
    kfile = os.path.join(workdir, "bar.k")
    deck.write_k(kfile, mesh)
    prob.time_end = prob.dt * 11.5
    deck.write_control(os.path.join(workdir, "Control.dat"), kfile, prob, ["BAR", "WALL"],
                       time_output=prob.dt * 3.5, out_vars="\n".join(VTK_VARS))
    return prob


def vtk_main():
    with tempfile.TemporaryDirectory() as wd:
        vtk_deck(wd)
        with open(os.path.join(wd, "mega.out"), "w") as fo:
            subprocess.run([REF_MEGA, "-j", "1"], cwd=wd, stdin=subprocess.DEVNULL, stdout=fo, stderr=fo,
                           check=True, timeout=600)
        frames = sorted(glob.glob(os.path.join(wd, "MEGAOUT_*.vtk")))
        np.savez_compressed(os.path.join(HERE, "vtk_frames.npz"),
                            names=np.array([os.path.basename(f) for f in frames]),
                            texts=np.array([open(f).read() for f in frames]))
        print([os.path.basename(f) for f in frames])


def damage_main():
    ref = load_ref()
    rng = np.random.default_rng(20260304)
    np.savez_compressed(os.path.join(HERE, "damage.npz"), **damage_vectors(ref, rng))
    mesh = deck.taylor_bar_mesh(H, length=LEN, seed=0)
    for mt, name in ((6, "run_dam6.npz"), (3, "run_dp3.npz")):
        prob = small(material=damage_material(mt))
        with tempfile.TemporaryDirectory() as wd:
            np.savez_compressed(os.path.join(HERE, name),
                                **mega_run(prob, mesh, 40, (1, 10, 40), wd, extra=("dam", "eps_shear")))


def final_state_vectors(workdir):
    """The CPU-runnable Taylor bar (spacing 0.5 mm, seed 0, 190 m/s, 80 us,
    Lagrangian FP64) run to the end by the reference program with one thread;
    the last frame's positions, displacement and plastic strain."""
    mesh = deck.taylor_bar_mesh(0.5e-3, seed=0)
    prob = deck.taylor_problem(0.5e-3, seed=0)
    nsteps = int(np.floor(80e-6 / prob.dt)) + 1
    kfile = os.path.join(workdir, "bar.k")
    deck.write_k(kfile, mesh)
    prob.time_end = prob.dt * (nsteps - 0.5)
    deck.write_control(os.path.join(workdir, "Control.dat"), kfile, prob, ["BAR", "WALL"],
                       time_output=prob.dt * (nsteps - 1.5),
                       out_vars="unformatted\ndisplacement\neps")
    with open(os.path.join(workdir, "mega.out"), "w") as fo:
        subprocess.run([REF_MEGA, "-j", "1"], cwd=workdir, stdin=subprocess.DEVNULL, stdout=fo, stderr=fo,
                       check=True, timeout=7200)
    last = sorted(glob.glob(os.path.join(workdir, "MEGAOUT_*.vtk")))[-1]
    fr = read_unformatted_vtk(last)
    return {"nsteps": np.array([nsteps]), "frame": np.array([int(last[-8:-4])]), "x": fr["points"],
            "u": np.stack([fr["displacement_x"], fr["displacement_y"], fr["displacement_z"]], 1),
            "eps": fr["eps"]}


def trajectory_vectors(workdir, lagrangian, nsteps=1700, every=100, threads=8):
    """The 181k-node Taylor bar (spacing 0.19 mm, jitter seed 1, 190 m/s, copper J2
    linear hardening) run by the reference program for ``nsteps`` steps with a
    full-precision frame every ``every`` steps, Lagrangian or semi-Lagrangian
    with bins rebuilt every 20 steps (PDSEARCH = ceiling(PDSTIME/DLT),
    MAIN.F90:430).  Stored per frame: the number of steps done, the bar's length max z (the wall plane is
    z = 0), the impact face's radius, the largest physical speed, the
    volume-weighted L2 norm of the plastic strain and the positions of every
    353rd node; per step, the kinetic energy MEGA logs to
    ENERGY_INFO.OUT (MAIN.F90:245,555, five digits); the number of steps finished and, if the
    program stopped itself, EXIT_PROGRAM's reason.  The run is left to go
    where it goes."""
    h = 0.19e-3
    mesh = deck.taylor_bar_mesh(h, seed=1)
    prob = deck.taylor_problem(h, seed=1, lagrangian=lagrangian, pdsearch_steps=20)
    kfile = os.path.join(workdir, "bar.k")
    done_mark = os.path.join(workdir, "finished")
    rc = 0
    if os.path.exists(done_mark):   # a finished run in a kept working directory is read again
        rc = int(open(done_mark).read())
    else:
        deck.write_k(kfile, mesh)
        prob.time_end = prob.dt * (nsteps - 0.5)
        deck.write_control(os.path.join(workdir, "Control.dat"), kfile, prob, ["BAR", "WALL"],
                           time_output=prob.dt * (every - 0.5), out_vars="unformatted\ndisplacement\nvelocity\neps")
        with open(os.path.join(workdir, "mega.out"), "w") as fo:
            rc = subprocess.run([REF_MEGA, "-j", str(threads)], cwd=workdir, stdin=subprocess.DEVNULL, stdout=fo,
                                stderr=fo, timeout=4 * 3600).returncode
        with open(done_mark, "w") as f:
            f.write(str(rc))
    # MEGA prints exodusStep and STEPS at a frame; STEPS counts from 0 and is
    # raised after the output (MAIN.F90:274,632-647), so the frame holds STEPS + 1 steps
    steps = {}
    with open(os.path.join(workdir, "mega.out")) as f:
        for ln in f:
            p = ln.split()
            if len(p) >= 4 and p[0].isdigit() and p[1].isdigit():
                steps[int(p[0])] = int(p[1]) + 1
    wall = mesh.sets["WALL"] - 1
    sample = np.arange(0, prob.n, 353)
    rows = {"step": [], "length": [], "radius": [], "vmax": [], "eps_l2": [], "x_sample": []}
    for k in sorted(steps):
        path = os.path.join(workdir, "MEGAOUT_%04d.vtk" % k)
        if k == 0 or not os.path.exists(path):
            continue
        fr = read_unformatted_vtk(path)
        x = fr["points"]
        v = np.stack([fr["velocity_x"], fr["velocity_y"], fr["velocity_z"]], 1)
        rows["step"].append(steps[k])
        rows["length"].append(x[:, 2].max())
        rows["radius"].append(np.sqrt(x[wall, 0] ** 2 + x[wall, 1] ** 2).max())
        rows["vmax"].append(np.sqrt((v * v).sum(1)).max())
        rows["eps_l2"].append(np.sqrt((prob.vol * fr["eps"] ** 2).sum()))
        rows["x_sample"].append(x[sample])
    ke = []   # ENERGY_INFO.OUT: STEPS, internal, kinetic, external, total energy (MAIN.F90:245,555)
    with open(os.path.join(workdir, "ENERGY_INFO.OUT")) as f:
        for ln in f:
            p = ln.split()
            if len(p) == 5 and p[0].isdigit():
                ke.append((int(p[0]), float(p[2])))
    with open(os.path.join(workdir, "TIME.OUT")) as f:   # one line per finished step (MAIN.F90:653)
        steps_done = sum(1 for ln in f if ln.strip())
    with open(os.path.join(workdir, "mega.out")) as f:
        txt = f.read()
    reason = ""   # EXIT_PROGRAM's reason line, if the program stopped itself (EXIT_PROGRAM.F90)
    if "FATAL ERROR" in txt:
        tail = [ln.strip() for ln in txt[txt.index("FATAL ERROR"):].splitlines()]
        reason = " ".join(t for t in tail if t and not t.startswith("*") and "FATAL" not in t)[:200]
    out = {k: np.array(v) for k, v in rows.items()}
    out.update({"sample": sample, "ke_step": np.array([s for s, _ in ke], dtype=np.int64),
                "ke": np.array([e for _, e in ke]), "n": np.array([prob.n]), "dt": np.array([prob.dt]),
                "lagrangian": np.array([lagrangian]), "pdsearch": np.array([prob.pdsearch]),
                "nsteps": np.array([nsteps]), "returncode": np.array([rc]),
                "steps_done": np.array([steps_done]), "stop_reason": np.array(reason)})
    return out


def pd_main():
    ref = load_ref()
    rng = np.random.default_rng(20260203)
    with tempfile.TemporaryDirectory() as wd:
        cwd = os.getcwd()
        os.chdir(wd)   # the reference's WARN appends to MEGA.log and WARN.log in the working directory
        open("MEGA.log", "a").close()
        open("WARN.log", "a").close()
        try:
            np.savez_compressed(os.path.join(HERE, "pd.npz"), **pd_vectors(ref, rng))
        finally:
            os.chdir(cwd)
    prob = small(peridynamics=1, mm_units=True)
    mesh = deck.taylor_bar_mesh(H * 1e3, radius=3.8e-3 * 1e3, length=LEN * 1e3, seed=0)
    assert np.array_equal(mesh.coo, prob.coo)
    with tempfile.TemporaryDirectory() as wd:
        np.savez_compressed(os.path.join(HERE, "run_pd.npz"), **mega_run(prob, mesh, 40, (1, 10, 40), wd))


def pd_damage_main():
    """The small peridynamic bar (mm deck) with the damage materials, 40 steps
    of the reference program, frames 1, 10 and 40 with damage and shear strain."""
    mesh = deck.taylor_bar_mesh(H * 1e3, radius=3.8e-3 * 1e3, length=LEN * 1e3, seed=0)
    for mt, name in ((6, "run_pd_dam6.npz"), (3, "run_pd_dp3.npz")):
        prob = small(peridynamics=1, mm_units=True, material=damage_material(mt))
        assert np.array_equal(mesh.coo, prob.coo) and prob.mat_types()[0] == mt
        with tempfile.TemporaryDirectory() as wd:
            np.savez_compressed(os.path.join(HERE, name),
                                **mega_run(prob, mesh, 40, (1, 10, 40), wd, extra=("dam", "eps_shear")))
    # the automatic step with the Drucker-Prager wave-speed factor (FINT_PD.F90:539-542)
    prob = small(peridynamics=1, mm_units=True, material=damage_material(3))
    np.savez_compressed(os.path.join(HERE, "run_pd_dp3_auto.npz"),
                        **pd_auto_run(prob, PD_DLT_FAC, extra=("dam", "eps_shear")))


PD_DLT_FAC = 0.3


def pd_auto_main():
    """The automatic time step (IAUTO_TS = 1) of CONSTRUCT_FINT_PD: DLT_FINT of
    the reference's call on pd.npz's deformed state, and 40 steps of the
    reference program on the small peridynamic bar (mm deck, DLT_FAC 0.3) with
    every step's TIME (TIME.OUT, MAIN.F90:653) and frames 1, 10 and 40.  The
    deck's TIME_END lies midway between the oracle's times after steps 39 and
    40; the program's own TIME.OUT says how many steps it took."""
    ref = load_ref()
    g = np.load(os.path.join(HERE, "pd.npz"))
    out = {}
    with tempfile.TemporaryDirectory() as wd:
        cwd = os.getcwd()
        os.chdir(wd)   # the reference's WARN appends to MEGA.log and WARN.log in the working directory
        open("MEGA.log", "a").close()
        open("WARN.log", "a").close()
        try:
            prob = small(peridynamics=1, mm_units=True)
            ref_set_control(ref, prob)
            ref.ref_set_auto_ts(1, PD_DLT_FAC)
            n = prob.n
            coo, win, vol = (np.ascontiguousarray(prob.coo), np.ascontiguousarray(prob.win),
                             np.ascontiguousarray(prob.vol))
            gn, st = g["mm_gn"], g["mm_gstack"]
            gs = (np.concatenate([[0], np.cumsum(gn)[:-1]]) + 1).astype(np.int32)
            nnz = int(gn.sum())
            props = np.ascontiguousarray(prob.props_table()[prob.mat_id])
            ebc = np.ascontiguousarray(prob.ebc_node, dtype=np.int32)
            state = np.zeros((n, 20))
            state[:, 2] = g["pd0_eps0"]
            s, e = g["pd0_stress0"].copy(), g["pd0_strain0"].copy()
            dtot = g["pd0_dtot"]
            f, wv = np.zeros((n, 3)), np.zeros(n)
            ref.ref_construct_fint_pd(n, win, vol, coo, np.ascontiguousarray(coo + dtot), props, state, s, e,
                                      np.ascontiguousarray(g["pd0_dinc"]), np.ascontiguousarray(dtot),
                                      np.full(n, 2, np.int32), ebc, np.ascontiguousarray(gn), gs, nnz,
                                      np.ascontiguousarray(st), g["mm_omega"].copy(),
                                      np.ascontiguousarray(g["mm_invk"]).copy(), int(gn.max()), 0, f, wv, prob.dt)
            out.update({"call_dlt_fac": np.array([PD_DLT_FAC]), "call_dlt": np.array([ref.ref_last_dlt_fint()]),
                        "call_wave": wv, "call_f": f})
        finally:
            os.chdir(cwd)
    out.update(pd_auto_run(small(peridynamics=1, mm_units=True), PD_DLT_FAC))
    np.savez_compressed(os.path.join(HERE, "pd_auto.npz"), **out)


def pd_auto_run(prob, dlt_fac, extra=()):
    """40 steps of the reference program on the small peridynamic bar (mm deck)
    with IAUTO_TS = 1: TIME_END midway between the oracle's times after steps
    39 and 40, every step's TIME from TIME.OUT, frames 1, 10 and 40."""
    from oracle_bind import OracleRun, load_oracle
    run = OracleRun(load_oracle(), prob, dlt_fac=dlt_fac).step(40)
    t = np.cumsum(run.dlt)
    mesh = deck.taylor_bar_mesh(H * 1e3, radius=3.8e-3 * 1e3, length=LEN * 1e3, seed=0)
    out = {}
    with tempfile.TemporaryDirectory() as wd:
        kfile = os.path.join(wd, "bar.k")
        deck.write_k(kfile, mesh)
        prob.time_end = 0.5 * (t[38] + t[39])
        deck.write_control(os.path.join(wd, "Control.dat"), kfile, prob, ["BAR", "WALL"], time_output=-1.0,
                           out_vars="unformatted\ndisplacement\nvelocity\nfint\nstress\nstrain\neps" +
                           "".join("\n" + v for v in extra), dlt_fac=dlt_fac)
        with open(os.path.join(wd, "mega.out"), "w") as fo:
            subprocess.run([REF_MEGA, "-j", "1"], cwd=wd, stdin=subprocess.DEVNULL, stdout=fo, stderr=fo,
                           check=True, timeout=3600)
        with open(os.path.join(wd, "TIME.OUT")) as f:
            times = [float(ln.split()[0].replace("D", "E")) for ln in f if ln.strip()]
        out.update({"run_dlt_fac": np.array([dlt_fac]), "run_time_end": np.array([prob.time_end]),
                    "run_time": np.array(times)})
        for k in (1, 10, 40):
            fr = read_unformatted_vtk(os.path.join(wd, "MEGAOUT_%04d.vtk" % k))
            out[f"step{k}_x"] = fr["points"]
            out[f"step{k}_u"] = np.stack([fr["displacement_x"], fr["displacement_y"], fr["displacement_z"]], 1)
            out[f"step{k}_vphy"] = np.stack([fr["velocity_x"], fr["velocity_y"], fr["velocity_z"]], 1)
            out[f"step{k}_fint"] = np.stack([fr["fint_x"], fr["fint_y"], fr["fint_z"]], 1)
            out[f"step{k}_stress"] = np.stack([fr[f"stress_{c}"] for c in ("xx", "yy", "zz", "yz", "xz", "xy")], 1)
            out[f"step{k}_eps"] = fr["eps"]
            for v in extra:
                out[f"step{k}_{v}"] = fr[v]
    return out


def main():
    if "--pd-auto" in sys.argv:
        pd_auto_main()
        return
    if "--final" in sys.argv:
        with tempfile.TemporaryDirectory() as wd:
            np.savez_compressed(os.path.join(HERE, "taylor_h05_final.npz"), **final_state_vectors(wd))
        return
    if "--pd-damage" in sys.argv:
        pd_damage_main()
        return
    if "--pd" in sys.argv:
        pd_main()
        return
    if "--trajectory" in sys.argv:
        # TRAJ_DIR keeps the working directory, so that a finished run is parsed again without rerunning
        lag = 0 if "--semi" in sys.argv else 1
        name = "traj_h019_semilag.npz" if lag == 0 else "traj_h019_lagrangian.npz"
        kw = dict(nsteps=500, every=50) if lag == 0 else dict(nsteps=1400, every=100)
        keep = os.environ.get("TRAJ_DIR")
        if keep:
            wd = os.path.join(keep, "semi" if lag == 0 else "lag")
            os.makedirs(wd, exist_ok=True)
            np.savez_compressed(os.path.join(HERE, name), **trajectory_vectors(wd, lag, **kw))
        else:
            with tempfile.TemporaryDirectory() as wd:
                np.savez_compressed(os.path.join(HERE, name), **trajectory_vectors(wd, lag, **kw))
        return
    if "--damage" in sys.argv:
        damage_main()
        return
    if "--vtk" in sys.argv:
        vtk_main()
        return
    ref = load_ref()
    rng = np.random.default_rng(20260101)
    np.savez_compressed(os.path.join(HERE, "dense.npz"), **dense_vectors(ref, rng))
    np.savez_compressed(os.path.join(HERE, "geometry.npz"), **geometry_vectors(ref))
    np.savez_compressed(os.path.join(HERE, "search_rk_fint.npz"), **search_rk_fint_vectors(ref, rng))
    mesh = deck.taylor_bar_mesh(H, length=LEN, seed=0)
    for lag, name in ((1, "run_lagrangian.npz"), (0, "run_semilagrangian.npz")):
        with tempfile.TemporaryDirectory() as wd:
            prob = small(lagrangian=lag)
            np.savez_compressed(os.path.join(HERE, name), **mega_run(prob, mesh, 40, (1, 10, 40), wd))
    for f in sorted(glob.glob(os.path.join(HERE, "*.npz"))):
        print(os.path.basename(f), os.path.getsize(f))


if __name__ == "__main__":
    main()
