"""The oracle against the golden fixtures written from the reference program
(tests/golden/make_golden.py).  With one thread the oracle keeps the
reference's operation order, so the bar is bit-exact equality."""
import ctypes as C
import os

import numpy as np
import pytest

from oracle_bind import OracleRun, control, oracle_fint, oracle_fint_pd, oracle_search, oracle_shapes, oracle_udfm
from taylor_amd import deck

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
H, LEN = 0.7e-3, 6e-3


def load(name):
    return np.load(os.path.join(GOLD, name))


def small(lagrangian=1, **kw):
    return deck.taylor_problem(H, length=LEN, seed=0, lagrangian=lagrangian, pdsearch_steps=3, **kw)


def test_dense_routines_bit_exact(oracle):
    g = load("dense.npz")
    for x, (phi, dphi) in zip(g["kernel_x"], g["kernel_out"]):
        a, b = C.c_double(), C.c_double()
        oracle.or_kernel(float(x), C.byref(a), C.byref(b))
        assert a.value == phi and b.value == dphi
    for a, r44, a3, r3, rinv, det in zip(g["m44_in"], g["m44_out"], g["inv3_in"], g["inv3_out"], g["inverse_out"], g["det_out"]):
        o = np.zeros(16)
        oracle.or_m44inv(np.ascontiguousarray(a), o)
        assert np.array_equal(o, r44)
        o = np.zeros(9)
        oracle.or_inv3(np.ascontiguousarray(a3), o)
        assert np.array_equal(o, r3)
        o = np.zeros(9)
        oracle.or_inverse(np.ascontiguousarray(a3), 3, o)
        assert np.array_equal(o, rinv)
        assert oracle.or_determinant(np.ascontiguousarray(a3)) == det
    for l, out in zip(g["hw_in"], g["hw_out"]):
        R, S, D = np.zeros(9), np.zeros(6), np.zeros(6)
        oracle.or_hughes_winget(np.ascontiguousarray(l), R, S, D)
        assert np.array_equal(np.concatenate([R, S, D]), out)
    for r, v, out in zip(g["hw_out"][:, :9], g["rot_in"], g["rot_out"]):
        w = v.copy()
        oracle.or_rotate_tensor(np.ascontiguousarray(r), w)
        assert np.array_equal(w, out)


@pytest.mark.parametrize("law", ["lin", "pow"])
def test_return_map_bit_exact(oracle, law):
    g = load("dense.npz")
    props = g[f"vm_{law}_props"]
    for p, e0, out in zip(g[f"vm_{law}_pred"], g[f"vm_{law}_eps0"], g[f"vm_{law}_out"]):
        s, st = np.zeros(6), np.zeros(20)
        st[2] = e0
        assert oracle.or_von_mises(s, np.ascontiguousarray(p), st, props) >= 0
        assert np.array_equal(s, out[:6]) and st[2] == out[6]


def test_geometry_preprocessing_bit_exact():
    g = load("geometry.npz")
    mesh = deck.Mesh(coo=g["coo"], elcon=g["elcon"], sets={})
    geo = deck.preprocess(mesh)
    assert np.array_equal(geo.win0, g["adial"])
    assert np.array_equal(geo.vol, g["vol"])
    assert np.array_equal(geo.sm_len, g["sm_len"])
    assert np.array_equal(geo.dist_max, g["dist_max"])
    # the generator itself is deterministic: same seed, same deck
    assert np.array_equal(deck.taylor_bar_mesh(H, length=LEN, seed=0).coo, g["coo"])


def test_bins_and_search_bit_exact(oracle):
    g = load("search_rk_fint.npz")
    prob = small()
    rp, col, ijk = oracle_search(oracle, prob)
    assert np.array_equal(ijk, g["bin_ijk"])
    assert np.array_equal(np.diff(rp), g["gn"])
    assert np.array_equal(rp[:-1] + 1, g["gstart"])
    assert np.array_equal(col + 1, g["gstack"])      # same entries in the same order
    # stale bins: positions moved, bin assignment kept
    rp2, col2, _ = oracle_search(oracle, prob, x=g["x_moved"], ijk=g["bin_ijk"])
    assert np.array_equal(np.diff(rp2), g["gn_stale"])
    assert np.array_equal(col2 + 1, g["gstack_stale"])


def test_rk_shapes_bit_exact(oracle):
    g = load("search_rk_fint.npz")
    prob = small()
    rp, col, _ = oracle_search(oracle, prob)
    shp, dshp = oracle_shapes(oracle, prob, rp, col)
    assert np.array_equal(shp, g["rk_shp"])
    assert np.array_equal(dshp, g["rk_dshp"])


@pytest.mark.parametrize("lag", [1, 0])
def test_construct_fint_bit_exact(oracle, lag):
    g = load("search_rk_fint.npz")
    prob = small(lagrangian=lag)
    n = prob.n
    rp, col, _ = oracle_search(oracle, prob)
    shp, dshp = g["rk_shp"].copy(), g["rk_dshp"].copy()
    s, e = g[f"fint{lag}_stress0"].copy(), g[f"fint{lag}_strain0"].copy()
    st = np.zeros((n, 20))
    st[:, 2] = g[f"fint{lag}_eps0"]
    dtot = g[f"fint{lag}_dtot"]
    f, fx, w, _ = oracle_fint(oracle, prob, rp, col, shp, dshp, 0, g[f"fint{lag}_dinc"], dtot, prob.coo + dtot, s, e, st)
    assert np.array_equal(f, g[f"fint{lag}_f"])
    assert np.array_equal(s, g[f"fint{lag}_stress"])
    assert np.array_equal(e, g[f"fint{lag}_strain"])
    assert np.array_equal(st[:, 2], g[f"fint{lag}_eps"])
    assert w == g[f"fint{lag}_work"][0]


@pytest.mark.parametrize("lag,name", [(1, "run_lagrangian.npz"), (0, "run_semilagrangian.npz")])
def test_whole_run_bit_exact(oracle, lag, name):
    """40 explicit steps of the reference program MEGA (frames 1, 10, 40)."""
    g = load(name)
    run = OracleRun(oracle, small(lagrangian=lag))
    done = 0
    for k in (1, 10, 40):
        run.step(k - done)
        done = k
        A = run.arr
        assert np.array_equal(A["x"], g[f"step{k}_x"])
        assert np.array_equal(A["dphy_tot"], g[f"step{k}_u"])
        assert np.array_equal(A["vphy"], g[f"step{k}_vphy"])
        assert np.array_equal(A["fint"], g[f"step{k}_fint"])
        assert np.array_equal(A["stress"], g[f"step{k}_stress"])
        assert np.array_equal(A["state"][:, 2], g[f"step{k}_eps"])


@pytest.mark.parametrize("tag,mm", [("si", False), ("mm", True)])
def test_pd_bonds_and_shape_tensor_bit_exact(oracle, tag, mm):
    """Bond lists (SOFT_SEARCH) and UDFM_SHAPE_TENSOR per node.  In SI units the
    shape tensor is ~1e-16, INVERSE stops at its first pivot and returns it
    unreduced; the restatement does the same."""
    g = load("pd.npz")
    prob = small(peridynamics=1, mm_units=mm)
    rp, col, _ = oracle_search(oracle, prob)
    assert np.array_equal(np.diff(rp), g[f"{tag}_gn"]) and np.array_equal(col + 1, g[f"{tag}_gstack"])
    w, invk = oracle_udfm(oracle, prob, rp, col)
    assert np.array_equal(w, g[f"{tag}_omega"])
    assert np.array_equal(invk, g[f"{tag}_invk"])


@pytest.mark.parametrize("linit", [1, 0])
def test_construct_fint_pd_bit_exact(oracle, linit):
    g = load("pd.npz")
    prob = small(peridynamics=1, mm_units=True)
    n = prob.n
    rp, col, _ = oracle_search(oracle, prob)
    w = np.zeros(int(rp[-1])) if linit else g["mm_omega"].copy()
    invk = np.zeros((n, 9)) if linit else np.ascontiguousarray(g["mm_invk"]).copy()
    s, e = g[f"pd{linit}_stress0"].copy(), g[f"pd{linit}_strain0"].copy()
    st = np.zeros((n, 20))
    st[:, 2] = g[f"pd{linit}_eps0"]
    dtot = g[f"pd{linit}_dtot"]
    f, wv, status = oracle_fint_pd(oracle, prob, rp, col, w, invk, linit, g[f"pd{linit}_dinc"], dtot,
                                   prob.coo + dtot, s, e, st)
    assert status == 0
    assert np.array_equal(f, g[f"pd{linit}_f"])
    assert np.array_equal(s, g[f"pd{linit}_stress"])
    assert np.array_equal(e, g[f"pd{linit}_strain"])
    assert np.array_equal(st[:, 2], g[f"pd{linit}_eps"])
    assert np.array_equal(wv, g[f"pd{linit}_wave"])
    if linit:
        assert np.array_equal(w, g["mm_omega"]) and np.array_equal(invk, g["mm_invk"])


def test_whole_run_pd_bit_exact(oracle):
    """40 explicit steps of the reference program with *PERIDYNAMICS 1 (mm deck)."""
    g = load("run_pd.npz")
    run = OracleRun(oracle, small(peridynamics=1, mm_units=True))
    done = 0
    for k in (1, 10, 40):
        run.step(k - done)
        done = k
        A = run.arr
        assert np.array_equal(A["x"], g[f"step{k}_x"])
        assert np.array_equal(A["dphy_tot"], g[f"step{k}_u"])
        assert np.array_equal(A["vphy"], g[f"step{k}_vphy"])
        assert np.array_equal(A["fint"], g[f"step{k}_fint"])
        assert np.array_equal(A["stress"], g[f"step{k}_stress"])
        assert np.array_equal(A["state"][:, 2], g[f"step{k}_eps"])


def test_pd_automatic_time_step_one_call_bit_exact(oracle):
    """DLT_FINT of the reference's CONSTRUCT_FINT_PD with AUTO_TS on (FINT_PD.F90:
    690-754) on the deformed yielding state of pd.npz: the shortest bond per
    particle at the current coordinates over its secant wave speed, smallest
    over the particles, times DLT_FAC."""
    g, a = load("pd.npz"), load("pd_auto.npz")
    prob = small(peridynamics=1, mm_units=True)
    n = prob.n
    rp, col, _ = oracle_search(oracle, prob)
    st = np.zeros((n, 20))
    st[:, 2] = g["pd0_eps0"]
    s, e = g["pd0_stress0"].copy(), g["pd0_strain0"].copy()
    dtot = g["pd0_dtot"]
    info = {}
    f, wv, status = oracle_fint_pd(oracle, prob, rp, col, g["mm_omega"].copy(), np.ascontiguousarray(g["mm_invk"]).copy(),
                                   0, g["pd0_dinc"], dtot, prob.coo + dtot, s, e, st,
                                   dlt_fac=float(a["call_dlt_fac"][0]), info=info)
    assert status == 0
    assert np.array_equal(wv, a["call_wave"]) and np.array_equal(f, a["call_f"])
    assert info["dlt"] == float(a["call_dlt"][0])
    # the rule itself: brute force over all pairs of the list at the current coordinates
    x = prob.coo + dtot
    best = np.full(n, np.inf)
    for i in range(n):
        j = col[rp[i]:rp[i + 1]]
        j = j[j != i]
        best[i] = np.sqrt(((x[j] - x[i]) ** 2).sum(1)).min()
    assert np.isclose((best / wv).min() * float(a["call_dlt_fac"][0]), info["dlt"], rtol=1e-14, atol=0)


def test_whole_run_pd_automatic_time_step_bit_exact(oracle):
    """40 steps of the reference program with IAUTO_TS = 1 (DLT_FAC 0.3) on the
    peridynamic mm deck: the first step has no predictor (MAIN.F90:348-350),
    each step's DLT comes from its own internal force (MAIN.F90:428); every
    step's TIME and frames 1, 10 and 40 bit-exact."""
    a = load("pd_auto.npz")
    run = OracleRun(oracle, small(peridynamics=1, mm_units=True), dlt_fac=float(a["run_dlt_fac"][0]))
    done = 0
    for k in (1, 10, 40):
        run.step(k - done)
        done = k
        A = run.arr
        assert np.array_equal(A["x"], a[f"step{k}_x"])
        assert np.array_equal(A["dphy_tot"], a[f"step{k}_u"])
        assert np.array_equal(A["vphy"], a[f"step{k}_vphy"])
        assert np.array_equal(A["fint"], a[f"step{k}_fint"])
        assert np.array_equal(A["stress"], a[f"step{k}_stress"])
        assert np.array_equal(A["state"][:, 2], a[f"step{k}_eps"])
    # TIME = TIME + DLT in step order, as the program logs it
    t, tt = [], 0.0
    for d in run.dlt:
        tt = tt + d
        t.append(tt)
    assert len(a["run_time"]) == 40 and np.array_equal(np.array(t), a["run_time"])
    # the first step did not move the particles: its predictor is skipped
    assert np.array_equal(a["step1_u"], np.zeros_like(a["step1_u"]))


@pytest.mark.parametrize("tag,fn", [("dam6", "or_von_mises_dam"), ("dp3", "or_druck_prag")])
def test_damage_models_bit_exact(oracle, tag, fn):
    """VON_MISES_DAM (type 6) and DRUCK_PRAG (type 3) of the compiled reference
    on 400 random states each, spanning their branches."""
    g = load("damage.npz")
    f = getattr(oracle, fn)
    for k in range(len(g[f"{tag}_props"])):
        s = g[f"{tag}_stress0"][k].copy()
        st = g[f"{tag}_state0"][k].copy()
        f(s, g[f"{tag}_total"][k].copy(), g[f"{tag}_inc"][k].copy(), st, g[f"{tag}_props"][k].copy())
        assert np.array_equal(s, g[f"{tag}_stress"][k]) and np.array_equal(st, g[f"{tag}_state"][k])


@pytest.mark.parametrize("mat_type,name", [(6, "run_dam6.npz"), (3, "run_dp3.npz")])
def test_whole_run_damage_bit_exact(oracle, mat_type, name):
    """40 explicit steps of the reference program with material type 6 / 3:
    frames 1, 10, 40 including the damage and the shear plastic strain."""
    from damage_util import damage_material
    g = load(name)
    run = OracleRun(oracle, small(material=damage_material(mat_type)))
    done = 0
    for k in (1, 10, 40):
        run.step(k - done)
        done = k
        A = run.arr
        assert np.array_equal(A["x"], g[f"step{k}_x"])
        assert np.array_equal(A["fint"], g[f"step{k}_fint"])
        assert np.array_equal(A["stress"], g[f"step{k}_stress"])
        assert np.array_equal(A["state"][:, 2], g[f"step{k}_eps"])
        assert np.array_equal(A["state"][:, 3], g[f"step{k}_dam"])
        assert np.array_equal(A["state"][:, 0], g[f"step{k}_eps_shear"])


@pytest.mark.parametrize("mat_type,name", [(6, "run_pd_dam6.npz"), (3, "run_pd_dp3.npz")])
def test_whole_run_pd_damage_bit_exact(oracle, mat_type, name):
    """40 steps of the reference program with *PERIDYNAMICS 1 (mm deck) and
    material type 6 / 3: CONSTRUCT_FINT_PD passes the damage models the same
    arguments as CONSTRUCT_FINT (FINT_PD.F90:493); frames 1, 10, 40 including
    the damage and the shear plastic strain."""
    from damage_util import damage_material
    g = load(name)
    run = OracleRun(oracle, small(peridynamics=1, mm_units=True, material=damage_material(mat_type)))
    done = 0
    for k in (1, 10, 40):
        run.step(k - done)
        done = k
        A = run.arr
        assert np.array_equal(A["x"], g[f"step{k}_x"])
        assert np.array_equal(A["vphy"], g[f"step{k}_vphy"])
        assert np.array_equal(A["fint"], g[f"step{k}_fint"])
        assert np.array_equal(A["stress"], g[f"step{k}_stress"])
        assert np.array_equal(A["state"][:, 2], g[f"step{k}_eps"])
        assert np.array_equal(A["state"][:, 3], g[f"step{k}_dam"])
        assert np.array_equal(A["state"][:, 0], g[f"step{k}_eps_shear"])
    assert (g["step40_dam"] > 0).sum() > 100   # damage under way in many particles


def test_whole_run_pd_drucker_prager_automatic_time_step_bit_exact(oracle):
    """40 steps of the reference program with *PERIDYNAMICS 1, material type 3
    and IAUTO_TS = 1: the Drucker-Prager wave speed is divided by 0.15
    (FINT_PD.F90:539-542) before the step is formed; every step's TIME and
    frames 1, 10, 40 bit-exact."""
    from damage_util import damage_material
    a = load("run_pd_dp3_auto.npz")
    run = OracleRun(oracle, small(peridynamics=1, mm_units=True, material=damage_material(3)),
                    dlt_fac=float(a["run_dlt_fac"][0]))
    done = 0
    for k in (1, 10, 40):
        run.step(k - done)
        done = k
        A = run.arr
        for key, ref in (("x", "x"), ("vphy", "vphy"), ("fint", "fint"), ("stress", "stress")):
            assert np.array_equal(A[key], a[f"step{k}_{ref}"])
        assert np.array_equal(A["state"][:, 3], a[f"step{k}_dam"])
    t, tt = [], 0.0
    for d in run.dlt:
        tt = tt + d
        t.append(tt)
    assert len(a["run_time"]) == 40 and np.array_equal(np.array(t), a["run_time"])
