"""GPU parity of the damage material types: 6 (VON_MISES_DAM,
CONSTITUTE_VONMISES_DAM.F90) and 3 (DRUCK_PRAG, CONSTITUTE_DRUCKPRAG.F90),
per point through mega_constitute against the oracle, and over 40 device
steps against the reference program's own frames
(tests/golden/run_dam6.npz, run_dp3.npz).  FP64 bars as for the J2 path;
the damage state is read back through the damage-state field."""
import os

import numpy as np
import pytest

from damage_util import damage_material
from taylor_amd import deck

pytestmark = pytest.mark.gpu

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def rel_err(a, b):
    a, b = np.asarray(a), np.asarray(b)
    return float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-300))


@pytest.fixture(scope="module")
def mega():
    from taylor_amd import mega as m
    m.load_library()
    return m


@pytest.mark.parametrize("mat_type", [6, 3])
@pytest.mark.parametrize("finite", [1, 0])
def test_constitute_damage_matches_oracle(mega, oracle, mat_type, finite):
    mat = damage_material(mat_type)
    rng = np.random.default_rng(10 + mat_type + finite)
    n = 3000
    scale = 3e8 if mat_type == 6 else 2e7
    L = rng.normal(size=(n, 9)) * 2e-3
    stress = rng.normal(size=(n, 6)) * scale
    strain = rng.normal(size=(n, 6)) * (0.03 if mat_type == 6 else 0.004)
    state = np.zeros((n, 20))
    state[:, 0] = rng.uniform(0, 0.05, n)
    state[:, 2] = rng.uniform(0, 0.2, n)
    state[:, 3] = rng.uniform(0, 0.5, n) * (rng.uniform(size=n) < 0.5)
    state[:, 4] = rng.uniform(0, 0.01, n) * (rng.uniform(size=n) < 0.3)
    s_ref, e_ref, st_ref = stress.copy(), strain.copy(), state.copy()
    for i in range(n):
        oracle.or_constitute_point(np.ascontiguousarray(L[i]), s_ref[i], e_ref[i], st_ref[i], mat.props, mat_type,
                                   finite, np.zeros(6))
    s, e, st = stress.copy(), strain.copy(), state.copy()
    mega.constitute(L, s, e, st, mat.props, mat_type, finite)
    print("changed damage", int((st_ref[:, 3] != state[:, 3]).sum()), "plastic", int((st_ref[:, 2] != state[:, 2]).sum()))
    assert rel_err(s, s_ref) < 1e-12
    assert rel_err(e, e_ref) < 1e-12
    for k in (0, 2, 3, 4):
        assert np.abs(st[:, k] - st_ref[:, k]).max() <= 1e-12 * max(1.0, np.abs(st_ref[:, k]).max())


@pytest.mark.parametrize("mat_type,name", [(6, "run_dam6.npz"), (3, "run_dp3.npz")])
def test_damage_steps_match_the_reference_program(mega, mat_type, name):
    """40 device steps in FP64 against frames 1, 10, 40 of the reference
    program with the same deck: positions, force, stress, plastic strain,
    damage and (type 3) shear plastic strain."""
    g = np.load(os.path.join(GOLD, name))
    prob = deck.taylor_problem(0.7e-3, length=6e-3, seed=0, lagrangian=1, pdsearch_steps=3,
                               material=damage_material(mat_type))
    m = mega.Mega(prob)
    done = 0
    for k in (1, 10, 40):
        m.step(k - done)
        done = k
        ds = m.download("damage_state")
        errs = {"x": rel_err(m.download("x"), g[f"step{k}_x"]),
                "fint": rel_err(m.download("fint"), g[f"step{k}_fint"]),
                "stress": rel_err(m.download("stress"), g[f"step{k}_stress"]),
                "eps": float(np.abs(m.download("eps") - g[f"step{k}_eps"]).max()),
                "dam": float(np.abs(ds[:, 2] - g[f"step{k}_dam"]).max()),
                "eps_shear": float(np.abs(ds[:, 0] - g[f"step{k}_eps_shear"]).max())}
        print(name, k, errs, flush=True)
        assert errs["x"] < 1e-12 and errs["fint"] < 1e-9 and errs["stress"] < 1e-9
        assert errs["eps"] < 1e-9 and errs["dam"] < 1e-9 and errs["eps_shear"] < 1e-9
    m.close()


@pytest.mark.parametrize("mat_type", [6, 3])
def test_damage_steps_mixed_precision(mega, mat_type):
    """Mixed precision (FP32 stress, return maps and damage state) against FP64
    over 40 steps: positions to 1e-6 of the bar, damage within 2e-2."""
    prob = deck.taylor_problem(0.7e-3, length=6e-3, seed=0, material=damage_material(mat_type))
    out = {}
    for prec in (mega.MEGA_FP64, mega.MEGA_MIXED):
        m = mega.Mega(prob, precision=prec)
        m.step(40)
        out[prec] = (m.download("x"), m.download("damage_state"))
        m.close()
    (x64, d64), (x32, d32) = out[mega.MEGA_FP64], out[mega.MEGA_MIXED]
    ex = float(np.abs(x32 - x64).max() / np.abs(x64).max())
    ed = float(np.abs(d32[:, 2] - d64[:, 2]).max())
    print("mixed vs FP64: x", ex, "damage", ed)
    assert ex < 1e-6 and ed < 2e-2


@pytest.mark.parametrize("mat_type", [6, 3])
@pytest.mark.parametrize("pd", [0, 1])
def test_construct_fint_carries_the_damage_state(mega, oracle, mat_type, pd):
    """One CONSTRUCT_FINT (or CONSTRUCT_FINT_PD) call through mega_fint on a
    deformed state with damage under way: the caller's STATE(1), (2), (4), (5)
    go in and come back updated, as GSTATE does in the reference
    (FINT.F90:809-821, FINT_PD.F90:497-503), against the oracle."""
    from oracle_bind import oracle_fint, oracle_fint_pd, oracle_search
    mm = bool(pd)   # the peridynamic deck in mm / tonne / s / MPa
    prob = deck.taylor_problem(0.7e-3, length=6e-3, seed=0, pdsearch_steps=3, peridynamics=pd, mm_units=mm,
                               material=damage_material(mat_type))
    n = prob.n
    L = 1e3 if mm else 1.0
    S = (1e-6 if mm else 1.0) * (3e8 if mat_type == 6 else 2e7)
    rng = np.random.default_rng(40 + mat_type + 2 * pd)
    m = mega.Mega(prob)
    m.search()
    rp, col, _ = oracle_search(oracle, prob)
    nnz = int(rp[-1])
    z = np.zeros((n, 3))
    zs = lambda w: np.zeros((n, w))   # noqa: E731
    if pd:
        omega, invk = np.zeros(nnz), np.zeros((n, 9))
        oracle_fint_pd(oracle, prob, rp, col, omega, invk, 1, z, z, prob.coo, zs(6), zs(6), zs(20))
    else:
        shp, dshp = np.zeros(nnz), np.zeros((nnz, 3))
        oracle_fint(oracle, prob, rp, col, shp, dshp, 1, z, z, prob.coo, zs(6), zs(6), zs(20))
    m.construct_fint(1, z, z, prob.coo, zs(6), zs(6), zs(20))
    dtot = rng.normal(size=(n, 3)) * 2e-6 * L
    dinc = (rng.normal(size=(n, 3)) * 4e-7 + np.array([0, 0, -1e-6])) * L
    stress = rng.normal(size=(n, 6)) * S
    strain = rng.normal(size=(n, 6)) * (0.01 if mat_type == 6 else 0.004)   # total strain norms within (K_I, K_C)
    state = np.zeros((n, 20))
    state[:, 0] = np.abs(rng.normal(size=n)) * 0.01    # EPS_SHEAR
    state[:, 1] = rng.normal(size=n) * 1e-3             # EPS_VOL (carried unchanged)
    state[:, 2] = np.abs(rng.normal(size=n)) * 0.02     # EPS
    state[:, 3] = rng.uniform(0.0, 0.5, size=n)         # DAM
    state[:, 4] = np.abs(rng.normal(size=n)) * 0.01     # DAM_SV
    x_cur = prob.coo + dtot
    s_r, e_r, st_r = stress.copy(), strain.copy(), state.copy()
    if pd:
        f_r, _, _ = oracle_fint_pd(oracle, prob, rp, col, omega, invk, 0, dinc, dtot, x_cur, s_r, e_r, st_r)
    else:
        f_r, _, _, _ = oracle_fint(oracle, prob, rp, col, shp, dshp, 0, dinc, dtot, x_cur, s_r, e_r, st_r)
    s_g, e_g, st_g = stress.copy(), strain.copy(), state.copy()
    f_g, _, _, status = m.construct_fint(0, dinc, dtot, x_cur, s_g, e_g, st_g)
    m.close()
    errs = {"fint": rel_err(f_g, f_r), "stress": rel_err(s_g, s_r)}
    for k, name in ((0, "eps_shear"), (1, "eps_vol"), (2, "eps"), (3, "dam"), (4, "dam_sv")):
        errs[name] = rel_err(st_g[:, k], st_r[:, k])
    changed = {name: int((st_r[:, k] != state[:, k]).sum()) for k, name in ((0, "eps_shear"), (3, "dam"), (4, "dam_sv"))}
    print(f"type {mat_type}, {'PD' if pd else 'RK'}: status {status}, changed {changed}, "
          + " ".join(f"{a} {b:.1e}" for a, b in errs.items()), flush=True)
    assert status == 0
    assert changed["dam"] > n // 10 and changed["dam_sv"] > n // 10
    if mat_type == 3:
        assert changed["eps_shear"] > n // 10
    assert np.array_equal(st_g[:, 1], state[:, 1])
    for k, v in errs.items():
        assert v < 1e-10, (k, v)
