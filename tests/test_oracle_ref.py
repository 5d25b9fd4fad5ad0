"""The oracle against the reference subroutines called live (oracle/_ref,
compiled from the reference's own sources) on fresh seeded inputs; skipped
where that build is absent.  Bar: bit-exact."""
import ctypes as C

import numpy as np
import pytest

from oracle_bind import control, oracle_fint, oracle_search, oracle_shapes, ref_set_control
from taylor_amd import deck


@pytest.mark.parametrize("seed", [3, 4])
def test_search_rk_fint_against_reference(oracle, ref, seed):
    prob = deck.taylor_problem(0.9e-3, length=5e-3, seed=seed, v0=250.0)
    ref_set_control(ref, prob)
    n = prob.n
    coo, win, d = np.ascontiguousarray(prob.coo), np.ascontiguousarray(prob.win), prob.dist_max
    geo, nb = np.zeros(6), np.zeros(3, np.int32)
    ref.ref_bin_geometry(n, coo, win, float(d[0]), float(d[1]), float(d[2]), geo, nb)
    ijk = np.zeros((n, 3), np.int32)
    ref.ref_get_bins(n, coo, ijk)
    dim = 200 * n
    gn, gs, st = np.zeros(n, np.int32), np.zeros(n, np.int32), np.zeros(dim, np.int32)
    mx, nz = C.c_int(), C.c_int()
    ref.ref_soft_search(n, coo, win, np.zeros((n, 6)), float(d[0]), float(d[1]), float(d[2]), gn, gs, dim, st,
                        C.byref(mx), C.byref(nz))
    rp, col, oijk = oracle_search(oracle, prob)
    assert np.array_equal(oijk, ijk) and np.array_equal(col + 1, st[:nz.value])
    shp, dshp = oracle_shapes(oracle, prob, rp, col)
    ebc = np.ascontiguousarray(prob.ebc_node, dtype=np.int32)
    for i in range(0, n, 7):
        a, b = rp[i], rp[i + 1]
        s1, d1 = np.zeros(b - a), np.zeros(3 * (b - a))
        ref.ref_rk1(coo[i].copy(), n, coo, win, np.ascontiguousarray(col[a:b] + 1), int(b - a), ebc, int(ebc[i]), s1, d1)
        assert np.array_equal(s1, shp[a:b]) and np.array_equal(d1.reshape(-1, 3), dshp[a:b])
    rng = np.random.default_rng(seed)
    dtot = rng.normal(size=(n, 3)) * 3e-6
    dinc = rng.normal(size=(n, 3)) * 5e-7
    stress = rng.normal(size=(n, 6)) * 4e8
    strain = rng.normal(size=(n, 6)) * 1e-3
    state = np.zeros((n, 20))
    state[:, 2] = np.abs(rng.normal(size=n)) * 0.03
    s1, e1, t1 = stress.copy(), strain.copy(), state.copy()
    f1, _, w1, _ = oracle_fint(oracle, prob, rp, col, shp.copy(), dshp.copy(), 0, dinc, dtot, coo + dtot, s1, e1, t1)
    s2, e2, t2 = stress.copy(), strain.copy(), state.copy()
    f2, fx2, wv, seq = np.zeros((n, 3)), np.zeros((n, 3)), np.zeros(n), np.zeros(n)
    w2 = C.c_double(0)
    props = np.ascontiguousarray(prob.props_table()[prob.mat_id])
    ref.ref_construct_fint(n, win, np.ascontiguousarray(prob.vol), coo, np.ascontiguousarray(coo + dtot), props, t2, s2, e2,
                           np.ascontiguousarray(dinc), np.ascontiguousarray(dtot), np.full(n, 2, np.int32), ebc,
                           np.diff(rp).astype(np.int32), (rp[:-1] + 1).astype(np.int32), int(rp[-1]),
                           (col + 1).astype(np.int32), shp.copy(), dshp.copy(), int(np.diff(rp).max()), 0,
                           f2, fx2, wv, np.zeros((n, 3)), np.ones(n, np.int32), seq, C.byref(w2), prob.dt)
    assert np.array_equal(f1, f2) and np.array_equal(s1, s2) and np.array_equal(t1[:, 2], t2[:, 2])
    assert w1 == w2.value
