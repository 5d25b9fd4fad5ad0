"""GPU parity of the device geometry preprocessing (mega_preprocess:
GET_ADIAL + GET_VOLUMES + GET_SMOOTHING_LENGTHS, PRE_MODEL.F90:248-262)
against the reference's own PRE_MODEL output (tests/golden/geometry.npz,
written by the compiled reference), against deck.preprocess on a larger bar,
and against closed forms on a tetrahedral lattice.  Bar: windows and
volumes bit-exact; smoothing lengths and XDIST_MAX etc. come from the
correctly rounded VOL**THRD, while the reference's C-library pow misrounds
about 0.1 % of such powers by one ulp; through the quotient of two powers
and the products after it that is at most 2 ulp, on a bounded number of
nodes (printed)."""
import numpy as np
import pytest

from taylor_amd import deck

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def mega():
    from taylor_amd import mega as m
    m.load_library()
    return m


def load(name):
    import os
    return np.load(os.path.join(os.path.dirname(__file__), "golden", name))


def within_ulps(a, b, k=2):
    """(all positive entries within k ulp of each other, number that differ)"""
    a, b = np.ascontiguousarray(a, np.float64), np.ascontiguousarray(b, np.float64)
    assert (a > 0).all() and (b > 0).all()
    dist = np.abs(a.view(np.int64) - b.view(np.int64))
    return bool((dist <= k).all()), int((dist > 0).sum())


def test_preprocess_equals_reference_pre_model(mega):
    g = load("geometry.npz")
    win, vol, sm, dmax = mega.preprocess(g["coo"], g["elcon"])
    assert np.array_equal(win, g["adial"])
    assert np.array_equal(vol, g["vol"])
    ok, ndiff = within_ulps(sm, g["sm_len"])
    print("smoothing lengths differing in the last bit:", ndiff, "of", sm.size, flush=True)
    assert ok and ndiff <= 0.01 * sm.size
    assert within_ulps(dmax, g["dist_max"])[0]


def test_preprocess_equals_host_restatement_on_a_larger_bar(mega):
    mesh = deck.taylor_bar_mesh(0.25e-3, length=12e-3, seed=4)
    ref = deck.preprocess(mesh)
    print("nodes", mesh.coo.shape[0], "elements", mesh.elcon.shape[0], flush=True)
    win, vol, sm, dmax = mega.preprocess(mesh.coo, mesh.elcon)
    assert np.array_equal(win, ref.win0)
    assert np.array_equal(vol, ref.vol)
    ok, ndiff = within_ulps(sm, ref.sm_len)
    print("smoothing lengths differing in the last bit:", ndiff, "of", sm.size, flush=True)
    assert ok and ndiff <= 0.01 * sm.size
    assert within_ulps(dmax, ref.dist_max)[0]


def kuhn_tets(nx, ny, nz, h):
    """Lattice of nx*ny*nz cubes of side h, each split into the 6 tetrahedra
    along its main diagonal; elements as 8 ids (the 4 corners, the last one
    repeated), 1-based."""
    ix, iy, iz = np.meshgrid(np.arange(nx + 1), np.arange(ny + 1), np.arange(nz + 1), indexing="ij")
    coo = np.stack([ix.ravel(), iy.ravel(), iz.ravel()], axis=1) * h
    nid = lambda i, j, k: (i * (ny + 1) + j) * (nz + 1) + k + 1
    els = []
    for i in range(nx):
        for j in range(ny):
            for k in range(nz):
                base = np.array([i, j, k])
                for perm in ([0, 1, 2], [0, 2, 1], [1, 0, 2], [1, 2, 0], [2, 0, 1], [2, 1, 0]):
                    p, ids = base.copy(), [nid(*base)]
                    for a in perm:
                        p[a] += 1
                        ids.append(nid(*p))
                    els.append(ids + [ids[-1]] * 4)
    return coo, np.array(els, dtype=np.int32)


def test_preprocess_tetrahedra_closed_forms(mega):
    h = 0.5   # lattice coordinates, differences and the determinants are exact
    nx, ny, nz = 5, 4, 3
    coo, el = kuhn_tets(nx, ny, nz, h)
    win, vol, sm, dmax = mega.preprocess(coo, el, hex=np.zeros(len(el), np.int32))
    # every tetrahedron spans its whole cube: extents (h, h, h) at every node
    assert np.array_equal(win, np.full_like(win, h))
    # each tetrahedron is h^3/6 and gives a quarter to each of its 4 corners
    share = 0.25 * ((h ** 3) / 6.0)
    cnt = np.bincount(el[:, :4].ravel() - 1, minlength=len(coo))
    assert np.allclose(vol, cnt * share, rtol=1e-15, atol=0)
    assert abs(vol.sum() - nx * ny * nz * h ** 3) < 1e-13
    assert cnt.max() == 24   # interior node of the Kuhn lattice
    # smoothing lengths: (2 h) * (vol / (2 h)^3)^(1/3) / 2 with THRD = 0.333333333333333
    ratio = np.array([v ** 0.333333333333333 / (8 * h ** 3) ** 0.333333333333333 for v in vol])
    assert np.allclose(sm[:, 0], h * ratio, rtol=1e-15)
    assert dmax[0] == sm[:, 0].max() and dmax[2] == sm[:, 5].max()


def test_preprocess_refuses_unknown_node_ids(mega):
    g = load("geometry.npz")
    el = g["elcon"].copy()
    el[3, 5] = g["coo"].shape[0] + 1
    with pytest.raises(mega.MegaError) as ex:
        mega.preprocess(g["coo"], el)
    assert ex.value.status == 6   # MEGA_ERR_ARG
