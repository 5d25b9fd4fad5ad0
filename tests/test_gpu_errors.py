"""Run-time stops of the reference, through the C-ABI on the device.

* GET_BINS stops when a bin reaches MAX_NEIGH = 1000 nodes
  (BIN_SEARCH.F90:47-52): nodes that left the body or became non-finite pile
  into one bin.  mega_search returns MEGA_ERR_BINS without searching, and a
  later search at valid positions rebuilds the bins and gives the lists of a
  fresh search.
* CONSTITUTION stops on a material type it has no routine for
  (CONSTITUTION.F90:79-84; type 4, VISCO_ELASTIC, is not on this path
  either): a step over such nodes returns MEGA_ERR_MATERIAL, in both
  precisions, on the RK path and on the peridynamic one (which calls the same
  CONSTITUTION, FINT_PD.F90:493).
(The J2 Newton failure, MEGA_ERR_RETURN_MAP, is forced in the parity tests.)
"""
import copy
import dataclasses

import numpy as np
import pytest

from taylor_amd import deck, mega

pytestmark = pytest.mark.gpu

BINS, MATERIAL = 3, 7


@pytest.fixture(scope="module")
def lib():
    return mega.load_library()


def small_bar():
    return deck.taylor_problem(0.7e-3, length=6e-3, seed=0, v0=190.0, pdsearch_steps=3)


def test_a_full_bin_is_refused_and_a_later_search_recovers(lib):
    prob = small_bar()
    assert prob.n >= 1000
    g = mega.Mega(prob)
    g.search()
    ref = g.neighbors()
    for name, x in (("non-finite", np.full((prob.n, 3), np.nan)),
                    ("one point", np.repeat(prob.coo[:1], prob.n, axis=0))):
        print(f"{prob.n} nodes in one bin ({name})", flush=True)
        with pytest.raises(mega.MegaError) as ex:
            g.search(x)
        assert ex.value.status == BINS
    g.search()   # bins rebuilt at the context's positions
    got = g.neighbors()
    g.close()
    for a, b in zip(got, ref):
        assert np.array_equal(a, b)


@pytest.mark.parametrize("mat_type", [4, 5])
@pytest.mark.parametrize("precision", [mega.MEGA_FP64, mega.MEGA_MIXED])
@pytest.mark.parametrize("path", ["rk", "pd"])
def test_a_material_without_a_routine_stops_the_step(lib, path, precision, mat_type):
    prob = small_bar() if path == "rk" else deck.pd_plate_problem(lx=3e-3, ly=3e-3, lz=1.2e-3, impact_radius=1e-3)
    other = copy.deepcopy(prob.materials[0])
    other.mat_type = mat_type
    mat_id = (np.arange(prob.n) % 7 == 3).astype(np.int32)
    bad = dataclasses.replace(prob, materials=[prob.materials[0], other], mat_id=mat_id)
    g = mega.Mega(bad, precision=precision)
    st = lib.mega_step(g.ctx, prob.dt, 2)
    g.close()
    print(f"{path}: type {mat_type} on {int(mat_id.sum())} of {prob.n} nodes, precision {precision}: status {st}",
          flush=True)
    assert st == MATERIAL
