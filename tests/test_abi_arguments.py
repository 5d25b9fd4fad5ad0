"""Refusals of the C-ABI (include/mega.h) that are decided from the arguments
alone, before any device call, so they run without a GPU.

The reference stops the program on the same conditions (EXIT_PROGRAM):
RK_DEGREE / RK_CONT outside what it implements (PRE_CONTROL.F90:384-400;
its Taylor-bar decks fix RK_DEGREE = 1, RK_CONT = 3, PRE_CONTROL.F90:782-787),
material types without a constitutive routine (CONSTITUTION.F90:79-84) and a
peridynamic run that is not Lagrangian (FINT_PD.F90:377-432 only forms the
Lagrangian shape tensors once).  Here each is a status code.
"""
import ctypes as C

import numpy as np
import pytest

from taylor_amd import mega

ARG, MATERIAL, BINS = 6, 7, 3


@pytest.fixture(scope="module")
def lib():
    return mega.load_library()


def arrays(n, lazy=False):
    """Inputs of mega_create for n nodes.  lazy: np.empty, whose pages are never
    touched (a refusal on the size alone reads none of them)."""
    f64 = (lambda k: np.empty(k)) if lazy else (lambda k: np.ones(k))
    i32 = (lambda k: np.empty(k, np.int32)) if lazy else (lambda k: np.ones(k, np.int32))
    return [f64(3 * n), f64(3 * n), f64(n), f64(3 * n), i32(3 * n), i32(n), i32(n), f64(30 * n), f64(3 * n)]


def control(**kw):
    c = mega.MegaControl()
    c.rk_degree, c.rk_cont, c.lagrangian, c.finite_strain, c.pdsearch = 1, 3, 1, 1, 1
    for k, v in kw.items():
        setattr(c, k, v)
    return c


def create(lib, ctrl, n, arrs, drop=None):
    ptrs = [None if k == drop else a.ctypes.data_as(C.c_void_p) for k, a in enumerate(arrs)]
    out = C.c_void_p(12345)
    st = lib.mega_create(C.byref(ctrl) if ctrl is not None else None, n, *ptrs, C.byref(out))
    return st, out.value


@pytest.mark.parametrize("bad", [
    {"rk_degree": 0}, {"rk_degree": 2}, {"rk_cont": 2}, {"rk_cont": -1},
    {"itype_int": 2}, {"itype_int": 1, "peridynamics": 1},
    {"precision": 2}, {"precision": -1},
    {"surrogate": 1},                                   # the learned return map runs in mixed precision only
    {"peridynamics": 1, "lagrangian": 0},               # CONSTRUCT_FINT_PD is Lagrangian only
    {"peridynamics": 1, "precision": 1, "surrogate": 1},
], ids=lambda d: ",".join(f"{k}={v}" for k, v in d.items()))
def test_create_refuses_unsupported_control(lib, bad):
    st, out = create(lib, control(**bad), 8, arrays(8))
    assert st == ARG, mega._status_string(st)
    assert out == 12345   # left as it was: no context was made


@pytest.mark.parametrize("drop", range(8))   # X0, win, vol, mass, ebc, ebc_node, mat_type, prop
def test_create_refuses_missing_arrays(lib, drop):
    st, _ = create(lib, control(), 8, arrays(8), drop=drop)
    assert st == ARG


def test_create_refuses_empty_and_negative_sizes(lib):
    for n in (0, -1):
        assert create(lib, control(), n, arrays(1))[0] == ARG
    assert create(lib, None, 8, arrays(8))[0] == ARG


def test_create_refuses_node_counts_beyond_the_search_index(lib):
    """27 per-source-bin counters per node are indexed in int32: the largest n
    is 79,536,431; one more is refused before any array is read."""
    n = 2 ** 31 // 27 + 1
    assert 27 * n > 2 ** 31 - 1 and 27 * (n - 1) <= 2 ** 31 - 1
    st, _ = create(lib, control(), n, arrays(n, lazy=True))
    assert st == BINS, mega._status_string(st)


def test_calls_without_a_context_are_refused(lib):
    buf = np.zeros(64)
    p = buf.ctypes.data_as(C.c_void_p)
    assert lib.mega_destroy(None) == 0
    assert lib.mega_search(None, None, 1) == ARG
    assert lib.mega_shapes(None, None) == ARG
    assert lib.mega_interp(None, p, p) == ARG
    assert lib.mega_step(None, 1e-9, 1) == ARG
    assert lib.mega_step_auto(None, 0.5, 1, p) == ARG
    assert lib.mega_download(None, 0, p) == ARG
    assert lib.mega_upload(None, 0, p) == ARG


def test_constitute_refuses_materials_without_a_routine(lib):
    """CONSTITUTION.F90:79-84 stops on a type it has no routine for; types 1, 2,
    3 and 6 are on this path."""
    n = 4
    lmat, stress, strain = np.zeros(9 * n), np.zeros(6 * n), np.zeros(6 * n)
    state, prop = np.zeros(20 * n), np.ones(30)
    args = [a.ctypes.data_as(C.c_void_p) for a in (lmat, stress, strain, state, prop)]
    for t in (0, 4, 5, 7):
        assert lib.mega_constitute(n, *args, t, 1, 0) == MATERIAL
    assert lib.mega_constitute(n, *args, 2, 1, 3) == ARG        # precision
    assert lib.mega_constitute(0, *args, 2, 1, 0) == ARG        # no points
    assert lib.mega_constitute(n, None, *args[1:], 2, 1, 0) == ARG
