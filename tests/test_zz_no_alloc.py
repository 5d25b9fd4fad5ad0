"""mega_preprocess_noalloc (no counterpart in the reference: the geometry
preprocessing of PRE_MODEL.F90:248-262 for callers that cannot allocate device
memory) against mega_preprocess on the same meshes, bit for bit: two calls in
a row on one GPU, the second with a smaller mesh (the fixed scratch keeps the
first call's contents), and a size beyond the scratch refused before any
device work."""
import ctypes as C

import numpy as np
import pytest

from taylor_amd import deck

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def mega():
    from taylor_amd import mega as m
    m.load_library()
    return m


class Hip:
    """Device buffers from the HIP runtime the library itself is linked
    against (/opt/rocm's libamdhip64), so that no second runtime is loaded."""

    def __init__(self):
        self.rt = C.CDLL("libamdhip64.so.7")
        self.rt.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        self.rt.hipFree.argtypes = [C.c_void_p]
        self.rt.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        self.rt.hipDeviceSynchronize.argtypes = []
        self.bufs = []

    def put(self, a):
        a = np.ascontiguousarray(a)
        p = C.c_void_p()
        assert self.rt.hipMalloc(C.byref(p), max(a.nbytes, 8)) == 0
        self.bufs.append(p)
        assert self.rt.hipMemcpy(p, a.ctypes.data_as(C.c_void_p), a.nbytes, 1) == 0   # host to device
        return p

    def get(self, p, like):
        out = np.empty_like(like)
        assert self.rt.hipMemcpy(out.ctypes.data_as(C.c_void_p), p, out.nbytes, 2) == 0   # device to host
        return out

    def free(self):
        for p in self.bufs:
            self.rt.hipFree(p)
        self.bufs = []


def run_noalloc(lib, coo, elcon, hex=None, with_sm=True):
    hip = Hip()
    try:
        n, numel = coo.shape[0], elcon.shape[0]
        win0, vol0, sm0 = np.full((n, 3), -1.0), np.full(n, -1.0), np.full((n, 6), -1.0)
        d_coo = hip.put(np.asarray(coo, np.float64))
        d_el = hip.put(np.asarray(elcon, np.int32))
        d_hex = None if hex is None else hip.put(np.asarray(hex, np.int32))
        d_win, d_vol = hip.put(win0), hip.put(vol0)
        d_sm = hip.put(sm0) if with_sm else None
        dmax = np.zeros(3)
        st = lib.mega_preprocess_noalloc(n, d_coo, numel, d_el, d_hex, d_win, d_vol, d_sm,
                                         dmax.ctypes.data_as(C.c_void_p))
        assert hip.rt.hipDeviceSynchronize() == 0
        return (st, hip.get(d_win, win0), hip.get(d_vol, vol0), (hip.get(d_sm, sm0) if with_sm else None), dmax)
    finally:
        hip.free()


def same_bits(a, b):
    """equal bit for bit (NaN where the existing entry point has NaN)"""
    a, b = np.ascontiguousarray(a, np.float64), np.ascontiguousarray(b, np.float64)
    return a.shape == b.shape and np.array_equal(a.view(np.int64), b.view(np.int64))


def test_noalloc_equals_preprocess_twice_then_smaller(mega):
    lib = mega.load_library()
    big = deck.taylor_bar_mesh(0.25e-3, length=12e-3, seed=4)
    small = deck.taylor_bar_mesh(0.5e-3, length=5e-3, seed=7)
    hex_small = (np.arange(small.elcon.shape[0]) % 3 != 0).astype(np.int32)   # every third a tetrahedron
    for label, mesh, hx, with_sm in (("bar 0.25 mm, all hexahedra", big, None, True),
                                     ("bar 0.5 mm, mixed hexahedra/tetrahedra", small, hex_small, True),
                                     ("bar 0.5 mm, no smoothing-length output", small, None, False)):
        print(f"noalloc: {label}: {mesh.coo.shape[0]} nodes, {mesh.elcon.shape[0]} elements", flush=True)
        win, vol, sm, dmax = mega.preprocess(mesh.coo, mesh.elcon, hx)
        st, w2, v2, s2, d2 = run_noalloc(lib, mesh.coo, mesh.elcon, hx, with_sm)
        assert st == 0, mega._status_string(st)
        assert same_bits(w2, win)
        assert same_bits(v2, vol)
        if with_sm:
            assert same_bits(s2, sm)
        assert same_bits(d2, dmax)


def test_noalloc_refuses_sizes_beyond_its_scratch(mega):
    lib = mega.load_library()
    mesh = deck.taylor_bar_mesh(0.7e-3, length=3e-3, seed=1)
    too_many = 524288 + 1
    st, *_ = run_noalloc(lib, mesh.coo, mesh.elcon)
    assert st == 0
    # refused on the host before any device work: the 16-double buffer is never read
    hip = Hip()
    buf = hip.put(np.zeros(16))
    dmax = np.zeros(3)
    for nump, numel in ((too_many, 10), (10, too_many)):
        st = lib.mega_preprocess_noalloc(nump, buf, numel, buf, None, buf, buf, None,
                                         dmax.ctypes.data_as(C.c_void_p))
        assert st == 8, mega._status_string(st)   # MEGA_ERR_SCRATCH
    hip.free()
    assert b"scratch" in lib.mega_status_string(8)
