"""ctypes bindings for the oracle (oracle/liboracle.so) and, where it was built,
the reference harness (oracle/_ref/libmega_ref.so).  Test infrastructure: only
tests, smoke() and bench.py's CPU baseline use these."""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_SO = os.path.join(ROOT, "oracle", "liboracle.so")
REF_SO = os.path.join(ROOT, "oracle", "_ref", "libmega_ref.so")
REF_MEGA = os.path.join(ROOT, "oracle", "_ref", "MEGA")

dp = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
ip = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")
lp = np.ctypeslib.ndpointer(dtype=np.int64, flags="C_CONTIGUOUS")


class OrControl(C.Structure):
    _fields_ = [("rk_degree", C.c_int), ("rk_cont", C.c_int), ("ql", C.c_int), ("ql_coef", C.c_double),
                ("ql_len", C.c_double), ("shsup", C.c_int), ("lagrangian", C.c_int),
                ("finite_strain", C.c_int), ("gravity", C.c_double * 3), ("peridynamics", C.c_int),
                ("itype_int", C.c_int)]


class OrFintIO(C.Structure):
    _fields_ = [("n", C.c_int),
                ("win", C.c_void_p), ("vol", C.c_void_p), ("coo0", C.c_void_p), ("coo_cur", C.c_void_p),
                ("props", C.c_void_p), ("mat_id", C.c_void_p), ("mat_type", C.c_void_p), ("ebc_node", C.c_void_p),
                ("dinc", C.c_void_p), ("dinc_tot", C.c_void_p), ("bodyforce", C.c_void_p),
                ("row_ptr", C.c_void_p), ("col", C.c_void_p), ("shp", C.c_void_p), ("dshp", C.c_void_p),
                ("stress", C.c_void_p), ("strain", C.c_void_p), ("state", C.c_void_p),
                ("fint", C.c_void_p), ("fext", C.c_void_p), ("wave_vel", C.c_void_p), ("strain_eq", C.c_void_p),
                ("invk", C.c_void_p),
                ("int_work", C.c_double), ("linit", C.c_int), ("nthreads", C.c_int), ("ret_fail", C.c_int64),
                ("auto_ts", C.c_int), ("dlt_fac", C.c_double), ("char_dist", C.c_void_p), ("dlt_fint", C.c_double)]


class OrModel(C.Structure):
    _fields_ = [("n", C.c_int), ("c", OrControl), ("dt", C.c_double), ("time_end", C.c_double),
                ("pdsearch", C.c_int), ("nthreads", C.c_int),
                ("coo0", C.c_void_p), ("win", C.c_void_p), ("vol", C.c_void_p), ("mass", C.c_void_p),
                ("sm_dist", C.c_void_p), ("ebc", C.c_void_p), ("ebc_node", C.c_void_p),
                ("mat_id", C.c_void_p), ("nmat", C.c_int), ("props", C.c_void_p), ("mat_type", C.c_void_p),
                ("x", C.c_void_p), ("v", C.c_void_p), ("a", C.c_void_p), ("d", C.c_void_p), ("dtot", C.c_void_p),
                ("dphy", C.c_void_p), ("dphy_tot", C.c_void_p), ("fint", C.c_void_p), ("fext", C.c_void_p),
                ("prforce", C.c_void_p), ("stress", C.c_void_p), ("strain", C.c_void_p), ("state", C.c_void_p),
                ("strain_eq", C.c_void_p), ("wave_vel", C.c_void_p), ("vphy", C.c_void_p), ("aphy", C.c_void_p),
                ("int_work", C.c_double), ("ke", C.c_double), ("ie", C.c_double), ("ee", C.c_double),
                ("time", C.c_double), ("step", C.c_int64), ("nnz", C.c_int64), ("nnz_cap", C.c_int64),
                ("row_ptr", C.POINTER(C.c_int64)), ("col", C.POINTER(C.c_int32)),
                ("shp", C.POINTER(C.c_double)), ("dshp", C.POINTER(C.c_double)), ("invk", C.POINTER(C.c_double)),
                ("geom", C.c_double * 6), ("nb", C.c_int32 * 3), ("ijk", C.POINTER(C.c_int32)),
                ("search_count", C.c_int), ("ret_fail", C.c_int64),
                ("auto_ts", C.c_int), ("dlt_fac", C.c_double), ("char_dist", C.POINTER(C.c_double))]


def _ptr(a):
    return a.ctypes.data_as(C.c_void_p) if a is not None else None


def load_oracle():
    lib = C.CDLL(ORACLE_SO)
    lib.or_kernel.argtypes = [C.c_double, C.POINTER(C.c_double), C.POINTER(C.c_double)]
    lib.or_m44inv.argtypes = [dp, dp]
    lib.or_inv3.argtypes = [dp, dp]
    lib.or_inverse.argtypes = [dp, C.c_int, dp]
    lib.or_inverse.restype = C.c_int
    lib.or_determinant.argtypes = [dp]
    lib.or_determinant.restype = C.c_double
    lib.or_hughes_winget.argtypes = [dp, dp, dp, dp]
    lib.or_rotate_tensor.argtypes = [dp, dp]
    lib.or_von_mises.argtypes = [dp, dp, dp, dp]
    lib.or_von_mises.restype = C.c_int
    for nm in ("or_von_mises_dam", "or_druck_prag"):
        getattr(lib, nm).argtypes = [dp, dp, dp, dp, dp]
        getattr(lib, nm).restype = C.c_int
    lib.or_estimate_moduli.argtypes = [dp, dp, C.POINTER(C.c_double), C.POINTER(C.c_double)]
    lib.or_constitute_point.argtypes = [dp, dp, dp, dp, dp, C.c_int, C.c_int, dp]
    lib.or_constitute_point.restype = C.c_int
    lib.or_rk1.argtypes = [dp, C.c_int, dp, dp, ip, ip, C.c_int, C.POINTER(OrControl), dp, dp]
    lib.or_bin_geometry.argtypes = [C.c_int, dp, dp, C.c_double, C.c_double, C.c_double, dp, ip]
    lib.or_get_bins.argtypes = [C.c_int, dp, dp, ip, ip]
    lib.or_soft_search.argtypes = [C.c_int, dp, dp, ip, ip, C.c_int, lp, C.c_void_p]
    lib.or_soft_search.restype = C.c_int64
    lib.or_construct_fint.argtypes = [C.POINTER(OrFintIO), C.POINTER(OrControl)]
    lib.or_construct_fint.restype = C.c_int
    lib.or_interp.argtypes = [C.c_int, lp, ip, dp, dp, dp]
    lib.or_udfm.argtypes = [dp, C.c_int, dp, dp, dp, ip, C.c_int, dp, dp]
    lib.or_dfm.argtypes = [dp, dp, dp, dp, ip, C.c_int, dp, dp, dp]
    lib.or_snni_grad.argtypes = [dp, dp, C.c_double, C.c_int, dp, dp, ip, ip, C.c_int, C.POINTER(OrControl), dp]
    lib.or_construct_fint_pd.argtypes = [C.POINTER(OrFintIO), C.POINTER(OrControl)]
    lib.or_construct_fint_pd.restype = C.c_int
    lib.or_model_init.argtypes = [C.POINTER(OrModel)]
    lib.or_model_init.restype = C.c_int
    lib.or_model_step.argtypes = [C.POINTER(OrModel)]
    lib.or_model_step.restype = C.c_int
    lib.or_model_free.argtypes = [C.POINTER(OrModel)]
    return lib


def have_ref() -> bool:
    return os.path.exists(REF_SO)


def load_ref():
    lib = C.CDLL(REF_SO)
    lib.ref_set_control.argtypes = [C.c_int] * 5 + [C.c_double, C.c_double] + [C.c_int] * 3 + [dp]
    lib.ref_mls_kernel0.argtypes = [C.c_double, C.c_double, C.c_int, C.POINTER(C.c_double), C.POINTER(C.c_double)]
    lib.ref_m44inv.argtypes = [dp, dp]
    lib.ref_inv3.argtypes = [dp, dp]
    lib.ref_inverse.argtypes = [dp, C.c_int, dp]
    lib.ref_determinant.argtypes = [dp, C.POINTER(C.c_double)]
    lib.ref_hughes_winget.argtypes = [dp, dp, dp, dp]
    lib.ref_rotate_tensor.argtypes = [dp, dp]
    lib.ref_von_mises.argtypes = [dp, dp, dp, dp, dp]
    for nm in ("ref_von_mises_dam", "ref_druck_prag"):
        getattr(lib, nm).argtypes = [dp, dp, dp, dp, dp, dp]
    lib.ref_estimate_moduli.argtypes = [dp, dp, C.POINTER(C.c_double), C.POINTER(C.c_double)]
    lib.ref_rk1.argtypes = [dp, C.c_int, dp, dp, ip, C.c_int, ip, C.c_int, dp, dp]
    lib.ref_bin_geometry.argtypes = [C.c_int, dp, dp, C.c_double, C.c_double, C.c_double, dp, ip]
    lib.ref_get_bins.argtypes = [C.c_int, dp, ip]
    lib.ref_soft_search.argtypes = [C.c_int, dp, dp, dp, C.c_double, C.c_double, C.c_double, ip, ip, C.c_int,
                                    ip, C.POINTER(C.c_int), C.POINTER(C.c_int)]
    lib.ref_construct_fint.argtypes = ([C.c_int, dp, dp, dp, dp, dp, dp, dp, dp, dp, dp, ip, ip, ip, ip,
                                        C.c_int, ip, dp, dp, C.c_int, C.c_int, dp, dp, dp, dp, ip, dp,
                                        C.POINTER(C.c_double), C.c_double])
    lib.ref_preprocess.argtypes = [C.c_int, C.c_int, ip, dp, dp, dp, dp, dp]
    lib.ref_udfm.argtypes = [dp, C.c_int, dp, dp, dp, ip, C.c_int, ip, C.c_int, dp, dp]
    lib.ref_construct_fint_pd.argtypes = ([C.c_int, dp, dp, dp, dp, dp, dp, dp, dp, dp, dp, ip, ip, ip, ip,
                                           C.c_int, ip, dp, dp, C.c_int, C.c_int, dp, dp, C.c_double])
    lib.ref_set_auto_ts.argtypes = [C.c_int, C.c_double]
    lib.ref_last_dlt_fint.restype = C.c_double
    return lib


def control(prob) -> OrControl:
    c = OrControl()
    c.rk_degree, c.rk_cont, c.ql, c.ql_coef, c.ql_len = 1, 3, int(prob.ql), float(prob.ql_coef), float(prob.ql_len)
    c.shsup, c.lagrangian, c.finite_strain = int(prob.shsup), int(prob.lagrangian), int(prob.finite_strain)
    for k in range(3):
        c.gravity[k] = float(prob.gravity[k])
    c.peridynamics = int(getattr(prob, "peridynamics", 0))
    c.itype_int = int(getattr(prob, "itype_int", 0))
    return c


def ref_set_control(ref, prob):
    ref.ref_set_control(1, 3, 0, 0, int(prob.ql), float(prob.ql_coef), float(prob.ql_len), int(prob.shsup),
                        int(prob.lagrangian), int(prob.finite_strain), np.ascontiguousarray(prob.gravity, dtype=np.float64))


class OracleRun:
    """Whole explicit run of the oracle (MAIN.F90:304-658) on a deck.Problem."""

    def __init__(self, lib, prob, nthreads: int = 1, dlt_fac: float | None = None):
        """dlt_fac: the automatic time step's DLT_FAC (IAUTO_TS = 1,
        peridynamics); each step's DLT is then kept in ``dlt``."""
        n = prob.n
        self.lib, self.prob = lib, prob
        f = lambda *s: np.zeros(s, dtype=np.float64)
        self.arr = dict(coo0=np.ascontiguousarray(prob.coo, dtype=np.float64), win=np.ascontiguousarray(prob.win),
                        vol=np.ascontiguousarray(prob.vol), mass=np.ascontiguousarray(prob.mass),
                        sm_dist=np.ascontiguousarray(prob.dist_max, dtype=np.float64),
                        ebc=np.ascontiguousarray(prob.ebc, dtype=np.int32),
                        ebc_node=np.ascontiguousarray(prob.ebc_node, dtype=np.int32),
                        mat_id=np.ascontiguousarray(prob.mat_id, dtype=np.int32),
                        props=np.ascontiguousarray(prob.props_table()), mat_type=prob.mat_types(),
                        x=f(n, 3), v=np.ascontiguousarray(prob.vinit, dtype=np.float64).copy(), a=f(n, 3), d=f(n, 3),
                        dtot=f(n, 3), dphy=f(n, 3), dphy_tot=f(n, 3), fint=f(n, 3), fext=f(n, 3), prforce=f(n, 3),
                        stress=f(n, 6), strain=f(n, 6), state=f(n, 20), strain_eq=f(n), wave_vel=f(n),
                        vphy=f(n, 3), aphy=f(n, 3))
        m = OrModel()
        m.n = n
        m.c = control(prob)
        m.dt, m.time_end, m.pdsearch, m.nthreads = prob.dt, prob.time_end, int(prob.pdsearch), nthreads
        for k, a in self.arr.items():
            setattr(m, k, _ptr(a))
        m.nmat = len(prob.materials)
        if dlt_fac is not None:
            m.auto_ts, m.dlt_fac = 1, float(dlt_fac)
        self.dlt = []
        self.m = m
        if lib.or_model_init(C.byref(m)) != 0:
            raise RuntimeError("oracle init failed")

    def step(self, k: int = 1):
        for _ in range(k):
            r = self.lib.or_model_step(C.byref(self.m))
            if r == -2:
                raise RuntimeError("unsupported material")
            self.dlt.append(self.m.dt)
        return self

    def csr(self):
        n, nnz = self.m.n, self.m.nnz
        rp = np.ctypeslib.as_array(self.m.row_ptr, shape=(n + 1,)).copy()
        col = np.ctypeslib.as_array(self.m.col, shape=(nnz,)).copy()
        shp = np.ctypeslib.as_array(self.m.shp, shape=(nnz,)).copy()
        dshp = np.ctypeslib.as_array(self.m.dshp, shape=(nnz * 3,)).reshape(nnz, 3).copy()
        return rp, col, shp, dshp

    def __del__(self):
        try:
            self.lib.or_model_free(C.byref(self.m))
        except Exception:
            pass


def oracle_search(lib, prob, x=None, ijk=None):
    """Bins from X0 geometry, cells of x (or given ijk), SOFT_SEARCH at x.
    Returns row_ptr (int64), col (int32, 0-based), ijk."""
    n = prob.n
    coo = np.ascontiguousarray(prob.coo, dtype=np.float64)
    x = coo if x is None else np.ascontiguousarray(x, dtype=np.float64)
    win = np.ascontiguousarray(prob.win)
    geom = np.zeros(6)
    nb = np.zeros(3, np.int32)
    d = prob.dist_max
    lib.or_bin_geometry(n, coo, win, float(d[0]), float(d[1]), float(d[2]), geom, nb)
    if ijk is None:
        ijk = np.zeros((n, 3), np.int32)
        lib.or_get_bins(n, x, geom, nb, ijk)
    rp = np.zeros(n + 1, np.int64)
    nnz = lib.or_soft_search(n, x, win, ijk, nb, int(prob.shsup), rp, None)
    col = np.zeros(max(nnz, 1), np.int32)
    lib.or_soft_search(n, x, win, ijk, nb, int(prob.shsup), rp, col.ctypes.data_as(C.c_void_p))
    return rp, col[:nnz], ijk


def oracle_shapes(lib, prob, rp, col, xs=None):
    """RK1 at every node's own point, nodal positions xs (default X0)."""
    n = prob.n
    xs = np.ascontiguousarray(prob.coo if xs is None else xs, dtype=np.float64)
    win = np.ascontiguousarray(prob.win)
    ebc = np.ascontiguousarray(prob.ebc_node, dtype=np.int32)
    c = control(prob)
    nnz = int(rp[-1])
    shp = np.zeros(nnz)
    dshp = np.zeros((nnz, 3))
    for i in range(n):
        a, b = int(rp[i]), int(rp[i + 1])
        ln = b - a
        s1 = np.zeros(ln)
        d1 = np.zeros(3 * ln)
        lib.or_rk1(xs[i].copy(), n, xs, win, ebc, np.ascontiguousarray(col[a:b]), ln, C.byref(c), s1, d1)
        shp[a:b] = s1
        dshp[a:b] = d1.reshape(ln, 3)
    return shp, dshp


def oracle_udfm(lib, prob, rp, col):
    """UDFM_SHAPE_TENSOR at every node: influence weights (nnz) and K^-1 (n, 9)."""
    n = prob.n
    coo = np.ascontiguousarray(prob.coo, dtype=np.float64)
    vol, win = np.ascontiguousarray(prob.vol), np.ascontiguousarray(prob.win)
    w = np.zeros(int(rp[-1]))
    invk = np.zeros((n, 9))
    for i in range(n):
        a, b = int(rp[i]), int(rp[i + 1])
        s1, k1 = np.zeros(b - a), np.zeros(9)
        lib.or_udfm(coo[i].copy(), n, coo, vol, win, np.ascontiguousarray(col[a:b]), b - a, s1, k1)
        w[a:b], invk[i] = s1, k1
    return w, invk


def oracle_fint_pd(lib, prob, rp, col, omega, invk, linit, dinc, dinc_tot, x_cur, stress, strain, state,
                   nthreads=1, dlt_fac=None, info=None):
    """One or_construct_fint_pd call; omega/invk/stress/strain/state updated in
    place.  Returns fint, wave_vel, status.  With dlt_fac the automatic time
    step is on and info["dlt"] receives DLT_FINT."""
    n = prob.n
    io = OrFintIO()
    keep = dict(win=np.ascontiguousarray(prob.win), vol=np.ascontiguousarray(prob.vol),
                coo0=np.ascontiguousarray(prob.coo, dtype=np.float64),
                coo_cur=np.ascontiguousarray(prob.coo if x_cur is None else x_cur, dtype=np.float64),
                props=np.ascontiguousarray(prob.props_table()), mat_id=np.ascontiguousarray(prob.mat_id, dtype=np.int32),
                mat_type=prob.mat_types(), ebc_node=np.ascontiguousarray(prob.ebc_node, dtype=np.int32),
                dinc=np.ascontiguousarray(dinc, dtype=np.float64), dinc_tot=np.ascontiguousarray(dinc_tot, dtype=np.float64),
                row_ptr=np.ascontiguousarray(rp, dtype=np.int64), col=np.ascontiguousarray(col, dtype=np.int32))
    fint, fext, wv = np.zeros((n, 3)), np.zeros((n, 3)), np.zeros(n)
    io.n = n
    for k, a in keep.items():
        setattr(io, k, _ptr(a))
    for k, a in dict(shp=omega, invk=invk, stress=stress, strain=strain, state=state).items():
        assert a.dtype == np.float64 and a.flags.c_contiguous
        setattr(io, k, _ptr(a))
    io.fint, io.fext, io.wave_vel = _ptr(fint), _ptr(fext), _ptr(wv)
    io.linit, io.nthreads = int(linit), nthreads
    cd = np.zeros(n)
    if dlt_fac is not None:
        io.auto_ts, io.dlt_fac, io.char_dist = 1, float(dlt_fac), _ptr(cd)
    st = lib.or_construct_fint_pd(C.byref(io), C.byref(control(prob)))
    if info is not None:
        info["dlt"] = io.dlt_fint
    return fint, wv, st


def oracle_fint(lib, prob, rp, col, shp, dshp, linit, dinc, dinc_tot, x_cur, stress, strain, state,
                nthreads=1):
    """One or_construct_fint call; shp/dshp/stress/strain/state updated in place.
    Returns fint, fext, int_work, status."""
    n = prob.n
    io = OrFintIO()
    keep = dict(win=np.ascontiguousarray(prob.win), vol=np.ascontiguousarray(prob.vol),
                coo0=np.ascontiguousarray(prob.coo, dtype=np.float64),
                coo_cur=np.ascontiguousarray(prob.coo if x_cur is None else x_cur, dtype=np.float64),
                props=np.ascontiguousarray(prob.props_table()), mat_id=np.ascontiguousarray(prob.mat_id, dtype=np.int32),
                mat_type=prob.mat_types(), ebc_node=np.ascontiguousarray(prob.ebc_node, dtype=np.int32),
                dinc=np.ascontiguousarray(dinc, dtype=np.float64), dinc_tot=np.ascontiguousarray(dinc_tot, dtype=np.float64),
                row_ptr=np.ascontiguousarray(rp, dtype=np.int64), col=np.ascontiguousarray(col, dtype=np.int32))
    fint = np.zeros((n, 3))
    fext = np.zeros((n, 3))
    wv = np.zeros(n)
    seq = np.zeros(n)
    io.n = n
    for k, a in keep.items():
        setattr(io, k, _ptr(a))
    for k, a in dict(shp=shp, dshp=dshp, stress=stress, strain=strain, state=state).items():
        assert a.dtype == np.float64 and a.flags.c_contiguous
        setattr(io, k, _ptr(a))
    io.fint, io.fext, io.wave_vel, io.strain_eq = _ptr(fint), _ptr(fext), _ptr(wv), _ptr(seq)
    io.bodyforce = None
    io.int_work, io.linit, io.nthreads = 0.0, int(linit), nthreads
    st = lib.or_construct_fint(C.byref(io), C.byref(control(prob)))
    return fint, fext, io.int_work, st
