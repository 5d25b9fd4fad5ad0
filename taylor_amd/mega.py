"""Python host for the MI355X explicit meshfree step: ctypes bindings of the
C-ABI in include/mega.h (taylor_amd/libmega_hip.so).

The methods follow the reference's stages and argument meanings:

    Mega.search          GET_BINS + SOFT_SEARCH        BIN_SEARCH.F90:2-328
    Mega.shapes          RK1 for all nodes             RK.f90:16-453
    Mega.interp          HANDELER DO_INTERP            HANDELER.F90:114-133
    Mega.construct_fint  CONSTRUCT_FINT + CONSTITUTION FINT.F90:16-1247
    Mega.step            MAIN time loop                MAIN.F90:304-658
    constitute           HUGHES_WINGET/ROTATE/VON_MISES per point
                                                       FINT.F90:758-801
    Mlp                  trainer/inference of the learned J2 return map
                         (no reference counterpart; stands in for VON_MISES,
                         CONSTITUTE_VONMISES.F90:21-158)
    Mega.set_surrogate / surrogate_stats / harvest / harvest_take

Errors come back as MegaError carrying the library status (the reference
stops the program instead).  The library must be built (taylor_amd/build.py);
nothing here falls back to a CPU path.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libmega_hip.so")

MEGA_FP64, MEGA_MIXED = 0, 1
MLP_PARAMS, MLP_NORM = 9671, 59
FIELDS = {"x": 0, "v": 1, "a": 2, "d": 3, "dtot": 4, "uphy": 5, "fint": 6, "fext": 7, "stress": 8,
          "strain": 9, "eps": 10, "strain_eq": 11, "wave_vel": 12, "vphy": 13, "damage_state": 14, "aphy": 15, "vphy_step": 16}
WIDTH = {"stress": 6, "strain": 6, "eps": 1, "strain_eq": 1, "wave_vel": 1, "damage_state": 4}


class MegaError(RuntimeError):
    def __init__(self, status: int, what: str):
        self.status = status
        super().__init__(f"{what}: status {status} ({_status_string(status)})")


class MegaControl(C.Structure):
    _fields_ = [("rk_degree", C.c_int32), ("rk_cont", C.c_int32), ("itype_int", C.c_int32), ("ql", C.c_int32),
                ("ql_coef", C.c_double), ("ql_len", C.c_double), ("shsup", C.c_int32), ("lagrangian", C.c_int32),
                ("finite_strain", C.c_int32), ("pdsearch", C.c_int32), ("gravity", C.c_double * 3),
                ("dist_max", C.c_double * 3), ("precision", C.c_int32), ("surrogate", C.c_int32),
                ("peridynamics", C.c_int32), ("reserved", C.c_int32)]


_lib = None


def load_library():
    """Load libmega_hip.so; raise if it was not built (no CPU fallback)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"{LIB_PATH} is missing: build it with python taylor_amd/build.py")
    lib = C.CDLL(LIB_PATH)
    vp, dp, ip, i32, i64 = C.c_void_p, C.POINTER(C.c_double), C.POINTER(C.c_int32), C.c_int32, C.c_int64
    sig = {
        "mega_create": [C.POINTER(MegaControl), i32, vp, vp, vp, vp, vp, vp, vp, vp, vp, C.POINTER(vp)],
        "mega_destroy": [vp],
        "mega_search": [vp, vp, i32],
        "mega_get_neighbors": [vp, vp, vp, vp, i64, C.POINTER(i64)],
        "mega_shapes": [vp, vp],
        "mega_get_shapes": [vp, vp, vp, i64],
        "mega_get_pd_shape": [vp, vp, vp, i64],
        "mega_interp": [vp, vp, vp],
        "mega_fint": [vp, i32, vp, vp, vp, vp, vp, vp, vp, vp, dp],
        "mega_constitute": [i32, vp, vp, vp, vp, vp, i32, i32, i32],
        "mega_preprocess": [i32, vp, i32, vp, vp, vp, vp, vp, vp],
        "mega_preprocess_noalloc": [i32, vp, i32, vp, vp, vp, vp, vp, vp],
        "mega_step": [vp, C.c_double, i32],
        "mega_step_auto": [vp, C.c_double, i32, vp],
        "mega_download": [vp, i32, vp],
        "mega_upload": [vp, i32, vp],
        "mega_energies": [vp, dp, dp, dp],
        "mega_stats": [vp, C.POINTER(i64), ip, C.POINTER(i64)],
        "mega_profile": [vp, i32],
        "mega_kernel_times": [vp, C.POINTER(C.c_char_p), dp, ip, i32],
        "mega_mlp_create": [C.c_uint64, C.POINTER(vp)],
        "mega_mlp_destroy": [vp],
        "mega_mlp_set_params": [vp, vp],
        "mega_mlp_get_params": [vp, vp],
        "mega_mlp_get_norm": [vp, vp],
        "mega_mlp_set_data": [vp, vp, i64, vp],
        "mega_mlp_train": [vp, i32, i32, C.c_double, C.c_double, vp],
        "mega_mlp_forward": [vp, vp, vp, i64],
        "mega_mlp_gradient": [vp, vp, vp, i64, vp, C.POINTER(C.c_float)],
        "mega_set_surrogate": [vp, vp, vp, vp, C.c_double],
        "mega_surrogate_stats": [vp, C.POINTER(i64), i32],
        "mega_harvest": [vp, i64, i32],
        "mega_harvest_take": [vp, vp, i64, C.POINTER(i64)],
    }
    for name, args in sig.items():
        f = getattr(lib, name)
        f.argtypes = args
        f.restype = C.c_int
    lib.mega_status_string.argtypes = [C.c_int]
    lib.mega_status_string.restype = C.c_char_p
    _lib = lib
    return lib


def _status_string(s: int) -> str:
    try:
        return load_library().mega_status_string(s).decode()
    except Exception:
        return "?"


def _p(a: np.ndarray | None):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


def _chk(status: int, what: str, allow=()):
    if status != 0 and status not in allow:
        raise MegaError(status, what)
    return status


def _f64(a, shape=None):
    a = np.ascontiguousarray(a, dtype=np.float64)
    return a if shape is None else a.reshape(shape)


def _i32(a):
    return np.ascontiguousarray(a, dtype=np.int32)


def make_control(prob, precision: int = MEGA_FP64, surrogate: int = 0) -> MegaControl:
    c = MegaControl()
    c.rk_degree, c.rk_cont, c.itype_int = 1, 3, int(getattr(prob, "itype_int", 0))
    c.ql, c.ql_coef, c.ql_len = int(prob.ql), float(prob.ql_coef), float(prob.ql_len)
    c.shsup, c.lagrangian, c.finite_strain = int(prob.shsup), int(prob.lagrangian), int(prob.finite_strain)
    c.pdsearch = int(prob.pdsearch)
    for k in range(3):
        c.gravity[k] = float(prob.gravity[k])
        c.dist_max[k] = float(prob.dist_max[k])
    c.precision, c.surrogate = int(precision), int(surrogate)
    c.peridynamics = int(getattr(prob, "peridynamics", 0))
    return c


class Mega:
    """One device context holding a model (positions, windows, state, lists)."""

    def __init__(self, prob, precision: int = MEGA_FP64, surrogate: int = 0):
        lib = load_library()
        self.lib, self.prob, self.n = lib, prob, prob.n
        self.ctrl = make_control(prob, precision, surrogate)
        n = prob.n
        props = prob.props_table()[prob.mat_id]
        mtype = prob.mat_types()[prob.mat_id]
        keep = [_f64(prob.coo), _f64(prob.win), _f64(prob.vol), _f64(prob.mass), _i32(prob.ebc),
                _i32(prob.ebc_node), _i32(mtype), _f64(props), _f64(prob.vinit)]
        ctx = C.c_void_p()
        _chk(lib.mega_create(C.byref(self.ctrl), n, *[_p(a) for a in keep], C.byref(ctx)), "mega_create")
        self.ctx = ctx

    def close(self):
        if getattr(self, "ctx", None):
            self.lib.mega_destroy(self.ctx)
            self.ctx = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # -- stages --------------------------------------------------------------
    def search(self, x=None, rebuild_bins: bool = True):
        xs = None if x is None else _f64(x)
        _chk(self.lib.mega_search(self.ctx, _p(xs), int(rebuild_bins)), "mega_search")

    def neighbors(self):
        """(GN, GSTART, GSTACK) in the reference's 1-based form."""
        nnz = C.c_int64()
        _chk(self.lib.mega_get_neighbors(self.ctx, None, None, None, 0, C.byref(nnz)), "mega_get_neighbors")
        gn = np.zeros(self.n, np.int32)
        gs = np.zeros(self.n, np.int32)
        st = np.zeros(max(nnz.value, 1), np.int32)
        _chk(self.lib.mega_get_neighbors(self.ctx, _p(gn), _p(gs), _p(st), st.size, C.byref(nnz)), "mega_get_neighbors")
        return gn, gs, st[:nnz.value]

    def shapes(self, x_eval=None):
        xs = None if x_eval is None else _f64(x_eval)
        return _chk(self.lib.mega_shapes(self.ctx, _p(xs)), "mega_shapes", allow=(2,))

    def get_shapes(self):
        gn, _, _ = self.neighbors()
        nnz = int(gn.sum())
        shp = np.zeros(max(nnz, 1))
        dshp = np.zeros((max(nnz, 1), 3))
        _chk(self.lib.mega_get_shapes(self.ctx, _p(shp), _p(dshp), nnz), "mega_get_shapes")
        return shp[:nnz], dshp[:nnz]

    def pd_shape(self):
        """(omega[nnz], invk[n, 9]) of the peridynamic shape tensors."""
        nnz, _, _ = self.stats()
        w, k = np.zeros(max(nnz, 1)), np.zeros((self.n, 9))
        _chk(self.lib.mega_get_pd_shape(self.ctx, _p(w), _p(k), nnz), "mega_get_pd_shape")
        return w[:nnz], k

    def interp(self, coef):
        coef = _f64(coef, (self.n, 3))
        out = np.zeros((self.n, 3))
        _chk(self.lib.mega_interp(self.ctx, _p(coef), _p(out)), "mega_interp")
        return out

    def construct_fint(self, linit, dinc, dinc_tot, x_cur, stress, strain, state, int_work=0.0):
        """One CONSTRUCT_FINT call; stress/strain/state are updated in place.
        Returns (fint, fext, int_work, status)."""
        for a, w in ((stress, 6), (strain, 6), (state, 20)):
            assert a.dtype == np.float64 and a.flags.c_contiguous and a.size == w * self.n
        fint = np.zeros((self.n, 3))
        fext = np.zeros((self.n, 3))
        wk = C.c_double(int_work)
        st = self.lib.mega_fint(self.ctx, int(linit), _p(_f64(dinc)), _p(_f64(dinc_tot)),
                                _p(None if x_cur is None else _f64(x_cur)), _p(stress), _p(strain), _p(state),
                                _p(fint), _p(fext), C.byref(wk))
        _chk(st, "mega_fint", allow=(1, 2))
        return fint, fext, wk.value, st

    def step(self, nsteps: int = 1, dt: float | None = None):
        st = self.lib.mega_step(self.ctx, float(dt or self.prob.dt), int(nsteps))
        _chk(st, "mega_step", allow=(1, 2))
        return st

    def step_auto(self, dlt_fac: float, nsteps: int = 1):
        """nsteps with the automatic time step (IAUTO_TS = 1, peridynamics);
        returns (status, each step's DLT)."""
        dlt = np.zeros(max(int(nsteps), 1))
        st = self.lib.mega_step_auto(self.ctx, float(dlt_fac), int(nsteps), _p(dlt))
        _chk(st, "mega_step_auto", allow=(1, 2))
        return st, dlt[:int(nsteps)]

    def download(self, name: str) -> np.ndarray:
        w = WIDTH.get(name, 3)
        out = np.zeros(self.n * w)
        _chk(self.lib.mega_download(self.ctx, FIELDS[name], _p(out)), "mega_download")
        return out.reshape(self.n, w) if w > 1 else out

    def upload(self, name: str, arr):
        _chk(self.lib.mega_upload(self.ctx, FIELDS[name], _p(_f64(arr))), "mega_upload")

    def energies(self):
        a, b, c = C.c_double(), C.c_double(), C.c_double()
        _chk(self.lib.mega_energies(self.ctx, C.byref(a), C.byref(b), C.byref(c)), "mega_energies")
        return a.value, b.value, c.value

    def stats(self):
        nnz, mx, fb = C.c_int64(), C.c_int32(), C.c_int64()
        _chk(self.lib.mega_stats(self.ctx, C.byref(nnz), C.byref(mx), C.byref(fb)), "mega_stats")
        return nnz.value, mx.value, fb.value

    def profile(self, on: bool | int = True):
        """Per-kernel HIP-event timing of step(): False/0 off, True/1 every
        kernel, 2 the force kernel only."""
        _chk(self.lib.mega_profile(self.ctx, int(on)), "mega_profile")

    def set_surrogate(self, params, norm, props, tol: float = 1e-3):
        """Enable the learned return map for J2 nodes of property row props[30]
        (params None disables)."""
        if params is None:
            _chk(self.lib.mega_set_surrogate(self.ctx, None, None, None, 0.0), "mega_set_surrogate")
            return
        self._surr = (np.ascontiguousarray(params, np.float32), np.ascontiguousarray(norm, np.float32),
                      _f64(props))
        _chk(self.lib.mega_set_surrogate(self.ctx, _p(self._surr[0]), _p(self._surr[1]), _p(self._surr[2]),
                                         float(tol)), "mega_set_surrogate")

    def surrogate_stats(self, reset: bool = False) -> dict:
        s = (C.c_int64 * 4)()
        _chk(self.lib.mega_surrogate_stats(self.ctx, s, int(reset)), "mega_surrogate_stats")
        return {"plastic": s[0], "outside_envelope": s[1], "evaluated": s[2], "rejected": s[3]}

    def harvest(self, capacity: int, every: int = 1):
        _chk(self.lib.mega_harvest(self.ctx, int(capacity), int(every)), "mega_harvest")

    def harvest_take(self, cap: int) -> tuple[np.ndarray, int]:
        """(tuples[m, 20] float32, number recorded since the last take)"""
        out = np.zeros((max(int(cap), 1), 20), np.float32)
        cnt = C.c_int64()
        _chk(self.lib.mega_harvest_take(self.ctx, _p(out), int(cap), C.byref(cnt)), "mega_harvest_take")
        m = min(cnt.value, int(cap))
        return out[:m].copy(), cnt.value

    def kernel_times(self) -> dict:
        cap = 32
        names = (C.c_char_p * cap)()
        ms = (C.c_double * cap)()
        cnt = (C.c_int32 * cap)()
        k = self.lib.mega_kernel_times(self.ctx, names, ms, cnt, cap)
        return {names[i].decode(): (ms[i], cnt[i]) for i in range(k)}


def constitute(lmat, stress, strain, state, props, mat_type: int = 2, finite_strain: int = 1,
               precision: int = MEGA_FP64) -> int:
    """Per-point HUGHES_WINGET + ROTATE_TENSOR + elastic trial + VON_MISES on the
    GPU (FINT.F90:758-801); stress/strain/state updated in place."""
    lib = load_library()
    n = lmat.shape[0]
    for a in (stress, strain, state):
        assert a.dtype == np.float64 and a.flags.c_contiguous
    st = lib.mega_constitute(n, _p(_f64(lmat)), _p(stress), _p(strain), _p(state), _p(_f64(props)),
                             int(mat_type), int(finite_strain), int(precision))
    return _chk(st, "mega_constitute", allow=(1,))


def preprocess(coo, elcon, hex=None):
    """GET_ADIAL + GET_VOLUMES + GET_SMOOTHING_LENGTHS on the GPU
    (PRE_MODEL.F90:248-262): coo (n, 3), elcon (numel, 8) 1-based node ids,
    hex (numel,) 1 = hexahedron / 0 = tetrahedron or None (all hexahedra).
    Returns (win (n, 3), vol (n,), sm_len (n, 6), dist_max (3,))."""
    lib = load_library()
    coo = _f64(coo)
    el = _i32(elcon)
    n, numel = coo.shape[0], el.shape[0]
    assert coo.shape == (n, 3) and el.shape == (numel, 8)
    hx = None if hex is None else _i32(hex)
    if hx is not None:
        assert hx.shape == (numel,)
    win, vol, sm, dmax = np.zeros((n, 3)), np.zeros(n), np.zeros((n, 6)), np.zeros(3)
    st = lib.mega_preprocess(n, _p(coo), numel, _p(el), _p(hx), _p(win), _p(vol), _p(sm), _p(dmax))
    _chk(st, "mega_preprocess")
    return win, vol, sm, dmax


def geometry(mesh):
    """deck.Geometry of a deck.Mesh from the device preprocessing (the same
    windows and volumes as deck.preprocess, bit for bit)."""
    from taylor_amd import deck
    win, vol, sm, dmax = preprocess(mesh.coo, mesh.elcon)
    return deck.Geometry(win0=win, vol=vol, sm_len=sm, dist_max=dmax)


class Mlp:
    """Trainer / batch inference of the 13-64-64-64-7 network on the GPU."""

    def __init__(self, seed: int = 0):
        self.lib = load_library()
        h = C.c_void_p()
        _chk(self.lib.mega_mlp_create(C.c_uint64(seed), C.byref(h)), "mega_mlp_create")
        self.h = h

    def close(self):
        if getattr(self, "h", None):
            self.lib.mega_mlp_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def params(self) -> np.ndarray:
        out = np.zeros(MLP_PARAMS, np.float32)
        _chk(self.lib.mega_mlp_get_params(self.h, _p(out)), "mega_mlp_get_params")
        return out

    @params.setter
    def params(self, p):
        p = np.ascontiguousarray(p, np.float32)
        assert p.size == MLP_PARAMS
        _chk(self.lib.mega_mlp_set_params(self.h, _p(p)), "mega_mlp_set_params")

    @property
    def norm(self) -> np.ndarray:
        out = np.zeros(MLP_NORM, np.float32)
        _chk(self.lib.mega_mlp_get_norm(self.h, _p(out)), "mega_mlp_get_norm")
        return out

    def set_data(self, tuples, props):
        t = np.ascontiguousarray(tuples, np.float32).reshape(-1, 20)
        _chk(self.lib.mega_mlp_set_data(self.h, _p(t), t.shape[0], _p(_f64(props))), "mega_mlp_set_data")

    def train(self, iters: int, batch: int = 8192, lr0: float = 2e-3, lr1: float = 2e-5) -> np.ndarray:
        hist = np.zeros(max(iters, 1), np.float32)
        _chk(self.lib.mega_mlp_train(self.h, int(iters), int(batch), float(lr0), float(lr1), _p(hist)),
             "mega_mlp_train")
        return hist[:iters]

    def forward(self, xhat) -> np.ndarray:
        x = np.ascontiguousarray(xhat, np.float32).reshape(-1, 13)
        y = np.zeros((x.shape[0], 7), np.float32)
        _chk(self.lib.mega_mlp_forward(self.h, _p(x), _p(y), x.shape[0]), "mega_mlp_forward")
        return y

    def gradient(self, xhat, that):
        x = np.ascontiguousarray(xhat, np.float32).reshape(-1, 13)
        t = np.ascontiguousarray(that, np.float32).reshape(-1, 7)
        g = np.zeros(MLP_PARAMS, np.float32)
        loss = C.c_float()
        _chk(self.lib.mega_mlp_gradient(self.h, _p(x), _p(t), x.shape[0], _p(g), C.byref(loss)),
             "mega_mlp_gradient")
        return g, loss.value


def write_model(path: str, prob, nsteps: int, chunk: int = 500, precision: int = MEGA_FP64) -> None:
    """Binary model file of the Fortran host (taylor_amd/host/mega_drive.f90):
    int32 n, nsteps, chunk; real64 dt; the mega_control struct; then the arrays
    mega_create takes, in its argument order."""
    import struct
    ctrl = make_control(prob, precision, 0)
    props = prob.props_table()[prob.mat_id]
    mtype = prob.mat_types()[prob.mat_id]
    with open(path, "wb") as f:
        f.write(struct.pack("<iiid", prob.n, int(nsteps), int(chunk), float(prob.dt)))
        f.write(bytes(ctrl))
        for a in (_f64(prob.coo), _f64(prob.win), _f64(prob.vol), _f64(prob.mass), _i32(prob.ebc),
                  _i32(prob.ebc_node), _i32(mtype), _f64(props), _f64(prob.vinit)):
            f.write(a.tobytes())
