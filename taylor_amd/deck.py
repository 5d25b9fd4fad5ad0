"""MODEL/CONTROL decks for the explicit meshfree step: synthetic Taylor-bar
generator, LS-DYNA ``.k`` / ``Control.dat`` writer and reader, and the one-time
geometry preprocessing the step consumes (windows, nodal volumes, smoothing
lengths, mass).

Reference behaviour restated here (file:line in the reference program):

* ``.k`` node, hex and node-set formats: ``(I8,3F16.0,2I8)``, ``(10I8)``,
  ``(8I10)``                                              PRE_MODEL.F90:163,178,199
* window = largest per-axis extent of any attached hex  PRE_MODEL_ADIAL.F90:64-114
* nodal volume = 1/8 of each attached hex's volume, hex volume from three
  determinants about the 1-7 diagonal                   PRE_MODEL_VOLUMES.F90:70-92,
                                                        PRE_MODEL_VOL2.F90:36-52
* smoothing lengths from the unscaled window and volume PRE_MODEL_SMOOTHING_LENGTHS.F90:49-99
* window scaled by the deck's normalised window after QL_LEN is fixed
                                                        PRE_CONTROL.F90:803-872
* Lame constants into props 28-30, mass = V * rho     READ.F90:496-507, PRE_CONTROL.F90:972-978
* EBC-node flag: any fixity equal to 1                  STATE_INIT.f90:102-109

The accumulation order of volumes follows the reference's element order, so the
values match the reference bit for bit.
"""
from __future__ import annotations

import math
import os
from dataclasses import dataclass, field

import numpy as np

MASK64 = (1 << 64) - 1


def splitmix64(seed: int, count: int) -> np.ndarray:
    """``count`` 64-bit outputs of SplitMix64 (Steele, Lea, Flood 2014) from ``seed``."""
    with np.errstate(over="ignore"):
        idx = np.arange(1, count + 1, dtype=np.uint64)
        z = np.uint64(seed & MASK64) + idx * np.uint64(0x9E3779B97F4A7C15)
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        z = z ^ (z >> np.uint64(31))
    return z


def uniform_pm(seed: int, count: int) -> np.ndarray:
    """Uniform deviates on [-1, 1) from the top 53 bits of SplitMix64."""
    u = (splitmix64(seed, count) >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)
    return 2.0 * u - 1.0


@dataclass
class Material:
    """One ``*MATERIALS`` entry: type and the 30-value property row (1-based in
    the reference: 1 nu, 2 E, 3 rho, 4 sigma_0, 5 B, 6 CE, 21 damping)."""
    mat_type: int
    props: np.ndarray

    @staticmethod
    def j2(E: float, nu: float, rho: float, sig0: float, B: float, CE: float, damping: float = 0.0) -> "Material":
        p = np.zeros(30)
        p[0], p[1], p[2], p[3], p[4], p[5] = nu, E, rho, sig0, B, CE
        p[20] = damping
        return Material(2, fill_lame(p))

    @staticmethod
    def j2_damage(E: float, nu: float, rho: float, sig0: float, B: float, CE: float, k_i: float, k_c: float,
                  dam_max: float, damping: float = 0.0) -> "Material":
        """Type 6, VON_MISES_DAM (CONSTITUTE_VONMISES_DAM.F90): J2 as type 2 plus
        isotropic damage with props 9-11 = K_I, K_C, DAM_MAX."""
        p = np.zeros(30)
        p[0], p[1], p[2], p[3], p[4], p[5] = nu, E, rho, sig0, B, CE
        p[8], p[9], p[10] = k_i, k_c, dam_max
        p[20] = damping
        return Material(6, fill_lame(p))

    @staticmethod
    def drucker_prager(E: float, nu: float, rho: float, dp_type: int, phi_deg: float, c: float, psi_deg: float,
                       t_cut: float, k_i: float, k_c: float, dam_max: float, damping: float = 0.0) -> "Material":
        """Type 3, DRUCK_PRAG (CONSTITUTE_DRUCKPRAG.F90): props 4-7 = DP_TYPE, PHI,
        C, PSI as the deck gives them, 8-11 = T_CUT, K_I, K_C, DAM_MAX; props
        25-27 prepared as the reader does (drucker_prager_prep)."""
        p = np.zeros(30)
        p[0], p[1], p[2] = nu, E, rho
        p[3], p[4], p[5], p[6] = float(dp_type), phi_deg, c, psi_deg
        p[7], p[8], p[9], p[10] = t_cut, k_i, k_c, dam_max
        p[20] = damping
        return Material(3, fill_lame(drucker_prager_prep(p)))

    @staticmethod
    def elastic(E: float, nu: float, rho: float) -> "Material":
        p = np.zeros(30)
        p[0], p[1], p[2] = nu, E, rho
        return Material(1, fill_lame(p))

    def in_mm(self) -> "Material":
        """The same material in mm / tonne / s / MPa: moduli and stresses x1e-6,
        density x1e-12; strains, angles, exponents and damping unchanged."""
        p = self.props
        E, nu, rho, damping = p[1] * 1e-6, p[0], p[2] * 1e-12, p[20]
        if self.mat_type == 1:
            return Material.elastic(E, nu, rho)
        if self.mat_type == 2:
            return Material.j2(E, nu, rho, p[3] * 1e-6, p[4], p[5], damping)
        if self.mat_type == 6:
            return Material.j2_damage(E, nu, rho, p[3] * 1e-6, p[4], p[5], p[8], p[9], p[10], damping)
        if self.mat_type == 3:
            return Material.drucker_prager(E, nu, rho, int(p[3]), p[4], p[5] * 1e-6, p[6], p[7] * 1e-6, p[8], p[9],
                                           p[10], damping)
        raise ValueError(f"material type {self.mat_type}: no unit conversion")


def drucker_prager_prep(p: np.ndarray) -> np.ndarray:
    """Props 25-27 = Q_PHI, K_PHI, Q_PSI of the Drucker-Prager cone from
    DP_TYPE, PHI (degrees), C = props 4-6 (READ.F90:420-483): 1 circumscribes,
    2 middle-circumscribes, 3 inscribes the Mohr-Coulomb surface, 4 takes
    Q_PHI/3, K_PHI, Q_PSI as given; Q_PSI is then reset to 0 (no dilatation)."""
    p = np.array(p, dtype=np.float64)
    dp_type, phi, c = int(p[3]), p[4], p[5]
    phi_rad = phi * 3.141592653589793 / 180.0
    srt32 = math.sqrt(3.0 / 2.0)
    s = math.sin(phi_rad)
    q_phi, k_phi = p[24], p[25]
    if dp_type == 1:
        q_phi = 2.0 * s / (srt32 * (3.0 - s)) * 3.0
        k_phi = 6.0 * c * math.cos(phi_rad) / (srt32 * (3.0 - s))
    elif dp_type == 2:
        q_phi = 2.0 * s / (srt32 * (3.0 + s)) * 3.0
        k_phi = 6.0 * c * math.cos(phi_rad) / (srt32 * (3.0 + s))
    elif dp_type == 3:
        q_phi = s / math.sqrt(9.0 + 3.0 * s * s) / math.sqrt(0.5) * 3.0
        k_phi = 3.0 * c * math.cos(phi_rad) / math.sqrt(9.0 + 3.0 * s * s) / math.sqrt(0.5)
    elif dp_type == 4:
        q_phi = p[4] * 3.0
        k_phi = p[5]
    p[24], p[25], p[26] = q_phi, k_phi, 0.0
    return p


def fill_lame(p: np.ndarray) -> np.ndarray:
    """Props 28-30 = lambda, mu, lambda + 2 mu (READ.F90:496-507)."""
    p = np.array(p, dtype=np.float64)
    nu, E = p[0], p[1]
    lam = E * nu / ((1.0 + nu) * (1.0 - 2.0 * nu))
    mu = E / (2.0 * (1.0 + nu))
    p[27], p[28], p[29] = lam, mu, lam + 2.0 * mu
    return p


def copper_linear() -> Material:
    """Copper, J2 with linear hardening H = 100 MPa as B = H/sigma_0, CE = 1."""
    return Material.j2(E=117e9, nu=0.35, rho=8930.0, sig0=400e6, B=100e6 / 400e6, CE=1.0)


def p_wave_speed(mat: Material) -> float:
    nu, E, rho = mat.props[0], mat.props[1], mat.props[2]
    K = E / (3.0 * (1.0 - 2.0 * nu))
    G = E / (2.0 * (1.0 + nu))
    return math.sqrt((K + 4.0 * G / 3.0) / rho)


@dataclass
class Mesh:
    coo: np.ndarray            # (n, 3) float64
    elcon: np.ndarray          # (nel, 8) int32, 1-based node ids
    sets: dict                 # name -> 1-based node ids (int32)


def taylor_bar_mesh(h: float, radius: float = 3.8e-3, length: float = 25.4e-3,
                    jitter: float = 0.05, seed: int = 0) -> Mesh:
    """Jittered cubic-lattice cylinder, axis along z, impact face at z = 0.

    Lattice cells whose centre lies within ``radius`` of the axis become 8-node
    hexes (LS-DYNA node order: n1-n4 bottom counter-clockwise, n5-n8 top);
    only nodes of some hex are kept, numbered 1..N in (z, y, x) lattice order.
    Every coordinate gets a uniform jitter of +/- ``jitter`` * h from SplitMix64
    (x, y, z per node in node order), except z on the impact face, which stays
    on the wall plane z = 0.  Coordinates are rounded through the deck's 16-wide
    text field so that the reference and this build read identical values.
    """
    nl = int(round(length / h))
    m = int(math.ceil(radius / h)) + 1
    ii = np.arange(-m, m)
    ci, cj = np.meshgrid(ii, ii, indexing="xy")
    ctr = ((ci + 0.5) * h) ** 2 + ((cj + 0.5) * h) ** 2 <= radius * radius
    cells_ij = np.stack([ci[ctr], cj[ctr]], axis=1)           # (ncell, 2) in (j, i) row-major order
    # node lattice points used by any cell of a layer
    used = np.zeros((2 * m + 1, 2 * m + 1), dtype=bool)       # [j+m, i+m]
    for di in (0, 1):
        for dj in (0, 1):
            used[cells_ij[:, 1] + m + dj, cells_ij[:, 0] + m + di] = True
    jj, iiu = np.nonzero(used)                                 # sorted by j then i
    per_layer = len(jj)
    lid = -np.ones_like(used, dtype=np.int64)
    lid[jj, iiu] = np.arange(per_layer)
    nnode = per_layer * (nl + 1)
    k = np.repeat(np.arange(nl + 1), per_layer)
    x = np.tile((iiu - m).astype(np.float64), nl + 1) * h
    y = np.tile((jj - m).astype(np.float64), nl + 1) * h
    z = k.astype(np.float64) * h
    jit = uniform_pm(seed, 3 * nnode).reshape(nnode, 3) * (jitter * h)
    coo = np.stack([x, y, z], axis=1) + jit
    coo[k == 0, 2] = 0.0
    coo = np.array([float("%16.8e" % v) for v in coo.ravel()]).reshape(nnode, 3)
    # hexes, layer by layer
    a = lid[cells_ij[:, 1] + m, cells_ij[:, 0] + m]
    b = lid[cells_ij[:, 1] + m, cells_ij[:, 0] + m + 1]
    c = lid[cells_ij[:, 1] + m + 1, cells_ij[:, 0] + m + 1]
    d = lid[cells_ij[:, 1] + m + 1, cells_ij[:, 0] + m]
    quad = np.stack([a, b, c, d], axis=1)
    els = []
    for layer in range(nl):
        lo = quad + layer * per_layer
        hi = quad + (layer + 1) * per_layer
        els.append(np.concatenate([lo, hi], axis=1))
    elcon = (np.concatenate(els, axis=0) + 1).astype(np.int32)
    bar = np.arange(1, nnode + 1, dtype=np.int32)
    wall = np.arange(1, per_layer + 1, dtype=np.int32)
    return Mesh(coo=coo, elcon=elcon, sets={"BAR": bar, "WALL": wall})


def plate_mesh(h: float, lx: float, ly: float, lz: float, jitter: float = 0.05, seed: int = 3) -> Mesh:
    """Jittered cubic-lattice box [0,lx]x[0,ly]x[0,lz] (peridynamic plate)."""
    nx, ny, nz = int(round(lx / h)), int(round(ly / h)), int(round(lz / h))
    gi, gj, gk = np.meshgrid(np.arange(nx + 1), np.arange(ny + 1), np.arange(nz + 1), indexing="ij")
    # node order (k, j, i)
    order = np.lexsort((gi.ravel(), gj.ravel(), gk.ravel()))
    I, J, K = gi.ravel()[order], gj.ravel()[order], gk.ravel()[order]
    n = len(I)
    coo = np.stack([I * h, J * h, K * h], axis=1).astype(np.float64)
    coo += uniform_pm(seed, 3 * n).reshape(n, 3) * (jitter * h)
    coo = np.array([float("%16.8e" % v) for v in coo.ravel()]).reshape(n, 3)
    nid = np.empty((nx + 1, ny + 1, nz + 1), dtype=np.int64)
    nid[I, J, K] = np.arange(n)
    ci, cj, ck = np.meshgrid(np.arange(nx), np.arange(ny), np.arange(nz), indexing="ij")
    o = np.lexsort((ci.ravel(), cj.ravel(), ck.ravel()))
    ci, cj, ck = ci.ravel()[o], cj.ravel()[o], ck.ravel()[o]
    el = np.stack([nid[ci, cj, ck], nid[ci + 1, cj, ck], nid[ci + 1, cj + 1, ck], nid[ci, cj + 1, ck],
                   nid[ci, cj, ck + 1], nid[ci + 1, cj, ck + 1], nid[ci + 1, cj + 1, ck + 1], nid[ci, cj + 1, ck + 1]],
                  axis=1)
    return Mesh(coo=coo, elcon=(el + 1).astype(np.int32), sets={"PLATE": np.arange(1, n + 1, dtype=np.int32)})


# ----------------------------------------------------------------------------
# geometry preprocessing (one-time work before the time loop)

def _det3_abs(A):
    """DETERMINANT.F90:18-35 on stacked 3x3 (..., 3, 3), columns as in the reference."""
    return np.abs(A[..., 0, 0] * (A[..., 1, 1] * A[..., 2, 2] - A[..., 2, 1] * A[..., 1, 2])
                  + A[..., 0, 1] * (A[..., 2, 0] * A[..., 1, 2] - A[..., 1, 0] * A[..., 2, 2])
                  + A[..., 0, 2] * (A[..., 1, 0] * A[..., 2, 1] - A[..., 2, 0] * A[..., 1, 1]))


def hex_volumes(coo: np.ndarray, elcon: np.ndarray) -> np.ndarray:
    """GET_VOL2 (PRE_MODEL_VOL2.F90:36-52) for every hex."""
    X = coo[elcon - 1]                      # (nel, 8, 3)
    p = [X[:, i, :] for i in range(8)]
    d71 = p[6] - p[0]
    A = np.stack([d71, p[1] - p[0], p[2] - p[5]], axis=2)
    B = np.stack([d71, p[4] - p[0], p[5] - p[7]], axis=2)
    C = np.stack([d71, p[3] - p[0], p[7] - p[2]], axis=2)
    return (_det3_abs(A) + _det3_abs(B) + _det3_abs(C)) / 6.0


@dataclass
class Geometry:
    win0: np.ndarray     # (n, 3) unscaled window (ADIAL)
    vol: np.ndarray      # (n,)
    sm_len: np.ndarray   # (n, 6)
    dist_max: np.ndarray  # (3,) XDIST_MAX, YDIST_MAX, ZDIST_MAX


def preprocess(mesh: Mesh) -> Geometry:
    coo, el = mesh.coo, mesh.elcon
    n = coo.shape[0]
    X = coo[el - 1]                                          # (nel, 8, 3)
    ext = X.max(axis=1) - X.min(axis=1)                      # per-axis max |xk - xj|
    win0 = np.zeros((n, 3))
    for k in range(8):                                       # GET_ADIAL keeps the max
        np.maximum.at(win0, el[:, k] - 1, ext)
    dv = hex_volumes(coo, el)
    vol = np.zeros(n)
    contrib = 0.125 * dv
    # reference order: element by element, corners 1..8
    np.add.at(vol, (el - 1).ravel(), np.repeat(contrib, 8))
    thrd = 0.333333333333333
    dims = win0 * 2.0
    vol_win = dims[:, 0] * dims[:, 1] * dims[:, 2]
    # scalar libm pow, as the reference calls it (vectorised pow may differ by an ulp)
    ratio = np.array([math.pow(a, thrd) / math.pow(b, thrd) for a, b in zip(vol.tolist(), vol_win.tolist())])
    ratio = ratio * 1.0
    sm = np.empty((n, 6))
    sm[:, 0] = sm[:, 1] = dims[:, 0] * ratio / 2.0
    sm[:, 2] = sm[:, 3] = dims[:, 1] * ratio / 2.0
    sm[:, 4] = sm[:, 5] = dims[:, 2] * ratio / 2.0
    dmax = np.array([max(0.0, sm[:, 0].max()), max(0.0, sm[:, 2].max()), max(0.0, sm[:, 5].max())])
    return Geometry(win0=win0, vol=vol, sm_len=sm, dist_max=dmax)


# ----------------------------------------------------------------------------
# the assembled problem handed to the step

@dataclass
class Problem:
    """Everything the explicit step needs, in the reference layout."""
    coo: np.ndarray          # (n, 3)
    win: np.ndarray          # (n, 3) scaled window
    vol: np.ndarray          # (n,)
    mass: np.ndarray         # (n, 3)
    ebc: np.ndarray          # (n, 3) int32 fixity
    ebc_node: np.ndarray     # (n,) int32
    vinit: np.ndarray        # (n, 3) initial velocity (0 on fixed dofs)
    mat_id: np.ndarray       # (n,) int32
    materials: list          # [Material]
    dist_max: np.ndarray     # (3,)
    sets: dict
    dt: float
    time_end: float
    ql: int = 1
    ql_coef: float = 0.001
    ql_len: float = 0.0
    lagrangian: int = 1
    finite_strain: int = 1
    shsup: int = 0
    pdsearch: int = 1
    norm_win: float = 1.84
    gravity: np.ndarray = field(default_factory=lambda: np.zeros(3))
    elcon: np.ndarray | None = None
    peridynamics: int = 0    # CONSTRUCT_FINT_PD instead of CONSTRUCT_FINT (*PERIDYNAMICS 1)
    itype_int: int = 0       # *INTEGRATION PARAMETERS: 0 direct nodal integration, 1 smoothed gradients

    @property
    def n(self) -> int:
        return self.coo.shape[0]

    def props_table(self) -> np.ndarray:
        return np.stack([m.props for m in self.materials]).astype(np.float64)

    def mat_types(self) -> np.ndarray:
        return np.array([m.mat_type for m in self.materials], dtype=np.int32)


def taylor_problem(h: float, v0: float = 190.0, seed: int = 0, norm_win: float = 1.84,
                   radius: float = 3.8e-3, length: float = 25.4e-3, time_end: float = 80e-6,
                   dt_frac: float = 0.3, lagrangian: int = 1, pdsearch_steps: int = 20,
                   material: Material | None = None, ql: int = 1, ql_coef: float = 0.001,
                   peridynamics: int = 0, mm_units: bool = False, itype_int: int = 0) -> Problem:
    """Taylor bar hitting a fixed plane: BAR gets v_z = -v0, WALL (impact face)
    gets fixity 0 0 1, dt = dt_frac * h / c_p.  peridynamics=1 selects the
    correspondence peridynamic force (Lagrangian only in the reference).
    Arguments are SI; mm_units=True writes the deck in the consistent
    mm / tonne / s / N / MPa system instead (lengths x1e3, stresses x1e-6,
    density x1e-12), which keeps the peridynamic shape tensor O(1): in SI at
    sub-mm spacing it is ~1e-16 and the reference's INVERSE treats it as
    singular (INVERSE.F90:41-45).  itype_int=1 selects the smoothed
    (stabilized non-conforming) gradients, Lagrangian only."""
    if mm_units:
        h, radius, length, v0 = h * 1e3, radius * 1e3, length * 1e3, v0 * 1e3
        material = (material or copper_linear()).in_mm()
    mesh = taylor_bar_mesh(h, radius=radius, length=length, seed=seed)
    mat = material or copper_linear()
    prob = assemble(mesh, {"BAR": (0, 0, 0)} | {"WALL": (0, 0, 1)}, {"BAR": (0.0, 0.0, -v0)}, mat,
                    norm_win=norm_win, dt=dt_frac * h / p_wave_speed(mat), time_end=time_end,
                    lagrangian=lagrangian, pdsearch=(1 if lagrangian else pdsearch_steps), ql=ql, ql_coef=ql_coef)
    prob.peridynamics = int(peridynamics)
    prob.itype_int = int(itype_int)
    return prob


def aluminium_powerlaw() -> Material:
    """The reference's shipped deck material (Control.dat *MATERIALS): J2 with
    power-law hardening sigma_0 (1 + B eps)^CE, E = 78.2 GPa, nu = 0.3,
    rho = 2700 kg/m^3, sigma_0 = 290 MPa, B = 125, CE = 0.1; VON_MISES then
    iterates Newton (CONSTITUTE_VONMISES.F90:94-119)."""
    return Material.j2(E=78.2e9, nu=0.3, rho=2700.0, sig0=0.29e9, B=125.0, CE=0.1)


def aluminium_elastic(mm_units: bool = False) -> Material:
    """Aluminium, elastic (type 1): E = 70 GPa, nu = 0.25, rho = 2700 kg/m^3;
    in mm / tonne / s / MPa with mm_units."""
    if mm_units:
        return Material.elastic(E=70e3, nu=0.25, rho=2.7e-9)
    return Material.elastic(E=70e9, nu=0.25, rho=2700.0)


def pd_plate_problem(h: float = 0.2e-3, lx: float = 100e-3, ly: float = 100e-3, lz: float = 5e-3, seed: int = 3,
                     v0: float = 300.0, impact_radius: float = 5e-3, horizon: float = 3.015, dt_frac: float = 0.3,
                     time_end: float = 20e-6, mm_units: bool = True, preprocess_fn=None) -> Problem:
    """Correspondence-peridynamic plate (CONSTRUCT_FINT_PD, *PERIDYNAMICS 1):
    a jittered cubic lattice [0,lx]x[0,ly]x[0,lz] (plate_mesh) of elastic
    aluminium, the four lateral faces clamped (set EDGE, fixity 1 1 1: the
    particles within half a spacing of them) and the particles within
    impact_radius of the plate's vertical centre axis (set IMPACT) given
    v_z = -v0 as *INITIAL CONDITION: the footprint of a striker, as the
    reference's cards can express it (the reference has no contact).
    Spherical support (*SUPPORT 1, SHSUP) of horizon ~ horizon * h: the
    normalised window is horizon / 1.086, 1.086 h being the mean nodal window
    of a lattice jittered by 5 %.  dt = dt_frac * h / c_p.  Arguments are SI;
    mm_units (default) writes the model in mm / tonne / s / MPa, where the
    shape tensor is O(1) (in SI at sub-mm spacing INVERSE takes it for
    singular, INVERSE.F90:41-45).  preprocess_fn(mesh) -> Geometry replaces
    the host preprocessing (e.g. taylor_amd.mega.geometry, on the device)."""
    s = 1e3 if mm_units else 1.0
    mesh = plate_mesh(h * s, lx * s, ly * s, lz * s, seed=seed)
    x, y = mesh.coo[:, 0], mesh.coo[:, 1]
    r = np.sqrt((x - 0.5 * lx * s) ** 2 + (y - 0.5 * ly * s) ** 2)
    ids = np.arange(1, mesh.coo.shape[0] + 1, dtype=np.int32)
    mesh.sets["IMPACT"] = ids[r <= impact_radius * s]
    half = 0.5 * h * s
    edge = (x < half) | (x > lx * s - half) | (y < half) | (y > ly * s - half)
    mesh.sets["EDGE"] = ids[edge]
    mat = aluminium_elastic(mm_units)
    prob = assemble(mesh, {"EDGE": (1, 1, 1)}, {"IMPACT": (0.0, 0.0, -v0 * s)}, mat, norm_win=horizon / 1.086,
                    dt=dt_frac * h * s / p_wave_speed(mat), time_end=time_end, lagrangian=1, pdsearch=1, ql=0,
                    shsup=1, geometry=None if preprocess_fn is None else preprocess_fn(mesh))
    prob.peridynamics = 1
    return prob


def assemble(mesh: Mesh, fixity: dict, velocity: dict, mat: Material, norm_win: float, dt: float,
             time_end: float, lagrangian: int = 1, pdsearch: int = 1, ql: int = 1, ql_coef: float = 0.001,
             shsup: int = 0, geometry: "Geometry | None" = None) -> Problem:
    geo = geometry if geometry is not None else preprocess(mesh)
    n = mesh.coo.shape[0]
    # QL length from the unscaled window (PRE_CONTROL.F90:803-828), then scale
    ql_len = float(geo.win0[0, 0])
    for w in (geo.win0[:, 0], geo.win0[:, 1], geo.win0[:, 2]):
        ql_len = min(ql_len, float(w.min()))
    ql_len = ql_len * 0.1
    win = geo.win0 * norm_win
    ebc = np.zeros((n, 3), dtype=np.int32)
    for name, fx in fixity.items():
        ids = mesh.sets[name] - 1
        for k in range(3):
            if fx[k]:
                ebc[ids, k] = fx[k]
    vinit = np.zeros((n, 3))
    for name, vv in velocity.items():
        vinit[mesh.sets[name] - 1] = np.array(vv, dtype=np.float64)
    vinit[ebc == 1] = 0.0
    ebc_node = (ebc == 1).any(axis=1).astype(np.int32)
    mass = np.repeat((geo.vol * mat.props[2])[:, None], 3, axis=1)
    return Problem(coo=mesh.coo, win=win, vol=geo.vol, mass=mass, ebc=ebc, ebc_node=ebc_node, vinit=vinit,
                   mat_id=np.zeros(n, dtype=np.int32), materials=[mat], dist_max=geo.dist_max,
                   sets=mesh.sets, dt=dt, time_end=time_end, ql=ql, ql_coef=ql_coef, ql_len=ql_len,
                   lagrangian=lagrangian, finite_strain=1, shsup=shsup, pdsearch=pdsearch,
                   norm_win=norm_win, elcon=mesh.elcon)


# ----------------------------------------------------------------------------
# deck text (reference formats)

def write_k(path: str, mesh: Mesh) -> None:
    with open(path, "w") as f:
        f.write("*KEYWORD\n$ synthetic deck\n*NODE\n$    nid               x               y               z      tc      rc\n")
        for i, (x, y, z) in enumerate(mesh.coo, start=1):
            f.write("%8d%16.8e%16.8e%16.8e%8d%8d\n" % (i, x, y, z, 0, 0))
        f.write("*ELEMENT_SOLID\n")
        for e, row in enumerate(mesh.elcon, start=1):
            f.write("%8d%8d" % (e, 1) + "".join("%8d" % v for v in row) + "\n")
        for sid, (name, ids) in enumerate(mesh.sets.items(), start=1):
            f.write("*SET_NODE_LIST\n%10d\n" % sid)
            for s in range(0, len(ids), 8):
                f.write("".join("%10d" % v for v in ids[s:s + 8]) + "\n")
        f.write("*END\n")


def read_k(path: str) -> Mesh:
    """LS-DYNA subset the reference reads (PRE_MODEL.F90:149-229)."""
    nodes, els, sets = {}, [], {}
    with open(path) as f:
        lines = f.read().splitlines()
    i = 0
    cur = None
    while i < len(lines):
        ln = lines[i].rstrip()
        if ln.startswith("*NODE"):
            i += 2 if i + 1 < len(lines) and lines[i + 1].startswith("$") else 1
            while i < len(lines) and not lines[i].startswith("*") and not lines[i].startswith("$"):
                s = lines[i]
                nodes[int(s[0:8])] = (float(s[8:24]), float(s[24:40]), float(s[40:56]))
                i += 1
            continue
        if ln.startswith("*ELEMENT_SOLID"):
            i += 1
            while i < len(lines) and not lines[i].startswith("*"):
                s = lines[i]
                if not s.startswith("$") and s.strip():
                    els.append([int(s[8 * k:8 * k + 8]) for k in range(2, 10)])
                i += 1
            continue
        if ln.startswith("*SET_NODE_LIST"):
            sid = int(lines[i + 1])
            cur = sets.setdefault(sid, [])
            i += 2
            while i < len(lines) and not lines[i].startswith("*"):
                s = lines[i]
                cur.extend(int(s[10 * k:10 * k + 10]) for k in range(len(s) // 10) if s[10 * k:10 * k + 10].strip()
                           and int(s[10 * k:10 * k + 10]) != 0)
                i += 1
            continue
        i += 1
    n = max(nodes)
    coo = np.array([nodes[k] for k in range(1, n + 1)])
    return Mesh(coo=coo, elcon=np.array(els, dtype=np.int32),
                sets={sid: np.array(v, dtype=np.int32) for sid, v in sets.items()})


def write_control(path: str, kfile: str, prob: Problem, set_names: list, time_output: float,
                  out_vars: str = "unformatted", dlt_fac: float | None = None) -> None:
    """Control.dat with the cards the reference reads (PRE_CONTROL.F90:132-588).
    dlt_fac: IAUTO_TS = 1 with that DLT_FAC in place of the fixed DT."""
    mat = prob.materials[0]
    p = mat.props
    rows = [p[0:5], p[5:10], p[10:15], p[15:20], p[20:25], np.zeros(5)]
    fx_set = [nm for nm in set_names if nm == "WALL"]
    txt = ["*MODEL", "# Name", os.path.basename(kfile), "",
           "*TIME PARAMETERS", "#TIME_END", "  %.17e" % prob.time_end, "#IAUTO_TS", "  %d" % (dlt_fac is not None), "#DT", "  %.17e" % (prob.dt if dlt_fac is None else dlt_fac), "",
           "*QL PARAMETERS", "# USE, COEFFICIENT", "   %d          %.17g" % (prob.ql, prob.ql_coef), "",
           "*RK PARAMETERS", "# RK_DEGREE, RK_CONT, RK_IMPL", "   1          3        0", "",
           "*PERIDYNAMICS", "# SWITCH", "  %d" % int(prob.peridynamics), "",
           "*HPC PARAMETERS", "#HPC_SCHEME", "1", "",
           "*GHOSTING PARAMETERS", "#GHOST_BUFFER", "0", "",
           "*INTEGRATION PARAMETERS", "#TYPE", str(int(prob.itype_int)), "",
           "*FORMULATION PARAMETERS", "# IFINITE_STRAIN, ILAGRANGIAN", "   %d   %d" % (prob.finite_strain, prob.lagrangian), ""]
    if not prob.lagrangian and prob.pdsearch > 1:
        txt += ["*PERIODIC SEARCH", "# TIME", "  %.17e" % (prob.pdsearch * prob.dt * (1 - 1e-9)), ""]
    if prob.shsup:
        txt += ["*SUPPORT", "# SWITCH", "  1", ""]
    for sid, nm in enumerate(set_names, start=1):
        txt += ["*SET NAME", "#SET NO", str(sid), "#NAME", nm, ""]
    txt += ["*BODY DEFINITION", "#SET", set_names[0], "#ID", "1", "#GRAV", "0.0 0.0 0.0", ""]
    for nm in fx_set:
        txt += ["*BOUNDARY CONDITION", "#SET", nm, "#FIXITY", "0 0 1", ""]
    v = prob.vinit[np.argmax(np.abs(prob.vinit[:, 2]))]
    txt += ["*WINDOW DEFINITION", "#SET", set_names[0], "#WINDOW", "%.17g" % prob.norm_win, "",
            "*INITIAL CONDITION", "#SET", set_names[0], "#VEL", "%.17g %.17g %.17g" % (0.0, 0.0, v[2]), "",
            "*MATERIALS", "#SET", set_names[0], "#TYPE", str(mat.mat_type), "#props"]
    for r in rows:
        txt.append(" ".join("%.17g" % x for x in r))
    txt += ["", "*OUTPUT PARAMETERS", "#TIME_OUTPUT", "  %.17e" % time_output, "",
            "*OUTPUT VARIABLES", out_vars, ""]
    with open(path, "w") as f:
        f.write("\n".join(txt) + "\n")
