"""CPU restatement of the learned J2 return map's network and its training
gradient -- TEST INFRASTRUCTURE ONLY: tests/ import it as the checker of the
GPU kernels in taylor_amd/csrc/mega_mlp.hip and mlp_mfma.hpp; nothing in the
product calls it.

The network has no counterpart in the reference (it stands in for VON_MISES,
CONSTITUTE_VONMISES.F90:21-158).  Layers: 13 -> 64 -> 64 -> 64 -> 7 with SiLU
after the three hidden layers; parameters are one flat FP32 vector, per layer
W[out][in] row-major then b[out].

With emulate=True every matrix-product operand (weights, activations, the
back-propagated errors) is rounded to bfloat16 (round to nearest even) before
the product, as the matrix cores take them, and products are accumulated in
float64; biases and the bias gradients stay unrounded.  With emulate=False the
same formulas run in plain float64 (the reference for finite differences and
for torch autograd in the tests).
"""
from __future__ import annotations

import numpy as np

KIN, KHID, KOUT = 13, 64, 7
SIZES = [(KHID, KIN), (KHID, KHID), (KHID, KHID), (KOUT, KHID)]
NPARAMS = sum(o * i + o for o, i in SIZES)   # 9671
IN_MEAN, IN_SCALE, OUT_SCALE, ENV_LO, ENV_HI, NNORM = 0, 13, 26, 33, 46, 59


def bf16(a) -> np.ndarray:
    """float -> bfloat16 (round to nearest even) -> float64."""
    f = np.ascontiguousarray(a, dtype=np.float32)
    u = f.view(np.uint32).astype(np.uint64)
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    return r.astype(np.uint32).view(np.float32).astype(np.float64)


def unpack(P):
    P = np.asarray(P, dtype=np.float64)
    out, k = [], 0
    for o, i in SIZES:
        W = P[k:k + o * i].reshape(o, i)
        k += o * i
        b = P[k:k + o]
        k += o
        out.append((W, b))
    return out


def pack(layers) -> np.ndarray:
    return np.concatenate([np.concatenate([W.ravel(), b]) for W, b in layers]).astype(np.float32)


def silu(z):
    return z / (1.0 + np.exp(-z))


def dsilu(z):
    s = 1.0 / (1.0 + np.exp(-z))
    return s * (1.0 + z * (1.0 - s))


def forward(P, x, emulate: bool = True):
    """y[n, 7] and the pre-activations/activations of each layer."""
    q = bf16 if emulate else (lambda a: np.asarray(a, dtype=np.float64))
    layers = unpack(P)
    a = np.asarray(x, dtype=np.float64)
    acts, zs = [a], []
    for li, (W, b) in enumerate(layers):
        z = q(a) @ q(W).T + b
        zs.append(z)
        if li < 3:
            a = silu(z)
            acts.append(a)
        else:
            a = z
    return a, zs, acts


def gradient(P, x, t, emulate: bool = True):
    """Gradient (flat, float64) and value of (1/n) sum ||y - t||^2."""
    q = bf16 if emulate else (lambda a: np.asarray(a, dtype=np.float64))
    layers = unpack(P)
    y, zs, acts = forward(P, x, emulate)
    n = y.shape[0]
    e = y - np.asarray(t, dtype=np.float64)
    loss = float((e * e).sum() / n)
    dz = 2.0 * e / n
    grads = [None] * 4
    for li in range(3, -1, -1):
        W, _ = layers[li]
        grads[li] = (q(dz).T @ q(acts[li]), dz.sum(axis=0))
        if li > 0:
            da = q(dz) @ q(W)
            dz = da * dsilu(zs[li - 1])
    g = np.concatenate([np.concatenate([gW.ravel(), gb]) for gW, gb in grads])
    return g, loss


def trial(tuples, props):
    """Elastic trial stress of harvested tuples (FINT.F90:789-794 with props
    28-30 = lambda, mu, lambda + 2 mu)."""
    u = np.asarray(tuples, dtype=np.float64)
    lam, mu, l2m = props[27], props[28], props[29]
    de, s = u[:, 0:6], u[:, 6:12]
    tr = s.copy()
    tr[:, 0] += l2m * de[:, 0] + lam * de[:, 1] + lam * de[:, 2]
    tr[:, 1] += lam * de[:, 0] + l2m * de[:, 1] + lam * de[:, 2]
    tr[:, 2] += lam * de[:, 0] + lam * de[:, 1] + l2m * de[:, 2]
    tr[:, 3:6] += mu * de[:, 3:6]
    return tr


def prepare(tuples, props):
    """Raw inputs x[n, 13], plastic-correction targets t[n, 7] and the
    normalisation block norm[59] of a harvested training set."""
    u = np.asarray(tuples, dtype=np.float32).astype(np.float64)
    x = u[:, 0:13]
    t = np.concatenate([trial(u, props) - u[:, 13:19], (u[:, 19] - u[:, 12])[:, None]], axis=1)
    norm = np.zeros(NNORM)
    mean = x.mean(axis=0)
    var = np.maximum((x * x).mean(axis=0) - mean * mean, 0.0)
    norm[IN_MEAN:IN_MEAN + 13] = mean
    norm[IN_SCALE:IN_SCALE + 13] = np.where(var > 0, 1.0 / np.sqrt(np.where(var > 0, var, 1.0)), 1.0)
    rms = np.sqrt((t * t).mean(axis=0))
    norm[OUT_SCALE:OUT_SCALE + 7] = np.where(rms > 0, rms, 1.0)
    norm[ENV_LO:ENV_LO + 13] = x.min(axis=0)
    norm[ENV_HI:ENV_HI + 13] = x.max(axis=0)
    return x, t, norm.astype(np.float32)


def normalise(x, t, norm):
    n = np.asarray(norm, dtype=np.float32)
    xh = ((np.asarray(x, np.float32) - n[IN_MEAN:IN_MEAN + 13]) * n[IN_SCALE:IN_SCALE + 13]).astype(np.float32)
    th = (np.asarray(t, np.float32) / n[OUT_SCALE:OUT_SCALE + 7]).astype(np.float32)
    return xh, th


def return_residual(pred, s, eps_n, eps, sig0, B, CE, mu):
    """Largest residual of the radial-return equations in units of
    sigma_y(eps): yield condition, flow direction, pressure, multiplier."""
    pred, s = np.asarray(pred, np.float64), np.asarray(s, np.float64)

    def dev(v):
        d = v.copy()
        d[:, :3] -= v[:, :3].sum(1, keepdims=True) / 3
        return d

    def tn(v):
        return np.sqrt((v[:, :3] ** 2).sum(1) + 2 * (v[:, 3:] ** 2).sum(1))

    a, b = dev(pred), dev(s)
    na, nb = tn(a), tn(b)
    sy = sig0 * (1.0 + B * eps) ** CE
    r_yield = np.abs(np.sqrt(1.5) * nb - sy)
    r_dir = tn(b - (nb / na)[:, None] * a)
    r_pres = np.abs(s[:, :3].sum(1) - pred[:, :3].sum(1)) / 3
    r_mult = np.abs(na - nb - 2 * mu * np.sqrt(1.5) * (eps - eps_n))
    return np.maximum.reduce([r_yield, r_dir, r_pres, r_mult]) / sy


def decode(pred, eps_n, y, norm, props, tol):
    """The step's use of the network output: stress = trial - correction,
    eqps = eqps_n + increment, accepted when the return-map residual is
    within tol and eqps does not decrease."""
    os_ = np.asarray(norm, np.float64)[OUT_SCALE:OUT_SCALE + 7]
    s = np.asarray(pred, np.float64) - os_[:6] * y[:, :6]
    e = np.asarray(eps_n, np.float64) + os_[6] * y[:, 6]
    res = return_residual(pred, s, eps_n, e, props[3], props[4], props[5], props[28])
    ok = (res <= tol) & (e >= eps_n)
    return s, e, ok
