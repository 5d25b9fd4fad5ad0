"""The CPU restatement of the learned return map's network (oracle/mlp_oracle.py)
pinned independently: bfloat16 rounding against torch's conversion, the
float64 gradient against torch autograd and finite differences, the training
targets against the closed-form linear-hardening radial return."""
import numpy as np
import torch

from mlp_util import j2_tuples, mlp_oracle as mo
from taylor_amd import deck


def test_bf16_rounding_matches_torch():
    rng = np.random.default_rng(0)
    x = np.concatenate([rng.normal(size=20000) * 10.0 ** rng.integers(-6, 6, size=20000),
                        [1.0, 1.0 + 2 ** -8, 1.0 + 3 * 2 ** -8, -1.0 - 2 ** -8, 0.0, 65504.0]]).astype(np.float32)
    ref = torch.from_numpy(x).to(torch.bfloat16).to(torch.float64).numpy()
    assert np.array_equal(mo.bf16(x), ref)
    # ties go to the even mantissa
    assert mo.bf16(np.float32(1.0 + 2 ** -8)) == 1.0
    assert mo.bf16(np.float32(1.0 + 3 * 2 ** -8)) == 1.0 + 2 ** -6


def _random_net(seed):
    rng = np.random.default_rng(seed)
    layers = []
    for o, i in mo.SIZES:
        a = np.sqrt(6.0 / (o + i))
        layers.append((rng.uniform(-a, a, size=(o, i)), rng.normal(size=o) * 0.1))
    return mo.pack(layers), rng


def test_gradient_matches_torch_autograd():
    P, rng = _random_net(1)
    x = rng.normal(size=(96, 13)).astype(np.float32)
    t = rng.normal(size=(96, 7)).astype(np.float32)
    g, loss = mo.gradient(P, x, t, emulate=False)
    Pt = torch.tensor(P.astype(np.float64), requires_grad=True)
    k, a = 0, torch.tensor(x.astype(np.float64))
    for li, (o, i) in enumerate(mo.SIZES):
        W = Pt[k:k + o * i].reshape(o, i)
        k += o * i
        b = Pt[k:k + o]
        k += o
        a = a @ W.T + b
        if li < 3:
            a = torch.nn.functional.silu(a)
    lt = ((a - torch.tensor(t.astype(np.float64))) ** 2).sum() / x.shape[0]
    lt.backward()
    assert abs(lt.item() - loss) < 1e-12 * abs(loss)
    assert np.allclose(g, Pt.grad.numpy(), rtol=1e-10, atol=1e-13)


def test_gradient_finite_differences():
    P, rng = _random_net(2)
    x = rng.normal(size=(40, 13)).astype(np.float32)
    t = rng.normal(size=(40, 7)).astype(np.float32)
    P = P.astype(np.float64)
    g, _ = mo.gradient(P, x, t, emulate=False)
    for k in rng.choice(mo.NPARAMS, size=40, replace=False):
        h = 1e-6
        Pp, Pm = P.copy(), P.copy()
        Pp[k] += h
        Pm[k] -= h
        fd = (mo.gradient(Pp, x, t, emulate=False)[1] - mo.gradient(Pm, x, t, emulate=False)[1]) / (2 * h)
        assert abs(fd - g[k]) <= 1e-6 * max(1.0, abs(g[k]))


def test_bf16_emulation_is_close_but_not_exact():
    P, rng = _random_net(3)
    x = rng.normal(size=(200, 13)).astype(np.float32)
    y1 = mo.forward(P, x, emulate=True)[0]
    y0 = mo.forward(P, x, emulate=False)[0]
    rel = np.abs(y1 - y0).max() / np.abs(y0).max()
    assert 1e-5 < rel < 5e-2


def test_targets_closed_form_linear_hardening(oracle):
    """t = (trial - sigma, eqps increment): with CE = 1 the radial return
    gives |t_dev| = 2 mu dgamma, eqps increment sqrt(2/3) dgamma with
    dgamma = f / (2 mu + 2/3 B sigma_0), and no pressure change."""
    mat = deck.copper_linear()
    p = mat.props
    rng = np.random.default_rng(4)
    tup = j2_tuples(oracle, p, 300, rng)
    x, t, norm = mo.prepare(tup, p)
    tr = mo.trial(tup, p)
    dev = tr.copy()
    dev[:, :3] -= tr[:, :3].sum(axis=1, keepdims=True) / 3
    tn = np.sqrt((dev[:, :3] ** 2).sum(1) + 2 * (dev[:, 3:] ** 2).sum(1))
    f = tn - np.sqrt(2 / 3) * p[3] * (1 + p[4] * tup[:, 12].astype(np.float64))
    dg = f / (2 * p[28] + 2 / 3 * p[4] * p[3])
    corr = np.sqrt((t[:, :3] ** 2).sum(1) + 2 * (t[:, 3:6] ** 2).sum(1))
    scale = p[3]
    assert np.all(np.abs(t[:, :3].sum(1)) < 1e-6 * scale * 3)          # deviatoric correction
    assert np.allclose(corr, 2 * p[28] * dg, rtol=1e-5, atol=1e-6 * scale)
    # eqps is stored in float32: its rounding bounds the increment's error
    assert np.allclose(t[:, 6], np.sqrt(2 / 3) * dg, rtol=1e-5, atol=2.5e-7)
    # normalisation block and envelope
    assert np.allclose(norm[mo.ENV_LO:mo.ENV_LO + 13], x.min(0)) and np.allclose(norm[mo.ENV_HI:mo.ENV_HI + 13], x.max(0))
    xh, th = mo.normalise(x, t, norm)
    assert np.allclose(xh.mean(0), 0, atol=1e-4) and np.allclose(xh.std(0), 1, atol=1e-3)
    assert np.allclose(np.sqrt((th ** 2).mean(0)), 1, atol=1e-4)


def test_decode_accepts_exact_answers_and_rejects_perturbed(oracle):
    mat = deck.copper_linear()
    p = mat.props
    rng = np.random.default_rng(6)
    tup = j2_tuples(oracle, p, 200, rng)
    x, t, norm = mo.prepare(tup, p)
    y = t / norm[mo.OUT_SCALE:mo.OUT_SCALE + 7]
    pred = mo.trial(tup, p)
    s, e, ok = mo.decode(pred, tup[:, 12].astype(np.float64), y, norm, p, 1e-3)
    assert ok.all()
    assert np.allclose(s, tup[:, 13:19], rtol=1e-5, atol=1e-6 * p[3])
    # a 1 % error in the correction's deviatoric size breaks the yield condition
    # wherever the correction exceeds 0.1 sigma_y
    y2 = y.copy()
    y2[:, :6] *= 1.01
    _, _, ok2 = mo.decode(pred, tup[:, 12].astype(np.float64), y2, norm, p, 1e-3)
    big = np.sqrt((t[:, :3] ** 2).sum(1) + 2 * (t[:, 3:6] ** 2).sum(1)) > 0.1 * p[3]
    assert big.any() and not ok2[big].any()
    # a rotated correction keeps the yield condition but breaks the flow
    # direction; a pressure change breaks the pressure equation
    y3 = y.copy()
    y3[:, 3:6] = -y3[:, 3:6]
    s3, e3, ok3 = mo.decode(pred, tup[:, 12].astype(np.float64), y3, norm, p, 1e-3)
    shear = np.abs(t[:, 3:6]).max(1) > 0.05 * p[3]
    assert shear.any() and not ok3[shear].any()
    y4 = y.copy()
    y4[:, :3] += 0.01 * p[3] / norm[mo.OUT_SCALE:mo.OUT_SCALE + 3]
    assert not mo.decode(pred, tup[:, 12].astype(np.float64), y4, norm, p, 1e-3)[2].any()
