"""The CPU restatement of the surrogate network (the gradient-check oracle
for the GPU kernels): finite differences pin its backward pass, closed forms
pin the loss."""
import ctypes as C

import numpy as np
import pytest

from conftest import oracle_engine


def random_params(n_params, scale=0.08, seed=0):
    rng = np.random.default_rng(seed)
    return (rng.standard_normal(n_params) * scale).astype(np.float32)


def layer_slices():
    """(offset, rows, row length, used inputs) of the five layers."""
    ln, lk = [256, 256, 256, 256, 192], [6, 256, 256, 256, 256]
    out, o = [], 0
    for n, k in zip(ln, lk):
        kp = 8 if k == 6 else k + 8
        out.append((o, n, kp, k))
        o += n * kp
    return out, o


def masked_params(w):
    """Zero the padding (columns after the bias, head rows >= 161)."""
    w = w.copy()
    for l, (o, n, kp, k) in enumerate(layer_slices()[0]):
        blk = w[o:o + n * kp].reshape(n, kp)
        blk[:, k + 1:] = 0
        if l == 4:
            blk[161:] = 0
    return w


def sample_rows(n, seed=1):
    rng = np.random.default_rng(seed)
    x = np.stack([rng.uniform(0.2, 1.0, n), rng.uniform(0.5, 1.0, n), rng.uniform(-0.6, 0.6, n),
                  rng.uniform(-0.6, 0.6, n), rng.uniform(-1, 1, n), rng.uniform(-1, 1, n)], 1).astype(np.float32)
    # deposited fractions: a third are full absorptions (1.0)
    fe = np.where(rng.uniform(size=n) < 1.0 / 3.0, 1.0, rng.uniform(0, 0.99, n))
    y = np.stack([(rng.uniform(size=n) < 0.7).astype(float), fe, rng.uniform(0, 1, n),
                  rng.normal(0, 0.3, n), rng.normal(0, 0.3, n)], 1).astype(np.float32)
    return x, y


def grad_d(e, x, y):
    n = e.surrogate_n_params()
    g = np.zeros(n, np.float64)
    loss = C.c_double()
    f = e.lib.gpet_cpu_surrogate_grad_d
    f.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_int64, C.POINTER(C.c_double), C.c_void_p]
    st = f(e.ctx, x.ctypes.data_as(C.c_void_p), y.ctypes.data_as(C.c_void_p), len(x), C.byref(loss),
           g.ctypes.data_as(C.c_void_p))
    assert st == 0
    return loss.value, g


@pytest.fixture()
def eng():
    from gpet_amd import engine
    e = oracle_engine(engine.make_options())
    f = e.lib.gpet_cpu_surrogate_set_exact
    f.argtypes = [C.c_void_p, C.c_int32]
    f(e.ctx, 1)
    yield e
    e.close()


def test_layout():
    _, n = layer_slices()
    assert n == 256 * 8 + 3 * 256 * 264 + 192 * 264 == 255488


def test_zero_network_closed_form(eng):
    """h = 0: BCE = log 2 per row; the mixture is 16 equal unit Gaussians at
    0 with full-absorption probability 1/2 each, so the event NLL is
    log 2 + sum_d (y_d^2 / 2 + log sqrt(2 pi)) over the four targets, or over
    the three positional ones for a full absorption (fE >= 0.999)."""
    eng.surrogate_set_params(np.zeros(layer_slices()[1], np.float32))
    x, y = sample_rows(64)
    h = eng.surrogate_forward(x)
    assert np.all(h == 0)
    loss, g = grad_d(eng, x, y)
    ev = y[:, 0] > 0.5
    terms = 0.5 * y[:, 1:].astype(float) ** 2 + 0.5 * np.log(2 * np.pi)
    pp = y[:, 1] >= np.float32(0.999)
    assert pp[ev].any() and (~pp[ev]).any()
    nll = np.log(2.0) + terms[:, 1:].sum(1) + np.where(pp, 0.0, terms[:, 0])
    expect = np.log(2.0) + np.where(ev, nll, 0.0).mean()
    assert loss == pytest.approx(expect, rel=1e-12)
    # dL/d(head bias of the event logit) = mean(sigmoid(0) - e)
    o, n, kp, k = layer_slices()[0][4]
    assert g[o + 0 * kp + k] == pytest.approx(np.mean(0.5 - y[:, 0]), rel=1e-12)


def test_gradient_finite_differences(eng):
    """Central differences on parameters of every layer (weights and biases,
    mixture weights, means and log sigmas)."""
    w = masked_params(random_params(layer_slices()[1], seed=3))
    x, y = sample_rows(24, seed=4)
    eng.surrogate_set_params(w)
    _, g = grad_d(eng, x, y)
    rng = np.random.default_rng(5)
    picks = []
    for l, (o, n, kp, k) in enumerate(layer_slices()[0]):
        rows = 161 if l == 4 else n
        for _ in range(8):
            picks.append(o + int(rng.integers(rows)) * kp + int(rng.integers(k + 1)))
    o4, _, kp4, _ = layer_slices()[0][4]
    for r in (0, 3, 20, 90, 150):      # event logit, a weight logit, a mean, a log sigma, a photopeak logit
        picks.append(o4 + r * kp4 + 256)
    for i in picks:
        eps = np.float32(1e-3)
        wp, wm = w.copy(), w.copy()
        wp[i] += eps
        wm[i] -= eps
        eng.surrogate_set_params(wp)
        lp, _ = grad_d(eng, x, y)
        eng.surrogate_set_params(wm)
        lm, _ = grad_d(eng, x, y)
        fd = (lp - lm) / (float(wp[i]) - float(wm[i]))
        assert abs(fd - g[i]) <= 2e-3 * max(abs(g[i]), 1e-2), (i, fd, g[i])
    # padding gets no gradient
    o, n, kp, k = layer_slices()[0][1]
    assert np.all(g[o:o + n * kp].reshape(n, kp)[:, k + 1:] == 0)
    assert np.all(g[o4:].reshape(192, kp4)[161:] == 0)


def test_bf16_mode_close_to_exact(eng):
    w = masked_params(random_params(layer_slices()[1], seed=6))
    x, _ = sample_rows(32, seed=7)
    eng.surrogate_set_params(w)
    h_exact = eng.surrogate_forward(x)
    eng.lib.gpet_cpu_surrogate_set_exact(eng.ctx, 0)
    h_bf = eng.surrogate_forward(x)
    scale = np.abs(h_exact).max()
    assert np.abs(h_bf - h_exact).max() < 0.05 * scale
    assert np.abs(h_bf - h_exact).max() > 0          # rounding is on
