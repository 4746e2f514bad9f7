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
    """Zero the padding (columns after the bias, head rows >= 190)."""
    w = w.copy()
    for l, (o, n, kp, k) in enumerate(layer_slices()[0]):
        blk = w[o:o + n * kp].reshape(n, kp)
        blk[:, k + 1:] = 0
        if l == 4:
            blk[190:] = 0
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
    """h = 0: BCE = log 2 per row; for an event the energy class is uniform
    over the 61 classes (log 61) and the position mixture is 16 equal unit
    Gaussians at 0, so its NLL is sum_d (y_d^2 / 2 + log sqrt(2 pi)) over
    depth, dy and dz, whatever the class."""
    eng.surrogate_set_params(np.zeros(layer_slices()[1], np.float32))
    x, y = sample_rows(64)
    h = eng.surrogate_forward(x)
    assert h.shape[1] == 190 and np.all(h == 0)
    loss, g = grad_d(eng, x, y)
    ev = y[:, 0] > 0.5
    nll = np.log(61.0) + (0.5 * y[:, 2:].astype(float) ** 2 + 0.5 * np.log(2 * np.pi)).sum(1)
    expect = np.log(2.0) + np.where(ev, nll, 0.0).mean()
    assert loss == pytest.approx(expect, rel=1e-12)
    # dL/d(head bias of the event logit) = mean(sigmoid(0) - e)
    o, n, kp, k = layer_slices()[0][4]
    assert g[o + 0 * kp + k] == pytest.approx(np.mean(0.5 - y[:, 0]), rel=1e-12)


def energy_classes(x0, fe):
    """The energy class of each row from explicit bin edges (float64): full
    absorption at fE >= 0.999; else 48 equal bins of [0, edge) and 12 of
    [edge, 0.999), edge = 2a / (1 + 2a) for an entry of a = E / 511 keV (the
    most one Compton scatter deposits).  Returns the classes and the rows'
    distance to the nearest edge relative to the bin width."""
    a = x0.astype(np.float64)
    edge = 2 * a / (1 + 2 * a)
    cls = np.zeros(len(a), int)
    near = np.full(len(a), np.inf)
    for i, (e, f) in enumerate(zip(edge, fe.astype(np.float64))):
        if f >= 0.999:
            continue
        lo = e * np.arange(49) / 48
        hi = e + (0.999 - e) * np.arange(13) / 12
        edges = np.concatenate([lo, hi[1:]])
        b = int(np.searchsorted(edges, f, side="right")) - 1
        cls[i] = 1 + b
        w = edges[b + 1] - edges[b]
        near[i] = min(f - edges[b], edges[b + 1] - f) / w
    return cls, near


def head_bias_params(bias):
    """Every weight zero, the head's biases `bias` (190): h = bias for every
    input (the hidden activations are SiLU(0) = 0)."""
    w = np.zeros(layer_slices()[1], np.float32)
    o, n, kp, k = layer_slices()[0][4]
    w[o + np.arange(190) * kp + k] = bias
    return w


def test_energy_class_cross_entropy(eng):
    """Class logits c / 7 (others zero): an event's class term is
    lse(logits) - logit[class], and dL/d(class-logit bias c) is the mean over
    rows of event * (softmax_c - [class == c]), so the gradient recovers the
    number of events in each class: both must follow the bin edges above
    (a shifted bin, bins of fE instead of fE / edge, or an off-by-one
    fails)."""
    bias = np.zeros(190, np.float32)
    bias[1:62] = np.arange(61) / 7.0
    eng.surrogate_set_params(head_bias_params(bias))
    x, y = sample_rows(400, seed=11)
    cls, near = energy_classes(x[:, 0], y[:, 1])
    assert near.min() > 1e-4
    ev = y[:, 0] > 0.5
    assert len(set(cls[ev])) > 40 and (cls[ev] == 0).any() and (cls[ev] > 48).any()
    loss, g = grad_d(eng, x, y)
    lg = bias[1:62].astype(float)
    lse = np.log(np.exp(lg).sum())
    pos = (0.5 * y[:, 2:].astype(float) ** 2 + 0.5 * np.log(2 * np.pi)).sum(1)
    nll = lse - lg[cls] + pos
    expect = np.log(2.0) + np.where(ev, nll, 0.0).mean()
    assert loss == pytest.approx(expect, rel=1e-12)
    o, n, kp, k = layer_slices()[0][4]
    gc = g[o + np.arange(1, 62) * kp + k]
    p = np.exp(lg - lse)
    counts = np.bincount(cls[ev], minlength=61)
    assert np.allclose(gc, (ev.sum() * p - counts) / len(x), rtol=1e-10, atol=1e-13)


def test_position_weights_follow_energy_type(eng):
    """Component 0 centred at depth 0.5 and the others at 0 (unit sigmas);
    the full-absorption weights put almost everything on component 0 and
    the continuum weights on the others, so the position NLL of an event is
    that of N(0.5, 1) or of N(0, 1) in depth by its energy type."""
    bias = np.zeros(190, np.float32)
    bias[62] = 30.0           # full absorption: component 0
    bias[78] = -30.0          # continuum: components 1..15
    bias[94] = 0.5            # mean depth of component 0
    eng.surrogate_set_params(head_bias_params(bias))
    x, y = sample_rows(200, seed=12)
    cls, _ = energy_classes(x[:, 0], y[:, 1])
    ev = y[:, 0] > 0.5
    loss, _ = grad_d(eng, x, y)
    yd = y[:, 2:].astype(float)
    mu = np.where(cls == 0, 0.5, 0.0)
    pos = 0.5 * (yd[:, 0] - mu) ** 2 + 0.5 * (yd[:, 1:] ** 2).sum(1) + 1.5 * np.log(2 * np.pi)
    nll = np.log(61.0) + pos
    expect = np.log(2.0) + np.where(ev, nll, 0.0).mean()
    assert loss == pytest.approx(expect, rel=1e-9)


def test_gradient_finite_differences(eng):
    """Central differences on parameters of every layer (weights and biases,
    class logits, both sets of mixture weights, means and log sigmas)."""
    w = masked_params(random_params(layer_slices()[1], seed=3))
    x, y = sample_rows(24, seed=4)
    eng.surrogate_set_params(w)
    _, g = grad_d(eng, x, y)
    rng = np.random.default_rng(5)
    picks = []
    for l, (o, n, kp, k) in enumerate(layer_slices()[0]):
        rows = 190 if l == 4 else n
        for _ in range(8):
            picks.append(o + int(rng.integers(rows)) * kp + int(rng.integers(k + 1)))
    o4, _, kp4, _ = layer_slices()[0][4]
    # event logit, class logits (full absorption, below and above the edge),
    # a full-absorption and a continuum weight logit, a mean, a log sigma
    for r in (0, 1, 7, 55, 64, 85, 100, 160):
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
    assert np.all(g[o4:].reshape(192, kp4)[190:] == 0)


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
