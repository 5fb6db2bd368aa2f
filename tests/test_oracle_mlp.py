"""The oracle's transport-surrogate network (CPU, FP64): the backpropagated
gradient of the loss is pinned by central finite differences of the loss
itself, plus closed forms that a dropped term or a transposed weight would
break.  (The oracle is the checker of the GPU training kernels.)"""
import numpy as np
import pytest

from oracle import pyoracle

H = 64


def shape(nin, nout):
    o = {}
    o["w1"] = 0
    o["b1"] = H * nin
    o["w2"] = o["b1"] + H
    o["b2"] = o["w2"] + H * H
    o["w3"] = o["b2"] + H
    o["b3"] = o["w3"] + H * H
    o["w4"] = o["b3"] + H
    o["b4"] = o["w4"] + nout * H
    o["count"] = o["b4"] + nout
    return o


def random_params(nin, nout, seed):
    rng = np.random.default_rng(seed)
    o = shape(nin, nout)
    p = np.zeros(o["count"], dtype=np.float32)
    p[o["w1"]:o["b1"]] = rng.uniform(-1, 1, H * nin) * np.sqrt(3.0 / nin)
    p[o["b1"]:o["w2"]] = rng.uniform(-0.2, 0.2, H)
    for w, b in (("w2", "b2"), ("w3", "b3")):
        p[o[w]:o[b]] = rng.uniform(-1, 1, H * H) * np.sqrt(3.0 / H)
        p[o[b]:o[b] + H] = rng.uniform(-0.2, 0.2, H)
    p[o["w4"]:o["b4"]] = rng.uniform(-1, 1, nout * H) * np.sqrt(3.0 / H)
    p[o["b4"]:] = rng.uniform(-0.2, 0.2, nout)
    return p


def samples(n, nin, nout, seed):
    rng = np.random.default_rng(seed)
    x = np.zeros((n, 16), dtype=np.float32)
    t = np.zeros((n, 16), dtype=np.float32)
    x[:, :nin] = rng.standard_normal((n, nin))
    t[:, :nout] = rng.standard_normal((n, nout))
    return x, t


def test_gradient_matches_finite_differences():
    nin = nout = 11
    p = random_params(nin, nout, 1)
    x, t = samples(8, nin, nout, 2)
    loss, g = pyoracle.mlp_loss_grad(p, nin, nout, x, t, emulate=False)
    o = shape(nin, nout)
    rng = np.random.default_rng(3)
    # parameters from every weight matrix and bias vector
    picks = []
    for lo, hi in ((o["w1"], o["b1"]), (o["b1"], o["w2"]), (o["w2"], o["b2"]), (o["b2"], o["w3"]),
                   (o["w3"], o["b3"]), (o["b3"], o["w4"]), (o["w4"], o["b4"]), (o["b4"], o["count"])):
        picks += list(rng.integers(lo, hi, 6))
    for i in picks:
        h = 1e-2
        pp, pm = p.copy(), p.copy()
        pp[i] = np.float32(p[i] + h)
        pm[i] = np.float32(p[i] - h)
        lp = pyoracle.mlp_loss_grad(pp, nin, nout, x, t, emulate=False)[0]
        lm = pyoracle.mlp_loss_grad(pm, nin, nout, x, t, emulate=False)[0]
        fd = (lp - lm) / (float(pp[i]) - float(pm[i]))
        assert abs(fd - g[i]) <= 1e-6 * max(1.0, np.abs(g).max()), (i, fd, g[i])


def test_loss_and_output_closed_forms():
    nin = nout = 5
    o = shape(nin, nout)
    p = random_params(nin, nout, 4)
    p[o["w4"]:o["b4"]] = 0.0  # output = b4 whatever the inputs
    x, t = samples(16, nin, nout, 5)
    y = pyoracle.mlp_forward(p, nin, nout, x, emulate=False)
    assert np.allclose(y[:, :nout], p[o["b4"]:][None, :], rtol=0, atol=1e-7)
    loss, g = pyoracle.mlp_loss_grad(p, nin, nout, x, t, emulate=False)
    b4 = p[o["b4"]:].astype(np.float64)
    tt = t[:, :nout].astype(np.float64)
    assert np.isclose(loss, np.mean((b4[None, :] - tt) ** 2), rtol=1e-12)
    assert np.allclose(g[o["b4"]:], 2 * np.sum(b4[None, :] - tt, axis=0) / (16 * nout), rtol=1e-12, atol=0)
    # with W4 = 0 no error reaches the hidden layers
    assert np.abs(g[:o["w4"]]).max() == 0.0


def test_emulated_gradient_close_to_exact():
    """BF16 rounding of operands changes the gradient by O(2^-8), not more."""
    nin = nout = 11
    p = random_params(nin, nout, 6)
    x, t = samples(64, nin, nout, 7)
    _, g = pyoracle.mlp_loss_grad(p, nin, nout, x, t, emulate=False)
    _, ge = pyoracle.mlp_loss_grad(p, nin, nout, x, t, emulate=True)
    rel = np.abs(ge - g).max() / np.abs(g).max()
    assert 1e-5 < rel < 5e-2, rel


def test_folded_inference_is_the_same_network():
    """The GPU's inference form (the sigmoid's exponent factor -log2 e folded
    into the layer-1 weights and the layer 1-3 biases, -ln 2 into the output
    weights) is the same function: unrounded it matches the exact network up
    to the float scaling of the parameters, and with the GPU's BF16 rounding
    to O(2^-8).  A factor missing from any one bias or weight block moves the
    outputs by O(1)."""
    nin = nout = 11
    p = random_params(nin, nout, 8)
    x, _ = samples(256, nin, nout, 9)
    y = pyoracle.mlp_forward(p, nin, nout, x, emulate=False)[:, :nout]
    yf = pyoracle.mlp_forward(p, nin, nout, x, emulate=2)[:, :nout]
    ye = pyoracle.mlp_forward(p, nin, nout, x, emulate=True)[:, :nout]
    scale = np.abs(y).max()
    assert np.abs(yf - y).max() / scale < 1e-6
    rel = np.abs(ye - y).max() / scale
    assert 1e-5 < rel < 5e-2, rel
