"""Pins of the CPU restatement of the Darcy correction network
(oracle/darcy_net.py): finite differences of the FP64 loss, the adjoint
identities of convolution, pooling and upsampling, and the identity kernel."""
import os
import sys

import numpy
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, os.path.join(ROOT, "oracle"))
import darcy_net as N  # noqa: E402


def test_param_count():
    assert N.param_count() == 284977
    assert N.init_params(numpy.random.default_rng(0)).size == N.param_count()


def test_identity_kernel_and_shift():
    rng = numpy.random.default_rng(1)
    x = rng.standard_normal((4, 5, 6, 3))
    w = numpy.zeros((27, 3, 3))
    w[13] = numpy.eye(3)  # tap (0, 0, 0)
    assert numpy.array_equal(N.conv(x, w), x)
    # tap (0, 0, +1) reads x[.., x+1, :], zero past the edge
    w2 = numpy.zeros((27, 3, 3))
    w2[14] = numpy.eye(3)
    y = N.conv(x, w2)
    assert numpy.array_equal(y[:, :, :-1], x[:, :, 1:])
    assert numpy.all(y[:, :, -1] == 0.0)


@pytest.mark.parametrize("wrap", [0, 1, 2, 3])
def test_conv_adjoint_identities(wrap):
    rng = numpy.random.default_rng(2)
    x = rng.standard_normal((4, 6, 8, 5))
    w = rng.standard_normal((27, 7, 5))
    y = rng.standard_normal((4, 6, 8, 7))
    lhs = numpy.sum(N.conv(x, w, wrap=wrap) * y)
    assert numpy.isclose(lhs, numpy.sum(x * N.conv_t(y, w, wrap=wrap)), rtol=1e-12)
    # weight gradient: d<conv(x, w), y>/dw = conv_wgrad(x, y)
    g = N.conv_wgrad(x, y, wrap=wrap)
    assert numpy.isclose(lhs, numpy.sum(g * w), rtol=1e-12)
    a = rng.standard_normal((4, 6, 8, 2))
    b = rng.standard_normal((2, 3, 4, 2))
    assert numpy.isclose(numpy.sum(N.pool(a) * b), numpy.sum(a * N.pool_t(b)), rtol=1e-12)
    assert numpy.isclose(numpy.sum(N.up(b) * a), numpy.sum(b * N.up_t(a)), rtol=1e-12)


def test_wrap_padding_is_periodic():
    rng = numpy.random.default_rng(4)
    x = rng.standard_normal((4, 6, 8, 2))
    w = rng.standard_normal((27, 3, 2))
    y = N.conv(x, w, wrap=3)
    # shifting the input by a period along x and y shifts the output alike
    xs = numpy.roll(numpy.roll(x, 3, axis=2), 2, axis=1)
    ys = numpy.roll(numpy.roll(y, 3, axis=2), 2, axis=1)
    assert numpy.allclose(N.conv(xs, w, wrap=3), ys, rtol=1e-12, atol=1e-12)
    # but not along z (zero padding)
    assert not numpy.allclose(N.conv(numpy.roll(x, 1, axis=0), w, wrap=3), numpy.roll(y, 1, axis=0))


@pytest.mark.parametrize("wrap", [0, 3])
def test_backward_matches_finite_differences(wrap):
    rng = numpy.random.default_rng(3)
    shape = (8, 8, 8)
    params = N.init_params(rng)
    # positive biases keep every ReLU away from its kink under the probes
    for w, b in N.split_params(params):
        b[:] = 0.05
    inp = rng.standard_normal(shape + (3,))
    R = rng.standard_normal(shape)
    out, cache = N.forward(params, inp, wrap=wrap)
    g = N.backward(cache, R)
    assert g.shape == params.shape

    def loss(p):
        return numpy.sum(N.forward(p, inp, wrap=wrap)[0] * R)

    # a few coordinates of every layer's weights and bias
    offs = []
    off = 0
    for (ci, co), t in zip(N.LAYERS, N.TAPS_OF):
        nw = t * co * ci
        offs += list(off + rng.choice(nw, 3, replace=False))
        offs.append(off + nw + rng.integers(co))
        off += nw + co
    h = 1e-6
    for k in offs:
        e = numpy.zeros_like(params)
        e[k] = h
        fd = (loss(params + e) - loss(params - e)) / (2 * h)
        assert fd == pytest.approx(g[k], rel=2e-4, abs=1e-7 * max(1.0, abs(fd))), k


def test_bf16_rounding():
    v = numpy.array([1.0, 1.0 + 2 ** -8, 1.0 + 3 * 2 ** -8, -2.5, 0.0, 3.0e38])
    r = N.bf16_round(v)
    assert r[0] == 1.0
    assert r[1] == 1.0            # tie to even
    assert r[2] == 1.0 + 2 ** -6  # tie to even (upwards)
    assert r[3] == -2.5 and r[4] == 0.0
    assert numpy.all(numpy.abs(r - v) <= numpy.abs(v) * 2 ** -8)
