"""CPU FP64 restatement of the Darcy correction network (test infrastructure).

This module is the checker for the network kernels of the HIP library
(sphere_amd/csrc/net_kernels.hip, net.hip): only tests/ and the CPU
baseline of bench.py may import it, and nothing in the product calls it.
The network has no counterpart in the reference; it is defined by
sphere_amd/csrc/net.hip:

    x0 = relu(c0(in))          3 -> 16, grid L0
    x1 = relu(c1(pool x0))    16 -> 32, L1 = L0 / 2
    x2 = relu(c2(pool x1))    32 -> 48, L2
    x3 = relu(c3(pool x2))    48 -> 64, L3
    a3 = relu(c4(x3))         64 -> 48;  y2 = up(a3) + x2
    a2 = relu(c5(y2))         48 -> 32;  y1 = up(a2) + x1
    a1 = relu(c6(y1))         32 -> 16;  y0 = up(a1) + x0
    h0 = relu(c7(y0))         16 -> 16
    out = h0 . w8 + b8        16 -> 1, pointwise

c0..c7 are 3x3x3 convolutions with padding 1 (zeros, or wrap-around along
the periodic axes), pool is 2x2x2 averaging,
up is nearest-neighbour upsampling.  Arrays are [nz, ny, nx, channels];
weights w[tap][cout][cin] with tap = (dz+1)*9 + (dy+1)*3 + (dx+1) (one
tap for the pointwise head), then b[cout], per layer, flattened in layer
order.

With bf16=True the forward pass rounds the convolution weights, the input
and every activation to bfloat16 (round to nearest even) where the kernels
store BF16 (the head's weights stay FP32, as in the kernels), and the
backward pass rounds each convolution's output gradient likewise, so that
the restatement follows the kernels' arithmetic up to FP32 accumulation
order.

What pins it (tests/test_oracle_net.py): central finite differences of the
FP64 loss against the backward pass for every layer, the adjoint identity
<conv(x), y> = <x, conv_T(y)>, the identity kernel, and pooling /
upsampling as adjoint pairs.
"""
import numpy

LAYERS = [(3, 16), (16, 32), (32, 48), (48, 64), (64, 48), (48, 32), (32, 16), (16, 16), (16, 1)]
TAPS_OF = [27] * 8 + [1]
TAPS = [(dz, dy, dx) for dz in (-1, 0, 1) for dy in (-1, 0, 1) for dx in (-1, 0, 1)]


def param_count():
    return sum(t * co * ci + co for (ci, co), t in zip(LAYERS, TAPS_OF))


def split_params(params):
    """[(w [taps, cout, cin], b [cout])] views of the flat parameter vector."""
    out = []
    off = 0
    for (ci, co), t in zip(LAYERS, TAPS_OF):
        w = params[off:off + t * co * ci].reshape(t, co, ci)
        off += t * co * ci
        b = params[off:off + co]
        off += co
        out.append((w, b))
    return out


def bf16_round(a):
    """Round float values to the nearest bfloat16 (ties to even), as FP64."""
    f = numpy.asarray(a, dtype=numpy.float32)
    u = f.view(numpy.uint32).astype(numpy.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) >> 16
    return (u.astype(numpy.uint32) << 16).view(numpy.float32).astype(numpy.float64)


def pad1(x, wrap=0):
    """One-cell padding of [nz, ny, nx, c]: zeros, or wrap-around along x
    (wrap bit 0) and y (bit 1)."""
    xp = numpy.pad(x, ((1, 1), (0, 0), (0, 0), (0, 0)))
    xp = numpy.pad(xp, ((0, 0), (1, 1), (0, 0), (0, 0)), mode="wrap" if wrap & 2 else "constant")
    return numpy.pad(xp, ((0, 0), (0, 0), (1, 1), (0, 0)), mode="wrap" if wrap & 1 else "constant")


def conv(x, w, b=None, wrap=0):
    """3x3x3 convolution, padding 1: out[v][co] = sum_tap,ci x[v+tap][ci] w[tap][co][ci]."""
    nz, ny, nx, _ = x.shape
    xp = pad1(x, wrap)
    out = numpy.zeros((nz, ny, nx, w.shape[1]))
    for t, (dz, dy, dx) in enumerate(TAPS):
        sl = xp[1 + dz:1 + dz + nz, 1 + dy:1 + dy + ny, 1 + dx:1 + dx + nx, :]
        out += sl @ w[t].T
    if b is not None:
        out += b
    return out


def conv_t(dy, w, wrap=0):
    """Adjoint of conv in x: dx[u][ci] = sum_tap,co dy[u - tap][co] w[tap][co][ci]."""
    nz, ny, nx, _ = dy.shape
    dp = pad1(dy, wrap)
    out = numpy.zeros((nz, ny, nx, w.shape[2]))
    for t, (dz, dyy, dx) in enumerate(TAPS):
        sl = dp[1 - dz:1 - dz + nz, 1 - dyy:1 - dyy + ny, 1 - dx:1 - dx + nx, :]
        out += sl @ w[t]
    return out


def conv_wgrad(x, dy, wrap=0):
    """dW[tap][co][ci] = sum_v dy[v][co] x[v+tap][ci]."""
    nz, ny, nx, ci = x.shape
    co = dy.shape[3]
    xp = pad1(x, wrap)
    g = numpy.zeros((27, co, ci))
    d2 = dy.reshape(-1, co)
    for t, (dz, dyy, dx) in enumerate(TAPS):
        sl = xp[1 + dz:1 + dz + nz, 1 + dyy:1 + dyy + ny, 1 + dx:1 + dx + nx, :]
        g[t] = d2.T @ sl.reshape(-1, ci)
    return g


_conv = conv


def pool(x):
    nz, ny, nx, c = x.shape
    return x.reshape(nz // 2, 2, ny // 2, 2, nx // 2, 2, c).mean(axis=(1, 3, 5))


def pool_t(d):
    return up(d) / 8.0


def up(x):
    return numpy.repeat(numpy.repeat(numpy.repeat(x, 2, axis=0), 2, axis=1), 2, axis=2)


def up_t(d):
    nz, ny, nx, c = d.shape
    return d.reshape(nz // 2, 2, ny // 2, 2, nx // 2, 2, c).sum(axis=(1, 3, 5))


def relu(x):
    return numpy.maximum(x, 0.0)


def forward(params, inp, bf16=False, wrap=0):
    """inp [nz, ny, nx, 3] -> (out [nz, ny, nx], cache for backward)."""
    r = bf16_round if bf16 else (lambda a: a)
    P = [(r(w) if w.shape[0] == 27 else w, b)
         for w, b in split_params(numpy.asarray(params, dtype=numpy.float64))]
    c = {"P": P, "wrap": wrap, "bf16": bf16}

    def conv(x, w, b):
        return _conv(x, w, b, wrap)

    c["in"] = r(inp)
    c["x0"] = r(relu(conv(c["in"], *P[0])))
    c["p0"] = r(pool(c["x0"]))
    c["x1"] = r(relu(conv(c["p0"], *P[1])))
    c["p1"] = r(pool(c["x1"]))
    c["x2"] = r(relu(conv(c["p1"], *P[2])))
    c["p2"] = r(pool(c["x2"]))
    c["x3"] = r(relu(conv(c["p2"], *P[3])))
    c["a3"] = r(relu(conv(c["x3"], *P[4])))
    c["y2"] = r(up(c["a3"]) + c["x2"])
    c["a2"] = r(relu(conv(c["y2"], *P[5])))
    c["y1"] = r(up(c["a2"]) + c["x1"])
    c["a1"] = r(relu(conv(c["y1"], *P[6])))
    c["y0"] = r(up(c["a1"]) + c["x0"])
    c["h0"] = r(relu(conv(c["y0"], *P[7])))
    out = c["h0"] @ P[8][0][0, 0] + P[8][1][0]
    return out, c


def backward(c, dout):
    """Gradient of sum(out * dout) with respect to the flat parameters."""
    P = c["P"]
    wrap = c["wrap"]
    # with bf16, the output gradient of each convolution is rounded as the
    # kernels store it (their BF16 operand)
    r = bf16_round if c["bf16"] else (lambda a: a)
    grads = [None] * 9

    def layer(k, x, dz):
        dz = r(dz)
        grads[k] = (conv_wgrad(x, dz, wrap), dz.sum(axis=(0, 1, 2)))
        return conv_t(dz, P[k][0], wrap)

    w8 = P[8][0][0, 0]
    grads[8] = ((numpy.tensordot(dout, c["h0"], axes=3))[None, None, :], numpy.array([dout.sum()]))
    dh0 = dout[..., None] * w8
    dy0 = layer(7, c["y0"], dh0 * (c["h0"] > 0))
    da1 = up_t(dy0)
    dy1 = layer(6, c["y1"], da1 * (c["a1"] > 0))
    da2 = up_t(dy1)
    dy2 = layer(5, c["y2"], da2 * (c["a2"] > 0))
    da3 = up_t(dy2)
    dx3 = layer(4, c["x3"], da3 * (c["a3"] > 0))
    dp2 = layer(3, c["p2"], dx3 * (c["x3"] > 0))
    dx2 = dy2 + pool_t(dp2)
    dp1 = layer(2, c["p1"], dx2 * (c["x2"] > 0))
    dx1 = dy1 + pool_t(dp1)
    dp0 = layer(1, c["p0"], dx1 * (c["x1"] > 0))
    dx0 = dy0 + pool_t(dp0)
    dz0 = r(dx0 * (c["x0"] > 0))
    grads[0] = (conv_wgrad(c["in"], dz0, wrap), dz0.sum(axis=(0, 1, 2)))
    flat = []
    for gw, gb in grads:
        flat.append(gw.ravel())
        flat.append(gb.ravel())
    return numpy.concatenate(flat)


def init_params(rng, scale=1.0):
    """He-style initial weights (std sqrt(2 / (taps cin))) and zero biases."""
    out = []
    for (ci, co), t in zip(LAYERS, TAPS_OF):
        std = scale * numpy.sqrt(2.0 / (t * ci))
        out.append(rng.standard_normal(t * co * ci) * std)
        out.append(numpy.zeros(co))
    return numpy.concatenate(out)
