"""Pins the FP64 network oracle (oracle/nn_oracle.py) on the CPU.

Each check is independent of the oracle's own formulas: central finite
differences, closed forms of special kernels, and adjoint identities.  A
planted mistake (a flipped kernel in the data gradient) is shown to fail.
"""
import numpy as np
import pytest

import nn_oracle as NN


def _rng(seed):
    return np.random.default_rng(seed)


def test_conv_identity_and_shift_kernels():
    x = _rng(1).standard_normal((2, 3, 6, 5))
    w = np.zeros((3, 3, 3, 3))
    for c in range(3):
        w[c, c, 1, 1] = 1.0
    assert np.allclose(NN.conv3_fwd(x, w, np.zeros(3)), x)
    # tap (ky, kx) = (1, 2) reads x[y, x + 1]: a left shift with zero fill
    w = np.zeros((1, 1, 3, 3))
    w[0, 0, 1, 2] = 1.0
    y = NN.conv3_fwd(x[:, :1], w, np.array([0.5]))
    ref = np.zeros_like(x[:, :1])
    ref[..., :-1] = x[:, :1, :, 1:]
    assert np.allclose(y, ref + 0.5)
    # all-ones kernel: 3x3 box sum with zero padding
    w = np.ones((1, 1, 3, 3))
    y = NN.conv3_fwd(x[:1, :1], w, np.zeros(1))[0, 0]
    xp = np.pad(x[0, 0], 1)
    box = sum(xp[dy:dy + 6, dx:dx + 5] for dy in range(3) for dx in range(3))
    assert np.allclose(y, box)


def test_pool_and_upsample_closed_forms():
    x = _rng(2).standard_normal((2, 3, 4, 6))
    assert np.allclose(NN.pool2(NN.up2(x)), x)
    assert np.allclose(NN.pool2(np.full((1, 1, 4, 4), 3.0)), 3.0)
    assert NN.up2(x).shape == (2, 3, 8, 12)


@pytest.mark.parametrize("op", ["conv", "pool", "up"])
def test_adjoint_dot_products(op):
    r = _rng(3)
    if op == "conv":
        x = r.standard_normal((2, 3, 5, 7))
        w = r.standard_normal((4, 3, 3, 3))
        y = r.standard_normal((2, 4, 5, 7))
        lhs = (NN.conv3_fwd(x, w, np.zeros(4)) * y).sum()
        dx, dw, db = NN.conv3_bwd(x, w, y)
        assert abs(lhs - (x * dx).sum()) <= 1e-10 * abs(lhs)
        # linear in w as well: <conv(x; w), y> = <w, dw>
        assert abs(lhs - (w * dw).sum()) <= 1e-10 * abs(lhs)
        assert np.allclose(db, y.sum(axis=(0, 2, 3)))
    elif op == "pool":
        x = r.standard_normal((2, 3, 6, 4))
        y = r.standard_normal((2, 3, 3, 2))
        assert abs((NN.pool2(x) * y).sum() - (x * NN.pool2_bwd(y)).sum()) <= 1e-12
    else:
        x = r.standard_normal((2, 3, 3, 2))
        y = r.standard_normal((2, 3, 6, 4))
        assert abs((NN.up2(x) * y).sum() - (x * NN.up2_bwd(y)).sum()) <= 1e-12


def _loss(x, flat, t, arch):
    y, cache = NN.unet_forward(x, flat, *arch)
    L, dy = NN.mse(y, t)
    return L, dy, cache


def test_unet_gradients_by_finite_differences():
    """every parameter tensor and the input, central differences in FP64."""
    arch = (2, 2, 3, 3)      # in, out, base, levels
    r = _rng(4)
    layers = NN.unet_layers(*arch)
    n = sum(co * ci * 9 + co for ci, co in layers)
    flat = r.standard_normal(n) * 0.3
    x = r.standard_normal((2, 2, 8, 8))
    t = r.standard_normal((2, 2, 8, 8))
    L, dy, cache = _loss(x, flat, t, arch)
    dx, g = NN.unet_backward(dy, flat, cache, *arch)
    h = 1e-6
    # a few coordinates of every layer's weights and bias
    o = 0
    idx = []
    for ci, co in layers:
        nw = co * ci * 9
        idx += list(r.choice(nw, size=3, replace=False) + o) + [o + nw + int(r.integers(co))]
        o += nw + co
    for i in idx:
        e = np.zeros(n)
        e[i] = h
        fd = (_loss(x, flat + e, t, arch)[0] - _loss(x, flat - e, t, arch)[0]) / (2 * h)
        assert abs(fd - g[i]) <= 1e-6 * max(1.0, abs(fd)) + 1e-8, (i, fd, g[i])
    for k in range(6):
        j = tuple(r.integers(s) for s in x.shape)
        e = np.zeros_like(x)
        e[j] = h
        fd = (_loss(x + e, flat, t, arch)[0] - _loss(x - e, flat, t, arch)[0]) / (2 * h)
        assert abs(fd - dx[j]) <= 1e-6 * max(1.0, abs(fd)) + 1e-8


def test_planted_flip_in_data_gradient_fails():
    """a conv data gradient that forgets to flip the kernel breaks the adjoint test."""
    r = _rng(5)
    x = r.standard_normal((1, 2, 5, 5))
    w = r.standard_normal((2, 2, 3, 3))
    y = r.standard_normal((1, 2, 5, 5))
    lhs = (NN.conv3_fwd(x, w, np.zeros(2)) * y).sum()
    wrong = NN.conv3_fwd(y, np.transpose(w, (1, 0, 2, 3)), np.zeros(2))   # no 180-degree flip
    assert abs(lhs - (x * wrong).sum()) > 1e-3 * abs(lhs)


def test_bf16_rounding_against_torch_and_ties():
    """the oracle's BF16 rounding: torch's bfloat16 conversion (round to nearest
    even) on random values of every magnitude, and the tie cases by hand"""
    torch = pytest.importorskip("torch")
    r = _rng(21)
    x = (r.standard_normal(200000) * np.exp(r.uniform(-60, 60, 200000))).astype(np.float32)
    # exact ties: a BF16 value plus half a BF16 ulp
    base = NN.bf16(r.standard_normal(1000)).astype(np.float32)
    ulp = np.abs(base) * 2.0 ** -7
    ties = (base + 0.5 * np.float32(2.0) ** np.floor(np.log2(ulp))).astype(np.float32)
    for v in (x, ties):
        want = torch.from_numpy(v).to(torch.bfloat16).to(torch.float64).numpy()
        assert np.array_equal(NN.bf16(v), want)
    # 1 + 2^-8 lies halfway between 1 and 1 + 2^-7: to the even mantissa, 1; 1 + 3 2^-8
    # lies halfway between 1 + 2^-7 and 1 + 2^-6: to 1 + 2^-6
    assert NN.bf16(np.float32(1 + 2.0 ** -8)) == 1.0
    assert NN.bf16(np.float32(1 + 3 * 2.0 ** -8)) == 1 + 2.0 ** -6
    assert NN.bf16(np.float32(1 + 2.0 ** -8 + 2.0 ** -20)) == 1 + 2.0 ** -7
    assert NN.bf16(np.float32(-1.5)) == -1.5


def test_bf16_convolution_is_the_convolution_of_rounded_operands():
    """bf16_ops rounds exactly the operands the kernels round (input and weights
    forward; output gradient with weights or input backward; bias never), and a
    plant that leaves the weights unrounded is seen"""
    r = _rng(22)
    x, w, b = r.standard_normal((2, 3, 5, 6)), r.standard_normal((4, 3, 3, 3)), r.standard_normal(4)
    dy = r.standard_normal((2, 4, 5, 6))
    assert np.array_equal(NN.conv3_fwd(x, w, b, True), NN.conv3_fwd(NN.bf16(x), NN.bf16(w), b))
    for got, want in zip(NN.conv3_bwd(x, w, dy, True), NN.conv3_bwd(NN.bf16(x), NN.bf16(w), NN.bf16(dy))):
        assert np.array_equal(got, want)
    assert not np.allclose(NN.conv3_fwd(x, w, b, True), NN.conv3_fwd(NN.bf16(x), w, b), rtol=1e-6, atol=0)
    # and it stays within BF16's rounding of the exact convolution
    e = np.abs(NN.conv3_fwd(x, w, b, True) - NN.conv3_fwd(x, w, b)).max()
    assert 0 < e <= 2 * 2.0 ** -8 * (np.abs(x).max() * np.abs(w).sum(axis=(1, 2, 3)).max())
