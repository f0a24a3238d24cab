"""FP64 numpy restatement of the learned-start network (test infrastructure only).

The reference has no learned part; this module is the CPU oracle for the
network the library trains with its own HIP kernels: 3x3 convolutions (zero
padding 1, cross-correlation), ReLU, 2x2 average pooling, nearest 2x
upsampling, channel concatenation, and the U-Net built from them, forward and
backward.  The flat parameter layout matches the library's: for each conv
layer in order, weights [Co][Ci][3][3] then bias [Co].

With bf16_ops the convolutions take their operands as the library's kernels do:
weights, inputs and output gradients rounded from their FP32 values to BF16
(round to nearest even), products summed exactly here (FP64), the bias FP32.

Pinned by tests/test_nn_oracle.py: central finite differences of the loss with
respect to inputs and every parameter tensor, closed forms (identity and shift
kernels, constant pooling, pool(up(x)) = x), the adjoint dot-product test of
every linear operator, and the BF16 rounding against torch's bfloat16
conversion and its tie cases.
"""
from __future__ import annotations

import numpy as np


def bf16(x):
    """x rounded to FP32, then to BF16 by round to nearest even (as float64)."""
    u = np.ascontiguousarray(np.asarray(x, np.float32)).view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


# ---------------------------------------------------------------------------
# layers  (x: [N][C][H][W])
# ---------------------------------------------------------------------------
def conv3_fwd(x, w, b, bf16_ops=False):
    if bf16_ops:
        x, w = bf16(x), bf16(w)
    N, Ci, H, W = x.shape
    Co = w.shape[0]
    xp = np.zeros((N, Ci, H + 2, W + 2))
    xp[:, :, 1:-1, 1:-1] = x
    y = np.broadcast_to(b[None, :, None, None], (N, Co, H, W)).astype(np.float64).copy()
    for ky in range(3):
        for kx in range(3):
            y += np.einsum("oi,nihw->nohw", w[:, :, ky, kx], xp[:, :, ky:ky + H, kx:kx + W])
    return y


def conv3_bwd(x, w, dy, bf16_ops=False):
    """returns dx, dw, db for y = conv3_fwd(x, w, b)."""
    if bf16_ops:
        x, w, dy = bf16(x), bf16(w), bf16(dy)
    N, Ci, H, W = x.shape
    xp = np.zeros((N, Ci, H + 2, W + 2))
    xp[:, :, 1:-1, 1:-1] = x
    dxp = np.zeros_like(xp)
    dw = np.zeros_like(w, dtype=np.float64)
    for ky in range(3):
        for kx in range(3):
            dw[:, :, ky, kx] = np.einsum("nohw,nihw->oi", dy, xp[:, :, ky:ky + H, kx:kx + W])
            dxp[:, :, ky:ky + H, kx:kx + W] += np.einsum("oi,nohw->nihw", w[:, :, ky, kx], dy)
    return dxp[:, :, 1:-1, 1:-1].copy(), dw, dy.sum(axis=(0, 2, 3))


def relu(x):
    return np.maximum(x, 0.0)


def pool2(x):
    N, C, H, W = x.shape
    return x.reshape(N, C, H // 2, 2, W // 2, 2).mean(axis=(3, 5))


def pool2_bwd(dy):
    return np.repeat(np.repeat(dy, 2, axis=2), 2, axis=3) * 0.25


def up2(x):
    return np.repeat(np.repeat(x, 2, axis=2), 2, axis=3)


def up2_bwd(dy):
    N, C, H, W = dy.shape
    return dy.reshape(N, C, H // 2, 2, W // 2, 2).sum(axis=(3, 5))


# ---------------------------------------------------------------------------
# U-Net
# ---------------------------------------------------------------------------
def unet_layers(in_ch, out_ch, base, levels):
    """(Ci, Co) of every conv layer in parameter order: encoder a/b per level,
    decoder a/b per level from the deepest skip up, then the head."""
    ch = [base << l for l in range(levels)]
    layers = []
    cin = in_ch
    for l in range(levels):
        layers += [(cin, ch[l]), (ch[l], ch[l])]
        cin = ch[l]
    for l in range(levels - 2, -1, -1):
        layers += [(ch[l + 1] + ch[l], ch[l]), (ch[l], ch[l])]
    layers.append((ch[0], out_ch))
    return layers


def unpack_params(flat, layers):
    out, o = [], 0
    for ci, co in layers:
        w = flat[o:o + co * ci * 9].reshape(co, ci, 3, 3)
        o += co * ci * 9
        b = flat[o:o + co]
        o += co
        out.append((w, b))
    assert o == flat.size
    return out


def unet_forward(x, flat, in_ch, out_ch, base, levels, bf16_ops=False):
    """returns (y, cache)."""
    layers = unet_layers(in_ch, out_ch, base, levels)
    P = unpack_params(np.asarray(flat, np.float64), layers)
    cache = {"x": [], "h": [], "e": [], "cat": [], "hd": [], "d": {}}
    k = 0
    inp = x
    for l in range(levels):
        if l > 0:
            inp = pool2(cache["e"][l - 1])
        cache["x"].append(inp)
        h = relu(conv3_fwd(inp, *P[k], bf16_ops)); k += 1
        e = relu(conv3_fwd(h, *P[k], bf16_ops)); k += 1
        cache["h"].append(h)
        cache["e"].append(e)
    d = cache["e"][levels - 1]
    cache["d"][levels - 1] = d
    for l in range(levels - 2, -1, -1):
        cat = np.concatenate([up2(d), cache["e"][l]], axis=1)
        hd = relu(conv3_fwd(cat, *P[k], bf16_ops)); k += 1
        d = relu(conv3_fwd(hd, *P[k], bf16_ops)); k += 1
        cache["cat"].append(cat)
        cache["hd"].append(hd)
        cache["d"][l] = d
    y = conv3_fwd(d, *P[k], bf16_ops)
    return y, cache


def unet_backward(dy, flat, cache, in_ch, out_ch, base, levels, bf16_ops=False):
    """returns (dx, dflat)."""
    layers = unet_layers(in_ch, out_ch, base, levels)
    P = unpack_params(np.asarray(flat, np.float64), layers)
    grads = [None] * len(layers)
    ch = [base << l for l in range(levels)]
    k = len(layers) - 1
    dd, dw, db = conv3_bwd(cache["d"][0], P[k][0], dy, bf16_ops)
    grads[k] = (dw, db)
    de_skip = [None] * levels
    # decoder, shallowest first (reverse of the forward order)
    for j, l in enumerate(range(0, levels - 1)):
        idx = len(cache["cat"]) - 1 - j           # cache index of level l
        ka = 2 * levels + 2 * idx
        kb = ka + 1
        dd = dd * (cache["d"][l] > 0)
        dhd, dw, db = conv3_bwd(cache["hd"][idx], P[kb][0], dd, bf16_ops)
        grads[kb] = (dw, db)
        dhd = dhd * (cache["hd"][idx] > 0)
        dcat, dw, db = conv3_bwd(cache["cat"][idx], P[ka][0], dhd, bf16_ops)
        grads[ka] = (dw, db)
        de_skip[l] = dcat[:, ch[l + 1]:]
        dd = up2_bwd(dcat[:, :ch[l + 1]])
    # encoder, deepest first
    de = dd
    dx = None
    for l in range(levels - 1, -1, -1):
        if l < levels - 1:
            de = de_skip[l] + pool2_bwd(dx)
        ka, kb = 2 * l, 2 * l + 1
        de = de * (cache["e"][l] > 0)
        dh, dw, db = conv3_bwd(cache["h"][l], P[kb][0], de, bf16_ops)
        grads[kb] = (dw, db)
        dh = dh * (cache["h"][l] > 0)
        dx, dw, db = conv3_bwd(cache["x"][l], P[ka][0], dh, bf16_ops)
        grads[ka] = (dw, db)
    dflat = np.concatenate([np.concatenate([g[0].ravel(), g[1].ravel()]) for g in grads])
    return dx, dflat


def mse(y, t):
    d = y - t
    return float((d * d).mean()), 2.0 * d / d.size
