"""CPU oracle of the learned warm start's network (test infrastructure: only tests/, the smoke
entry point and the bench's CPU leg use it, as the checker; the product never calls it).

The reference has no learned part (its solver starts from zero, sparse_compressive_holo.cu:258-259);
this restates, in float64 numpy, the network the library's warmstart kernels implement:

    x16 = bf16(in_scale * (Re, Im))  (16 channels, 2..15 zero)
    e0 = relu(conv(x16)); e_k = relu(conv(pool(e_{k-1}))) for k = 1..3  (16, 32, 64, 128 channels)
    d2 = relu(conv(e2 ++ up(e3))); d1 = relu(conv(e1 ++ up(d2))); d0 = relu(conv(e0 ++ up(d1)))
    out = out_scale * (Wo d0 + bo)
    loss = sum_v w_v |out_v - t_v|^2 / (2 V out_scale^2),  w_v = 1 + pos_weight where t_v != 0

conv = 3x3x3 correlation with zero padding, tap t = (dz+1)*9 + (dy+1)*3 + (dx+1), weights
[27][Cout][Cin]; pool = 2x2x2 mean; up = nearest 2x2x2; ++ = channel concatenation.  With
``bf16=True`` every tensor the GPU stores in BF16 (input, activations, pre-activation
gradients, weights) is rounded to BF16 (round to nearest even) at the same point; sums are
float64.  Backward is the hand-derived chain rule; Adam is the textbook update.

Parity is pinned in tests/test_unet_oracle_cpu.py: conv against scipy.ndimage.correlate, the
backward pass against central finite differences of the float64 forward (bf16=False), pooling
and upsampling as adjoint pairs, Adam against a hand-worked two-step case.
"""
from __future__ import annotations

import numpy as np

LEVELS = 4
CHAN = (16, 32, 64, 128)
TAPS = 27
# (level, C0, C1, Cout) of e0 e1 e2 e3 d2 d1 d0
LAYERS = ((0, 16, 0, 16), (1, 16, 0, 32), (2, 32, 0, 64), (3, 64, 0, 128), (2, 64, 128, 64), (1, 32, 64, 32),
          (0, 16, 32, 16))
B1, B2, EPS = 0.9, 0.999, 1e-8


def bf16(a):
    """Round float values to BF16 (round to nearest even), returned as float64."""
    f = np.asarray(a, np.float32)
    u = f.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


def bf16_bits(a):
    """BF16 bit patterns (uint16) of float values, round to nearest even."""
    f = np.asarray(a, np.float32)
    u = f.view(np.uint32).astype(np.uint64)
    return (((u + 0x7FFF + ((u >> 16) & 1)) >> 16) & 0xFFFF).astype(np.uint16)


def from_bits(b):
    return (np.asarray(b, np.uint16).astype(np.uint32) << 16).view(np.float32).astype(np.float64)


def tap_offsets():
    return [(t // 9 - 1, (t // 3) % 3 - 1, t % 3 - 1) for t in range(TAPS)]


def shift(a, dz, dy, dx):
    """s[z, y, x] = a[z + dz, y + dy, x + dx], zero outside (a: [D][H][W][C])."""
    D, H, W = a.shape[:3]
    s = np.zeros_like(a)
    zs, zd = slice(max(dz, 0), D + min(dz, 0)), slice(max(-dz, 0), D + min(-dz, 0))
    ys, yd = slice(max(dy, 0), H + min(dy, 0)), slice(max(-dy, 0), H + min(-dy, 0))
    xs, xd = slice(max(dx, 0), W + min(dx, 0)), slice(max(-dx, 0), W + min(-dx, 0))
    s[zd, yd, xd] = a[zs, ys, xs]
    return s


def up(a):
    """Nearest 2x2x2 upsampling of [D][H][W][C]."""
    return a.repeat(2, 0).repeat(2, 1).repeat(2, 2)


def up_adjoint(g):
    """Adjoint of up: the sum over each voxel's 8 children."""
    D, H, W, C = g.shape
    return g.reshape(D // 2, 2, H // 2, 2, W // 2, 2, C).sum(axis=(1, 3, 5))


def pool(a):
    """2x2x2 average (the GPU sums in float32 and scales by 1/8)."""
    return up_adjoint(a) * 0.125


def conv(x, w, b=None):
    """out[v][co] = b[co] + sum_t sum_ci x[v + t][ci] w[t][co][ci] (x: [D][H][W][Cin])."""
    out = np.zeros(x.shape[:3] + (w.shape[1],))
    for t, (dz, dy, dx) in enumerate(tap_offsets()):
        out += shift(x, dz, dy, dx) @ w[t].T
    if b is not None:
        out += b
    return out


def conv_dgrad(g, w):
    """grad_in[v][ci] = sum_t sum_co g[v - t][co] w[t][co][ci]."""
    out = np.zeros(g.shape[:3] + (w.shape[2],))
    for t, (dz, dy, dx) in enumerate(tap_offsets()):
        out += shift(g, -dz, -dy, -dx) @ w[t]
    return out


def conv_wgrad(x, g):
    """dW[t][co][ci] = sum_v g[v][co] x[v + t][ci], db[co] = sum_v g[v][co]."""
    Cin, Cout = x.shape[3], g.shape[3]
    gv = g.reshape(-1, Cout)
    dw = np.zeros((TAPS, Cout, Cin))
    for t, (dz, dy, dx) in enumerate(tap_offsets()):
        dw[t] = gv.T @ shift(x, dz, dy, dx).reshape(-1, Cin)
    return dw, gv.sum(0)


# Nearest upsampling followed by a 3x3x3 convolution, per output parity class: along one axis an
# output at parity p reads the low-resolution voxels (v >> 1) + p + o - 1, o in {0, 1}, with the
# taps S[(p, o)] summed (p = 0: o = 0 takes tap -1, o = 1 taps 0 and +1; p = 1: o = 0 taps -1 and
# 0, o = 1 tap +1): 8 taps on the low-resolution grid instead of 27 on the fine one.
PARITY_TAPS = {(0, 0): (-1,), (0, 1): (0, 1), (1, 0): (-1, 0), (1, 1): (1,)}


def parity_weights(w, c0, rounding=True):
    """Combined weights of the upsampled channels (c0: ...) of w [27][Cout][Cin], summed in
    float64 from w's values and rounded to BF16 (rounding): [pz][py][px][oz][oy][ox][Cout][Cin - c0]."""
    w = np.asarray(w, np.float64)
    out = np.zeros((2, 2, 2, 2, 2, 2, w.shape[1], w.shape[2] - c0))
    for pz in range(2):
        for py in range(2):
            for px in range(2):
                for oz in range(2):
                    for oy in range(2):
                        for ox in range(2):
                            acc = np.zeros(out.shape[6:])
                            for tz in PARITY_TAPS[(pz, oz)]:
                                for ty in PARITY_TAPS[(py, oy)]:
                                    for tx in PARITY_TAPS[(px, ox)]:
                                        acc += w[(tz + 1) * 9 + (ty + 1) * 3 + (tx + 1), :, c0:]
                            out[pz, py, px, oz, oy, ox] = acc
    return bf16(out) if rounding else out


def conv_up_parity(xl, wp):
    """conv(up(xl), w) restricted to the upsampled channels, with parity_weights wp (xl: the
    low-resolution [D/2][H/2][W/2][C1])."""
    Dl, Hl, Wl, _ = xl.shape
    out = np.zeros((2 * Dl, 2 * Hl, 2 * Wl, wp.shape[6]))
    for pz in range(2):
        for py in range(2):
            for px in range(2):
                acc = np.zeros((Dl, Hl, Wl, wp.shape[6]))
                for oz in range(2):
                    for oy in range(2):
                        for ox in range(2):
                            acc += shift(xl, pz + oz - 1, py + oy - 1, px + ox - 1) @ wp[pz, py, px, oz, oy, ox].T
                out[pz::2, py::2, px::2] = acc
    return out


def param_count():
    n = 0
    for _, c0, c1, co in LAYERS:
        n += TAPS * co * (c0 + c1) + co
    return n + 34


def unpack(params):
    """Per-layer (W [27][Cout][Cin], b [Cout]) and the head (Wo [2][16], bo [2]) of a flat array."""
    p = np.asarray(params, np.float64)
    out, off = [], 0
    for _, c0, c1, co in LAYERS:
        n = TAPS * co * (c0 + c1)
        out.append((p[off:off + n].reshape(TAPS, co, c0 + c1), p[off + n:off + n + co]))
        off += n + co
    return out, (p[off:off + 32].reshape(2, 16), p[off + 32:off + 34])


def pack(layer_grads, head_grads):
    parts = []
    for dw, db in layer_grads:
        parts += [np.ravel(dw), np.ravel(db)]
    parts += [np.ravel(head_grads[0]), np.ravel(head_grads[1])]
    return np.concatenate(parts)


def forward(params, vol, in_scale, out_scale, rounding=True, parity=False):
    """Network output (complex [D][H][W]) and the activations, for volume vol [D][H][W] complex.
    parity: the decoder convolutions of up to 32 output channels take their upsampled channels
    through parity_weights (the inference path), else through the 27 BF16 taps on the upsampled
    grid (the training path, and the inference path of the 64-channel decoder layer)."""
    r = bf16 if rounding else (lambda a: np.asarray(a, np.float64))
    layers, (wo, bo) = unpack(params)
    wl = [(r(w), b) for w, b in layers]
    D, H, W = vol.shape
    x = np.zeros((D, H, W, 16))
    if rounding:
        x[..., 0] = bf16(np.float32(vol.real) * np.float32(in_scale))
        x[..., 1] = bf16(np.float32(vol.imag) * np.float32(in_scale))
    else:
        x[..., 0] = vol.real * in_scale
        x[..., 1] = vol.imag * in_scale
    act = {"x16": x}

    def layer(i, inp):
        w, b = wl[i]
        return r(np.maximum(conv(inp, w, b), 0.0))

    e = [layer(0, x)]
    p = [None]
    for k in range(1, LEVELS):
        p.append(r(pool(e[k - 1])))
        e.append(layer(k, p[k]))
    def dec(i, skip, low):
        if not parity or wl[i][0].shape[1] > 32:
            return layer(i, np.concatenate([skip, up(low)], axis=3))
        w, b = wl[i]
        c0 = skip.shape[3]
        wp = parity_weights(layers[i][0], c0, rounding)
        return r(np.maximum(conv(skip, w[:, :, :c0], b) + conv_up_parity(low, wp), 0.0))

    d2 = dec(4, e[2], e[3])
    d1 = dec(5, e[1], d2)
    d0 = dec(6, e[0], d1)
    o = d0 @ wo.T + bo
    act.update(e=e, p=p, d=[d0, d1, d2], o=o)
    return out_scale * (o[..., 0] + 1j * o[..., 1]), act


def loss_grad(params, vol, target, in_scale, out_scale, rounding=True, pos_weight=0.0):
    """Loss and the flat parameter gradient (same layout as params)."""
    r = bf16 if rounding else (lambda a: np.asarray(a, np.float64))
    layers, (wo, bo) = unpack(params)
    wl = [r(w) for w, _ in layers]
    _, a = forward(params, vol, in_scale, out_scale, rounding)
    e, p, (d0, d1, d2), o = a["e"], a["p"], a["d"], a["o"]
    V = d0.shape[0] * d0.shape[1] * d0.shape[2]
    ts = np.stack([target.real, target.imag], -1) / out_scale
    diff = o - ts
    wv = np.where((target.real != 0) | (target.imag != 0), 1.0 + pos_weight, 1.0)[..., None]
    loss = float(np.sum(wv * diff * diff) / (2.0 * V))
    g = wv * diff / V
    dwo = g.reshape(-1, 2).T @ d0.reshape(-1, 16)
    dbo = g.reshape(-1, 2).sum(0)
    grads = [None] * 7

    def mask(act_, post):
        return r(np.where(act_ > 0, post, 0.0))

    gd0 = mask(d0, g @ wo)
    grads[6] = conv_wgrad(np.concatenate([e[0], up(d1)], 3), gd0)
    dg0 = conv_dgrad(gd0, wl[6])
    gd1 = mask(d1, up_adjoint(dg0[..., 16:48]))
    grads[5] = conv_wgrad(np.concatenate([e[1], up(d2)], 3), gd1)
    dg1 = conv_dgrad(gd1, wl[5])
    gd2 = mask(d2, up_adjoint(dg1[..., 32:96]))
    grads[4] = conv_wgrad(np.concatenate([e[2], up(e[3])], 3), gd2)
    dg2 = conv_dgrad(gd2, wl[4])
    ge3 = mask(e[3], up_adjoint(dg2[..., 64:192]))
    grads[3] = conv_wgrad(p[3], ge3)
    dp3 = conv_dgrad(ge3, wl[3])
    ge2 = mask(e[2], dg2[..., :64] + up(dp3) * 0.125)
    grads[2] = conv_wgrad(p[2], ge2)
    dp2 = conv_dgrad(ge2, wl[2])
    ge1 = mask(e[1], dg1[..., :32] + up(dp2) * 0.125)
    grads[1] = conv_wgrad(p[1], ge1)
    dp1 = conv_dgrad(ge1, wl[1])
    ge0 = mask(e[0], dg0[..., :16] + up(dp1) * 0.125)
    grads[0] = conv_wgrad(a["x16"], ge0)
    return loss, pack(grads, (dwo, dbo))


def adam(params, grads, m, v, t, lr):
    """Adam step t (1-based): returns new params, m, v."""
    m = B1 * m + (1 - B1) * grads
    v = B2 * v + (1 - B2) * grads * grads
    mh = m / (1 - B1 ** t)
    vh = v / (1 - B2 ** t)
    return params - lr * mh / (np.sqrt(vh) + EPS), m, v


def init_params(seed=0):
    """He-normal weights (the input layer's fan-in is its 2 used channels), zero biases."""
    rng = np.random.default_rng(seed)
    parts = []
    for i, (_, c0, c1, co) in enumerate(LAYERS):
        cin = c0 + c1
        used = 2 if i == 0 else cin
        w = np.zeros((TAPS, co, cin))
        w[:, :, :used] = rng.normal(size=(TAPS, co, used)) * np.sqrt(2.0 / (TAPS * used))
        parts += [w.ravel(), np.zeros(co)]
    parts += [rng.normal(size=32) * 0.25, np.zeros(2)]
    return np.concatenate(parts)
