"""CPU checks that pin the warm-start network's oracle (oracle/unet3d.py): the convolution
against scipy.ndimage.correlate, its input and weight gradients as adjoints of the forward map,
pooling/upsampling as an adjoint pair, the whole backward pass against central finite
differences of the float64 forward, Adam against a hand-worked case, BF16 rounding against
known bit patterns; and a planted mistake (a dropped skip gradient) that the finite-difference
check catches."""
from __future__ import annotations

import numpy as np
import pytest
from scipy import ndimage

from oracle import unet3d as U


def test_bf16_rounding_known_answers():
    vals = np.array([1.0, 1.0 + 2.0 ** -8, 1.0 + 3 * 2.0 ** -8, -2.5, 65504.0, 1e-30], np.float32)
    bits = U.bf16_bits(vals)
    # 1 + 2^-8 is a tie between 0x3F80 and 0x3F81 -> even; 1 + 3*2^-8 -> tie -> 0x3F82
    assert bits.tolist() == [0x3F80, 0x3F80, 0x3F82, 0xC020, 0x4780, 0x0DA2]
    assert np.array_equal(U.from_bits(bits), U.bf16(vals))


def test_conv_matches_scipy_correlate():
    rng = np.random.default_rng(1)
    x = rng.normal(size=(6, 5, 7, 3))
    w = rng.normal(size=(27, 2, 3))
    b = rng.normal(size=2)
    got = U.conv(x, w, b)
    for co in range(2):
        exp = np.full(x.shape[:3], b[co])
        for ci in range(3):
            k = w[:, co, ci].reshape(3, 3, 3)  # tap t = (dz+1)*9 + (dy+1)*3 + (dx+1)
            exp += ndimage.correlate(x[..., ci], k, mode="constant", cval=0.0)
        np.testing.assert_allclose(got[..., co], exp, rtol=1e-12, atol=1e-12)


def test_conv_gradients_are_adjoints():
    rng = np.random.default_rng(2)
    x = rng.normal(size=(4, 6, 5, 3))
    w = rng.normal(size=(27, 4, 3))
    g = rng.normal(size=(4, 6, 5, 4))
    lhs = np.sum(U.conv(x, w) * g)
    assert np.isclose(lhs, np.sum(x * U.conv_dgrad(g, w)), rtol=1e-12)
    dw, db = U.conv_wgrad(x, g)
    assert np.isclose(lhs, np.sum(w * dw), rtol=1e-12)
    assert np.allclose(db, g.reshape(-1, 4).sum(0))


def test_pool_and_upsample_are_adjoint():
    rng = np.random.default_rng(3)
    a = rng.normal(size=(4, 6, 8, 2))
    c = rng.normal(size=(2, 3, 4, 2))
    assert np.isclose(np.sum(U.up(c) * a), np.sum(c * U.up_adjoint(a)), rtol=1e-12)
    np.testing.assert_allclose(U.pool(a)[1, 2, 3], a[2:4, 4:6, 6:8].mean(axis=(0, 1, 2)))


def _problem(seed, D=8, H=8, W=8):
    rng = np.random.default_rng(seed)
    vol = rng.normal(size=(D, H, W)) + 1j * rng.normal(size=(D, H, W))
    tgt = 0.3 * (rng.normal(size=(D, H, W)) + 1j * rng.normal(size=(D, H, W)))
    p = U.init_params(seed)
    # non-zero biases so every layer's bias gradient is exercised
    layers, _ = U.unpack(p)
    off = 0
    for (w, b), (_, c0, c1, co) in zip(layers, U.LAYERS):
        off += w.size
        p[off:off + co] = rng.normal(size=co) * 0.05
        off += co
    p[-2:] = [0.1, -0.2]
    return vol, tgt, p


def _fd_check(loss_grad, p, idx, vol, tgt, h=1e-6, **kw):
    _, g = loss_grad(p, vol, tgt, 0.7, 1.3, rounding=False, **kw)
    for i in idx:
        pp, pm = p.copy(), p.copy()
        pp[i] += h
        pm[i] -= h
        lp, _ = loss_grad(pp, vol, tgt, 0.7, 1.3, rounding=False, **kw)
        lm, _ = loss_grad(pm, vol, tgt, 0.7, 1.3, rounding=False, **kw)
        fd = (lp - lm) / (2 * h)
        assert abs(fd - g[i]) <= 1e-5 * max(1e-3, abs(fd)) + 1e-9, (i, fd, g[i])


def _sample_indices(rng, n_each=3):
    """Parameter indices in every layer's weights and bias and in the head."""
    idx, off = [], 0
    for i, (_, c0, c1, co) in enumerate(U.LAYERS):
        nw = U.TAPS * co * (c0 + c1)
        if i == 0:  # the input layer's used channels 0, 1
            cand = [((t * co) + o) * 16 + c for t in range(27) for o in range(co) for c in range(2)]
            idx += list(off + rng.choice(cand, n_each, replace=False))
        else:
            idx += list(off + rng.integers(0, nw, n_each))
        idx += list(off + nw + rng.integers(0, co, 1))
        off += nw + co
    idx += list(off + rng.integers(0, 34, 3))
    return idx


def test_backward_matches_finite_differences():
    vol, tgt, p = _problem(4)
    _fd_check(U.loss_grad, p, _sample_indices(np.random.default_rng(5)), vol, tgt)


def test_weighted_loss_backward_matches_finite_differences():
    vol, tgt, p = _problem(8)
    tgt[np.random.default_rng(9).random(tgt.shape) > 0.1] = 0  # a sparse target
    _fd_check(U.loss_grad, p, _sample_indices(np.random.default_rng(10), 2), vol, tgt, pos_weight=7.0)
    l0, _ = U.loss_grad(p, vol, tgt, 0.7, 1.3, rounding=False)
    l7, _ = U.loss_grad(p, vol, tgt, 0.7, 1.3, rounding=False, pos_weight=7.0)
    o, _ = U.forward(p, vol, 0.7, 1.3, rounding=False)
    nz = (tgt != 0)
    extra = 7.0 * np.sum(np.abs(o[nz] - tgt[nz]) ** 2) / (2 * vol.size * 1.3 ** 2)
    assert np.isclose(l7 - l0, extra, rtol=1e-10)


def test_finite_differences_catch_a_dropped_skip_gradient(monkeypatch):
    """Planted mistake: the decoder's skip half of d0's input gradient is dropped; the check
    on the input layer's weights must fail."""
    vol, tgt, p = _problem(6)
    orig = U.conv_dgrad

    def broken(g, w):
        out = orig(g, w)
        if out.shape[3] == 48:  # d0's input gradient: e0 (16) ++ up(d1) (32)
            out[..., :16] = 0.0
        return out

    monkeypatch.setattr(U, "conv_dgrad", broken)
    with pytest.raises(AssertionError):
        _fd_check(U.loss_grad, p, [5 * 16 * 16 + 3 * 16 + 1, 7 * 16 * 16 + 2 * 16], vol, tgt)


def test_rounded_gradients_stay_close_to_float64():
    vol, tgt, p = _problem(7)
    l64, g64 = U.loss_grad(p, vol, tgt, 0.7, 1.3, rounding=False)
    l16, g16 = U.loss_grad(p, vol, tgt, 0.7, 1.3, rounding=True)
    assert abs(l16 - l64) <= 2e-2 * abs(l64)
    assert np.linalg.norm(g16 - g64) <= 5e-2 * np.linalg.norm(g64)


def test_adam_two_steps_by_hand():
    p, m, v = np.array([1.0, -2.0]), np.zeros(2), np.zeros(2)
    g1, g2 = np.array([0.5, -0.1]), np.array([0.25, 0.3])
    p1, m, v = U.adam(p, g1, m, v, 1, 0.01)
    # step 1: m-hat = g, v-hat = g^2, so the update is lr * sign(g) (up to eps)
    np.testing.assert_allclose(p1, p - 0.01 * np.sign(g1), rtol=1e-7)
    p2, m, v = U.adam(p1, g2, m, v, 2, 0.01)
    m2 = 0.9 * 0.1 * g1 + 0.1 * g2
    v2 = 0.999 * 0.001 * g1 ** 2 + 0.001 * g2 ** 2
    exp = p1 - 0.01 * (m2 / (1 - 0.81)) / (np.sqrt(v2 / (1 - 0.999 ** 2)) + 1e-8)
    np.testing.assert_allclose(p2, exp, rtol=1e-12)


def test_param_count_matches_layout():
    assert U.param_count() == 727010 - 27 * 16 * 2 + 27 * 16 * 16
    p = U.init_params(0)
    assert p.size == U.param_count()
    layers, head = U.unpack(p)
    assert [w.shape for w, _ in layers][0] == (27, 16, 16)
    assert np.all(layers[0][0][:, :, 2:] == 0)


def test_parity_class_upsampled_conv_equals_conv_of_the_upsampled_input():
    """conv(up(x), w) over the upsampled channels equals the 8-tap parity-class form on the low
    grid with the taps summed per class (exact in float64, borders included); with BF16-rounded
    combined weights it differs by the weights' rounding only."""
    rng = np.random.default_rng(3)
    xl = rng.standard_normal((3, 4, 5, 6))
    w = rng.standard_normal((27, 5, 2 + 6))
    ref = U.conv(U.up(xl), w[:, :, 2:])
    got = U.conv_up_parity(xl, U.parity_weights(w, 2, rounding=False))
    assert np.max(np.abs(got - ref)) <= 1e-12 * np.max(np.abs(ref))
    got_r = U.conv_up_parity(xl, U.parity_weights(w, 2))
    assert np.max(np.abs(got_r - ref)) <= 2 ** -8 * 27 * np.max(np.abs(w)) * np.max(np.abs(xl)) * 6
    # a class's weights: parity (0, 0, 0) offset (1, 1, 1) holds taps {0, +1}^3 summed
    pw = U.parity_weights(w, 2, rounding=False)
    taps = [(tz + 1) * 9 + (ty + 1) * 3 + (tx + 1) for tz in (0, 1) for ty in (0, 1) for tx in (0, 1)]
    assert np.allclose(pw[0, 0, 0, 1, 1, 1], w[taps, :, 2:].sum(0))
    assert np.allclose(pw[1, 0, 1, 0, 1, 1], w[[(tz + 1) * 9 + (ty + 1) * 3 + 2 for tz in (-1, 0) for ty in (0, 1)], :, 2:].sum(0))


def test_parity_forward_is_the_network_within_bf16():
    """The inference form of the network (parity-class decoder convolutions) against the training
    form on the same parameters: within BF16 rounding of the outputs."""
    rng = np.random.default_rng(4)
    p = U.init_params(seed=5)
    vol = (rng.standard_normal((8, 16, 16)) + 1j * rng.standard_normal((8, 16, 16))) * 0.3
    a, _ = U.forward(p, vol, 0.9, 1.3)
    b, _ = U.forward(p, vol, 0.9, 1.3, parity=True)
    assert np.linalg.norm(a - b) <= 2e-2 * np.linalg.norm(a)
    c, _ = U.forward(p, vol, 0.9, 1.3, rounding=False)
    d, _ = U.forward(p, vol, 0.9, 1.3, rounding=False, parity=True)
    assert np.linalg.norm(c - d) <= 1e-12 * np.linalg.norm(c)
