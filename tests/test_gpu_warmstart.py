"""GPU tests of the learned warm start's kernels and network against the CPU oracle
(oracle/unet3d.py, float64 with the GPU's BF16 rounding points):
  - the MFMA convolution forward (BF16 out, bias + ReLU, concatenated half-resolution input),
    input gradient and weight/bias gradient on random BF16 tensors, odd plane sizes and every
    output-channel blocking (16, 48, 64, 128);
  - the whole network's loss and parameter gradients (every layer) and its prediction;
  - Adam steps that lower the loss, and the parameter round trip;
  - rihvr_reconstruct_learned's start decision (objectives of the network's, the A^T b and the
    zero start) and its solve against the plain solver's."""
from __future__ import annotations

import numpy as np
import pytest

from oracle import unet3d as U

pytestmark = pytest.mark.gpu


def _bits(torch, a):
    return torch.from_numpy(U.bf16_bits(a).view(np.int16)).cuda()


def _vals(t):
    return U.from_bits(t.cpu().numpy().view(np.uint16))


def _rel(a, b):
    return float(np.linalg.norm(np.ravel(a) - np.ravel(b)) / max(np.linalg.norm(np.ravel(b)), 1e-30))


CONV_CASES = [
    # (D, H, W, C0, C1, Cout)
    (4, 6, 20, 16, 32, 48),     # W not a multiple of 16: M-tiles span rows; concatenated input
    (2, 18, 30, 32, 0, 16),     # plane of 540 voxels: three workgroups, the last one partial
    (4, 4, 8, 64, 128, 64),     # d2's shape
    (2, 8, 8, 64, 0, 128),      # two output-channel blocks
]


@pytest.mark.parametrize("case", range(len(CONV_CASES)))
def test_conv_forward_matches_oracle(pkg, gpu_available, case):
    import torch
    D, H, W, C0, C1, Cout = CONV_CASES[case]
    print(f"conv forward case {CONV_CASES[case]}", flush=True)
    rng = np.random.default_rng(case)
    x0 = U.bf16(rng.normal(size=(D, H, W, C0)))
    x1 = U.bf16(rng.normal(size=(D // 2, H // 2, W // 2, C1))) if C1 else None
    w = U.bf16(rng.normal(size=(27, Cout, C0 + C1)) * 0.1)
    b = rng.normal(size=Cout).astype(np.float32)
    got = pkg.conv3d_bf16_forward(_bits(torch, x0), _bits(torch, w), torch.from_numpy(b).cuda(),
                                  None if x1 is None else _bits(torch, x1), relu=True)
    x = x0 if x1 is None else np.concatenate([x0, U.up(x1)], 3)
    exact = np.maximum(U.conv(x, w, b.astype(np.float64)), 0.0)
    # bound: the BF16 rounding of the output (half an ulp, at most 2^-8 relative) plus the FP32
    # accumulation error of 27 * Cin terms in another order, bounded by 2e-5 of the sum of the
    # terms' magnitudes (sums that cancel to small outputs are where an ulp comparison fails)
    mag = U.conv(np.abs(x), np.abs(w), np.abs(b).astype(np.float64))
    g = U.from_bits(got.cpu().numpy().view(np.uint16))
    err = np.abs(g - exact)
    bound = 2.0 ** -8 * np.abs(exact) + 2e-5 * mag
    print(f"max err/bound {np.max(err / bound):.3f}; exact BF16 {np.mean(U.bf16_bits(exact) == U.bf16_bits(g)):.4f}",
          flush=True)
    assert np.all(err <= bound)
    assert np.mean(U.bf16_bits(exact) == U.bf16_bits(g)) > 0.98


@pytest.mark.parametrize("case", range(len(CONV_CASES)))
def test_conv_dgrad_matches_oracle(pkg, gpu_available, case):
    import torch
    D, H, W, C0, C1, Cout = CONV_CASES[case]
    Cin = C0 + C1
    rng = np.random.default_rng(10 + case)
    g = U.bf16(rng.normal(size=(D, H, W, Cout)))
    w = U.bf16(rng.normal(size=(27, Cout, Cin)) * 0.1)
    wT = np.ascontiguousarray(w[::-1].transpose(0, 2, 1))  # wT[26 - t][ci][co] = w[t][co][ci]
    got = pkg.conv3d_bf16_dgrad(_bits(torch, g), _bits(torch, wT)).cpu().numpy()
    exp = U.conv_dgrad(g, w)
    assert _rel(got, exp) <= 1e-5


@pytest.mark.parametrize("case", range(len(CONV_CASES)))
def test_conv_wgrad_matches_oracle(pkg, gpu_available, case):
    import torch
    D, H, W, C0, C1, Cout = CONV_CASES[case]
    rng = np.random.default_rng(20 + case)
    x0 = U.bf16(rng.normal(size=(D, H, W, C0)))
    x1 = U.bf16(rng.normal(size=(D // 2, H // 2, W // 2, C1))) if C1 else None
    g = U.bf16(rng.normal(size=(D, H, W, Cout)))
    dw, db = pkg.conv3d_bf16_wgrad(_bits(torch, x0), _bits(torch, g), None if x1 is None else _bits(torch, x1))
    x = x0 if x1 is None else np.concatenate([x0, U.up(x1)], 3)
    edw, edb = U.conv_wgrad(x, g)
    assert _rel(dw.cpu().numpy(), edw) <= 1e-5
    assert _rel(db.cpu().numpy(), edb) <= 1e-5


def _net_problem(seed, shape=(16, 16, 24)):
    rng = np.random.default_rng(seed)
    vol = (rng.normal(size=shape) + 1j * rng.normal(size=shape)).astype(np.complex64)
    tgt = (0.2 * (rng.normal(size=shape) + 1j * rng.normal(size=shape))).astype(np.complex64)
    p = U.init_params(seed)
    layers, _ = U.unpack(p)
    off = 0
    for (w, b), (_, c0, c1, co) in zip(layers, U.LAYERS):
        off += w.size
        p[off:off + co] = rng.normal(size=co) * 0.05
        off += co
    p[-2:] = [0.05, -0.1]
    return vol, tgt, p.astype(np.float32)


def _layer_slices():
    out, off = [], 0
    for i, (_, c0, c1, co) in enumerate(U.LAYERS):
        n = 27 * co * (c0 + c1)
        out.append((f"W{i}", slice(off, off + n)))
        out.append((f"b{i}", slice(off + n, off + n + co)))
        off += n + co
    out.append(("head", slice(off, off + 34)))
    return out


def test_network_gradients_match_oracle(pkg, gpu_available):
    import torch
    vol, tgt, p = _net_problem(3)
    print("network gradient case 16x16x24", flush=True)
    with pkg.WarmStartNet(*vol.shape) as net:
        net.set_params(p, 0.8, 1.5)
        loss, g = net.loss_grad(torch.from_numpy(vol).cuda(), torch.from_numpy(tgt).cuda())
    el, eg = U.loss_grad(p.astype(np.float64), vol.astype(np.complex128), tgt.astype(np.complex128), 0.8, 1.5)
    print(f"loss GPU {loss:.9g} oracle {el:.9g}", flush=True)
    assert abs(loss - el) <= 1e-3 * abs(el)
    for name, sl in _layer_slices():
        r = _rel(g[sl], eg[sl])
        print(f"  {name}: rel-L2 {r:.2e}", flush=True)
        assert r <= 3e-2, name


def test_weighted_loss_gradients_match_oracle(pkg, gpu_available):
    """The loss weight on the target's non-zero voxels (a sparse target, as the solver's)."""
    import torch
    vol, tgt, p = _net_problem(6)
    rng = np.random.default_rng(60)
    tgt[rng.random(tgt.shape) > 0.02] = 0
    with pkg.WarmStartNet(*vol.shape) as net:
        net.set_params(p, 0.8, 1.5)
        net.set_loss_weight(40.0)
        loss, g = net.loss_grad(torch.from_numpy(vol).cuda(), torch.from_numpy(tgt).cuda())
    el, eg = U.loss_grad(p.astype(np.float64), vol.astype(np.complex128), tgt.astype(np.complex128), 0.8, 1.5,
                         pos_weight=40.0)
    assert abs(loss - el) <= 1e-3 * abs(el)
    assert _rel(g, eg) <= 3e-2


@pytest.mark.parametrize("shape", [(8, 32, 16), (8, 24, 136), (16, 16, 64)])
def test_network_prediction_matches_oracle(pkg, gpu_available, shape):
    import torch
    vol, _, p = _net_problem(4, shape=shape)
    with pkg.WarmStartNet(*vol.shape) as net:
        net.set_params(p, 0.8, 1.5)
        got = net.predict(torch.from_numpy(vol).cuda()).cpu().numpy()
        a, i, o = net.get_params()
    # the prediction takes the decoder's upsampled channels in parity-class form (8 combined
    # BF16 taps on the half-resolution grid); the training form rounds the 27 taps separately
    exp, _ = U.forward(p.astype(np.float64), vol.astype(np.complex128), 0.8, 1.5, parity=True)
    plain, _ = U.forward(p.astype(np.float64), vol.astype(np.complex128), 0.8, 1.5)
    print(f"prediction vs oracle: {_rel(got, exp):.3e} (parity form), {_rel(got, plain):.3e} (training form)",
          flush=True)
    assert _rel(got, exp) <= 1e-2
    assert _rel(got, plain) <= 3e-2
    assert np.array_equal(a, p) and i == np.float32(0.8) and o == np.float32(1.5)


def test_adam_lowers_the_loss_and_matches_oracle_step(pkg, gpu_available):
    import torch
    vol, tgt, p = _net_problem(5)
    v, t = torch.from_numpy(vol).cuda(), torch.from_numpy(tgt).cuda()
    with pkg.WarmStartNet(*vol.shape) as net:
        net.set_params(p, 0.8, 1.5)
        l0, g = net.loss_grad(v, t)
        net.adam_step(1e-3)
        p1, _, _ = net.get_params()
        losses = [net.train_step(v, t, 1e-3) for _ in range(30)]
    # the first Adam step moves every parameter with a non-zero gradient by lr (|m-hat / sqrt(v-hat)| = 1)
    e1, _, _ = U.adam(p.astype(np.float64), g.astype(np.float64), np.zeros(p.size), np.zeros(p.size), 1, 1e-3)
    np.testing.assert_allclose(p1, e1, rtol=1e-6, atol=1e-7)
    print(f"loss {l0:.5g} -> {losses[-1]:.5g}", flush=True)
    assert losses[-1] < 0.9 * l0


def _small_cfg():
    return dict(width=64, height=64, num_planes=16, start_plane=1000.0, plane_stepsize=40.0, wavelength=0.633,
                resolution=2.0, num_inverse_iterations=40, regularization_param=0.01,
                regularization_TV_param=0.01, num_tv_iterations=5, mode=7)


def test_zero_network_start_equals_the_plain_solve(pkg, scenes, gpu_available):
    """A network whose output is zero everywhere (all parameters 0) proposes the zero start: the
    thresholded start's objective is the zero start's 0.5 ||b||^2 (every threshold leaves it
    zero), the A^T b start has zero data residual (A A^T = Nz I on a real hologram), the decision
    follows the rule, and either way the solve equals the plain solver's solve to the same target
    bit for bit."""
    import torch
    c = _small_cfg()
    holo, _ = scenes.disks_scene(scenes.optics_of(c), 8, 2.0, 4.0, 11, 0.01)
    p = np.zeros(pkg.net_param_count(), np.float32)
    with pkg.SparseCompressiveHolo(pkg.make_params(**c)) as rec, pkg.WarmStartNet(16, 64, 64) as net:
        net.set_params(p, 1.0, 1.0)
        plain = rec.reconstructBatch(holo[None])
        target = (1 + 1e-3) * plain[0]["f_final"]
        st = rec.reconstructLearned(net, torch.from_numpy(holo).cuda(), f_target=target)
        vol = rec.getVolume()
        ref = rec.reconstructBatch(holo[None], f_target=[target], max_iterations=c["num_inverse_iterations"])
        vref = rec.getVolume()
    print(st, flush=True)
    b2 = 0.5 * float(np.sum(np.abs(holo.astype(np.complex128)) ** 2))
    assert abs(st["f_zero"] - b2) <= 1e-6 * b2
    assert np.isnan(st["f_net"]) and np.isnan(st["resid_net"])  # not evaluated
    assert st["tau"] == 0.0 and abs(st["f_thresholded"] - b2) <= 1e-6 * b2
    assert st["resid_adjoint"] <= 1e-9 * b2
    assert st["fallback"] == (0 if st["f_thresholded"] <= min(st["f_adjoint"], st["f_zero"]) else 1)
    assert st["iterations"] == ref[0]["iterations"]
    assert np.array_equal(vol, vref)


def test_learned_start_decision_follows_the_residual_rule(pkg, scenes, gpu_available):
    """With a randomly initialised network: tau is q max|x_net| for the q in {0.02, 0.05}
    whose thresholded start has the least objective (each candidate's objective recomputed here
    by rihvr_begin on a soft threshold formed by torch from net.predict(A^T b)); the decision is
    the rule's (the thresholded start when its objective is not above the A^T b start's nor the
    zero start's, else the zero start); the objective the solver computes at the chosen start
    equals the one the decision used; without a target the solve runs max_iterations."""
    import torch
    c = _small_cfg()
    holo, _ = scenes.disks_scene(scenes.optics_of(c), 8, 2.0, 4.0, 12, 0.01)
    h = torch.from_numpy(holo).cuda()
    shape = (c["num_planes"], c["height"], c["width"])
    with pkg.SparseCompressiveHolo(pkg.make_params(**c)) as rec, pkg.WarmStartNet(16, 64, 64, seed=3) as net:
        st = rec.reconstructLearned(net, h, f_target=0.0, max_iterations=5)
        vol_learned = rec.getVolume()
        g = torch.empty(shape, dtype=torch.complex64, device="cuda")
        rec.adjoint(h, out=g)
        x = net.predict(g)
        torch.cuda.synchronize()
        xmax = float(torch.abs(x).max())
        fq = {}
        for q in (0.02, 0.05):
            tau = np.float32(q * xmax)
            m = torch.abs(x)
            xs = x * torch.clamp((m - float(tau)) / torch.clamp(m, min=1e-30), min=0)
            fq[q] = (float(tau), rec.begin(h, warm_start=xs.contiguous()), xs.contiguous())
        # the learned solve started from its candidate in a ring slot with the candidate's line
        # map; the same start passed from outside (no map: every line read) gives the same solve
        best_q = min(fq, key=lambda k: fq[k][1])
        ref = None
        if st["fallback"] == 0:
            rec.reconstructBatch(h[None], warm_start=fq[best_q][2], max_iterations=5)
            ref = rec.getVolume()
    fq = {k: v[:2] for k, v in fq.items()}
    print(st, fq, flush=True)
    if ref is not None:
        assert np.linalg.norm(vol_learned - ref) <= 1e-5 * np.linalg.norm(ref)
    best = min(fq.values(), key=lambda v: v[1])
    assert abs(st["f_thresholded"] - best[1]) <= 1e-5 * best[1]
    assert any(abs(st["tau"] - t) <= 1e-6 * t and abs(f - best[1]) <= 1e-5 * best[1] for t, f in fq.values())
    assert st["fallback"] == (0 if st["f_thresholded"] <= min(st["f_adjoint"], st["f_zero"]) else 1)
    f_chosen = st["f_zero"] if st["fallback"] else st["f_thresholded"]
    assert abs(st["f_start"] - f_chosen) <= 1e-4 * f_chosen
    assert st["iterations"] == 5
