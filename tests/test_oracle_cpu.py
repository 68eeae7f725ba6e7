"""CPU tests: the oracle pinned by an independent numpy restatement, committed
golden vectors, closed forms and invariants (K1-K8), brute force on tiny inputs,
and mistakes planted in the oracle's restatement that the checks must catch.
"""
import pathlib
import sys

import numpy as np
import pytest

HERE = pathlib.Path(__file__).resolve().parent
sys.path.insert(0, str(HERE / "golden"))
import numpy_restatement as R  # noqa: E402

GOLD = np.load(HERE / "golden" / "golden_small.npz")


def rel(a, b):
    return float(np.linalg.norm(np.asarray(a) - np.asarray(b)) / max(np.linalg.norm(np.asarray(b)), 1e-300))


def params(orc, W, H, Nz, z0=1000.0, dz=50.0, lam=0.633, dx=2.0, its=5, reg=0.01, tv=0.01, ntv=5, mode=7):
    return orc.make_params(W, H, Nz, z0, dz, lam, dx, its, reg, tv, ntv, mode)


# --------------------------------------------------------------------------- FFT
@pytest.mark.parametrize("H,W", [(16, 16), (32, 48), (28, 20), (26, 39), (64, 64), (7, 13)])
def test_fft2_matches_numpy(orc, H, W):
    rng = np.random.default_rng(H * 100 + W)
    a = rng.standard_normal((H, W)) + 1j * rng.standard_normal((H, W))
    for prec, tol in (("f64", 1e-13), ("f32", 3e-7)):
        f = orc.fft2(a, -1, prec)
        assert np.abs(f - np.fft.fft2(a)).max() / np.abs(np.fft.fft2(a)).max() < tol
        fi = orc.fft2(a, +1, prec)
        ref = np.fft.ifft2(a) * H * W
        assert np.abs(fi - ref).max() / np.abs(ref).max() < tol


def test_fft2_brute_force_dft(orc):
    """Dense DFT matrices on a tiny non-power-of-2 grid."""
    H, W = 6, 10
    rng = np.random.default_rng(1)
    a = rng.standard_normal((H, W)) + 1j * rng.standard_normal((H, W))
    FH = np.exp(-2j * np.pi * np.outer(np.arange(H), np.arange(H)) / H)
    FW = np.exp(-2j * np.pi * np.outer(np.arange(W), np.arange(W)) / W)
    np.testing.assert_allclose(orc.fft2(a, -1, "f64"), FH @ a @ FW.T, rtol=0, atol=1e-12)


# ----------------------------------------------------------------- tables / params
def test_kappa_and_planes_match_restatement(orc):
    p = params(orc, 32, 32, 4)
    np.testing.assert_allclose(orc.kappa(p), GOLD["A_kappa"], rtol=1e-15, atol=0)
    np.testing.assert_array_equal(orc.plane_list(p), GOLD["A_planes"])


def test_kappa_evanescent_clamp(orc):
    """Q13: spatial frequencies past lambda/dx give kappa = 0 (H = 1)."""
    p = params(orc, 32, 32, 1, lam=1.5, dx=1.0)   # (lambda/dx)^2 f^2 > 1 at the band edge
    k = orc.kappa(p)
    assert k[16, 16] == 0.0 and k[0, 0] == pytest.approx(-2 * np.pi / float(np.float32(1.5)), rel=1e-15)


def test_plane_list_is_float32(orc):
    p = params(orc, 16, 16, 5, z0=0.1, dz=0.7)
    z = orc.plane_list(p)
    want = [np.float32(0.1) + np.float32(np.float32(i) * np.float32(0.7)) for i in range(5)]
    np.testing.assert_array_equal(z, np.array(want, np.float32))


# ------------------------------------------------------------- operators (golden)
@pytest.mark.parametrize("case,shape,optics", [
    ("A", (32, 32, 4), dict(z0=1000.0, dz=50.0, lam=0.633, dx=2.0)),
    ("B", (48, 32, 3), dict(z0=2500.0, dz=20.0, lam=0.632, dx=1.1)),
])
def test_operators_match_golden(orc, case, shape, optics):
    W, H, Nz = shape
    p = params(orc, W, H, Nz, **optics)
    x, r = GOLD[f"{case}_x"], GOLD[f"{case}_r"]
    # FP64 oracle == FP64 restatement
    assert rel(orc.forward(p, x, "f64"), GOLD[f"{case}_fwd_f64"]) < 1e-12
    assert rel(orc.adjoint(p, r, "f64"), GOLD[f"{case}_adj_f64"]) < 1e-12
    # FP32 oracle == restatement with float-rounded phase arguments, to FP32 precision
    assert rel(orc.forward(p, x, "f32"), GOLD[f"{case}_fwd_round"]) < 2e-6
    assert rel(orc.adjoint(p, r, "f32"), GOLD[f"{case}_adj_round"]) < 2e-6


def test_phase_rounding_is_visible_at_large_z():
    """Q2 pin: at z ~ 2500 um the float rounding of z*kappa moves A by >> FP32 noise,
    so an oracle that dropped the rounding would fail test_operators_match_golden."""
    d = rel(GOLD["B_fwd_round"], GOLD["B_fwd_f64"])
    assert d > 1e-4, d
    d2 = rel(GOLD["B_adj_round"], GOLD["B_adj_f64"])
    assert d2 > 1e-4, d2


def test_prox_matches_golden(orc):
    x0, g = GOLD["A_prox_x0"], GOLD["A_prox_g"]
    p7 = params(orc, 32, 32, 4, reg=0.05, tv=0.02, ntv=5, mode=7)
    assert rel(orc.prox(p7, x0, g, 0.25, "f64"), GOLD["A_prox_fl2d"]) < 1e-12
    assert rel(orc.prox(p7, x0, g, 0.25, "f32"), GOLD["A_prox_fl2d"]) < 1e-6
    p5 = params(orc, 32, 32, 4, reg=0.05, tv=0.0, mode=5)
    assert rel(orc.prox(p5, x0, g, 0.25, "f64"), GOLD["A_prox_l1"]) < 1e-14


def test_solve_matches_golden(orc):
    p = params(orc, 32, 32, 4, its=6, reg=0.01, tv=0.01, ntv=5)
    vol, est, res, st = orc.solve(p, GOLD["A_holo"], None, "f64")
    assert rel(vol, GOLD["A_solve_x"]) < 1e-9
    assert rel(est, GOLD["A_solve_est"]) < 1e-9
    np.testing.assert_allclose(st["f"], GOLD["A_solve_f"], rtol=1e-9)
    assert st["f_initial"] == pytest.approx(float(GOLD["A_solve_f_initial"]), rel=1e-12)
    assert list(st["restarts"]) == list(GOLD["A_solve_restarts"])
    assert list(st["backtracks"]) == list(GOLD["A_solve_backtracks"])
    # FP32 oracle: within FP32 rounding of the rounded-argument restatement; its
    # distance to the FP64 operator is the float rounding of z*kappa (Q2)
    vol32, _, _, st32 = orc.solve(p, GOLD["A_holo"], None, "f32")
    assert rel(vol32, GOLD["A_solve_x_round"]) < 1e-5
    np.testing.assert_allclose(st32["f"], GOLD["A_solve_f_round"], rtol=1e-5)
    assert 1e-5 < rel(vol32, GOLD["A_solve_x"]) < 1e-2
    assert list(st32["restarts"]) == list(GOLD["A_solve_restarts"])


# ------------------------------------------------------------ closed forms K1..K8
def test_K1_adjoint_identity(orc):
    rng = np.random.default_rng(7)
    p = params(orc, 32, 48, 5, z0=2500.0, dz=20.0, lam=0.632, dx=1.1)
    x = rng.standard_normal((5, 48, 32)) + 1j * rng.standard_normal((5, 48, 32))
    r = rng.standard_normal((48, 32)) + 0j
    for prec, tol in (("f64", 1e-12), ("f32", 5e-6)):
        lhs = np.sum(orc.forward(p, x, prec).real * r.real)
        rhs = np.sum((np.conj(x) * orc.adjoint(p, r, prec)).real)
        assert abs(lhs - rhs) <= tol * abs(lhs), (prec, lhs, rhs)


def test_K2_constant_plane_maps_to_its_real_part(orc):
    p = params(orc, 32, 32, 4, z0=2500.0, dz=20.0, lam=0.632, dx=1.1)
    x = np.zeros((4, 32, 32), complex)
    x[2] = 0.7 - 0.3j
    np.testing.assert_allclose(orc.forward(p, x, "f64"), np.full((32, 32), 0.7 + 0j), atol=1e-13)


def test_K3_single_fourier_mode(orc):
    W = H = 32
    p = params(orc, W, H, 3, z0=1000.0, dz=50.0, lam=0.633, dx=2.0)
    kx, ky = 3, 5
    yy, xx = np.mgrid[0:H, 0:W]
    mode = np.exp(2j * np.pi * (kx * xx / W + ky * yy / H))
    x = np.zeros((3, H, W), complex)
    x[1] = mode
    z = float(np.float32(1000.0) + np.float32(50.0))
    lam = float(np.float32(0.633))
    f2 = (kx / W) ** 2 + (ky / H) ** 2
    kap = -2 * np.pi * np.sqrt(1 - f2 * (lam / float(np.float32(2.0))) ** 2) / lam
    want = (np.exp(-2j * np.pi * z / lam) * np.exp(-1j * z * kap) * mode).real + 0j
    np.testing.assert_allclose(orc.forward(p, x, "f64"), want, atol=1e-11)


def test_K4_operator_norm_bound(orc):
    """||A||^2 <= Nz (power iteration on A^T A), tight on the DC mode."""
    Nz = 4
    p = params(orc, 32, 32, Nz, z0=2500.0, dz=20.0, lam=0.632, dx=1.1)
    rng = np.random.default_rng(3)
    v = rng.standard_normal((Nz, 32, 32)) + 1j * rng.standard_normal((Nz, 32, 32))
    for _ in range(30):
        v = orc.adjoint(p, orc.forward(p, v, "f64"), "f64")
        lam_est = np.linalg.norm(v)
        v /= lam_est
    assert lam_est <= Nz * (1 + 1e-9)
    # DC: e^{-i phi_z} and H_{-z}(0) = e^{+i phi_z} cancel, so x_z = 1 gives A x = Nz
    dc = np.ones((Nz, 32, 32), complex)
    ratio = np.linalg.norm(orc.forward(p, dc, "f64")) ** 2 / np.linalg.norm(dc) ** 2
    assert ratio == pytest.approx(Nz, rel=1e-12)


def test_K5_tv_prox_invariants(orc):
    rng = np.random.default_rng(9)
    Nz, H, W = 2, 16, 32
    p = params(orc, W, H, Nz, reg=0.0, tv=0.05, ntv=5)
    # constant input is returned unchanged (div of a constant-input dual stays 0)
    c = np.full((Nz, H, W), 0.3 - 0.2j)
    np.testing.assert_allclose(orc.prox(p, c, np.zeros_like(c), 0.5, "f64"), c, atol=1e-15)
    # n_tv = 0 returns x-hat
    p0 = params(orc, W, H, Nz, reg=0.0, tv=0.05, ntv=0)
    x0 = rng.standard_normal((Nz, H, W)) + 1j * rng.standard_normal((Nz, H, W))
    g = rng.standard_normal((Nz, H, W)) + 1j * rng.standard_normal((Nz, H, W))
    np.testing.assert_allclose(orc.prox(p0, x0, g, 0.5, "f64"), x0 - 0.5 * g, atol=1e-15)
    # the TV stage is equivariant under a global phase (joint re/im projection)
    th = np.exp(0.7j)
    a = orc.prox(p, x0 * th, g * th, 0.5, "f64")
    b = orc.prox(p, x0, g, 0.5, "f64") * th
    np.testing.assert_allclose(a, b, atol=1e-12)
    # ... but the per-component soft threshold is not
    ps = params(orc, W, H, Nz, reg=0.5, tv=0.05, ntv=5)
    a = orc.prox(ps, x0 * th, g * th, 0.5, "f64")
    b = orc.prox(ps, x0, g, 0.5, "f64") * th
    assert np.abs(a - b).max() > 1e-3


def test_K5_tv_prox_brute_force_against_rof_dual(orc):
    """On an 8x8 plane, many dual iterations converge to the ROF minimiser of
    0.5||y - xhat||^2 + mu * TV(y) (periodic, isotropic over re/im); check the
    optimality of the result by comparing its objective with perturbed points."""
    rng = np.random.default_rng(11)
    H = W = 16
    p = params(orc, W, H, 1, reg=0.0, tv=0.1, ntv=400)
    xh = rng.standard_normal((1, H, W)) + 1j * rng.standard_normal((1, H, W))
    t = 1.0
    y = orc.prox(p, xh, np.zeros_like(xh), t, "f64")
    mu = t * 0.1

    def F(v):
        a = np.roll(v, 1, axis=2) - v
        b = np.roll(v, 1, axis=1) - v
        return 0.5 * np.sum(np.abs(v - xh) ** 2) + mu * np.sum(np.sqrt(np.abs(a) ** 2 + np.abs(b) ** 2))

    f0 = F(y)
    for _ in range(20):
        d = 1e-3 * (rng.standard_normal(y.shape) + 1j * rng.standard_normal(y.shape))
        assert F(y + d) >= f0 - 1e-6
    assert F(y) < F(xh)


def test_K6_soft_threshold_closed_form(orc):
    rng = np.random.default_rng(4)
    Nz, H, W = 1, 16, 16
    thr_reg = 0.3
    p = params(orc, W, H, Nz, reg=thr_reg, tv=0.0, mode=5)
    x0 = (rng.standard_normal((Nz, H, W)) + 1j * rng.standard_normal((Nz, H, W))).astype(np.complex64)
    y = orc.prox(p, x0, np.zeros_like(x0), 1.0, "f32")
    T = np.float32(thr_reg)

    def st(v):
        return np.sign(v) * np.maximum(np.abs(v) - T, 0)

    np.testing.assert_allclose(y.real, st(x0.real), rtol=3e-7, atol=1e-7)
    np.testing.assert_allclose(y.imag, st(x0.imag), rtol=3e-7, atol=1e-7)


def test_K7_alpha_sequence_and_holo_linearity(orc):
    """FISTA alpha_{k+1} = (1 + sqrt(1 + 4 alpha_k^2)) / 2 from 1 gives momentum steps
    0, 0.2818, 0.4340, ...; the solver's hologram momentum by linearity equals A of the
    returned volume."""
    a = [1.0]
    for _ in range(4):
        a.append((1 + np.sqrt(1 + 4 * a[-1] ** 2)) / 2)
    assert a[1] == pytest.approx(1.6180339887, rel=1e-10)
    assert a[2] == pytest.approx(2.1935270853, rel=1e-9)
    p = params(orc, 32, 32, 4, its=6, reg=0.01, tv=0.01)
    vol, est, res, st = orc.solve(p, GOLD["A_holo"], None, "f64")
    np.testing.assert_allclose(est, orc.forward(p, vol, "f64"), atol=1e-10)


def test_K8_single_plane_at_zero(orc):
    rng = np.random.default_rng(2)
    p = params(orc, 16, 32, 1, z0=0.0, dz=1.0)
    x = rng.standard_normal((1, 32, 16)) + 1j * rng.standard_normal((1, 32, 16))
    np.testing.assert_allclose(orc.forward(p, x, "f64"), x[0].real + 0j, atol=1e-13)
    r = rng.standard_normal((32, 16)) + 0j
    np.testing.assert_allclose(orc.adjoint(p, r, "f64")[0], r, atol=1e-13)


# --------------------------------------------------------------- reductions
def test_restart_sum_quirk(orc):
    """Q7: the reference's restart reduction weights offsets 0..511 of each
    1024-chunk by d_re + 2 d_im and offsets 512..1023 by d_im only
    (sparse_compressive_holo.cu:4411-4412).  Pinned with single-voxel inputs."""
    def one(idx, dre_on, dim_on):
        x0 = np.zeros((1, 32, 64), np.complex64)
        x1 = np.zeros_like(x0)
        y0 = np.zeros_like(x0)
        flat = x0.reshape(-1)
        x1f, y0f = x1.reshape(-1), y0.reshape(-1)
        # d = (x0 - x1)(x1 - y0) per component: x0 = 3, x1 = 1, y0 = 0 gives d = 2
        flat[idx] = (3.0 if dre_on else 0.0) + 1j * (3.0 if dim_on else 0.0)
        x1f[idx] = (1.0 if dre_on else 0.0) + 1j * (1.0 if dim_on else 0.0)
        return orc.restart_sum_f32(x0, x1, y0)

    assert one(5, True, False) == 2.0           # first half: d_re counted once
    assert one(5, False, True) == 4.0           # first half: d_im counted twice
    assert one(512 + 5, True, False) == 0.0     # second half: d_re dropped
    assert one(512 + 5, False, True) == 2.0     # second half: d_im once
    assert one(1024 + 700, False, True) == 2.0  # the pattern repeats per 1024-chunk
    rng = np.random.default_rng(5)
    x0, x1, y0 = ((rng.standard_normal((2, 32, 64)) + 1j * rng.standard_normal((2, 32, 64))).astype(np.complex64)
                  for _ in range(3))
    got = orc.restart_sum_f32(x0, x1, y0)
    want = R.restart_sum(x0.astype(complex), x1.astype(complex), y0.astype(complex))
    assert got == pytest.approx(want, rel=1e-4)


def test_reduction_trees_close_to_double(orc):
    rng = np.random.default_rng(6)
    v = (rng.standard_normal((3, 64, 64)) + 1j * rng.standard_normal((3, 64, 64))).astype(np.complex64)
    v[v.real > 1.5] = 0
    assert orc.summag_f32(v) == pytest.approx(float(np.sum(np.abs(v.astype(complex)) ** 2)), rel=1e-5)
    assert orc.l1_f32(v) == pytest.approx(float(np.sum(np.abs(v.astype(complex)))), rel=1e-5)
    assert orc.tv2d_f32(v) == pytest.approx(R.tv2d(v.astype(complex)), rel=1e-5)


# ------------------------------------------------------------ planted mistakes
def test_planted_mistakes_are_caught():
    """Mistakes planted in a copy of the restatement (on the CPU) change the
    checked quantities well beyond the tolerances used above."""
    W, H, Nz = 32, 32, 4
    opt = dict(W=W, H=H, nz=Nz, z0=1000.0, dz=50.0, wavelength=0.633, resolution=2.0)
    op = R.Operator(round_args=False, **opt)
    x, r = GOLD["A_x"], GOLD["A_r"]
    good = op.forward(x)
    # (a) wrong sign of the transfer-function exponent in A
    bad = np.zeros_like(good)
    S = 0
    for i in range(Nz):
        z = -op.z[i]
        S = S + np.exp(-1j * z * op.kap) * np.fft.fft2(x[i] * np.exp(-2j * np.pi * z / op.lam)) / op.N
    bad = (np.fft.ifft2(S) * op.N).real
    assert rel(bad, good) > 1e-2
    # (b) dropped per-plane phase factor in A^T
    g_good = op.adjoint(r)
    g_bad = g_good * np.exp(-2j * np.pi * op.z[:, None, None] / op.lam)
    assert rel(g_bad, g_good) > 1e-2
    # (c) forward instead of backward differences in the TV projection
    x0, gr = GOLD["A_prox_x0"], GOLD["A_prox_g"]
    good_p = GOLD["A_prox_fl2d"]
    xh = x0 - 0.25 * gr
    gx = np.zeros_like(xh)
    gy = np.zeros_like(xh)
    mu, sc = 0.25 * 0.02, 1 / (12 * 0.25 * 0.02)
    for _ in range(5):
        y = xh - mu * ((np.roll(gx, -1, 2) - gx) + (np.roll(gy, -1, 1) - gy))
        ax = gx + sc * (np.roll(y, -1, 2) - y)
        ay = gy + sc * (np.roll(y, -1, 1) - y)
        n = np.maximum(np.sqrt(np.abs(ax) ** 2 + np.abs(ay) ** 2), 1)
        gx, gy = ax / n, ay / n
    y = R.soft(xh - mu * ((np.roll(gx, -1, 2) - gx) + (np.roll(gy, -1, 1) - gy)), 0.25 * 0.05)
    assert rel(y, good_p) > 1e-3
    # (d) mode 6 with z not periodic (plane 0 has no previous plane) or with the
    # isotropic projection left on: both fail the closed form of the TV stage
    rng = np.random.default_rng(31)
    Nz6, H6, W6 = 6, 16, 32
    x6 = rng.standard_normal((Nz6, H6, W6)) + 1j * rng.standard_normal((Nz6, H6, W6)) + 0.5
    closed = np.fft.ifftn(np.fft.fftn(x6) * (1 - _laplace_symbol(Nz6, H6, W6) / 12.0) ** 3)
    for variant in ("open_z", "projected"):
        mu6, sc6 = 0.1, 1 / (12 * 0.1)
        gx, gy, gz = (np.zeros_like(x6) for _ in range(3))

        def est():
            gzn = np.roll(gz, -1, 0)
            if variant == "open_z":
                gzn[-1] = 0
            return x6 - mu6 * ((np.roll(gx, -1, 2) - gx) + (np.roll(gy, -1, 1) - gy) + (gzn - gz))
        for _ in range(3):
            y = est()
            yp = np.roll(y, 1, 0)
            if variant == "open_z":
                yp[0] = y[0]
            ax = gx + sc6 * (np.roll(y, 1, 2) - y)
            ay = gy + sc6 * (np.roll(y, 1, 1) - y)
            az = gz + sc6 * (yp - y)
            n = 1.0
            if variant == "projected":
                n = np.maximum(np.sqrt(np.abs(ax) ** 2 + np.abs(ay) ** 2 + np.abs(az) ** 2), 1)
            gx, gy, gz = ax / n, ay / n, az / n
        assert rel(est(), closed) > 1e-3, variant


# ------------------------------------------------------------------ mode 6 (3D)
def _laplace_symbol(Nz, H, W):
    """Eigenvalues of div(grad) for the reference's stencils (backward difference
    gradient, forward difference divergence, periodic): sum_d 4 sin^2(pi k_d / N_d)."""
    kz = np.fft.fftfreq(Nz)[:, None, None]
    ky = np.fft.fftfreq(H)[None, :, None]
    kx = np.fft.fftfreq(W)[None, None, :]
    return 4 * (np.sin(np.pi * kz) ** 2 + np.sin(np.pi * ky) ** 2 + np.sin(np.pi * kx) ** 2)


@pytest.mark.parametrize("ntv", [1, 3, 5])
def test_mode6_tv_stage_closed_form(orc, ntv):
    """With the normaliser forced to 1 the dual iteration is linear and, as
    mu * sc = 1/12, telescopes: y = (I - div grad / 12)^n x.  With no exact zeros
    and zero L1 weight the soft thresholds are the identity."""
    rng = np.random.default_rng(31)
    Nz, H, W = 6, 16, 32
    x0 = rng.standard_normal((Nz, H, W)) + 1j * rng.standard_normal((Nz, H, W)) + 0.5
    g = rng.standard_normal((Nz, H, W)) + 1j * rng.standard_normal((Nz, H, W))
    p = params(orc, W, H, Nz, reg=0.0, tv=0.07, ntv=ntv, mode=6)
    t = 0.25
    xh = x0 - t * g
    exp = np.fft.ifftn(np.fft.fftn(xh) * (1 - _laplace_symbol(Nz, H, W) / 12.0) ** ntv)
    assert rel(orc.prox(p, x0, g, t, "f64"), exp) < 1e-13
    assert rel(orc.prox(p, x0, g, t, "f32"), exp) < 2e-6


def test_mode6_invariants(orc):
    rng = np.random.default_rng(32)
    Nz, H, W = 4, 16, 16
    x0 = rng.standard_normal((Nz, H, W)) + 1j * rng.standard_normal((Nz, H, W))
    g = rng.standard_normal((Nz, H, W)) + 1j * rng.standard_normal((Nz, H, W))
    # n_tv = 0: two half thresholds compose to one full threshold (mode 5's prox)
    p6 = params(orc, W, H, Nz, reg=0.8, tv=0.05, ntv=0, mode=6)
    p5 = params(orc, W, H, Nz, reg=0.8, tv=0.0, mode=5)
    np.testing.assert_allclose(orc.prox(p6, x0, g, 0.5, "f64"), orc.prox(p5, x0, g, 0.5, "f64"), atol=1e-14)
    # a constant volume passes the TV stage unchanged
    c = np.full((Nz, H, W), 0.3 - 0.2j)
    pc = params(orc, W, H, Nz, reg=0.0, tv=0.05, ntv=5, mode=6)
    np.testing.assert_allclose(orc.prox(pc, c, np.zeros_like(c), 0.5, "f64"), c, atol=1e-15)
    # z is periodic: rolling the input along z rolls the output
    a = orc.prox(pc, np.roll(x0, 1, axis=0), np.roll(g, 1, axis=0), 0.5, "f64")
    np.testing.assert_allclose(a, np.roll(orc.prox(pc, x0, g, 0.5, "f64"), 1, axis=0), atol=1e-14)


def test_mode6_prox_and_solve_match_restatement(orc):
    rng = np.random.default_rng(33)
    Nz, H, W = 4, 16, 32
    x0 = rng.standard_normal((Nz, H, W)) + 1j * rng.standard_normal((Nz, H, W))
    g = rng.standard_normal((Nz, H, W)) + 1j * rng.standard_normal((Nz, H, W))
    p = params(orc, W, H, Nz, reg=0.3, tv=0.02, ntv=5, mode=6)
    exp = R.prox_fl3d(x0, g, 0.25, 0.3, 0.02, 5)
    assert rel(orc.prox(p, x0, g, 0.25, "f64"), exp) < 1e-12
    assert rel(orc.prox(p, x0, g, 0.25, "f32"), exp) < 2e-6
    # whole solve in FP64 against the numpy solve
    holo = GOLD["A_holo"]
    Hh, Wh = holo.shape
    ps = params(orc, Wh, Hh, 4, its=6, reg=0.01, tv=0.01, ntv=3, mode=6)
    op = R.Operator(Wh, Hh, 4, 1000.0, 50.0, 0.633, 2.0)
    ref, _, _, log = R.solve(op, holo, 6, 0.01, 0.01, 3, mode=6)
    vol, _, _, st = orc.solve(ps, holo, None, "f64")
    assert rel(vol, ref) < 1e-9
    np.testing.assert_allclose(st["f"], log["f"], rtol=1e-9)
    assert list(st["restarts"]) == list(log["restarts"])


def test_tv3d_norm(orc):
    rng = np.random.default_rng(34)
    v = (rng.standard_normal((5, 16, 32)) + 1j * rng.standard_normal((5, 16, 32))).astype(np.complex64)
    v[1] = 0  # sparse planes as the solver produces
    assert orc.tv3d_f32(v) == pytest.approx(R.tv3d(v.astype(np.complex128)), rel=1e-6)
    # the z term is what separates it from the 2D norm
    assert orc.tv3d_f32(v) > orc.tv2d_f32(v) * 1.05


def test_unsupported_mode_is_invalid_argument(orc):
    p = params(orc, 16, 16, 2, mode=3)
    with pytest.raises(RuntimeError, match="status 7"):
        orc.solve(p, np.zeros((16, 16), np.complex64))
