"""GPU parity: librihvr (HIP, gfx950) against the CPU oracle's FP32 restatement.

Tolerances (relative L2 over the complex arrays):
  operators A, A^T and one prox call : 2e-5 (FP32 FFT / stencil rounding only)
  full solves                          : 1e-4 (the parity bar for a solve), identical
                                         thresholded support except voxels whose
                                         magnitude is within 1e-4 of the threshold,
                                         identical backtrack/restart decisions.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def rel(a, b):
    return float(np.linalg.norm(np.asarray(a) - np.asarray(b)) / max(np.linalg.norm(np.asarray(b)), 1e-30))


def cfg(W, H, Nz, its=5, reg=0.01, tv=0.01, ntv=5, z0=1000.0, dz=50.0, lam=0.633, dx=2.0, mode=7):
    return dict(width=W, height=H, num_planes=Nz, start_plane=z0, plane_stepsize=dz, wavelength=lam,
                resolution=dx, num_inverse_iterations=its, regularization_param=reg,
                regularization_TV_param=tv, num_tv_iterations=ntv, mode=mode)


def rand_vol(rng, Nz, H, W, sparsity=None):
    v = (rng.standard_normal((Nz, H, W)) + 1j * rng.standard_normal((Nz, H, W))).astype(np.complex64)
    if sparsity is not None:
        v *= (rng.random((Nz, H, W)) < sparsity)
    return v


# (512, 1024, 3): 1024-point column transforms of a plane narrower than its height
SHAPES = [(32, 32, 4), (64, 48, 3), (256, 256, 6), (448, 448, 2), (640, 576, 2), (1024, 1024, 2),
          (1024, 1024, 5), (512, 1024, 3)]


@pytest.mark.parametrize("W,H,Nz", SHAPES)
def test_forward_matches_oracle(pkg, orc, gpu_available, W, H, Nz):
    c = cfg(W, H, Nz, z0=2500.0, dz=20.0, lam=0.632, dx=1.1)
    rng = np.random.default_rng(W * 7 + H)
    x = rand_vol(rng, Nz, H, W, sparsity=0.05)
    with pkg.SparseCompressiveHolo(pkg.make_params(**c)) as rec:
        est = rec.forward(x)
    ref = orc.forward(orc.make_params(**c), x, "f32")
    assert np.all(est.imag == 0)
    assert rel(est, ref) < 2e-5, rel(est, ref)


@pytest.mark.parametrize("W,H,Nz", SHAPES)
def test_adjoint_matches_oracle(pkg, orc, gpu_available, W, H, Nz):
    c = cfg(W, H, Nz, z0=2500.0, dz=20.0, lam=0.632, dx=1.1)
    rng = np.random.default_rng(W * 11 + H)
    r = rng.standard_normal((H, W)).astype(np.complex64)
    with pkg.SparseCompressiveHolo(pkg.make_params(**c)) as rec:
        g = rec.adjoint(r)
    ref = orc.adjoint(orc.make_params(**c), r, "f32")
    assert rel(g, ref) < 2e-5, rel(g, ref)


def test_adjoint_identity_gpu(pkg, gpu_available):
    """<A x, r> = sum Re(conj(x) A^T r) for the GPU operator pair (K1)."""
    c = cfg(256, 256, 8, z0=2500.0, dz=20.0, lam=0.632, dx=1.1)
    rng = np.random.default_rng(5)
    x = rand_vol(rng, 8, 256, 256)
    r = rng.standard_normal((256, 256)).astype(np.complex64)
    with pkg.SparseCompressiveHolo(pkg.make_params(**c)) as rec:
        ax = rec.forward(x)
        atr = rec.adjoint(r)
    lhs = float(np.sum(ax.real.astype(np.float64) * r.real))
    rhs = float(np.sum((np.conj(x.astype(np.complex128)) * atr).real))
    assert abs(lhs - rhs) <= 1e-5 * (abs(lhs) + 1e-12) + 1e-3, (lhs, rhs)


@pytest.mark.parametrize("ntv", [0, 1, 3, 5, 8])
@pytest.mark.parametrize("W,H", [(32, 32), (64, 48), (256, 128), (1024, 64)])
def test_prox_fl2d_matches_oracle(pkg, orc, gpu_available, ntv, W, H):
    Nz = 3
    c = cfg(W, H, Nz, reg=0.05, tv=0.02, ntv=ntv)
    rng = np.random.default_rng(ntv * 100 + W + H)
    x0 = rand_vol(rng, Nz, H, W, sparsity=0.2) * np.float32(0.5)
    g = rand_vol(rng, Nz, H, W) * np.float32(2.0)
    t = 1.0 / Nz
    with pkg.SparseCompressiveHolo(pkg.make_params(**c)) as rec:
        y = rec.prox(x0, g, t)
    ref = orc.prox(orc.make_params(**c), x0, g, t, "f32")
    assert rel(y, ref) < 2e-5, rel(y, ref)
    # thresholded support: only voxels at the threshold may differ
    thr = t * 0.05
    nz_g = (y.real != 0) | (y.imag != 0)
    nz_r = (ref.real != 0) | (ref.imag != 0)
    diff = nz_g != nz_r
    assert diff.sum() <= max(2, int(1e-4 * y.size)), diff.sum()
    del thr


def test_prox_l1_matches_oracle(pkg, orc, gpu_available):
    W, H, Nz = 64, 64, 4
    c = cfg(W, H, Nz, reg=0.1, tv=0.0, mode=5)
    rng = np.random.default_rng(3)
    x0 = rand_vol(rng, Nz, H, W, sparsity=0.3)
    g = rand_vol(rng, Nz, H, W)
    with pkg.SparseCompressiveHolo(pkg.make_params(**c)) as rec:
        y = rec.prox(x0, g, 0.25)
    ref = orc.prox(orc.make_params(**c), x0, g, 0.25, "f32")
    assert rel(y, ref) < 1e-6, rel(y, ref)


def _solve_both(pkg, orc, c, holo, warm=None):
    with pkg.SparseCompressiveHolo(pkg.make_params(**c)) as rec:
        st = rec.inverseReconstruct(holo, warm_start=warm)
        vol = rec.getVolume()
        est, res = rec.getEstimate()
    ref, rest, rres, rst = orc.solve(orc.make_params(**c), holo, warm, "f32")
    return vol, st, ref, rst, (est, res, rest, rres)


def _support_mismatch(a, b, thr_mag):
    na = (a.real != 0) | (a.imag != 0)
    nb = (b.real != 0) | (b.imag != 0)
    return int((na != nb).sum())


def test_solve_small_matches_oracle(pkg, orc, scenes, gpu_available):
    c = cfg(64, 64, 8, its=10)
    holo, _ = scenes.disks_scene(scenes.optics_of(c), 6, 2.0, 4.0, 11, 0.01)
    vol, st, ref, rst, (est, res, rest, rres) = _solve_both(pkg, orc, c, holo)
    assert rel(vol, ref) < 1e-4, rel(vol, ref)
    assert list(st["backtracks"]) == list(rst["backtracks"])
    assert list(st["restarts"]) == list(rst["restarts"])
    assert _support_mismatch(vol, ref, 0) <= max(2, int(1e-4 * vol.size))
    assert abs(st["f_final"] - rst["f_final"]) <= 1e-4 * abs(rst["f_final"])
    np.testing.assert_allclose(st["f"], rst["f"], rtol=1e-4)
    assert rel(est, rest) < 1e-4 and rel(res, rres) < 1e-3


def test_solve_c1_matches_oracle(pkg, orc, scenes, gpu_available):
    """The 256^2 x 32 workload (50 FISTA x 5 TV iterations) at full size."""
    c = dict(scenes.C1, mode=7)
    holo = scenes.c1_hologram()
    vol, st, ref, rst, _ = _solve_both(pkg, orc, c, holo)
    r = rel(vol, ref)
    print(f"C1 rel-L2 vs FP32 oracle: {r:.3e}; support mismatches {_support_mismatch(vol, ref, 0)} "
          f"of {vol.size}; nnz {st['nnz']} vs {rst['nnz']}; backtracks {st['total_backtracks']} vs "
          f"{rst['total_backtracks']}")
    assert r < 1e-4, r
    assert list(st["restarts"]) == list(rst["restarts"])
    assert list(st["backtracks"]) == list(rst["backtracks"])
    assert _support_mismatch(vol, ref, 0) <= max(2, int(1e-5 * vol.size))


def test_solve_mode5_and_warm_start(pkg, orc, scenes, gpu_available):
    c = cfg(64, 64, 6, its=6, tv=0.0, reg=0.02)  # TV weight 0 -> mode 5 (:192-195)
    holo, _ = scenes.disks_scene(scenes.optics_of(c), 5, 2.0, 4.0, 12, 0.01)
    rng = np.random.default_rng(1)
    warm = rand_vol(rng, 6, 64, 64, sparsity=0.01) * np.float32(0.1)
    vol, st, ref, rst, _ = _solve_both(pkg, orc, c, holo, warm)
    assert rel(vol, ref) < 1e-4, rel(vol, ref)
    assert list(st["restarts"]) == list(rst["restarts"])


@pytest.mark.parametrize("ntv", [0, 1, 5])
@pytest.mark.parametrize("W,H,Nz", [(32, 32, 3), (64, 48, 5), (256, 128, 4)])
def test_prox_fl3d_matches_oracle(pkg, orc, gpu_available, ntv, W, H, Nz):
    """Mode 6 (proximalFusedLasso, :3422-3475): half thresholds around the 3D,
    z-periodic TV stage."""
    c = cfg(W, H, Nz, reg=0.05, tv=0.02, ntv=ntv, mode=6)
    rng = np.random.default_rng(ntv * 1000 + W + H + Nz)
    x0 = rand_vol(rng, Nz, H, W, sparsity=0.2) * np.float32(0.5)
    g = rand_vol(rng, Nz, H, W) * np.float32(2.0)
    t = 1.0 / Nz
    with pkg.SparseCompressiveHolo(pkg.make_params(**c)) as rec:
        y = rec.prox(x0, g, t)
    ref = orc.prox(orc.make_params(**c), x0, g, t, "f32")
    assert rel(y, ref) < 2e-5, rel(y, ref)
    diff = ((y.real != 0) | (y.imag != 0)) != ((ref.real != 0) | (ref.imag != 0))
    assert diff.sum() <= max(2, int(1e-4 * y.size)), diff.sum()


def test_solve_mode6_matches_oracle(pkg, orc, scenes, gpu_available):
    c = cfg(64, 64, 8, its=8, reg=0.01, tv=0.01, ntv=3, mode=6)
    holo, _ = scenes.disks_scene(scenes.optics_of(c), 6, 2.0, 4.0, 14, 0.01)
    vol, st, ref, rst, (est, res, rest, rres) = _solve_both(pkg, orc, c, holo)
    assert rel(vol, ref) < 1e-4, rel(vol, ref)
    assert list(st["restarts"]) == list(rst["restarts"])
    assert list(st["backtracks"]) == list(rst["backtracks"])
    assert _support_mismatch(vol, ref, 0) <= max(2, int(1e-4 * vol.size))
    # the objective carries the 3D TV norm (calcTVNorm)
    np.testing.assert_allclose(st["f"], rst["f"], rtol=1e-4)
    assert abs(st["f_final"] - rst["f_final"]) <= 1e-4 * abs(rst["f_final"])


def test_pipelined_iteration_matches_one_stream_and_oracle(pkg, orc, scenes, gpu_available, monkeypatch):
    """With plane groups the first pass of an iteration runs A^T of group g+1 on a
    second stream while the prox and A of group g run on the main one: same kernels
    and partial layout as one stream, so the volumes agree bit for bit."""
    c = cfg(128, 128, 96, its=6)
    holo, _ = scenes.disks_scene(scenes.optics_of(c), 12, 2.0, 4.0, 15, 0.01)
    monkeypatch.setenv("RIHVR_PIPE_PLANES", "32")  # three groups on two streams
    with pkg.SparseCompressiveHolo(pkg.make_params(**c)) as rec:
        st = rec.inverseReconstruct(holo)
        a = rec.getVolume()
    monkeypatch.setenv("RIHVR_PIPE_PLANES", "0")  # one group: no second stream (the default)
    with pkg.SparseCompressiveHolo(pkg.make_params(**c)) as rec:
        st1 = rec.inverseReconstruct(holo)
        b = rec.getVolume()
    np.testing.assert_array_equal(a, b)
    np.testing.assert_array_equal(st["f"], st1["f"])
    ref, _, _, rst = orc.solve(orc.make_params(**c), holo, None, "f32")
    assert rel(a, ref) < 1e-4, rel(a, ref)
    assert list(st["restarts"]) == list(rst["restarts"])


@pytest.mark.parametrize("W,H,Nz,mode", [(448, 448, 6, 7), (640, 576, 4, 7), (832, 832, 2, 7), (576, 576, 4, 6),
                                          (256, 256, 16, 5)])
def test_solve_other_sizes_match_oracle(pkg, orc, scenes, gpu_available, W, H, Nz, mode):
    """Zero-padded sizes of the reference's CLI (448 = 384 + 2*32, 576, 832 = 768 + 64,
    non-square 640x576) through radix-7/9/13/5 FFT plans, and modes 6 and 5."""
    tv = 0.0 if mode == 5 else 0.01
    c = cfg(W, H, Nz, its=6, reg=0.01, tv=tv, ntv=3, mode=mode)
    holo, _ = scenes.disks_scene(scenes.optics_of(c), 8, 2.0, 5.0, 16 + W + Nz, 0.01)
    vol, st, ref, rst, (est, res, rest, rres) = _solve_both(pkg, orc, c, holo)
    assert rel(vol, ref) < 1e-4, rel(vol, ref)
    assert list(st["restarts"]) == list(rst["restarts"])
    assert _support_mismatch(vol, ref, 0) <= max(2, int(1e-4 * vol.size))
    assert rel(est, rest) < 1e-4


@pytest.mark.parametrize("ntv", [1, 8, 10])
def test_solve_mode7_tv_iterations_match_oracle(pkg, orc, scenes, gpu_available, ntv):
    """Mode 7 with 1, 8 and 10 TV iterations on a 128x96 plane: the prox's halo (64 - 2 n_tv
    output columns per wave, wrapped periodically at the plane's edges), several row chunks
    per plane, and the occupancy maps of a sparse solve (rows of y_{k-1} whose lines are all
    empty skipped, the map bits that stand for y_{k-1} in the reduction-skip test)."""
    c = cfg(128, 96, 6, its=8, reg=0.1, tv=0.05, ntv=ntv)  # 2-3 % of the voxels non-zero
    holo, _ = scenes.disks_scene(scenes.optics_of(c), 6, 2.0, 4.0, 21 + ntv, 0.01)
    vol, st, ref, rst, (est, res, rest, rres) = _solve_both(pkg, orc, c, holo)
    assert rel(vol, ref) < 1e-4, rel(vol, ref)
    assert list(st["restarts"]) == list(rst["restarts"])
    assert list(st["backtracks"]) == list(rst["backtracks"])
    assert _support_mismatch(vol, ref, 0) <= max(2, int(1e-4 * vol.size))
    np.testing.assert_allclose(st["f"], rst["f"], rtol=1e-4)
    assert 0 < st["nnz"] < vol.size // 10  # sparse: the maps are in use


def test_solve_complex_hologram_matches_oracle(pkg, orc, scenes, gpu_available):
    """A hologram with an imaginary part (the API takes complex64): the estimate stays real,
    so the residual's imaginary part is -Im(b) and its spectrum is not Hermitian-symmetric;
    the solve, its objective and the returned estimate and residual follow the oracle."""
    c = cfg(64, 64, 6, its=8, reg=0.05, tv=0.02)
    holo, _ = scenes.disks_scene(scenes.optics_of(c), 6, 2.0, 4.0, 31, 0.01)
    rng = np.random.default_rng(32)
    holo = (holo + 1j * (0.1 * rng.standard_normal(holo.shape))).astype(np.complex64)
    vol, st, ref, rst, (est, res, rest, rres) = _solve_both(pkg, orc, c, holo)
    assert rel(vol, ref) < 1e-4, rel(vol, ref)
    assert list(st["restarts"]) == list(rst["restarts"])
    assert list(st["backtracks"]) == list(rst["backtracks"])
    np.testing.assert_allclose(st["f"], rst["f"], rtol=1e-4)
    assert np.all(est.imag == 0)
    assert rel(est, rest) < 1e-4 and rel(res, rres) < 1e-3
    np.testing.assert_allclose(res.imag, -holo.imag, rtol=0, atol=1e-6)


def test_device_pointers_match_host(pkg, scenes, gpu_available):
    import torch
    c = cfg(64, 64, 4, its=3)
    holo, _ = scenes.disks_scene(scenes.optics_of(c), 4, 2.0, 4.0, 13, 0.01)
    with pkg.SparseCompressiveHolo(pkg.make_params(**c)) as rec:
        rec.inverseReconstruct(holo)
        a = rec.getVolume()
        hd = torch.from_numpy(holo).cuda()
        rec.inverseReconstruct(hd)
        out = torch.empty((4, 64, 64), dtype=torch.complex64, device="cuda")
        rec.getVolume(out)
        torch.cuda.synchronize()
    np.testing.assert_array_equal(a, out.cpu().numpy())


def test_error_codes(pkg, gpu_available):
    with pytest.raises(pkg.RihvrError) as e:
        pkg.SparseCompressiveHolo(pkg.make_params(**cfg(40, 32, 2)))
    assert e.value.status == 7
    with pytest.raises(pkg.RihvrError) as e:
        pkg.SparseCompressiveHolo(pkg.make_params(**cfg(32, 32, 2, mode=3)))
    assert e.value.status == 7  # modes other than 5, 6, 7: HOLO_ERROR_INVALID_ARGUMENT (:178-189)
    with pytest.raises(pkg.RihvrError) as e:
        pkg.SparseCompressiveHolo(pkg.make_params(**cfg(4112, 32, 2)))  # rows beyond the 4096 FFT limit
    assert e.value.status == 7
    with pkg.SparseCompressiveHolo(pkg.make_params(**cfg(32, 32, 2))) as rec:
        bad = np.zeros((32, 32), np.complex64)
        bad[3, 4] = np.nan
        with pytest.raises(pkg.RihvrError) as e:
            rec.inverseReconstruct(bad)
        assert e.value.status == 10
        with pytest.raises(pkg.RihvrError) as e:
            rec.getPlane(0)
        assert e.value.status == 4


def test_stepwise_equals_reconstruct(pkg, scenes, gpu_available):
    """begin + iterate(a) + iterate(b) + finish == reconstruct(a + b), bit for bit."""
    c = cfg(64, 64, 6, its=7)
    holo, _ = scenes.disks_scene(scenes.optics_of(c), 5, 2.0, 4.0, 21, 0.01)
    with pkg.SparseCompressiveHolo(pkg.make_params(**c)) as rec:
        st = rec.inverseReconstruct(holo)
        a = rec.getVolume()
        f0 = rec.begin(holo)
        s1 = rec.iterate(3, with_objective=True)
        rec.set_timing(True)
        s2 = rec.iterate(4, with_objective=True)
        tm = rec.get_timing()
        rec.set_timing(False)
        fin = rec.finish()
        b = rec.getVolume()
    np.testing.assert_array_equal(a, b)
    assert f0 == st["f_initial"]
    assert fin["f_final"] == st["f_final"]
    # the last f of an iterate call comes from a separate reduction pass (other order)
    np.testing.assert_allclose(np.concatenate([s1["f"], s2["f"]]), st["f"], rtol=1e-6)
    assert list(np.concatenate([s1["restarts"], s2["restarts"]])) == list(st["restarts"])
    assert tm["prox"][0] >= 4 and tm["prox"][1] > 0
    assert tm["fft_rows_fwdacc"][0] >= 4


def test_iterate_without_begin_is_invalid_state(pkg, gpu_available):
    with pkg.SparseCompressiveHolo(pkg.make_params(**cfg(32, 32, 2))) as rec:
        with pytest.raises(pkg.RihvrError) as e:
            rec.iterate(1)
        assert e.value.status == 4
