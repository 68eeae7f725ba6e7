"""GPU parity beyond the fixed-length FFT kernels and the fused prox's TV-iteration limit:
hologram sizes the reference accepts (it plans any size, cufftPlan2d(H, W),
sparse_compressive_holo.cu:98; its README's Dandelion case pads 1024 by 64 on each side to
1152, README.md:46-57) run on run-time FFT plans, and mode 7 with more TV iterations than the
fused kernel holds (the reference loops any n_tv, :3693-3702) runs the multi-pass prox.  Each
against the FP32 oracle, whose FFT handles any length (radices 4, 2 and any prime):
  - A and A^T at 1152^2, 1280 x 1024, 2048^2, 272 x 208 (factors 17 and 13), 2032 x 32 (a
    factor 127, one direct-DFT pass), 1040 x 64 (W > 1024: 32-voxel occupancy lines): 2e-5;
  - solves at 1152^2, 2048^2, 272 x 208 and 1040 x 64 with identical decisions: 1e-4;
  - the prox and solves with n_tv = 12 and 20 on fixed- and run-time-plan sizes: 2e-5 / 1e-4;
  - a 1024^2 x 1000-plane solve (the README's Rotating Rods depth; volumes above 2^32 bytes)
    against the oracle for 2 iterations.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def rel(a, b):
    return float(np.linalg.norm(np.asarray(a) - np.asarray(b)) / max(np.linalg.norm(np.asarray(b)), 1e-30))


def cfg(W, H, Nz, its=5, reg=0.01, tv=0.01, ntv=5, z0=2500.0, dz=20.0, lam=0.632, dx=1.1, mode=7):
    return dict(width=W, height=H, num_planes=Nz, start_plane=z0, plane_stepsize=dz, wavelength=lam,
                resolution=dx, num_inverse_iterations=its, regularization_param=reg,
                regularization_TV_param=tv, num_tv_iterations=ntv, mode=mode)


def rand_vol(rng, Nz, H, W, sparsity=None):
    v = (rng.standard_normal((Nz, H, W)) + 1j * rng.standard_normal((Nz, H, W))).astype(np.complex64)
    if sparsity is not None:
        v *= (rng.random((Nz, H, W)) < sparsity)
    return v


OP_SHAPES = [(1152, 1152, 2), (1280, 1024, 2), (2048, 2048, 2), (272, 208, 3), (2032, 32, 2), (1040, 64, 3)]


@pytest.mark.parametrize("W,H,Nz", OP_SHAPES)
def test_operators_at_any_size_match_oracle(pkg, orc, gpu_available, W, H, Nz):
    print(f"operators {W}x{H}x{Nz}", flush=True)
    c = cfg(W, H, Nz)
    rng = np.random.default_rng(W * 3 + H)
    x = rand_vol(rng, Nz, H, W, sparsity=0.05)
    r = rng.standard_normal((H, W)).astype(np.complex64)
    with pkg.SparseCompressiveHolo(pkg.make_params(**c)) as rec:
        est = rec.forward(x)
        g = rec.adjoint(r)
    p = orc.make_params(**c)
    assert np.all(est.imag == 0)
    assert rel(est, orc.forward(p, x, "f32")) < 2e-5
    assert rel(g, orc.adjoint(p, r, "f32")) < 2e-5


def _solve_both(pkg, orc, c, holo):
    with pkg.SparseCompressiveHolo(pkg.make_params(**c)) as rec:
        st = rec.inverseReconstruct(holo)
        vol = rec.getVolume()
        est, res = rec.getEstimate()
    ref, rest, rres, rst = orc.solve(orc.make_params(**c), holo, None, "f32")
    return vol, st, ref, rst, est, rest


def _check_solve(vol, st, ref, rst, est, rest, tol=1e-4):
    assert rel(vol, ref) < tol, rel(vol, ref)
    assert list(st["backtracks"]) == list(rst["backtracks"])
    assert list(st["restarts"]) == list(rst["restarts"])
    np.testing.assert_allclose(st["f"], rst["f"], rtol=1e-4)
    assert rel(est, rest) < tol


@pytest.mark.parametrize("W,H,Nz,its", [(1152, 1152, 4, 6), (2048, 2048, 4, 4), (272, 208, 6, 12), (1040, 64, 4, 12)])
def test_solve_at_any_size_matches_oracle(pkg, orc, scenes, gpu_available, W, H, Nz, its):
    print(f"solve {W}x{H}x{Nz}, {its} iterations", flush=True)
    c = cfg(W, H, Nz, its=its, reg=0.05, tv=0.02)
    holo, _ = scenes.tracers_rods_scene(scenes.optics_of(c), max(8, W * H // 20000), 2, W + H, 0.02)
    _check_solve(*_solve_both(pkg, orc, c, holo))


@pytest.mark.parametrize("ntv", [11, 12, 20])
@pytest.mark.parametrize("W,H", [(64, 48), (272, 208)])
def test_prox_many_tv_iterations_matches_oracle(pkg, orc, gpu_available, ntv, W, H):
    Nz = 3
    c = cfg(W, H, Nz, reg=0.05, tv=0.02, ntv=ntv, z0=1000.0, dz=50.0, lam=0.633, dx=2.0)
    rng = np.random.default_rng(ntv * 100 + W + H)
    x0 = rand_vol(rng, Nz, H, W, sparsity=0.2) * np.float32(0.5)
    g = rand_vol(rng, Nz, H, W) * np.float32(2.0)
    t = 1.0 / Nz
    with pkg.SparseCompressiveHolo(pkg.make_params(**c)) as rec:
        y = rec.prox(x0, g, t)
    ref = orc.prox(orc.make_params(**c), x0, g, t, "f32")
    assert rel(y, ref) < 2e-5, rel(y, ref)


@pytest.mark.parametrize("W,H,Nz,ntv", [(128, 96, 6, 12), (272, 208, 4, 20)])
def test_solve_many_tv_iterations_matches_oracle(pkg, orc, scenes, gpu_available, W, H, Nz, ntv):
    print(f"solve n_tv={ntv} {W}x{H}x{Nz}", flush=True)
    c = cfg(W, H, Nz, its=12, reg=0.01, tv=0.01, ntv=ntv, z0=1000.0, dz=50.0, lam=0.633, dx=2.0)
    holo, _ = scenes.disks_scene(scenes.optics_of(c), 6, 2.0, 4.0, 21 + ntv, 0.01)
    _check_solve(*_solve_both(pkg, orc, c, holo))


@pytest.mark.timeout(900)
def test_thousand_plane_solve_matches_oracle(pkg, orc, scenes, gpu_available):
    """1024^2 x 1000 planes (each volume 8 GiB), 2 FISTA x 5 TV iterations."""
    import torch
    c = cfg(1024, 1024, 1000, its=2, reg=0.15, tv=0.1, z0=2500.0, dz=5.0)
    print("solve 1024x1024x1000, 2 iterations", flush=True)
    holo = scenes.c2_hologram()
    vol = torch.empty((1000, 1024, 1024), dtype=torch.complex64, device="cuda")
    with pkg.SparseCompressiveHolo(pkg.make_params(**c)) as rec:
        st = rec.inverseReconstruct(torch.from_numpy(holo).cuda())
        rec.getVolume(out=vol)
    torch.cuda.synchronize()
    ref, _, _, rst = orc.solve(orc.make_params(**c), holo, None, "f32")
    # compare plane by plane on the host (the difference and the reference norms in float64)
    num = den = 0.0
    for z0 in range(0, 1000, 100):
        a = vol[z0:z0 + 100].cpu().numpy().astype(np.complex128)
        b = ref[z0:z0 + 100].astype(np.complex128)
        num += float(np.sum(np.abs(a - b) ** 2))
        den += float(np.sum(np.abs(b) ** 2))
    r = (num / max(den, 1e-300)) ** 0.5
    print(f"1024^2 x 1000: rel-L2 {r:.3e}, nnz {st['nnz']} vs {rst['nnz']}", flush=True)
    assert r < 1e-4
    np.testing.assert_allclose(st["f"], rst["f"], rtol=1e-4)
    del vol
    torch.cuda.empty_cache()
