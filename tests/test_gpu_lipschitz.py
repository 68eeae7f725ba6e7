"""rihvr_estimate_lipschitz against what the operator fixes: A A^T = Nz I on real holograms, so
A^T A is Nz times the projection onto a subspace of real dimension N = W H inside the volume's
2 Nz N.  One secant step (the reference's estimate, ||A^T A d|| / ||d||) of a random d is then
Nz ||P d|| / ||d||, about Nz / sqrt(2 Nz) = sqrt(Nz / 2); from the second step the power iteration
sits on the range and returns Nz.  A solve after the call is unchanged."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def cfg(W, H, Nz):
    return dict(width=W, height=H, num_planes=Nz, start_plane=1000.0, plane_stepsize=50.0, wavelength=0.633,
                resolution=2.0, num_inverse_iterations=6, regularization_param=0.01,
                regularization_TV_param=0.01, num_tv_iterations=5, mode=7)


@pytest.mark.parametrize("W,H,Nz", [(128, 128, 16), (272, 208, 9), (256, 256, 64)])
def test_lipschitz_matches_the_operator(pkg, scenes, gpu_available, W, H, Nz):
    c = cfg(W, H, Nz)
    holo, _ = scenes.disks_scene(scenes.optics_of(c), 6, 2.0, 4.0, 5, 0.01)
    with pkg.SparseCompressiveHolo(pkg.make_params(**c)) as rec:
        st0 = rec.inverseReconstruct(holo)
        v0 = rec.getVolume()
        l1 = rec.estimateLipschitz(seed=7, iterations=1)
        l3 = rec.estimateLipschitz(seed=7, iterations=3)
        l3b = rec.estimateLipschitz(seed=8, iterations=3)
        st1 = rec.inverseReconstruct(holo)
        v1 = rec.getVolume()
    print(f"{W}x{H}x{Nz}: secant {l1:.5g} (sqrt(Nz/2) = {np.sqrt(Nz / 2):.5g}), power {l3:.7g} / {l3b:.7g}",
          flush=True)
    assert abs(l1 - np.sqrt(Nz / 2)) <= 0.05 * np.sqrt(Nz / 2)
    assert abs(l3 - Nz) <= 1e-4 * Nz and abs(l3b - Nz) <= 1e-4 * Nz
    assert np.array_equal(v0, v1) and st0["f_final"] == st1["f_final"]
