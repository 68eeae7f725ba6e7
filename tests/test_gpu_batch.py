"""GPU tests of rihvr_reconstruct_batch: a frame sequence, warm starts from the
previous frame kept on the device, and the objective-target stop rule, each
checked against the single-frame path (itself checked against the oracle)."""
from __future__ import annotations

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def cfg(W, H, Nz, its=8, reg=0.01, tv=0.01, ntv=5, mode=7):
    return dict(width=W, height=H, num_planes=Nz, start_plane=1000.0, plane_stepsize=50.0, wavelength=0.633,
                resolution=2.0, num_inverse_iterations=its, regularization_param=reg,
                regularization_TV_param=tv, num_tv_iterations=ntv, mode=mode)


@pytest.fixture(scope="module")
def frames(scenes):
    c = cfg(64, 64, 6)
    return c, scenes.sequence(scenes.optics_of(c), 3, n_tracers=12, n_rods=1, seed=41, sigma=0.01)


def test_batch_plain_equals_single_frames(pkg, frames, gpu_available):
    c, holos = frames
    with pkg.SparseCompressiveHolo(pkg.make_params(**c)) as rec:
        vols = np.zeros((3, 6, 64, 64), np.complex64)
        fs = rec.reconstructBatch(holos, volumes=vols)
        for k in range(3):
            st = rec.inverseReconstruct(holos[k])
            np.testing.assert_array_equal(vols[k], rec.getVolume())
            assert fs[k]["iterations"] == c["num_inverse_iterations"]
            assert fs[k]["f_final"] == st["f_final"]
            assert fs[k]["nnz"] == st["nnz"]


def test_batch_warm_from_previous(pkg, frames, gpu_available):
    c, holos = frames
    with pkg.SparseCompressiveHolo(pkg.make_params(**c)) as rec:
        vols = np.zeros((3, 6, 64, 64), np.complex64)
        rec.reconstructBatch(holos, warm_from_previous=True, volumes=vols)
        last = rec.getVolume()
        np.testing.assert_array_equal(last, vols[2])
        rec.inverseReconstruct(holos[0])
        np.testing.assert_array_equal(rec.getVolume(), vols[0])
        for k in (1, 2):
            rec.inverseReconstruct(holos[k], warm_start=vols[k - 1])
            np.testing.assert_array_equal(rec.getVolume(), vols[k])


def test_batch_objective_target(pkg, frames, gpu_available):
    c, holos = frames
    c20 = dict(c, num_inverse_iterations=20)
    with pkg.SparseCompressiveHolo(pkg.make_params(**c20)) as rec:
        st = rec.inverseReconstruct(holos[0])
        f = np.asarray(st["f"])
        # first iteration where f drops clearly below its predecessor, target in between
        k = next(i for i in range(3, 20) if f[i] < f[i - 1] * (1 - 1e-4))
        target = 0.5 * (f[k] + f[k - 1])
        stop = next(i for i in range(20) if f[i] <= target)
        fs = rec.reconstructBatch(holos[:1], f_target=[target], max_iterations=20)
        got = rec.getVolume()
        assert fs[0]["iterations"] == stop + 1 and fs[0]["reached"]
        assert fs[0]["f_final"] <= target
        rec.begin(holos[0])
        rec.iterate(stop + 1)
        rec.finish()
        np.testing.assert_array_equal(got, rec.getVolume())
        # an unreachable target runs to the cap and says so
        fs = rec.reconstructBatch(holos[:1], f_target=[1e-30], max_iterations=4)
        assert fs[0]["iterations"] == 4 and not fs[0]["reached"]


def test_batch_mode6_and_device_arrays(pkg, frames, gpu_available):
    import torch
    c, holos = frames
    c6 = dict(c, mode=6, num_inverse_iterations=4)
    with pkg.SparseCompressiveHolo(pkg.make_params(**c6)) as rec:
        hd = torch.from_numpy(holos).cuda()
        vd = torch.zeros((3, 6, 64, 64), dtype=torch.complex64, device="cuda")
        rec.reconstructBatch(hd, warm_from_previous=True, volumes=vd)
        torch.cuda.synchronize()
        vh = np.zeros((3, 6, 64, 64), np.complex64)
        rec.reconstructBatch(holos, warm_from_previous=True, volumes=vh)
    np.testing.assert_array_equal(vd.cpu().numpy(), vh)


def test_backtracks_with_work_queued_ahead(pkg, frames, gpu_available, monkeypatch):
    """Without an objective target the solver queues each accepted iteration's work before it
    reads the scalars; a backtrack discards that work.  Forced backtracks (iterations 1 and 4)
    give the same volume as the path that waits for the scalars (an unreachable target)."""
    c, holos = frames
    c12 = dict(c, num_inverse_iterations=12)
    monkeypatch.setenv("RIHVR_TEST_BACKTRACKS", str((1 << 1) | (1 << 4)))
    with pkg.SparseCompressiveHolo(pkg.make_params(**c12)) as rec:
        st = rec.inverseReconstruct(holos[0])
        ahead = rec.getVolume()
        fs = rec.reconstructBatch(holos[:1], f_target=[1e-300], max_iterations=12)
        waited = rec.getVolume()
    assert st["total_backtracks"] == 2 and fs[0]["total_backtracks"] == 2
    assert list(st["backtracks"][:6]) == [0, 1, 0, 0, 1, 0]
    np.testing.assert_array_equal(ahead, waited)
    monkeypatch.delenv("RIHVR_TEST_BACKTRACKS")
    with pkg.SparseCompressiveHolo(pkg.make_params(**c12)) as rec:
        plain = rec.inverseReconstruct(holos[0])
    assert plain["total_backtracks"] == 0


def test_scene_propagation_on_the_gpu_matches_the_host(scenes, gpu_available):
    """The batched workloads' holograms come from propagate_sum_torch (complex128 on the GPU):
    the same sum as the host's propagate_sum, to 1e-12 relative, and the same seeded holograms."""
    c = cfg(96, 80, 10)
    o = scenes.optics_of(c)
    rng = np.random.default_rng(5)
    masks = (rng.random((10, 80, 96)) < 0.01).astype(np.float32)
    masks[3] = 0  # an empty plane is skipped by both
    a = scenes.propagate_sum(masks, o)
    b = scenes.propagate_sum_torch(masks, o)
    assert np.linalg.norm(a - b) <= 1e-12 * np.linalg.norm(a)
    h0, _ = scenes.tracers_rods_scene(o, 30, 2, 17, 0.02)
    h1, _ = scenes.tracers_rods_scene(o, 30, 2, 17, 0.02, propagate=scenes.propagate_sum_torch)
    assert np.max(np.abs(h0 - h1)) <= 1e-6
