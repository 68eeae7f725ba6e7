"""GPU tests of rihvr_prepare_holograms (the input contract on the device for a batch of
frames) against rihvr_prepare_hologram, the host function the CPU tests pin against a
numpy restatement: every frame bit for bit, for 8-bit, 16-bit and float frames, mean
normalisation and background removal, the centre/size and rectangle crops (inside the
image and past its borders), zero padding, odd frame and output sizes, host and device
buffers; and the errors."""
from __future__ import annotations

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _fp(pkg, pad=0, center=(-1, -1), size=-1, roi=(0, 0, 0, 0), remove_bg=False):
    fp = pkg.FileParams()
    fp.solver.zero_padding = pad
    fp.center_x, fp.center_y = center
    fp.roi_size = size
    fp.roi_x, fp.roi_y, fp.roi_w, fp.roi_h = roi
    fp.remove_background = 1 if remove_bg else 0
    return fp


def frames_of(dtype, n, h, w, seed):
    rng = np.random.default_rng(seed)
    if dtype == np.float32:
        return (rng.random((n, h, w)) * 300).astype(np.float32)
    hi = 256 if dtype == np.uint8 else 65536
    return rng.integers(0, hi, size=(n, h, w)).astype(dtype)


CASES = [
    dict(),                                        # whole square image, no padding
    dict(pad=16),
    dict(center=(20, 30), size=32, pad=4),
    dict(center=(2, 3), size=24),
    dict(roi=(-5, -3, 40, 40), pad=8),             # past the top-left corner: zero fill
    dict(roi=(50, 40, 32, 32)),                    # past the bottom-right corner
]


@pytest.mark.parametrize("dtype", [np.uint8, np.uint16, np.float32])
@pytest.mark.parametrize("case", range(len(CASES)))
@pytest.mark.parametrize("remove_bg", [False, True])
def test_device_equals_host(pkg, gpu_available, dtype, case, remove_bg):
    kw = CASES[case]
    fp = _fp(pkg, remove_bg=remove_bg, **kw)
    fr = frames_of(dtype, 3, 64, 64, seed=case)
    bg = (np.random.default_rng(99).random((64, 64)) * 200).astype(np.float32) if remove_bg else None
    if bg is not None:
        bg[0, :5] = 0.0  # sqrt(B) below 0.1 takes the floor
    got = pkg.prepare_holograms(fp, fr, bg)
    for k in range(3):
        exp = pkg.prepare_hologram(fp, fr[k].astype(np.float32), bg)
        np.testing.assert_array_equal(got[k].view(np.uint32), exp.view(np.uint32))


@pytest.mark.parametrize("dtype", [np.uint8, np.uint16, np.float32])
@pytest.mark.parametrize("remove_bg", [False, True])
def test_odd_sizes_equal_host(pkg, gpu_available, dtype, remove_bg):
    # 37 x 45 frames: a frame size that is no multiple of 16 bytes (the mean pass's unaligned
    # frames and its tail), and an odd output width 35 (one pixel per thread in the write pass)
    fp = _fp(pkg, remove_bg=remove_bg, center=(20, 18), size=33, pad=1)
    fr = frames_of(dtype, 5, 37, 45, seed=11)
    bg = (np.random.default_rng(12).random((37, 45)) * 200).astype(np.float32) if remove_bg else None
    got = pkg.prepare_holograms(fp, fr, bg)
    assert got.shape == (5, 35, 35)
    for k in range(5):
        exp = pkg.prepare_hologram(fp, fr[k].astype(np.float32), bg)
        np.testing.assert_array_equal(got[k].view(np.uint32), exp.view(np.uint32))


def test_many_frames(pkg, gpu_available):
    # 65,537 frames: the mean and write passes take two launches of at most 65,535 frames
    fp = _fp(pkg)
    fr = frames_of(np.uint8, 65537, 4, 4, seed=13)
    got = pkg.prepare_holograms(fp, fr)
    assert got.shape == (65537, 4, 4)
    for k in (0, 1, 65534, 65535, 65536):
        exp = pkg.prepare_hologram(fp, fr[k].astype(np.float32))
        np.testing.assert_array_equal(got[k].view(np.uint32), exp.view(np.uint32))


def test_device_buffers_and_batch_solve(pkg, gpu_available):
    import torch
    fp = _fp(pkg, pad=32)
    fr = frames_of(np.uint8, 4, 64, 64, seed=7)
    dev = torch.from_numpy(fr).cuda()
    out = torch.empty((4, 128, 128), dtype=torch.complex64, device="cuda")
    pkg.prepare_holograms(fp, dev, out=out)
    torch.cuda.synchronize()
    host = pkg.prepare_holograms(fp, fr)
    np.testing.assert_array_equal(out.cpu().numpy(), host)
    # the prepared batch feeds the solver directly from the device
    p = pkg.make_params(128, 128, 4, 1000.0, 50.0, 0.633, 2.0, 3, 0.01, 0.01, 2)
    with pkg.SparseCompressiveHolo(p) as rec:
        v_dev = np.zeros((4, 4, 128, 128), np.complex64)
        rec.reconstructBatch(out, volumes=v_dev)
        v_host = np.zeros_like(v_dev)
        rec.reconstructBatch(host, volumes=v_host)
    np.testing.assert_array_equal(v_dev, v_host)


def test_background_on_device(pkg, gpu_available):
    import torch
    fp = _fp(pkg, remove_bg=True, size=48, center=(30, 30))
    fr = frames_of(np.uint16, 2, 64, 64, seed=3)
    bg = (np.random.default_rng(4).random((64, 64)) * 5000).astype(np.float32)
    got = pkg.prepare_holograms(fp, torch.from_numpy(fr).cuda(), torch.from_numpy(bg).cuda())
    exp = pkg.prepare_holograms(fp, fr, bg)
    np.testing.assert_array_equal(got, exp)


def test_errors(pkg, gpu_available):
    fr = frames_of(np.uint8, 1, 32, 32, seed=1)
    with pytest.raises(pkg.RihvrError) as e:
        pkg.prepare_holograms(_fp(pkg, remove_bg=True), fr)
    assert e.value.status == 3                       # background missing
    with pytest.raises(pkg.RihvrError) as e:
        pkg.prepare_holograms(_fp(pkg, roi=(0, 0, 16, 8)), fr)
    assert e.value.status == 7                       # non-square crop
    with pytest.raises(ValueError):
        pkg.prepare_holograms(_fp(pkg), fr.astype(np.int64))


@pytest.mark.parametrize("dtype", [np.uint16, np.float32])
def test_full_size_frames_match_numpy_oracle(pkg, gpu_available, dtype):
    """1024^2 16-bit and float frames straight against the numpy restatement of the input
    contract (tests/test_io_cpu.py): the device sums a frame's mean in block order, the host
    in raster order, and for these pixel types the double sum is not exact, so values agree
    with the host function to 1 float ulp and with numpy to FP32 rounding."""
    from test_io_cpu import _prepare_np
    fp = _fp(pkg, pad=0)
    fr = frames_of(dtype, 2, 1024, 1024, seed=21)
    got = pkg.prepare_holograms(fp, fr)
    for k in range(2):
        exp_np = _prepare_np(fr[k].astype(np.float32))
        np.testing.assert_allclose(got[k], exp_np, rtol=2e-6, atol=2e-6)
        host = pkg.prepare_hologram(fp, fr[k].astype(np.float32))
        ulps = np.abs(got[k].real.view(np.int32).astype(np.int64) - host.real.view(np.int32).astype(np.int64))
        assert int(ulps.max()) <= 1, int(ulps.max())
        assert np.all(got[k].imag == 0)
