"""GPU tests of the background estimation (rihvr_background_mean / _median5) and of
the make-background executable against the CPU restatement in oracle/background.py."""
from __future__ import annotations

import subprocess

import numpy as np
import pytest

from imgfiles import tiff_bytes
from test_background_cpu import read_png

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def bgo():
    import oracle.background as b
    return b


def _frames(n, h=48, w=64, seed=3):
    rng = np.random.default_rng(seed)
    base = rng.integers(60, 200, size=(h, w)).astype(np.float64)
    return np.clip(base[None] + rng.normal(0, 12, size=(n, h, w)), 0, 255).round().astype(np.uint8)


@pytest.mark.parametrize("n, h, w", [(7, 48, 64), (21, 48, 64), (21, 37, 45)])
def test_background_mean_matches_oracle(pkg, bgo, gpu_available, n, h, w):
    # 21 frames run the sum pass's batches of 8 loads and its remainder; 37 x 45 pixels
    # (no multiple of 4) take the one-pixel-per-thread passes
    f = _frames(n, h, w).astype(np.float32)
    bg, st = pkg.background_mean(f)
    rbg, rst = bgo.mean_background(f)
    np.testing.assert_array_equal(bg, rbg)  # same float summation order
    for k in ("alpha", "beta", "minimum", "average"):
        assert st[k] == pytest.approx(rst[k], rel=1e-12, abs=1e-12), k
    with pytest.raises(pkg.RihvrError) as e:
        pkg.background_mean(np.zeros((0, 4, 4), np.float32))
    assert e.value.status == 3


def test_background_mean_many_frames(pkg, bgo, gpu_available):
    # 65,537 frames: the statistics pass's frame rows (frames + 1) take two launches of at
    # most 65,535 grid rows each
    f = np.random.default_rng(8).integers(0, 256, size=(65537, 4, 4)).astype(np.float32)
    bg, st = pkg.background_mean(f)
    rbg, rst = bgo.mean_background(f)
    np.testing.assert_array_equal(bg, rbg)
    for k in ("alpha", "beta", "minimum", "average"):
        assert st[k] == pytest.approx(rst[k], rel=1e-12, abs=1e-12), k


def test_background_median5_matches_oracle(pkg, bgo, gpu_available):
    f = _frames(5, seed=4).astype(np.float32)
    np.testing.assert_array_equal(pkg.background_median5(f), bgo.median5(f))


def _write_frames(d, frames):
    for i, fr in enumerate(frames):
        (d / f"f_{i:04d}.tif").write_bytes(tiff_bytes(fr))


def _exe(pkg):
    exe = pkg.PKG_DIR / "_build" / "make-background"
    assert exe.exists(), "make-background is not built"
    return exe


def test_make_background_mean_end_to_end(pkg, bgo, tmp_path, gpu_available):
    frames = _frames(12, seed=5)
    _write_frames(tmp_path, frames)
    out = tmp_path / "out"
    out.mkdir()
    p = tmp_path / "bg.txt"
    # the layout the reference GUI writes (src/QTGUI/bgcompute.cpp:26-34)
    p.write_text(f"num_images            1001 12\nstart_image           1002 0\nbg_start_image        1003 2\n"
                 f"bg_end_image          1004 11\ninput_pathn           2002 {tmp_path}/\n"
                 f"input_filename_format 2003 f_%04d.tif\noutput_pathn          2004 {out}\n"
                 f"use_median_background 7006 0\n")
    r = subprocess.run([str(_exe(pkg)), "-F", str(p)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
    ids, step = bgo.frame_ids(0, 12, 2, 11)
    assert f"Used {len(ids)} images for background calculation" in r.stdout
    rbg, rst = bgo.mean_background(frames[ids].astype(np.float32))
    np.testing.assert_array_equal(read_png(out / "background.png"), bgo.to_u8(rbg))
    lines = (out / "background_stats.txt").read_text().splitlines()
    assert lines == ["%f" % rst[k] for k in ("alpha", "beta", "minimum", "average")]


def test_make_background_median_end_to_end(pkg, bgo, tmp_path, gpu_available):
    frames = _frames(9, seed=6)
    _write_frames(tmp_path, frames)
    out = tmp_path / "out"
    out.mkdir()
    p = tmp_path / "bg.txt"
    p.write_text(f"start_image 1002 0\nbg_start_image 1003 0\nbg_end_image 1004 9\n"
                 f"input_pathn 2002 {tmp_path}/\ninput_filename_format 2003 f_%04d.tif\n"
                 f"output_pathn 2004 {out}\nuse_median_background 7006 1\n")
    r = subprocess.run([str(_exe(pkg)), "-F", str(p)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
    sched = bgo.median_schedule(0, 9)
    for oid, stack in sched:
        exp = bgo.to_u8(bgo.median5(frames[stack].astype(np.float32)))
        np.testing.assert_array_equal(read_png(out / f"background_{oid:04d}.png"), exp)
    assert sorted(q.name for q in out.glob("background_*.png")) == [f"background_{o:04d}.png" for o, _ in sched]


def _moving_spot_frames(n, h, w, seed):
    """A static scene with a spot that moves each frame and camera noise (8-bit values)."""
    rng = np.random.default_rng(seed)
    base = rng.integers(50, 190, size=(h, w)).astype(np.float64)
    fr = np.clip(base[None] + rng.normal(0, 6, size=(n, h, w)), 0, 255)
    for i in range(n):
        y, x = (7 * i) % (h - 4), (11 * i) % (w - 4)
        fr[i, y:y + 3, x:x + 3] = 250
    return fr.round().astype(np.float32)


@pytest.mark.parametrize("n, h, w, k, lo, hi", [(12, 64, 64, 4, 2, 8), (7, 48, 32, 1, 1, 100), (16, 96, 80, 2, 5, 12)])
@pytest.mark.parametrize("device", [False, True])
def test_background_correlation_matches_oracle(pkg, bgo, gpu_available, n, h, w, k, lo, hi, device):
    print(f"ccbg case {n}x{h}x{w} k={k} min={lo} max={hi} device={device}", flush=True)
    fr = _moving_spot_frames(n, h, w, seed=n + k)
    arg = fr
    if device:
        import torch
        arg = torch.from_numpy(fr).cuda()
    bg, used, enh = pkg.background_correlation(arg, lo, hi, k, enhanced=True)
    rbg, rused, renh = bgo.correlation_backgrounds(fr, lo, hi, k, enhanced=True)
    assert list(used) == list(rused)
    np.testing.assert_array_equal(bg, rbg)
    np.testing.assert_array_equal(enh, renh)


def test_ccbg_end_to_end(pkg, bgo, tmp_path, gpu_available):
    fr = _moving_spot_frames(9, 64, 64, seed=8).astype(np.uint8)
    inp = tmp_path / "in"
    inp.mkdir()
    for i in range(9):
        (inp / f"im_{i + 1:04d}.tif").write_bytes(tiff_bytes(fr[i]))
    out = tmp_path / "bgs"
    p = tmp_path / "params.txt"
    p.write_text(f"# ccbg parameters\ninputFolder = {inp}\noutputFolder = {out}\nfilePattern = im_%04d.tif\n"
                 f"startIdx = 1\nendIdx = 9\nminFrames = 3\nmaxFrames = 6\nresizePct = 0.25\nsaveEnhanced = true\n")
    exe = _exe(pkg).parent / "ccbg"
    r = subprocess.run([str(exe), str(p)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
    rbg, rused, renh = bgo.correlation_backgrounds(fr.astype(np.float32), 3, 6, 4, enhanced=True)
    for i in range(9):
        np.testing.assert_array_equal(read_png(out / f"background_{i + 1:04d}.png"), rbg[i])
        np.testing.assert_array_equal(read_png(out / f"enhanced_{i + 1:04d}.png"), renh[i])
        assert f"Saved background for image {i + 1} using {rused[i]} frames + enhanced image" in r.stdout
