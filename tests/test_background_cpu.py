"""CPU tests of the background layer: the PNG writer (decoded with zlib), the
parameter-file flag 7006, the oracle's background restatement pinned by closed
forms and by the reference's frame schedules, and make-background's error paths."""
from __future__ import annotations

import pathlib
import struct
import subprocess
import zlib

import numpy as np
import pytest

from imgfiles import tiff_bytes

ROOT = pathlib.Path(__file__).resolve().parent.parent


@pytest.fixture(scope="module")
def bgo():
    import oracle.background as b
    return b


def read_png(path):
    b = pathlib.Path(path).read_bytes()
    assert b[:8] == b"\x89PNG\r\n\x1a\n"
    pos, chunks = 8, []
    while pos < len(b):
        (n,) = struct.unpack(">I", b[pos:pos + 4])
        typ, data = b[pos + 4:pos + 8], b[pos + 8:pos + 8 + n]
        (crc,) = struct.unpack(">I", b[pos + 8 + n:pos + 12 + n])
        assert crc == zlib.crc32(typ + data) & 0xFFFFFFFF
        chunks.append((typ, data))
        pos += 12 + n
    assert chunks[0][0] == b"IHDR" and chunks[-1][0] == b"IEND"
    w, h, depth, ctype, comp, filt, inter = struct.unpack(">IIBBBBB", chunks[0][1])
    assert (depth, ctype, comp, filt, inter) == (8, 0, 0, 0, 0)
    raw = zlib.decompress(b"".join(d for t, d in chunks if t == b"IDAT"))
    rows = np.frombuffer(raw, np.uint8).reshape(h, w + 1)
    assert np.all(rows[:, 0] == 0)
    return rows[:, 1:].copy()


@pytest.mark.parametrize("shape", [(1, 1), (7, 13), (300, 260)])  # the last spans several stored blocks
def test_png_roundtrip(pkg, tmp_path, shape):
    px = np.random.default_rng(1).integers(0, 256, size=shape).astype(np.uint8)
    p = tmp_path / "x.png"
    pkg.write_png8(p, px)
    np.testing.assert_array_equal(read_png(p), px)


def test_param_median_flag(pkg, tmp_path):
    # the layout the reference GUI writes for make-background (src/QTGUI/bgcompute.cpp:26-34)
    text = ("num_images            1001 50\nstart_image           1002 0\nbg_start_image        1003 5\n"
            "bg_end_image          1004 30\ninput_pathn           2002 /d/\n"
            "input_filename_format 2003 f_%04d.tif\noutput_pathn          2004 /o\n"
            "use_median_background 7006 1\n")
    p = tmp_path / "bg.txt"
    p.write_text(text)
    fp, rc = pkg.read_param_file(p)
    assert rc == 0 and fp.use_median_background == 1
    assert (fp.image_count, fp.start_image, fp.background_start, fp.background_end) == (50, 0, 5, 30)


def test_mean_background_closed_form(bgo):
    """f_k = c_k + A_k p with p a +-1 checkerboard: B = mean(c) + mean(A) p, I_k - B has
    mean c_k - mean(c) and standard deviation |A_k - mean(A)|."""
    c = np.array([100.0, 104.0, 96.0, 120.0])
    A = np.array([8.0, 4.0, 12.0, 0.0])
    y, x = np.mgrid[0:16, 0:24]
    p = np.where((x + y) % 2 == 0, 1.0, -1.0)
    frames = np.stack([ck + ak * p for ck, ak in zip(c, A)]).astype(np.float32)
    bg, st = bgo.mean_background(frames)
    np.testing.assert_array_equal(bg, (c.mean() + A.mean() * p).astype(np.float32))
    ms = np.mean(np.abs(A - A.mean()))
    assert st["alpha"] == pytest.approx(0.0 - 5.0 * ms, abs=1e-12)
    assert st["beta"] == pytest.approx(255.0 / (2 * 5 * ms))
    assert st["average"] == pytest.approx(c.mean()) and st["minimum"] == 0.0


def test_background_frame_schedules(bgo):
    ids, step = bgo.frame_ids(0, 50, 10, 20)
    assert ids == list(range(10, 19)) and step == 1          # end - start - 1 frames
    ids, step = bgo.frame_ids(3, 7, -1, -1)
    assert ids == list(range(3, 10)) and step == 1           # image_count from start_image
    ids, step = bgo.frame_ids(0, 0, 0, 1001)
    assert step == 5 and len(ids) == 200 and ids[-1] == 995  # more than 500: about 200 frames
    # the median window: outputs 0-2 from frames 0-4, then frame i replaces slot (i - 3) % 5
    sched = bgo.median_schedule(0, 8)
    assert [o for o, _ in sched] == [0, 1, 2, 3, 4, 5]
    assert sched[3][1] == [3, 1, 2, 3, 4] and sched[5][1] == [3, 4, 5, 3, 4]
    v = np.array([[5.0], [1.0], [4.0], [2.0], [3.0]], np.float32)[:, :, None]
    assert bgo.median5(v).item() == 3.0


def _exe(pkg):
    exe = pkg.PKG_DIR / "_build" / "make-background"
    if not exe.exists():
        pytest.skip("make-background not built")
    return exe


def test_make_background_error_paths(pkg, tmp_path):
    r = subprocess.run([str(_exe(pkg))], capture_output=True, text=True, timeout=60)
    assert r.returncode == 255 and "User must give input parameter file using -F flag" in r.stdout
    p = tmp_path / "bg.txt"
    p.write_text(f"input_fn 2001 {tmp_path}/none_%04d.tif\noutput_pathn 2004 {tmp_path}\n")
    r = subprocess.run([str(_exe(pkg)), "-F", str(p)], capture_output=True, text=True, timeout=60)
    assert r.returncode == 255 and "Detected HoloError: 2" in r.stdout
    (tmp_path / "f_0000.tif").write_bytes(tiff_bytes(np.zeros((16, 16), np.uint8)))
    p.write_text(f"input_fn 2001 {tmp_path}/f_%04d.tif\noutput_pathn 2004 {tmp_path}\n"
                 "bg_start 1003 0\nbg_end 1004 3\nuse_median_background 7006 true\n")
    r = subprocess.run([str(_exe(pkg)), "-F", str(p)], capture_output=True, text=True, timeout=60)
    assert r.returncode == 255
    assert "Median background calculation requires at least 5 frames" in r.stdout
    assert "Detected HoloError: 11" in r.stdout


# ---- correlation background (ccbg.cpp) ------------------------------------------------
def test_box_resize_is_the_block_mean():
    from oracle import background as B
    rng = np.random.default_rng(3)
    f = rng.integers(0, 256, size=(2, 8, 12)).astype(np.float32)
    got = B.box_resize(f, 4)
    exp = f.reshape(2, 2, 4, 3, 4).mean(axis=(2, 4))
    np.testing.assert_allclose(got, exp, rtol=0, atol=1e-5)
    assert np.array_equal(B.box_resize(f, 1), f)


def test_entropy_known_answers():
    from oracle import background as B
    assert B.entropy_f32(np.bincount([7] * 100, minlength=256), 100) == 0.0
    two = np.bincount([0] * 50 + [9] * 50, minlength=256)
    assert abs(float(B.entropy_f32(two, 100)) - np.log(2.0)) < 1e-6


def test_frame_count_rule_by_hand():
    from oracle import background as B
    # entropies 1, 2, 3, 2, 1: smoothed 1.5, 2, 2.333, 2, 1.5 -> first fall at index 3
    assert B.pick_frame_count([1, 2, 3, 2, 1], 1, 5) == 3
    # rising throughout: no fall -> 1, then the clamps
    assert B.pick_frame_count([1, 2, 3, 4, 5], 1, 5) == 1
    assert B.pick_frame_count([1, 2, 3, 4, 5], 4, 5) == 3
    assert B.pick_frame_count([1, 2, 3, 2, 1], 9, 5) == 4  # min_frames beyond the ranking


def test_ranking_groups_similar_frames():
    """Frames 0-2 are pattern A plus noise, 3-5 pattern B plus noise: each frame's two best
    matches are the rest of its group; a frame and an affine copy of it correlate to 1."""
    from oracle import background as B
    rng = np.random.default_rng(4)
    a, b = rng.random((16, 16)) * 200, rng.random((16, 16)) * 200
    fr = np.stack([a + rng.normal(0, 3, a.shape) for _ in range(3)] +
                  [b + rng.normal(0, 3, b.shape) for _ in range(3)]).astype(np.float32)
    order = B.ranking(fr)
    for i in range(6):
        assert set(order[i][:2]) == ({0, 1, 2} - {i} if i < 3 else {3, 4, 5} - {i})
    x = np.stack([a, 2 * a + 7]).astype(np.float32)
    xm = x.reshape(2, -1).astype(np.float64)
    xm -= xm.mean(axis=1, keepdims=True)
    c = (xm[0] @ xm[1]) / (np.sqrt((xm[0] @ xm[0]) * (xm[1] @ xm[1])) + 1e-8)
    assert abs(c - 1.0) < 1e-9


def test_correlation_background_of_a_static_scene():
    """A static background with a spot that moves every frame: with enough frames averaged
    the spot drops out, and the frames used respect the clamps."""
    from oracle import background as B
    rng = np.random.default_rng(5)
    base = rng.integers(60, 180, size=(32, 32)).astype(np.float32)
    fr = np.repeat(base[None], 10, axis=0).copy()
    for i in range(10):
        fr[i, 3 * i % 30: 3 * i % 30 + 2, 5:7] = 255
    bg, used, _ = B.correlation_backgrounds(fr, 3, 8, 4)
    assert np.all((used >= 3) & (used <= 8))
    assert np.abs(bg.astype(int) - base.astype(int)).max() <= 255 // 3 + 1
