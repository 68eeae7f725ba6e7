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
