"""CPU tests of the host-side input/output of librihvr.so and of the
sparse-inverse-recon executable's error paths (no device work).

The parameter-file tests follow HoloUi::readFile (src/core/holo_ui.cpp:171-486)
and the file the reference GUI writes (src/QTGUI/reconstructioncompute.cpp:46-111);
the preprocessing tests restate main_sparse_inverse.cu:101-124 with numpy
(rescale hologram.cu:401-439, normalize_mean0 :332-358, removeBackground
:1365-1432, crop :768-880, zeroPad :882-945).
"""
from __future__ import annotations

import ctypes
import pathlib
import struct
import subprocess

import numpy as np
import pytest

from imgfiles import pgm_bytes, read_tiff, tiff_bytes

GUI_FILE = """num_images            1001 1
start_image           1002 3
input_pathn           2002 /data/holo/
input_filename_format 2003 img_%04d.tif
output_pathn          2004 /tmp/out
start_plane_z         3001 2500
plane_step_dz         3002 2
num_planes            3003 1000
wavelength            3004 0.632
resolution            3005 1.1
roi_x                 3006 300
roi_y                 3007 200
roi_size              3009 256
zero_padding          3010 0
inverse_iterations    6001 100
regularization_tau    6002 0.15
regularization_TV     6004 0.1
num_TV_iterations     6005 5
output_planes         7001 true
output_each_step      7005 false
background_image      2005 /data/bg.tif
segment_min_voxels    5009 2
segment_min_intensity 5010 1
segment_close_size    5011 3
"""


def _write(tmp_path, text, name="params.txt"):
    p = tmp_path / name
    p.write_bytes(text.encode())
    return p


def _s(field) -> str:
    return bytes(field).split(b"\0", 1)[0].decode()


# --------------------------------------------------------------------------
# parameter file
# --------------------------------------------------------------------------
def test_param_defaults_missing_file(pkg, tmp_path):
    fp, rc = pkg.read_param_file(tmp_path / "none.txt")
    assert rc == 2  # RIHVR_E_BAD_FILENAME
    s = fp.solver
    # HoloUi::HoloUi defaults (holo_ui.cpp:10-62)
    assert (s.start_plane, s.plane_stepsize, s.num_planes) == (0.0, 1.0, 256)
    assert (s.wavelength, s.resolution, s.zero_padding) == (632.0, 1000.0, 256)
    assert (s.num_inverse_iterations, s.num_tv_iterations) == (500, 5)
    assert s.regularization_param == pytest.approx(0.01) and s.regularization_TV_param == 0.0
    assert s.mode == pkg.MODE_FASTA_FUSED_LASSO_2D
    assert (fp.center_x, fp.center_y, fp.center_z, fp.roi_size) == (-1.0, -1.0, -1.0, -1)
    assert (fp.roi_x, fp.roi_y, fp.roi_w, fp.roi_h) == (0, 0, 0, 0)
    assert (fp.background_start, fp.background_end) == (-1, -1)
    assert fp.output_planes == 0 and fp.output_steps == 0 and fp.remove_background == 0


def test_param_gui_file(pkg, tmp_path):
    fp, rc = pkg.read_param_file(_write(tmp_path, GUI_FILE))
    assert rc == 0 and fp.unknown_id == 0
    s = fp.solver
    assert (fp.image_count, fp.start_image) == (1, 3)
    assert _s(fp.image_path) == "/data/holo/"
    assert _s(fp.image_formatted_name) == "img_%04d.tif"
    assert _s(fp.image_filename) == "/data/holo/img_%04d.tif"  # path + format (no 2001 line)
    assert _s(fp.output_path) == "/tmp/out"
    assert (s.start_plane, s.plane_stepsize, s.num_planes) == (2500.0, 2.0, 1000)
    assert s.wavelength == np.float32(0.632) and s.resolution == np.float32(1.1)
    assert (fp.center_x, fp.center_y, fp.roi_size) == (300.0, 200.0, 256)
    assert s.zero_padding == 0 and s.num_inverse_iterations == 100 and s.num_tv_iterations == 5
    assert s.regularization_param == np.float32(0.15) and s.regularization_TV_param == np.float32(0.1)
    assert fp.output_planes == 1 and fp.output_steps == 0
    assert fp.remove_background == 1 and _s(fp.background_filename) == "/data/bg.tif"


def test_param_full_filename_and_comments(pkg, tmp_path):
    text = ("% a comment line\n# another\ninput_pathn 2002 /p/\ninput_fn 2001 /abs/holo.tif\n"
            "input_filename_format 2003 ignored_%d.tif\nnum_planes 3003 12.9\n")
    fp, rc = pkg.read_param_file(_write(tmp_path, text))
    assert rc == 0
    assert _s(fp.image_filename) == "/abs/holo.tif"  # a 2001 line is used as is
    assert fp.solver.num_planes == 12  # float value converted to int (truncation)


def test_param_flag_reads_previous_value(pkg, tmp_path):
    # "%f" fails on a word, so `value` keeps the previous line's number: a flag
    # line with a word other than true/false takes that number (holo_ui.cpp:195, 437-446)
    text = "num_TV_iterations 6005 4\noutput_each_step 7005 maybe\noutput_planes 7001 0\n"
    fp, rc = pkg.read_param_file(_write(tmp_path, text))
    assert rc == 0 and fp.output_steps == 1 and fp.output_planes == 0
    text = "zero_padding 3010 0\noutput_each_step 7005 yes\noutput_planes 7001 true\n"
    fp, _ = pkg.read_param_file(_write(tmp_path, text))
    assert fp.output_steps == 0 and fp.output_planes == 1


def test_param_unknown_id_stops(pkg, tmp_path):
    text = "input_pathn 2002 /p/\nnum_planes 3003 7\nbogus 9999 1\nnum_inverse_iterations 6001 3\n"
    fp, rc = pkg.read_param_file(_write(tmp_path, text))
    assert rc == 9 and fp.unknown_id == 9999  # RIHVR_E_INVALID_FILE
    assert fp.solver.num_planes == 7  # read before the unknown line
    assert fp.solver.num_inverse_iterations == 500  # not reached
    assert _s(fp.image_filename) == ""  # the reference returns before composing it


def test_param_blank_line_and_no_trailing_newline(pkg, tmp_path):
    # a blank line re-applies the previous id and values; the last line without a
    # newline is read once; trailing '\r' of CRLF files stays in string values
    text = "num_planes 3003 9\n\nwavelength 3004 0.5\r\noutput_pathn 2004 /o\r\nresolution 3005 2"
    fp, rc = pkg.read_param_file(_write(tmp_path, text))
    assert rc == 0
    assert fp.solver.num_planes == 9 and fp.solver.wavelength == np.float32(0.5)
    assert fp.solver.resolution == 2.0
    assert _s(fp.output_path) == "/o\r"


def test_param_other_executables_ids_accepted(pkg, tmp_path):
    text = "".join(f"x {i} 1\n" for i in (4001, 4002, 5001, 5002, 5003, 5004, 5005, 5006, 5007, 5008,
                                           5009, 5010, 5011, 6003, 7002, 7003, 7004, 7006, 2006))
    fp, rc = pkg.read_param_file(_write(tmp_path, text))
    assert rc == 0 and fp.unknown_id == 0


def test_file_params_struct_matches_header(pkg):
    hdr = (pathlib.Path(__file__).resolve().parent.parent / "include" / "rihvr.h").read_text()
    body = hdr[hdr.index("typedef struct rihvr_file_params"):]
    body = body[:body.index("} rihvr_file_params;")]
    names = []
    import re
    for line in body.splitlines()[1:]:
        line = line.split("/*")[0].strip().rstrip(";")
        if not line:
            continue
        decl = line.split(None, 1)[1]
        for d in decl.split(","):
            names.append(re.sub(r"\[.*\]", "", d).strip())
    assert names == [f[0] for f in pkg.FileParams._fields_]


# --------------------------------------------------------------------------
# image decoding
# --------------------------------------------------------------------------
@pytest.mark.parametrize("big_endian", [False, True])
@pytest.mark.parametrize("dtype", [np.uint8, np.uint16])
@pytest.mark.parametrize("rps", [None, 3])
def test_read_tiff_gray(pkg, tmp_path, big_endian, dtype, rps):
    rng = np.random.default_rng(5)
    img = rng.integers(0, np.iinfo(dtype).max + 1, size=(13, 17)).astype(dtype)
    p = tmp_path / "g.tif"
    p.write_bytes(tiff_bytes(img, big_endian=big_endian, rows_per_strip=rps))
    got = pkg.read_image(p)
    assert got.dtype == np.float32 and got.shape == img.shape
    np.testing.assert_array_equal(got, img.astype(np.float32))


def test_read_tiff_rgb_to_gray(pkg, tmp_path):
    # cv::cvtColor BGR2GRAY of 8-bit data: pure red 76, green 150, blue 29; equal
    # channels give the channel value back (the weights sum to one)
    img = np.zeros((2, 4, 3), np.uint8)
    img[0, 0] = (255, 0, 0)
    img[0, 1] = (0, 255, 0)
    img[0, 2] = (0, 0, 255)
    img[0, 3] = (255, 255, 255)
    for i, v in enumerate((0, 1, 128, 254)):
        img[1, i] = (v, v, v)
    p = tmp_path / "rgb.tif"
    p.write_bytes(tiff_bytes(img))
    got = pkg.read_image(p)
    np.testing.assert_array_equal(got, [[76, 150, 29, 255], [0, 1, 128, 254]])


@pytest.mark.parametrize("dtype", [np.uint8, np.uint16])
def test_read_pgm(pkg, tmp_path, dtype):
    img = (np.arange(35).reshape(5, 7) * (1 if dtype == np.uint8 else 1000)).astype(dtype)
    p = tmp_path / "g.pgm"
    p.write_bytes(pgm_bytes(img))
    np.testing.assert_array_equal(pkg.read_image(p), img.astype(np.float32))


def test_read_image_errors(pkg, tmp_path):
    with pytest.raises(pkg.RihvrError) as e:
        pkg.read_image(tmp_path / "missing.tif")
    assert e.value.status == 2
    bad = tmp_path / "x.bin"
    bad.write_bytes(b"not an image at all")
    with pytest.raises(pkg.RihvrError) as e:
        pkg.read_image(bad)
    assert e.value.status == 1
    lzw = tmp_path / "lzw.tif"
    lzw.write_bytes(tiff_bytes(np.zeros((4, 4), np.uint8), compression=5))
    with pytest.raises(pkg.RihvrError) as e:
        pkg.read_image(lzw)
    assert e.value.status == 1
    trunc = tmp_path / "t.tif"
    trunc.write_bytes(tiff_bytes(np.ones((8, 8), np.uint8))[:-10])
    with pytest.raises(pkg.RihvrError) as e:
        pkg.read_image(trunc)
    assert e.value.status == 9


def _tiff_with_size(W, H) -> bytes:
    """An 8x8 TIFF whose ImageWidth / ImageLength tags say W x H (LONG values)."""
    b = bytearray(tiff_bytes(np.ones((8, 8), np.uint8)))
    n = struct.unpack_from("<H", b, 8)[0]
    for i in range(n):
        e = 10 + 12 * i
        tag = struct.unpack_from("<H", b, e)[0]
        if tag in (256, 257):
            struct.pack_into("<HII", b, e + 2, 4, 1, W if tag == 256 else H)
    return bytes(b)


def _size_of(pkg, path):
    w, h = ctypes.c_int32(0), ctypes.c_int32(0)
    rc = pkg.lib().rihvr_read_image(str(path).encode(), None, ctypes.byref(w), ctypes.byref(h))
    return rc, w.value, h.value


def test_read_image_size_limits(pkg, tmp_path):
    """Sizes beyond cv::imread's default limits (side 2^20, 2^30 pixels) are refused as an
    unsupported image (status 1), so a reported size is always positive and fits int32: headers
    of 2^31 (negative as int32), 2^20 + 1 and 2^20 x 2^11 (2^31 pixels); 2^20 x 2^10 is the
    largest accepted and its missing raster is an invalid file (9).  PGM: a 23-digit width and
    a width of 2^20 + 1 are refused the same way."""
    p = tmp_path / "s.tif"
    for W, H in [(1 << 31, 8), (8, 0xFFFFFFFF), ((1 << 20) + 1, 8), (1 << 20, 1 << 11)]:
        p.write_bytes(_tiff_with_size(W, H))
        assert _size_of(pkg, p)[0] == 1, (W, H)
    p.write_bytes(_tiff_with_size(1 << 20, 1 << 10))
    assert _size_of(pkg, p) == (0, 1 << 20, 1 << 10)
    out = np.empty(8, np.float32)  # never written: the raster check comes first
    w, h = ctypes.c_int32(0), ctypes.c_int32(0)
    assert pkg.lib().rihvr_read_image(str(p).encode(), ctypes.c_void_p(out.ctypes.data), ctypes.byref(w),
                                      ctypes.byref(h)) == 9
    g = tmp_path / "s.pgm"
    for head in (b"P5\n99999999999999999999999 10\n255\n", b"P5 1048577 1 255\n"):
        g.write_bytes(head + bytes(64))
        assert _size_of(pkg, g)[0] == 1, head


def test_writers_and_preparation_refuse_sizes_beyond_the_readers(pkg, tmp_path):
    """The writers, rihvr_write_hologram and rihvr_prepare_hologram refuse sizes the readers
    refuse (status 7, invalid argument) before touching the pixels (a 1-element buffer stands for them)."""
    L = pkg.lib()
    one_u8 = np.zeros(1, np.uint8)
    one_f = np.zeros(1, np.float32)
    p = str(tmp_path / "big.tif").encode()
    for W, H in [((1 << 20) + 1, 1), (1 << 16, 1 << 16), (0x7FFFFFFF, 0x7FFFFFFF)]:
        assert L.rihvr_write_tiff8(p, ctypes.c_void_p(one_u8.ctypes.data), W, H) == 7
        assert L.rihvr_write_png8(p, ctypes.c_void_p(one_u8.ctypes.data), W, H) == 7
        assert L.rihvr_write_hologram(p, ctypes.c_void_p(one_f.ctypes.data), W, H, 1, 0.0, 1.0) == 7
        ow, oh = ctypes.c_int32(0), ctypes.c_int32(0)
        assert L.rihvr_prepare_hologram(ctypes.byref(_fp(pkg)), ctypes.c_void_p(one_f.ctypes.data), None, W, H, None,
                                        ctypes.byref(ow), ctypes.byref(oh)) == 7
    assert not (tmp_path / "big.tif").exists()


# --------------------------------------------------------------------------
# preprocessing (the executable's input contract)
# --------------------------------------------------------------------------
def _fp(pkg, pad=0, center=(-1, -1), size=-1, roi=(0, 0, 0, 0), remove_bg=False):
    fp = pkg.FileParams()
    fp.solver.zero_padding = pad
    fp.center_x, fp.center_y = center
    fp.roi_size = size
    fp.roi_x, fp.roi_y, fp.roi_w, fp.roi_h = roi
    fp.remove_background = 1 if remove_bg else 0
    return fp


def _normalise_np(img, bg=None):
    inv255 = np.float32(1.0 / 255.0)
    h = img.astype(np.float32) * inv255
    if bg is not None:
        b = bg.astype(np.float32) * inv255
        sb = np.maximum(np.sqrt(b), np.float32(0.1))
        return (h - b) / sb
    inv_mean = np.float32(1.0 / h.astype(np.float64).mean())
    return h * inv_mean - np.float32(1.0)


def _crop_np(h, center, size, roi):
    H, W = h.shape
    if roi[2] * roi[3] > 0:
        x, y, w, hh = roi
    else:
        cx, cy = int(center[0]), int(center[1])
        w = hh = size
        # C's size / 2 truncates toward zero
        x = max(cx - int(size / 2), 0)
        y = max(cy - int(size / 2), 0)
        if x + size > W:
            x = W - size
        if y + size > H:
            y = H - size
        if W < size:
            x, w = 0, W
        if H < size:
            y, hh = 0, H
        if size < 1:
            x, y, w, hh = 0, 0, W, H
    top = max(0, -y)
    y = max(y, 0)
    bottom = max(0, y + hh - (H + top))
    left = max(0, -x)
    x = max(x, 0)
    # the reference sizes the right border with the ROI height (hologram.cu:848-852)
    right = max(0, x + hh - (W + left)) if x + w > W + left else 0
    big = np.pad(h, ((top, bottom), (left, right)))
    assert y + hh <= big.shape[0] and x + w <= big.shape[1]
    return big[y:y + hh, x:x + w]


def _prepare_np(img, bg=None, pad=0, center=(-1, -1), size=-1, roi=(0, 0, 0, 0)):
    c = _crop_np(_normalise_np(img, bg), center, size, roi)
    out = np.zeros((c.shape[0] + 2 * pad, c.shape[1] + 2 * pad), np.complex64)
    out[pad:pad + c.shape[0], pad:pad + c.shape[1]] = c
    return out


@pytest.fixture(scope="module")
def image():
    rng = np.random.default_rng(11)
    return rng.integers(1, 256, size=(48, 48)).astype(np.float32)


def test_prepare_mean_normalise_and_pad(pkg, image):
    got = pkg.prepare_hologram(_fp(pkg, pad=8), image)
    exp = _prepare_np(image, pad=8)
    assert got.shape == (64, 64)
    np.testing.assert_allclose(got, exp, rtol=2e-7, atol=2e-7)
    assert np.all(got.imag == 0)
    # the mean of the interior is zero by construction
    assert abs(got[8:56, 8:56].real.astype(np.float64).mean()) < 1e-6


def test_prepare_background(pkg, image):
    rng = np.random.default_rng(12)
    bg = rng.integers(0, 256, size=image.shape).astype(np.float32)
    bg[0, :4] = 0  # sqrt(B) below 0.1 takes the floor
    got = pkg.prepare_hologram(_fp(pkg, remove_bg=True), image, bg)
    np.testing.assert_allclose(got, _prepare_np(image, bg=bg), rtol=2e-7, atol=2e-7)
    with pytest.raises(pkg.RihvrError) as e:
        pkg.prepare_hologram(_fp(pkg, remove_bg=True), image)
    assert e.value.status == 3  # RIHVR_E_MISSING_DATA


@pytest.mark.parametrize("center,size", [((20, 30), 16), ((2, 3), 16), ((46, 47), 16), ((24, 24), 64),
                                         ((24, 24), 0), ((24, 24), -1), ((11, 40), 31)])
def test_prepare_crop_centre(pkg, image, center, size):
    got = pkg.prepare_hologram(_fp(pkg, pad=2, center=center, size=size), image)
    exp = _prepare_np(image, pad=2, center=center, size=size)
    np.testing.assert_allclose(got, exp, rtol=2e-7, atol=2e-7)


@pytest.mark.parametrize("roi", [(4, 6, 20, 20), (-5, -3, 16, 16), (40, 40, 16, 16), (0, 0, 48, 48)])
def test_prepare_crop_rect(pkg, image, roi):
    got = pkg.prepare_hologram(_fp(pkg, roi=roi), image)
    exp = _prepare_np(image, roi=roi)
    np.testing.assert_allclose(got, exp, rtol=2e-7, atol=2e-7)


def test_prepare_crop_rect_zero_fill(pkg, image):
    # a rectangle past the bottom-right corner is filled with zeros
    got = pkg.prepare_hologram(_fp(pkg, roi=(40, 40, 16, 16)), image)
    assert np.all(got[8:, :] == 0) and np.all(got[:, 8:] == 0)
    assert np.any(got[:8, :8] != 0)


def test_prepare_non_square_rejected(pkg, image):
    # zeroPad allocates rows/cols swapped (hologram.cu:905): only square crops pass
    with pytest.raises(pkg.RihvrError) as e:
        pkg.prepare_hologram(_fp(pkg, roi=(0, 0, 16, 8)), image)
    assert e.value.status == 7


# --------------------------------------------------------------------------
# writers
# --------------------------------------------------------------------------
def test_write_tiff8_roundtrip(pkg, tmp_path):
    rng = np.random.default_rng(3)
    px = rng.integers(0, 256, size=(9, 14)).astype(np.uint8)
    p = tmp_path / "w.tif"
    pkg.write_tiff8(p, px)
    np.testing.assert_array_equal(read_tiff(p), px)
    np.testing.assert_array_equal(pkg.read_image(p), px.astype(np.float32))


def test_write_hologram_minmax_and_range(pkg, tmp_path):
    rng = np.random.default_rng(4)
    holo = (rng.standard_normal((16, 16)) + 1j * rng.standard_normal((16, 16))).astype(np.complex64)
    p = tmp_path / "h.tif"
    pkg.write_hologram(p, holo)
    v = holo.real.astype(np.float64)
    exp = np.clip(np.rint(((v - v.min()) * (255.0 / (v.max() - v.min()))).astype(np.float32)), 0, 255)
    got = read_tiff(p)
    np.testing.assert_array_equal(got, exp.astype(np.uint8))
    assert got.min() == 0 and got.max() == 255
    lo, hi = -0.5, 0.5
    pkg.write_hologram(p, holo, lo, hi)
    exp = np.clip(np.rint((255 * (v - lo) / (hi - lo)).astype(np.float32)), 0, 255).astype(np.uint8)
    np.testing.assert_array_equal(read_tiff(p), exp)  # values outside lo..hi saturate


def test_write_hologram_ties_to_even(pkg, tmp_path):
    # 255 (v - 0) / (255 - 0) = v: x.5 values round to the even neighbour
    holo = np.array([[0.5, 1.5, 2.5, 3.5, 254.5, 300.0, -3.0, 7.25]], np.complex64)
    p = tmp_path / "t.tif"
    pkg.write_hologram(p, holo, 0.0, 255.0)
    np.testing.assert_array_equal(read_tiff(p), [[0, 2, 2, 4, 254, 255, 0, 7]])


# --------------------------------------------------------------------------
# the executable (paths that end before any device work)
# --------------------------------------------------------------------------
def _cli(pkg):
    exe = pkg.PKG_DIR / "_build" / "sparse-inverse-recon"
    if not exe.exists():
        pytest.skip("sparse-inverse-recon not built")
    return exe


def test_cli_requires_param_file(pkg):
    r = subprocess.run([str(_cli(pkg))], capture_output=True, text=True, timeout=60)
    assert r.returncode == 255
    assert "User must give input parameter file using -F flag" in r.stdout


def test_cli_unknown_option(pkg):
    r = subprocess.run([str(_cli(pkg)), "-X"], capture_output=True, text=True, timeout=60)
    assert r.returncode == 255 and "Unknown input argument" in r.stdout


def test_cli_missing_image(pkg, tmp_path):
    p = _write(tmp_path, f"input_fn 2001 {tmp_path}/nothing_%04d.tif\noutput_pathn 2004 {tmp_path}\n")
    r = subprocess.run([str(_cli(pkg)), "-F", str(p), "-S", "1"], capture_output=True, text=True, timeout=60)
    assert r.returncode == 255
    assert f"Unable to read file: {tmp_path}/nothing_0000.tif" in r.stdout
    assert "Detected HoloError: 2" in r.stdout
    assert "Save planes in subfolder flag set to 1" in r.stdout


def test_file_params_layout_matches_c(pkg, tmp_path):
    src = tmp_path / "sz.c"
    src.write_text('#include <stddef.h>\n#include <stdio.h>\n#include "rihvr.h"\n'
                   'int main(void){printf("%zu %zu %zu", sizeof(rihvr_file_params),'
                   ' offsetof(rihvr_file_params, remove_background), offsetof(rihvr_file_params, unknown_id));'
                   'return 0;}\n')
    exe = tmp_path / "sz"
    inc = pathlib.Path(__file__).resolve().parent.parent / "include"
    subprocess.run(["gcc", "-std=c99", f"-I{inc}", str(src), "-o", str(exe)], check=True, timeout=60)
    out = subprocess.run([str(exe)], capture_output=True, text=True, check=True).stdout.split()
    F = pkg.FileParams
    assert [int(v) for v in out] == [ctypes.sizeof(F), F.remove_background.offset, F.unknown_id.offset]
