"""Malformed input files through the host readers (no device work): seeded mutations of valid
TIFF and PGM files (byte changes, extreme header values, truncation, extension, repeated
slices), random parameter files and random sparse-voxel files.  Every call must return a
status, never crash or write past the caller's buffer; an image that reads must report a
positive size and fill exactly the buffer that size describes (a guard region after it stays
untouched).  The readers are Hologram::read (hologram.cu:139-199), HoloUi::readFile
(holo_ui.cpp:171-486) and SparseVolume::loadData (sparse_volume.cu:562-682).
"""
from __future__ import annotations

import ctypes
import random

import numpy as np
import pytest

from imgfiles import pgm_bytes, tiff_bytes

N_FILES = 400
GUARD = 64


def _seeds():
    out = []
    for dt in (np.uint8, np.uint16):
        img = (np.arange(12 * 9).reshape(9, 12) * 7 % 250).astype(dt)
        out += [tiff_bytes(img), tiff_bytes(img, big_endian=True, rows_per_strip=2), pgm_bytes(img)]
    out.append(tiff_bytes(np.zeros((5, 7, 3), np.uint8)))
    return out


def _mutate(rng: random.Random, b: bytes) -> bytes:
    b = bytearray(b)
    for _ in range(rng.randint(1, 6)):
        k = rng.randint(0, 4)
        if k == 0 and b:
            b[rng.randrange(len(b))] = rng.randrange(256)
        elif k == 1 and len(b) > 4:
            i = rng.randrange(len(b) - 4)
            v = rng.choice([0, 1, 0xFFFF, 0xFFFFFFFF, 0x7FFFFFFF, 0x8000, rng.randrange(1 << 32)])
            b[i:i + 4] = v.to_bytes(4, "little")
        elif k == 2:
            b = b[:rng.randrange(len(b) + 1)]
        elif k == 3:
            b += bytes(rng.randrange(256) for _ in range(rng.randint(1, 64)))
        elif len(b) > 8:
            i = rng.randrange(len(b))
            b[i:i] = b[i:min(len(b), i + rng.randint(1, 32))]
    return bytes(b)


@pytest.mark.parametrize("seed", [11, 12])
def test_mutated_images(pkg, tmp_path, seed):
    rng = random.Random(seed)
    seeds = _seeds()
    p = tmp_path / "f.img"
    path = str(p).encode()
    read = refused = 0
    for _ in range(N_FILES):
        p.write_bytes(_mutate(rng, rng.choice(seeds)))
        w, h = ctypes.c_int32(0), ctypes.c_int32(0)
        rc = pkg.lib().rihvr_read_image(path, None, ctypes.byref(w), ctypes.byref(h))
        if rc != 0:
            refused += 1
            continue
        assert w.value > 0 and h.value > 0
        if w.value * h.value > 1 << 22:
            continue
        n = w.value * h.value
        buf = np.full(n + GUARD, -7.0, np.float32)
        rc = pkg.lib().rihvr_read_image(path, ctypes.c_void_p(buf.ctypes.data), ctypes.byref(w), ctypes.byref(h))
        assert np.all(buf[n:] == -7.0)
        if rc == 0:
            read += 1
            assert w.value * h.value == n
    # the mutations leave many files readable and make many unreadable
    assert read > N_FILES // 10 and refused > N_FILES // 10, (read, refused)


PARAM_IDS = [1001, 1002, 1003, 1004, 2001, 2002, 2003, 2004, 2005, 3001, 3002, 3003, 3004, 3005, 3006, 3010,
             3014, 5009, 5010, 5011, 6001, 6002, 6004, 6005, 7001, 7005, 7006, 9999]


def test_random_parameter_files(pkg, tmp_path, capfd):
    rng = random.Random(13)
    p = tmp_path / "params.txt"
    for _ in range(N_FILES):
        lines = []
        for _ in range(rng.randint(0, 30)):
            k = rng.randint(0, 5)
            if k == 0:
                lines.append("")
            elif k == 1:
                lines.append("#" + "x" * rng.randint(0, 100))
            elif k == 2:
                val = rng.choice(["", "1", "-5", "1e40", "nan", "abc", "x" * 5000, "0.5"])
                lines.append(f"name {rng.choice(PARAM_IDS)} {val}")
            elif k == 3:
                lines.append("".join(chr(rng.randrange(1, 256)) for _ in range(rng.randint(0, 200))))
            elif k == 4:
                lines.append(f"name {rng.choice(PARAM_IDS)}")
            else:
                lines.append(f"a b c d {rng.randint(-10, 10)}")
        p.write_text("\n".join(lines), encoding="latin-1")
        try:
            pkg.read_param_file(p)
        except pkg.RihvrError:
            pass
    capfd.readouterr()  # the parser's messages about unknown lines


def test_random_voxel_files(pkg, tmp_path):
    rng = random.Random(14)
    p = tmp_path / "inverse_voxels.csv"
    path = str(p).encode()
    W, H, NZ = 32, 32, 8
    for _ in range(N_FILES):
        rows = ["x, y, z, value" if rng.random() < 0.9 else "garbage"]
        for _ in range(rng.randint(0, 40)):
            k = rng.randint(0, 3)
            if k == 0:
                rows.append(f"{rng.randint(-3, 40)}, {rng.randint(-3, 40)}, {rng.randint(-3, 10)}, {rng.random()}")
            elif k == 1:
                rows.append(",".join(str(rng.randint(-5, 50)) for _ in range(rng.randint(0, 6))))
            elif k == 2:
                rows.append("".join(chr(rng.randrange(32, 127)) for _ in range(rng.randint(0, 80))))
            else:
                rows.append(f"{rng.randint(0, 31)}, {rng.randint(0, 31)}, {rng.randint(0, 7)}, {rng.random() * 100}")
        p.write_text("\n".join(rows) + ("\n" if rng.random() < 0.5 else ""))
        cnt = ctypes.c_int64(0)
        rc = pkg.lib().rihvr_read_sparse_voxels(path, W, H, NZ, ctypes.byref(cnt), None, None, None)
        if rc != 0:
            continue
        n = cnt.value
        assert 0 <= n <= len(rows)
        pl = np.full(n + GUARD, -7, np.int32)
        ix = np.full(n + GUARD, -7, np.int32)
        iv = np.full(n + GUARD, -7.0, np.float32)
        rc = pkg.lib().rihvr_read_sparse_voxels(path, W, H, NZ, ctypes.byref(cnt), ctypes.c_void_p(pl.ctypes.data),
                                                ctypes.c_void_p(ix.ctypes.data), ctypes.c_void_p(iv.ctypes.data))
        assert rc == 0 and cnt.value == n
        assert np.all(pl[n:] == -7) and np.all(ix[n:] == -7) and np.all(iv[n:] == -7.0)
        assert np.all((pl[:n] >= 0) & (pl[:n] < NZ)) and np.all((ix[:n] >= 0) & (ix[:n] < W * H))
