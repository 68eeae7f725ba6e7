"""Minimal TIFF / PGM encoders and a TIFF decoder for the tests.

Written from the TIFF 6.0 baseline layout (header, one IFD, strip offsets and
byte counts) independently of the library's own reader and writer, so that a
file round trip checks one against the other.
"""
from __future__ import annotations

import struct

import numpy as np

_TYPE_SIZE = {3: 2, 4: 4}


def tiff_bytes(img: np.ndarray, big_endian=False, rows_per_strip=None, photometric=None,
               compression=1) -> bytes:
    """Uncompressed TIFF of a [H][W] uint8/uint16 or [H][W][3] uint8 (RGB) array."""
    img = np.asarray(img)
    e = ">" if big_endian else "<"
    H, W = img.shape[:2]
    spp = 3 if img.ndim == 3 else 1
    bps = img.dtype.itemsize * 8
    rps = rows_per_strip or H
    raster = img.astype(img.dtype.newbyteorder(e)).tobytes()
    row_bytes = W * spp * img.dtype.itemsize
    strips = [raster[y * row_bytes:min(H, y + rps) * row_bytes] for y in range(0, H, rps)]
    n_tags = 10
    ifd_off = 8
    extra_off = ifd_off + 2 + 12 * n_tags + 4
    extra = b""
    # strip offsets / counts arrays (always out of line when more than one)
    offs_pos = extra_off + len(extra)
    data_start = extra_off + 8 * len(strips) + (6 if spp == 3 else 0)
    offsets, p = [], data_start
    for s in strips:
        offsets.append(p)
        p += len(s)
    extra += b"".join(struct.pack(e + "I", o) for o in offsets)
    counts_pos = extra_off + len(extra)
    extra += b"".join(struct.pack(e + "I", len(s)) for s in strips)
    bps_pos = extra_off + len(extra)
    if spp == 3:
        extra += struct.pack(e + "HHH", 8, 8, 8)

    def tag(t, typ, cnt, val):
        if typ == 3 and cnt == 1:
            return struct.pack(e + "HHIHH", t, typ, cnt, val, 0)
        return struct.pack(e + "HHII", t, typ, cnt, val)

    phot = photometric if photometric is not None else (2 if spp == 3 else 1)
    tags = [
        tag(256, 4, 1, W), tag(257, 4, 1, H),
        tag(258, 3, 3, bps_pos) if spp == 3 else tag(258, 3, 1, bps),
        tag(259, 3, 1, compression), tag(262, 3, 1, phot),
        tag(273, 4, len(strips), offs_pos) if len(strips) > 1 else tag(273, 4, 1, offsets[0]),
        tag(277, 3, 1, spp), tag(278, 4, 1, rps),
        tag(279, 4, len(strips), counts_pos) if len(strips) > 1 else tag(279, 4, 1, len(strips[0])),
        tag(284, 3, 1, 1),
    ]
    head = (b"MM" if big_endian else b"II") + struct.pack(e + "HI", 42, ifd_off)
    ifd = struct.pack(e + "H", n_tags) + b"".join(tags) + struct.pack(e + "I", 0)
    return head + ifd + extra + b"".join(strips)


def pgm_bytes(img: np.ndarray) -> bytes:
    img = np.asarray(img)
    H, W = img.shape
    maxv = 255 if img.dtype == np.uint8 else 65535
    head = f"P5\n# test image\n{W} {H}\n{maxv}\n".encode()
    return head + img.astype(">u2" if maxv > 255 else np.uint8).tobytes()


def read_tiff(path) -> np.ndarray:
    """Decodes an uncompressed single-strip-or-more 8-bit grayscale TIFF."""
    b = open(path, "rb").read()
    e = "<" if b[:2] == b"II" else ">"
    (magic, ifd) = struct.unpack(e + "HI", b[2:8])
    assert magic == 42
    (n,) = struct.unpack(e + "H", b[ifd:ifd + 2])
    tags = {}
    for i in range(n):
        t, typ, cnt, raw = struct.unpack(e + "HHI4s", b[ifd + 2 + 12 * i: ifd + 14 + 12 * i])
        size = _TYPE_SIZE[typ]
        if cnt * size <= 4:
            data = raw
        else:
            (off,) = struct.unpack(e + "I", raw)
            data = b[off:off + cnt * size]
        tags[t] = list(struct.unpack(e + ("H" if size == 2 else "I") * cnt, data[:cnt * size]))
    W, H = tags[256][0], tags[257][0]
    assert tags.get(259, [1])[0] == 1 and tags.get(258, [8])[0] == 8 and tags.get(277, [1])[0] == 1
    raster = b"".join(b[o:o + c] for o, c in zip(tags[273], tags[279]))
    return np.frombuffer(raster, np.uint8)[:W * H].reshape(H, W).copy()
