"""CPU restatement of the reference's sparse segmentation (the segmentation executable).

TEST INFRASTRUCTURE ONLY: imported by tests/ as the checker of the library's
segmentation functions and of the segmentation executable, and by bench.py's
cpu_baseline leg of the segmentation workload.

- load_voxels: SparseVolume::loadData (src/core/sparse_volume.cu:562-682): the
  "x, y, z, value" header (first 14 characters), lines "%d, %d, %d, %e"; the
  plane counts come from each line's z, then plane z takes the next count[z] lines
  of the file in order; the in-plane index is y*W + x with the reference's bounds
  test (x > W, y > H, z > depth is an error, so x == W wraps to the next row).
  Indices outside the plane or volume and negative coordinates, undefined behaviour
  in the reference, are INVALID_DATA here; so is a malformed line.
- binarize: SparseSegmentation::binarize (sparse_segmentation.cu:39-111) on the
  dense planes: value.x = sqrt(intensity) as loaded, intensity = x*x in float,
  max intensity over the loaded voxels (calcMaxIntensity, sparse_volume.cu:482-516),
  thr = min_intensity * max / 256 in double; binary = intensity > thr, the complex
  data masked to the voxels above thr.  A voxel listed twice keeps its larger
  intensity, NaN ranking lowest (the reference's result depends on its scatter
  order).
- close: simpleMorph dilate then erode (sparse_segmentation.cu:145-315): every
  non-zero source adds 1 to each in-bounds target within the ball sx^2+sy^2+sz^2 <=
  cs^2; erode keeps counts > num_strel - 0.1.  cs <= 0 leaves the volume unchanged.
- label: labelConnectedComponents (:411-532): the 27 neighbours of a voxel are the
  in-plane flat indices i + ys*W + xs (0 <= index < W*H, so rows wrap at the x
  borders) on planes z-1..z+1; the iteration takes each voxel's label down to the
  smallest neighbour label until nothing changes, so every voxel ends with the
  smallest global index z*W*H + i of its component.  Planes -1 and Nz, read out of
  bounds by the reference, have no voxels here.
- extract_objects: ObjectCloud(SparseVolume) + includeIntensity
  (object_cloud.cu:44-58, 127-170, 219-289): objects in order of first appearance in
  the z-major raster, voxels of each in raster order; values = |masked complex|^2.
  Blob3d statistics (blob3d.cu:79-105, 198-330): centroid = int sums / n in float,
  weighted centroid = float sums of w*x in voxel order / float sum of w, max value,
  mean value = double accumulation cast to float / float n, bounds.  The reference's
  cap of 20*1024 objects (object_cloud.h:22; beyond it writes past its arrays) does
  not apply here.
- size_filter: ObjectCloud::blobSizeFilter (object_cloud.cu:313-328): an object with
  fewer voxels than min_area is replaced by the last one, which is tested next.
- centroid_csv: ObjectCloud::writeCentroids (object_cloud.cu:408-470), the three
  formats the executable writes (main_segmentation.cu:112-125).
"""
from __future__ import annotations

import ctypes
import ctypes.util

import numpy as np

INVALID_FILE = 9
INVALID_DATA = 10

_libc = ctypes.CDLL(ctypes.util.find_library("c") or "libc.so.6")
_libc.strtof.restype = ctypes.c_float
_libc.strtof.argtypes = [ctypes.c_char_p, ctypes.POINTER(ctypes.c_char_p)]


class SegError(Exception):
    def __init__(self, code, msg=""):
        super().__init__(f"{code}: {msg}")
        self.code = code


def _strtof(s: str) -> np.float32:
    """C's strtof, as fscanf's %e converts (correctly rounded to float)."""
    return np.float32(_libc.strtof(s.strip().encode(), None))


def load_voxels(text: str, W: int, H: int, Nz: int):
    """-> (plane int32[n], index int32[n], intensity float32[n]) in the order the
    reference stores them (plane by plane, file order within a plane)."""
    if text[:14] != "x, y, z, value":
        raise SegError(INVALID_FILE, "header")
    lines = [ln for ln in text[14:].split("\n") if ln.strip()]
    rows = []
    for ln in lines:
        parts = ln.split(",")
        if len(parts) != 4:
            raise SegError(INVALID_FILE, ln)
        try:
            x, y, z = (int(p) for p in parts[:3])
        except ValueError as e:
            raise SegError(INVALID_FILE, ln) from e
        if x > W or y > H or z > Nz:
            raise SegError(INVALID_DATA, ln)
        if x < 0 or y < 0 or z < 0 or z >= Nz or y * W + x >= W * H:
            raise SegError(INVALID_DATA, ln)
        rows.append((x, y, z, _strtof(parts[3])))
    counts = np.zeros(Nz, np.int64)
    for r in rows:
        counts[r[2]] += 1
    plane, index, inten = [], [], []
    k = 0
    for z in range(Nz):
        for _ in range(counts[z]):
            x, y, _z, v = rows[k]
            k += 1
            plane.append(z)
            index.append(y * W + x)
            inten.append(v)
    return (np.asarray(plane, np.int32), np.asarray(index, np.int32), np.asarray(inten, np.float32))


def binarize(plane, index, intensity, W, H, Nz, min_intensity):
    """-> (binary bool [Nz,H,W], weight float32 [Nz,H,W], max_intensity, thr)."""
    with np.errstate(invalid="ignore"):
        amp = np.sqrt(np.asarray(intensity, np.float32))  # value.x = sqrt(intensity)
    inten = (amp * amp).astype(np.float32)
    # a repeated voxel keeps its larger intensity; NaN (from a negative or NaN value)
    # ranks below every number and above an absent voxel
    key = np.zeros((Nz, H * W), np.int64)
    for z, i, v in zip(plane, index, inten):
        k = 1 if np.isnan(v) else int(np.float32(v).view(np.uint32)) + 2
        key[z, i] = max(key[z, i], k)
    dense = np.where(key == 1, np.uint32(0x7FC00000), np.maximum(key - 2, 0)).astype(np.uint32).view(np.float32)
    dense = np.where(key == 0, np.float32(0), dense)
    mx = 0.0
    for v in inten:
        if float(v) > mx:
            mx = float(v)
    thr = float(min_intensity) * mx / 256.0
    dense = dense.reshape(Nz, H, W)
    with np.errstate(invalid="ignore"):
        binary = dense.astype(np.float64) > thr
    weight = np.where(binary, dense, np.float32(0)).astype(np.float32)
    return binary, weight, mx, thr


def ball_offsets(cs: int):
    return [(sz, sy, sx) for sz in range(-cs, cs + 1) for sx in range(-cs, cs + 1) for sy in range(-cs, cs + 1)
            if sx * sx + sy * sy + sz * sz <= cs * cs]


def _morph_counts(src: np.ndarray, cs: int) -> np.ndarray:
    """counts[t] = number of non-zero sources p with t = p + (sx, sy, -sz) in bounds."""
    Nz, H, W = src.shape
    cnt = np.zeros(src.shape, np.float32)
    nz = src != 0

    def span(d, n):  # (source, target) slices of a shift by d along an axis of length n
        if abs(d) >= n:
            return None
        return slice(max(0, -d), n - max(0, d)), slice(max(0, d), n - max(0, -d))

    for sz, sy, sx in ball_offsets(cs):
        sp = [span(d, n) for d, n in ((-sz, Nz), (sy, H), (sx, W))]  # target - source
        if None in sp:
            continue
        (zs, zt), (ys, yt), (xs, xt) = sp
        cnt[zt, yt, xt] += nz[zs, ys, xs]
    return cnt


def close(binary: np.ndarray, cs: int) -> np.ndarray:
    """-> bool volume after dilate + erode (the counts only matter through != 0)."""
    if cs <= 0:
        return binary.astype(bool).copy()
    num_strel = len(ball_offsets(cs))
    dil = _morph_counts(binary, cs)
    ero = _morph_counts(dil, cs)
    ero = np.where(ero > num_strel - 0.1, ero, 0)
    return ero != 0


NONE = np.int64(-1)


def label(mask: np.ndarray) -> np.ndarray:
    """-> int64 labels [Nz,H,W]: smallest global index of the component, -1 off the mask."""
    Nz, H, W = mask.shape
    P = H * W
    fg = mask.reshape(Nz, P)
    big = np.int64(np.iinfo(np.int64).max)
    L = np.where(fg, np.arange(Nz * P, dtype=np.int64).reshape(Nz, P), big)
    offsets = [ys * W + xs for xs in (-1, 0, 1) for ys in (-1, 0, 1)]
    while True:
        new = L.copy()
        for dz in (-1, 0, 1):
            if abs(dz) >= Nz:
                continue
            for o in offsets:
                if abs(o) >= P:
                    continue
                sh = np.full((Nz, P), big, np.int64)
                zt, zsrc = slice(max(0, -dz), Nz - max(0, dz)), slice(max(0, dz), Nz - max(0, -dz))
                it, isrc = slice(max(0, -o), P - max(0, o)), slice(max(0, o), P - max(0, -o))
                sh[zt, it] = L[zsrc, isrc]  # sh[z, i] = L[z + dz, i + o]
                new = np.minimum(new, sh)
        new = np.where(fg, new, big)
        if np.array_equal(new, L):
            break
        L = new
    return np.where(fg, L, NONE).reshape(Nz, H, W)


def extract_objects(labels: np.ndarray, weight: np.ndarray):
    """-> list of dicts, in order of first appearance in the raster."""
    Nz, H, W = labels.shape
    flat = labels.reshape(-1)
    wflat = weight.reshape(-1).astype(np.float32)
    order = {}
    objs = []
    for g in np.flatnonzero(flat >= 0):
        lab = int(flat[g])
        if lab not in order:
            order[lab] = len(objs)
            objs.append(dict(label=lab, voxels=[], values=[]))
        o = objs[order[lab]]
        o["voxels"].append(int(g))
        o["values"].append(wflat[g])
    out = []
    P = H * W
    for o in objs:
        n = len(o["voxels"])
        xs = [g % P % W for g in o["voxels"]]
        ys = [g % P // W for g in o["voxels"]]
        zs = [g // P for g in o["voxels"]]
        fn = np.float32(n)
        cent = [np.float32(np.float32(sum(c)) / fn) for c in (xs, ys, zs)]
        with np.errstate(invalid="ignore", divide="ignore"):
            sx = sy = sz = sw = np.float32(0)
            for x, y, z, w in zip(xs, ys, zs, o["values"]):
                sx = np.float32(sx + np.float32(w * np.float32(x)))
                sy = np.float32(sy + np.float32(w * np.float32(y)))
                sz = np.float32(sz + np.float32(w * np.float32(z)))
                sw = np.float32(sw + w)
            wcent = [np.float32(sx / sw), np.float32(sy / sw), np.float32(sz / sw)]
        acc = 0.0
        for w in o["values"]:
            acc += float(w)
        mean = np.float32(np.float32(acc) / fn)
        out.append(dict(label=o["label"], num_voxels=n, centroid=cent, weighted_centroid=wcent,
                        max_value=max(o["values"]), mean_value=mean,
                        bounds=[min(xs), max(xs), min(ys), max(ys), min(zs), max(zs)]))
    return out


def size_filter(objects, min_area):
    objs = list(objects)
    n, removed = 0, 0
    while n < len(objs):
        if objs[n]["num_voxels"] < min_area:
            objs[n] = objs[-1]
            objs.pop()
            removed += 1
        else:
            n += 1
    return objs, removed


def fmt_f(x) -> str:
    x = np.float32(x)
    if np.isnan(x):
        return "-nan" if np.signbit(x) else "nan"
    return "%f" % float(x)


def fmt_e(x) -> str:
    x = np.float32(x)
    if np.isnan(x):
        return "-nan" if np.signbit(x) else "nan"
    return "%e" % float(x)


def centroid_csv(objects, weighted=False, verbose=False) -> str:
    if not verbose:
        lines = ["x (pixels), y (pixels), z (planes), # of voxels"]
        for o in objects:
            c = o["weighted_centroid"] if weighted else o["centroid"]
            lines.append("%s, %s, %s, %d" % (fmt_f(c[0]), fmt_f(c[1]), fmt_f(c[2]), o["num_voxels"]))
    else:
        lines = ["x (pixels), y (pixels), z (planes), weighted x, weighted y, weighted z, # of voxels"
                 ", max value, mean value, xmin, xmax, ymin, ymax, zmin, zmax"]
        for o in objects:
            c, w = o["centroid"], o["weighted_centroid"]
            lines.append(", ".join([fmt_f(v) for v in c] + [fmt_f(v) for v in w] + ["%d" % o["num_voxels"]]
                                   + [fmt_e(o["max_value"]), fmt_e(o["mean_value"])]
                                   + ["%d" % b for b in o["bounds"]]))
    return "\n".join(lines) + "\n"


def segment(plane, index, intensity, W, H, Nz, min_intensity=1.0, close_size=3, min_area=2):
    """The executable's chain (main_segmentation.cu:101-125) after loading."""
    binary, weight, mx, thr = binarize(plane, index, intensity, W, H, Nz, min_intensity)
    closed = close(binary, close_size)
    labels = label(closed)
    objs = extract_objects(labels, weight)
    kept, removed = size_filter(objs, min_area)
    return dict(binary=binary, weight=weight, max_intensity=mx, thr=thr, closed=closed, labels=labels,
                objects=objs, kept=kept, removed=removed)
