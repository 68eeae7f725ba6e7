"""CPU tests of the segmentation layer: the oracle's restatement of the reference's
segmentation chain pinned by library routines (scipy.ndimage's binary closing and
26-connected labelling), closed forms and hand-worked cases of the reference's quirks;
the sparse-volume reader and the 5009-5011 parameters of the C ABI against it; the
layout of rihvr_object; and the segmentation executable's error paths."""
from __future__ import annotations

import ctypes
import pathlib
import subprocess

import numpy as np
import pytest
import scipy.ndimage as ndi

ROOT = pathlib.Path(__file__).resolve().parent.parent


@pytest.fixture(scope="module")
def so():
    import oracle.segmentation as s
    return s


def ball(cs):
    se = np.zeros((2 * cs + 1,) * 3, bool)
    for z in range(-cs, cs + 1):
        for y in range(-cs, cs + 1):
            for x in range(-cs, cs + 1):
                se[z + cs, y + cs, x + cs] = x * x + y * y + z * z <= cs * cs
    return se


# ---- close: scipy's binary closing with border value 0 -----------------------------
@pytest.mark.parametrize("cs", [0, 1, 2, 3, 4])
def test_close_is_scipy_dilation_then_erosion(so, cs):
    rng = np.random.default_rng(100 + cs)
    for trial in range(6):
        Nz, H, W = (int(v) for v in rng.integers(3, 14, 3))
        b = rng.random((Nz, H, W)) < rng.uniform(0.02, 0.35)
        got = so.close(b, cs)
        if cs == 0:
            ref = b
        else:
            se = ball(cs)
            ref = ndi.binary_erosion(ndi.binary_dilation(b, se, border_value=0), se, border_value=0)
        np.testing.assert_array_equal(got, ref)


def test_ball_sizes(so):
    # number of lattice points in the ball of radius r (OEIS A000605): 1, 7, 33, 123, 257
    assert [len(so.ball_offsets(r)) for r in range(5)] == [1, 7, 33, 123, 257]


def test_close_removes_border_voxels_and_fills_gaps(so):
    # a ball closes a one-voxel gap in a bar but erodes everything within cs of the border
    b = np.zeros((9, 9, 20), bool)
    b[3:6, 3:6, 2:9] = True
    b[3:6, 3:6, 10:18] = True
    c = so.close(b, 1)
    assert c[4, 4, 9]                     # gap filled
    assert not c[:, :, 0].any() and not c[0].any() and not c[-1].any()
    assert c[4, 4, 2:18].all()


# ---- label: scipy's 26-connected labelling away from the x borders ------------------
def test_label_matches_scipy_26_connectivity(so):
    rng = np.random.default_rng(7)
    for trial in range(12):
        Nz, H, W = (int(v) for v in rng.integers(2, 12, 3) + np.array([0, 0, 2]))
        m = rng.random((Nz, H, W)) < rng.uniform(0.05, 0.4)
        m[:, :, 0] = m[:, :, -1] = False     # rows cannot wrap
        L = so.label(m)
        lab, n = ndi.label(m, np.ones((3, 3, 3), bool))
        flat_l, flat_s = L.reshape(-1), lab.reshape(-1)
        assert np.all(flat_l[flat_s == 0] == -1)
        for k in range(1, n + 1):
            idx = np.flatnonzero(flat_s == k)
            assert np.all(flat_l[idx] == idx.min())


def test_label_rows_wrap_at_the_x_borders(so):
    """The reference's neighbours are flat in-plane indices i + dy*W + dx: (W-1, y) touches
    (0, y+1) (offset +1), (0, y) touches (W-1, y) (offset W-1) and (W-1, y-2) (offset -W-1),
    and the same across planes."""
    W, H = 6, 5
    cases = [((0, 1, W - 1), (0, 2, 0), True),
             ((0, 2, 0), (0, 2, W - 1), True),
             ((0, 3, 0), (0, 1, W - 1), True),
             ((0, 1, W - 1), (1, 2, 0), True),
             ((0, 0, 0), (0, 0, W - 1), True),      # flat offset W-1 from (0, 0) is (W-1, 0)
             ((0, 0, 0), (0, 3, 0), False),
             ((0, H - 1, W - 1), (1, 0, 0), False)]  # in-plane index P is outside the plane
    for a, b, joined in cases:
        m = np.zeros((2, H, W), bool)
        m[a] = m[b] = True
        L = so.label(m)
        assert (L[a] == L[b]) == joined, (a, b)
        assert L[a] == min(np.ravel_multi_index(a, m.shape), np.ravel_multi_index(b, m.shape)) or not joined


def test_label_is_the_smallest_index_of_each_component(so):
    m = np.zeros((3, 6, 7), bool)
    m[2, 1, 1] = m[1, 2, 2] = m[0, 3, 3] = True   # a diagonal chain through the planes
    m[0, 5, 5] = True
    L = so.label(m)
    assert L[2, 1, 1] == L[1, 2, 2] == L[0, 3, 3] == np.ravel_multi_index((0, 3, 3), m.shape)
    assert L[0, 5, 5] == np.ravel_multi_index((0, 5, 5), m.shape)


# ---- binarize -------------------------------------------------------------------
def test_binarize_threshold_and_reloaded_intensity(so):
    plane = np.array([0, 0, 1], np.int32)
    index = np.array([3, 5, 7], np.int32)
    inten = np.array([4.0, 1.0, 2.0], np.float32)
    b, w, mx, thr = so.binarize(plane, index, inten, 4, 3, 2, 64.0)
    assert mx == 4.0 and thr == 1.0                   # 64 * 4 / 256
    assert b.sum() == 2 and b[0, 0, 3] and b[1, 1, 3] and not b[0, 1, 1]   # 1.0 > 1.0 is false
    r2 = np.float32(np.sqrt(np.float32(2.0))) ** 2    # sqrt(2)^2 in float is not 2
    assert r2 != np.float32(2.0) and w[1, 1, 3] == r2 and w[0, 0, 3] == 4.0
    # a negative threshold makes every voxel of the volume an object voxel
    b, w, mx, thr = so.binarize(plane, index, inten, 4, 3, 2, -1.0)
    assert thr < 0 and b.all() and w[0, 0, 0] == 0


def test_binarize_repeated_voxel_keeps_larger(so):
    b, w, mx, thr = so.binarize(np.array([0, 0]), np.array([1, 1]), np.array([9.0, 16.0], np.float32), 2, 1, 1, 1.0)
    assert w[0, 0, 1] == 16.0 and mx == 16.0


# ---- objects ----------------------------------------------------------------------
def test_objects_closed_forms(so):
    lab = np.full((6, 8, 9), -1, np.int64)
    wgt = np.zeros((6, 8, 9), np.float32)
    lab[1:4, 2:5, 3:6] = 1          # a 3x3x3 cube, centre (x, y, z) = (4, 3, 2)
    lab[0, 7, 0] = 5                # a single voxel that comes first in z-major order
    wgt[1:4, 2:5, 3:6] = 2.0
    wgt[0, 7, 0] = 3.0
    objs = so.extract_objects(lab, wgt)
    assert [o["num_voxels"] for o in objs] == [1, 27]
    cube = objs[1]
    assert cube["centroid"] == [4.0, 3.0, 2.0] and cube["weighted_centroid"] == [4.0, 3.0, 2.0]
    assert cube["bounds"] == [3, 5, 2, 4, 1, 3] and cube["max_value"] == 2.0 and cube["mean_value"] == 2.0
    assert objs[0]["centroid"] == [0.0, 7.0, 0.0] and objs[0]["max_value"] == 3.0


def test_weighted_centroid_against_float64(so):
    rng = np.random.default_rng(3)
    lab = np.full((4, 10, 12), -1, np.int64)
    sel = rng.random(lab.shape) < 0.4
    lab[sel] = 0
    wgt = (rng.random(lab.shape) * sel).astype(np.float32)
    o = so.extract_objects(lab, wgt)[0]
    z, y, x = np.nonzero(sel)
    w = wgt[sel].astype(np.float64)
    ref = [(w * c).sum() / w.sum() for c in (x, y, z)]
    np.testing.assert_allclose(o["weighted_centroid"], ref, rtol=1e-5)
    np.testing.assert_allclose(o["centroid"], [x.mean(), y.mean(), z.mean()], rtol=1e-6)
    np.testing.assert_allclose(o["mean_value"], w.mean(), rtol=1e-6)
    assert o["num_voxels"] == sel.sum()


def test_size_filter_swaps_with_last(so):
    objs = [dict(num_voxels=n, tag=t) for t, n in enumerate([1, 5, 1, 7, 2])]
    kept, removed = so.size_filter(objs, 2)
    assert [o["num_voxels"] for o in kept] == [2, 5, 7] and removed == 2
    assert [o["tag"] for o in kept] == [4, 1, 3]


def test_centroid_csv_formats(so):
    o = dict(centroid=[np.float32(1.5), np.float32(2.0), np.float32(0.25)],
             weighted_centroid=[np.float32(1.25), np.float32(2.0), np.float32(0.5)], num_voxels=7,
             max_value=np.float32(3.0), mean_value=np.float32(0.125), bounds=[1, 2, 3, 4, 0, 1])
    assert so.centroid_csv([o]) == "x (pixels), y (pixels), z (planes), # of voxels\n1.500000, 2.000000, 0.250000, 7\n"
    assert so.centroid_csv([o], weighted=True).endswith("1.250000, 2.000000, 0.500000, 7\n")
    v = so.centroid_csv([o], verbose=True).split("\n")
    assert v[0] == ("x (pixels), y (pixels), z (planes), weighted x, weighted y, weighted z, # of voxels, "
                    "max value, mean value, xmin, xmax, ymin, ymax, zmin, zmax")
    assert v[1] == ("1.500000, 2.000000, 0.250000, 1.250000, 2.000000, 0.500000, 7, 3.000000e+00, "
                    "1.250000e-01, 1, 2, 3, 4, 0, 1")
    nan = np.float32(np.nan)
    assert so.fmt_f(-nan) == "-nan" and so.fmt_f(nan) == "nan"   # glibc prints the NaN's sign


# ---- the sparse volume reader --------------------------------------------------------
SPARSE = """x, y, z, value
1, 0, 1, 2.500000e+00
3, 2, 0, 1.000000e+00
4, 1, 0, 7.000000e-01
0, 3, 1, 4.000000e+00
"""


def test_load_voxels_plane_assignment_follows_counts(so):
    """Plane z takes the next count[z] rows: a file out of z order keeps row order."""
    plane, index, inten = so.load_voxels(SPARSE, 4, 4, 2)
    # counts: z=0 has 2 rows, z=1 has 2 rows -> rows 0,1 go to plane 0, rows 2,3 to plane 1
    assert plane.tolist() == [0, 0, 1, 1]
    assert index.tolist() == [1, 2 * 4 + 3, 1 * 4 + 4, 3 * 4]   # x == W wraps to the next row
    assert inten.tolist() == [np.float32(2.5), 1.0, np.float32(0.7), 4.0]


@pytest.mark.parametrize("text,code", [
    ("x,y,z,value\n1, 1, 0, 1e0\n", 9),
    ("x, y, z, value\n5, 1, 0, 1e0\n", 10),       # x > W
    ("x, y, z, value\n1, 1, 3, 1e0\n", 10),       # z > depth
    ("x, y, z, value\n1, 1, 2, 1e0\n", 10),       # z == depth: stored out of range by the reference
    ("x, y, z, value\n0, 4, 0, 1e0\n", 10),       # index >= W*H
    ("x, y, z, value\n-1, 1, 0, 1e0\n", 10),
    ("x, y, z, value\n1, 1, 0\n", 9),
])
def test_load_voxels_errors(so, pkg, tmp_path, text, code):
    with pytest.raises(so.SegError) as e:
        so.load_voxels(text, 4, 4, 2)
    assert e.value.code == code
    p = tmp_path / "v.csv"
    p.write_text(text)
    with pytest.raises(pkg.RihvrError) as e2:
        pkg.read_sparse_voxels(p, 4, 4, 2)
    assert e2.value.status == code


def test_read_sparse_voxels_matches_oracle(so, pkg, tmp_path):
    rng = np.random.default_rng(11)
    W, H, Nz = 13, 9, 5
    rows = []
    for _ in range(200):
        x, y, z = int(rng.integers(0, W + 1)), int(rng.integers(0, H - 1)), int(rng.integers(0, Nz))
        rows.append("%d, %d, %d, %e" % (x, y, z, float(rng.random() * 10.0 ** int(rng.integers(-3, 3)))))
    text = "x, y, z, value\n" + "\n".join(rows) + "\n"
    p = tmp_path / "v.csv"
    p.write_text(text)
    got = pkg.read_sparse_voxels(p, W, H, Nz)
    ref = so.load_voxels(text, W, H, Nz)
    for g, r in zip(got, ref):
        np.testing.assert_array_equal(g, r)
    (tmp_path / "e.csv").write_text("x, y, z, value\n")
    assert all(len(a) == 0 for a in pkg.read_sparse_voxels(tmp_path / "e.csv", W, H, Nz))
    with pytest.raises(pkg.RihvrError) as e:
        pkg.read_sparse_voxels(tmp_path / "missing.csv", W, H, Nz)
    assert e.value.status == 2


# ---- parameters and layout ------------------------------------------------------------
def test_segment_parameters(pkg, tmp_path):
    p = tmp_path / "p.txt"
    p.write_text("num_planes 3003 4\n")
    fp, rc = pkg.read_param_file(p)
    assert rc == 0 and (fp.segment_min_area, fp.segment_min_intensity, fp.segment_close_size) == (2, 1.0, 3)
    p.write_text("segment_min_area 5009 12\nsegment_min_intensity 5010 2.5\nsegment_close_size 5011 0\n")
    fp, rc = pkg.read_param_file(p)
    assert rc == 0 and (fp.segment_min_area, fp.segment_min_intensity, fp.segment_close_size) == (12, 2.5, 0)


def test_object_layout_matches_c(pkg, tmp_path):
    src = tmp_path / "sz.c"
    src.write_text('#include <stddef.h>\n#include <stdio.h>\n#include "rihvr.h"\n'
                   'int main(void){printf("%zu %zu %zu %zu %zu", sizeof(rihvr_object),'
                   ' offsetof(rihvr_object, num_voxels), offsetof(rihvr_object, label),'
                   ' offsetof(rihvr_object, bounds), offsetof(rihvr_file_params, segment_min_intensity));'
                   'return 0;}\n')
    exe = tmp_path / "sz"
    subprocess.run(["gcc", "-std=c99", f"-I{ROOT / 'include'}", str(src), "-o", str(exe)], check=True, timeout=60)
    out = [int(v) for v in subprocess.run([str(exe)], capture_output=True, text=True, check=True).stdout.split()]
    O = pkg.Object
    assert out == [ctypes.sizeof(O), O.num_voxels.offset, O.label.offset, O.bounds.offset,
                   pkg.FileParams.segment_min_intensity.offset]


# ---- the executable's error paths (no device work is reached) --------------------------
def _exe(pkg):
    return pkg.LIB_PATH.parent / "segmentation"


def test_segmentation_usage_and_missing_inputs(pkg, tmp_path):
    r = subprocess.run([str(_exe(pkg))], capture_output=True, text=True, timeout=60)
    assert r.returncode == 255 and "User must give input parameter file using -F flag" in r.stdout
    p = tmp_path / "p.txt"
    p.write_text(f"input_fn 2001 {tmp_path}/img_%04d.tif\noutput_pathn 2004 {tmp_path}\n")
    r = subprocess.run([str(_exe(pkg)), "-F", str(p)], capture_output=True, text=True, timeout=60)
    assert r.returncode == 255
    assert f"Unable to read file: {tmp_path}/img_0000.tif" in r.stdout and "Detected HoloError: 2" in r.stdout
