"""GPU parity of the segmentation chain (librihvr's rihvr_segmentation_* through the C
ABI) against the oracle's restatement of the reference's SparseSegmentation /
ObjectCloud / Blob3d: every stage bit for bit (binary volume, values, closed volume,
labels, objects, centroid files) on seeded volumes; the dense-volume loader; edge cases
(empty input, a negative threshold, repeated voxels, bad entries, call order); the
segmentation executable end to end on a reconstruction's voxel file; and, at
1024 x 1024 x 128, size-independent properties of planted spheres."""
from __future__ import annotations

import subprocess

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def so():
    import oracle.segmentation as s
    return s


def blob_volume(shape, n_blobs, seed, noise=0.002):
    """A sparse voxel list: balls of random radius and intensity plus scattered voxels."""
    Nz, H, W = shape
    rng = np.random.default_rng(seed)
    inten = np.zeros(shape, np.float32)
    zz, yy, xx = np.mgrid[0:Nz, 0:H, 0:W]
    for _ in range(n_blobs):
        c = [rng.uniform(0, n) for n in shape]
        r = rng.uniform(0.8, 3.0)
        d2 = (zz - c[0]) ** 2 + (yy - c[1]) ** 2 + (xx - c[2]) ** 2
        inten = np.maximum(inten, np.where(d2 <= r * r, rng.uniform(50, 255) * np.exp(-d2 / (r * r)), 0)
                           .astype(np.float32))
    sel = rng.random(shape) < noise
    inten[sel] = np.maximum(inten[sel], rng.uniform(0, 80, sel.sum()).astype(np.float32))
    z, y, x = np.nonzero(inten)
    order = rng.permutation(len(z))  # the list order does not matter
    return (z[order].astype(np.int32), (y[order] * W + x[order]).astype(np.int32),
            inten[z, y, x][order].astype(np.float32))


SHAPES = [(9, 24, 31), (6, 40, 64), (5, 13, 300), (4, 20, 600)]


def run_chain(pkg, shape, vox, min_intensity, cs):
    Nz, H, W = shape
    seg = pkg.SparseSegmentation(W, H, Nz, voxels=vox)
    mx, thr = seg.binarize(min_intensity)
    binary, values = seg.mask(), seg.values()
    seg.close(cs)
    closed = seg.mask()
    seg.labelConnectedComponents()
    labels = seg.labels()
    cloud = seg.extractObjects()
    return seg, cloud, dict(mx=mx, thr=thr, binary=binary, values=values, closed=closed, labels=labels)


def same_float(a, b):
    a, b = np.float32(a), np.float32(b)
    return a.view(np.uint32) == b.view(np.uint32)


def assert_objects_equal(got, ref):
    assert len(got) == len(ref)
    for g, r in zip(got, ref):
        assert g["num_voxels"] == r["num_voxels"] and g["label"] == r["label"] and g["bounds"] == r["bounds"]
        for k in ("centroid", "weighted_centroid"):
            assert all(same_float(a, b) for a, b in zip(g[k], r[k])), (k, g[k], r[k])
        assert same_float(g["max_value"], r["max_value"]) and same_float(g["mean_value"], r["mean_value"])


@pytest.mark.parametrize("shape", SHAPES)
@pytest.mark.parametrize("cs", [0, 1, 3])
def test_chain_matches_oracle(pkg, so, gpu_available, shape, cs):
    Nz, H, W = shape
    vox = blob_volume(shape, 10, seed=hash((shape, cs)) % 1000)
    print(f"segmentation chain {shape} cs={cs} n={len(vox[0])}", flush=True)
    seg, cloud, st = run_chain(pkg, shape, vox, 1.0, cs)
    ref = so.segment(*vox, W, H, Nz, min_intensity=1.0, close_size=cs, min_area=2)
    assert st["mx"] == ref["max_intensity"] and st["thr"] == ref["thr"]
    np.testing.assert_array_equal(st["binary"], ref["binary"].astype(np.uint8))
    np.testing.assert_array_equal(st["values"].view(np.uint32), ref["weight"].view(np.uint32))
    np.testing.assert_array_equal(st["closed"], ref["closed"].astype(np.uint8))
    np.testing.assert_array_equal(st["labels"].astype(np.int64), np.where(ref["labels"] < 0, 0xFFFFFFFF, ref["labels"]))
    assert_objects_equal(cloud.objects(), ref["objects"])
    removed = cloud.blobSizeFilter(2)
    assert removed == ref["removed"]
    assert_objects_equal(cloud.objects(), ref["kept"])
    seg.destroy()


def test_centroid_files_match_oracle(pkg, so, gpu_available, tmp_path):
    shape = (8, 48, 56)
    Nz, H, W = shape
    vox = blob_volume(shape, 14, seed=5)
    seg, cloud, st = run_chain(pkg, shape, vox, 2.0, 2)
    cloud.blobSizeFilter(3)
    ref = so.segment(*vox, W, H, Nz, min_intensity=2.0, close_size=2, min_area=3)
    for weighted, verbose in ((False, False), (True, False), (False, True)):
        f = tmp_path / f"c_{weighted}_{verbose}.csv"
        cloud.writeCentroids(f, weighted, verbose)
        assert f.read_text() == so.centroid_csv(ref["kept"], weighted, verbose)
    seg.destroy()


def test_thresholds_and_sizes(pkg, so, gpu_available):
    shape = (7, 30, 33)
    Nz, H, W = shape
    vox = blob_volume(shape, 12, seed=77, noise=0.01)
    for mi in (0.0, 0.5, 20.0, 300.0):
        seg, cloud, st = run_chain(pkg, shape, vox, mi, 3)
        ref = so.segment(*vox, W, H, Nz, min_intensity=mi, close_size=3, min_area=4)
        np.testing.assert_array_equal(st["closed"], ref["closed"].astype(np.uint8))
        assert_objects_equal(cloud.objects(), ref["objects"])
        cloud.blobSizeFilter(4)
        assert_objects_equal(cloud.objects(), ref["kept"])
        seg.destroy()


def test_rows_wrap_like_the_reference(pkg, so, gpu_available):
    """Object voxels on the x borders join across rows as the reference's flat offsets do."""
    shape = (3, 6, 8)
    Nz, H, W = shape
    pts = [(0, 1, W - 1), (0, 2, 0), (1, 4, 0), (1, 4, W - 1), (2, 5, W - 1), (0, 5, W - 1), (1, 0, 0)]
    z = np.array([p[0] for p in pts], np.int32)
    idx = np.array([p[1] * W + p[2] for p in pts], np.int32)
    vox = (z, idx, np.full(len(pts), 100.0, np.float32))
    seg, cloud, st = run_chain(pkg, shape, vox, 1.0, 0)
    ref = so.segment(*vox, W, H, Nz, min_intensity=1.0, close_size=0, min_area=1)
    np.testing.assert_array_equal(st["labels"].astype(np.int64), np.where(ref["labels"] < 0, 0xFFFFFFFF, ref["labels"]))
    assert_objects_equal(cloud.objects(), ref["objects"])
    seg.destroy()


def test_dense_loader(pkg, so, gpu_available):
    shape = (6, 32, 40)
    Nz, H, W = shape
    rng = np.random.default_rng(9)
    vol = np.zeros(shape, np.complex64)
    sel = rng.random(shape) < 0.03
    vol[sel] = (rng.standard_normal(sel.sum()) + 1j * rng.standard_normal(sel.sum())).astype(np.complex64)
    vol[2:4, 10:14, 10:15] = 3 + 4j
    seg = pkg.SparseSegmentation(W, H, Nz, dense=vol)
    mx, thr = seg.binarize(30.0)
    values = seg.values()
    seg.close(1)
    seg.labelConnectedComponents()
    cloud = seg.extractObjects()
    z, y, x = np.nonzero(vol)
    v = vol[z, y, x]
    inten = (v.real * v.real + v.imag * v.imag).astype(np.float32)
    ref = so.segment(z.astype(np.int32), (y * W + x).astype(np.int32), inten, W, H, Nz, min_intensity=30.0,
                     close_size=1)
    assert mx == ref["max_intensity"] and thr == ref["thr"]
    np.testing.assert_array_equal(values.view(np.uint32), ref["weight"].view(np.uint32))
    assert_objects_equal(cloud.objects(), ref["objects"])
    seg.destroy()


def test_dense_loads_reuse_odd_unaligned(pkg, so, gpu_available):
    # one context, three loads: an odd voxel count (the tail of the two-voxel loads), a device
    # volume starting 8 bytes past a 16-byte boundary (one voxel per load), and a sparser
    # volume after a denser one (the keys of the rows the last load flagged are cleared)
    import torch
    shape = (3, 7, 9)
    Nz, H, W = shape
    seg = pkg.SparseSegmentation(W, H, Nz)
    rng = np.random.default_rng(21)
    for k, (density, unaligned) in enumerate([(0.3, False), (0.3, True), (0.02, False)]):
        vol = np.zeros(shape, np.complex64)
        sel = rng.random(shape) < density
        vol[sel] = (rng.standard_normal(sel.sum()) + 1j * rng.standard_normal(sel.sum())).astype(np.complex64)
        if unaligned:
            buf = torch.zeros(vol.size + 1, dtype=torch.complex64, device="cuda")
            buf[1:] = torch.from_numpy(vol.ravel()).cuda()
            dev = buf[1:].view(shape)
            assert dev.data_ptr() % 16 == 8
            seg.loadDense(dev)
        else:
            seg.loadDense(vol)
        mx, thr = seg.binarize(10.0)
        values = seg.values()
        z, y, x = np.nonzero(vol)
        v = vol[z, y, x]
        inten = (v.real * v.real + v.imag * v.imag).astype(np.float32)
        ref = so.segment(z.astype(np.int32), (y * W + x).astype(np.int32), inten, W, H, Nz, min_intensity=10.0,
                         close_size=1)
        assert mx == ref["max_intensity"] and thr == ref["thr"], k
        np.testing.assert_array_equal(values.view(np.uint32), ref["weight"].view(np.uint32))
    seg.destroy()


def test_two_closings(pkg, so, gpu_available):
    # the first closing after binarize skips the rows binarize found empty; a second closing
    # may see bits in such rows and must read every row
    shape = (8, 24, 70)
    Nz, H, W = shape
    rng = np.random.default_rng(5)
    vol = np.zeros(shape, np.complex64)
    sel = rng.random(shape) < 0.02
    vol[sel] = (rng.standard_normal(sel.sum()) + 1j * rng.standard_normal(sel.sum())).astype(np.complex64)
    vol[3, 5, 10:60:3] = 4 + 0j  # dotted lines with empty rows between them
    vol[3, 9, 10:60:3] = 4 + 0j
    seg = pkg.SparseSegmentation(W, H, Nz, dense=vol)
    seg.binarize(5.0)
    seg.close(2)
    seg.close(2)
    z, y, x = np.nonzero(vol)
    v = vol[z, y, x]
    inten = (v.real * v.real + v.imag * v.imag).astype(np.float32)
    binary, _w, _m, _t = so.binarize(z.astype(np.int32), (y * W + x).astype(np.int32), inten, W, H, Nz, 5.0)
    exp = so.close(so.close(binary, 2), 2)
    np.testing.assert_array_equal(seg.mask().reshape(shape) != 0, exp)
    seg.destroy()


def test_device_lists(pkg, so, gpu_available):
    import torch
    shape = (5, 20, 24)
    Nz, H, W = shape
    vox = blob_volume(shape, 6, seed=3)
    seg = pkg.SparseSegmentation(W, H, Nz)
    seg.load(*(torch.from_numpy(a).cuda() for a in vox))
    seg.binarize(1.0)
    seg.close(2)
    seg.labelConnectedComponents()
    got = seg.extractObjects().objects()
    ref = so.segment(*vox, W, H, Nz, min_intensity=1.0, close_size=2)
    assert_objects_equal(got, ref["objects"])
    seg.destroy()


def test_edge_cases(pkg, so, gpu_available):
    W, H, Nz = 10, 8, 3
    empty = (np.zeros(0, np.int32), np.zeros(0, np.int32), np.zeros(0, np.float32))
    seg, cloud, st = run_chain(pkg, (Nz, H, W), empty, 1.0, 3)
    assert st["mx"] == 0.0 and cloud.getNumObjects() == 0 and not st["closed"].any()
    seg.destroy()
    # a negative threshold: every voxel of the volume belongs to the one object
    vox = (np.array([1], np.int32), np.array([17], np.int32), np.array([5.0], np.float32))
    seg, cloud, st = run_chain(pkg, (Nz, H, W), vox, -1.0, 0)
    objs = cloud.objects()
    ref = so.segment(*vox, W, H, Nz, min_intensity=-1.0, close_size=0)
    assert len(objs) == 1 and objs[0]["num_voxels"] == W * H * Nz
    assert_objects_equal(objs, ref["objects"])
    seg.destroy()
    # repeated voxels keep the larger intensity; NaN ranks lowest
    vox = (np.array([0, 0, 1, 1], np.int32), np.array([5, 5, 9, 9], np.int32),
           np.array([4.0, 9.0, np.nan, 16.0], np.float32))
    seg, cloud, st = run_chain(pkg, (Nz, H, W), vox, 1.0, 0)
    ref = so.segment(*vox, W, H, Nz, min_intensity=1.0, close_size=0)
    np.testing.assert_array_equal(st["values"].view(np.uint32), ref["weight"].view(np.uint32))
    assert st["values"][0, 0, 5] == 9.0 and st["values"][1, 0, 9] == 16.0
    seg.destroy()


def test_errors(pkg, gpu_available):
    W, H, Nz = 10, 8, 3
    with pkg.SparseSegmentation(W, H, Nz) as seg:
        with pytest.raises(pkg.RihvrError) as e:
            seg.binarize(1.0)
        assert e.value.status == 4                      # nothing loaded
        with pytest.raises(pkg.RihvrError) as e:
            seg.load(np.array([3], np.int32), np.array([0], np.int32), np.array([1.0], np.float32))
        assert e.value.status == 10                     # plane outside the volume
        with pytest.raises(pkg.RihvrError) as e:
            seg.load(np.array([0], np.int32), np.array([W * H], np.int32), np.array([1.0], np.float32))
        assert e.value.status == 10
        seg.load(np.array([0], np.int32), np.array([3], np.int32), np.array([1.0], np.float32))
        with pytest.raises(pkg.RihvrError) as e:
            seg.labelConnectedComponents()
        assert e.value.status == 4
        seg.binarize(1.0)
        with pytest.raises(pkg.RihvrError) as e:
            seg.close(64)
        assert e.value.status == 7
        with pytest.raises(pkg.RihvrError) as e:
            seg.binarize(1.0)
        assert e.value.status == 4                      # binarize consumes the load
    with pytest.raises(pkg.RihvrError) as e:
        pkg.SparseSegmentation(1 << 16, 1 << 16, 1)
    assert e.value.status == 7


def test_executable_end_to_end(pkg, so, scenes, gpu_available, tmp_path):
    """sparse-inverse-recon's voxel file of a simulated scene, then segmentation -F."""
    from imgfiles import tiff_bytes
    W = H = 64
    Nz = 8
    cfg = dict(width=W, height=H, num_planes=Nz, start_plane=1000.0, plane_stepsize=100.0, wavelength=0.633,
               resolution=2.0, num_inverse_iterations=20, regularization_param=0.02, regularization_TV_param=0.01,
               num_tv_iterations=5)
    opt = scenes.optics_of(cfg)
    masks = np.zeros((Nz, H, W), np.float32)
    masks[2, 20:25, 18:23] = 1.0
    masks[5, 40:44, 44:47] = 1.0
    masks[6, 10:12, 50:52] = 1.0
    holo = scenes.hologram_from_masks(masks, opt, 0.0, np.random.default_rng(1))
    with pkg.SparseCompressiveHolo(pkg.make_params(**cfg)) as rec:
        rec.inverseReconstruct(holo)
        rec.saveSparse(tmp_path / "inverse_voxels_0007.csv")
    img = np.clip(np.round(128 + 40 * holo.real), 0, 255).astype(np.uint8)
    (tmp_path / "img_0007.tif").write_bytes(tiff_bytes(img))
    p = tmp_path / "params.txt"
    p.write_text(f"start_image 1002 7\ninput_fn 2001 {tmp_path}/img_%04d.tif\noutput_pathn 2004 {tmp_path}\n"
                 f"num_planes 3003 {Nz}\nzero_padding 3010 0\nsegment_min_area 5009 3\n"
                 f"segment_min_intensity 5010 8\nsegment_close_size 5011 1\n")
    exe = pkg.LIB_PATH.parent / "segmentation"
    r = subprocess.run([str(exe), "-F", str(p)], capture_output=True, text=True, timeout=120)
    print(r.stdout[-2000:], r.stderr[-2000:], flush=True)
    assert r.returncode == 0
    text = (tmp_path / "inverse_voxels_0007.csv").read_text()
    vox = so.load_voxels(text, W, H, Nz)
    ref = so.segment(*vox, W, H, Nz, min_intensity=8.0, close_size=1, min_area=3)
    assert len(ref["kept"]) >= 1
    assert f"max intensity = {ref['max_intensity']:f}, thr = {ref['thr']:f}" in r.stdout
    assert f"Number of particles: {len(ref['objects'])}" in r.stdout
    assert (tmp_path / "centroids_0007.csv").read_text() == so.centroid_csv(ref["kept"])
    assert (tmp_path / "centroids_weighted_0007.csv").read_text() == so.centroid_csv(ref["kept"], weighted=True)
    assert (tmp_path / "centroids_verbose_0007.csv").read_text() == so.centroid_csv(ref["kept"], verbose=True)


def test_planted_spheres_full_size(pkg, so, gpu_available):
    """1024 x 1024 x 128: 300 separated balls of radius 3 with equal intensity come out as
    300 objects with the ball's closed voxel count and its exact centre."""
    W = H = 1024
    Nz = 128
    r, cs = 3, 3
    # the closed shape of one ball, from the oracle on a small volume (translation invariant)
    small = 21
    zz, yy, xx = np.mgrid[0:small, 0:small, 0:small] - small // 2
    ballm = (zz ** 2 + yy ** 2 + xx ** 2) <= r * r
    z0, y0, x0 = np.nonzero(ballm)
    ref1 = so.segment(z0.astype(np.int32), (y0 * small + x0).astype(np.int32), np.full(len(z0), 100, np.float32),
                      small, small, small, min_intensity=1.0, close_size=cs)
    assert len(ref1["objects"]) == 1
    n_closed = ref1["objects"][0]["num_voxels"]
    rng = np.random.default_rng(2026)
    centres = set()
    grid = [(z, y, x) for z in range(12, Nz - 12, 26) for y in range(20, H - 20, 60) for x in range(20, W - 20, 60)]
    pick = rng.choice(len(grid), 300, replace=False)
    planes, idx = [], []
    for k in pick:
        cz, cy, cx = grid[k]
        centres.add((cx, cy, cz))
        planes.append(z0 - small // 2 + cz)
        idx.append((y0 - small // 2 + cy) * W + (x0 - small // 2 + cx))
    vox = (np.concatenate(planes).astype(np.int32), np.concatenate(idx).astype(np.int32),
           np.full(300 * len(z0), 100.0, np.float32))
    print(f"planted spheres {Nz}x{H}x{W}, {len(vox[0])} voxels", flush=True)
    with pkg.SparseSegmentation(W, H, Nz, voxels=vox) as seg:
        seg.binarize(1.0)
        seg.close(cs)
        seg.labelConnectedComponents()
        objs = seg.extractObjects().objects()
        t = seg.timing()
    print("stage ms", t, flush=True)
    assert len(objs) == 300
    assert all(o["num_voxels"] == n_closed for o in objs)
    assert {tuple(int(c) for c in o["centroid"]) for o in objs} == centres
    assert all(float(c) == int(c) for o in objs for c in o["centroid"])
    labels = [o["label"] for o in objs]
    assert labels == sorted(labels)   # first appearance in the raster
