"""GPU tests of the result writers and of the sparse-inverse-recon executable.

Writers: SparseVolume::saveData (sparse_volume.cu:518-560), saveProjections
(:684-813) and SparseCompressiveHolo::savePlanes (sparse_compressive_holo.cu:
1039-1194), checked against numpy computations on the volume read back through
rihvr_get_volume (8-bit outputs to one grey level: the reference's float
expression order is not fixed by OpenCV's API).  The executable
(main_sparse_inverse.cu:43-198) is run end to end on a synthetic 8-bit hologram
and its voxel list is checked against the FP32 oracle solving the same
preprocessed hologram.
"""
from __future__ import annotations

import pathlib
import subprocess

import numpy as np
import pytest

from imgfiles import read_tiff, tiff_bytes

pytestmark = pytest.mark.gpu


def cfg(W, H, Nz, its=8, reg=0.01, tv=0.01, ntv=5, z0=1000.0, dz=50.0, lam=0.633, dx=2.0):
    return dict(width=W, height=H, num_planes=Nz, start_plane=z0, plane_stepsize=dz, wavelength=lam,
                resolution=dx, num_inverse_iterations=its, regularization_param=reg,
                regularization_TV_param=tv, num_tv_iterations=ntv, mode=7)


@pytest.fixture(scope="module")
def solved(pkg, scenes, gpu_available):
    c = cfg(64, 64, 8)
    holo, _ = scenes.disks_scene(scenes.optics_of(c), 6, 2.0, 4.0, 21, 0.01)
    h = pkg.SparseCompressiveHolo(pkg.make_params(**c))
    h.inverseReconstruct(holo.astype(np.complex64))
    vol = h.getVolume()
    assert np.count_nonzero(vol) > 0
    yield h, vol
    h.destroy()


def _read_csv(path):
    lines = pathlib.Path(path).read_text().splitlines()
    assert lines[0] == "x, y, z, value"
    rows = [ln.split(", ") for ln in lines[1:]]
    return np.array([[int(r[0]), int(r[1]), int(r[2])] for r in rows], np.int64).reshape(-1, 3), \
        np.array([float(r[3]) for r in rows], np.float64)


def test_save_sparse(solved, tmp_path):
    h, vol = solved
    p = tmp_path / "voxels.csv"
    h.saveSparse(p)
    idx, val = _read_csv(p)
    z, y, x = np.nonzero(vol)  # C order: z, then y, then x
    np.testing.assert_array_equal(idx, np.stack([x, y, z], 1))
    inten = (vol.real[z, y, x] ** 2 + vol.imag[z, y, x] ** 2).astype(np.float32)
    np.testing.assert_allclose(val, inten, rtol=1e-6)
    assert "e" in pathlib.Path(p).read_text().splitlines()[1].split(", ")[3]  # %e format


def _mag_range(vol):
    m = np.sqrt(vol.real.astype(np.float32) ** 2 + vol.imag.astype(np.float32) ** 2)
    return m, float(m.min()), float(m.max())


def _close_u8(got, exp):
    d = np.abs(got.astype(int) - exp.astype(int))
    assert d.max() <= 1, d.max()
    assert (d > 0).mean() < 0.01


def test_save_projections(solved, tmp_path):
    h, vol = solved
    out = str(tmp_path) + "/"
    h.saveProjections(out, "cmb")
    m, lo, hi = _mag_range(vol)
    n = ((m - np.float32(lo)) / np.float32(hi - lo)).astype(np.float32)
    xy = np.maximum(np.zeros_like(n[0]), n[1:].max(0))  # plane 0 is left out of xy
    exp = {"xy": xy, "xz": n.max(1), "yz": n.max(2).T}
    for k, e in exp.items():
        got = read_tiff(tmp_path / f"cmb_{k}.tif")
        assert got.shape == e.shape
        _close_u8(got, np.clip(np.rint(e * np.float32(255)), 0, 255))
    h.saveProjections(out, "step7", True)
    for k in exp:
        assert (tmp_path / f"{k}_step7.tif").exists()


def test_save_planes(solved, tmp_path):
    h, vol = solved
    prefix = str(tmp_path / "inverse")
    h.savePlanes(prefix)
    Nz = vol.shape[0]
    m, lo, hi = _mag_range(vol)
    ri = np.sqrt(hi) / 10
    ph = np.arctan2(vol.imag.astype(np.float64), vol.real.astype(np.float64))
    ph = np.where(ph < 0, ph + 2 * np.pi, ph) / (2 * np.pi)
    for z in (0, Nz // 2, Nz - 1):
        u = lambda a: np.clip(np.rint(np.asarray(a, np.float32) * np.float32(255)), 0, 255)  # noqa: E731
        _close_u8(read_tiff(f"{prefix}_{z:04d}.tif"), u((m[z] - lo) / (hi - lo)))
        _close_u8(read_tiff(f"{prefix}_phase_{z:04d}.tif"), u(ph[z]))
        _close_u8(read_tiff(f"{prefix}_real_{z:04d}.tif"), u((vol[z].real + ri) / (2 * ri)))
        _close_u8(read_tiff(f"{prefix}_imag_{z:04d}.tif"), u((vol[z].imag + ri) / (2 * ri)))
    assert len(list(tmp_path.glob("inverse_*.tif"))) == 4 * Nz


def test_snapshot_leaves_solve_unchanged(pkg, scenes, gpu_available):
    c = cfg(64, 48, 5, its=6)
    holo, _ = scenes.disks_scene(scenes.optics_of(c), 5, 2.0, 4.0, 22, 0.01)
    holo = holo.astype(np.complex64)
    with pkg.SparseCompressiveHolo(pkg.make_params(**c)) as a:
        a.inverseReconstruct(holo)
        ref = a.getVolume()
        ref_est, ref_res = a.getEstimate()
    with pkg.SparseCompressiveHolo(pkg.make_params(**c)) as b:
        with pytest.raises(pkg.RihvrError):
            b.snapshot()  # no active solve
        b.begin(holo)
        b.iterate(3)
        b.snapshot()
        mid = b.getVolume()
        b.iterate(3)
        with pytest.raises(pkg.RihvrError) as e:
            b.getVolume()  # the snapshot is superseded by the next iteration
        assert e.value.status == 4
        b.snapshot()
        snap6 = b.getVolume()
        b.finish()
        vol = b.getVolume()
        est, res = b.getEstimate()
    np.testing.assert_array_equal(vol, ref)
    np.testing.assert_array_equal(snap6, ref)
    np.testing.assert_array_equal(est, ref_est)
    np.testing.assert_array_equal(res, ref_res)
    assert not np.array_equal(mid, ref)


# --------------------------------------------------------------------------
# the executable
# --------------------------------------------------------------------------
def _exe(pkg):
    exe = pkg.PKG_DIR / "_build" / "sparse-inverse-recon"
    assert exe.exists(), "sparse-inverse-recon is not built"
    return exe


def _param_file(path, img_dir, out_dir, c, pad, output_planes=True, output_steps=False):
    # the layout the reference GUI writes (src/QTGUI/reconstructioncompute.cpp:46-111)
    text = (f"num_images            1001 1\nstart_image           1002 3\n"
            f"input_pathn           2002 {img_dir}/\ninput_filename_format 2003 holo_%04d.tif\n"
            f"output_pathn          2004 {out_dir}/\n"
            f"start_plane_z         3001 {c['start_plane']}\nplane_step_dz         3002 {c['plane_stepsize']}\n"
            f"num_planes            3003 {c['num_planes']}\nwavelength            3004 {c['wavelength']}\n"
            f"resolution            3005 {c['resolution']}\nzero_padding          3010 {pad}\n"
            f"inverse_iterations    6001 {c['num_inverse_iterations']}\n"
            f"regularization_tau    6002 {c['regularization_param']}\n"
            f"regularization_TV     6004 {c['regularization_TV_param']}\n"
            f"num_TV_iterations     6005 {c['num_tv_iterations']}\n"
            f"output_planes         7001 {'true' if output_planes else 'false'}\n"
            f"output_each_step      7005 {'true' if output_steps else 'false'}\n")
    path.write_text(text)


@pytest.fixture(scope="module")
def holo_image(scenes, tmp_path_factory):
    c = cfg(48, 48, 6)
    b, _ = scenes.disks_scene(scenes.optics_of(c), 5, 2.0, 3.5, 23, 0.01)
    img = np.clip(np.rint(128.0 * (1.0 + b.real)), 0, 255).astype(np.uint8)
    d = tmp_path_factory.mktemp("img")
    (d / "holo_0003.tif").write_bytes(tiff_bytes(img))
    return d, img


def test_cli_end_to_end_matches_oracle(pkg, orc, holo_image, tmp_path, gpu_available):
    img_dir, img = holo_image
    c = cfg(64, 64, 6, its=8)  # the 48^2 image zero-padded by 8
    out = tmp_path / "out"
    out.mkdir()
    pf = tmp_path / "params.txt"
    _param_file(pf, img_dir, out, c, pad=8)
    print(f"running {_exe(pkg)} -F {pf}", flush=True)
    r = subprocess.run([str(_exe(pkg)), "-F", str(pf)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
    assert "Done. Total time:" in r.stdout
    # the same preprocessing and solve through the Python binding and the oracle
    fp, rc = pkg.read_param_file(pf)
    assert rc == 0
    holo = pkg.prepare_hologram(fp, pkg.read_image(img_dir / "holo_0003.tif"))
    assert holo.shape == (64, 64)
    ref, _, _, rst = orc.solve(orc.make_params(**c), holo)
    idx, val = _read_csv(out / "inverse_voxels.csv")
    z, y, x = np.nonzero(ref)
    got = np.zeros(ref.shape, np.float64)
    got[idx[:, 2], idx[:, 1], idx[:, 0]] = val
    inten = np.abs(ref.astype(np.complex128)) ** 2
    mismatch = int(((got != 0) != (inten != 0)).sum())
    assert mismatch <= 2, mismatch
    assert np.linalg.norm(got - inten) <= 2e-4 * np.linalg.norm(inten)
    # files of the run
    for f in ("enh.tif", "estimate_final.tif", "residual_final.tif", "all_metrics.csv"):
        assert (out / f).exists(), f
    assert (out / "all_metrics.csv").read_text() == "it, f, sparsity, stepsize, total_backtracks\n"
    for k in ("xy", "xz", "yz"):
        assert (out / f"cmb_{k}.tif").exists()
    assert len(list(out.glob("inverse_0*.tif"))) == c["num_planes"]
    enh = read_tiff(out / "enh.tif")
    assert enh.shape == (64, 64) and enh.min() == 0 and enh.max() == 255


def test_cli_output_steps_and_subfolder(pkg, holo_image, tmp_path, gpu_available):
    img_dir, _ = holo_image
    c = cfg(48, 48, 4, its=3)
    out = tmp_path / "o"
    out.mkdir()
    pf = tmp_path / "params.txt"
    _param_file(pf, img_dir, out, c, pad=0, output_planes=True, output_steps=True)
    print(f"running {_exe(pkg)} -F {pf} -S 1", flush=True)
    r = subprocess.run([str(_exe(pkg)), "-F", str(pf), "-S", "1"], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
    lines = (out / "all_metrics.csv").read_text().splitlines()
    assert lines[0] == "it, f, sparsity, stepsize, total_backtracks" and len(lines) == 4
    # per-iteration objective values equal the stepwise solve's
    fp, _ = pkg.read_param_file(pf)
    holo = pkg.prepare_hologram(fp, pkg.read_image(img_dir / "holo_0003.tif"))
    with pkg.SparseCompressiveHolo(pkg.make_params(**c)) as h:
        h.begin(holo)
        for it in range(3):
            st = h.iterate(1, with_objective=True)
            fields = lines[1 + it].split(", ")
            assert int(fields[0]) == it
            assert float(fields[1]) == pytest.approx(st["f"][0], abs=1e-6, rel=1e-5)
            assert int(fields[4]) == st["total_backtracks"]
    for it in range(3):
        for k in ("xy", "xz", "yz"):
            assert (out / f"{k}_cmb_iter{it:03d}_end.tif").exists()
        assert (out / f"estimate_iter{it:03d}.tif").exists() and (out / f"residual_iter{it:03d}.tif").exists()
    sub = out / "0003"
    assert len(list(sub.glob("inverse_*.tif"))) == 4 * c["num_planes"]


def test_cli_background_removal_matches_pipeline(pkg, orc, holo_image, tmp_path, gpu_available):
    """The remove_background branch (2005 line): (I - B) / max(sqrt(B), 0.1)."""
    img_dir, img = holo_image
    rng = np.random.default_rng(5)
    bg = np.clip(img.astype(np.float64) * 0.0 + 120 + rng.normal(0, 3, img.shape), 0, 255).astype(np.uint8)
    (img_dir / "bg.tif").write_bytes(tiff_bytes(bg))
    c = cfg(48, 48, 4, its=5)
    out = tmp_path / "o"
    out.mkdir()
    pf = tmp_path / "params.txt"
    _param_file(pf, img_dir, out, c, pad=0, output_planes=False)
    with open(pf, "a") as f:
        f.write(f"background_image      2005 {img_dir}/bg.tif\n")
    r = subprocess.run([str(_exe(pkg)), "-F", str(pf)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
    assert "Removing background" in r.stdout
    fp, rc = pkg.read_param_file(pf)
    assert rc == 0 and fp.remove_background == 1
    holo = pkg.prepare_hologram(fp, pkg.read_image(img_dir / "holo_0003.tif"), pkg.read_image(img_dir / "bg.tif"))
    ref, _, _, _ = orc.solve(orc.make_params(**c), holo)
    idx, val = _read_csv(out / "inverse_voxels.csv")
    got = np.zeros(ref.shape, np.float64)
    got[idx[:, 2], idx[:, 1], idx[:, 0]] = val
    inten = np.abs(ref.astype(np.complex128)) ** 2
    assert int(((got != 0) != (inten != 0)).sum()) <= 2
    assert np.linalg.norm(got - inten) <= 2e-4 * max(np.linalg.norm(inten), 1e-30)
