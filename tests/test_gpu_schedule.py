"""Schedules, the autotuner and the batched multi-size call on the GPU.

  - every launch choice a schedule holds (the fixed-length or run-time-plan FFTs, the run-time
    plans' largest radix 16/8/4, the z chunks of A and A^T, the prox's rows per wave, graphs on
    or off) solves to the FP32 oracle's result (1e-4 rel-L2, identical backtracks and restarts):
    a schedule changes speed, not the mathematics;
  - rihvr_get_schedule reports what a context runs with;
  - rihvr_autotune: the best schedule is the fastest of those it measured, and solves to the
    same result as the default schedule (1e-4);
  - rihvr_reconstruct_multi over holograms of three sizes, two parameter sets sharing one size,
    an objective target and host and device arrays, with the default, autotuned and given
    schedules: each item equals the single-hologram batch call with the schedule the item
    reports, bit for bit.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def rel(a, b):
    return float(np.linalg.norm(np.asarray(a) - np.asarray(b)) / max(np.linalg.norm(np.asarray(b)), 1e-30))


def cfg(W, H, Nz, its=10, reg=0.01, tv=0.01, ntv=5, mode=7):
    return dict(width=W, height=H, num_planes=Nz, start_plane=1000.0, plane_stepsize=50.0, wavelength=0.633,
                resolution=2.0, num_inverse_iterations=its, regularization_param=reg,
                regularization_TV_param=tv, num_tv_iterations=ntv, mode=mode)


SCHEDULES = [
    dict(fft_path=0),
    dict(fft_path=1, max_radix=16),
    dict(fft_path=1, max_radix=8),
    dict(fft_path=1, max_radix=4),
    dict(zchunk_fwd=3, zchunk_adj=5),
    dict(zchunk_fwd=16, zchunk_adj=1),
    dict(prox_rows=24),
    dict(graphs=0, prox_rows=7),
    dict(fft_path=1, max_radix=8, graphs=0, zchunk_fwd=2),
]


@pytest.mark.parametrize("k", range(len(SCHEDULES)))
@pytest.mark.parametrize("shape", [(128, 96, 12), (272, 208, 5)])
def test_every_schedule_solves_to_the_oracle(pkg, orc, scenes, gpu_available, k, shape):
    W, H, Nz = shape
    sd = SCHEDULES[k]
    print(f"schedule {sd} at {W}x{H}x{Nz}", flush=True)
    c = cfg(W, H, Nz, its=10)
    holo, _ = scenes.disks_scene(scenes.optics_of(c), 6, 2.0, 4.0, 70 + W, 0.01)
    with pkg.SparseCompressiveHolo(pkg.make_params(**c), pkg.Schedule.make(**sd)) as rec:
        got = rec.schedule()
        st = rec.inverseReconstruct(holo)
        vol = rec.getVolume()
    fixed = W == 128
    assert got["fft_path"] == (sd.get("fft_path", 0) if fixed else 1)
    assert got["graphs"] == sd.get("graphs", 1)
    for key, default in (("zchunk_fwd", 32), ("zchunk_adj", 8)):
        assert got[key] == min(sd.get(key, default), Nz), key
    if "prox_rows" in sd:
        assert got["prox_rows"] == sd["prox_rows"]
    ref, _, _, rst = orc.solve(orc.make_params(**c), holo, None, "f32")
    assert rel(vol, ref) < 1e-4, rel(vol, ref)
    assert list(st["backtracks"]) == list(rst["backtracks"])
    assert list(st["restarts"]) == list(rst["restarts"])
    np.testing.assert_allclose(st["f"], rst["f"], rtol=1e-4)


def test_autotune_picks_the_fastest_and_solves_the_same(pkg, scenes, gpu_available):
    c = cfg(256, 256, 32, its=15)
    holo, _ = scenes.disks_scene(scenes.optics_of(c), 10, 2.0, 4.0, 91, 0.01)
    p = pkg.make_params(**c)
    best, tried = pkg.autotune(p, holo, iterations=5)
    for t in tried:
        print(t, flush=True)
    assert len(tried) >= 8
    assert all(t["ms_per_iteration"] > 0 for t in tried)
    assert best.ms_per_iteration == min(t["ms_per_iteration"] for t in tried)
    with pkg.SparseCompressiveHolo(p) as rec:
        rec.inverseReconstruct(holo)
        v0 = rec.getVolume()
    with pkg.SparseCompressiveHolo(p, best) as rec:
        rec.inverseReconstruct(holo)
        v1 = rec.getVolume()
    assert rel(v1, v0) < 1e-4
    # the synthetic problem when no hologram is given
    best2, tried2 = pkg.autotune(p, None, iterations=3)
    assert best2.ms_per_iteration > 0 and len(tried2) >= 8


def test_autotune_rejects_bad_arguments(pkg, gpu_available):
    p = pkg.make_params(**cfg(64, 64, 4))
    with pytest.raises(pkg.RihvrError):
        pkg.autotune(p, None, iterations=0)


@pytest.mark.parametrize("autotune", [0, 1, 2])
@pytest.mark.parametrize("device", [False, True])
def test_multi_equals_single_solves(pkg, scenes, gpu_available, autotune, device):
    import torch
    specs = [cfg(128, 96, 12, reg=0.01, its=8), cfg(272, 208, 5, reg=0.02, its=6),
             cfg(128, 96, 12, reg=0.004, tv=0.02, its=9), cfg(64, 64, 8, reg=0.01, tv=0.0, mode=7, its=7),
             cfg(272, 208, 5, reg=0.01, its=12)]
    targets = [0.0, 0.0, 0.0, 0.0, None]
    holos = [scenes.disks_scene(scenes.optics_of(c), 6, 2.0, 4.0, 300 + k, 0.01)[0] for k, c in enumerate(specs)]
    # the last item stops at an objective target: the plain solve's f after 6 iterations
    with pkg.SparseCompressiveHolo(pkg.make_params(**dict(specs[4], num_inverse_iterations=6))) as rec:
        targets[4] = rec.inverseReconstruct(holos[4])["f_final"] * (1 + 1e-6)
    print(f"multi: autotune={autotune} device={device}", flush=True)
    items = []
    for k, (c, h, tg) in enumerate(zip(specs, holos, targets)):
        vol_shape = (c["num_planes"], c["height"], c["width"])
        if device:
            hh = torch.from_numpy(h).cuda()
            vv = torch.empty(vol_shape, dtype=torch.complex64, device="cuda")
        else:
            hh, vv = h, np.empty(vol_shape, np.complex64)
        given = pkg.Schedule.make(fft_path=1, max_radix=8, zchunk_fwd=4, zchunk_adj=3, prox_rows=40, graphs=k % 2)
        items.append(dict(params=pkg.make_params(**c), holo=hh, volume=vv, f_target=tg, schedule=given))
    if device:
        torch.cuda.synchronize()
    res = pkg.reconstruct_multi(items, autotune=autotune)
    if device:
        torch.cuda.synchronize()
    for k, (c, h, tg, r) in enumerate(zip(specs, holos, targets, res)):
        sch = r["schedule"]
        if autotune == 2:  # the first item of the group gives the group's schedule
            first = [j for j in range(len(specs)) if specs[j]["width"] == c["width"] and
                     specs[j]["num_planes"] == c["num_planes"]][0]
            assert sch["fft_path"] == 1 and sch["max_radix"] == 8 and sch["graphs"] == first % 2
            assert sch["zchunk_fwd"] == min(4, c["num_planes"]) and sch["prox_rows"] == 40
        s = pkg.Schedule.make(sch["fft_path"], sch["max_radix"], sch["zchunk_fwd"], sch["zchunk_adj"],
                              sch["prox_rows"], sch["graphs"])
        with pkg.SparseCompressiveHolo(pkg.make_params(**c), s) as rec:
            fs = rec.reconstructBatch(h[None], f_target=None if not tg else [tg],
                                      max_iterations=c["num_inverse_iterations"])[0]
            vol = rec.getVolume()
        got = items[k]["volume"]
        got = got.cpu().numpy() if device else got
        assert np.array_equal(got.view(np.uint32), vol.view(np.uint32)), k
        assert r["iterations"] == fs["iterations"] and r["f_final"] == fs["f_final"] and r["nnz"] == fs["nnz"], k
        assert r["reached"]
    assert res[4]["iterations"] <= 6


def test_multi_items_match_the_oracle_per_size_class(pkg, orc, scenes, gpu_available):
    """One item per size class of a mixed batch (two fixed-plan sizes, a run-time-plan size, three
    depths) against the FP32 oracle: 1e-4 rel-L2 with the oracle's decisions."""
    specs = [cfg(128, 96, 12, reg=0.01, its=8), cfg(272, 208, 5, reg=0.02, its=6), cfg(64, 64, 8, reg=0.004, its=7)]
    holos = [scenes.disks_scene(scenes.optics_of(c), 6, 2.0, 4.0, 400 + k, 0.01)[0] for k, c in enumerate(specs)]
    items = [dict(params=pkg.make_params(**c), holo=h, volume=np.empty((c["num_planes"], c["height"], c["width"]),
                                                                         np.complex64))
             for c, h in zip(specs, holos)]
    res = pkg.reconstruct_multi(items, autotune=1)
    for c, h, it, r in zip(specs, holos, items, res):
        ref, _, _, rst = orc.solve(orc.make_params(**c), h, None, "f32")
        print(f"multi {c['width']}x{c['height']}x{c['num_planes']}: rel {rel(it['volume'], ref):.2e}, "
              f"schedule {r['schedule']}", flush=True)
        assert rel(it["volume"], ref) < 1e-4
        assert r["iterations"] == c["num_inverse_iterations"]
        assert sum(rst["backtracks"]) == r["total_backtracks"]
