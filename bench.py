"""Benchmark: FISTA iterations per second of the plain FASTA/FISTA + 2D-TV solve
on the 1024x1024x128 FP32 volume (the reference's default mode 7), on one MI355X.

The unit of work is one FISTA iteration of SparseCompressiveHolo::inverseReconstruct
(prox with A^T inside, A, residual, FFT(r), reductions, momentum; backtracks that
occur are inside it).  A bench step is one whole solve of the workload's 100
iterations, so the rate covers the dense early iterations and the sparse late ones
alike: the hologram is generated synthetically (500 tracers + 20 rods, seeded) and
copied to HBM; per step the solve is set up (rihvr_begin), its 100 iterations are
timed between device synchronisations (rihvr_iterate), and it is finished.  W untimed
warm-up solves come first.  value = 100 K / (summed time of the K iterate calls).

A second leg ("frames") measures s/frame to an objective tolerance on a short
time sequence at the same optics, cold start against a warm start from the
previous frame (rihvr_reconstruct_batch).

Prints one JSON line.  --gpus N > 1 exits: this build runs on one GPU.
"""
from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

import __graft_entry__ as entry  # noqa: E402

METRIC = "FISTA iter/s @1024^2x128 FP32 (%HBM roof); s/frame to tol, warm-start vs plain"
HBM_PEAK_GBPS = 8000.0          # MI355X HBM3E
FP32_PEAK_TFLOPS = 157.3        # MI355X FP32 vector (vendor figure)


def algorithmic_bytes(name: str, W: int, H: int, Nz: int, nparts: int, f_line: float = 1.0) -> float:
    """Compulsory HBM bytes of one launch of each solver kernel (complex64 = 8 B).  f_line:
    the fraction of 16-voxel lines of the volume that hold a non-zero (the solver reads and
    writes only those lines of y; see the line occupancy maps)."""
    N, V = W * H, W * H * Nz
    return {
        # reads g (dense) and the occupied lines of y_{k-1}, y_{k-2}; writes those of y_k
        "prox": 8.0 * V + 3 * 8.0 * V * f_line,
        # reads the occupied lines of y_k, writes the column-transformed planes (upper bound)
        "fft_cols_fwd": 8.0 * V * f_line + 8.0 * V,
        # reads the column-transformed planes and kappa rows, writes nparts partial spectra
        "fft_rows_fwdacc": 8.0 * V + 8.0 * N * nparts + 8.0 * N * nparts,
        # reads R-hat and kappa (16 B / pixel), writes the row-transformed planes
        "fft_rows_adj": 8.0 * V + 16.0 * N,
        # reads the row-transformed planes, writes g
        "fft_cols_adj": 16.0 * V,
    }.get(name, float("nan"))


def prox_flops_per_voxel(n_tv: int) -> int:
    """FP32 operations per voxel of the mode-7 prox (fma = 2): momentum point and x-hat 10,
    first TV level 25 (forward differences 6, the y-gradient 6, the 4-component norm 7,
    max and rsqrt 2, the projection 4), each further level 35 (+ the divergence 6 and the
    estimate 4), final estimate 10, soft threshold 12, zero canonicalisation 2.  The line
    search, restart and objective sums (about 36 more) are counted out: they run only on
    rows holding a non-zero."""
    return 10 + 25 + 35 * (n_tv - 1) + 10 + 12 + 2 if n_tv > 0 else 10 + 12 + 2


# the kernel's limiting unit: the prox's TV pipeline is issue-bound on the vector ALUs
# (measured VALU issue in profiles/c2_hbm_counters.json), the FFT passes on HBM
KERNEL_BOUND = {"prox": "valu"}


def iteration_bytes(W, H, Nz):
    """B = 3V + 8H per FISTA iteration (read y_{k-1}, y_{k-2}, write y_k; 8 holograms)."""
    N = W * H
    return 3 * 8.0 * N * Nz + 8 * 8.0 * N


def iteration_flops(W, H, Nz):
    """F = 10 Nz N log2 N + 10 N log2 N + 200 N Nz."""
    N = W * H
    return 10 * Nz * N * math.log2(N) + 10 * N * math.log2(N) + 200 * N * Nz


def cpu_baseline(cfg: dict, holo, iters_hi=2, iters_lo=1):
    """FP32 oracle (CPU restatement of the reference) on the host cores.  Per-iteration
    time = (t(iters_hi) - t(iters_lo)) / (iters_hi - iters_lo), which cancels the
    solve set-up."""
    import oracle
    times = {}
    for n in (iters_lo, iters_hi):
        p = oracle.make_params(**dict(cfg, num_inverse_iterations=n))
        t0 = time.perf_counter()
        oracle.solve(p, holo, None, "f32")
        times[n] = time.perf_counter() - t0
    per_it = (times[iters_hi] - times[iters_lo]) / (iters_hi - iters_lo)
    return dict(value=1.0 / per_it, unit="iter/s", cores=oracle.get_threads(), kind="port",
                sample=(f"{cfg['width']}x{cfg['height']}x{cfg['num_planes']} FP32 oracle, "
                        f"{iters_hi - iters_lo} FISTA iterations (difference of a {iters_hi}- and a "
                        f"{iters_lo}-iteration solve, {times[iters_hi] + times[iters_lo]:.1f} s total)"))


WEIGHTS = os.path.join(ROOT, "rihvr---regularized-inverse-holographic-volume-reconstruction_amd", "weights",
                       "warmstart_c2.npz")


def load_net(pkg, shape):
    """The committed warm-start network (trained at the 1024^2 x 128 workload's optics)."""
    with np.load(WEIGHTS, allow_pickle=False) as z:
        params, i_s, o_s = z["params"], float(z["in_scale"]), float(z["out_scale"])
    net = pkg.WarmStartNet(*shape)
    net.set_params(params, i_s, o_s)
    return net


def frames_leg(pkg, scenes, cfg, n_frames: int, tol: float = 1e-3, seed: int = 20261203):
    """s/frame to an objective tolerance on a time sequence at the workload's optics:
    plain cold-start FISTA for num_inverse_iterations gives f_plain per frame; then the cold
    start, the learned warm start (rihvr_reconstruct_learned with the committed network: A^T b,
    network, the start test, FISTA) and the previous frame's volume as start all run until
    f <= (1 + tol) f_plain or the same cap.  Frames are device-resident."""
    import torch
    optics = scenes.optics_of(cfg)
    holos = scenes.sequence(optics, n_frames, n_tracers=500, n_rods=20, seed=seed, sigma=0.02)
    hd = torch.from_numpy(holos).cuda()
    cap = cfg["num_inverse_iterations"]
    rec = pkg.SparseCompressiveHolo(pkg.make_params(**cfg))
    net = load_net(pkg, (cfg["num_planes"], cfg["height"], cfg["width"]))
    try:
        def run(**kw):
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            fs = rec.reconstructBatch(hd, **kw)
            torch.cuda.synchronize()
            return fs, time.perf_counter() - t0
        plain, t_plain = run()
        target = [(1.0 + tol) * f["f_final"] for f in plain]
        cold, t_cold = run(f_target=target, max_iterations=cap)
        warm, t_warm = run(f_target=target, max_iterations=cap, warm_from_previous=True)
        learned = []
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for k in range(n_frames):
            learned.append(rec.reconstructLearned(net, hd[k], f_target=target[k], max_iterations=cap))
        torch.cuda.synchronize()
        t_learned = time.perf_counter() - t0
    finally:
        net.destroy()
        rec.destroy()

    def leg(fs, t):
        return {"s_per_frame": t / n_frames, "iterations": [f["iterations"] for f in fs],
                "reached": [f["reached"] for f in fs], "f_final": [f["f_final"] for f in fs]}
    # frame 0 has no predecessor: the temporal comparison is over frames 1..n-1
    it_cold = sum(f["iterations"] for f in cold[1:])
    it_warm = sum(f["iterations"] for f in warm[1:])
    lrn = leg(learned, t_learned)
    lrn.update(fallback=[f["fallback"] for f in learned], tau=[f["tau"] for f in learned],
               f_thresholded=[f["f_thresholded"] for f in learned],
               ms_network=[f["ms_network"] for f in learned], ms_adjoint=[f["ms_adjoint"] for f in learned],
               ms_solve=[f["ms_solve"] for f in learned])
    return {
        "workload": (f"{n_frames} frames, {cfg['width']}x{cfg['height']}x{cfg['num_planes']} at the same optics, "
                     "500 tracers + 20 rods advected by a seeded divergence-free 16-mode flow (rms 3 px/frame)"),
        "tolerance": f"f <= (1 + {tol}) * f_plain({cap})",
        "plain_fixed": dict(leg(plain, t_plain), iterations_each=cap),
        "plain_to_tol": leg(cold, t_cold),
        "learned_to_tol": lrn,
        "warm_previous_to_tol": leg(warm, t_warm),
        "learned_speedup_vs_plain_to_tol": (t_cold / t_learned) if t_learned else None,
        "learned_speedup_vs_plain_fixed": (t_plain / t_learned) if t_learned else None,
        "iterations_ratio_cold_over_warm_frames_1_on": (it_cold / it_warm) if it_warm else None,
        "network": "weights/warmstart_c2.npz (BF16 MFMA encoder-decoder, 16-128 channels, 4 levels)",
    }


C2_F64 = os.path.join(ROOT, "tests", "golden", "c2_oracle_f64.npz")


def fp64_error(vol, holo):
    """Relative L2 error of the returned volume (device, complex64) against the FP64 oracle's
    solution of the same hologram: the committed fixture written by tests/golden/make_c2_golden.py
    (its non-zero voxels).  None when the fixture is absent or belongs to another input."""
    import torch
    if not os.path.exists(C2_F64):
        return None
    with np.load(C2_F64, allow_pickle=False) as z:
        if complex(z["holo_sum"]) != complex(np.sum(holo.astype(np.complex128))):
            return None
        idx, val = z["idx"].astype(np.int64), z["val"]
    d = vol.reshape(-1).to(torch.complex128)
    ref = torch.from_numpy(val).to(d.device)
    d[torch.from_numpy(idx).to(d.device)] -= ref
    return float(torch.linalg.vector_norm(d) / torch.linalg.vector_norm(ref))


def c5_bench(pkg, scenes, n_frames: int, tol: float = 1e-3, seed: int = 910001):
    """A stream of holograms at the 1024^2 x 128 workload's optics with 300-800 tracers and
    rods each (independent scenes, not a sequence): plain cold start against the learned
    start, both to (1 + tol) f_plain(100)."""
    import torch
    ws = entry._load_submodule("warmstart")
    with np.load(WEIGHTS, allow_pickle=False) as z:
        params, i_s, o_s = z["params"], float(z["in_scale"]), float(z["out_scale"])
    with pkg.WarmStartNet(128, 1024, 1024) as net:
        net.set_params(params, i_s, o_s)
        def progress(k):
            if (k + 1) % 16 == 0 or k + 1 == n_frames:
                print(f"c5: {k + 1} of {n_frames} holograms done", file=sys.stderr, flush=True)

        s = ws.evaluate(net, "c2", 1024, n_frames, seed, log=None, propagate=scenes.propagate_sum_torch,
                        progress=progress)
    rows = s.pop("rows")
    s["iterations_cold_mean"] = float(np.mean(s["cold_iterations"]))
    s["iterations_learned_mean"] = float(np.mean(s["learned_iterations"]))
    s["ms_network_mean"] = float(np.mean([r["ms_network"] for r in rows]))
    # every learned result is measured by its objective against the plain solver's f(100)
    s["f_learned_over_f_plain"] = [r["f_learned"] / r["f_plain"] for r in rows]
    return {"metric": "s/frame to (1 + 1e-3) f_plain(100), learned warm start vs plain, 1024^2x128 FP32",
            "value": s["learned_s_per_frame"], "unit": "s/frame", "n_gpus": 1, "steps": n_frames, "warmup": 0,
            "ms_per_step": 1e3 * s["learned_s_per_frame"], "higher_is_better": False, "scaling": "weak",
            "vs_baseline": None, "dtype": "bf16 network / f32 FISTA",
            "data": "synthetic (seeded 300-800 tracers and rods per hologram)",
            "config": {"workload": f"stream of {n_frames} holograms 1024x1024x128 (C2 optics), learned start vs plain"},
            "summary": s}


C3_SEQ = dict(width=512, height=512, num_planes=100, start_plane=1000.0, plane_stepsize=25.0, wavelength=0.633,
              resolution=2.0, num_inverse_iterations=50, regularization_param=0.01, regularization_TV_param=0.01,
              num_tv_iterations=5, mode=7)


def c3_mixed_specs(n: int, seed: int = 930001):
    """The mixed batch: n holograms, each of side 384, 512 or 768 padded by 32 or 64 zeros on each
    side, Nz 40, 64 or 100, L1 weight log-uniform in [0.001, 0.01], TV 0.01, at the sequence's
    optics (dz 25 um, dx 2 um, lambda 0.633 um), with 1000 tracers per 512^2 of field and 10 rods."""
    rng = np.random.default_rng(seed)
    out = []
    for k in range(n):
        side = int(rng.choice([384, 512, 768]))
        pad = int(rng.choice([32, 64]))
        nz = int(rng.choice([40, 64, 100]))
        reg = float(np.exp(rng.uniform(np.log(0.001), np.log(0.01))))
        c = dict(C3_SEQ, width=side + 2 * pad, height=side + 2 * pad, num_planes=nz, regularization_param=reg,
                 zero_padding=pad)
        out.append((c, side, pad, int(round(1000 * side * side / 512 ** 2)), 10, seed + 1 + k))
    return out


def c3_bench(pkg, scenes, n_seq: int = 64, n_mixed: int = 48):
    """The batched workload: (1) a 64-frame 512^2 x 100 sequence (1000 tracers advected by the
    seeded 16-mode divergence-free flow, 3 px/frame rms) solved frame after frame on one context
    (rihvr_reconstruct_batch, 50 FISTA x 5 TV each, cold starts); (2) a mixed batch of 48 holograms
    of uneven padded sizes, Nz and weights in one rihvr_reconstruct_multi call.  Each part runs with
    the default schedule and with the schedule rihvr_autotune picks per geometry (tuned once on the
    geometry's first hologram, tuning time reported apart); the queued iterations are
    graph-captured in both.  Holograms are generated on the GPU (torch, complex128) and are
    device-resident before the timed calls."""
    import torch
    prop = scenes.propagate_sum_torch
    out = {}
    # (1) the sequence
    seq = scenes.sequence(scenes.optics_of(C3_SEQ), n_seq, n_tracers=1000, n_rods=0, seed=920001, sigma=0.02,
                          propagate=prop)
    hs = torch.from_numpy(seq).cuda()
    p = pkg.make_params(**{k: v for k, v in C3_SEQ.items()})
    t0 = time.perf_counter()
    best, tried = pkg.autotune(p, hs[0].contiguous(), iterations=10)
    tune_s = time.perf_counter() - t0
    legs = {}
    for name, sch in (("default", None), ("autotuned", best)):
        with pkg.SparseCompressiveHolo(p, sch) as rec:
            rec.reconstructBatch(hs[:1])  # warm-up: first launches, graph captures
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            fs = rec.reconstructBatch(hs)
            torch.cuda.synchronize()
            dt = time.perf_counter() - t0
            legs[name] = {"s_total": dt, "frames_per_s": n_seq / dt,
                          "iter_per_s": sum(f["iterations"] for f in fs) / dt, "schedule": rec.schedule()}
    out["sequence"] = {"workload": f"{n_seq} frames 512x512x100, 1000 tracers, 50 FISTA x 5 TV each",
                       "tune_s": tune_s, "candidates": len(tried), **legs,
                       "autotuned_speedup": legs["default"]["s_total"] / legs["autotuned"]["s_total"]}
    del hs
    # (2) the mixed batch
    specs = c3_mixed_specs(n_mixed)
    holos = []
    for c, side, pad, n_tr, n_rods, seed in specs:
        inner = dict(c, width=side, height=side)
        h, _ = scenes.tracers_rods_scene(scenes.optics_of(inner), n_tr, n_rods, seed, 0.02, propagate=prop)
        holos.append(torch.from_numpy(np.pad(h, pad)).cuda())
    tuned = {}
    t0 = time.perf_counter()
    for (c, *_), h in zip(specs, holos):
        key = (c["width"], c["num_planes"])
        if key not in tuned:
            tuned[key] = pkg.autotune(pkg.make_params(**{k: v for k, v in c.items()}), h, iterations=10)[0]
    tune_s = time.perf_counter() - t0
    total_it = sum(c["num_inverse_iterations"] for c, *_ in specs)
    voxels = sum(c["width"] * c["height"] * c["num_planes"] * c["num_inverse_iterations"] for c, *_ in specs)
    legs = {}
    for name, mode in (("default", 0), ("autotuned", 2)):
        items = [dict(params=pkg.make_params(**{k: v for k, v in c.items()}), holo=h,
                      schedule=tuned[(c["width"], c["num_planes"])]) for (c, *_), h in zip(specs, holos)]
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        res = pkg.reconstruct_multi(items, autotune=mode)
        torch.cuda.synchronize()
        dt = time.perf_counter() - t0
        legs[name] = {"s_total": dt, "holograms_per_s": n_mixed / dt, "iter_per_s": total_it / dt,
                      "voxel_iterations_per_s": voxels / dt, "all_reached": all(r["reached"] for r in res)}
    out["mixed"] = {"workload": (f"{n_mixed} holograms: sides 384/512/768 + 2 x (32 or 64) zero padding, Nz 40/64/100, "
                                 "L1 0.001-0.01, TV 0.01, 50 FISTA x 5 TV each"),
                    "groups": len(tuned), "tune_s": tune_s, **legs,
                    "schedules": {f"{w}x{w}x{nz}": s.as_dict() for (w, nz), s in sorted(tuned.items())},
                    "autotuned_speedup": legs["default"]["s_total"] / legs["autotuned"]["s_total"]}
    total = out["sequence"]["autotuned"]["s_total"] + out["mixed"]["autotuned"]["s_total"]
    return {"metric": "holograms/s, batched sequence + mixed-size batch, autotuned schedules, FP32",
            "value": (n_seq + n_mixed) / total, "unit": "holograms/s", "n_gpus": 1, "steps": n_seq + n_mixed,
            "warmup": 1, "ms_per_step": 1e3 * total / (n_seq + n_mixed), "higher_is_better": True,
            "scaling": "weak", "vs_baseline": None, "dtype": "f32",
            "data": "synthetic (seeded tracers and rods, angular-spectrum forward model, sigma 0.02)",
            "config": {"workload": f"{n_seq}-frame 512^2x100 sequence + {n_mixed} mixed-size holograms"},
            "summary": out}


def c4_bench(pkg, steps: int, warmup: int, crop=(64, 128, 128), n_pairs: int = 16):
    """Training throughput of the warm-start network at the 256^2 x 64 training workload's
    optics: pairs from the plain solver (200 iterations), Adam steps on crops; the dominant
    cost is the BF16 MFMA convolutions (forward, input and weight gradients), whose FLOPs per
    step are counted from the architecture."""
    import torch
    ws = entry._load_submodule("warmstart")
    xs, ys = ws.make_pairs("c4", 256, n_pairs, 4000001, 200, log=None)
    i_s, o_s, frac = ws.scales(xs, ys)
    D, H, W = crop
    V = D * H * W
    conv_fwd = 0.0
    for lv, c0, c1, co in ((1, 16, 0, 32), (2, 32, 0, 64), (3, 64, 0, 128), (2, 64, 128, 64), (1, 32, 64, 32),
                           (0, 16, 32, 16)):
        conv_fwd += 2.0 * 27 * (c0 + c1) * co * V / 8 ** lv
    first = 2.0 * 27 * 2 * 16 * V  # the input layer's two real channels
    flops_step = 3 * conv_fwd + 2 * first  # forward + input gradients + weight gradients
    with pkg.WarmStartNet(D, H, W, seed=1) as net:
        p, _, _ = net.get_params()
        net.set_params(p, i_s, o_s)
        net.set_loss_weight((1 - frac) / (2 * max(frac, 1e-9)))
        rng = np.random.default_rng(0)
        bx = torch.empty((D, H, W), dtype=torch.complex64, device="cuda")
        by = torch.empty_like(bx)

        def step():
            i = int(rng.integers(0, n_pairs))
            z0 = int(rng.integers(0, xs.shape[1] - D + 1))
            y0 = int(rng.integers(0, 256 - H + 1))
            x0 = int(rng.integers(0, 256 - W + 1))
            bx.copy_(xs[i, z0:z0 + D, y0:y0 + H, x0:x0 + W])
            by.copy_(ys[i, z0:z0 + D, y0:y0 + H, x0:x0 + W])
            return net.train_step(bx, by, 1e-3)
        for _ in range(warmup):
            step()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(steps):
            loss = step()
        torch.cuda.synchronize()
        dt = (time.perf_counter() - t0) / steps
    achieved = flops_step / dt / 1e12
    return {"metric": "warm-start training steps/s (Adam, BF16 MFMA convolutions), 256^2x64 optics",
            "value": 1.0 / dt, "unit": "steps/s", "n_gpus": 1, "steps": steps, "warmup": warmup,
            "ms_per_step": 1e3 * dt, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
            "dtype": "bf16 (fp32 accumulation and master weights)",
            "data": f"synthetic ({n_pairs} pairs from the plain solver, 200 iterations, C4 optics)",
            "config": {"workload": f"one Adam step on a {D}x{H}x{W} crop", "last_loss": loss},
            "roofline": {"bound": "mfma", "kernel": "train step (convolutions)", "achieved": achieved,
                         "peak": 2500.0, "unit": "TFLOP/s", "frac": achieved / 2500.0, "traffic": None,
                         "algorithmic_work_per_launch": flops_step}}


def background_bench(pkg, steps: int, warmup: int, n_frames: int = 200, size: int = 1024):
    """make-background's mean background over n_frames device-resident 1024^2 frames
    (rihvr_background_mean): frames/s, and the two passes' compulsory bytes (read every
    frame twice, write the background) per call.  The call's own device allocations and
    the copy of the background to the host are inside the timed region."""
    import torch
    g = torch.Generator(device="cuda").manual_seed(20261204)
    frames = torch.randint(0, 256, (n_frames, size, size), generator=g, device="cuda").float()
    for _ in range(warmup):
        pkg.background_mean(frames)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        pkg.background_mean(frames)
    torch.cuda.synchronize()
    dt = (time.perf_counter() - t0) / steps
    nbytes = 2.0 * n_frames * size * size * 4 + size * size * 4
    return {"metric": f"background frames/s (mean + statistics of {n_frames} frames {size}^2, FP32)",
            "value": n_frames / dt, "unit": "frames/s", "n_gpus": 1, "steps": steps, "warmup": warmup,
            "ms_per_step": 1e3 * dt, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
            "dtype": "f32", "data": "synthetic (uniform grey values)",
            "config": {"workload": f"make-background mean path, {n_frames} frames of {size}x{size}"},
            "achieved_GBps_wall": nbytes / dt / 1e9, "algorithmic_bytes_per_call": nbytes}


def preprocess_bench(pkg, steps: int, warmup: int, n_frames: int = 100, size: int = 1024):
    """The input contract on the device (rihvr_prepare_holograms) for n_frames 8-bit
    device-resident frames of size^2 into complex64 holograms on the device: frames/s.
    Compulsory bytes: the frames twice (mean pass, write pass) + the holograms."""
    import torch
    g = torch.Generator(device="cuda").manual_seed(20261205)
    frames = torch.randint(0, 256, (n_frames, size, size), generator=g, device="cuda").to(torch.uint8)
    out = torch.empty((n_frames, size, size), dtype=torch.complex64, device="cuda")
    fp = pkg.FileParams()
    fp.roi_size = -1
    for _ in range(warmup):
        pkg.prepare_holograms(fp, frames, out=out)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        pkg.prepare_holograms(fp, frames, out=out)
    torch.cuda.synchronize()
    dt = (time.perf_counter() - t0) / steps
    nbytes = n_frames * size * size * (2.0 * 1 + 8)
    return {"metric": f"preprocessing frames/s ({n_frames} 8-bit frames {size}^2 -> complex64 holograms)",
            "value": n_frames / dt, "unit": "frames/s", "n_gpus": 1, "steps": steps, "warmup": warmup,
            "ms_per_step": 1e3 * dt, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
            "dtype": "u8->f32", "data": "synthetic (uniform grey values)",
            "config": {"workload": f"rescale /255, I/mean - 1, no crop, no padding; {n_frames} frames of {size}x{size}"},
            "achieved_GBps_wall": nbytes / dt / 1e9, "algorithmic_bytes_per_call": nbytes}


def segmentation_bench(pkg, scenes, steps: int, warmup: int, iterations: int = 100, cpu: bool = True):
    """The segmentation executable's chain on the 1024^2 x 128 reconstruction of the
    seeded tracers-and-rods hologram (iterations of plain FISTA+TV, mode 7), read from
    the device volume: load, binarize (min intensity 1), close (radius 3), label,
    extract, size filter (2) per step, the host pass over the object voxels included.
    Per-stage device times come from HIP events on the segmentation stream."""
    import torch
    cfg = dict(scenes.C2, mode=7)
    W, H, Nz = cfg["width"], cfg["height"], cfg["num_planes"]
    N = W * H * Nz
    p = pkg.make_params(**dict(cfg, num_inverse_iterations=iterations))
    vol = torch.empty((Nz, H, W), dtype=torch.complex64, device="cuda")
    with pkg.SparseCompressiveHolo(p) as rec:
        rec.inverseReconstruct(torch.from_numpy(scenes.c2_hologram()).cuda())
        rec.getVolume(out=vol)
    torch.cuda.synchronize()
    nnz = int(torch.count_nonzero(vol).item())
    seg = pkg.SparseSegmentation(W, H, Nz)

    def chain():
        seg.loadDense(vol)
        seg.binarize(1.0)
        seg.close(3)
        seg.labelConnectedComponents()
        cloud = seg.extractObjects()
        n_obj = cloud.getNumObjects()
        removed = cloud.blobSizeFilter(2)
        return n_obj, removed

    for _ in range(warmup):
        chain()
    torch.cuda.synchronize()
    stage = dict.fromkeys(("load", "binarize", "close", "label", "extract"), 0.0)
    t0 = time.perf_counter()
    for _ in range(steps):
        n_obj, removed = chain()
        for k, v in seg.timing().items():
            stage[k] += v / steps
    torch.cuda.synchronize()
    dt = (time.perf_counter() - t0) / steps
    fg = int(seg.mask().sum())
    seg.destroy()
    # compulsory HBM bytes per stage (N voxels, bit volume N/8 bytes, H*Nz row flags of 4
    # bytes): load reads the complex volume and the row flags (keys are written only where
    # non-zero); binarize reads the row flags and writes the bit volume (the keys of the
    # flagged rows, data-dependent, are not counted); close reads and writes the bit volume
    # twice; label reads the bit volume three times (labels are touched only at object
    # voxels); extract reads it three times (counts, roots, scatter)
    rows = H * Nz
    stage_bytes = {"load": 8.0 * N + 4.0 * rows, "binarize": 4.0 * rows + N / 8, "close": 4 * N / 8,
                   "label": 3 * N / 8, "extract": 3 * N / 8}
    dom = max(stage, key=stage.get)
    achieved = stage_bytes[dom] / (stage[dom] * 1e-3) / 1e9
    out = {"metric": "segmentation volumes/s (1024^2x128 reconstruction -> centroids)",
           "value": 1.0 / dt, "unit": "volumes/s", "n_gpus": 1, "steps": steps, "warmup": warmup,
           "ms_per_step": 1e3 * dt, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
           "dtype": "f32/u8/u32", "data": f"synthetic (C2 tracers+rods hologram, {iterations} FISTA+TV iterations)",
           "config": {"workload": f"segmentation chain of the {W}x{H}x{Nz} reconstruction: min intensity 1, "
                                  f"close radius 3, min area 2", "nonzero_voxels": nnz, "object_voxels": fg,
                      "objects": n_obj, "objects_removed_by_size": removed},
           "stages_ms": stage, "gpu_ms": sum(stage.values()),
           "roofline": {"bound": "hbm", "kernel": f"stage:{dom}", "achieved": achieved, "peak": HBM_PEAK_GBPS,
                        "unit": "GB/s", "frac": achieved / HBM_PEAK_GBPS, "traffic": None,
                        "avg_launch_ms": stage[dom], "algorithmic_bytes_per_launch": stage_bytes[dom]},
           "chain_GBps": sum(stage_bytes.values()) / (sum(stage.values()) * 1e-3) / 1e9}
    if cpu:
        out["cpu_baseline"] = segmentation_cpu_baseline(vol)
    return out


def torch_argmax_abs(vol):
    return vol.abs().reshape(-1).argmax().item()


def segmentation_cpu_baseline(vol, box=(32, 128, 128)):
    """The oracle's restatement of the chain (numpy, one process) on a box of the
    same reconstruction; reported per voxel and scaled to volumes/s of the full size."""
    import oracle.segmentation as so
    Nz, H, W = box
    fullN = vol.numel()
    # the box around the brightest voxel, inside the volume
    c = np.unravel_index(int(torch_argmax_abs(vol)), tuple(vol.shape))
    z0, y0, x0 = (int(min(max(c[k] - n // 2, 0), vol.shape[k] - n)) for k, n in enumerate(box))
    sub = vol[z0:z0 + Nz, y0:y0 + H, x0:x0 + W].cpu().numpy()
    z, y, x = np.nonzero(sub)
    v = sub[z, y, x]
    inten = (v.real * v.real + v.imag * v.imag).astype(np.float32)
    t0 = time.perf_counter()
    so.segment(z.astype(np.int32), (y * W + x).astype(np.int32), inten, W, H, Nz, 1.0, 3, 2)
    dt = time.perf_counter() - t0
    per_voxel = dt / (Nz * H * W)
    return {"value": 1.0 / (per_voxel * fullN), "unit": "volumes/s", "cores": 1, "kind": "port",
            "sample": f"oracle chain on the {W}x{H}x{Nz} box around the brightest voxel of the same reconstruction "
                      f"({len(z)} non-zero voxels, {dt:.1f} s), scaled per voxel"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10,
                    help="timed steps; a step of the c2/c1 workloads is one whole solve")
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--workload", choices=["c2", "c1", "c3", "c4", "c5", "background", "segmentation", "preprocess"],
                    default="c2")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--frames", type=int, default=3,
                    help="frames of the s/frame-to-tolerance leg (0 skips it)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; --gpus must be 1",
                          "n_gpus_requested": args.gpus}))
        sys.exit(2)

    import torch
    if not torch.cuda.is_available():
        print(json.dumps({"error": "no GPU visible"}))
        sys.exit(1)

    pkg = entry.load_package()
    if args.workload == "background":
        print(json.dumps(background_bench(pkg, max(1, min(args.steps, 20)), min(args.warmup, 3))), flush=True)
        return
    if args.workload == "preprocess":
        print(json.dumps(preprocess_bench(pkg, max(1, min(args.steps, 20)), min(args.warmup, 3))), flush=True)
        return
    scenes = entry._load_submodule("scenes")
    if args.workload == "c3":
        print(json.dumps(c3_bench(pkg, scenes)), flush=True)
        return
    if args.workload == "c5":
        print(json.dumps(c5_bench(pkg, scenes, max(1, args.steps))), flush=True)
        return
    if args.workload == "c4":
        print(json.dumps(c4_bench(pkg, max(1, args.steps * 10), max(1, args.warmup))), flush=True)
        return
    if args.workload == "segmentation":
        print(json.dumps(segmentation_bench(pkg, scenes, max(1, min(args.steps, 20)), min(args.warmup, 3),
                                            cpu=not args.no_cpu)), flush=True)
        return
    cfg = dict(scenes.C2 if args.workload == "c2" else scenes.C1, mode=7)
    holo = scenes.c2_hologram() if args.workload == "c2" else scenes.c1_hologram()
    W, H, Nz = cfg["width"], cfg["height"], cfg["num_planes"]

    n_it = cfg["num_inverse_iterations"]
    rec = pkg.SparseCompressiveHolo(pkg.make_params(**cfg))
    holo_dev = torch.from_numpy(holo).cuda()
    torch.cuda.synchronize()
    # warm-up solves; the first one with every kernel timed names the dominant kernel
    dom_name = "prox"
    for w in range(args.warmup):
        rec.begin(holo_dev)
        if w == 0:
            rec.set_timing(True)
        rec.iterate(n_it)
        if w == 0:
            warm_timing = rec.get_timing()
            dom_name = max(warm_timing.items(), key=lambda kv: kv[1][1])[0]
            rec.set_timing(False)
        rec.finish()
    # the timed region: K whole solves of n_it iterations each; the solve's set-up (begin:
    # hologram check, measurement spectrum, zero start, initial residual) and its finish
    # (returned momentum point, final objective) run between the timed spans.  HIP events
    # bracket the dominant kernel's launches only (two per launch, on the solver's stream).
    rec.set_timing(True, only=dom_name)
    elapsed = 0.0
    setup_s = 0.0
    bt_timed = rs_timed = 0
    for _ in range(args.steps):
        torch.cuda.synchronize()
        ta = time.perf_counter()
        rec.begin(holo_dev)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        st = rec.iterate(n_it)
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        fin = rec.finish()
        torch.cuda.synchronize()
        elapsed += t1 - t0
        setup_s += (t0 - ta) + (time.perf_counter() - t1)
        bt_timed += int(np.sum(st["backtracks"]))
        rs_timed += int(np.sum(st["restarts"]))
    timing = rec.get_timing()
    # after the timed region: one more solve with every kernel timed, for the breakdown
    rec.set_timing(True)
    rec.begin(holo_dev)
    rec.iterate(n_it)
    breakdown = rec.get_timing()
    rec.set_timing(False)
    fin = rec.finish()
    n_after = n_it
    # fraction of 16-voxel lines holding a non-zero in the returned volume
    vol = torch.empty((Nz, H, W), dtype=torch.complex64, device="cuda")
    rec.getVolume(out=vol)
    f_line = float((vol != 0).reshape(Nz, H, W // 16, 16).any(-1).float().mean().item())
    err64 = fp64_error(vol, holo) if args.workload == "c2" else None
    del vol
    schedule = rec.schedule()
    rec.destroy()

    ips = args.steps * n_it / elapsed
    zchunk = schedule["zchunk_fwd"]  # planes per partial spectrum
    nparts = (Nz + zchunk - 1) // zchunk
    dom_launches, dom_ms = timing[dom_name]
    dom_avg_ms = dom_ms / max(1, dom_launches)
    dom_bytes = algorithmic_bytes(dom_name, W, H, Nz, nparts, f_line)
    bound = KERNEL_BOUND.get(dom_name, "hbm")
    if bound == "valu":
        dom_work = prox_flops_per_voxel(cfg["num_tv_iterations"]) * float(W * H * Nz)
        achieved = dom_work / (dom_avg_ms * 1e-3) / 1e12
        peak, unit = FP32_PEAK_TFLOPS, "TFLOP/s"
    else:
        dom_work = dom_bytes
        achieved = dom_bytes / (dom_avg_ms * 1e-3) / 1e9
        peak, unit = HBM_PEAK_GBPS, "GB/s"
    it_bytes = iteration_bytes(W, H, Nz)
    it_gbps = it_bytes * ips / 1e9
    gpu_ms_after = sum(v[1] for v in breakdown.values())
    # HBM traffic of the dominant kernel from the committed rocprofv3 counter passes
    traffic = None
    valu_issue = None
    counters = os.path.join(ROOT, "profiles", "c2_hbm_counters.json")
    if args.workload == "c2" and os.path.exists(counters):
        kc = json.load(open(counters)).get("kernels", {}).get(dom_name)
        traffic = kc["hbm_bytes"] if kc else None
        valu_issue = kc.get("valu_issue_frac") if kc else None

    out = {
        "metric": METRIC,
        "value": ips,
        "unit": "iter/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1e3 * elapsed / args.steps,
        "step": (f"one whole {n_it}-iteration solve (rihvr_iterate({n_it}) after rihvr_begin; value = "
                 f"{n_it} x steps / the summed iterate time)"),
        "s_per_frame_with_setup": (elapsed + setup_s) / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded 500 tracers + 20 rods, angular-spectrum forward model, sigma 0.02)",
        "config": {
            "workload": f"plain FISTA+TV (mode 7), {W}x{H}x{Nz} FP32, z0={cfg['start_plane']} dz={cfg['plane_stepsize']} "
                        f"dx={cfg['resolution']} lambda={cfg['wavelength']}, L1 {cfg['regularization_param']}, "
                        f"TV {cfg['regularization_TV_param']}, {cfg['num_tv_iterations']} TV iterations",
            "width": W, "height": H, "num_planes": Nz,
        },
        "schedule": schedule,
        "roofline": {
            "bound": bound,
            "kernel": dom_name,
            "achieved": achieved,
            "peak": peak,
            "unit": unit,
            "frac": achieved / peak,
            "traffic": traffic,
            "traffic_source": "profiles/c2_hbm_counters.json (rocprofv3 FETCH_SIZE x2.0 + WRITE_SIZE x1.0, timed-region iterations)" if traffic else None,
            "avg_launch_ms": dom_avg_ms,
            "algorithmic_work_per_launch": dom_work,
            "algorithmic_bytes_per_launch": dom_bytes,
            "hbm_achieved_GBps": dom_bytes / (dom_avg_ms * 1e-3) / 1e9,
            "valu_issue_frac": valu_issue,
            "note": ("FP32 operations counted per voxel (prox_flops_per_voxel) against the published FP32 vector "
                     "peak, which assumes packed dual-issue FP32; this build compiles without packed FP32, so "
                     "its issue ceiling is half of it; valu_issue_frac is the measured fraction of that ceiling"
                     if bound == "valu" else None),
            "line_fraction": f_line,
        },
        "iteration_roofline": {
            "bytes_per_iteration": it_bytes,
            "achieved_GBps": it_gbps,
            "hbm_frac": it_gbps / HBM_PEAK_GBPS,
            "fp32_tflops": iteration_flops(W, H, Nz) * ips / 1e12,
            "fp32_frac": iteration_flops(W, H, Nz) * ips / 1e12 / FP32_PEAK_TFLOPS,
        },
        "kernels_ms_per_iteration": {k: v[1] / n_after for k, v in sorted(breakdown.items(), key=lambda kv: -kv[1][1])
                                     if v[0] > 0},
        "kernels_breakdown_source": f"every kernel timed over one more {n_after}-iteration solve after the timed region",
        "gpu_ms_per_iteration_after": gpu_ms_after / n_after,
        "solver": {"iterations_per_solve": n_it, "solves_total": args.warmup + args.steps + 1,
                   "backtracks_timed": bt_timed, "restarts_timed": rs_timed,
                   "stepsize": st["stepsize"], "f_final": fin["f_final"], "sparsity": fin["sparsity"],
                   "error_vs_fp64_oracle": err64,
                   "error_vs_fp64_oracle_source": "relative L2 of the returned volume against the FP64 oracle's "
                                                  "solution of the same hologram (tests/golden/c2_oracle_f64.npz)"},
    }
    if args.frames > 0:
        out["frames"] = frames_leg(pkg, scenes, cfg, args.frames)
    if not args.no_cpu:
        try:
            out["cpu_baseline"] = cpu_baseline(cfg, holo)
        except Exception as e:  # reported, never substituted for the GPU value
            out["cpu_baseline"] = {"error": repr(e)}
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
