"""Training and evaluation of the learned warm start (rihvr_net_* / rihvr_reconstruct_learned).

Pairs come from the plain solver itself: for a seeded synthetic hologram b of tracers and rods,
the network's input is A^T b (rihvr_adjoint) and its target the volume the plain FISTA+TV solve
returns after ``target_iterations`` iterations.  Training is Adam on the network's BF16 MFMA
kernels (rihvr_net_train_step), one pair per step, on random (z, y, x) crops of the pairs.
Everything stays on the device; the pairs of a 256^2 x 128 set take 64 MiB each.

Command line (one GPU), from the repository root:
    python <package dir>/warmstart.py train --optics c2 --size 256 --pairs 256 --epochs 8 --out w.npz
    python <package dir>/warmstart.py eval --weights w.npz --size 1024 --frames 4
"""
from __future__ import annotations

import argparse
import json
import math
import sys
import time

import numpy as np

if __package__ in (None, ""):
    import pathlib
    sys.path.insert(0, str(pathlib.Path(__file__).resolve().parents[1]))
    import __graft_entry__ as _entry  # noqa: E402
    pkg = _entry.load_package()
    scenes = _entry._load_submodule("scenes")
else:  # imported as part of the package
    from . import scenes  # noqa: F401
    import importlib
    pkg = importlib.import_module(__package__)

# Optics of the workloads: C2/C5 (1024^2 x 128 at dx 1.1 um) and C4 (256^2 x 64 at dx 2 um)
OPTICS = {
    "c2": dict(num_planes=128, start_plane=2500.0, plane_stepsize=20.0, wavelength=0.632, resolution=1.1,
               regularization_param=0.15, regularization_TV_param=0.1, num_tv_iterations=5),
    "c4": dict(num_planes=64, start_plane=1000.0, plane_stepsize=40.0, wavelength=0.633, resolution=2.0,
               regularization_param=0.01, regularization_TV_param=0.01, num_tv_iterations=5),
}
# objects per 1024^2 of field (C5: 300-800 tracers and rods; C4: 50-300 disks and rods per 256^2)
DENSITY = {"c2": (300, 800, 1024 * 1024), "c4": (50, 300, 256 * 256)}


def config(optics: str, size: int, iterations: int = 100) -> dict:
    return dict(OPTICS[optics], width=size, height=size, num_inverse_iterations=iterations, mode=7)


def scene(optics: str, cfg: dict, seed: int, propagate=None):
    """A seeded hologram of tracers and rods at the workload's object density, the rods a
    fraction drawn uniformly from 2 % to 30 % of the objects (propagate: the forward model's
    plane sum, numpy by default; scenes.propagate_sum_torch runs it on the GPU)."""
    rng = np.random.default_rng(seed)
    lo, hi, area = DENSITY[optics]
    n = int(round(rng.uniform(lo, hi) * cfg["width"] * cfg["height"] / area))
    n_rods = max(1, int(round(n * rng.uniform(0.02, 0.30))))
    o = scenes.optics_of(cfg)
    kw = {} if propagate is None else {"propagate": propagate}
    if optics == "c4":
        holo, _ = scenes.tracers_rods_scene(o, n - n_rods, n_rods, seed, 0.02, tracer_diam=(4.0, 12.0), **kw)
    else:
        holo, _ = scenes.tracers_rods_scene(o, n - n_rods, n_rods, seed, 0.02, **kw)
    return holo


def make_pairs(optics: str, size: int, n_pairs: int, seed: int, target_iterations: int = 200, log=print):
    """(inputs, targets): device complex64 tensors [n][Nz][size][size] (A^T b, x_target)."""
    import torch
    cfg = config(optics, size, target_iterations)
    shape = (cfg["num_planes"], size, size)
    xs = torch.empty((n_pairs,) + shape, dtype=torch.complex64, device="cuda")
    ys = torch.empty_like(xs)
    t0 = time.time()
    with pkg.SparseCompressiveHolo(pkg.make_params(**cfg)) as rec:
        for i in range(n_pairs):
            h = torch.from_numpy(scene(optics, cfg, seed + i)).cuda()
            rec.adjoint(h, out=xs[i])
            rec.inverseReconstruct(h)
            rec.getVolume(out=ys[i])
            if log and (i + 1) % max(1, n_pairs // 8) == 0:
                log(f"pairs {i + 1}/{n_pairs} ({time.time() - t0:.0f} s)")
    torch.cuda.synchronize()
    return xs, ys


def scales(xs, ys):
    """in_scale = 1 / rms(A^T b); out_scale = rms of the targets' non-zero voxels; the
    fraction of non-zero target voxels."""
    import torch
    in_rms = float(torch.sqrt(torch.mean(torch.abs(xs[: min(len(xs), 16)]) ** 2)).item())
    t = torch.abs(ys[: min(len(ys), 16)])
    nz = t[t > 0]
    out_rms = float(torch.sqrt(torch.mean(nz ** 2)).item()) if nz.numel() else 1.0
    return 1.0 / max(in_rms, 1e-30), max(out_rms, 1e-30), nz.numel() / max(1, t.numel())


def train(net, xs, ys, epochs: int, lr: float, crop, seed: int = 0, log=print, lr_min_frac=0.1):
    """Adam over random crops (crop = (Dz, Hc, Wc), multiples of 8) with a cosine learning
    rate from lr to lr_min_frac * lr.  Returns the per-epoch mean losses."""
    import torch
    rng = np.random.default_rng(seed)
    n = len(xs)
    D, H, W = xs.shape[1:]
    cz, cy, cx = crop
    bx = torch.empty((cz, cy, cx), dtype=torch.complex64, device="cuda")
    by = torch.empty_like(bx)
    total = epochs * n
    step = 0
    hist = []
    t0 = time.time()
    for ep in range(epochs):
        order = rng.permutation(n)
        acc = 0.0
        for i in order:
            z0 = int(rng.integers(0, D - cz + 1))
            y0 = int(rng.integers(0, H - cy + 1))
            x0 = int(rng.integers(0, W - cx + 1))
            bx.copy_(xs[i, z0:z0 + cz, y0:y0 + cy, x0:x0 + cx])
            by.copy_(ys[i, z0:z0 + cz, y0:y0 + cy, x0:x0 + cx])
            a = lr * (lr_min_frac + (1 - lr_min_frac) * 0.5 * (1 + math.cos(math.pi * step / max(1, total))))
            acc += net.train_step(bx, by, a)
            step += 1
        hist.append(acc / n)
        if log:
            log(f"epoch {ep + 1}/{epochs}: loss {hist[-1]:.5g} ({time.time() - t0:.0f} s)")
    return hist


def evaluate(net, optics: str, size: int, n_frames: int, seed: int, iterations: int = 100, tol: float = 1e-3,
             log=print, propagate=None, progress=None):
    """Per frame: the plain solve's f(iterations) sets the target f <= (1 + tol) f_plain; then
    plain cold start and the learned start each run to it (capped at iterations).  Returns a
    summary with s/frame and iteration counts.  progress(k) is called after frame k."""
    import torch
    cfg = config(optics, size, iterations)
    rows = []
    with pkg.SparseCompressiveHolo(pkg.make_params(**cfg)) as rec:
        for k in range(n_frames):
            h = torch.from_numpy(scene(optics, cfg, seed + k, propagate)).cuda()
            plain = rec.reconstructBatch(h[None])[0]
            target = (1.0 + tol) * plain["f_final"]
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            cold = rec.reconstructBatch(h[None], f_target=[target], max_iterations=iterations)[0]
            torch.cuda.synchronize()
            t_cold = time.perf_counter() - t0
            t0 = time.perf_counter()
            lr_ = rec.reconstructLearned(net, h, f_target=target, max_iterations=iterations)
            torch.cuda.synchronize()
            t_learn = time.perf_counter() - t0
            rows.append(dict(f_plain=plain["f_final"], cold_iterations=cold["iterations"], cold_s=t_cold,
                             learned_iterations=lr_["iterations"], learned_s=t_learn,
                             fallback=lr_["fallback"],
                             reached=lr_["reached"], f_learned=lr_["f_final"],
                             f_adjoint=lr_["f_adjoint"], f_zero=lr_["f_zero"],
                             tau=lr_["tau"], f_thresholded=lr_["f_thresholded"],
                             ms_network=lr_["ms_network"], ms_adjoint=lr_["ms_adjoint"], ms_solve=lr_["ms_solve"]))
            if log:
                log(json.dumps(rows[-1]))
            if progress:
                progress(k)
    s = dict(frames=n_frames, size=size, optics=optics, tolerance=tol,
             cold_s_per_frame=float(np.mean([r["cold_s"] for r in rows])),
             learned_s_per_frame=float(np.mean([r["learned_s"] for r in rows])),
             cold_iterations=[r["cold_iterations"] for r in rows],
             learned_iterations=[r["learned_iterations"] for r in rows],
             fallback_rate=float(np.mean([r["fallback"] for r in rows])),
             reached=[r["reached"] for r in rows], rows=rows)
    s["speedup"] = s["cold_s_per_frame"] / s["learned_s_per_frame"]
    return s


def main(argv=None):
    ap = argparse.ArgumentParser()
    sub = ap.add_subparsers(dest="cmd", required=True)
    t = sub.add_parser("train")
    t.add_argument("--optics", choices=list(OPTICS), default="c2")
    t.add_argument("--size", type=int, default=256)
    t.add_argument("--pairs", type=int, default=256)
    t.add_argument("--target-iterations", type=int, default=200)
    t.add_argument("--epochs", type=int, default=8)
    t.add_argument("--lr", type=float, default=2e-3)
    t.add_argument("--crop", type=int, nargs=3, default=None)
    t.add_argument("--seed", type=int, default=1)
    t.add_argument("--pos-weight", type=float, default=-1.0,
                   help="loss weight of the target's non-zero voxels (-1: (1 - f) / (2 f), f their fraction)")
    t.add_argument("--out", required=True)
    t.add_argument("--eval-frames", type=int, default=4)
    e = sub.add_parser("eval")
    e.add_argument("--weights", required=True)
    e.add_argument("--optics", choices=list(OPTICS), default="c2")
    e.add_argument("--size", type=int, default=256)
    e.add_argument("--frames", type=int, default=4)
    e.add_argument("--seed", type=int, default=900001)
    a = ap.parse_args(argv)

    def log(m):
        print(m, flush=True)

    import torch
    torch.zeros(1, device="cuda")  # the GPU context is torch's before the library's first call

    if a.cmd == "train":
        xs, ys = make_pairs(a.optics, a.size, a.pairs, a.seed * 100000, a.target_iterations, log)
        i_s, o_s, frac = scales(xs, ys)
        pw = a.pos_weight if a.pos_weight >= 0 else (1.0 - frac) / (2.0 * max(frac, 1e-9))
        log(f"in_scale {i_s:.4g} out_scale {o_s:.4g} non-zero fraction {frac:.3g} pos_weight {pw:.4g}; "
            f"loss of a zero prediction {frac * (1 + pw) / 2:.4g}")
        crop = tuple(a.crop) if a.crop else tuple(xs.shape[1:])
        with pkg.WarmStartNet(*crop) as net:
            net.init(a.seed)
            p, _, _ = net.get_params()
            net.set_params(p, i_s, o_s)
            net.set_loss_weight(pw)
            hist = train(net, xs, ys, a.epochs, a.lr, crop, a.seed, log)
            net.save(a.out, optics=a.optics, size=a.size, pairs=a.pairs, epochs=a.epochs, losses=np.array(hist),
                     pos_weight=pw)
            params, i_s, o_s = net.get_params()
        del xs, ys
        if a.eval_frames > 0:
            shape = (OPTICS[a.optics]["num_planes"], a.size, a.size)
            with pkg.WarmStartNet(*shape) as net:
                net.set_params(params, i_s, o_s)
                log(json.dumps({k: v for k, v in evaluate(net, a.optics, a.size, a.eval_frames, 900001, log=log).items()
                                if k != "rows"}))
    else:
        with np.load(a.weights, allow_pickle=False) as z:
            params, i_s, o_s = z["params"], float(z["in_scale"]), float(z["out_scale"])
        shape = (OPTICS[a.optics]["num_planes"], a.size, a.size)
        with pkg.WarmStartNet(*shape) as net:
            net.set_params(params, i_s, o_s)
            log(json.dumps({k: v for k, v in evaluate(net, a.optics, a.size, a.frames, a.seed, log=log).items()
                            if k != "rows"}))


if __name__ == "__main__":
    main()
