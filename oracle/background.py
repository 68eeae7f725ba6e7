"""CPU restatement of the reference's background estimation (make-background).

TEST INFRASTRUCTURE ONLY: imported by tests/ as the checker of the library's
background functions and of the make-background executable.

- mean_background: HoloSequence::useBackgroundSubtraction
  (src/core/holo_sequence.cu:309-437): B = running float sum of the frames in order,
  times float(1/n) (cv::Mat /= n); per frame cv::meanStdDev of I - B (float
  subtraction, double moments, population standard deviation); alpha = mean of the
  means - 5 * mean of the std devs, beta = 255 / (2 * mean std * 5) with
  bg_scale1 = 5 and bg_min_grey = 0 (:22-23, :414-415); average = cv::mean(B).
- frame_ids: the frames loadNextChunk visits (:138-160) for useBackgroundSubtraction's
  start/count (:332-352), with one chunk (the whole set fits in memory).
- median_schedule: median_background_calc (main_create_background.cu:104-193): the
  frames in the five-slot stack for each output id, with the reference's slot order.
"""
from __future__ import annotations

import numpy as np


def mean_background(frames):
    f = np.asarray(frames, np.float32)
    n = f.shape[0]
    s = f[0].copy()
    for k in range(1, n):
        s = (s + f[k]).astype(np.float32)
    bg = (s * np.float32(1.0 / n)).astype(np.float32)
    means, stds = [], []
    for k in range(n):
        d = (f[k] - bg).astype(np.float32).astype(np.float64)
        m = d.mean()
        var = (d * d).mean() - m * m
        means.append(m)
        stds.append(np.sqrt(max(var, 0.0)))
    mm, ms = float(np.mean(means)), float(np.mean(stds))
    stats = dict(alpha=mm - 5.0 * ms, beta=255.0 / (2.0 * ms * 5.0), minimum=0.0,
                 average=float(bg.astype(np.float64).mean()))
    return bg, stats


def frame_ids(start_image, image_count, background_start, background_end):
    start, total = start_image, image_count
    if background_start >= 0 and background_end >= 0:
        start, total = background_start, background_end - background_start - 1
    step = total // 200 if total > 500 else 1
    count = total // step if total > 0 else 0
    return [start + k * step for k in range(count)], step


def median_schedule(background_start, background_end):
    """[(output id, [frame ids in the stack])] in the order the reference writes them."""
    num = background_end - background_start
    stack = [background_start + i for i in range(5)]
    out = [(background_start + i, list(stack)) for i in range(3)]
    for i in range(3, num - 2):
        stack[(i - 3) % 5] = background_start + i
        out.append((background_start + i, list(stack)))
    return out


def median5(frames5):
    return np.median(np.asarray(frames5, np.float32), axis=0).astype(np.float32)


def to_u8(v):
    """convertTo(CV_8U): round half to even, saturate."""
    return np.clip(np.rint(np.asarray(v, np.float64)), 0, 255).astype(np.uint8)


# ---- correlation background (ccbg.cpp computeBackgrounds) -------------------------------
def box_resize(frames, k):
    """cv::resize INTER_AREA at the integer factor k: each k x k block's values added row by
    row in float32, times float32(1 / k^2)."""
    f = np.asarray(frames, np.float32)
    s = np.zeros(f.shape[:-2] + (f.shape[-2] // k, f.shape[-1] // k), np.float32)
    for dy in range(k):
        for dx in range(k):
            s = (s + f[..., dy::k, dx::k]).astype(np.float32)
    return (s * np.float32(1.0 / (k * k))).astype(np.float32)


def entropy_f32(hist, n_pixels):
    """computeEntropy: p = float(h) / N; e -= p * log(p) over the non-empty bins in order (float32)."""
    e = np.float32(0.0)
    for h in hist:
        if h > 0:
            p = np.float32(h) / np.float32(n_pixels)
            e = np.float32(e - np.float32(p * np.float32(np.log(p))))
    return e


def pick_frame_count(entropies, min_frames, mmax):
    """The 3-point moving average of the entropies, the first index whose smoothed entropy falls,
    clamped to [min_frames - 1, mmax - 1] (and to mmax - 1): the index of the last frame used."""
    ent = [np.float32(e) for e in entropies]
    sm = []
    for i in range(len(ent)):
        s, c = ent[i], 1
        if i > 0:
            s, c = np.float32(s + ent[i - 1]), c + 1
        if i + 1 < len(ent):
            s, c = np.float32(s + ent[i + 1]), c + 1
        sm.append(np.float32(s / np.float32(c)))
    opt = 1
    for i in range(1, len(sm)):
        if np.float32(sm[i] - sm[i - 1]) < 0:
            opt = i
            break
    opt = max(min_frames - 1, min(opt, mmax - 1))
    return min(opt, mmax - 1)


def ranking(small):
    """The other frames of each frame by normalised cross-correlation, highest first (ties by
    frame number); correlations of the centred images summed in float64."""
    x = np.asarray(small, np.float64).reshape(len(small), -1)
    x = x - x.mean(axis=1, keepdims=True)
    g = x @ x.T
    n = len(x)
    out = []
    for f in range(n):
        c = [(g[f, j] / (np.sqrt(g[f, f] * g[j, j]) + 1e-8), j) for j in range(n) if j != f]
        c.sort(key=lambda t: -t[0])  # stable: equal correlations keep frame order
        out.append([j for _, j in c])
    return out


def correlation_backgrounds(frames, min_frames, max_frames, k, enhanced=False):
    """(backgrounds uint8 [n][H][W], frames used [n], enhanced uint8 or None)."""
    f = np.asarray(frames, np.float32)
    n = f.shape[0]
    small = box_resize(f, k)
    P = small.shape[1] * small.shape[2]
    xs = small.reshape(n, -1)
    order = ranking(small)
    mmax = min(max_frames, n - 1)
    bgs = np.zeros(f.shape, np.uint8)
    enh = np.zeros(f.shape, np.uint8) if enhanced else None
    used = np.zeros(n, np.int32)
    for i in range(n):
        bg = np.zeros(P, np.float32)
        ents = []
        for m in range(mmax):
            bg = (bg + xs[order[i][m]]).astype(np.float32)
            avg = (bg / np.float32(m + 1)).astype(np.float32)
            v = (xs[i] - avg).astype(np.float32).astype(np.float64)
            r = np.trunc(v + np.copysign(0.5, v))  # std::round: halves away from zero
            b = np.clip(r, 0, 255).astype(np.int64)
            ents.append(entropy_f32(np.bincount(b, minlength=256), P))
        opt = pick_frame_count(ents, min_frames, mmax)
        used[i] = opt + 1
        s = np.zeros(f.shape[1:], np.float32)
        for kk in range(opt + 1):
            s = (s + f[order[i][kk]]).astype(np.float32)
        bf = (s / np.float32(opt + 1)).astype(np.float32)
        bgs[i] = np.clip(np.rint(bf), 0, 255).astype(np.uint8)
        if enhanced:
            d = (f[i] - bf).astype(np.float32)
            lo, hi = float(d.min()), float(d.max())
            sc = 255.0 / (hi - lo) if hi > lo else 0.0
            vf = (d.astype(np.float64) * sc - lo * sc).astype(np.float32)
            enh[i] = np.clip(np.rint(vf), 0, 255).astype(np.uint8)
    return bgs, used, enh
