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
