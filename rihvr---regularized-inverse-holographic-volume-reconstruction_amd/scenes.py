"""Seeded synthetic holograms for the benchmark and test workloads.

Objects (opaque disks, point tracers, rods) are rasterised into per-plane
opacity masks on the reconstruction grid z_i = z0 + i*dz.  Each plane is
propagated to the hologram plane with the angular-spectrum kernel
exp(i 2 pi z (sqrt(1 - lambda^2 f^2 / dx^2) - 1) / lambda) -- the same
convention as the solver's forward operator, so objects refocus at their
plane -- and superposed on a unit background (first-Born model):
    U = 1 + sum_i IFFT( FFT(-m_i) * P_{z_i} ),   I = |U|^2 + N(0, sigma^2).
The solver input is b = I / mean(I) - 1 (the reference's normalize_mean0,
src/core/hologram.cu:332-358), complex64 with zero imaginary part.
"""
from __future__ import annotations

import dataclasses
import os

import numpy as np

try:
    import scipy.fft as _fft

    def _fft2(a):
        return _fft.fft2(a, workers=os.cpu_count() or 1)

    def _ifft2(a):
        return _fft.ifft2(a, workers=os.cpu_count() or 1)
except ImportError:  # pragma: no cover
    _fft2, _ifft2 = np.fft.fft2, np.fft.ifft2


@dataclasses.dataclass
class Optics:
    width: int
    height: int
    num_planes: int
    start_plane: float
    plane_stepsize: float
    wavelength: float
    resolution: float

    def planes(self):
        z0 = np.float32(self.start_plane)
        dz = np.float32(self.plane_stepsize)
        return np.array([z0 + np.float32(np.float32(i) * dz) for i in range(self.num_planes)], np.float64)


def _draw_disc(mask, cx, cy, r):
    H, W = mask.shape
    rr = int(np.ceil(r)) + 1
    ys = np.arange(int(np.floor(cy)) - rr, int(np.floor(cy)) + rr + 2)
    xs = np.arange(int(np.floor(cx)) - rr, int(np.floor(cx)) + rr + 2)
    yy, xx = np.meshgrid(ys, xs, indexing="ij")
    inside = (xx + 0.5 - cx) ** 2 + (yy + 0.5 - cy) ** 2 <= r * r
    if not inside.any():  # a tracer always covers at least its centre pixel
        inside = (xx == int(np.floor(cx))) & (yy == int(np.floor(cy)))
    mask[yy[inside] % H, xx[inside] % W] = 1.0


def propagate_sum(masks, optics: Optics):
    """sum_i IFFT(FFT(-m_i) P_{z_i}) for the planes that hold objects."""
    H, W = optics.height, optics.width
    lam, dx = float(np.float32(optics.wavelength)), float(np.float32(optics.resolution))
    fx = np.fft.fftfreq(W)[None, :]
    fy = np.fft.fftfreq(H)[:, None]
    root = np.sqrt(np.maximum(1.0 - (fx * fx + fy * fy) * (lam / dx) ** 2, 0.0))
    acc = np.zeros((H, W), np.complex128)
    zs = optics.planes()
    for i in range(optics.num_planes):
        m = masks[i]
        if not m.any():
            continue
        P = np.exp(1j * 2 * np.pi * zs[i] * (root - 1.0) / lam)
        acc += _fft2(-m.astype(np.float64)) * P
    return _ifft2(acc)


def propagate_sum_torch(masks, optics: Optics):
    """propagate_sum evaluated with torch on the GPU in complex128 (the same sum; for large
    synthetic batches, where the host FFTs take seconds per hologram)."""
    import torch
    H, W = optics.height, optics.width
    lam, dx = float(np.float32(optics.wavelength)), float(np.float32(optics.resolution))
    fx = torch.fft.fftfreq(W, dtype=torch.float64, device="cuda")[None, :]
    fy = torch.fft.fftfreq(H, dtype=torch.float64, device="cuda")[:, None]
    root = torch.sqrt(torch.clamp(1.0 - (fx * fx + fy * fy) * (lam / dx) ** 2, min=0.0))
    occupied = np.flatnonzero(masks.reshape(masks.shape[0], -1).any(1))
    zs = torch.from_numpy(optics.planes()[occupied]).cuda()
    acc = torch.zeros((H, W), dtype=torch.complex128, device="cuda")
    for k0 in range(0, len(occupied), 16):
        idx = occupied[k0:k0 + 16]
        m = torch.from_numpy(np.ascontiguousarray(masks[idx])).cuda().to(torch.float64)
        ph = (2 * np.pi / lam) * zs[k0:k0 + 16, None, None] * (root - 1.0)[None]
        acc += (torch.fft.fft2(-m) * torch.polar(torch.ones_like(ph), ph)).sum(0)
    return torch.fft.ifft2(acc).cpu().numpy()


def hologram_from_masks(masks, optics: Optics, sigma: float, rng, propagate=None):
    U = 1.0 + (propagate or propagate_sum)(masks, optics)
    inten = np.abs(U) ** 2 + rng.normal(0.0, sigma, size=U.shape)
    b = inten / inten.mean() - 1.0
    return b.astype(np.float32).astype(np.complex64)


def disks_scene(optics: Optics, n_disks: int, rmin: float, rmax: float, seed: int, sigma: float):
    """C1-style scene: opaque disks at uniform (x, y, plane)."""
    rng = np.random.default_rng(seed)
    masks = np.zeros((optics.num_planes, optics.height, optics.width), np.float32)
    for _ in range(n_disks):
        r = rng.uniform(rmin, rmax)
        cx, cy = rng.uniform(0, optics.width), rng.uniform(0, optics.height)
        zi = rng.integers(0, optics.num_planes)
        _draw_disc(masks[zi], cx, cy, r)
    return hologram_from_masks(masks, optics, sigma, rng), masks


def tracers_rods_scene(optics: Optics, n_tracers: int, n_rods: int, seed: int, sigma: float,
                       rod_len=(20.0, 80.0), rod_width=3.0, tracer_diam=(1.0, 2.0), propagate=None):
    """C2-style scene: point tracers plus rods of random 3D orientation, the rods
    rasterised into per-plane occluding slices (discs of the rod width sampled
    every half pixel along the axis; z in plane units = pixels * dx / dz)."""
    rng = np.random.default_rng(seed)
    H, W, Nz = optics.height, optics.width, optics.num_planes
    masks = np.zeros((Nz, H, W), np.float32)
    for _ in range(n_tracers):
        d = rng.uniform(*tracer_diam)
        cx, cy = rng.uniform(0, W), rng.uniform(0, H)
        zi = rng.integers(0, Nz)
        _draw_disc(masks[zi], cx, cy, d / 2)
    z_per_px = optics.resolution / optics.plane_stepsize
    for _ in range(n_rods):
        length = rng.uniform(*rod_len)
        v = rng.normal(size=3)
        v /= np.linalg.norm(v)
        cx, cy, cz = rng.uniform(0, W), rng.uniform(0, H), rng.uniform(0, Nz)
        for s in np.arange(-length / 2, length / 2 + 1e-9, 0.5):
            x, y = cx + s * v[0], cy + s * v[1]
            zi = int(np.floor(cz + s * v[2] * z_per_px))
            if 0 <= zi < Nz:
                _draw_disc(masks[zi], x, y, rod_width / 2)
    return hologram_from_masks(masks, optics, sigma, rng, propagate), masks


class FlowField:
    """Seeded divergence-free in-plane velocity field from a stream function of
    random Fourier modes on the periodic W x H domain, scaled to an rms speed in
    pixels per frame."""

    def __init__(self, W, H, rng, n_modes=16, rms=3.0, kmax=3):
        self.W, self.H = W, H
        ks = []
        while len(ks) < n_modes:
            kx, ky = rng.integers(-kmax, kmax + 1, size=2)
            if kx or ky:
                ks.append((kx, ky))
        self.k = np.array(ks, np.float64) * (2 * np.pi / np.array([W, H], np.float64))
        self.a = rng.normal(size=n_modes)
        self.ph = rng.uniform(0, 2 * np.pi, size=n_modes)
        gy, gx = np.mgrid[0:H:8, 0:W:8]
        u, v = self._raw(gx.ravel().astype(np.float64), gy.ravel().astype(np.float64))
        self.scale = rms / max(np.sqrt(np.mean(u * u + v * v)), 1e-12)

    def _raw(self, x, y):
        arg = np.outer(x, self.k[:, 0]) + np.outer(y, self.k[:, 1]) + self.ph
        c = np.cos(arg) * self.a
        # psi = sum a sin(k.r + ph):  u = dpsi/dy, v = -dpsi/dx
        return c @ self.k[:, 1], -(c @ self.k[:, 0])

    def __call__(self, x, y):
        u, v = self._raw(np.asarray(x, np.float64), np.asarray(y, np.float64))
        return u * self.scale, v * self.scale


def sequence(optics: Optics, n_frames: int, n_tracers: int, n_rods: int, seed: int, sigma: float,
             rms=3.0, return_tracks=False, tracer_diam=(1.0, 2.0), rod_len=(20.0, 80.0), rod_width=3.0,
             propagate=None):
    """Time sequence of holograms: tracers and rods (fixed depth and orientation)
    advected between frames by a FlowField.  Returns holograms [n][H][W] complex64
    (and the tracer tracks [n][n_tracers][3] = x, y, plane with return_tracks)."""
    rng = np.random.default_rng(seed)
    H, W, Nz = optics.height, optics.width, optics.num_planes
    flow = FlowField(W, H, rng, rms=rms)
    tx, ty = rng.uniform(0, W, n_tracers), rng.uniform(0, H, n_tracers)
    tz = rng.integers(0, Nz, n_tracers)
    td = rng.uniform(*tracer_diam, n_tracers)
    rx, ry, rz = rng.uniform(0, W, n_rods), rng.uniform(0, H, n_rods), rng.uniform(0, Nz, n_rods)
    rl = rng.uniform(*rod_len, n_rods)
    rv = rng.normal(size=(n_rods, 3))
    rv /= np.linalg.norm(rv, axis=1, keepdims=True)
    z_per_px = optics.resolution / optics.plane_stepsize
    holos = np.zeros((n_frames, H, W), np.complex64)
    tracks = np.zeros((n_frames, n_tracers, 3))
    for f in range(n_frames):
        masks = np.zeros((Nz, H, W), np.float32)
        for i in range(n_tracers):
            _draw_disc(masks[tz[i]], tx[i], ty[i], td[i] / 2)
        for i in range(n_rods):
            for s_ in np.arange(-rl[i] / 2, rl[i] / 2 + 1e-9, 0.5):
                zi = int(np.floor(rz[i] + s_ * rv[i, 2] * z_per_px))
                if 0 <= zi < Nz:
                    _draw_disc(masks[zi], rx[i] + s_ * rv[i, 0], ry[i] + s_ * rv[i, 1], rod_width / 2)
        holos[f] = hologram_from_masks(masks, optics, sigma, rng, propagate)
        tracks[f] = np.stack([tx, ty, tz.astype(np.float64)], 1)
        u, v = flow(tx, ty)
        tx, ty = (tx + u) % W, (ty + v) % H
        u, v = flow(rx, ry)
        rx, ry = (rx + u) % W, (ry + v) % H
    return (holos, tracks) if return_tracks else holos


# Workloads (the parameter sets of the reference's CLI parameter file)
C1 = dict(width=256, height=256, num_planes=32, start_plane=1000.0, plane_stepsize=50.0,
          wavelength=0.633, resolution=2.0, num_inverse_iterations=50, regularization_param=0.01,
          regularization_TV_param=0.01, num_tv_iterations=5)
C2 = dict(width=1024, height=1024, num_planes=128, start_plane=2500.0, plane_stepsize=20.0,
          wavelength=0.632, resolution=1.1, num_inverse_iterations=100, regularization_param=0.15,
          regularization_TV_param=0.1, num_tv_iterations=5)
SEED_C1 = 20261201
SEED_C2 = 20261202


def optics_of(cfg: dict) -> Optics:
    return Optics(cfg["width"], cfg["height"], cfg["num_planes"], cfg["start_plane"],
                  cfg["plane_stepsize"], cfg["wavelength"], cfg["resolution"])


def c1_hologram(seed=SEED_C1, cfg=C1):
    return disks_scene(optics_of(cfg), 20, 2.0, 6.0, seed, 0.01)[0]


def c2_hologram(seed=SEED_C2, cfg=C2):
    return tracers_rods_scene(optics_of(cfg), 500, 20, seed, 0.02)[0]
