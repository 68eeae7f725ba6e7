"""Independent numpy restatement of the reference's sparse inverse path.

Written separately from the C++ oracle (different FFT: numpy's pocketfft; array
code instead of loops) so that the two pin each other.  Used by
make_golden.py to write the committed fixtures and by the CPU tests.

``round_args=True`` rounds every phase argument to float32 before sin/cos, the
reference's sincosf(float(double)) semantics (sparse_compressive_holo.cu:1849-1855,
1872-1879; sparse_gradient.cu:85-92, 111-117); arithmetic stays float64.
``round_args=False`` is the FP64 operator.
"""
from __future__ import annotations

import numpy as np


def plane_list(z0, dz, nz):
    """z_i = z0 + i*dz evaluated in float32 (src/core/holo_ui.cpp:478-483)."""
    z0 = np.float32(z0)
    dz = np.float32(dz)
    return np.array([z0 + np.float32(np.float32(i) * dz) for i in range(nz)], np.float32)


def kappa(W, H, wavelength, resolution):
    """-2*pi*sqrt(max(0, 1 - f^2 (lambda/dx)^2))/lambda on numpy's fftfreq grid
    (sparse_compressive_holo.cu:38-56)."""
    lam = float(np.float32(wavelength))
    dx = float(np.float32(resolution))
    fx = np.fft.fftfreq(W)[None, :]
    fy = np.fft.fftfreq(H)[:, None]
    s = 1.0 - (fx * fx + fy * fy) * (lam / dx) ** 2
    return -2.0 * np.pi * np.sqrt(np.maximum(s, 0.0)) / lam


def _phase(arg, round_args):
    a = np.asarray(arg, np.float64)
    if round_args:
        a = a.astype(np.float32).astype(np.float64)
    return np.cos(a) + 1j * np.sin(a)


class Operator:
    def __init__(self, W, H, nz, z0, dz, wavelength, resolution, round_args=False):
        self.W, self.H, self.nz = W, H, nz
        self.z = plane_list(z0, dz, nz).astype(np.float64)
        self.lam = float(np.float32(wavelength))
        self.kap = kappa(W, H, wavelength, resolution)
        self.round_args = round_args
        self.N = W * H

    def forward(self, vol):
        """A (SparseCompressiveHolo::backPropagate, :1935-2077)."""
        S = np.zeros((self.H, self.W), np.complex128)
        for i in range(self.nz):
            z = -self.z[i]
            e = _phase(2 * np.pi * z / self.lam, self.round_args)
            Hm = _phase(z * self.kap, self.round_args)
            S += Hm * np.fft.fft2(vol[i] * e) / self.N
        est = np.fft.ifft2(S) * self.N
        return est.real + 0j

    def adjoint(self, holo):
        """A^T per plane (SparseGradient::reconstruct/getPlane, sparse_gradient.cu:47-165)."""
        R = np.fft.fft2(holo)
        out = np.empty((self.nz, self.H, self.W), np.complex128)
        for i in range(self.nz):
            z = self.z[i]
            Hp = _phase(z * self.kap, self.round_args)
            g = np.fft.ifft2(Hp * R / self.N) * self.N
            out[i] = g * _phase(2 * np.pi * z / self.lam, self.round_args)
        return out


def soft(v, thr):
    """block_soft_threshold_kernel per component (:3065-3110)."""
    def one(a):
        y = np.maximum(np.abs(a) - thr, 0.0)
        return (y / (y + thr)) * a
    return one(v.real) + 1j * one(v.imag)


def f32(v):
    """Parameters are float fields promoted to double (Q3: src/core/holo_ui.h:97-121)."""
    return float(np.float32(v))


def prox_fl2d(x0, g, t, reg, reg_tv, n_tv):
    """proximalFusedLasso2d (:3652-3732), vectorised over planes (axis 0)."""
    reg, reg_tv = f32(reg), f32(reg_tv)
    xh = x0 - t * g
    gx = np.zeros_like(xh)
    gy = np.zeros_like(xh)
    mu = t * reg_tv
    sc = 1.0 / (12.0 * t * reg_tv)

    def estimate():
        # evaluation order of tv_denoised_estimate_2d_kernel: ((gx+ - gx) + gy+) - gy
        div = ((np.roll(gx, -1, axis=2) - gx) + np.roll(gy, -1, axis=1)) - gy
        return xh - mu * div

    for _ in range(n_tv):
        y = estimate()
        ax = gx + sc * (np.roll(y, 1, axis=2) - y)
        ay = gy + sc * (np.roll(y, 1, axis=1) - y)
        nrm = np.sqrt(((ax.real * ax.real + ax.imag * ax.imag) + ay.real * ay.real) + ay.imag * ay.imag)
        nrm = np.maximum(nrm, 1.0)
        gx = ax / nrm
        gy = ay / nrm
    y = estimate()
    if reg > 0:
        y = soft(y, t * reg)
    return y


def prox_fl3d(x0, g, t, reg, reg_tv, n_tv):
    """proximalFusedLasso, mode 6 (:3422-3475, :3867-4047): soft threshold at half the
    weight, 3D dual iterations with periodic x, y and z and the normaliser forced to 1
    (:3991), soft threshold at half the weight again.  Volume axes (z, y, x)."""
    reg, reg_tv = f32(reg), f32(reg_tv)
    thr = t * (reg / 2.0)
    x1 = soft(x0 - t * g, thr)
    gx = np.zeros_like(x1)
    gy = np.zeros_like(x1)
    gz = np.zeros_like(x1)
    mu = t * reg_tv
    sc = 1.0 / (12.0 * t * reg_tv)

    def estimate():
        # tv_denoised_estimate_kernel evaluation order: ((((gx+ - gx) + gy+) - gy) + gz_next) - gz
        div = ((((np.roll(gx, -1, axis=2) - gx) + np.roll(gy, -1, axis=1)) - gy) + np.roll(gz, -1, axis=0)) - gz
        return x1 - mu * div

    for _ in range(n_tv):
        y = estimate()
        gx = gx + sc * (np.roll(y, 1, axis=2) - y)
        gy = gy + sc * (np.roll(y, 1, axis=1) - y)
        gz = gz + sc * (np.roll(y, 1, axis=0) - y)  # previous plane, plane Nz-1 before plane 0
    return soft(estimate(), thr)


def tv3d(vol):
    """calcTVNorm (sparse_volume.cu:236-363): z-periodic backward differences."""
    a = np.roll(vol, 1, axis=2) - vol
    b = np.roll(vol, 1, axis=1) - vol
    c = np.roll(vol, 1, axis=0) - vol
    return float(np.sum(np.sqrt(np.abs(a) ** 2 + np.abs(b) ** 2 + np.abs(c) ** 2)))


def prox_l1(x0, g, t, reg):
    """proximalL1 (:3274-3392)."""
    return soft(x0 - t * g, t * f32(reg))


def tv2d(vol):
    """calcTVNorm2d (sparse_volume.cu:365-446)."""
    a = np.roll(vol, 1, axis=2) - vol
    b = np.roll(vol, 1, axis=1) - vol
    return float(np.sum(np.sqrt(np.abs(a) ** 2 + np.abs(b) ** 2)))


def restart_sum(x0, x1, y0):
    """sparse_fistaRestart_kernel weights (:4393-4432): within each 1024-chunk of a
    plane, offsets <512 add d_re + 2 d_im, offsets >=512 add d_im."""
    dre = ((x0.real - x1.real) * (x1.real - y0.real)).reshape(x0.shape[0], -1)
    dim = ((x0.imag - x1.imag) * (x1.imag - y0.imag)).reshape(x0.shape[0], -1)
    off = np.arange(dre.shape[1]) % 1024
    lo = off < 512
    return float(np.sum(np.where(lo, dre + 2 * dim, dim)))


def solve(op: Operator, holo, iters, reg, reg_tv, n_tv, mode=7, warm=None):
    """inverseReconstruct (:173-633) in float64.  Returns (x1, est, r, log)."""
    if reg_tv == 0.0:
        mode = 5
    nz = op.nz
    t = 1.0 / nz
    x1 = np.zeros((nz, op.H, op.W), np.complex128) if warm is None else warm.astype(np.complex128)
    x0 = x1.copy()
    y0 = x1.copy()
    alpha1 = 1.0

    def prox(x0_, G_, t_):
        if mode == 7:
            return prox_fl2d(x0_, G_, t_, reg, reg_tv, n_tv)
        if mode == 6:
            return prox_fl3d(x0_, G_, t_, reg, reg_tv, n_tv)
        return prox_l1(x0_, G_, t_, reg)

    def objective(r, x):
        f = 0.5 * np.sum(np.abs(r) ** 2) + f32(reg) * np.sum(np.abs(x))
        if mode == 7:
            f += f32(reg_tv) * tv2d(x)
        elif mode == 6:
            f += f32(reg_tv) * tv3d(x)
        return float(f)

    est = op.forward(x1)
    r = est - holo
    f = objective(r, x1)
    prev_f_resid = 0.5 * float(np.sum(np.abs(r) ** 2))
    G = op.adjoint(r)
    cur = np.zeros_like(est)
    log = dict(f=[], f_initial=f, ratio=[], lim=[], backtracks=[], restarts=[], restart_sum=[],
               f_resid=[])
    for _ in range(iters):
        x0 = x1.copy()
        alpha0 = alpha1
        x1 = prox(x0, G, t)
        est = op.forward(x1)
        r = est - holo
        d = x1 - x0
        lim = float(np.sum(d.real * G.real + d.imag * G.imag)) + float(np.sum(np.abs(d) ** 2)) / (2 * t)
        f_resid = 0.5 * float(np.sum(np.abs(r) ** 2))
        ratio = f_resid / (prev_f_resid + lim)
        bt = 0
        while (ratio > 1.0 or ratio < 0) and bt < 2:
            t *= 0.5
            x1 = prox(x0, G, t)
            est = op.forward(x1)
            r = est - holo
            d = x1 - x0
            lim = float(np.sum(d.real * G.real + d.imag * G.imag)) + float(np.sum(np.abs(d) ** 2)) / (2 * t)
            f_resid = 0.5 * float(np.sum(np.abs(r) ** 2))
            ratio = f_resid / (prev_f_resid + lim)
            bt += 1
        rs = restart_sum(x0, x1, y0)
        if rs > 0:
            alpha0 = 1.0
        alpha1 = (1.0 + np.sqrt(1.0 + 4.0 * alpha0 * alpha0)) / 2.0
        s = (alpha0 - 1.0) / alpha1
        y1 = x1
        x1 = y1 + s * (y1 - y0)
        y0 = y1
        prev = cur
        cur = est
        est = cur + s * (cur - prev)
        r = est - holo
        G = op.adjoint(r)
        f = objective(r, x1)
        prev_f_resid = 0.5 * float(np.sum(np.abs(r) ** 2))
        log["f"].append(f)
        log["f_resid"].append(prev_f_resid)
        log["ratio"].append(ratio)
        log["lim"].append(lim)
        log["backtracks"].append(bt)
        log["restarts"].append(int(rs > 0))
        log["restart_sum"].append(rs)
    log["stepsize"] = t
    return x1, est, r, log
