"""ctypes binding of the CPU oracle (oracle/rihvr_oracle.cpp).

TEST INFRASTRUCTURE ONLY.  Imported by tests/, by ``__graft_entry__.smoke()``
and by the ``cpu_baseline`` leg of bench.py, always as the checker or as the
CPU baseline -- never as the product path.
"""
from __future__ import annotations

import ctypes
import os
import pathlib

import numpy as np

_HERE = pathlib.Path(__file__).resolve().parent
LIB_PATH = _HERE / "_build" / "librihvr_oracle.so"


class Params(ctypes.Structure):
    """Mirror of ``rihvr_params`` (include/rihvr.h)."""

    _fields_ = [
        ("width", ctypes.c_int32),
        ("height", ctypes.c_int32),
        ("num_planes", ctypes.c_int32),
        ("start_plane", ctypes.c_float),
        ("plane_stepsize", ctypes.c_float),
        ("wavelength", ctypes.c_float),
        ("resolution", ctypes.c_float),
        ("zero_padding", ctypes.c_int32),
        ("num_inverse_iterations", ctypes.c_int32),
        ("regularization_param", ctypes.c_float),
        ("regularization_TV_param", ctypes.c_float),
        ("num_tv_iterations", ctypes.c_int32),
        ("mode", ctypes.c_int32),
    ]


_P = ctypes.POINTER


class Stats(ctypes.Structure):
    """Mirror of ``rihvr_stats`` (include/rihvr.h)."""

    _fields_ = [
        ("iterations", ctypes.c_int32),
        ("total_backtracks", ctypes.c_int32),
        ("stepsize", ctypes.c_double),
        ("f_initial", ctypes.c_double),
        ("f_final", ctypes.c_double),
        ("sparsity", ctypes.c_double),
        ("nnz", ctypes.c_int64),
        ("f", _P(ctypes.c_double)),
        ("f_resid", _P(ctypes.c_double)),
        ("ratio", _P(ctypes.c_double)),
        ("lim", _P(ctypes.c_double)),
        ("backtracks", _P(ctypes.c_int32)),
        ("restarts", _P(ctypes.c_int32)),
        ("restart_sum", _P(ctypes.c_double)),
        ("sparsity_it", _P(ctypes.c_double)),
    ]


def make_params(width, height, num_planes, start_plane, plane_stepsize, wavelength,
                resolution, num_inverse_iterations, regularization_param,
                regularization_TV_param, num_tv_iterations, mode=7, zero_padding=0) -> Params:
    return Params(int(width), int(height), int(num_planes), float(start_plane),
                  float(plane_stepsize), float(wavelength), float(resolution),
                  int(zero_padding), int(num_inverse_iterations), float(regularization_param),
                  float(regularization_TV_param), int(num_tv_iterations), int(mode))


class StatsBuffers:
    """Owns the per-iteration arrays a ``Stats`` points to."""

    def __init__(self, n_iter: int):
        n = max(1, int(n_iter))
        self.f = np.zeros(n, np.float64)
        self.f_resid = np.zeros(n, np.float64)
        self.ratio = np.zeros(n, np.float64)
        self.lim = np.zeros(n, np.float64)
        self.backtracks = np.zeros(n, np.int32)
        self.restarts = np.zeros(n, np.int32)
        self.restart_sum = np.zeros(n, np.float64)
        self.sparsity_it = np.zeros(n, np.float64)
        d = lambda a: a.ctypes.data_as(_P(ctypes.c_double))  # noqa: E731
        i = lambda a: a.ctypes.data_as(_P(ctypes.c_int32))  # noqa: E731
        self.stats = Stats(0, 0, 0.0, 0.0, 0.0, 0.0, 0, d(self.f), d(self.f_resid), d(self.ratio),
                           d(self.lim), i(self.backtracks), i(self.restarts), d(self.restart_sum),
                           d(self.sparsity_it))

    def as_dict(self) -> dict:
        s = self.stats
        n = s.iterations
        return dict(iterations=n, total_backtracks=s.total_backtracks, stepsize=s.stepsize,
                    f_initial=s.f_initial, f_final=s.f_final, sparsity=s.sparsity, nnz=s.nnz,
                    f=self.f[:n].copy(), f_resid=self.f_resid[:n].copy(),
                    ratio=self.ratio[:n].copy(), lim=self.lim[:n].copy(),
                    backtracks=self.backtracks[:n].copy(), restarts=self.restarts[:n].copy(),
                    restart_sum=self.restart_sum[:n].copy())


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            raise RuntimeError(f"oracle library missing: {LIB_PATH} (run make -C oracle)")
        L = ctypes.CDLL(str(LIB_PATH))
        vp = ctypes.c_void_p
        pp = _P(Params)
        L.ro_version.restype = ctypes.c_int
        L.ro_set_threads.argtypes = [ctypes.c_int]
        L.ro_get_threads.restype = ctypes.c_int
        for nm in ("ro_fft2_f32", "ro_fft2_f64"):
            getattr(L, nm).argtypes = [vp, ctypes.c_int, ctypes.c_int, ctypes.c_int]
            getattr(L, nm).restype = ctypes.c_int
        L.ro_kappa.argtypes = [pp, vp]
        L.ro_plane_list.argtypes = [pp, vp]
        for nm in ("ro_forward_f32", "ro_forward_f64", "ro_adjoint_f32", "ro_adjoint_f64"):
            getattr(L, nm).argtypes = [pp, vp, vp]
            getattr(L, nm).restype = ctypes.c_int
        for nm in ("ro_prox_f32", "ro_prox_f64"):
            getattr(L, nm).argtypes = [pp, vp, vp, ctypes.c_double, vp]
            getattr(L, nm).restype = ctypes.c_int
        for nm in ("ro_solve_f32", "ro_solve_f64"):
            getattr(L, nm).argtypes = [pp, vp, vp, vp, vp, vp, _P(Stats)]
            getattr(L, nm).restype = ctypes.c_int
        L.ro_summag_f32.argtypes = [vp, ctypes.c_long]
        L.ro_summag_f32.restype = ctypes.c_double
        L.ro_l1_f32.argtypes = [vp, ctypes.c_long]
        L.ro_l1_f32.restype = ctypes.c_double
        L.ro_tv2d_f32.argtypes = [vp, ctypes.c_int, ctypes.c_int, ctypes.c_int]
        L.ro_tv2d_f32.restype = ctypes.c_double
        L.ro_tv3d_f32.argtypes = [vp, ctypes.c_int, ctypes.c_int, ctypes.c_int]
        L.ro_tv3d_f32.restype = ctypes.c_double
        L.ro_restart_sum_f32.argtypes = [vp, vp, vp, ctypes.c_long, ctypes.c_int]
        L.ro_restart_sum_f32.restype = ctypes.c_double
        _lib = L
    return _lib


def set_threads(n: int) -> None:
    lib().ro_set_threads(int(n))


def get_threads() -> int:
    return int(lib().ro_get_threads())


def _c(a, dtype):
    """Contiguous complex array of the given complex dtype."""
    return np.ascontiguousarray(a, dtype=dtype)


def _ptr(a):
    return ctypes.c_void_p(a.ctypes.data) if a is not None else None


def _suffix(precision):
    if precision not in ("f32", "f64"):
        raise ValueError(precision)
    return precision, (np.complex64 if precision == "f32" else np.complex128)


def _check(rc, what):
    if rc != 0:
        raise RuntimeError(f"{what} failed with status {rc}")


def fft2(a, sign=-1, precision="f32"):
    sfx, ct = _suffix(precision)
    a = _c(a, ct).copy()
    H, W = a.shape
    _check(getattr(lib(), f"ro_fft2_{sfx}")(_ptr(a), H, W, int(sign)), "fft2")
    return a


def kappa(p: Params):
    out = np.zeros((p.height, p.width), np.float64)
    _check(lib().ro_kappa(ctypes.byref(p), _ptr(out)), "kappa")
    return out


def plane_list(p: Params):
    out = np.zeros(p.num_planes, np.float32)
    _check(lib().ro_plane_list(ctypes.byref(p), _ptr(out)), "plane_list")
    return out


def forward(p: Params, vol, precision="f32"):
    sfx, ct = _suffix(precision)
    vol = _c(vol, ct)
    out = np.zeros((p.height, p.width), ct)
    _check(getattr(lib(), f"ro_forward_{sfx}")(ctypes.byref(p), _ptr(vol), _ptr(out)), "forward")
    return out


def adjoint(p: Params, holo, precision="f32"):
    sfx, ct = _suffix(precision)
    holo = _c(holo, ct)
    out = np.zeros((p.num_planes, p.height, p.width), ct)
    _check(getattr(lib(), f"ro_adjoint_{sfx}")(ctypes.byref(p), _ptr(holo), _ptr(out)), "adjoint")
    return out


def prox(p: Params, x0, grad, t, precision="f32"):
    sfx, ct = _suffix(precision)
    x0 = _c(x0, ct)
    grad = _c(grad, ct)
    out = np.zeros_like(x0)
    _check(getattr(lib(), f"ro_prox_{sfx}")(ctypes.byref(p), _ptr(x0), _ptr(grad), float(t),
                                            _ptr(out)), "prox")
    return out


def solve(p: Params, holo, warm=None, precision="f32"):
    """Returns (volume, estimate, residual, stats dict)."""
    sfx, ct = _suffix(precision)
    holo = _c(holo, ct)
    warm = _c(warm, ct) if warm is not None else None
    out = np.zeros((p.num_planes, p.height, p.width), ct)
    est = np.zeros((p.height, p.width), ct)
    res = np.zeros((p.height, p.width), ct)
    sb = StatsBuffers(p.num_inverse_iterations)
    rc = getattr(lib(), f"ro_solve_{sfx}")(ctypes.byref(p), _ptr(holo), _ptr(warm), _ptr(out),
                                           _ptr(est), _ptr(res), ctypes.byref(sb.stats))
    _check(rc, "solve")
    return out, est, res, sb.as_dict()


def summag_f32(a):
    a = _c(a, np.complex64)
    return float(lib().ro_summag_f32(_ptr(a), a.size))


def l1_f32(a):
    a = _c(a, np.complex64)
    return float(lib().ro_l1_f32(_ptr(a), a.size))


def tv2d_f32(vol):
    vol = _c(vol, np.complex64)
    Nz, H, W = vol.shape
    return float(lib().ro_tv2d_f32(_ptr(vol), Nz, H, W))


def tv3d_f32(vol):
    vol = _c(vol, np.complex64)
    Nz, H, W = vol.shape
    return float(lib().ro_tv3d_f32(_ptr(vol), Nz, H, W))


def restart_sum_f32(x0, x1, y0):
    x0, x1, y0 = (_c(a, np.complex64) for a in (x0, x1, y0))
    Nz = x0.shape[0]
    return float(lib().ro_restart_sum_f32(_ptr(x0), _ptr(x1), _ptr(y0), x0[0].size, Nz))


def cores() -> int:
    return os.cpu_count() or 1
