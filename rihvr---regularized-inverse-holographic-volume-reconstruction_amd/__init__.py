"""MI355X-native sparse regularized-inverse holographic volume reconstruction.

Python binding of librihvr.so (C ABI in include/rihvr.h).  The class
:class:`SparseCompressiveHolo` mirrors the reference's in-process API
(src/core/sparse_compressive_holo.h:25-190): construct with the parameter set,
``inverseReconstruct(holo, mode)``, ``getPlane(z)``, ``setStepsize(step)``,
``destroy()``.  Arrays are complex64 numpy arrays (host) or torch tensors on
the GPU (device pointers are passed through without copies).

The library has no CPU fallback: if librihvr.so is missing or fails to load,
every call raises.
"""
from __future__ import annotations

import ctypes
import pathlib

import numpy as np

PKG_DIR = pathlib.Path(__file__).resolve().parent
LIB_PATH = PKG_DIR / "_build" / "librihvr.so"

HOST, DEVICE = 0, 1
MODE_FASTA_L1 = 5
MODE_FASTA_FUSED_LASSO = 6
MODE_FASTA_FUSED_LASSO_2D = 7

STATUS_NAMES = {
    0: "RIHVR_OK", 1: "RIHVR_E_UNKNOWN_TYPE", 2: "RIHVR_E_BAD_FILENAME", 3: "RIHVR_E_MISSING_DATA",
    4: "RIHVR_E_INVALID_STATE", 5: "RIHVR_E_OUT_OF_MEMORY", 6: "RIHVR_E_UNKNOWN_MODE",
    7: "RIHVR_E_INVALID_ARGUMENT", 8: "RIHVR_E_UNKNOWN_ERROR", 9: "RIHVR_E_INVALID_FILE",
    10: "RIHVR_E_INVALID_DATA", 11: "RIHVR_E_CRITICAL_ASSUMPTION", 12: "RIHVR_E_INFINITE_LOOP",
    13: "RIHVR_E_DEVICE",
}

# Every function include/rihvr.h declares.
EXPORTED = ("rihvr_abi_version", "rihvr_status_string", "rihvr_create", "rihvr_destroy",
            "rihvr_reconstruct", "rihvr_begin", "rihvr_iterate", "rihvr_finish", "rihvr_set_timing",
            "rihvr_get_timing", "rihvr_get_volume", "rihvr_get_plane", "rihvr_set_stepsize",
            "rihvr_get_estimate", "rihvr_forward", "rihvr_adjoint", "rihvr_prox", "rihvr_get_params",
            "rihvr_snapshot", "rihvr_read_param_file", "rihvr_read_image", "rihvr_prepare_hologram",
            "rihvr_write_tiff8", "rihvr_write_hologram", "rihvr_save_sparse", "rihvr_save_projections",
            "rihvr_save_planes", "rihvr_reconstruct_batch", "rihvr_write_png8", "rihvr_background_mean",
            "rihvr_background_median5", "rihvr_read_sparse_voxels", "rihvr_segmentation_create",
            "rihvr_segmentation_destroy", "rihvr_segmentation_load", "rihvr_segmentation_load_dense",
            "rihvr_segmentation_binarize", "rihvr_segmentation_close", "rihvr_segmentation_label",
            "rihvr_segmentation_extract", "rihvr_segmentation_size_filter", "rihvr_segmentation_objects",
            "rihvr_segmentation_write_centroids", "rihvr_segmentation_mask", "rihvr_segmentation_values",
            "rihvr_segmentation_labels", "rihvr_segmentation_timing", "rihvr_prepare_holograms",
            "rihvr_prepare_holograms_noalloc", "rihvr_net_param_count", "rihvr_net_create", "rihvr_net_destroy",
            "rihvr_net_init", "rihvr_net_set_params", "rihvr_net_get_params", "rihvr_net_predict",
            "rihvr_net_loss_grad", "rihvr_net_adam_step", "rihvr_net_train_step", "rihvr_net_sync",
            "rihvr_conv3d_bf16_forward", "rihvr_conv3d_bf16_dgrad", "rihvr_conv3d_bf16_wgrad",
            "rihvr_reconstruct_learned", "rihvr_net_set_loss_weight", "rihvr_background_correlation",
            "rihvr_create_scheduled", "rihvr_get_schedule", "rihvr_autotune", "rihvr_reconstruct_multi",
            "rihvr_estimate_lipschitz")
PATH_MAX = 4096


class RihvrError(RuntimeError):
    def __init__(self, status: int, what: str):
        super().__init__(f"{what}: {STATUS_NAMES.get(status, status)}")
        self.status = status


class Params(ctypes.Structure):
    """``rihvr_params`` (include/rihvr.h)."""

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
    """``rihvr_stats`` (include/rihvr.h)."""

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


class KernelTime(ctypes.Structure):
    """``rihvr_kernel_time`` (include/rihvr.h)."""

    _fields_ = [("name", ctypes.c_char * 32), ("launches", ctypes.c_int64), ("total_ms", ctypes.c_double)]


class FrameStats(ctypes.Structure):
    """Mirror of ``rihvr_frame_stats`` (include/rihvr.h)."""

    _fields_ = [
        ("iterations", ctypes.c_int32),
        ("total_backtracks", ctypes.c_int32),
        ("reached", ctypes.c_int32),
        ("pad_", ctypes.c_int32),
        ("f_initial", ctypes.c_double),
        ("f_final", ctypes.c_double),
        ("nnz", ctypes.c_int64),
    ]


class Schedule(ctypes.Structure):
    """Mirror of ``rihvr_schedule`` (include/rihvr.h): launch choices of a context."""

    _fields_ = [
        ("fft_path", ctypes.c_int32),
        ("max_radix", ctypes.c_int32),
        ("zchunk_fwd", ctypes.c_int32),
        ("zchunk_adj", ctypes.c_int32),
        ("prox_rows", ctypes.c_int32),
        ("graphs", ctypes.c_int32),
        ("ms_per_iteration", ctypes.c_double),
    ]

    @classmethod
    def make(cls, fft_path=0, max_radix=16, zchunk_fwd=0, zchunk_adj=0, prox_rows=0, graphs=1):
        return cls(int(fft_path), int(max_radix), int(zchunk_fwd), int(zchunk_adj), int(prox_rows), int(graphs), 0.0)

    def as_dict(self) -> dict:
        return {k: getattr(self, k) for k, _ in self._fields_}


class BackgroundStats(ctypes.Structure):
    """Mirror of ``rihvr_background_stats`` (include/rihvr.h)."""

    _fields_ = [("alpha", ctypes.c_double), ("beta", ctypes.c_double), ("minimum", ctypes.c_double),
                ("average", ctypes.c_double)]


class FileParams(ctypes.Structure):
    """Mirror of ``rihvr_file_params`` (include/rihvr.h): the parameter file."""

    _fields_ = [
        ("solver", Params),
        ("image_count", ctypes.c_int32),
        ("start_image", ctypes.c_int32),
        ("background_start", ctypes.c_int32),
        ("background_end", ctypes.c_int32),
        ("image_filename", ctypes.c_char * PATH_MAX),
        ("image_path", ctypes.c_char * PATH_MAX),
        ("image_formatted_name", ctypes.c_char * PATH_MAX),
        ("output_path", ctypes.c_char * PATH_MAX),
        ("background_filename", ctypes.c_char * PATH_MAX),
        ("remove_background", ctypes.c_int32),
        ("center_x", ctypes.c_float),
        ("center_y", ctypes.c_float),
        ("center_z", ctypes.c_float),
        ("roi_size", ctypes.c_int32),
        ("roi_x", ctypes.c_int32),
        ("roi_y", ctypes.c_int32),
        ("roi_w", ctypes.c_int32),
        ("roi_h", ctypes.c_int32),
        ("output_planes", ctypes.c_int32),
        ("output_steps", ctypes.c_int32),
        ("use_median_background", ctypes.c_int32),
        ("segment_min_area", ctypes.c_int32),
        ("segment_min_intensity", ctypes.c_double),
        ("segment_close_size", ctypes.c_int32),
        ("unknown_id", ctypes.c_int32),
    ]


class Object(ctypes.Structure):
    """Mirror of ``rihvr_object`` (include/rihvr.h): one segmented object."""

    _fields_ = [
        ("centroid", ctypes.c_float * 3),
        ("weighted_centroid", ctypes.c_float * 3),
        ("max_value", ctypes.c_float),
        ("mean_value", ctypes.c_float),
        ("num_voxels", ctypes.c_int64),
        ("label", ctypes.c_int64),
        ("bounds", ctypes.c_int32 * 6),
    ]

    def as_dict(self) -> dict:
        return dict(centroid=list(self.centroid), weighted_centroid=list(self.weighted_centroid),
                    max_value=self.max_value, mean_value=self.mean_value, num_voxels=self.num_voxels,
                    label=self.label, bounds=list(self.bounds))


class LearnedStats(ctypes.Structure):
    """Mirror of ``rihvr_learned_stats`` (include/rihvr.h)."""

    _fields_ = [
        ("iterations", ctypes.c_int32),
        ("fallback", ctypes.c_int32),
        ("reached", ctypes.c_int32),
        ("total_backtracks", ctypes.c_int32),
        ("resid_net", ctypes.c_double),
        ("resid_adjoint", ctypes.c_double),
        ("f_net", ctypes.c_double),
        ("f_adjoint", ctypes.c_double),
        ("f_zero", ctypes.c_double),
        ("f_start", ctypes.c_double),
        ("f_final", ctypes.c_double),
        ("ms_adjoint", ctypes.c_double),
        ("ms_network", ctypes.c_double),
        ("ms_solve", ctypes.c_double),
        ("tau", ctypes.c_double),
        ("f_thresholded", ctypes.c_double),
    ]

    def as_dict(self) -> dict:
        return {k: getattr(self, k) for k, _ in self._fields_}


class MultiItem(ctypes.Structure):
    """Mirror of ``rihvr_multi_item`` (include/rihvr.h)."""

    _fields_ = [
        ("params", Params),
        ("holo", ctypes.c_void_p),
        ("volume", ctypes.c_void_p),
        ("f_target", ctypes.c_double),
        ("stats", FrameStats),
        ("schedule", Schedule),
    ]


def make_params(width, height, num_planes, start_plane, plane_stepsize, wavelength, resolution,
                num_inverse_iterations, regularization_param, regularization_TV_param,
                num_tv_iterations, mode=MODE_FASTA_FUSED_LASSO_2D, zero_padding=0) -> Params:
    return Params(int(width), int(height), int(num_planes), float(start_plane), float(plane_stepsize),
                  float(wavelength), float(resolution), int(zero_padding), int(num_inverse_iterations),
                  float(regularization_param), float(regularization_TV_param), int(num_tv_iterations),
                  int(mode))


class StatsBuffers:
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
                    f=self.f[:n].copy(), f_resid=self.f_resid[:n].copy(), ratio=self.ratio[:n].copy(),
                    lim=self.lim[:n].copy(), backtracks=self.backtracks[:n].copy(),
                    restarts=self.restarts[:n].copy(), restart_sum=self.restart_sum[:n].copy(),
                    sparsity_it=self.sparsity_it[:n].copy())


_lib = None


def lib():
    """Loads librihvr.so (raises if it is missing: there is no fallback)."""
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            raise RuntimeError(f"librihvr.so is not built ({LIB_PATH}); run __graft_entry__.build()")
        L = ctypes.CDLL(str(LIB_PATH))
        vp, ip = ctypes.c_void_p, ctypes.c_int
        st = ctypes.c_int
        L.rihvr_abi_version.restype = ctypes.c_int
        L.rihvr_status_string.argtypes = [ctypes.c_int]
        L.rihvr_status_string.restype = ctypes.c_char_p
        L.rihvr_create.argtypes = [_P(Params), _P(vp)]
        L.rihvr_create.restype = st
        L.rihvr_destroy.argtypes = [vp]
        L.rihvr_destroy.restype = st
        L.rihvr_reconstruct.argtypes = [vp, vp, ip, vp, ip, _P(Stats)]
        L.rihvr_reconstruct.restype = st
        L.rihvr_begin.argtypes = [vp, vp, ip, vp, ip, _P(ctypes.c_double)]
        L.rihvr_begin.restype = st
        L.rihvr_iterate.argtypes = [vp, ctypes.c_int32, _P(Stats)]
        L.rihvr_iterate.restype = st
        L.rihvr_finish.argtypes = [vp, _P(Stats)]
        L.rihvr_finish.restype = st
        L.rihvr_set_timing.argtypes = [vp, ctypes.c_int32]
        L.rihvr_set_timing.restype = st
        L.rihvr_get_timing.argtypes = [vp, _P(KernelTime), ctypes.c_int32, _P(ctypes.c_int32)]
        L.rihvr_get_timing.restype = st
        L.rihvr_get_volume.argtypes = [vp, vp, ip]
        L.rihvr_get_volume.restype = st
        L.rihvr_get_plane.argtypes = [vp, ctypes.c_int32, vp, ip]
        L.rihvr_get_plane.restype = st
        L.rihvr_set_stepsize.argtypes = [vp, ctypes.c_double]
        L.rihvr_set_stepsize.restype = st
        L.rihvr_get_estimate.argtypes = [vp, vp, vp, ip]
        L.rihvr_get_estimate.restype = st
        L.rihvr_forward.argtypes = [vp, vp, ip, vp, ip]
        L.rihvr_forward.restype = st
        L.rihvr_adjoint.argtypes = [vp, vp, ip, vp, ip]
        L.rihvr_adjoint.restype = st
        L.rihvr_prox.argtypes = [vp, vp, vp, ctypes.c_double, vp, ip]
        L.rihvr_prox.restype = st
        L.rihvr_get_params.argtypes = [vp, _P(Params)]
        L.rihvr_get_params.restype = st
        L.rihvr_snapshot.argtypes = [vp]
        L.rihvr_snapshot.restype = st
        cp, i32p = ctypes.c_char_p, _P(ctypes.c_int32)
        L.rihvr_read_param_file.argtypes = [cp, _P(FileParams)]
        L.rihvr_read_param_file.restype = st
        L.rihvr_read_image.argtypes = [cp, vp, i32p, i32p]
        L.rihvr_read_image.restype = st
        L.rihvr_prepare_hologram.argtypes = [_P(FileParams), vp, vp, ctypes.c_int32, ctypes.c_int32, vp, i32p,
                                             i32p]
        L.rihvr_prepare_hologram.restype = st
        L.rihvr_write_tiff8.argtypes = [cp, vp, ctypes.c_int32, ctypes.c_int32]
        L.rihvr_write_tiff8.restype = st
        L.rihvr_write_hologram.argtypes = [cp, vp, ctypes.c_int32, ctypes.c_int32, ctypes.c_int32,
                                           ctypes.c_double, ctypes.c_double]
        L.rihvr_write_hologram.restype = st
        L.rihvr_save_sparse.argtypes = [vp, cp]
        L.rihvr_save_sparse.restype = st
        L.rihvr_save_projections.argtypes = [vp, cp, cp, ctypes.c_int32]
        L.rihvr_save_projections.restype = st
        L.rihvr_save_planes.argtypes = [vp, cp]
        L.rihvr_save_planes.restype = st
        L.rihvr_reconstruct_batch.argtypes = [vp, vp, ip, ctypes.c_int32, vp, ip, ctypes.c_int32,
                                              _P(ctypes.c_double), ctypes.c_int32, vp, ip, _P(FrameStats)]
        L.rihvr_reconstruct_batch.restype = st
        L.rihvr_create_scheduled.argtypes = [_P(Params), _P(Schedule), _P(vp)]
        L.rihvr_create_scheduled.restype = st
        L.rihvr_get_schedule.argtypes = [vp, _P(Schedule)]
        L.rihvr_get_schedule.restype = st
        L.rihvr_autotune.argtypes = [_P(Params), vp, ip, ctypes.c_int32, _P(Schedule), _P(Schedule), ctypes.c_int32,
                                     _P(ctypes.c_int32)]
        L.rihvr_autotune.restype = st
        L.rihvr_reconstruct_multi.argtypes = [_P(MultiItem), ctypes.c_int32, ip, ip, ctypes.c_int32]
        L.rihvr_reconstruct_multi.restype = st
        L.rihvr_estimate_lipschitz.argtypes = [vp, ctypes.c_int32, ctypes.c_int32, _P(ctypes.c_double)]
        L.rihvr_estimate_lipschitz.restype = st
        L.rihvr_write_png8.argtypes = [cp, vp, ctypes.c_int32, ctypes.c_int32]
        L.rihvr_write_png8.restype = st
        L.rihvr_background_mean.argtypes = [vp, ctypes.c_int32, ctypes.c_int32, ctypes.c_int32, ip, vp,
                                            _P(BackgroundStats)]
        L.rihvr_background_mean.restype = st
        L.rihvr_background_median5.argtypes = [vp, ctypes.c_int32, ctypes.c_int32, ip, vp]
        L.rihvr_background_median5.restype = st
        i64, i64p = ctypes.c_int64, _P(ctypes.c_int64)
        L.rihvr_read_sparse_voxels.argtypes = [cp, ctypes.c_int32, ctypes.c_int32, ctypes.c_int32, i64p, vp, vp,
                                               vp]
        L.rihvr_read_sparse_voxels.restype = st
        L.rihvr_segmentation_create.argtypes = [ctypes.c_int32, ctypes.c_int32, ctypes.c_int32, _P(vp)]
        L.rihvr_segmentation_create.restype = st
        L.rihvr_segmentation_destroy.argtypes = [vp]
        L.rihvr_segmentation_destroy.restype = None
        L.rihvr_segmentation_load.argtypes = [vp, i64, vp, vp, vp, ip]
        L.rihvr_segmentation_load.restype = st
        L.rihvr_segmentation_load_dense.argtypes = [vp, vp, ip]
        L.rihvr_segmentation_load_dense.restype = st
        L.rihvr_segmentation_binarize.argtypes = [vp, ctypes.c_double, _P(ctypes.c_double), _P(ctypes.c_double)]
        L.rihvr_segmentation_binarize.restype = st
        L.rihvr_segmentation_close.argtypes = [vp, ctypes.c_int32]
        L.rihvr_segmentation_close.restype = st
        L.rihvr_segmentation_label.argtypes = [vp]
        L.rihvr_segmentation_label.restype = st
        L.rihvr_segmentation_extract.argtypes = [vp, i64p]
        L.rihvr_segmentation_extract.restype = st
        L.rihvr_segmentation_size_filter.argtypes = [vp, ctypes.c_int32, i64p]
        L.rihvr_segmentation_size_filter.restype = st
        L.rihvr_segmentation_objects.argtypes = [vp, vp, i64, i64p]
        L.rihvr_segmentation_objects.restype = st
        L.rihvr_segmentation_write_centroids.argtypes = [vp, cp, ctypes.c_int32, ctypes.c_int32]
        L.rihvr_segmentation_write_centroids.restype = st
        for name in ("rihvr_segmentation_mask", "rihvr_segmentation_values", "rihvr_segmentation_labels"):
            getattr(L, name).argtypes = [vp, vp, ip]
            getattr(L, name).restype = st
        L.rihvr_segmentation_timing.argtypes = [vp, _P(ctypes.c_double)]
        L.rihvr_segmentation_timing.restype = st
        L.rihvr_prepare_holograms.argtypes = [_P(FileParams), vp, ctypes.c_int32, ctypes.c_int32, vp, ctypes.c_int32,
                                              ctypes.c_int32, ip, vp, ip, i32p, i32p]
        L.rihvr_prepare_holograms.restype = st
        L.rihvr_prepare_holograms_noalloc.argtypes = [_P(FileParams), vp, ctypes.c_int32, ctypes.c_int32, vp,
                                                      ctypes.c_int32, ctypes.c_int32, vp, i32p, i32p]
        L.rihvr_prepare_holograms_noalloc.restype = st
        i32 = ctypes.c_int32
        L.rihvr_net_param_count.argtypes = [i64p]
        L.rihvr_net_create.argtypes = [i32, i32, i32, _P(vp)]
        L.rihvr_net_destroy.argtypes = [vp]
        L.rihvr_net_init.argtypes = [vp, ctypes.c_uint64]
        L.rihvr_net_set_params.argtypes = [vp, vp, ctypes.c_float, ctypes.c_float]
        L.rihvr_net_get_params.argtypes = [vp, vp, _P(ctypes.c_float), _P(ctypes.c_float)]
        L.rihvr_net_predict.argtypes = [vp, vp, vp]
        L.rihvr_net_loss_grad.argtypes = [vp, vp, vp, _P(ctypes.c_double), vp]
        L.rihvr_net_adam_step.argtypes = [vp, ctypes.c_float]
        L.rihvr_net_set_loss_weight.argtypes = [vp, ctypes.c_float]
        L.rihvr_background_correlation.argtypes = [vp, i32, i32, i32, ip, i32, i32, i32, vp, vp, vp]
        L.rihvr_background_correlation.restype = st
        L.rihvr_net_train_step.argtypes = [vp, vp, vp, ctypes.c_float, _P(ctypes.c_double)]
        L.rihvr_net_sync.argtypes = [vp]
        L.rihvr_conv3d_bf16_forward.argtypes = [vp, i32, vp, i32, i32, i32, i32, vp, vp, i32, i32, vp]
        L.rihvr_conv3d_bf16_dgrad.argtypes = [vp, i32, i32, i32, i32, vp, i32, vp]
        L.rihvr_conv3d_bf16_wgrad.argtypes = [vp, i32, vp, i32, i32, i32, i32, vp, i32, vp]
        L.rihvr_reconstruct_learned.argtypes = [vp, vp, vp, ip, ctypes.c_double, i32, _P(LearnedStats)]
        for name in ("rihvr_net_param_count", "rihvr_net_create", "rihvr_net_destroy", "rihvr_net_init",
                     "rihvr_net_set_params", "rihvr_net_get_params", "rihvr_net_predict", "rihvr_net_loss_grad",
                     "rihvr_net_adam_step", "rihvr_net_train_step", "rihvr_net_sync", "rihvr_conv3d_bf16_forward",
                     "rihvr_conv3d_bf16_dgrad", "rihvr_conv3d_bf16_wgrad", "rihvr_reconstruct_learned",
                     "rihvr_net_set_loss_weight"):
            getattr(L, name).restype = st
        _lib = L
    return _lib


def _check(rc: int, what: str):
    if rc != 0:
        raise RihvrError(rc, what)


def _path(p) -> bytes:
    return str(p).encode()


# -- input and output of the reconstruction executable (host only) ----------
def read_param_file(path):
    """HoloUi::readFile (holo_ui.cpp:171-486).  Returns (FileParams, status):
    status 0, RIHVR_E_INVALID_FILE (9) when parsing stopped at an unknown id
    (FileParams.unknown_id; the values read before it are kept), or
    RIHVR_E_BAD_FILENAME (2) when the file cannot be opened (defaults kept)."""
    fp = FileParams()
    rc = lib().rihvr_read_param_file(_path(path), ctypes.byref(fp))
    return fp, int(rc)


def read_image(path) -> np.ndarray:
    """Raw grayscale values ([H][W] float32) of an uncompressed TIFF or PGM file."""
    w, h = ctypes.c_int32(0), ctypes.c_int32(0)
    _check(lib().rihvr_read_image(_path(path), None, ctypes.byref(w), ctypes.byref(h)), "rihvr_read_image")
    out = np.empty((h.value, w.value), np.float32)
    _check(lib().rihvr_read_image(_path(path), ctypes.c_void_p(out.ctypes.data), ctypes.byref(w),
                                  ctypes.byref(h)), "rihvr_read_image")
    return out


def prepare_hologram(fp, image, background=None) -> np.ndarray:
    """The executable's input contract (main_sparse_inverse.cu:101-124): [H'][W'] complex64."""
    img = np.ascontiguousarray(image, np.float32)
    bg = None if background is None else np.ascontiguousarray(background, np.float32)
    if bg is not None and bg.shape != img.shape:
        raise ValueError("background must have the image's shape")
    h, w = img.shape
    bp = None if bg is None else ctypes.c_void_p(bg.ctypes.data)
    ow, oh = ctypes.c_int32(0), ctypes.c_int32(0)
    _check(lib().rihvr_prepare_hologram(ctypes.byref(fp), ctypes.c_void_p(img.ctypes.data), bp, w, h, None,
                                        ctypes.byref(ow), ctypes.byref(oh)), "rihvr_prepare_hologram")
    out = np.empty((oh.value, ow.value), np.complex64)
    _check(lib().rihvr_prepare_hologram(ctypes.byref(fp), ctypes.c_void_p(img.ctypes.data), bp, w, h,
                                        ctypes.c_void_p(out.ctypes.data), ctypes.byref(ow), ctypes.byref(oh)),
           "rihvr_prepare_hologram")
    return out


PIXEL_TYPES = {"uint8": 1, "uint16": 2, "float32": 4}


def prepare_holograms(fp, frames, background=None, out=None):
    """rihvr_prepare_holograms: the input contract on the device for frames [n][H][W] of
    raw grey values (numpy or CUDA tensor; uint8, uint16 or float32) -> complex64
    holograms [n][H'][W'] (numpy, or written into ``out``, a CUDA tensor)."""
    if _is_torch(frames):
        if not frames.is_cuda or not frames.is_contiguous():
            raise ValueError("device frames must be contiguous CUDA tensors")
        dtype = str(frames.dtype).replace("torch.", "")
        fptr, floc, keep = ctypes.c_void_p(frames.data_ptr()), DEVICE, frames
    else:
        keep = np.ascontiguousarray(frames)
        dtype = str(keep.dtype)
        fptr, floc = ctypes.c_void_p(keep.ctypes.data), HOST
    if dtype not in PIXEL_TYPES:
        raise ValueError(f"pixel type {dtype}")
    n, h, w = keep.shape
    bkeep = None
    bptr = None
    if background is not None:
        # the background goes where the frames are: a contiguous float32 (h, w) CUDA tensor on
        # the frames' device with device frames, a float32 (h, w) host array with host frames
        if floc == DEVICE:
            import torch
            if (not _is_torch(background) or not background.is_cuda or background.dtype != torch.float32
                    or not background.is_contiguous() or tuple(background.shape) != (h, w)
                    or background.device != keep.device):
                raise ValueError(f"with device frames the background must be a contiguous float32 CUDA tensor "
                                 f"of shape {(h, w)} on {keep.device}")
            bkeep = background
            bptr = ctypes.c_void_p(background.data_ptr())
        else:
            if _is_torch(background):
                raise ValueError("with host frames the background must be a host array")
            bkeep = np.ascontiguousarray(background, np.float32)
            if bkeep.shape != (h, w):
                raise ValueError(f"background shape {bkeep.shape} != frame shape {(h, w)}")
            bptr = ctypes.c_void_p(bkeep.ctypes.data)
    ow, oh = ctypes.c_int32(), ctypes.c_int32()
    L = lib()
    _check(L.rihvr_prepare_holograms(ctypes.byref(fp), fptr, PIXEL_TYPES[dtype], n, bptr, w, h, floc, None, HOST,
                                     ctypes.byref(ow), ctypes.byref(oh)), "rihvr_prepare_holograms")
    shape = (n, oh.value, ow.value)
    if out is None:
        out = np.empty(shape, np.complex64)
    op, oloc, _k = _arg(out, shape, writable=True)
    _check(L.rihvr_prepare_holograms(ctypes.byref(fp), fptr, PIXEL_TYPES[dtype], n, bptr, w, h, floc, op, oloc,
                                     ctypes.byref(ow), ctypes.byref(oh)), "rihvr_prepare_holograms")
    return out


def prepare_holograms_noalloc(fp, frames, out, background=None):
    """rihvr_prepare_holograms_noalloc: rihvr_prepare_holograms for device frames (a contiguous
    uint8, uint16 or float32 CUDA tensor [n][H][W]) into ``out`` (a contiguous complex64 CUDA
    tensor [n][H'][W']), with no device allocation; the batch is bounded by the library's 1 MiB
    scratch (RihvrError RIHVR_E_INVALID_ARGUMENT beyond it)."""
    import torch
    if not _is_torch(frames) or not frames.is_cuda or not frames.is_contiguous():
        raise ValueError("frames must be a contiguous CUDA tensor")
    dtype = str(frames.dtype).replace("torch.", "")
    if dtype not in PIXEL_TYPES:
        raise ValueError(f"pixel type {dtype}")
    n, h, w = frames.shape
    bptr = None
    if background is not None:
        if (not _is_torch(background) or not background.is_cuda or background.dtype != torch.float32
                or not background.is_contiguous() or tuple(background.shape) != (h, w)
                or background.device != frames.device):
            raise ValueError(f"the background must be a contiguous float32 CUDA tensor of shape {(h, w)} "
                             f"on {frames.device}")
        bptr = ctypes.c_void_p(background.data_ptr())
    ow, oh = ctypes.c_int32(), ctypes.c_int32()
    L = lib()
    fptr = ctypes.c_void_p(frames.data_ptr())
    _check(L.rihvr_prepare_holograms_noalloc(ctypes.byref(fp), fptr, PIXEL_TYPES[dtype], n, bptr, w, h, None,
                                             ctypes.byref(ow), ctypes.byref(oh)), "rihvr_prepare_holograms_noalloc")
    shape = (n, oh.value, ow.value)
    if (not _is_torch(out) or not out.is_cuda or out.dtype != torch.complex64 or not out.is_contiguous()
            or tuple(out.shape) != shape or out.device != frames.device):
        raise ValueError(f"out must be a contiguous complex64 CUDA tensor of shape {shape} on {frames.device}")
    _check(L.rihvr_prepare_holograms_noalloc(ctypes.byref(fp), fptr, PIXEL_TYPES[dtype], n, bptr, w, h,
                                             ctypes.c_void_p(out.data_ptr()), ctypes.byref(ow), ctypes.byref(oh)),
           "rihvr_prepare_holograms_noalloc")
    return out


def write_tiff8(path, pixels):
    px = np.ascontiguousarray(pixels, np.uint8)
    h, w = px.shape
    _check(lib().rihvr_write_tiff8(_path(path), ctypes.c_void_p(px.ctypes.data), w, h), "rihvr_write_tiff8")


def write_png8(path, pixels):
    px = np.ascontiguousarray(pixels, np.uint8)
    h, w = px.shape
    _check(lib().rihvr_write_png8(_path(path), ctypes.c_void_p(px.ctypes.data), w, h), "rihvr_write_png8")


def _frames_arg(frames):
    """(pointer, location, keepalive, shape) for float32 frames: numpy (host) or a CUDA tensor."""
    if _is_torch(frames):
        import torch
        if frames.dtype != torch.float32 or not frames.is_contiguous() or not frames.is_cuda:
            raise ValueError("device frames must be contiguous float32 CUDA tensors")
        return ctypes.c_void_p(frames.data_ptr()), DEVICE, frames, tuple(frames.shape)
    f = np.ascontiguousarray(frames, np.float32)
    return ctypes.c_void_p(f.ctypes.data), HOST, f, f.shape


def background_mean(frames):
    """rihvr_background_mean: frames [n][H][W] raw grey values (numpy, or a float32 CUDA
    tensor) -> (background [H][W] float32, stats dict: alpha, beta, minimum, average)."""
    fp, loc, _keep, (n, h, w) = _frames_arg(frames)
    bg = np.empty((h, w), np.float32)
    st = BackgroundStats()
    _check(lib().rihvr_background_mean(fp, n, w, h, loc, ctypes.c_void_p(bg.ctypes.data), ctypes.byref(st)),
           "rihvr_background_mean")
    return bg, dict(alpha=st.alpha, beta=st.beta, minimum=st.minimum, average=st.average)


def background_median5(frames5):
    fp, loc, _keep, shape = _frames_arg(frames5)
    if shape[0] != 5:
        raise ValueError("five frames")
    _, h, w = shape
    out = np.empty((h, w), np.float32)
    _check(lib().rihvr_background_median5(fp, w, h, loc, ctypes.c_void_p(out.ctypes.data)),
           "rihvr_background_median5")
    return out


def background_correlation(frames, min_frames=5, max_frames=100, resize_factor=4, enhanced=False):
    """rihvr_background_correlation (ccbg.cpp): frames [n][H][W] raw grey values (numpy, or a
    float32 CUDA tensor) -> (backgrounds uint8 [n][H][W], frames used [n], enhanced or None)."""
    fp, loc, _keep, (n, h, w) = _frames_arg(frames)
    bg = np.empty((n, h, w), np.uint8)
    used = np.empty(n, np.int32)
    enh = np.empty((n, h, w), np.uint8) if enhanced else None
    _check(lib().rihvr_background_correlation(fp, n, w, h, loc, int(min_frames), int(max_frames), int(resize_factor),
                                              ctypes.c_void_p(bg.ctypes.data),
                                              ctypes.c_void_p(enh.ctypes.data) if enhanced else None,
                                              ctypes.c_void_p(used.ctypes.data)), "rihvr_background_correlation")
    return bg, used, enh


def write_hologram(path, holo, lo=None, hi=None):
    """Hologram::write: min/max stretch of the real part, or 255 (v - lo) / (hi - lo)."""
    a = np.ascontiguousarray(holo, np.complex64)
    h, w = a.shape
    minmax = lo is None
    _check(lib().rihvr_write_hologram(_path(path), ctypes.c_void_p(a.ctypes.data), w, h, 1 if minmax else 0,
                                      0.0 if minmax else float(lo), 0.0 if minmax else float(hi)),
           "rihvr_write_hologram")


def _is_torch(a) -> bool:
    return type(a).__module__.startswith("torch")


def _arg(a, shape, writable=False):
    """(pointer, location, keepalive) for a complex64 numpy array or CUDA torch tensor."""
    if a is None:
        return None, HOST, None
    if _is_torch(a):
        import torch
        if a.dtype != torch.complex64 or not a.is_contiguous() or not a.is_cuda:
            raise ValueError("device arrays must be contiguous complex64 CUDA tensors")
        if tuple(a.shape) != tuple(shape):
            raise ValueError(f"shape {tuple(a.shape)} != {tuple(shape)}")
        return ctypes.c_void_p(a.data_ptr()), DEVICE, a
    arr = np.asarray(a)
    if arr.dtype != np.complex64 or not arr.flags.c_contiguous:
        if writable:
            raise ValueError("output arrays must be contiguous complex64")
        arr = np.ascontiguousarray(arr, dtype=np.complex64)
    if tuple(arr.shape) != tuple(shape):
        raise ValueError(f"shape {tuple(arr.shape)} != {tuple(shape)}")
    return ctypes.c_void_p(arr.ctypes.data), HOST, arr


class SparseCompressiveHolo:
    """Mirror of umnholo::SparseCompressiveHolo over the C ABI."""

    def __init__(self, params: Params, schedule: "Schedule" = None):
        self.params = params
        self.shape_holo = (params.height, params.width)
        self.shape_vol = (params.num_planes, params.height, params.width)
        self._ctx = ctypes.c_void_p()
        if schedule is None:
            _check(lib().rihvr_create(ctypes.byref(params), ctypes.byref(self._ctx)), "rihvr_create")
        else:
            _check(lib().rihvr_create_scheduled(ctypes.byref(params), ctypes.byref(schedule), ctypes.byref(self._ctx)),
                   "rihvr_create_scheduled")
        self.last_stats = None

    def estimateLipschitz(self, seed=0, iterations=1) -> float:
        """rihvr_estimate_lipschitz: ||A^T A|| by `iterations` power steps from a seeded volume."""
        L = ctypes.c_double(0.0)
        _check(lib().rihvr_estimate_lipschitz(self._ctx, int(seed), int(iterations), ctypes.byref(L)),
               "rihvr_estimate_lipschitz")
        return L.value

    def schedule(self) -> dict:
        """rihvr_get_schedule: the launch choices this context runs with."""
        s = Schedule()
        _check(lib().rihvr_get_schedule(self._ctx, ctypes.byref(s)), "rihvr_get_schedule")
        return s.as_dict()

    # -- lifecycle --------------------------------------------------------
    def destroy(self):
        if self._ctx:
            lib().rihvr_destroy(self._ctx)
            self._ctx = ctypes.c_void_p()

    def __del__(self):
        try:
            self.destroy()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.destroy()

    # -- operations ---------------------------------------------------------
    def inverseReconstruct(self, holo, mode=None, warm_start=None) -> dict:
        if mode is not None and int(mode) != self.params.mode:
            raise ValueError("mode is fixed at construction (rihvr_params.mode)")
        hp, hl, _k1 = _arg(holo, self.shape_holo)
        wp, wl, _k2 = _arg(warm_start, self.shape_vol)
        sb = StatsBuffers(self.params.num_inverse_iterations)
        _check(lib().rihvr_reconstruct(self._ctx, hp, hl, wp, wl, ctypes.byref(sb.stats)),
               "rihvr_reconstruct")
        self.last_stats = sb.as_dict()
        return self.last_stats

    def reconstructLearned(self, net: "WarmStartNet", holo, f_target=0.0, max_iterations=0) -> dict:
        """rihvr_reconstruct_learned: the network's start (A^T b -> net), the fallback test
        against the A^T b start, then FISTA to f_target (<= 0: none) or max_iterations."""
        hp, hl, _k = _arg(holo, self.shape_holo)
        st = LearnedStats()
        _check(lib().rihvr_reconstruct_learned(self._ctx, net._net, hp, hl, float(f_target), int(max_iterations),
                                               ctypes.byref(st)), "rihvr_reconstruct_learned")
        return st.as_dict()

    def reconstructBatch(self, holos, warm_start=None, warm_from_previous=False, f_target=None,
                         max_iterations=0, volumes=None) -> list:
        """rihvr_reconstruct_batch: frames [n][H][W] one after the other; see the
        header for the start and stop rules.  Returns one dict per frame."""
        n = int(holos.shape[0])
        hp, hl, _k1 = _arg(holos, (n,) + self.shape_holo)
        wp, wl, _k2 = _arg(warm_start, self.shape_vol)
        tg = None
        if f_target is not None:
            tg_arr = np.ascontiguousarray(f_target, np.float64)
            if tg_arr.shape != (n,):
                raise ValueError("f_target needs one value per frame")
            tg = tg_arr.ctypes.data_as(_P(ctypes.c_double))
        vp_, vl, _k3 = (None, HOST, None) if volumes is None else _arg(volumes, (n,) + self.shape_vol,
                                                                       writable=True)
        fs = (FrameStats * max(1, n))()
        _check(lib().rihvr_reconstruct_batch(self._ctx, hp, hl, n, wp, wl, 1 if warm_from_previous else 0, tg,
                                             int(max_iterations), vp_, vl, fs), "rihvr_reconstruct_batch")
        return [dict(iterations=f.iterations, total_backtracks=f.total_backtracks, reached=bool(f.reached),
                     f_initial=f.f_initial, f_final=f.f_final, nnz=f.nnz) for f in fs[:n]]

    # -- stepwise solve (rihvr_begin / rihvr_iterate / rihvr_finish) -------
    def begin(self, holo, warm_start=None) -> float:
        hp, hl, _k1 = _arg(holo, self.shape_holo)
        wp, wl, _k2 = _arg(warm_start, self.shape_vol)
        f0 = ctypes.c_double(0.0)
        _check(lib().rihvr_begin(self._ctx, hp, hl, wp, wl, ctypes.byref(f0)), "rihvr_begin")
        return f0.value

    def iterate(self, n: int, with_objective: bool = False) -> dict:
        sb = StatsBuffers(n)
        if not with_objective:
            sb.stats.f = _P(ctypes.c_double)()
            sb.stats.f_resid = _P(ctypes.c_double)()
        _check(lib().rihvr_iterate(self._ctx, int(n), ctypes.byref(sb.stats)), "rihvr_iterate")
        return sb.as_dict()

    def finish(self) -> dict:
        sb = StatsBuffers(1)
        _check(lib().rihvr_finish(self._ctx, ctypes.byref(sb.stats)), "rihvr_finish")
        d = sb.as_dict()
        return dict(f_final=d["f_final"], nnz=d["nnz"], sparsity=d["sparsity"], stepsize=d["stepsize"],
                    total_backtracks=d["total_backtracks"], f_initial=d["f_initial"])

    def set_timing(self, enable: bool, only: str | None = None):
        """Per-kernel event timing; with ``only`` (a name of get_timing) that kernel alone."""
        code = 1 if enable else 0
        if enable and only is not None:
            code = 2 + list(self.get_timing().keys()).index(only)
        _check(lib().rihvr_set_timing(self._ctx, code), "rihvr_set_timing")

    def get_timing(self) -> dict:
        n = ctypes.c_int32(0)
        _check(lib().rihvr_get_timing(self._ctx, None, 0, ctypes.byref(n)), "rihvr_get_timing")
        arr = (KernelTime * n.value)()
        _check(lib().rihvr_get_timing(self._ctx, arr, n.value, ctypes.byref(n)), "rihvr_get_timing")
        return {k.name.decode(): (int(k.launches), float(k.total_ms)) for k in arr}

    def getVolume(self, out=None):
        out = np.empty(self.shape_vol, np.complex64) if out is None else out
        p, loc, _k = _arg(out, self.shape_vol, writable=True)
        _check(lib().rihvr_get_volume(self._ctx, p, loc), "rihvr_get_volume")
        return out

    def getPlane(self, z: int, out=None):
        out = np.empty(self.shape_holo, np.complex64) if out is None else out
        p, loc, _k = _arg(out, self.shape_holo, writable=True)
        _check(lib().rihvr_get_plane(self._ctx, int(z), p, loc), "rihvr_get_plane")
        return out

    def getParams(self) -> Params:
        p = Params()
        _check(lib().rihvr_get_params(self._ctx, ctypes.byref(p)), "rihvr_get_params")
        return p

    def snapshot(self):
        """Makes the current momentum point of a stepwise solve readable (getVolume,
        getPlane, getEstimate, save*) without ending the solve."""
        _check(lib().rihvr_snapshot(self._ctx), "rihvr_snapshot")

    def saveSparse(self, csv_path):
        _check(lib().rihvr_save_sparse(self._ctx, _path(csv_path)), "rihvr_save_sparse")

    def saveProjections(self, output_path, prefix, prefix_as_suffix=False):
        _check(lib().rihvr_save_projections(self._ctx, _path(output_path), _path(prefix),
                                            1 if prefix_as_suffix else 0), "rihvr_save_projections")

    def savePlanes(self, prefix):
        _check(lib().rihvr_save_planes(self._ctx, _path(prefix)), "rihvr_save_planes")

    def setStepsize(self, step: float):
        _check(lib().rihvr_set_stepsize(self._ctx, float(step)), "rihvr_set_stepsize")

    def getEstimate(self):
        est = np.empty(self.shape_holo, np.complex64)
        res = np.empty(self.shape_holo, np.complex64)
        _check(lib().rihvr_get_estimate(self._ctx, ctypes.c_void_p(est.ctypes.data),
                                        ctypes.c_void_p(res.ctypes.data), HOST), "rihvr_get_estimate")
        return est, res

    def forward(self, vol, out=None):
        vp, vl, _k = _arg(vol, self.shape_vol)
        out = np.empty(self.shape_holo, np.complex64) if out is None else out
        op, ol, _k2 = _arg(out, self.shape_holo, writable=True)
        _check(lib().rihvr_forward(self._ctx, vp, vl, op, ol), "rihvr_forward")
        return out

    def adjoint(self, holo, out=None):
        hp, hl, _k = _arg(holo, self.shape_holo)
        out = np.empty(self.shape_vol, np.complex64) if out is None else out
        op, ol, _k2 = _arg(out, self.shape_vol, writable=True)
        _check(lib().rihvr_adjoint(self._ctx, hp, hl, op, ol), "rihvr_adjoint")
        return out

    def prox(self, x0, grad, stepsize, out=None):
        xp, xl, _k = _arg(x0, self.shape_vol)
        gp, gl, _k2 = _arg(grad, self.shape_vol)
        if xl != gl:
            raise ValueError("x0 and grad must be in the same memory")
        if out is None:
            out = np.empty(self.shape_vol, np.complex64)
        op, ol, _k3 = _arg(out, self.shape_vol, writable=True)
        if ol != xl:
            raise ValueError("out must be in the same memory as x0")
        _check(lib().rihvr_prox(self._ctx, xp, gp, float(stepsize), op, ol), "rihvr_prox")
        return out


# -- segmentation and centroids (the segmentation executable) ---------------
# -- learned warm start ------------------------------------------------------------
def autotune(params: Params, holo=None, iterations=20, capacity=64):
    """rihvr_autotune: (best Schedule, [every measured candidate as a dict])."""
    hp, hl, _k = _arg(holo, (params.height, params.width))
    best = Schedule()
    tried = (Schedule * capacity)()
    n = ctypes.c_int32(0)
    _check(lib().rihvr_autotune(ctypes.byref(params), hp, hl, int(iterations), ctypes.byref(best), tried, capacity,
                                ctypes.byref(n)), "rihvr_autotune")
    return best, [t.as_dict() for t in tried[:min(n.value, capacity)]]


def reconstruct_multi(items, autotune=0):
    """rihvr_reconstruct_multi over items: dicts with "params" (Params), "holo" ([H][W] complex64),
    optional "volume" (a writable [Nz][H][W] complex64 output), "f_target" and "schedule" (a
    Schedule, read with autotune=2).  autotune: 0 (or False) the defaults, 1 (or True) tune each
    group, 2 the groups' given schedules.  Inputs and outputs must all be host arrays or all CUDA
    tensors.  Returns one dict per item (the frame statistics and the schedule it ran with)."""
    n = len(items)
    arr = (MultiItem * max(1, n))()
    keep = []
    hloc = vloc = None
    for k, it in enumerate(items):
        p = it["params"]
        hp, hl, kh = _arg(it["holo"], (p.height, p.width))
        keep.append(kh)
        if hloc is not None and hl != hloc:
            raise ValueError("holograms must all be host arrays or all device tensors")
        hloc = hl
        arr[k].params = p
        arr[k].holo = hp
        vol = it.get("volume")
        if vol is not None:
            vp_, vl, kv = _arg(vol, (p.num_planes, p.height, p.width), writable=True)
            keep.append(kv)
            if vloc is not None and vl != vloc:
                raise ValueError("volumes must all be host arrays or all device tensors")
            vloc = vl
            arr[k].volume = vp_
        arr[k].f_target = float(it.get("f_target", 0.0))
        if it.get("schedule") is not None:
            arr[k].schedule = it["schedule"]
    _check(lib().rihvr_reconstruct_multi(arr, n, HOST if hloc is None else hloc, HOST if vloc is None else vloc,
                                         int(autotune)), "rihvr_reconstruct_multi")
    out = []
    for a in arr[:n]:
        f = a.stats
        out.append(dict(iterations=f.iterations, total_backtracks=f.total_backtracks, reached=bool(f.reached),
                        f_initial=f.f_initial, f_final=f.f_final, nnz=f.nnz, schedule=a.schedule.as_dict()))
    return out


def net_param_count() -> int:
    n = ctypes.c_int64()
    _check(lib().rihvr_net_param_count(ctypes.byref(n)), "rihvr_net_param_count")
    return n.value


def _dev(t, dtype, shape=None, what="tensor"):
    """Device pointer of a contiguous CUDA tensor of the given torch dtype (and shape)."""
    if not _is_torch(t) or not t.is_cuda or not t.is_contiguous() or t.dtype != dtype:
        raise ValueError(f"{what} must be a contiguous {dtype} CUDA tensor")
    if shape is not None and tuple(t.shape) != tuple(shape):
        raise ValueError(f"{what} shape {tuple(t.shape)} != {tuple(shape)}")
    return ctypes.c_void_p(t.data_ptr())


class WarmStartNet:
    """The learned warm start's network (rihvr_net_*): volumes are [Nz][H][W] complex64 CUDA
    tensors; parameters travel as float32 numpy arrays."""

    def __init__(self, num_planes, height, width, seed=None):
        self.shape = (int(num_planes), int(height), int(width))
        self._net = ctypes.c_void_p()
        _check(lib().rihvr_net_create(*self.shape, ctypes.byref(self._net)), "rihvr_net_create")
        self.nparams = net_param_count()
        if seed is not None:
            self.init(seed)

    def destroy(self):
        if self._net:
            lib().rihvr_net_destroy(self._net)
            self._net = ctypes.c_void_p()

    def __del__(self):
        try:
            self.destroy()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.destroy()

    def init(self, seed: int):
        _check(lib().rihvr_net_init(self._net, int(seed)), "rihvr_net_init")

    def set_params(self, params, in_scale=1.0, out_scale=1.0):
        a = np.ascontiguousarray(params, np.float32)
        if a.size != self.nparams:
            raise ValueError(f"{a.size} parameters, expected {self.nparams}")
        _check(lib().rihvr_net_set_params(self._net, ctypes.c_void_p(a.ctypes.data), float(in_scale),
                                          float(out_scale)), "rihvr_net_set_params")

    def get_params(self):
        a = np.empty(self.nparams, np.float32)
        i, o = ctypes.c_float(), ctypes.c_float()
        _check(lib().rihvr_net_get_params(self._net, ctypes.c_void_p(a.ctypes.data), ctypes.byref(i),
                                          ctypes.byref(o)), "rihvr_net_get_params")
        return a, i.value, o.value

    def save(self, path, **meta):
        p, i, o = self.get_params()
        np.savez(path, params=p, in_scale=np.float32(i), out_scale=np.float32(o), shape=np.array(self.shape),
                 **{k: np.asarray(v) for k, v in meta.items()})

    def load(self, path):
        with np.load(path, allow_pickle=False) as z:
            self.set_params(z["params"], float(z["in_scale"]), float(z["out_scale"]))

    def predict(self, vol, out=None):
        import torch
        if out is None:
            out = torch.empty(self.shape, dtype=torch.complex64, device=vol.device)
        _check(lib().rihvr_net_predict(self._net, _dev(vol, torch.complex64, self.shape, "vol"),
                                       _dev(out, torch.complex64, self.shape, "out")), "rihvr_net_predict")
        return out

    def loss_grad(self, vol, target, with_grads=True):
        import torch
        loss = ctypes.c_double()
        g = np.empty(self.nparams, np.float32) if with_grads else None
        _check(lib().rihvr_net_loss_grad(self._net, _dev(vol, torch.complex64, self.shape, "vol"),
                                         _dev(target, torch.complex64, self.shape, "target"), ctypes.byref(loss),
                                         ctypes.c_void_p(g.ctypes.data) if g is not None else None),
               "rihvr_net_loss_grad")
        return loss.value, g

    def set_loss_weight(self, pos_weight):
        _check(lib().rihvr_net_set_loss_weight(self._net, float(pos_weight)), "rihvr_net_set_loss_weight")

    def adam_step(self, lr):
        _check(lib().rihvr_net_adam_step(self._net, float(lr)), "rihvr_net_adam_step")

    def train_step(self, vol, target, lr) -> float:
        import torch
        loss = ctypes.c_double()
        _check(lib().rihvr_net_train_step(self._net, _dev(vol, torch.complex64, self.shape, "vol"),
                                          _dev(target, torch.complex64, self.shape, "target"), float(lr),
                                          ctypes.byref(loss)), "rihvr_net_train_step")
        return loss.value

    def sync(self):
        _check(lib().rihvr_net_sync(self._net), "rihvr_net_sync")


def conv3d_bf16_forward(in0, weight, bias=None, in1=None, relu=False):
    """rihvr_conv3d_bf16_forward on int16 CUDA tensors of BF16 bit patterns: in0 [D][H][W][C0],
    in1 [D/2][H/2][W/2][C1] or None, weight [27][Cout][C0+C1], bias float32 [Cout] or None."""
    import torch
    D, H, W, C0 = in0.shape
    C1 = 0 if in1 is None else in1.shape[3]
    cout = weight.shape[1]
    out = torch.empty((D, H, W, cout), dtype=torch.int16, device=in0.device)
    _check(lib().rihvr_conv3d_bf16_forward(
        _dev(in0, torch.int16), C0, None if in1 is None else _dev(in1, torch.int16), C1, D, H, W,
        _dev(weight, torch.int16, (27, cout, C0 + C1)), None if bias is None else _dev(bias, torch.float32, (cout,)),
        cout, 1 if relu else 0, _dev(out, torch.int16)), "rihvr_conv3d_bf16_forward")
    return out


def conv3d_bf16_dgrad(grad_out, weight_t):
    """rihvr_conv3d_bf16_dgrad: grad_out [D][H][W][Cout] (BF16 bits), weight_t [27][Cin][Cout]."""
    import torch
    D, H, W, cout = grad_out.shape
    cin = weight_t.shape[1]
    out = torch.empty((D, H, W, cin), dtype=torch.float32, device=grad_out.device)
    _check(lib().rihvr_conv3d_bf16_dgrad(_dev(grad_out, torch.int16), cout, D, H, W,
                                         _dev(weight_t, torch.int16, (27, cin, cout)), cin,
                                         _dev(out, torch.float32)), "rihvr_conv3d_bf16_dgrad")
    return out


def conv3d_bf16_wgrad(in0, grad_out, in1=None):
    """rihvr_conv3d_bf16_wgrad: returns (dW [27][Cout][Cin] float32, db [Cout] float32)."""
    import torch
    D, H, W, C0 = in0.shape
    C1 = 0 if in1 is None else in1.shape[3]
    cout = grad_out.shape[3]
    out = torch.empty(27 * cout * (C0 + C1) + cout, dtype=torch.float32, device=in0.device)
    _check(lib().rihvr_conv3d_bf16_wgrad(_dev(in0, torch.int16), C0,
                                         None if in1 is None else _dev(in1, torch.int16), C1, D, H, W,
                                         _dev(grad_out, torch.int16, (D, H, W, cout)), cout, _dev(out, torch.float32)),
           "rihvr_conv3d_bf16_wgrad")
    n = 27 * cout * (C0 + C1)
    return out[:n].reshape(27, cout, C0 + C1), out[n:]


def read_sparse_voxels(path, width, height, num_planes):
    """SparseVolume::loadData (sparse_volume.cu:562-682) of an "x, y, z, value" file:
    -> (plane int32[n], index int32[n] = y*W + x, intensity float32[n])."""
    n = ctypes.c_int64(0)
    _check(lib().rihvr_read_sparse_voxels(_path(path), int(width), int(height), int(num_planes), ctypes.byref(n),
                                          None, None, None), "rihvr_read_sparse_voxels")
    plane = np.empty(n.value, np.int32)
    index = np.empty(n.value, np.int32)
    inten = np.empty(n.value, np.float32)
    _check(lib().rihvr_read_sparse_voxels(_path(path), int(width), int(height), int(num_planes), ctypes.byref(n),
                                          ctypes.c_void_p(plane.ctypes.data), ctypes.c_void_p(index.ctypes.data),
                                          ctypes.c_void_p(inten.ctypes.data)), "rihvr_read_sparse_voxels")
    return plane, index, inten


def _list_arg(a, dtype):
    """(pointer, location, keepalive) of a 1-D voxel-list array: numpy or a CUDA tensor."""
    if _is_torch(a):
        if not a.is_cuda or not a.is_contiguous() or a.element_size() != 4:
            raise ValueError("device voxel lists must be contiguous 4-byte CUDA tensors")
        return ctypes.c_void_p(a.data_ptr()), DEVICE, a
    arr = np.ascontiguousarray(a, dtype)
    return ctypes.c_void_p(arr.ctypes.data), HOST, arr


class SparseSegmentation:
    """Mirror of umnholo::SparseSegmentation (src/core/sparse_segmentation.h) over the C ABI:
    construct with the volume size and its voxels (read_sparse_voxels' lists, or a dense
    complex64 [Nz][H][W] volume), then binarize(min_intensity), close(close_size),
    labelConnectedComponents(), extractObjects() -> ObjectCloud."""

    def __init__(self, width, height, num_planes, voxels=None, dense=None):
        self.shape = (int(num_planes), int(height), int(width))
        self._h = ctypes.c_void_p()
        _check(lib().rihvr_segmentation_create(int(width), int(height), int(num_planes), ctypes.byref(self._h)),
               "rihvr_segmentation_create")
        self.max_intensity = None
        self.threshold = None
        if voxels is not None:
            self.load(*voxels)
        elif dense is not None:
            self.loadDense(dense)

    def destroy(self):
        if self._h:
            lib().rihvr_segmentation_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.destroy()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.destroy()

    def load(self, plane, index, intensity):
        pp, pl, _k1 = _list_arg(plane, np.int32)
        ip_, il, _k2 = _list_arg(index, np.int32)
        vp, vl, _k3 = _list_arg(intensity, np.float32)
        n = len(_k1)
        if not (pl == il == vl) or len(_k2) != n or len(_k3) != n:
            raise ValueError("plane, index and intensity must be equal-length lists in the same memory")
        _check(lib().rihvr_segmentation_load(self._h, n, pp, ip_, vp, pl), "rihvr_segmentation_load")

    def loadDense(self, volume):
        p, loc, _k = _arg(volume, self.shape)
        _check(lib().rihvr_segmentation_load_dense(self._h, p, loc), "rihvr_segmentation_load_dense")

    def binarize(self, min_intensity=1.0):
        mx, thr = ctypes.c_double(), ctypes.c_double()
        _check(lib().rihvr_segmentation_binarize(self._h, float(min_intensity), ctypes.byref(mx), ctypes.byref(thr)),
               "rihvr_segmentation_binarize")
        self.max_intensity, self.threshold = mx.value, thr.value
        return self.max_intensity, self.threshold

    def close(self, close_size=3):
        _check(lib().rihvr_segmentation_close(self._h, int(close_size)), "rihvr_segmentation_close")

    def labelConnectedComponents(self):
        _check(lib().rihvr_segmentation_label(self._h), "rihvr_segmentation_label")

    def extractObjects(self) -> "ObjectCloud":
        n = ctypes.c_int64()
        _check(lib().rihvr_segmentation_extract(self._h, ctypes.byref(n)), "rihvr_segmentation_extract")
        return ObjectCloud(self)

    def _copy(self, fn, dtype, what):
        out = np.empty(self.shape, dtype)
        _check(fn(self._h, ctypes.c_void_p(out.ctypes.data), HOST), what)
        return out

    def mask(self) -> np.ndarray:
        return self._copy(lib().rihvr_segmentation_mask, np.uint8, "rihvr_segmentation_mask")

    def values(self) -> np.ndarray:
        return self._copy(lib().rihvr_segmentation_values, np.float32, "rihvr_segmentation_values")

    def labels(self) -> np.ndarray:
        return self._copy(lib().rihvr_segmentation_labels, np.uint32, "rihvr_segmentation_labels")

    def timing(self) -> dict:
        ms = (ctypes.c_double * 5)()
        _check(lib().rihvr_segmentation_timing(self._h, ms), "rihvr_segmentation_timing")
        return dict(zip(("load", "binarize", "close", "label", "extract"), list(ms)))


class ObjectCloud:
    """Mirror of umnholo::ObjectCloud (src/core/object_cloud.h) for the objects of a
    SparseSegmentation: getNumObjects(), objects(), blobSizeFilter(min_size),
    writeCentroids(filename, use_weighted_centroid, verbose)."""

    def __init__(self, seg: SparseSegmentation):
        self._seg = seg

    def getNumObjects(self) -> int:
        n = ctypes.c_int64()
        _check(lib().rihvr_segmentation_objects(self._seg._h, None, 0, ctypes.byref(n)), "rihvr_segmentation_objects")
        return n.value

    def objects(self) -> list:
        n = self.getNumObjects()
        arr = (Object * max(n, 1))()
        got = ctypes.c_int64()
        _check(lib().rihvr_segmentation_objects(self._seg._h, arr, n, ctypes.byref(got)),
               "rihvr_segmentation_objects")
        return [arr[i].as_dict() for i in range(n)]

    def blobSizeFilter(self, min_size) -> int:
        removed = ctypes.c_int64()
        _check(lib().rihvr_segmentation_size_filter(self._seg._h, int(min_size), ctypes.byref(removed)),
               "rihvr_segmentation_size_filter")
        return removed.value

    def writeCentroids(self, filename, use_weighted_centroid=False, verbose=False):
        _check(lib().rihvr_segmentation_write_centroids(self._seg._h, _path(filename), 1 if use_weighted_centroid else 0,
                                                        1 if verbose else 0), "rihvr_segmentation_write_centroids")
