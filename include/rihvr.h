/*
 * rihvr.h -- C ABI of the MI355X-native sparse regularized-inverse holographic
 * volume reconstruction (FASTA/FISTA + fused lasso), built for gfx950.
 *
 * Every entry point replaces one call of the reference's in-process API
 * (class umnholo::SparseCompressiveHolo, src/core/sparse_compressive_holo.h:25-190)
 * and keeps its parameter set, layouts and error codes:
 *
 *   hologram : [H][W] interleaved complex64 (re, im), x fastest, already
 *              normalised and zero-padded (src/core/hologram.cu:183-188, :905-908;
 *              imaginary part is zero for a measured hologram).
 *   volume   : [Nz][H][W] interleaved complex64, plane z at offset z*H*W,
 *              element (x, y) at y*W + x (src/core/sparse_volume.h:208-209 stores
 *              the same index in COO form; the dense layout holds the same values).
 *
 * Pointers are plain addresses.  Each data argument carries a location flag
 * (RIHVR_HOST or RIHVR_DEVICE); device pointers must live on the context's GPU.
 * Calls are blocking: when a call returns, its outputs are complete.
 * No entry point throws; every one returns an rihvr_status.
 * One context per host thread; a context owns its device memory, its HIP stream
 * and its FFT schedules.  Contexts share no state.
 */
#ifndef RIHVR_H
#define RIHVR_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define RIHVR_ABI_VERSION 1

/* Status codes.  1..12 are the reference's umnholo::HoloError values
 * (src/core/umnholo.h:132-146), thrown by the reference where this ABI returns. */
typedef enum rihvr_status {
    RIHVR_OK = 0,
    RIHVR_E_UNKNOWN_TYPE = 1,
    RIHVR_E_BAD_FILENAME = 2,
    RIHVR_E_MISSING_DATA = 3,
    RIHVR_E_INVALID_STATE = 4,
    RIHVR_E_OUT_OF_MEMORY = 5,
    RIHVR_E_UNKNOWN_MODE = 6,
    RIHVR_E_INVALID_ARGUMENT = 7,
    RIHVR_E_UNKNOWN_ERROR = 8,
    RIHVR_E_INVALID_FILE = 9,
    RIHVR_E_INVALID_DATA = 10,
    RIHVR_E_CRITICAL_ASSUMPTION = 11,
    RIHVR_E_INFINITE_LOOP = 12,
    /* No HoloError counterpart: a HIP runtime call failed (the reference throws
     * CUDA_ERROR_UNKNOWN from CHECK_FOR_ERROR, src/core/umnholo.h:14-23). */
    RIHVR_E_DEVICE = 13
} rihvr_status;

/* Regularisation modes: umnholo::CompressiveHoloMode values accepted by
 * SparseCompressiveHolo::inverseReconstruct (src/core/compressive_holo.h:20-29,
 * src/core/sparse_compressive_holo.cu:178-188). */
typedef enum rihvr_mode {
    RIHVR_MODE_FASTA_L1 = 5,              /* proximalL1 */
    RIHVR_MODE_FASTA_FUSED_LASSO = 6,     /* 3D TV + L1 (not built; returns RIHVR_E_UNKNOWN_MODE) */
    RIHVR_MODE_FASTA_FUSED_LASSO_2D = 7   /* per-plane 2D TV + L1: the CLI default */
} rihvr_mode;

/* Data location flags. */
enum { RIHVR_HOST = 0, RIHVR_DEVICE = 1 };

/* Reconstruction parameters: the umnholo::Parameters fields the inverse path
 * reads (src/core/holo_ui.h:83-145), with the same float types (values are
 * promoted float -> double exactly where the reference promotes them).
 * Comment numbers are the parameter-file ids (src/core/holo_ui.h:22-78). */
typedef struct rihvr_params {
    int32_t width;                  /* padded hologram width W  (multiple of 16) */
    int32_t height;                 /* padded hologram height H (multiple of 16) */
    int32_t num_planes;             /* 3003 Nz */
    float   start_plane;            /* 3001 z0 (um) */
    float   plane_stepsize;         /* 3002 dz (um); plane z_i = z0 + i*dz in float */
    float   wavelength;             /* 3004 lambda (um) */
    float   resolution;             /* 3005 dx (um / pixel) */
    int32_t zero_padding;           /* 3010: informational, the hologram passed in is padded */
    int32_t num_inverse_iterations; /* 6001 FISTA iterations */
    float   regularization_param;   /* 6002 L1 (sparsity) weight */
    float   regularization_TV_param;/* 6004 TV weight; 0 switches mode 7 to mode 5 */
    int32_t num_tv_iterations;      /* 6005 TV dual iterations per prox */
    int32_t mode;                   /* rihvr_mode */
} rihvr_params;

/* Per-solve statistics: the values the reference prints per iteration
 * (src/core/sparse_compressive_holo.cu:561-565) and its control decisions.
 * Per-iteration arrays are optional (NULL = not recorded); when given they
 * must hold num_inverse_iterations entries. */
typedef struct rihvr_stats {
    int32_t iterations;        /* iterations run */
    int32_t total_backtracks;  /* sum of backtracks */
    double  stepsize;          /* final step size */
    double  f_initial;         /* objective at the start point (:273-280) */
    double  f_final;           /* objective at the returned point */
    double  sparsity;          /* nnz / (W*H*Nz) of the returned volume */
    int64_t nnz;               /* voxels with re != 0 or im != 0 */
    double  *f;                /* objective after iteration it (at the momentum point, :548) */
    double  *f_resid;          /* 0.5*||r||^2 at the momentum point after iteration it */
    double  *ratio;            /* last backtracking ratio of iteration it (:403, :442) */
    double  *lim;              /* line-search limit of the accepted step (:4381) */
    int32_t *backtracks;       /* backtracks taken in iteration it (<= 2, :423) */
    int32_t *restarts;         /* 1 when the restart test fired in iteration it (:471-475) */
    double  *restart_sum;      /* the restart reduction value (:4481) */
} rihvr_stats;

typedef struct rihvr_ctx rihvr_ctx;

int         rihvr_abi_version(void);
const char* rihvr_status_string(rihvr_status s);

/* Replaces SparseCompressiveHolo::SparseCompressiveHolo(Hologram)
 * (src/core/sparse_compressive_holo.cu:58-135): validates the parameter set,
 * allocates the volumes on the current HIP device and builds the FFT schedules.
 * Errors: RIHVR_E_INVALID_ARGUMENT (W or H not a multiple of 16, Nz < 1,
 * non-positive wavelength/resolution, unsupported FFT length), RIHVR_E_UNKNOWN_MODE,
 * RIHVR_E_OUT_OF_MEMORY, RIHVR_E_DEVICE. */
rihvr_status rihvr_create(const rihvr_params* params, rihvr_ctx** out);

/* Replaces SparseCompressiveHolo::destroy (:137-145).  NULL is accepted. */
rihvr_status rihvr_destroy(rihvr_ctx* ctx);

/* Replaces SparseCompressiveHolo::inverseReconstruct(holo, mode)
 * (:173-633).  holo: [H][W] complex64 at holo_loc.  warm_start: NULL for the
 * reference's zero start (:258-259, sparse_volume.cu:44-48), else a [Nz][H][W]
 * complex64 start volume at warm_loc (x0 = y0 = x1 = warm_start).
 * stats may be NULL.  The result (the final FISTA momentum point x1, :483-489)
 * stays in the context; read it with rihvr_get_volume / rihvr_get_plane.
 * Errors: RIHVR_E_INVALID_DATA (non-finite hologram value),
 * RIHVR_E_INVALID_ARGUMENT, RIHVR_E_DEVICE. */
rihvr_status rihvr_reconstruct(rihvr_ctx* ctx, const float* holo, int holo_loc,
                               const float* warm_start, int warm_loc,
                               rihvr_stats* stats);

/* Stepwise form of rihvr_reconstruct, split where the reference's
 * inverseReconstruct splits (sparse_compressive_holo.cu:173-633):
 *   rihvr_begin   -- the set-up before the loop (:238-296): start point, initial
 *                    residual, objective (*f_initial, may be NULL) and gradient;
 *   rihvr_iterate -- n_iter passes of the loop body (:299-607), continuing the
 *                    FISTA state of the previous call; per-iteration arrays of
 *                    `stats` are indexed from 0 for this call; asking for
 *                    stats->f or stats->f_resid costs one extra reduction pass;
 *   rihvr_finish  -- forms the returned momentum point (:483-489) and the final
 *                    objective (:608-633).
 * rihvr_reconstruct(h, w, st) == begin(h, w) + iterate(num_inverse_iterations) + finish.
 * rihvr_iterate / rihvr_finish without an active begin: RIHVR_E_INVALID_STATE. */
rihvr_status rihvr_begin(rihvr_ctx* ctx, const float* holo, int holo_loc,
                         const float* warm_start, int warm_loc, double* f_initial);
rihvr_status rihvr_iterate(rihvr_ctx* ctx, int32_t n_iter, rihvr_stats* stats);
rihvr_status rihvr_finish(rihvr_ctx* ctx, rihvr_stats* stats);

/* Per-kernel GPU timing with HIP events on the context's stream: the
 * counterpart of the reference's DECLARE_TIMING / SAVE_TIMING phase timers
 * (src/core/umnholo.h:79-124, compiled out by default there).
 * rihvr_set_timing(ctx, 1) resets the counters and starts recording; 0 stops.
 * rihvr_get_timing writes up to `capacity` records and sets *count to the number
 * of timed kernels. */
typedef struct rihvr_kernel_time {
    char    name[32];
    int64_t launches;
    double  total_ms;
} rihvr_kernel_time;
rihvr_status rihvr_set_timing(rihvr_ctx* ctx, int32_t enable);
rihvr_status rihvr_get_timing(rihvr_ctx* ctx, rihvr_kernel_time* out, int32_t capacity,
                              int32_t* count);

/* Copies the reconstructed volume ([Nz][H][W] complex64) out.  Replaces reading
 * x1 through SparseVolume::getPlane for every plane / saveSparse (:874-879). */
rihvr_status rihvr_get_volume(rihvr_ctx* ctx, float* out, int out_loc);

/* Replaces SparseCompressiveHolo::getPlane(size_t) (sparse_compressive_holo.h:92):
 * copies plane z ([H][W] complex64).  RIHVR_E_INVALID_ARGUMENT if z is out of range. */
rihvr_status rihvr_get_plane(rihvr_ctx* ctx, int32_t z, float* out, int out_loc);

/* Replaces SparseCompressiveHolo::setStepsize(double) (sparse_compressive_holo.h:94).
 * As in the reference, inverseReconstruct then starts from 1/Nz regardless
 * (:247): the forced step only skips the Lipschitz estimate, whose result the
 * reference discards.  Kept for API parity. */
rihvr_status rihvr_set_stepsize(rihvr_ctx* ctx, double step);

/* Final estimated hologram and residual of the last solve ([H][W] complex64),
 * the data the reference writes to estimate_final.tif / residual_final.tif
 * (:621-624).  Either pointer may be NULL. */
rihvr_status rihvr_get_estimate(rihvr_ctx* ctx, float* estimate, float* residual,
                                int out_loc);

/* Forward operator A of the solver: est = Re{ IFFT( sum_z H_{-z} . FFT(x_z e^{-i phi_z}) / (W*H) ) }
 * (SparseCompressiveHolo::backPropagate, :1935-2077).  volume [Nz][H][W] -> holo [H][W]
 * (imaginary part written as 0). */
rihvr_status rihvr_forward(rihvr_ctx* ctx, const float* volume, int vol_loc,
                           float* holo_out, int holo_loc);

/* Adjoint A^T of the solver for every plane: g_z = e^{+i phi_z} IFFT(H_z . FFT(r)) / (W*H)
 * (SparseGradient::reconstruct + getPlane, src/core/sparse_gradient.cu:47-165).
 * holo [H][W] -> volume [Nz][H][W]. */
rihvr_status rihvr_adjoint(rihvr_ctx* ctx, const float* holo, int holo_loc,
                           float* volume_out, int vol_loc);

/* One proximal step of the context's mode on every plane: out = prox(x0 - t*grad)
 * (proximalFusedLasso2d :3652-3732 for mode 7, proximalL1 :3274-3392 for mode 5).
 * All volumes [Nz][H][W] complex64 at loc. */
rihvr_status rihvr_prox(rihvr_ctx* ctx, const float* x0, const float* grad,
                        double stepsize, float* out, int loc);

#ifdef __cplusplus
}
#endif

#endif /* RIHVR_H */
