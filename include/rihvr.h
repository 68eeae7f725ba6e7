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

#define RIHVR_ABI_VERSION 4

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
    RIHVR_MODE_FASTA_FUSED_LASSO = 6,     /* proximalFusedLasso: L1/2, 3D z-periodic TV, L1/2 */
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
    double  *sparsity_it;      /* nnz / (W*H*Nz) of the momentum point after iteration it (:563) */
} rihvr_stats;

typedef struct rihvr_ctx rihvr_ctx;

int         rihvr_abi_version(void);
const char* rihvr_status_string(rihvr_status s);

/* Replaces SparseCompressiveHolo::SparseCompressiveHolo(Hologram)
 * (src/core/sparse_compressive_holo.cu:58-135): validates the parameter set,
 * allocates the volumes on the current HIP device and builds the FFT schedules.
 * Errors: RIHVR_E_INVALID_ARGUMENT (W or H not a multiple of 16, Nz < 1,
 * non-positive wavelength/resolution, unsupported FFT length, a mode other than
 * 5, 6, 7 -- the reference's inverseReconstruct throws HOLO_ERROR_INVALID_ARGUMENT
 * for those, :178-189), RIHVR_E_OUT_OF_MEMORY, RIHVR_E_DEVICE.  Any mode with a
 * zero TV weight runs as mode 5 (:192-195). */
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

/* Per-frame result of rihvr_reconstruct_batch. */
typedef struct rihvr_frame_stats {
    int32_t iterations;        /* iterations run for this frame */
    int32_t total_backtracks;
    int32_t reached;           /* 1 if f_final <= the frame's target (or no target) */
    int32_t pad_;
    double  f_initial;         /* objective at the start point */
    double  f_final;           /* objective at the returned volume */
    int64_t nnz;               /* non-zero voxels of the returned volume */
} rihvr_frame_stats;

/* A sequence of holograms of this context's geometry (holos: [n_frames][H][W]
 * complex64 at holo_loc), reconstructed one after the other on the device.  No
 * counterpart call in the reference, whose executable reconstructs one image per
 * run (main_sparse_inverse.cu:101-104); each frame is the reference's
 * inverseReconstruct with these start and stop rules:
 *   start: frame 0 from warm_start (NULL: zero, as the reference); frame k > 0 from
 *          zero, or with warm_from_previous != 0 from frame k-1's returned volume
 *          (kept on the device, no copy);
 *   stop:  f_target NULL -- max_iterations iterations (<= 0: num_inverse_iterations);
 *          else frame k stops after the first iteration whose objective at the
 *          momentum point is <= f_target[k], or at max_iterations.
 * volumes ([n_frames][Nz][H][W] at vol_loc) may be NULL; the last frame's result
 * stays readable through rihvr_get_volume.  frame_stats ([n_frames]) may be NULL. */
rihvr_status rihvr_reconstruct_batch(rihvr_ctx* ctx, const float* holos, int holo_loc, int32_t n_frames,
                                     const float* warm_start, int warm_loc, int32_t warm_from_previous,
                                     const double* f_target, int32_t max_iterations, float* volumes,
                                     int vol_loc, rihvr_frame_stats* frame_stats);

/* ---- Schedules and the batched call (no counterpart in the reference, which runs one
 * hologram per process, src/QTGUI/rihvr.cpp:2414-2906).  A schedule holds the launch choices
 * that change speed but not the mathematics (the sums' order changes with them, so results
 * agree to rounding across schedules and bit for bit within one). */
typedef struct rihvr_schedule {
    int32_t fft_path;    /* 0: the fixed-length FFT kernels where the size has them, 1: run-time plans */
    int32_t max_radix;   /* largest radix of the run-time plans: 16, 8 or 4 */
    int32_t zchunk_fwd;  /* planes per partial spectrum of A */
    int32_t zchunk_adj;  /* planes per workgroup of A^T's first pass */
    int32_t prox_rows;   /* output rows per wave of the fused mode-7 prox (0: the device rule) */
    int32_t graphs;      /* 1: the queued iterations run as captured HIP graphs */
    double  ms_per_iteration;  /* measured by rihvr_autotune (else 0) */
} rihvr_schedule;

/* rihvr_create with a schedule (NULL: the defaults; environment variables RIHVR_ZCHUNK,
 * RIHVR_ZCHUNK_ADJ, RIHVR_PROX_ROWS, RIHVR_GRAPHS apply only without one).  Errors as rihvr_create,
 * and RIHVR_E_INVALID_ARGUMENT for a schedule field out of range. */
rihvr_status rihvr_create_scheduled(const rihvr_params* params, const rihvr_schedule* schedule, rihvr_ctx** out);
/* The schedule a context runs with. */
rihvr_status rihvr_get_schedule(const rihvr_ctx* ctx, rihvr_schedule* out);
/* Measures schedules for params by coordinate search: the FFT path and the run-time plans'
 * largest radix, then zchunk_fwd, zchunk_adj, prox_rows (mode 7 with n_tv <= 10) and graphs,
 * each kept when it is faster.  Each candidate is a fresh context solving holo ([H][W] complex64
 * at holo_loc; NULL: A x for a pseudo-random x with about 1 voxel in 1000 non-zero) from the zero
 * start: 12 untimed iterations, then `iterations` timed ones.  The iterations' cost follows the
 * iterates' sparsity (the occupancy maps skip empty lines), so a hologram of the workload times
 * the workload best.  best receives the fastest; tried (capacity entries, may be NULL) the
 * candidates in the order measured, each with its time; *n_tried how many were measured (those
 * beyond capacity are not stored).  Runs on the current device, one context at a time. */
rihvr_status rihvr_autotune(const rihvr_params* params, const float* holo, int holo_loc, int32_t iterations,
                            rihvr_schedule* best, rihvr_schedule* tried, int32_t capacity, int32_t* n_tried);

/* One hologram of a batch: its own parameter set (geometry, optics, weights, iterations and
 * mode), its input and output, and its result. */
typedef struct rihvr_multi_item {
    rihvr_params params;
    const float* holo;         /* [H][W] complex64 at the call's holo_loc */
    float* volume;             /* [Nz][H][W] complex64 at the call's vol_loc, or NULL */
    double f_target;           /* > 0: stop at f <= f_target (momentum point), else run
                                  num_inverse_iterations */
    rihvr_frame_stats stats;   /* out */
    rihvr_schedule schedule;   /* in (autotune == 2, read from a group's first item): the group's
                                  schedule; out: the schedule the item ran with */
} rihvr_multi_item;

/* Reconstructs a batch of holograms of any mix of sizes and parameter sets in one call: items
 * sharing geometry and optics (W, H, Nz, z0, dz, lambda, dx, n_tv, mode) share one context, built
 * once per group and destroyed after it.  The group's schedule: autotune 0, the defaults; 1, the
 * one rihvr_autotune picks on the group's first hologram (10 timed iterations); 2, the schedule
 * field of the group's first item (schedules tuned once and kept by the caller).  Each item is
 * the reference's inverseReconstruct from the zero start with its own
 * weights and iteration count.  Items keep their order within a group; groups run in order of
 * their first item.  Each result equals the single-hologram solve with the same schedule bit
 * for bit. */
rihvr_status rihvr_reconstruct_multi(rihvr_multi_item* items, int32_t n_items, int holo_loc, int vol_loc,
                                     int32_t autotune);

/* SparseCompressiveHolo::estimateLipschitz (src/core/sparse_compressive_holo.cu:2205-2464): the
 * Lipschitz constant of the data term's gradient, ||A^T A||.  The reference's secant estimate
 * ||grad f(x2) - grad f(x1)|| / ||x2 - x1|| over two random volumes is ||A^T A d|| / ||d|| for
 * d = x2 - x1; here d is a pseudo-random volume from seed and each of `iterations` (>= 1) steps
 * applies A^T A to the normalised last result (power iteration), so iterations = 1 is the
 * reference's one secant and more converge to ||A||^2 (= Nz for real holograms, A A^T = Nz I).
 * At least 1e-6.  The reference discards its estimate (the step is 1/Nz, :247); so does the
 * solver here: the call informs and changes no solve.  Invalidates the context's result. */
rihvr_status rihvr_estimate_lipschitz(rihvr_ctx* ctx, int32_t seed, int32_t iterations, double* lipschitz);

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
 * rihvr_set_timing(ctx, 1) resets the counters and starts recording every kernel;
 * enable = 2 + k records only kernel k (the k-th record of rihvr_get_timing), so that
 * one kernel is timed with two events per launch; 0 stops.
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

/* The parameter set the context was created with. */
rihvr_status rihvr_get_params(const rihvr_ctx* ctx, rihvr_params* out);

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

/* Makes the current point of an active solve (the momentum point x1 after the
 * last rihvr_iterate, and its estimated hologram and residual) readable through
 * rihvr_get_volume / rihvr_get_plane / rihvr_get_estimate / rihvr_save_*, without
 * ending the solve; the next rihvr_iterate or rihvr_finish supersedes it.  The
 * reference's per-iteration outputs with output_steps set read x1 there
 * (sparse_compressive_holo.cu:571-597).  RIHVR_E_INVALID_STATE without an active solve. */
rihvr_status rihvr_snapshot(rihvr_ctx* ctx);

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

/* ------------------------------------------------------------------------
 * Host-side input and output of the reconstruction CLI (no device work).
 * ------------------------------------------------------------------------ */

/* The parameter file of the reference CLI: lines "name id value"; lines that start
 * with '%' or '#' are comments (src/core/holo_ui.cpp:171-486).  Defaults are
 * those of HoloUi::HoloUi (holo_ui.cpp:10-62); mode is 7 (main_sparse_inverse.cu:138). */
#define RIHVR_PATH_MAX 4096
typedef struct rihvr_file_params {
    rihvr_params solver;              /* width/height stay 0: they come from the image */
    int32_t image_count;              /* 1001 */
    int32_t start_image;              /* 1002 */
    int32_t background_start;         /* 1003 */
    int32_t background_end;           /* 1004 */
    char    image_filename[RIHVR_PATH_MAX];       /* 2001, else 2002 + 2003; a printf format */
    char    image_path[RIHVR_PATH_MAX];           /* 2002 */
    char    image_formatted_name[RIHVR_PATH_MAX]; /* 2003 */
    char    output_path[RIHVR_PATH_MAX];          /* 2004 */
    char    background_filename[RIHVR_PATH_MAX];  /* 2005 (sets remove_background) */
    int32_t remove_background;
    float   center_x, center_y, center_z; /* 3006, 3007, 3008 */
    int32_t roi_size;                 /* 3009 */
    int32_t roi_x, roi_y, roi_w, roi_h; /* 3011..3014 */
    int32_t output_planes;            /* 7001 */
    int32_t output_steps;             /* 7005 */
    int32_t use_median_background;    /* 7006 (make-background) */
    int32_t segment_min_area;         /* 5009 (segmentation), default 2 */
    double  segment_min_intensity;    /* 5010 (segmentation), default 1 */
    int32_t segment_close_size;       /* 5011 (segmentation), default 3 */
    int32_t unknown_id;               /* id of the line that stopped parsing, else 0 */
} rihvr_file_params;

/* Replaces HoloUi::readFile (holo_ui.cpp:171-486).  As in the reference, parsing
 * stops at the first unknown id and the values read so far are kept: the call then
 * returns RIHVR_E_INVALID_FILE with out->unknown_id set.  A missing file:
 * RIHVR_E_BAD_FILENAME. */
rihvr_status rihvr_read_param_file(const char* path, rihvr_file_params* out);

/* Reads a grayscale image as raw values (Hologram::read, hologram.cu:139-199 via
 * cv::imread): uncompressed baseline TIFF (8/16-bit gray, 8-bit RGB converted with
 * BGR->gray weights, strips, either byte order) or binary PGM (P5, 8/16-bit).
 * Call with pixels == NULL to get the size.  RIHVR_E_BAD_FILENAME if unreadable,
 * RIHVR_E_UNKNOWN_TYPE for an unsupported encoding or a size beyond cv::imread's default
 * limits (width or height above 2^20, more than 2^30 pixels), RIHVR_E_INVALID_FILE for a
 * raster shorter than its header says. */
rihvr_status rihvr_read_image(const char* path, float* pixels, int32_t* width, int32_t* height);

/* The input contract of main_sparse_inverse.cu:101-124: rescale to 0..1 (/255),
 * then I/mean(I) - 1 (normalize_mean0, hologram.cu:332-358), or with a background
 * (I - B) / max(sqrt(B), 0.1) (removeBackground, hologram.cu:1365-1432); crop
 * (ROI rectangle, else centre/size; hologram.cu:768-880) and zero-pad by
 * fp->solver.zero_padding (hologram.cu:882-945).  Output: [H'][W'] complex64 with
 * zero imaginary part; call with out == NULL for the size.  The reference's
 * zeroPad allocates rows/cols swapped (hologram.cu:905), so only square images
 * pass it: a non-square crop returns RIHVR_E_INVALID_ARGUMENT. */
rihvr_status rihvr_prepare_hologram(const rihvr_file_params* fp, const float* image,
                                    const float* background, int32_t width, int32_t height,
                                    float* out, int32_t* out_width, int32_t* out_height);

/* The same input contract on the device for a batch of frames: frames [n][height][width]
 * of raw grey values (pixel_type RIHVR_PIXEL_U8, _U16 or _F32) at in_loc, background
 * [height][width] float32 raw grey values at in_loc when fp->remove_background; out:
 * [n][H'][W'] complex64 at out_loc (call with out == NULL for the size).  Each frame equals
 * rihvr_prepare_hologram's output for it bit for bit for 8-bit frames, whose double sum for the
 * mean is exact in any order; for 16-bit and float frames the device adds the mean's terms in
 * block order and the host in raster order, so a value may differ by the rounding of that sum
 * (at most 1 float ulp).  The errors are that function's,
 * RIHVR_E_UNKNOWN_TYPE for another pixel_type, and RIHVR_E_INVALID_ARGUMENT for an output
 * taller than 65535 rows (the size query still answers).  Runs on the current HIP device. */
#define RIHVR_PIXEL_U8 1
#define RIHVR_PIXEL_U16 2
#define RIHVR_PIXEL_F32 4
rihvr_status rihvr_prepare_holograms(const rihvr_file_params* fp, const void* frames, int32_t pixel_type,
                                     int32_t n_frames, const float* background, int32_t width, int32_t height,
                                     int in_loc, float* out, int out_loc, int32_t* out_width, int32_t* out_height);

/* No counterpart in the reference.  rihvr_prepare_holograms for callers that cannot allocate
 * device memory (code under HIP graph capture, tight frame loops, harnesses that forbid
 * hipMalloc): frames, background and out are device pointers on the current HIP device, and the
 * call allocates nothing.  The mean pass's partial sums and the frames' inverse means live in a
 * 1 MiB scratch array that each GPU holds for the life of the process, which bounds the batch:
 * n_frames * 8 * ceil(width * height / (16384 / bytes per pixel)) bytes of partial sums (none
 * with a background), rounded up to 256, plus 4 * n_frames bytes must fit in 1 MiB (for example
 * 2,000 8-bit or 500 float frames of 1024^2), else RIHVR_E_INVALID_ARGUMENT.  Calls that
 * overlap on one GPU share that scratch and race; the call waits for its work before returning.
 * Results equal rihvr_prepare_holograms' bit for bit; its other errors are that function's. */
rihvr_status rihvr_prepare_holograms_noalloc(const rihvr_file_params* fp, const void* frames, int32_t pixel_type,
                                             int32_t n_frames, const float* background, int32_t width,
                                             int32_t height, float* out, int32_t* out_width, int32_t* out_height);

/* 8-bit grayscale uncompressed TIFF writer (the reference writes with cv::imwrite).  The
 * writers, rihvr_write_hologram and rihvr_prepare_hologram take the sizes rihvr_read_image
 * accepts (width and height up to 2^20, at most 2^30 pixels): RIHVR_E_INVALID_ARGUMENT beyond. */
rihvr_status rihvr_write_tiff8(const char* path, const uint8_t* pixels, int32_t width,
                               int32_t height);

/* Hologram::write (hologram.cu:1587-1652): the real part of a [H][W] complex64
 * hologram as 8-bit; scale_to_minmax != 0 stretches min..max to 0..255 (enh.tif),
 * else 255 * (v - lo) / (hi - lo) (estimate_final.tif, residual_final.tif). */
rihvr_status rihvr_write_hologram(const char* path, const float* holo, int32_t width,
                                  int32_t height, int32_t scale_to_minmax, double lo, double hi);

/* Outputs of the last solve (SparseVolume::saveData, sparse_volume.cu:518-560):
 * "x, y, z, value" rows with value = re^2 + im^2 (%e) for every non-zero voxel,
 * in z, then y*W + x order (the reference's order is that of its atomic compaction). */
rihvr_status rihvr_save_sparse(rihvr_ctx* ctx, const char* csv_path);

/* SparseVolume::saveProjections (sparse_volume.cu:684-813), MAX method: magnitude
 * normalised by its global min/max; xy (planes 1.. only, as the reference), xz and
 * yz maximum projections.  File names as the reference: <output_path><prefix>_xy.tif
 * (no separator is inserted), or with prefix_as_suffix <output_path>xy_<prefix>.tif;
 * likewise _xz and _yz.  A file that cannot be written: RIHVR_E_UNKNOWN_ERROR (the
 * reference's imwrite check). */
rihvr_status rihvr_save_projections(rihvr_ctx* ctx, const char* output_path, const char* prefix,
                                    int32_t prefix_as_suffix);

/* SparseCompressiveHolo::savePlanes (sparse_compressive_holo.cu:1039-1194):
 * <prefix>_%04d.tif (magnitude), _phase_%04d.tif, _real_%04d.tif, _imag_%04d.tif.
 * A file that cannot be written: RIHVR_E_UNKNOWN_ERROR. */
rihvr_status rihvr_save_planes(rihvr_ctx* ctx, const char* path_prefix);

/* 8-bit grayscale PNG writer (stored deflate blocks; the reference writes its
 * background images with cv::imwrite to .png). */
rihvr_status rihvr_write_png8(const char* path, const uint8_t* pixels, int32_t width, int32_t height);

/* ---- Background estimation (the reference's make-background executable,
 * src/core/main_create_background.cu, and HoloSequence::useBackgroundSubtraction,
 * holo_sequence.cu:309-437).  Frames are raw grey values, [n][H][W] float32, at
 * frames_loc (RIHVR_HOST or RIHVR_DEVICE); the work runs on the current HIP device. */
typedef struct rihvr_background_stats {
    double alpha;    /* mean of the frames' means of (I - B) - 5 * mean of their std devs */
    double beta;     /* 255 / (2 * 5 * mean of the std devs) */
    double minimum;  /* the reference's bg_min_grey, always 0 */
    double average;  /* mean grey value of the background B */
} rihvr_background_stats;

/* B = (sum of the frames, added in frame order in float) * float(1/n); then per frame the
 * population mean and standard deviation of I - B (cv::meanStdDev) and the summary
 * statistics above.  n >= 1, else RIHVR_E_MISSING_DATA.  background: [H][W] float32 on
 * the host; stats may be NULL. */
rihvr_status rihvr_background_mean(const float* frames, int32_t n_frames, int32_t width, int32_t height,
                                   int frames_loc, float* background, rihvr_background_stats* stats);

/* Per-pixel median of five frames (median5plane_kernel, main_create_background.cu:13-94);
 * out: [H][W] float32 on the host. */
rihvr_status rihvr_background_median5(const float* frames5, int32_t width, int32_t height, int frames_loc,
                                      float* out);

/* The reference's correlation background (ccbg.cpp computeBackgrounds): for each frame f of
 * frames [n][H][W] (raw grey values, float32 at frames_loc), the other frames ranked by their
 * normalised cross-correlation with f on images reduced by resize_factor (k x k box averages,
 * cv::resize INTER_AREA at resizePct = 1/k; W and H multiples of k); for m = 1 .. min(max_frames,
 * n - 1) the entropy of f minus the mean of its m best frames (256 bins of the rounded value,
 * clamped to 0..255); the first m where the 3-point-smoothed entropy falls, clamped to
 * [min_frames, min(max_frames, n - 1)] (the reference reads past its ranking when min_frames
 * >= n; here the count stops at n - 1); the background is the mean of that many full-resolution
 * frames as 8 bits (convertTo: rounded to even, saturated).  backgrounds: [n][H][W] on the host;
 * enhanced (may be NULL): normalize(f - background, 0..255, NORM_MINMAX) as 8 bits; frames_used
 * (may be NULL): [n].  The correlations are summed in double (the reference in float), so frames
 * whose correlations agree to float rounding may rank in another order; equal correlations rank
 * by frame number.  n < 2: RIHVR_E_MISSING_DATA. */
rihvr_status rihvr_background_correlation(const float* frames, int32_t n_frames, int32_t width, int32_t height,
                                          int frames_loc, int32_t min_frames, int32_t max_frames,
                                          int32_t resize_factor, uint8_t* backgrounds, uint8_t* enhanced,
                                          int32_t* frames_used);

/* ---- Segmentation and object centroids (the reference's segmentation executable,
 * src/core/main_segmentation.cu:83-125): SparseSegmentation (sparse_segmentation.cu)
 * binarize -> close -> labelConnectedComponents, then ObjectCloud (object_cloud.cu)
 * with Blob3d statistics (blob3d.cu) and blobSizeFilter.  A handle holds one volume of
 * W x H x Nz voxels (W*H*Nz < 2^32 - 1) in dense device buffers; the dense passes run
 * on the current HIP device on the handle's own stream; the object statistics, whose
 * float sums the reference adds in voxel order, are one pass over the compacted
 * object voxels on the host.  The call order is load -> binarize -> [close] -> label ->
 * extract -> [size_filter]; a call out of that order returns RIHVR_E_INVALID_STATE. */
typedef struct rihvr_segmentation rihvr_segmentation;

typedef struct rihvr_object {
    float   centroid[3];           /* x, y, z: integer sums / n in float (Blob3d::getCentroid, blob3d.cu:198-226) */
    float   weighted_centroid[3];  /* float sums of w*x in voxel order / float sum of w (:228-265) */
    float   max_value;             /* largest voxel value (getMaxValue, :320-324) */
    float   mean_value;            /* double sum cast to float / float n (getMeanValue, :326-330) */
    int64_t num_voxels;
    int64_t label;                 /* smallest global voxel index z*W*H + y*W + x of the object */
    int32_t bounds[6];             /* xmin, xmax, ymin, ymax, zmin, zmax (findBounds, :79-105) */
} rihvr_object;

/* SparseVolume::loadData (sparse_volume.cu:562-682): "x, y, z, value" rows of a sparse
 * volume file (the first 14 characters must be that header, else RIHVR_E_INVALID_FILE).
 * As the reference, plane z takes the next count[z] rows of the file, count[z] being the
 * number of rows whose z is z; index = y*W + x, and the reference's bounds test x > W,
 * y > H, z > depth is RIHVR_E_INVALID_DATA (x == W wraps to the next row).  Rows the
 * reference would store out of range (negative, index >= W*H, z >= Nz) and malformed rows:
 * RIHVR_E_INVALID_DATA / RIHVR_E_INVALID_FILE.  Call with plane == NULL to get *count;
 * then plane, index (int32) and intensity (float, the file's value) hold *count entries.
 * A missing file: RIHVR_E_BAD_FILENAME. */
rihvr_status rihvr_read_sparse_voxels(const char* path, int32_t width, int32_t height, int32_t num_planes,
                                      int64_t* count, int32_t* plane, int32_t* index, float* intensity);

/* SparseSegmentation's state; create allocates the dense buffers (about 12.3 bytes per
 * voxel: u32 values, u32 labels, u32 work, two bit-packed binary volumes).  RIHVR_E_INVALID_ARGUMENT for a non-positive size or W*H*Nz >= 2^32 - 1. */
rihvr_status rihvr_segmentation_create(int32_t width, int32_t height, int32_t num_planes,
                                       rihvr_segmentation** out);
void rihvr_segmentation_destroy(rihvr_segmentation* seg);

/* SparseSegmentation(SparseVolume) after loadData: the voxel list (plane, in-plane index,
 * intensity) at loc; the volume's value is sqrt(intensity) as loadData stores it.  A voxel
 * listed twice keeps its larger intensity (NaN ranking lowest).  An entry outside the
 * volume: RIHVR_E_INVALID_DATA. */
rihvr_status rihvr_segmentation_load(rihvr_segmentation* seg, int64_t count, const int32_t* plane,
                                     const int32_t* index, const float* intensity, int loc);

/* The same from a dense [Nz][H][W] complex64 volume at loc (for example a solver's
 * rihvr_get_volume output): intensity = re^2 + im^2 in float for every non-zero voxel, as
 * rihvr_save_sparse writes it but without the file's 7-digit rounding. */
rihvr_status rihvr_segmentation_load_dense(rihvr_segmentation* seg, const float* volume, int loc);

/* SparseSegmentation::binarize (sparse_segmentation.cu:63-111): max intensity over the
 * loaded voxels, thr = min_intensity * max / 256 (both returned; either pointer may be
 * NULL); voxels with intensity > thr form the binary volume and keep their intensity as
 * value, the others get value 0. */
rihvr_status rihvr_segmentation_binarize(rihvr_segmentation* seg, double min_intensity, double* max_intensity,
                                         double* threshold);

/* SparseSegmentation::close (:145-315): dilate then erode with the ball
 * sx^2 + sy^2 + sz^2 <= close_size^2; the erosion keeps a voxel only where the whole ball
 * lies inside the volume and inside the dilation.  close_size <= 0 leaves the volume as it
 * is; close_size > 63: RIHVR_E_INVALID_ARGUMENT (the ball's x reach is kept within one
 * 64-voxel word of the packed binary volume). */
rihvr_status rihvr_segmentation_close(rihvr_segmentation* seg, int32_t close_size);

/* labelConnectedComponents (:411-532): 26-neighbour components where the in-plane
 * neighbours are the flat indices i + dy*W + dx inside 0..W*H-1 (so rows wrap at the x
 * borders, as the reference's offsets do); every voxel's label is the smallest global
 * index of its component. */
rihvr_status rihvr_segmentation_label(rihvr_segmentation* seg);

/* SparseSegmentation::extractObjects (:731-739): ObjectCloud(SparseVolume) and
 * includeIntensity (object_cloud.cu:44-58, 127-289): objects in order of first appearance
 * in the z-major raster, each voxel's value its binarized intensity.  No cap on the
 * number of objects (the reference's 20*1024 cap, object_cloud.h:22, overruns its
 * arrays). */
rihvr_status rihvr_segmentation_extract(rihvr_segmentation* seg, int64_t* num_objects);

/* ObjectCloud::blobSizeFilter (object_cloud.cu:313-328): objects with fewer than min_area
 * voxels are replaced by the last object, which is tested next. */
rihvr_status rihvr_segmentation_size_filter(rihvr_segmentation* seg, int32_t min_area, int64_t* num_removed);

/* The current objects: up to capacity into out (may be NULL); *num_objects = their count. */
rihvr_status rihvr_segmentation_objects(rihvr_segmentation* seg, rihvr_object* out, int64_t capacity,
                                        int64_t* num_objects);

/* ObjectCloud::writeCentroids (object_cloud.cu:408-470): weighted != 0 writes the weighted
 * centroids, verbose != 0 the long format (both centroids, count, max, mean, bounds).
 * A file that cannot be opened: RIHVR_E_BAD_FILENAME. */
rihvr_status rihvr_segmentation_write_centroids(rihvr_segmentation* seg, const char* path, int32_t weighted,
                                                int32_t verbose);

/* Copies of the dense state ([Nz][H][W], to loc): the current binary volume (after
 * binarize or close), the voxel values (after binarize) and the labels (after label;
 * 0xFFFFFFFF off the objects). */
rihvr_status rihvr_segmentation_mask(rihvr_segmentation* seg, uint8_t* out, int loc);
rihvr_status rihvr_segmentation_values(rihvr_segmentation* seg, float* out, int loc);
rihvr_status rihvr_segmentation_labels(rihvr_segmentation* seg, uint32_t* out, int loc);

/* Device time (ms, HIP events on the handle's stream) of the last load, binarize, close,
 * label and extract calls, in that order: ms[5]. */
rihvr_status rihvr_segmentation_timing(rihvr_segmentation* seg, double* ms);

/* ---- Learned warm start.  No counterpart in the reference, whose solver always starts from
 * zero (sparse_compressive_holo.cu:258-259, sparse_volume.cu:44-48).  A 3D encoder-decoder
 * network maps the back-propagated volume A^T b of a hologram (SparseGradient::reconstruct,
 * sparse_gradient.cu:47-165, with r = b) to a start volume for FISTA; it is trained on the
 * solver's own converged volumes.  Architecture (fixed): 4 levels of 16, 32, 64, 128 channels,
 * one 3x3x3 convolution + ReLU per level on the way down (2x2x2 average pooling between levels)
 * and on the way up (nearest 2x2x2 upsampling concatenated with the skip), a 1x1x1 head to
 * (Re, Im).  Input scaled by in_scale, output by out_scale.  Activations are BF16, convolutions
 * run on MFMA with FP32 accumulation, master weights and Adam state are FP32.  Volumes are
 * [Nz][H][W] complex64 on the current HIP device; Nz, H, W must be multiples of 8. */
typedef struct rihvr_net rihvr_net;

/* Number of FP32 parameters: per convolution e0, e1, e2, e3, d2, d1, d0 the weights
 * [27][Cout][Cin] (tap = (dz+1)*9 + (dy+1)*3 + (dx+1); e0's Cin is 16 of which channels 0, 1
 * are Re, Im) then the bias [Cout]; then the head's [2][16] weights and [2] bias. */
rihvr_status rihvr_net_param_count(int64_t* count);
/* Allocates a network's inference buffers for Nz x H x W volumes (about 40 bytes per voxel;
 * training buffers, about 3 times that, are added by the first gradient call).  Parameters
 * start at zero: call rihvr_net_init or rihvr_net_set_params. */
rihvr_status rihvr_net_create(int32_t num_planes, int32_t height, int32_t width, rihvr_net** out);
rihvr_status rihvr_net_destroy(rihvr_net* net);
/* He-normal weights from a seeded generator, zero biases; resets Adam. */
rihvr_status rihvr_net_init(rihvr_net* net, uint64_t seed);
/* Host FP32 parameters (rihvr_net_param_count of them) and the input/output scales; resets Adam. */
rihvr_status rihvr_net_set_params(rihvr_net* net, const float* params, float in_scale, float out_scale);
rihvr_status rihvr_net_get_params(rihvr_net* net, float* params, float* in_scale, float* out_scale);
/* volume_out = the network's start volume for volume_in (device pointers). */
rihvr_status rihvr_net_predict(rihvr_net* net, const float* volume_in, float* volume_out);
/* Loss weight: voxels where the target is non-zero count 1 + pos_weight times (default 0: plain
 * mean squared error).  The solver's volumes are sparse (a few voxels in 10^4 are non-zero), so
 * the plain error is minimised by predicting zeros; the weight keeps the objects in the loss. */
rihvr_status rihvr_net_set_loss_weight(rihvr_net* net, float pos_weight);
/* Loss sum_v w_v |out_v - target_v|^2 / (2 V out_scale^2) (w_v = 1 + pos_weight where
 * target_v != 0, else 1) and its gradient with respect to every parameter (kept for
 * rihvr_net_adam_step; copied to the host array grads when not NULL). */
rihvr_status rihvr_net_loss_grad(rihvr_net* net, const float* volume_in, const float* target, double* loss,
                                 float* grads);
/* One Adam step (beta1 0.9, beta2 0.999, eps 1e-8) with the gradient of the last loss_grad. */
rihvr_status rihvr_net_adam_step(rihvr_net* net, float lr);
/* rihvr_net_loss_grad then rihvr_net_adam_step. */
rihvr_status rihvr_net_train_step(rihvr_net* net, const float* volume_in, const float* target, float lr,
                                  double* loss);
/* Waits for the network's queued work. */
rihvr_status rihvr_net_sync(rihvr_net* net);

/* The network's convolution kernels on device tensors (BF16 as uint16 bit patterns,
 * channels-last [D][H][W][C], channel counts multiples of 16, zero padding):
 *   forward: out[v][co] = act(bias[co] + sum_tap sum_ci in[v + tap][ci] w[tap][co][ci]), where the
 *            input is in0 (c0 channels) followed by in1 (c1 channels at half resolution, read at
 *            (z/2, y/2, x/2); NULL when c1 = 0); act = ReLU when relu != 0; bias may be NULL;
 *   dgrad:   grad_in[v][ci] = sum_tap sum_co grad_out[v - tap][co] w[tap][co][ci], with weight_t the
 *            flipped transpose wT[26 - tap][ci][co] (FP32 out);
 *   wgrad:   grad_weight_bias = dW[tap][co][ci] = sum_v grad_out[v][co] in[v + tap][ci], then
 *            db[co] = sum_v grad_out[v][co] (FP32, 27*cout*cin + cout values). */
rihvr_status rihvr_conv3d_bf16_forward(const uint16_t* in0, int32_t c0, const uint16_t* in1, int32_t c1, int32_t depth,
                                       int32_t height, int32_t width, const uint16_t* weight, const float* bias,
                                       int32_t cout, int32_t relu, uint16_t* out);
rihvr_status rihvr_conv3d_bf16_dgrad(const uint16_t* grad_out, int32_t cout, int32_t depth, int32_t height,
                                     int32_t width, const uint16_t* weight_t, int32_t cin, float* grad_in);
rihvr_status rihvr_conv3d_bf16_wgrad(const uint16_t* in0, int32_t c0, const uint16_t* in1, int32_t c1, int32_t depth,
                                     int32_t height, int32_t width, const uint16_t* grad_out, int32_t cout,
                                     float* grad_weight_bias);

/* Result of rihvr_reconstruct_learned. */
typedef struct rihvr_learned_stats {
    int32_t iterations;     /* FISTA iterations run from the chosen start */
    int32_t fallback;       /* 0: the network's start soft(x_net, tau), 1: the zero start */
    int32_t reached;        /* 1: the objective target was reached */
    int32_t total_backtracks;
    double  resid_net;      /* NaN: 0.5 ||A x_net - b||^2 of the raw prediction is not evaluated */
    double  resid_adjoint;  /* 0.5 ||A (t A^T b) - b||^2 of the A^T b start, t = 1/Nz (zero for a
                               real hologram: A A^T = Nz I on real holograms) */
    double  f_net;          /* NaN: the raw prediction's objective is not evaluated (the start test
                               needs only the thresholded candidates') */
    double  f_adjoint;      /* objective at the A^T b start */
    double  f_zero;         /* objective at the zero start, 0.5 ||b||^2 */
    double  f_start;        /* objective at the chosen start */
    double  f_final;        /* objective at the returned volume */
    double  ms_adjoint;     /* wall time of A^T b */
    double  ms_network;     /* wall time of the network's prediction, the threshold choice and the
                               start test */
    double  ms_solve;       /* wall time of FISTA from the chosen start */
    double  tau;            /* the soft threshold applied to x_net: q max|x_net|, q in {0.02, 0.05},
                               the one whose start has the least objective */
    double  f_thresholded;  /* objective at soft(x_net, tau) */
} rihvr_learned_stats;

/* Reconstructs one hologram with the network's start: A^T b -> network -> x_net ->
 * soft(x_net, tau) (complex soft threshold; tau = q max|x_net| for the q in {0.02, 0.05}
 * whose start has the least objective: the network's output holds small values at every voxel,
 * which FISTA's first prox would remove).  That start is used when its objective is not above
 * the objectives at the A^T b start (t A^T b, t = 1/Nz) and at the zero start; else the zero
 * start of the reference (fallback).  Then FISTA (the context's mode and parameters)
 * until the objective at the momentum point is <= f_target (f_target <= 0: no target) or
 * max_iterations (<= 0: num_inverse_iterations).  The network must have been created for the
 * context's Nz x H x W.  The result is read as after rihvr_reconstruct. */
rihvr_status rihvr_reconstruct_learned(rihvr_ctx* ctx, rihvr_net* net, const float* holo, int holo_loc,
                                       double f_target, int32_t max_iterations, rihvr_learned_stats* stats);

#ifdef __cplusplus
}
#endif

#endif /* RIHVR_H */
