/*
 * rihvr_oracle.h -- CPU restatement of the reference's sparse inverse
 * reconstruction, used ONLY as a test oracle and as the CPU baseline leg of
 * bench.py.  Nothing in the product library links or calls it.
 *
 * Two instantiations of one template:
 *   *_f32 : the reference's FP32 semantics -- float storage, phase arguments
 *           computed in double and rounded to float before sin/cos (the
 *           reference's sincosf(float(double)) calls), double arithmetic exactly
 *           where the reference kernels promote to double, float reduction trees
 *           with the reference's summation order and host accumulation types.
 *           sin/cos of the float argument are accurate (computed in double),
 *           standing in for the fast-math hardware approximation the
 *           reference is built with (src/core/Makefile:24-30).
 *   *_f64 : the same operator and solver in double throughout, phase arguments
 *           not rounded, reductions in double: the FP64 host oracle.
 *
 * Complex arrays are interleaved (re, im) of float (f32) or double (f64).
 */
#ifndef RIHVR_ORACLE_H
#define RIHVR_ORACLE_H

#include "../include/rihvr.h"

#ifdef __cplusplus
extern "C" {
#endif

int  ro_version(void);
void ro_set_threads(int n);   /* OpenMP threads; <= 0 keeps the default */
int  ro_get_threads(void);

/* 2D FFT of one [H][W] plane, in place.  sign = -1 forward, +1 inverse,
 * unnormalised (cuFFT convention). */
int ro_fft2_f32(float* data, int H, int W, int sign);
int ro_fft2_f64(double* data, int H, int W, int sign);

/* kappa(x, y) table [H][W] (sparse_rs_exponent_kernel). */
int ro_kappa(const rihvr_params* p, double* out);
/* plane positions z_i as the reference computes them (float). */
int ro_plane_list(const rihvr_params* p, float* out);

/* A: volume [Nz][H][W] -> holo [H][W] (imag 0). */
int ro_forward_f32(const rihvr_params* p, const float* vol, float* holo);
int ro_forward_f64(const rihvr_params* p, const double* vol, double* holo);

/* A^T for all planes: holo [H][W] -> volume [Nz][H][W]. */
int ro_adjoint_f32(const rihvr_params* p, const float* holo, float* vol);
int ro_adjoint_f64(const rihvr_params* p, const double* holo, double* vol);

/* One prox call of p->mode on every plane: out = prox(x0 - t*grad). */
int ro_prox_f32(const rihvr_params* p, const float* x0, const float* grad,
                double t, float* out);
int ro_prox_f64(const rihvr_params* p, const double* x0, const double* grad,
                double t, double* out);

/* Reference reductions (FP32 tree semantics for f32). */
double ro_summag_f32(const float* data, long n);         /* sum re^2+im^2 */
double ro_l1_f32(const float* vol, long n);              /* sum |v| (modulus) */
double ro_tv2d_f32(const float* vol, int Nz, int H, int W);
/* calcTVNorm (sparse_volume.cu:303-363): 3D, z-periodic */
double ro_tv3d_f32(const float* vol, int Nz, int H, int W);
double ro_restart_sum_f32(const float* x0, const float* x1, const float* y0, long n_plane, int Nz);

/* Full solve.  warm may be NULL.  out: [Nz][H][W].  estimate/residual: [H][W]
 * outputs (may be NULL).  Returns an rihvr_status value. */
int ro_solve_f32(const rihvr_params* p, const float* holo, const float* warm,
                 float* out, float* estimate, float* residual, rihvr_stats* st);
int ro_solve_f64(const rihvr_params* p, const double* holo, const double* warm,
                 double* out, double* estimate, double* residual, rihvr_stats* st);

#ifdef __cplusplus
}
#endif
#endif
