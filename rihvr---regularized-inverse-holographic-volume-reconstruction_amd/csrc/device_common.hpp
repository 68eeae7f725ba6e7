// device_common.hpp -- device helpers shared by the gfx950 kernels.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace rihvr {

__device__ __forceinline__ float2 cadd(float2 a, float2 b) { return make_float2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ float2 csub(float2 a, float2 b) { return make_float2(a.x - b.x, a.y - b.y); }
__device__ __forceinline__ float2 cmul(float2 a, float2 b) {
    return make_float2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__device__ __forceinline__ float2 cscale(float2 a, float s) { return make_float2(a.x * s, a.y * s); }

// sin and cos of a float argument, accurate to about 1 ulp for |a| < 2^20.
// The reference feeds sincosf() with float-rounded phase arguments of up to
// ~5e4 rad (sparse_compressive_holo.cu:1872-1879); only the float rounding of
// the argument is part of its arithmetic, so this computes sin/cos of exactly
// that float value.  Range reduction: Cody-Waite in double with a two-part
// pi/2 (exact to ~1e-16 rad for these magnitudes); kernels: the single
// precision minimax polynomials of the Cephes library (S. L. Moshier, sinf.c /
// cosf.c) on |r| <= pi/4.
__device__ __forceinline__ void sincos_accurate(float a, float* s, float* c) {
    const double kTwoOverPi = 0.63661977236758134308;
    const double kPio2Hi = 1.5707963267948966192e+00;
    const double kPio2Lo = 6.1232339957367660359e-17;
    double ad = (double)a;
    double q = rint(ad * kTwoOverPi);
    double r = fma(-q, kPio2Hi, ad);
    r = fma(-q, kPio2Lo, r);
    float x = (float)r;
    int iq = (int)q;
    float z = x * x;
    float sp = ((-1.9515295891e-4f * z + 8.3321608736e-3f) * z - 1.6666654611e-1f) * z * x + x;
    float cp = ((2.443315711809948e-5f * z - 1.388731625493765e-3f) * z + 4.166664568298827e-2f) * z * z
               - 0.5f * z + 1.0f;
    int j = iq & 3;
    float sv = (j & 1) ? cp : sp;
    float cv = (j & 1) ? sp : cp;
    if (j == 1 || j == 2) cv = -cv;
    if (j >= 2) sv = -sv;
    *s = sv;
    *c = cv;
}

// sin and cos of a float argument through the hardware units: the argument is
// reduced to revolutions in double (exact to ~1e-12 rev for |a| < 2^20), rounded
// to float in [-0.5, 0.5] and fed to v_sin_f32 / v_cos_f32 (which take
// revolutions).  Absolute error ~1e-6; used where that is far below the
// quantity's other error sources.
__device__ __forceinline__ void sincos_hw(float a, float* s, float* c) {
    const double kInvTwoPi = 0.15915494309189533577;
    double rev = (double)a * kInvTwoPi;
    rev -= rint(rev);
    const float r = (float)rev;
    *s = __builtin_amdgcn_sinf(r);
    *c = __builtin_amdgcn_cosf(r);
}

// e^{i * float(z * kappa)}: the reference's transfer-function factor H_z
// (sparse_rs_phase_mult_kernel, sparse_compressive_holo.cu:1863-1888; and
// sparse_rs_phase_mult_kernel2, sparse_gradient.cu:77-101): the double product
// is rounded to float, then sin/cos are taken.
#ifndef RIHVR_TRANSFER_SINCOS
#define RIHVR_TRANSFER_SINCOS sincos_hw
#endif
__device__ __forceinline__ float2 transfer(double z, double kappa) {
    float arg = (float)(z * kappa);
    float s, c;
    RIHVR_TRANSFER_SINCOS(arg, &s, &c);
    return make_float2(c, s);
}

// --- wave64 / workgroup reductions in double ---------------------------------
__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}

// Sums `v` over the workgroup (blockDim.x threads, a multiple of 64); the
// result is valid in thread 0.  `scratch` must hold blockDim.x/64 doubles.
__device__ __forceinline__ double block_sum(double v, double* scratch) {
    v = wave_sum(v);
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    __syncthreads();
    if (lane == 0) scratch[wid] = v;
    __syncthreads();
    double r = 0.0;
    if (threadIdx.x == 0) {
        const int nw = (blockDim.x + 63) >> 6;
        for (int i = 0; i < nw; ++i) r += scratch[i];
    }
    return r;
}

}  // namespace rihvr
