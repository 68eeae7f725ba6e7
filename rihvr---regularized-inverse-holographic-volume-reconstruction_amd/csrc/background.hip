// background.hip -- background estimation of the reference's make-background
// executable on the GPU: the mean background with the frame statistics of
// HoloSequence::useBackgroundSubtraction (src/core/holo_sequence.cu:309-437) and
// the median of five frames (median5plane_kernel, main_create_background.cu:13-94).
// Elementwise and reduction work, HBM-bound: one pass over the frames for the
// sum, one for the statistics.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <vector>

#include "../../include/rihvr.h"
#include "device_common.hpp"

namespace {

constexpr int kThreads = 256;
constexpr int kItems = 32;  // pixels per thread in the moment pass
constexpr int kBatch = 8;   // frame loads issued before their adds in the sum pass

// B = (f_0 + f_1 + ... + f_{n-1}) * scale, summed in frame order in float as the
// reference's running CuMat sum does; scale = float(1/n) as cv::Mat /= n.  V = float4
// takes four pixels per thread (N % 4 == 0, 16-byte aligned frames), V = float one.
__device__ __forceinline__ float vadd(float a, float b) { return a + b; }
__device__ __forceinline__ float4 vadd(float4 a, float4 b) {
    return make_float4(a.x + b.x, a.y + b.y, a.z + b.z, a.w + b.w);
}
__device__ __forceinline__ float vscale(float a, float s) { return a * s; }
__device__ __forceinline__ float4 vscale(float4 a, float s) { return make_float4(a.x * s, a.y * s, a.z * s, a.w * s); }

template <typename V>
__global__ void __launch_bounds__(kThreads) k_bg_sum(const V* __restrict__ frames, int n, long NV, float scale,
                                                     V* __restrict__ bg) {
    const long i = (long)blockIdx.x * kThreads + threadIdx.x;
    if (i >= NV) return;
    V s = frames[i];
    int k = 1;
    for (; k + kBatch <= n; k += kBatch) {
        V v[kBatch];
#pragma unroll
        for (int j = 0; j < kBatch; ++j) v[j] = frames[(size_t)(k + j) * NV + i];
#pragma unroll
        for (int j = 0; j < kBatch; ++j) s = vadd(s, v[j]);
    }
    for (; k < n; ++k) s = vadd(s, frames[(size_t)k * NV + i]);
    bg[i] = vscale(s, scale);
}

// Per frame (k0 + blockIdx.y) and block: sum of d and d^2 over its pixels, d = f - B in
// float (the Mat subtraction), accumulated in double (cv::meanStdDev).  Block 0 of
// frame row n (k == n) sums B itself for cv::mean(B).  V as in k_bg_sum.
__device__ __forceinline__ void moments_add(float f, float b, double& s1, double& s2) {
    const float d = f - b;
    s1 += (double)d;
    s2 += (double)d * (double)d;
}
__device__ __forceinline__ void moments_add(float4 f, float4 b, double& s1, double& s2) {
    moments_add(f.x, b.x, s1, s2);
    moments_add(f.y, b.y, s1, s2);
    moments_add(f.z, b.z, s1, s2);
    moments_add(f.w, b.w, s1, s2);
}
__device__ __forceinline__ double pixel_sum(float b) { return (double)b; }
__device__ __forceinline__ double pixel_sum(float4 b) {
    return (((double)b.x + (double)b.y) + (double)b.z) + (double)b.w;
}

template <typename V>
__global__ void __launch_bounds__(kThreads) k_bg_moments(const V* __restrict__ frames, int n, int k0, long NV,
                                                         const V* __restrict__ bg, double* __restrict__ partial) {
    constexpr int items = kItems / (sizeof(V) / sizeof(float));
    __shared__ double red[kThreads / 64];
    const int k = k0 + blockIdx.y;
    double s1 = 0.0, s2 = 0.0;
    for (int it = 0; it < items; ++it) {
        const long i = ((long)blockIdx.x * items + it) * kThreads + threadIdx.x;
        if (i >= NV) break;
        if (k < n) {
            moments_add(frames[(size_t)k * NV + i], bg[i], s1, s2);
        } else {
            s1 += pixel_sum(bg[i]);
        }
    }
    const double a = rihvr::block_sum(s1, red);
    const double b = rihvr::block_sum(s2, red);
    if (threadIdx.x == 0) {
        double* p = partial + ((size_t)k * gridDim.x + blockIdx.x) * 2;
        p[0] = a;
        p[1] = b;
    }
}

// Median of five by a 9-comparator network (the same comparisons as the reference's
// kernel; the median does not depend on the network).
__device__ __forceinline__ void cswap(float& a, float& b) {
    const float lo = fminf(a, b), hi = fmaxf(a, b);
    a = lo;
    b = hi;
}
__global__ void __launch_bounds__(kThreads) k_median5(const float* __restrict__ f, long N, float* __restrict__ out) {
    const long i = (long)blockIdx.x * kThreads + threadIdx.x;
    if (i >= N) return;
    float a0 = f[i], a1 = f[N + i], a2 = f[2 * N + i], a3 = f[3 * N + i], a4 = f[4 * N + i];
    cswap(a0, a1);
    cswap(a3, a4);
    cswap(a2, a4);
    cswap(a2, a3);
    cswap(a1, a4);
    cswap(a0, a3);
    cswap(a0, a2);
    cswap(a1, a3);
    cswap(a1, a2);
    out[i] = a2;
}

struct DevBuf {
    void* p = nullptr;
    ~DevBuf() {
        if (p) (void)hipFree(p);
    }
};

#define BG_HIP(call)                                                                               \
    do {                                                                                           \
        hipError_t e_ = (call);                                                                    \
        if (e_ != hipSuccess) {                                                                    \
            std::fprintf(stderr, "rihvr: HIP error %s at %s:%d\n", hipGetErrorString(e_), __FILE__, \
                         __LINE__);                                                                \
            return e_ == hipErrorOutOfMemory ? RIHVR_E_OUT_OF_MEMORY : RIHVR_E_DEVICE;              \
        }                                                                                          \
    } while (0)

// frames on the device: the caller's pointer, or a copy of the host array
rihvr_status device_frames(const float* frames, size_t count, int loc, DevBuf& hold, const float** out) {
    if (loc == RIHVR_DEVICE) {
        *out = frames;
        return RIHVR_OK;
    }
    BG_HIP(hipMalloc(&hold.p, count * sizeof(float)));
    BG_HIP(hipMemcpy(hold.p, frames, count * sizeof(float), hipMemcpyHostToDevice));
    *out = (const float*)hold.p;
    return RIHVR_OK;
}

}  // namespace

extern "C" {

rihvr_status rihvr_background_mean(const float* frames, int32_t n_frames, int32_t width, int32_t height,
                                   int frames_loc, float* background, rihvr_background_stats* stats) {
    if (!frames || !background || width <= 0 || height <= 0) return RIHVR_E_INVALID_ARGUMENT;
    if (n_frames < 1) return RIHVR_E_MISSING_DATA;
    const long N = (long)width * height;
    DevBuf hf, hb, hp;
    const float* df = nullptr;
    rihvr_status st = device_frames(frames, (size_t)n_frames * N, frames_loc, hf, &df);
    if (st) return st;
    BG_HIP(hipMalloc(&hb.p, N * sizeof(float)));
    float* dbg = (float*)hb.p;
    // four pixels per thread when every frame starts on a 16-byte boundary (hipMalloc'd B does)
    const bool wide = N % 4 == 0 && (size_t)df % 16 == 0;
    const long NV = wide ? N / 4 : N;
    const float scale = (float)(1.0 / n_frames);
    const unsigned g1 = (unsigned)((NV + kThreads - 1) / kThreads);
    if (wide)
        hipLaunchKernelGGL(k_bg_sum<float4>, dim3(g1), dim3(kThreads), 0, 0, (const float4*)df, n_frames, NV, scale,
                           (float4*)dbg);
    else
        hipLaunchKernelGGL(k_bg_sum<float>, dim3(g1), dim3(kThreads), 0, 0, df, n_frames, NV, scale, dbg);
    BG_HIP(hipGetLastError());
    const unsigned g2 = (unsigned)((N + (long)kThreads * kItems - 1) / ((long)kThreads * kItems));
    const size_t np = (size_t)(n_frames + 1) * g2 * 2;
    BG_HIP(hipMalloc(&hp.p, np * sizeof(double)));
    // frame rows 0..n (row n: B itself) in launches of at most 65535 (the grid's y limit)
    for (int k0 = 0; k0 <= n_frames; k0 += 65535) {
        const dim3 grid(g2, (unsigned)std::min(n_frames + 1 - k0, 65535));
        if (wide)
            hipLaunchKernelGGL(k_bg_moments<float4>, grid, dim3(kThreads), 0, 0, (const float4*)df, n_frames, k0, NV,
                               (const float4*)dbg, (double*)hp.p);
        else
            hipLaunchKernelGGL(k_bg_moments<float>, grid, dim3(kThreads), 0, 0, df, n_frames, k0, NV, dbg,
                               (double*)hp.p);
    }
    BG_HIP(hipGetLastError());
    std::vector<double> part(np);
    BG_HIP(hipMemcpy(part.data(), hp.p, np * sizeof(double), hipMemcpyDeviceToHost));
    BG_HIP(hipMemcpy(background, dbg, N * sizeof(float), hipMemcpyDeviceToHost));
    if (stats) {
        // cv::meanStdDev per frame: mean = s/N, std = sqrt(max(sq/N - mean^2, 0));
        // then cv::mean over the frames of each (holo_sequence.cu:394-415)
        double sum_mean = 0.0, sum_std = 0.0;
        for (int k = 0; k < n_frames; ++k) {
            double s1 = 0.0, s2 = 0.0;
            for (unsigned b = 0; b < g2; ++b) {
                s1 += part[((size_t)k * g2 + b) * 2];
                s2 += part[((size_t)k * g2 + b) * 2 + 1];
            }
            const double m = s1 / (double)N;
            const double var = s2 / (double)N - m * m;
            sum_mean += m;
            sum_std += std::sqrt(var > 0.0 ? var : 0.0);
        }
        double sb = 0.0;
        for (unsigned b = 0; b < g2; ++b) sb += part[((size_t)n_frames * g2 + b) * 2];
        const double mean_mean = sum_mean / n_frames, mean_std = sum_std / n_frames;
        const double scale1 = 5.0;  // bg_scale1 (holo_sequence.cu:22)
        stats->alpha = mean_mean - scale1 * mean_std;
        stats->beta = 255 / (2 * mean_std * scale1);
        stats->minimum = 0.0;  // bg_min_grey (holo_sequence.cu:23)
        stats->average = sb / (double)N;
    }
    return RIHVR_OK;
}

rihvr_status rihvr_background_median5(const float* frames5, int32_t width, int32_t height, int frames_loc,
                                      float* out) {
    if (!frames5 || !out || width <= 0 || height <= 0) return RIHVR_E_INVALID_ARGUMENT;
    const long N = (long)width * height;
    DevBuf hf, ho;
    const float* df = nullptr;
    rihvr_status st = device_frames(frames5, 5 * (size_t)N, frames_loc, hf, &df);
    if (st) return st;
    BG_HIP(hipMalloc(&ho.p, N * sizeof(float)));
    hipLaunchKernelGGL(k_median5, dim3((unsigned)((N + kThreads - 1) / kThreads)), dim3(kThreads), 0, 0, df, N,
                       (float*)ho.p);
    BG_HIP(hipGetLastError());
    BG_HIP(hipMemcpy(out, ho.p, N * sizeof(float), hipMemcpyDeviceToHost));
    return RIHVR_OK;
}

}  // extern "C"
