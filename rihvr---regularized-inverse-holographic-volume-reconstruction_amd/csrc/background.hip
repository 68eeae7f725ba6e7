// background.hip -- background estimation of the reference's make-background
// executable on the GPU: the mean background with the frame statistics of
// HoloSequence::useBackgroundSubtraction (src/core/holo_sequence.cu:309-437) and
// the median of five frames (median5plane_kernel, main_create_background.cu:13-94).
// Elementwise and reduction work, HBM-bound: one pass over the frames for the
// sum, one for the statistics.  And the reference's correlation background (ccbg.cpp):
// per frame, the mean of its best-correlated frames, as many as an entropy rule picks.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <numeric>
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

// ---- correlation background (ccbg.cpp) -------------------------------------------
// Box average of k x k pixels (cv::resize INTER_AREA at an integer factor): the block's values
// added row by row in float, times float(1 / k^2).  out: [n][H/k][W/k]
__global__ void k_cc_resize(const float* __restrict__ f, int W, int H, int k, long n_out, float* __restrict__ out) {
    const long i = (long)blockIdx.x * kThreads + threadIdx.x;
    if (i >= n_out) return;
    const int w = W / k, h = H / k;
    const int x = (int)(i % w), y = (int)((i / w) % h);
    const long fr = i / ((long)w * h);
    const float* src = f + (size_t)fr * W * H + (size_t)(y * k) * W + x * k;
    float s = 0.0f;
    for (int dy = 0; dy < k; ++dy)
        for (int dx = 0; dx < k; ++dx) s += src[(size_t)dy * W + dx];
    out[i] = s * (1.0f / (float)(k * k));
}

// mean[a] = sum_p x[a][p] / P in double (one block per frame)
__global__ void __launch_bounds__(kThreads) k_cc_mean(const float* __restrict__ x, long P, double* __restrict__ mean) {
    __shared__ double red[kThreads / 64];
    const float* xa = x + (size_t)blockIdx.x * P;
    double s = 0.0;
    for (long p = threadIdx.x; p < P; p += kThreads) s += (double)xa[p];
    s = rihvr::block_sum(s, red);
    if (threadIdx.x == 0) mean[blockIdx.x] = s / (double)P;
}

// gram[a][b] = sum_p (x[a][p] - m_a)(x[b][p] - m_b) in double, for a <= b (block (b, a))
__global__ void __launch_bounds__(kThreads) k_cc_gram(const float* __restrict__ x, long P, int n,
                                                      const double* __restrict__ mean, double* __restrict__ gram) {
    __shared__ double red[kThreads / 64];
    const int a = blockIdx.y, b = blockIdx.x;
    if (a > b) return;  // uniform per block
    const float* xa = x + (size_t)a * P;
    const float* xb = x + (size_t)b * P;
    const double ma = mean[a], mb = mean[b];
    double s = 0.0;
    for (long p = threadIdx.x; p < P; p += kThreads) s += ((double)xa[p] - ma) * ((double)xb[p] - mb);
    s = rihvr::block_sum(s, red);
    if (threadIdx.x == 0) gram[(size_t)a * n + b] = gram[(size_t)b * n + a] = s;
}

// Entropy histograms of frame f = blockIdx.y against the running mean of its ranked frames:
// for m = 0 .. mmax-1, avg = (x[order[0]] + ... + x[order[m]]) (float, in order) / float(m + 1),
// v = x[f] - avg, bin = clamp(round(v), 0, 255) (std::round: halves away from zero).
// hist: [n][mmax][256] counts (cleared by the caller).  One thread per pixel; LDS counts for
// 32 values of m at a time.
constexpr int kCcChunk = 32;
__global__ void __launch_bounds__(kThreads) k_cc_hist(const float* __restrict__ x, long P, const int* __restrict__ order,
                                                      int mmax, unsigned* __restrict__ hist) {
    __shared__ unsigned h[kCcChunk * 256];
    const int f = blockIdx.y;
    const long p = (long)blockIdx.x * kThreads + threadIdx.x;
    const bool in = p < P;
    const float xf = in ? x[(size_t)f * P + p] : 0.0f;
    const int* ord = order + (size_t)f * mmax;
    float bg = 0.0f;
    for (int m0 = 0; m0 < mmax; m0 += kCcChunk) {
        const int m1 = min(mmax, m0 + kCcChunk);
        for (int e = threadIdx.x; e < kCcChunk * 256; e += kThreads) h[e] = 0u;
        __syncthreads();
        for (int m = m0; m < m1; ++m) {
            if (in) {
                bg += x[(size_t)ord[m] * P + p];
                const float avg = bg / (float)(m + 1);
                const float v = xf - avg;
                const int b = min(255, max(0, (int)roundf(v)));
                atomicAdd(&h[(m - m0) * 256 + b], 1u);
            }
        }
        __syncthreads();
        unsigned* g = hist + ((size_t)f * mmax + m0) * 256;
        for (int e = threadIdx.x; e < (m1 - m0) * 256; e += kThreads)
            if (h[e]) atomicAdd(&g[e], h[e]);
        __syncthreads();
    }
}

// Backgrounds: bgF = (sum of the first used[f] ranked full frames, in order, in float) / float(used[f]);
// out8 = saturate(round half to even(bgF)) (Mat::convertTo to 8U); bgF kept for the enhanced image.
__global__ void k_cc_final(const float* __restrict__ full, long N, const int* __restrict__ order, int mmax,
                           const int* __restrict__ used, float* __restrict__ bgf, unsigned char* __restrict__ out8) {
    const int f = blockIdx.y;
    const long p = (long)blockIdx.x * kThreads + threadIdx.x;
    if (p >= N) return;
    const int u = used[f];
    const int* ord = order + (size_t)f * mmax;
    float s = 0.0f;
    for (int k = 0; k < u; ++k) s += full[(size_t)ord[k] * N + p];
    const float b = s / (float)u;
    if (bgf) bgf[(size_t)f * N + p] = b;
    const float r = rintf(b);
    out8[(size_t)f * N + p] = (unsigned char)fminf(fmaxf(r, 0.0f), 255.0f);
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

rihvr_status rihvr_background_correlation(const float* frames, int32_t n_frames, int32_t width, int32_t height,
                                          int frames_loc, int32_t min_frames, int32_t max_frames,
                                          int32_t resize_factor, uint8_t* backgrounds, uint8_t* enhanced,
                                          int32_t* frames_used) {
    if (!frames || !backgrounds || width <= 0 || height <= 0 || resize_factor < 1 || max_frames < 1 ||
        min_frames < 1)
        return RIHVR_E_INVALID_ARGUMENT;
    if (n_frames < 2) return RIHVR_E_MISSING_DATA;
    if (width % resize_factor || height % resize_factor) return RIHVR_E_INVALID_ARGUMENT;
    const int n = n_frames;
    const long N = (long)width * height;
    const int k = resize_factor;
    const long P = (long)(width / k) * (height / k);
    DevBuf hf, hx, hm, hg, ho, hh, hu, hb, h8;
    const float* df = nullptr;
    rihvr_status st = device_frames(frames, (size_t)n * N, frames_loc, hf, &df);
    if (st) return st;
    // correlation images and the centred Gram matrix (computeCorrelation on the resized frames)
    BG_HIP(hipMalloc(&hx.p, (size_t)n * P * sizeof(float)));
    const long nres = (long)n * P;
    hipLaunchKernelGGL(k_cc_resize, dim3((unsigned)((nres + kThreads - 1) / kThreads)), dim3(kThreads), 0, 0, df,
                       width, height, k, nres, (float*)hx.p);
    BG_HIP(hipMalloc(&hm.p, n * sizeof(double)));
    BG_HIP(hipMalloc(&hg.p, (size_t)n * n * sizeof(double)));
    hipLaunchKernelGGL(k_cc_mean, dim3(n), dim3(kThreads), 0, 0, (const float*)hx.p, P, (double*)hm.p);
    if (n > 65535) return RIHVR_E_INVALID_ARGUMENT;
    hipLaunchKernelGGL(k_cc_gram, dim3(n, n), dim3(kThreads), 0, 0, (const float*)hx.p, P, n, (const double*)hm.p,
                       (double*)hg.p);
    BG_HIP(hipGetLastError());
    std::vector<double> gram((size_t)n * n);
    BG_HIP(hipMemcpy(gram.data(), hg.p, gram.size() * sizeof(double), hipMemcpyDeviceToHost));
    // ranking: the other frames by correlation, highest first (ties: lower frame number first)
    const int mmax = std::min(max_frames, n - 1);
    std::vector<int> order((size_t)n * mmax);
    for (int f = 0; f < n; ++f) {
        std::vector<std::pair<double, int>> sc;
        for (int j = 0; j < n; ++j) {
            if (j == f) continue;
            const double c = gram[(size_t)f * n + j] / (std::sqrt(gram[(size_t)f * n + f] * gram[(size_t)j * n + j]) + 1e-8);
            sc.emplace_back(c, j);
        }
        std::stable_sort(sc.begin(), sc.end(), [](const auto& a, const auto& b) { return a.first > b.first; });
        for (int m = 0; m < mmax; ++m) order[(size_t)f * mmax + m] = sc[m].second;
    }
    BG_HIP(hipMalloc(&ho.p, order.size() * sizeof(int)));
    BG_HIP(hipMemcpy(ho.p, order.data(), order.size() * sizeof(int), hipMemcpyHostToDevice));
    // entropy of the frame minus each running mean (computeEntropy on the resized images)
    const size_t nh = (size_t)n * mmax * 256;
    BG_HIP(hipMalloc(&hh.p, nh * sizeof(unsigned)));
    BG_HIP(hipMemset(hh.p, 0, nh * sizeof(unsigned)));
    hipLaunchKernelGGL(k_cc_hist, dim3((unsigned)((P + kThreads - 1) / kThreads), n), dim3(kThreads), 0, 0,
                       (const float*)hx.p, P, (const int*)ho.p, mmax, (unsigned*)hh.p);
    BG_HIP(hipGetLastError());
    std::vector<unsigned> hist(nh);
    BG_HIP(hipMemcpy(hist.data(), hh.p, nh * sizeof(unsigned), hipMemcpyDeviceToHost));
    std::vector<int> used(n);
    for (int f = 0; f < n; ++f) {
        // entropy in float as computeEntropy: p = float(h) / N; e -= p * log(p)
        std::vector<float> ent(mmax);
        for (int m = 0; m < mmax; ++m) {
            float e = 0.0f;
            const unsigned* hm_ = hist.data() + ((size_t)f * mmax + m) * 256;
            for (int b = 0; b < 256; ++b)
                if (hm_[b] > 0) {
                    const float pr = float(hm_[b]) / (float)P;
                    e -= pr * std::log(pr);
                }
            ent[m] = e;
        }
        // 3-point moving average, first negative gradient, clamp to [min - 1, mmax - 1]
        std::vector<float> sm(mmax);
        for (int i = 0; i < mmax; ++i) {
            float sum = ent[i];
            int cnt = 1;
            if (i > 0) { sum += ent[i - 1]; ++cnt; }
            if (i + 1 < mmax) { sum += ent[i + 1]; ++cnt; }
            sm[i] = sum / cnt;
        }
        int opt = 1;
        for (int i = 1; i < mmax; ++i)
            if (sm[i] - sm[i - 1] < 0) {
                opt = i;
                break;
            }
        opt = std::max(min_frames - 1, std::min(opt, mmax - 1));
        opt = std::min(opt, mmax - 1);  // the reference reads past its ranking when min_frames >= n
        used[f] = opt + 1;
        if (frames_used) frames_used[f] = opt + 1;
    }
    BG_HIP(hipMalloc(&hu.p, n * sizeof(int)));
    BG_HIP(hipMemcpy(hu.p, used.data(), n * sizeof(int), hipMemcpyHostToDevice));
    BG_HIP(hipMalloc(&h8.p, (size_t)n * N));
    if (enhanced) BG_HIP(hipMalloc(&hb.p, (size_t)n * N * sizeof(float)));
    hipLaunchKernelGGL(k_cc_final, dim3((unsigned)((N + kThreads - 1) / kThreads), n), dim3(kThreads), 0, 0, df, N,
                       (const int*)ho.p, mmax, (const int*)hu.p, (float*)hb.p, (unsigned char*)h8.p);
    BG_HIP(hipGetLastError());
    BG_HIP(hipMemcpy(backgrounds, h8.p, (size_t)n * N, hipMemcpyDeviceToHost));
    if (enhanced) {
        // enhanced = normalize(I - bgF, 0, 255, NORM_MINMAX) as 8 bits, on the host
        std::vector<float> bgf((size_t)N), img((size_t)N);
        for (int f = 0; f < n; ++f) {
            BG_HIP(hipMemcpy(bgf.data(), (const float*)hb.p + (size_t)f * N, N * sizeof(float), hipMemcpyDeviceToHost));
            BG_HIP(hipMemcpy(img.data(), df + (size_t)f * N, N * sizeof(float), hipMemcpyDeviceToHost));
            float lo = INFINITY, hi = -INFINITY;
            for (long i = 0; i < N; ++i) {
                const float d = img[i] - bgf[i];
                img[i] = d;
                lo = std::min(lo, d);
                hi = std::max(hi, d);
            }
            const double sc = hi > lo ? 255.0 / ((double)hi - (double)lo) : 0.0;
            const double sh = -(double)lo * sc;
            uint8_t* o = enhanced + (size_t)f * N;
            for (long i = 0; i < N; ++i) {
                const float v = std::nearbyint((float)((double)img[i] * sc + sh));
                o[i] = (uint8_t)std::min(255.0f, std::max(0.0f, v));
            }
        }
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
