// preprocess.hip -- the reconstruction's input contract (main_sparse_inverse.cu:101-124)
// on the device for a batch of frames: raw grey values (8-bit, 16-bit or float) ->
// rescale /255 (hologram.cu:401-439) -> I/mean - 1 (normalize_mean0, :332-358) or
// (I - B)/max(sqrt(B), 0.1) (removeBackground, :1365-1432) -> crop (:768-880) -> zero pad
// (:882-945), written as complex64 holograms ready for the solver.  The arithmetic is
// rihvr_prepare_hologram's (float products and quotients, the mean summed in double).
// The mean's terms are added in block order here and in raster order on the host: for
// 8-bit frames that double sum is exact, so each frame equals the host function's output
// bit for bit; for 16-bit and float frames a value may differ from it by the rounding of
// the sum (at most 1 float ulp).  HBM-bound: one pass over the frames
// for the means, one pass writing the holograms.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdio>
#include <mutex>
#include <vector>

#include "../../include/rihvr.h"
#include "device_common.hpp"
#include "io_internal.hpp"

namespace {

constexpr int kThreads = 256;
constexpr int kItems = 4;  // 16-byte groups per thread in the mean pass
constexpr int kMaxGridYZ = 65535;
constexpr int kPix = 4;    // output pixels per thread in the write pass when rows allow it

template <typename T>
__device__ __forceinline__ float grey(const T* p, long i) {
    return (float)p[i];
}

// per frame (blockIdx.y) and block: double sum of the rescaled pixels, read 16 bytes at a time
template <typename T>
__global__ void __launch_bounds__(kThreads) k_pre_sum(const T* __restrict__ frames, long N, float inv255,
                                                      double* __restrict__ partial) {
    constexpr int G = 16 / sizeof(T);  // pixels per group
    __shared__ double red[kThreads / 64];
    const T* f = frames + (size_t)blockIdx.y * N;
    double s = 0.0;
    const long ng = N / G;  // whole groups; frames start at multiples of N elements
    const bool aligned = ((size_t)f % 16) == 0;
    for (int it = 0; it < kItems; ++it) {
        const long q = ((long)blockIdx.x * kItems + it) * kThreads + threadIdx.x;
        if (q >= ng) break;
        if (aligned) {
            union {
                uint4 u;
                T px[G];
            } v;
            v.u = ((const uint4*)f)[q];
#pragma unroll
            for (int k = 0; k < G; ++k) s += (double)((float)v.px[k] * inv255);
        } else {
#pragma unroll
            for (int k = 0; k < G; ++k) s += (double)(grey(f, G * q + k) * inv255);
        }
    }
    // the last N mod G pixels
    if (blockIdx.x == 0 && threadIdx.x < (unsigned)(N - G * ng)) s += (double)(grey(f, G * ng + threadIdx.x) * inv255);
    s = rihvr::block_sum(s, red);
    if (threadIdx.x == 0) partial[(size_t)blockIdx.y * gridDim.x + blockIdx.x] = s;
}

// inv_mean[f] = float(1 / (sum / N)), the partial sums added in order
__global__ void k_pre_mean(const double* __restrict__ partial, int blocks, int n_frames, long N,
                           float* __restrict__ inv_mean) {
    const int f = blockIdx.x * blockDim.x + threadIdx.x;
    if (f >= n_frames) return;
    double s = 0.0;
    for (int b = 0; b < blocks; ++b) s += partial[(size_t)f * blocks + b];
    inv_mean[f] = (float)(1.0 / (s / (double)N));
}

struct Geometry {
    int W, H;            // image
    int src_x0, src_y0;  // source of the crop's first pixel
    int rw, rh, pad, ow, oh;
};

// output pixels PIX * (blockIdx.x * kThreads + threadIdx.x) + [0, PIX) of row blockIdx.y of frame
// blockIdx.z + f0; PIX = kPix needs ow % kPix == 0 and a 16-byte aligned output (two complex per store).
// Every pixel's loads are issued unconditionally (pixel 0 of the frame stands in for a pixel outside
// the crop or the image) so that they overlap; the select afterwards keeps zero for those pixels.
template <typename T, int PIX>
__global__ void __launch_bounds__(kThreads) k_pre_write(const T* __restrict__ frames,
                                                        const float* __restrict__ background, Geometry g,
                                                        int f0, float inv255,
                                                        const float* __restrict__ inv_mean,
                                                        float2* __restrict__ out) {
#pragma clang fp contract(off)  // the host function's separate float roundings
    const int x = PIX * (blockIdx.x * kThreads + threadIdx.x);
    if (x >= g.ow) return;
    const int y = blockIdx.y;
    const int f = f0 + blockIdx.z;
    const T* fr = frames + (size_t)f * g.W * g.H;
    const int cy = y - g.pad, sy = g.src_y0 + cy;
    const bool row_in = cy >= 0 && cy < g.rh && sy >= 0 && sy < g.H;
    bool in[PIX];
    long si[PIX];
#pragma unroll
    for (int k = 0; k < PIX; ++k) {
        const int cx = x + k - g.pad, sx = g.src_x0 + cx;
        in[k] = row_in && cx >= 0 && cx < g.rw && sx >= 0 && sx < g.W;
        si[k] = in[k] ? (long)sy * g.W + sx : 0;
    }
    float h[PIX], v[PIX];
#pragma unroll
    for (int k = 0; k < PIX; ++k) h[k] = grey(fr, si[k]) * inv255;
    if (background) {
        float bk[PIX];
#pragma unroll
        for (int k = 0; k < PIX; ++k) bk[k] = background[si[k]];
#pragma unroll
        for (int k = 0; k < PIX; ++k) {
            const float b = bk[k] * inv255;
            const float sq = sqrtf(b);
            const float sb = sq < 0.1f ? 0.1f : sq;  // std::max(sqrt(b), 0.1f)
            v[k] = in[k] ? (h[k] - b) / sb : 0.0f;
        }
    } else {
        const float im = inv_mean[f];
#pragma unroll
        for (int k = 0; k < PIX; ++k) v[k] = in[k] ? h[k] * im - 1.0f : 0.0f;
    }
    float2* o = out + ((size_t)f * g.oh + y) * g.ow + x;
    if constexpr (PIX == 1) {
        *o = make_float2(v[0], 0.0f);
    } else {
#pragma unroll
        for (int k = 0; k < PIX; k += 2) ((float4*)o)[k / 2] = make_float4(v[k], 0.0f, v[k + 1], 0.0f);
    }
}

#define PRE_HIP(call)                                                                              \
    do {                                                                                           \
        hipError_t e_ = (call);                                                                    \
        if (e_ != hipSuccess) {                                                                    \
            std::fprintf(stderr, "rihvr: HIP error %s at %s:%d\n", hipGetErrorString(e_), __FILE__, \
                         __LINE__);                                                                \
            return e_ == hipErrorOutOfMemory ? RIHVR_E_OUT_OF_MEMORY : RIHVR_E_DEVICE;              \
        }                                                                                          \
    } while (0)

struct DevBuf {
    void* p = nullptr;
    ~DevBuf() {
        if (p) (void)hipFree(p);
    }
};

// The two passes with caller-provided scratch: partial (n * sum_blocks<T>(N) doubles, unused with
// a background) and inv_mean (n floats).  Launches on the null stream and waits for it.
template <typename T>
long sum_blocks(long N) {
    const long per_block = (long)kThreads * kItems * (16 / sizeof(T));
    return std::max(1L, (N + per_block - 1) / per_block);
}

template <typename T>
rihvr_status run_with(const T* frames, const float* bg, const Geometry& g, int n, float* out_dev, double* partial,
                      float* inv_mean) {
    const float inv255 = (float)(1.0 / 255.0);  // cv::Mat / 255 scales by the float reciprocal
    const long N = (long)g.W * g.H;
    if (!bg) {
        const int blocks = (int)sum_blocks<T>(N);
        for (int f0 = 0; f0 < n; f0 += kMaxGridYZ)
            hipLaunchKernelGGL(k_pre_sum<T>, dim3(blocks, std::min(n - f0, kMaxGridYZ)), dim3(kThreads), 0, 0,
                               frames + (size_t)f0 * N, N, inv255, partial + (size_t)f0 * blocks);
        hipLaunchKernelGGL(k_pre_mean, dim3((n + 63) / 64), dim3(64), 0, 0, (const double*)partial, blocks, n, N,
                           inv_mean);
        PRE_HIP(hipGetLastError());
    }
    // frames in launches of at most kMaxGridYZ (the grid's y and z limit)
    const bool wide = g.ow % kPix == 0 && (size_t)out_dev % 16 == 0;
    const int pix = wide ? kPix : 1;
    for (int f0 = 0; f0 < n; f0 += kMaxGridYZ) {
        const dim3 grid((unsigned)((g.ow + kThreads * pix - 1) / (kThreads * pix)), (unsigned)g.oh,
                        (unsigned)std::min(n - f0, kMaxGridYZ));
        if (wide)
            hipLaunchKernelGGL((k_pre_write<T, kPix>), grid, dim3(kThreads), 0, 0, frames, bg, g, f0, inv255,
                               (const float*)inv_mean, (float2*)out_dev);
        else
            hipLaunchKernelGGL((k_pre_write<T, 1>), grid, dim3(kThreads), 0, 0, frames, bg, g, f0, inv255,
                               (const float*)inv_mean, (float2*)out_dev);
    }
    PRE_HIP(hipGetLastError());
    PRE_HIP(hipDeviceSynchronize());
    return RIHVR_OK;
}

template <typename T>
rihvr_status run(const T* frames, const float* bg, const Geometry& g, int n, float* out_dev) {
    DevBuf hm, hp;
    PRE_HIP(hipMalloc(&hm.p, (size_t)n * sizeof(float)));
    if (!bg) PRE_HIP(hipMalloc(&hp.p, (size_t)n * sum_blocks<T>((long)g.W * g.H) * sizeof(double)));
    return run_with(frames, bg, g, n, out_dev, (double*)hp.p, (float*)hm.p);
}

// Scratch of rihvr_prepare_holograms_noalloc: the mean pass's per-frame partial sums, then the
// frames' inverse means.  Each GPU the library runs on holds its own copy until the process ends.
constexpr size_t kScratchBytes = size_t(1) << 20;
__device__ __align__(256) unsigned char scratch[kScratchBytes];
constexpr int kMaxDevices = 64;
std::mutex scratch_mu;
void* scratch_ptr[kMaxDevices] = {};

// The device address of this device's scratch (the array's name is a host-side stand-in).
rihvr_status device_scratch(void** out) {
    int dev = 0;
    PRE_HIP(hipGetDevice(&dev));
    if (dev < 0 || dev >= kMaxDevices) return RIHVR_E_INVALID_ARGUMENT;
    std::lock_guard<std::mutex> lock(scratch_mu);
    if (!scratch_ptr[dev]) {
        void* p = nullptr;
        PRE_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        scratch_ptr[dev] = p;
    }
    *out = scratch_ptr[dev];
    return RIHVR_OK;
}

rihvr_status check_args(const rihvr_file_params* fp, const void* frames, int32_t pixel_type, int32_t n_frames,
                        const float* background, int32_t width, int32_t height, int32_t* out_width,
                        int32_t* out_height, rihvr::CropGeometry* cg) {
    if (!fp || !frames || width <= 0 || height <= 0 || n_frames < 1 || !out_width || !out_height)
        return RIHVR_E_INVALID_ARGUMENT;
    if (pixel_type != RIHVR_PIXEL_U8 && pixel_type != RIHVR_PIXEL_U16 && pixel_type != RIHVR_PIXEL_F32)
        return RIHVR_E_UNKNOWN_TYPE;
    if (fp->remove_background && !background) return RIHVR_E_MISSING_DATA;
    const rihvr_status gs = rihvr::crop_geometry(fp, width, height, cg);
    if (gs) return gs;
    *out_width = cg->ow;
    *out_height = cg->oh;
    return RIHVR_OK;
}

}  // namespace

extern "C" {

rihvr_status rihvr_prepare_holograms(const rihvr_file_params* fp, const void* frames, int32_t pixel_type,
                                     int32_t n_frames, const float* background, int32_t width, int32_t height,
                                     int in_loc, float* out, int out_loc, int32_t* out_width, int32_t* out_height) {
    rihvr::CropGeometry cg;
    const rihvr_status as = check_args(fp, frames, pixel_type, n_frames, background, width, height, out_width,
                                       out_height, &cg);
    if (as) return as;
    if (!out) return RIHVR_OK;
    if (cg.oh > kMaxGridYZ) return RIHVR_E_INVALID_ARGUMENT;  // one output row per grid row
    const Geometry g{width, height, cg.src_x0, cg.src_y0, cg.rw, cg.rh, cg.pad, cg.ow, cg.oh};
    const size_t in_bytes = (size_t)n_frames * width * height * (size_t)pixel_type;
    const size_t out_bytes = (size_t)n_frames * cg.ow * cg.oh * 8;
    DevBuf hin, hbg, hout;
    const void* din = frames;
    const float* dbg = fp->remove_background ? background : nullptr;
    if (in_loc != RIHVR_DEVICE) {
        PRE_HIP(hipMalloc(&hin.p, in_bytes));
        PRE_HIP(hipMemcpy(hin.p, frames, in_bytes, hipMemcpyHostToDevice));
        din = hin.p;
        if (dbg) {
            PRE_HIP(hipMalloc(&hbg.p, (size_t)width * height * sizeof(float)));
            PRE_HIP(hipMemcpy(hbg.p, dbg, (size_t)width * height * sizeof(float), hipMemcpyHostToDevice));
            dbg = (const float*)hbg.p;
        }
    }
    float* dout = out;
    if (out_loc != RIHVR_DEVICE) {
        PRE_HIP(hipMalloc(&hout.p, out_bytes));
        dout = (float*)hout.p;
    }
    rihvr_status st;
    if (pixel_type == RIHVR_PIXEL_U8)
        st = run((const uint8_t*)din, dbg, g, n_frames, dout);
    else if (pixel_type == RIHVR_PIXEL_U16)
        st = run((const uint16_t*)din, dbg, g, n_frames, dout);
    else
        st = run((const float*)din, dbg, g, n_frames, dout);
    if (st) return st;
    if (out_loc != RIHVR_DEVICE) PRE_HIP(hipMemcpy(out, dout, out_bytes, hipMemcpyDeviceToHost));
    return RIHVR_OK;
}

rihvr_status rihvr_prepare_holograms_noalloc(const rihvr_file_params* fp, const void* frames, int32_t pixel_type,
                                             int32_t n_frames, const float* background, int32_t width,
                                             int32_t height, float* out, int32_t* out_width, int32_t* out_height) {
    rihvr::CropGeometry cg;
    const rihvr_status as = check_args(fp, frames, pixel_type, n_frames, background, width, height, out_width,
                                       out_height, &cg);
    if (as) return as;
    if (!out) return RIHVR_OK;
    if (cg.oh > kMaxGridYZ) return RIHVR_E_INVALID_ARGUMENT;
    const long N = (long)width * height;
    const long blocks = pixel_type == RIHVR_PIXEL_U8    ? sum_blocks<uint8_t>(N)
                        : pixel_type == RIHVR_PIXEL_U16 ? sum_blocks<uint16_t>(N)
                                                        : sum_blocks<float>(N);
    // partial sums (none with a background), then the inverse means at a 256-byte boundary
    const size_t part_bytes = fp->remove_background ? 0 : (size_t)n_frames * blocks * sizeof(double);
    const size_t mean_off = (part_bytes + 255) & ~size_t(255);
    if (mean_off + (size_t)n_frames * sizeof(float) > kScratchBytes) {
        std::fprintf(stderr,
                     "rihvr_prepare_holograms_noalloc: %d frames of %dx%d need %zu scratch bytes, more than the "
                     "%zu this entry point holds\n",
                     n_frames, width, height, mean_off + (size_t)n_frames * sizeof(float), kScratchBytes);
        return RIHVR_E_INVALID_ARGUMENT;
    }
    void* sp = nullptr;
    const rihvr_status ss = device_scratch(&sp);
    if (ss) return ss;
    double* partial = (double*)sp;
    float* inv_mean = (float*)((unsigned char*)sp + mean_off);
    const Geometry g{width, height, cg.src_x0, cg.src_y0, cg.rw, cg.rh, cg.pad, cg.ow, cg.oh};
    const float* dbg = fp->remove_background ? background : nullptr;
    if (pixel_type == RIHVR_PIXEL_U8)
        return run_with((const uint8_t*)frames, dbg, g, n_frames, out, partial, inv_mean);
    if (pixel_type == RIHVR_PIXEL_U16)
        return run_with((const uint16_t*)frames, dbg, g, n_frames, out, partial, inv_mean);
    return run_with((const float*)frames, dbg, g, n_frames, out, partial, inv_mean);
}

}  // extern "C"
