// segmentation.hip -- segmentation of a reconstructed volume into objects and their
// centroids: the chain of the reference's segmentation executable
// (src/core/main_segmentation.cu:83-125) over SparseSegmentation
// (sparse_segmentation.cu), ObjectCloud (object_cloud.cu) and Blob3d (blob3d.cu).
//
// The reference walks sparse per-plane lists on the host (bubble-sorted indices, a
// binary search per neighbour, one kernel launch per voxel for includeIntensity).
// Here every pass over the volume is a dense, HBM-streaming kernel over
// [Nz][H][W] buffers:
//   load      scatter of the voxel list into a u32 key volume (atomicMax keeps the
//             larger intensity of a repeated voxel) + max intensity
//   binarize  key -> binary u8 volume and float values (in place of the keys)
//   close     ball dilation then erosion, each as a row pass (distance along x to the
//             nearest set / unset voxel, capped at r+1) and a (y, z) pass over the
//             ball's rows: voxel t is dilated iff some row (dy, dz) of the ball has a
//             set voxel within its half-width w(dy, dz) = floor(sqrt(r^2-dy^2-dz^2))
//   label     union-find with atomicMin links to the smaller root (every root is the
//             smallest index of its tree), then a flattening pass: each voxel ends
//             with the smallest global index of its component
//   extract   ordered stream compaction of the object voxels (raster order), object
//             numbers = rank of the root among the roots (first appearance order)
// The object statistics need the reference's float sums in voxel order; they are one
// pass over the compacted voxels on the host.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <vector>

#include "../../include/rihvr.h"

namespace {

constexpr int kThreads = 256;
constexpr uint32_t kNone = 0xFFFFFFFFu;
constexpr int kItems = 16;                        // voxels per thread in the compaction
constexpr long kCompactSpan = (long)kThreads * kItems;  // voxels per compaction block
constexpr int kMaxClose = 254;                    // distances are stored in a byte

// key of a binarized intensity: 0 = no voxel, 1 = NaN, else float bits + 2 (v >= 0)
__device__ __forceinline__ uint32_t intensity_key(float v) { return v != v ? 1u : __float_as_uint(v) + 2u; }
__device__ __forceinline__ float key_value(uint32_t k) {
    return k == 0 ? 0.0f : (k == 1 ? __uint_as_float(0x7FC00000u) : __uint_as_float(k - 2u));
}

__device__ __forceinline__ uint32_t wave_max_u32(uint32_t v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v = max(v, (uint32_t)__shfl_xor((int)v, off, 64));
    return v;
}

// value.x = sqrt(intensity) as loadData stores it; the binarized intensity is x*x
// (no fused multiply-add: the reference's host and device products round once).
// sqrtf is the correctly rounded square root (v_sqrt_f32 plus its refinement step;
// the bare instruction, which __fsqrt_rn emits, can be one ulp off).
__device__ __forceinline__ float reloaded_intensity(float stored) {
    const float a = sqrtf(stored);
    return a * a;
}

__global__ void __launch_bounds__(kThreads) k_seg_scatter(const int32_t* __restrict__ plane,
                                                          const int32_t* __restrict__ index,
                                                          const float* __restrict__ inten, long n, int Nz, long P,
                                                          uint32_t* __restrict__ key, uint32_t* __restrict__ maxbits,
                                                          uint32_t* __restrict__ bad) {
    uint32_t m = 0;
    for (long i = (long)blockIdx.x * kThreads + threadIdx.x; i < n; i += (long)gridDim.x * kThreads) {
        const int z = plane[i];
        const int idx = index[i];
        const float v = reloaded_intensity(inten[i]);
        if (z < 0 || z >= Nz || idx < 0 || (long)idx >= P) {
            atomicAdd(bad, 1u);
            continue;
        }
        atomicMax(&key[(size_t)z * P + idx], intensity_key(v));
        if (v > 0.0f) m = max(m, __float_as_uint(v));  // v >= 0: bits order as values; NaN skipped
    }
    m = wave_max_u32(m);
    if ((threadIdx.x & 63) == 0 && m) atomicMax(maxbits, m);
}

// the same from a dense complex volume: intensity = re^2 + im^2 of the non-zero voxels
__global__ void __launch_bounds__(kThreads) k_seg_dense(const float2* __restrict__ vol, long N,
                                                        uint32_t* __restrict__ key, uint32_t* __restrict__ maxbits) {
#pragma clang fp contract(off)  // re^2 + im^2 with two roundings, as the file's writer computes it
    uint32_t m = 0;
    for (long i = (long)blockIdx.x * kThreads + threadIdx.x; i < N; i += (long)gridDim.x * kThreads) {
        const float2 c = vol[i];
        uint32_t k = 0;
        if (c.x != 0.0f || c.y != 0.0f) {
            const float stored = c.x * c.x + c.y * c.y;
            const float v = reloaded_intensity(stored);
            k = intensity_key(v);
            if (v > 0.0f) m = max(m, __float_as_uint(v));
        }
        key[i] = k;
    }
    m = wave_max_u32(m);
    if ((threadIdx.x & 63) == 0 && m) atomicMax(maxbits, m);
}

// binary = intensity > thr (float compared in double, binarize_kernel :39-50); the key
// becomes the value of the masked volume (mask_kernel :52-61): intensity or 0
__global__ void __launch_bounds__(kThreads) k_seg_binarize(uint32_t* __restrict__ key_value_bits,
                                                           uint8_t* __restrict__ mask, long N, double thr) {
    for (long i = (long)blockIdx.x * kThreads + threadIdx.x; i < N; i += (long)gridDim.x * kThreads) {
        const float v = key_value(key_value_bits[i]);
        const bool b = (double)v > thr;
        mask[i] = b ? 1 : 0;
        key_value_bits[i] = b ? __float_as_uint(v) : 0u;
    }
}

// Row pass of the closing: dst = distance along x to the nearest "set" voxel, capped at
// cap.  Dilation: set = src != 0, outside the row is not set.  Erosion: set = src == 0
// or outside the row (a ball that leaves the volume fails the erosion).  One block per
// 256-voxel segment of a row, the segment and a halo of cap voxels staged in LDS.
__global__ void __launch_bounds__(kThreads) k_seg_rowdist(const uint8_t* __restrict__ src,
                                                          uint8_t* __restrict__ dst, int W, long rows, int cap,
                                                          int erode) {
    __shared__ uint8_t s[kThreads + 2 * (kMaxClose + 1)];
    const long segs = (W + kThreads - 1) / kThreads;
    for (long b = blockIdx.x; b < rows * segs; b += gridDim.x) {
        const long row = b / segs;
        const int x0 = (int)(b - row * segs) * kThreads;
        const uint8_t* r = src + row * (long)W;
        __syncthreads();
        for (int j = threadIdx.x; j < kThreads + 2 * cap; j += kThreads) {
            const int x = x0 - cap + j;
            const bool inside = x >= 0 && x < W;
            s[j] = erode ? (!inside || r[inside ? x : 0] == 0) : (inside && r[inside ? x : 0] != 0);
        }
        __syncthreads();
        const int x = x0 + threadIdx.x;
        if (x < W) {
            int d = cap;
            const int c = cap + threadIdx.x;
            for (int k = 0; k < cap; ++k)
                if (s[c - k] | s[c + k]) {
                    d = k;
                    break;
                }
            dst[row * (long)W + x] = (uint8_t)d;
        }
    }
}

__device__ __forceinline__ int isqrt_floor(int n) {
    int r = (int)sqrtf((float)n);
    while (r * r > n) --r;
    while ((r + 1) * (r + 1) <= n) ++r;
    return r;
}

// (y, z) pass of the closing over the ball's rows (dy, dz), half-width w = floor(sqrt(
// r^2 - dy^2 - dz^2)).  Dilation: some in-volume row with dist <= w.  Erosion: every row
// inside the volume with dist > w (dist = distance to the nearest unset voxel or border).
__global__ void __launch_bounds__(kThreads) k_seg_ball(const uint8_t* __restrict__ dist,
                                                       uint8_t* __restrict__ out, int W, int H, int Nz, int r,
                                                       int erode) {
    const long N = (long)W * H * Nz;
    const long plane = (long)W * H;
    for (long i = (long)blockIdx.x * kThreads + threadIdx.x; i < N; i += (long)gridDim.x * kThreads) {
        const int z = (int)(i / plane);
        const long rem = i - (long)z * plane;
        const int y = (int)(rem / W);
        const int x = (int)(rem - (long)y * W);
        bool res = erode != 0;
        for (int dz = -r; dz <= r; ++dz) {
            const int zz = z + dz;
            const int rz = r * r - dz * dz;
            const int wy = isqrt_floor(rz);
            if (zz < 0 || zz >= Nz) {
                if (erode) {
                    res = false;
                    break;
                }
                continue;
            }
            bool stop = false;
            for (int dy = -wy; dy <= wy; ++dy) {
                const int yy = y + dy;
                if (yy < 0 || yy >= H) {
                    if (erode) {
                        res = false;
                        stop = true;
                        break;
                    }
                    continue;
                }
                const int w = isqrt_floor(rz - dy * dy);
                const int d = dist[((long)zz * H + yy) * W + x];
                if (d <= w) {  // dilation: a set voxel in reach; erosion: an unset one
                    res = !erode;
                    stop = true;
                    break;
                }
            }
            if (stop) break;
        }
        out[i] = res ? 1 : 0;
    }
}

// --- connected components ---------------------------------------------------
__device__ __forceinline__ uint32_t load_label(const uint32_t* lab, uint32_t i) {
    return __hip_atomic_load(lab + i, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ uint32_t find_root(const uint32_t* lab, uint32_t x) {
    uint32_t p = load_label(lab, x);
    while (p != x) {
        x = p;
        p = load_label(lab, x);
    }
    return x;
}
// link the larger root under the smaller; labels only decrease, so every loop ends
__device__ void unite(uint32_t* lab, uint32_t a, uint32_t b) {
    for (;;) {
        a = find_root(lab, a);
        b = find_root(lab, b);
        if (a == b) return;
        if (a < b) {
            const uint32_t t = a;
            a = b;
            b = t;
        }
        const uint32_t old = atomicMin(lab + a, b);
        if (old == a) return;
        a = old;
    }
}

__global__ void __launch_bounds__(kThreads) k_ccl_init(const uint8_t* __restrict__ mask, uint32_t* __restrict__ lab,
                                                       long N) {
    for (long i = (long)blockIdx.x * kThreads + threadIdx.x; i < N; i += (long)gridDim.x * kThreads)
        lab[i] = mask[i] ? (uint32_t)i : kNone;
}

// Each object voxel unites with its 13 earlier neighbours (the other 13 unite with it):
// in-plane flat offsets dy*W + dx kept inside 0..P-1 (labelConnectedComponents'
// offsets, :452-480), on the previous plane all nine, on the same plane the four below 0.
__global__ void __launch_bounds__(kThreads) k_ccl_merge(const uint8_t* __restrict__ mask, uint32_t* lab, int W,
                                                        long P, long N) {
    const long off[4] = {-(long)W - 1, -(long)W, -(long)W + 1, -1};
    for (long v = (long)blockIdx.x * kThreads + threadIdx.x; v < N; v += (long)gridDim.x * kThreads) {
        if (!mask[v]) continue;
        const long z = v / P;
        const long i = v - z * P;
        for (int k = 0; k < 4; ++k) {
            const long j = i + off[k];
            if (j >= 0 && mask[v + off[k]]) unite(lab, (uint32_t)v, (uint32_t)(v + off[k]));
        }
        if (z > 0) {
            for (int dy = -1; dy <= 1; ++dy)
                for (int dx = -1; dx <= 1; ++dx) {
                    const long o = (long)dy * W + dx;
                    const long j = i + o;
                    if (j >= 0 && j < P && mask[v - P + o]) unite(lab, (uint32_t)v, (uint32_t)(v - P + o));
                }
        }
    }
}

__global__ void __launch_bounds__(kThreads) k_ccl_flatten(uint32_t* lab, long N) {
    for (long v = (long)blockIdx.x * kThreads + threadIdx.x; v < N; v += (long)gridDim.x * kThreads) {
        const uint32_t l = lab[v];
        if (l != kNone && l != (uint32_t)v) lab[v] = find_root(lab, l);
    }
}

// --- ordered compaction -----------------------------------------------------
// A block covers kCompactSpan consecutive voxels; item k of thread t is voxel
// base + k*256 + t, so (k, t) order is raster order.
__device__ __forceinline__ uint32_t block_prefix(bool f, uint32_t* wave_counts, uint32_t* total) {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const uint64_t bal = __ballot(f);
    const uint32_t below = (uint32_t)__popcll(bal & ((1ull << lane) - 1ull));
    if (lane == 0) wave_counts[wave] = (uint32_t)__popcll(bal);
    __syncthreads();
    uint32_t before = 0, all = 0;
    for (int w = 0; w < kThreads / 64; ++w) {
        const uint32_t c = wave_counts[w];
        before += w < wave ? c : 0;
        all += c;
    }
    __syncthreads();
    *total = all;
    return before + below;
}

__global__ void __launch_bounds__(kThreads) k_cmp_count(const uint32_t* __restrict__ lab, long N,
                                                        uint32_t* __restrict__ counts) {
    __shared__ uint32_t wc[kThreads / 64];
    const long base = (long)blockIdx.x * kCompactSpan;
    uint32_t nobj = 0, nroot = 0;
    for (int k = 0; k < kItems; ++k) {
        const long v = base + (long)k * kThreads + threadIdx.x;
        const uint32_t l = v < N ? lab[v] : kNone;
        nobj += l != kNone;
        nroot += l == (uint32_t)v;
    }
    uint32_t t = nobj;
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) t += (uint32_t)__shfl_xor((int)t, off, 64);
    uint32_t u = nroot;
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) u += (uint32_t)__shfl_xor((int)u, off, 64);
    __shared__ uint32_t wr[kThreads / 64];
    if ((threadIdx.x & 63) == 0) {
        wc[threadIdx.x >> 6] = t;
        wr[threadIdx.x >> 6] = u;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t a = 0, b = 0;
        for (int w = 0; w < kThreads / 64; ++w) {
            a += wc[w];
            b += wr[w];
        }
        counts[2 * blockIdx.x] = a;
        counts[2 * blockIdx.x + 1] = b;
    }
}

// exclusive scan of the (objects, roots) block counts in one workgroup of 1024 threads;
// totals[0..1] = the sums
__global__ void __launch_bounds__(1024) k_cmp_scan(const uint32_t* __restrict__ counts, long nblocks,
                                                   uint32_t* __restrict__ offsets, uint32_t* __restrict__ totals) {
    __shared__ uint32_t sa[1024], sb[1024];
    const long per = (nblocks + 1023) / 1024;
    const long b0 = (long)threadIdx.x * per;
    const long b1 = min(nblocks, b0 + per);
    uint32_t a = 0, b = 0;
    for (long j = b0; j < b1; ++j) {
        a += counts[2 * j];
        b += counts[2 * j + 1];
    }
    sa[threadIdx.x] = a;
    sb[threadIdx.x] = b;
    __syncthreads();
    for (int off = 1; off < 1024; off <<= 1) {  // inclusive Hillis-Steele scan
        const uint32_t xa = threadIdx.x >= (unsigned)off ? sa[threadIdx.x - off] : 0u;
        const uint32_t xb = threadIdx.x >= (unsigned)off ? sb[threadIdx.x - off] : 0u;
        __syncthreads();
        sa[threadIdx.x] += xa;
        sb[threadIdx.x] += xb;
        __syncthreads();
    }
    uint32_t ra = sa[threadIdx.x] - a, rb = sb[threadIdx.x] - b;
    for (long j = b0; j < b1; ++j) {
        offsets[2 * j] = ra;
        offsets[2 * j + 1] = rb;
        ra += counts[2 * j];
        rb += counts[2 * j + 1];
    }
    if (threadIdx.x == 1023) {
        totals[0] = sa[1023];
        totals[1] = sb[1023];
    }
}

// rank of every root among the roots in raster order (its object number)
__global__ void __launch_bounds__(kThreads) k_cmp_roots(const uint32_t* __restrict__ lab, long N,
                                                        const uint32_t* __restrict__ offsets,
                                                        uint32_t* __restrict__ rank) {
    __shared__ uint32_t wc[kThreads / 64];
    const long base = (long)blockIdx.x * kCompactSpan;
    uint32_t run = offsets[2 * blockIdx.x + 1];
    for (int k = 0; k < kItems; ++k) {
        const long v = base + (long)k * kThreads + threadIdx.x;
        const bool f = v < N && lab[v] == (uint32_t)v;
        uint32_t all;
        const uint32_t pos = block_prefix(f, wc, &all);
        if (f) rank[v] = run + pos;
        run += all;
    }
}

// object voxels in raster order: global index, object number, value
__global__ void __launch_bounds__(kThreads) k_cmp_scatter(const uint32_t* __restrict__ lab,
                                                          const float* __restrict__ value, long N,
                                                          const uint32_t* __restrict__ offsets,
                                                          const uint32_t* __restrict__ rank,
                                                          uint32_t* __restrict__ out_index,
                                                          uint32_t* __restrict__ out_object,
                                                          float* __restrict__ out_value) {
    __shared__ uint32_t wc[kThreads / 64];
    const long base = (long)blockIdx.x * kCompactSpan;
    uint32_t run = offsets[2 * blockIdx.x];
    for (int k = 0; k < kItems; ++k) {
        const long v = base + (long)k * kThreads + threadIdx.x;
        const uint32_t l = v < N ? lab[v] : kNone;
        const bool f = l != kNone;
        uint32_t all;
        const uint32_t pos = block_prefix(f, wc, &all);
        if (f) {
            out_index[run + pos] = (uint32_t)v;
            out_object[run + pos] = rank[l];
            out_value[run + pos] = value[v];
        }
        run += all;
    }
}

#define SEG_HIP(call)                                                                              \
    do {                                                                                           \
        hipError_t e_ = (call);                                                                    \
        if (e_ != hipSuccess) {                                                                    \
            std::fprintf(stderr, "rihvr: HIP error %s at %s:%d\n", hipGetErrorString(e_), __FILE__, \
                         __LINE__);                                                                \
            return e_ == hipErrorOutOfMemory ? RIHVR_E_OUT_OF_MEMORY : RIHVR_E_DEVICE;              \
        }                                                                                          \
    } while (0)

unsigned grid_for(long n) {
    const long b = (n + kThreads - 1) / kThreads;
    return (unsigned)std::max(1L, std::min(b, 256L * 1024L));
}

enum Stage { EMPTY = 0, LOADED, BINARIZED, LABELED, EXTRACTED };
enum Timer { T_LOAD = 0, T_BINARIZE, T_CLOSE, T_LABEL, T_EXTRACT, T_COUNT };

}  // namespace

struct rihvr_segmentation {
    int W = 0, H = 0, Nz = 0;
    long P = 0, N = 0;
    int device = 0;
    hipStream_t stream = nullptr;
    uint32_t* key = nullptr;     // keys, then values (float bits)
    uint8_t* mask = nullptr;     // binary volume
    uint8_t* work = nullptr;     // 4N bytes: two byte volumes for the closing, or the root ranks
    uint32_t* lab = nullptr;     // labels
    uint32_t* counts = nullptr;  // 2 per compaction block
    uint32_t* offsets = nullptr;
    uint32_t* small = nullptr;   // [0] max bits, [1] bad entries, [2..3] totals
    long nblocks = 0;
    void* list = nullptr;        // staging of a host voxel list or volume
    size_t list_bytes = 0;
    hipEvent_t ev[T_COUNT][2] = {};
    bool timed[T_COUNT] = {};
    Stage stage = EMPTY;
    std::vector<rihvr_object> objects;
};

namespace {

rihvr_status stage_buffer(rihvr_segmentation* s, size_t bytes) {
    if (s->list_bytes >= bytes) return RIHVR_OK;
    if (s->list) SEG_HIP(hipFree(s->list));
    s->list = nullptr;
    s->list_bytes = 0;
    SEG_HIP(hipMalloc(&s->list, bytes));
    s->list_bytes = bytes;
    return RIHVR_OK;
}

struct DeviceGuard {
    int prev = -1;
    explicit DeviceGuard(int d) {
        if (hipGetDevice(&prev) != hipSuccess) prev = -1;
        if (prev != d) (void)hipSetDevice(d);
    }
    ~DeviceGuard() {
        int cur = -1;
        if (prev >= 0 && hipGetDevice(&cur) == hipSuccess && cur != prev) (void)hipSetDevice(prev);
    }
};

rihvr_status begin_timer(rihvr_segmentation* s, int t) {
    SEG_HIP(hipEventRecord(s->ev[t][0], s->stream));
    return RIHVR_OK;
}
rihvr_status end_timer(rihvr_segmentation* s, int t) {
    SEG_HIP(hipEventRecord(s->ev[t][1], s->stream));
    s->timed[t] = true;
    return RIHVR_OK;
}

rihvr_status finish_load(rihvr_segmentation* s) {
    uint32_t bad = 0;
    SEG_HIP(hipMemcpyAsync(&bad, s->small + 1, sizeof(bad), hipMemcpyDeviceToHost, s->stream));
    SEG_HIP(hipStreamSynchronize(s->stream));
    s->objects.clear();
    if (bad) {
        s->stage = EMPTY;
        return RIHVR_E_INVALID_DATA;
    }
    s->stage = LOADED;
    return RIHVR_OK;
}

// One pass over the compacted object voxels in raster order: Blob3d's statistics with
// its arithmetic (int sums, float products and sums in voxel order, double sum of the
// values).  Each object's first voxel is its root, so object numbers appear in order.
rihvr_status accumulate_objects(const rihvr_segmentation* s, const std::vector<uint32_t>& gidx,
                                const std::vector<uint32_t>& obj, const std::vector<float>& val,
                                std::vector<rihvr_object>& out) {
#pragma clang fp contract(off)
    struct Acc {
        uint32_t sx = 0, sy = 0, sz = 0;  // cv::Point3i sums, two's complement wrap
        float wx = 0, wy = 0, wz = 0, ws = 0, mx = 0;
        double sum = 0;
        int64_t n = 0, label = 0;
        int32_t b[6] = {0, 0, 0, 0, 0, 0};
    };
    std::vector<Acc> acc;
    for (size_t j = 0; j < gidx.size(); ++j) {
        const uint32_t g = gidx[j];
        const uint32_t o = obj[j];
        if (o > acc.size()) return RIHVR_E_UNKNOWN_ERROR;
        const int32_t z = (int32_t)(g / (uint32_t)s->P);
        const uint32_t r = g - (uint32_t)z * (uint32_t)s->P;
        const int32_t y = (int32_t)(r / (uint32_t)s->W);
        const int32_t x = (int32_t)(r - (uint32_t)y * (uint32_t)s->W);
        const float w = val[j];
        if (o == acc.size()) {
            acc.emplace_back();
            Acc& a = acc.back();
            a.label = g;
            a.mx = w;
            a.b[0] = a.b[1] = x;
            a.b[2] = a.b[3] = y;
            a.b[4] = a.b[5] = z;
        }
        Acc& a = acc[o];
        a.n += 1;
        a.sx += (uint32_t)x;
        a.sy += (uint32_t)y;
        a.sz += (uint32_t)z;
        const float px = w * (float)x;
        const float py = w * (float)y;
        const float pz = w * (float)z;
        a.wx = a.wx + px;
        a.wy = a.wy + py;
        a.wz = a.wz + pz;
        a.ws = a.ws + w;
        if (a.mx < w) a.mx = w;
        a.sum = a.sum + (double)w;
        a.b[0] = std::min(a.b[0], x);
        a.b[1] = std::max(a.b[1], x);
        a.b[2] = std::min(a.b[2], y);
        a.b[3] = std::max(a.b[3], y);
        a.b[4] = std::min(a.b[4], z);
        a.b[5] = std::max(a.b[5], z);
    }
    out.resize(acc.size());
    for (size_t o = 0; o < acc.size(); ++o) {
        const Acc& a = acc[o];
        rihvr_object& r = out[o];
        std::memset(&r, 0, sizeof(r));
        const float fn = (float)a.n;
        r.centroid[0] = (float)(int32_t)a.sx / fn;
        r.centroid[1] = (float)(int32_t)a.sy / fn;
        r.centroid[2] = (float)(int32_t)a.sz / fn;
        r.weighted_centroid[0] = a.wx / a.ws;
        r.weighted_centroid[1] = a.wy / a.ws;
        r.weighted_centroid[2] = a.wz / a.ws;
        r.max_value = a.mx;
        r.mean_value = (float)a.sum / fn;
        r.num_voxels = a.n;
        r.label = a.label;
        std::memcpy(r.bounds, a.b, sizeof(a.b));
    }
    return RIHVR_OK;
}

rihvr_status copy_out(rihvr_segmentation* s, void* dst, const void* src, size_t bytes, int loc) {
    SEG_HIP(hipMemcpyAsync(dst, src, bytes, loc == RIHVR_DEVICE ? hipMemcpyDeviceToDevice : hipMemcpyDeviceToHost,
                           s->stream));
    SEG_HIP(hipStreamSynchronize(s->stream));
    return RIHVR_OK;
}

}  // namespace

extern "C" {

rihvr_status rihvr_segmentation_create(int32_t width, int32_t height, int32_t num_planes,
                                       rihvr_segmentation** out) {
    if (!out) return RIHVR_E_INVALID_ARGUMENT;
    *out = nullptr;
    if (width <= 0 || height <= 0 || num_planes <= 0) return RIHVR_E_INVALID_ARGUMENT;
    const long N = (long)width * height * num_planes;
    if (N >= (long)kNone) return RIHVR_E_INVALID_ARGUMENT;
    auto* s = new rihvr_segmentation();
    s->W = width;
    s->H = height;
    s->Nz = num_planes;
    s->P = (long)width * height;
    s->N = N;
    s->nblocks = (N + kCompactSpan - 1) / kCompactSpan;
    rihvr_status st = RIHVR_OK;
    auto fail = [&](hipError_t e) {
        std::fprintf(stderr, "rihvr: HIP error %s in rihvr_segmentation_create\n", hipGetErrorString(e));
        st = e == hipErrorOutOfMemory ? RIHVR_E_OUT_OF_MEMORY : RIHVR_E_DEVICE;
    };
    hipError_t e = hipGetDevice(&s->device);
    if (e == hipSuccess) e = hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking);
    if (e == hipSuccess) e = hipMalloc((void**)&s->key, (size_t)N * 4);
    if (e == hipSuccess) e = hipMalloc((void**)&s->mask, (size_t)N);
    if (e == hipSuccess) e = hipMalloc((void**)&s->work, (size_t)N * 4);
    if (e == hipSuccess) e = hipMalloc((void**)&s->lab, (size_t)N * 4);
    if (e == hipSuccess) e = hipMalloc((void**)&s->counts, (size_t)s->nblocks * 8);
    if (e == hipSuccess) e = hipMalloc((void**)&s->offsets, (size_t)s->nblocks * 8);
    if (e == hipSuccess) e = hipMalloc((void**)&s->small, 16 * sizeof(uint32_t));
    for (int t = 0; t < T_COUNT && e == hipSuccess; ++t)
        for (int k = 0; k < 2 && e == hipSuccess; ++k) e = hipEventCreate(&s->ev[t][k]);
    if (e != hipSuccess) {
        fail(e);
        rihvr_segmentation_destroy(s);
        return st;
    }
    *out = s;
    return RIHVR_OK;
}

void rihvr_segmentation_destroy(rihvr_segmentation* s) {
    if (!s) return;
    DeviceGuard g(s->device);
    if (s->stream) (void)hipStreamSynchronize(s->stream);
    for (void* p : {(void*)s->key, (void*)s->mask, (void*)s->work, (void*)s->lab, (void*)s->counts,
                    (void*)s->offsets, (void*)s->small, s->list})
        if (p) (void)hipFree(p);
    for (int t = 0; t < T_COUNT; ++t)
        for (int k = 0; k < 2; ++k)
            if (s->ev[t][k]) (void)hipEventDestroy(s->ev[t][k]);
    if (s->stream) (void)hipStreamDestroy(s->stream);
    delete s;
}

rihvr_status rihvr_segmentation_load(rihvr_segmentation* s, int64_t count, const int32_t* plane,
                                     const int32_t* index, const float* intensity, int loc) {
    if (!s || count < 0 || (count > 0 && (!plane || !index || !intensity))) return RIHVR_E_INVALID_ARGUMENT;
    DeviceGuard g(s->device);
    s->stage = EMPTY;
    std::fill(s->timed, s->timed + T_COUNT, false);
    rihvr_status st = begin_timer(s, T_LOAD);
    if (st) return st;
    SEG_HIP(hipMemsetAsync(s->key, 0, (size_t)s->N * 4, s->stream));
    SEG_HIP(hipMemsetAsync(s->small, 0, 16 * sizeof(uint32_t), s->stream));
    if (count > 0) {
        const int32_t* dp = plane;
        const int32_t* di = index;
        const float* dv = intensity;
        if (loc != RIHVR_DEVICE) {
            const size_t b = (size_t)count * 4;
            st = stage_buffer(s, 3 * b);
            if (st) return st;
            char* base = (char*)s->list;
            SEG_HIP(hipMemcpyAsync(base, plane, b, hipMemcpyHostToDevice, s->stream));
            SEG_HIP(hipMemcpyAsync(base + b, index, b, hipMemcpyHostToDevice, s->stream));
            SEG_HIP(hipMemcpyAsync(base + 2 * b, intensity, b, hipMemcpyHostToDevice, s->stream));
            dp = (const int32_t*)base;
            di = (const int32_t*)(base + b);
            dv = (const float*)(base + 2 * b);
        }
        hipLaunchKernelGGL(k_seg_scatter, dim3(grid_for(count)), dim3(kThreads), 0, s->stream, dp, di, dv,
                           (long)count, s->Nz, s->P, s->key, s->small, s->small + 1);
        SEG_HIP(hipGetLastError());
    }
    st = end_timer(s, T_LOAD);
    if (st) return st;
    return finish_load(s);
}

rihvr_status rihvr_segmentation_load_dense(rihvr_segmentation* s, const float* volume, int loc) {
    if (!s || !volume) return RIHVR_E_INVALID_ARGUMENT;
    DeviceGuard g(s->device);
    s->stage = EMPTY;
    std::fill(s->timed, s->timed + T_COUNT, false);
    const float* dv = volume;
    if (loc != RIHVR_DEVICE) {
        rihvr_status st = stage_buffer(s, (size_t)s->N * 8);
        if (st) return st;
        SEG_HIP(hipMemcpyAsync(s->list, volume, (size_t)s->N * 8, hipMemcpyHostToDevice, s->stream));
        dv = (const float*)s->list;
    }
    rihvr_status st = begin_timer(s, T_LOAD);
    if (st) return st;
    SEG_HIP(hipMemsetAsync(s->small, 0, 16 * sizeof(uint32_t), s->stream));
    hipLaunchKernelGGL(k_seg_dense, dim3(grid_for(s->N)), dim3(kThreads), 0, s->stream, (const float2*)dv, s->N,
                       s->key, s->small);
    SEG_HIP(hipGetLastError());
    st = end_timer(s, T_LOAD);
    if (st) return st;
    return finish_load(s);
}

rihvr_status rihvr_segmentation_binarize(rihvr_segmentation* s, double min_intensity, double* max_intensity,
                                         double* threshold) {
    if (!s) return RIHVR_E_INVALID_ARGUMENT;
    if (s->stage != LOADED) return RIHVR_E_INVALID_STATE;
    DeviceGuard g(s->device);
    uint32_t bits = 0;
    SEG_HIP(hipMemcpyAsync(&bits, s->small, sizeof(bits), hipMemcpyDeviceToHost, s->stream));
    SEG_HIP(hipStreamSynchronize(s->stream));
    float fmax;
    std::memcpy(&fmax, &bits, sizeof(fmax));
    const double mx = (double)fmax;
    const double thr = min_intensity * mx / 256.0;  // sparse_segmentation.cu:69-71
    rihvr_status st = begin_timer(s, T_BINARIZE);
    if (st) return st;
    hipLaunchKernelGGL(k_seg_binarize, dim3(grid_for(s->N)), dim3(kThreads), 0, s->stream, s->key, s->mask, s->N,
                       thr);
    SEG_HIP(hipGetLastError());
    st = end_timer(s, T_BINARIZE);
    if (st) return st;
    SEG_HIP(hipStreamSynchronize(s->stream));
    if (max_intensity) *max_intensity = mx;
    if (threshold) *threshold = thr;
    s->stage = BINARIZED;
    s->timed[T_CLOSE] = s->timed[T_LABEL] = s->timed[T_EXTRACT] = false;
    return RIHVR_OK;
}

rihvr_status rihvr_segmentation_close(rihvr_segmentation* s, int32_t close_size) {
    if (!s) return RIHVR_E_INVALID_ARGUMENT;
    if (s->stage != BINARIZED) return RIHVR_E_INVALID_STATE;
    if (close_size > kMaxClose) return RIHVR_E_INVALID_ARGUMENT;
    if (close_size <= 0) return RIHVR_OK;  // simpleMorph returns at once (:225)
    DeviceGuard g(s->device);
    const long rows = (long)s->H * s->Nz;
    const long segs = (s->W + kThreads - 1) / kThreads;
    const unsigned row_grid = (unsigned)std::min(rows * segs, 1024L * 1024L);
    uint8_t* dist = s->work;
    uint8_t* dil = s->work + s->N;
    const int cap = close_size + 1;
    rihvr_status st = begin_timer(s, T_CLOSE);
    if (st) return st;
    hipLaunchKernelGGL(k_seg_rowdist, dim3(row_grid), dim3(kThreads), 0, s->stream, s->mask, dist, s->W, rows, cap,
                       0);
    hipLaunchKernelGGL(k_seg_ball, dim3(grid_for(s->N)), dim3(kThreads), 0, s->stream, dist, dil, s->W, s->H, s->Nz,
                       close_size, 0);
    hipLaunchKernelGGL(k_seg_rowdist, dim3(row_grid), dim3(kThreads), 0, s->stream, dil, dist, s->W, rows, cap, 1);
    hipLaunchKernelGGL(k_seg_ball, dim3(grid_for(s->N)), dim3(kThreads), 0, s->stream, dist, s->mask, s->W, s->H,
                       s->Nz, close_size, 1);
    SEG_HIP(hipGetLastError());
    st = end_timer(s, T_CLOSE);
    if (st) return st;
    SEG_HIP(hipStreamSynchronize(s->stream));
    return RIHVR_OK;
}

rihvr_status rihvr_segmentation_label(rihvr_segmentation* s) {
    if (!s) return RIHVR_E_INVALID_ARGUMENT;
    if (s->stage != BINARIZED) return RIHVR_E_INVALID_STATE;
    DeviceGuard g(s->device);
    rihvr_status st = begin_timer(s, T_LABEL);
    if (st) return st;
    const unsigned grid = grid_for(s->N);
    hipLaunchKernelGGL(k_ccl_init, dim3(grid), dim3(kThreads), 0, s->stream, s->mask, s->lab, s->N);
    hipLaunchKernelGGL(k_ccl_merge, dim3(grid), dim3(kThreads), 0, s->stream, s->mask, s->lab, s->W, s->P, s->N);
    hipLaunchKernelGGL(k_ccl_flatten, dim3(grid), dim3(kThreads), 0, s->stream, s->lab, s->N);
    SEG_HIP(hipGetLastError());
    st = end_timer(s, T_LABEL);
    if (st) return st;
    SEG_HIP(hipStreamSynchronize(s->stream));
    s->stage = LABELED;
    return RIHVR_OK;
}

rihvr_status rihvr_segmentation_extract(rihvr_segmentation* s, int64_t* num_objects) {
    if (!s) return RIHVR_E_INVALID_ARGUMENT;
    if (s->stage != LABELED) return RIHVR_E_INVALID_STATE;
    DeviceGuard g(s->device);
    rihvr_status st = begin_timer(s, T_EXTRACT);
    if (st) return st;
    const unsigned nb = (unsigned)s->nblocks;
    hipLaunchKernelGGL(k_cmp_count, dim3(nb), dim3(kThreads), 0, s->stream, s->lab, s->N, s->counts);
    hipLaunchKernelGGL(k_cmp_scan, dim3(1), dim3(1024), 0, s->stream, s->counts, s->nblocks, s->offsets,
                       s->small + 2);
    uint32_t* rank = (uint32_t*)s->work;
    hipLaunchKernelGGL(k_cmp_roots, dim3(nb), dim3(kThreads), 0, s->stream, s->lab, s->N, s->offsets, rank);
    SEG_HIP(hipGetLastError());
    uint32_t totals[2] = {0, 0};
    SEG_HIP(hipMemcpyAsync(totals, s->small + 2, sizeof(totals), hipMemcpyDeviceToHost, s->stream));
    SEG_HIP(hipStreamSynchronize(s->stream));
    const size_t nvox = totals[0];
    if (nvox > 0) {
        st = stage_buffer(s, nvox * 12);
        if (st) return st;
        uint32_t* oi = (uint32_t*)s->list;
        uint32_t* oo = oi + nvox;
        float* ov = (float*)(oo + nvox);
        hipLaunchKernelGGL(k_cmp_scatter, dim3(nb), dim3(kThreads), 0, s->stream, s->lab, (const float*)s->key,
                           s->N, s->offsets, rank, oi, oo, ov);
        SEG_HIP(hipGetLastError());
    }
    st = end_timer(s, T_EXTRACT);
    if (st) return st;
    std::vector<uint32_t> gi(nvox), ob(nvox);
    std::vector<float> va(nvox);
    if (nvox > 0) {
        const uint32_t* oi = (const uint32_t*)s->list;
        SEG_HIP(hipMemcpyAsync(gi.data(), oi, nvox * 4, hipMemcpyDeviceToHost, s->stream));
        SEG_HIP(hipMemcpyAsync(ob.data(), oi + nvox, nvox * 4, hipMemcpyDeviceToHost, s->stream));
        SEG_HIP(hipMemcpyAsync(va.data(), oi + 2 * nvox, nvox * 4, hipMemcpyDeviceToHost, s->stream));
    }
    SEG_HIP(hipStreamSynchronize(s->stream));
    st = accumulate_objects(s, gi, ob, va, s->objects);
    if (st) return st;
    if (s->objects.size() != totals[1]) return RIHVR_E_UNKNOWN_ERROR;
    s->stage = EXTRACTED;
    if (num_objects) *num_objects = (int64_t)s->objects.size();
    return RIHVR_OK;
}

rihvr_status rihvr_segmentation_size_filter(rihvr_segmentation* s, int32_t min_area, int64_t* num_removed) {
    if (!s) return RIHVR_E_INVALID_ARGUMENT;
    if (s->stage != EXTRACTED) return RIHVR_E_INVALID_STATE;
    std::vector<rihvr_object>& o = s->objects;
    int64_t removed = 0;
    for (size_t n = 0; n < o.size();) {
        if (o[n].num_voxels < min_area) {
            o[n] = o.back();
            o.pop_back();
            ++removed;
        } else {
            ++n;
        }
    }
    if (num_removed) *num_removed = removed;
    return RIHVR_OK;
}

rihvr_status rihvr_segmentation_objects(rihvr_segmentation* s, rihvr_object* out, int64_t capacity,
                                        int64_t* num_objects) {
    if (!s || capacity < 0) return RIHVR_E_INVALID_ARGUMENT;
    if (s->stage != EXTRACTED) return RIHVR_E_INVALID_STATE;
    const int64_t n = (int64_t)s->objects.size();
    if (out)
        for (int64_t i = 0; i < std::min(n, capacity); ++i) out[i] = s->objects[i];
    if (num_objects) *num_objects = n;
    return RIHVR_OK;
}

rihvr_status rihvr_segmentation_write_centroids(rihvr_segmentation* s, const char* path, int32_t weighted,
                                                int32_t verbose) {
    if (!s || !path) return RIHVR_E_INVALID_ARGUMENT;
    if (s->stage != EXTRACTED) return RIHVR_E_INVALID_STATE;
    FILE* f = std::fopen(path, "w");
    if (!f) return RIHVR_E_BAD_FILENAME;
    if (!verbose) {
        std::fprintf(f, "x (pixels), y (pixels), z (planes), # of voxels\n");
    } else {
        std::fprintf(f, "x (pixels), y (pixels), z (planes), weighted x, weighted y, weighted z, # of voxels, "
                        "max value, mean value, xmin, xmax, ymin, ymax, zmin, zmax\n");
    }
    for (const rihvr_object& o : s->objects) {
        const int n = (int)o.num_voxels;  // printed with %d (object_cloud.cu:447)
        if (!verbose) {
            const float* c = weighted ? o.weighted_centroid : o.centroid;
            std::fprintf(f, "%f, %f, %f, %d\n", c[0], c[1], c[2], n);
        } else {
            std::fprintf(f, "%f, %f, %f, %f, %f, %f, %d, %e, %e, %d, %d, %d, %d, %d, %d\n", o.centroid[0],
                         o.centroid[1], o.centroid[2], o.weighted_centroid[0], o.weighted_centroid[1],
                         o.weighted_centroid[2], n, o.max_value, o.mean_value, o.bounds[0], o.bounds[1], o.bounds[2],
                         o.bounds[3], o.bounds[4], o.bounds[5]);
        }
    }
    return std::fclose(f) == 0 ? RIHVR_OK : RIHVR_E_BAD_FILENAME;
}

rihvr_status rihvr_segmentation_mask(rihvr_segmentation* s, uint8_t* out, int loc) {
    if (!s || !out) return RIHVR_E_INVALID_ARGUMENT;
    if (s->stage < BINARIZED) return RIHVR_E_INVALID_STATE;
    DeviceGuard g(s->device);
    return copy_out(s, out, s->mask, (size_t)s->N, loc);
}

rihvr_status rihvr_segmentation_values(rihvr_segmentation* s, float* out, int loc) {
    if (!s || !out) return RIHVR_E_INVALID_ARGUMENT;
    if (s->stage < BINARIZED) return RIHVR_E_INVALID_STATE;
    DeviceGuard g(s->device);
    return copy_out(s, out, s->key, (size_t)s->N * 4, loc);
}

rihvr_status rihvr_segmentation_labels(rihvr_segmentation* s, uint32_t* out, int loc) {
    if (!s || !out) return RIHVR_E_INVALID_ARGUMENT;
    if (s->stage < LABELED) return RIHVR_E_INVALID_STATE;
    DeviceGuard g(s->device);
    return copy_out(s, out, s->lab, (size_t)s->N * 4, loc);
}

rihvr_status rihvr_segmentation_timing(rihvr_segmentation* s, double* ms) {
    if (!s || !ms) return RIHVR_E_INVALID_ARGUMENT;
    DeviceGuard g(s->device);
    for (int t = 0; t < T_COUNT; ++t) {
        ms[t] = 0.0;
        if (!s->timed[t]) continue;
        float v = 0.0f;
        SEG_HIP(hipEventSynchronize(s->ev[t][1]));
        SEG_HIP(hipEventElapsedTime(&v, s->ev[t][0], s->ev[t][1]));
        ms[t] = v;
    }
    return RIHVR_OK;
}

}  // extern "C"
