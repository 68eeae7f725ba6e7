// segmentation.hip -- segmentation of a reconstructed volume into objects and their
// centroids: the chain of the reference's segmentation executable
// (src/core/main_segmentation.cu:83-125) over SparseSegmentation
// (sparse_segmentation.cu), ObjectCloud (object_cloud.cu) and Blob3d (blob3d.cu).
//
// The reference walks sparse per-plane lists on the host (bubble-sorted indices, a
// binary search per neighbour, one kernel launch per voxel for includeIntensity).
// Here every pass is a dense, HBM-streaming kernel over [Nz][H][W] buffers, with the
// binary volume bit-packed (64 voxels per word, one wave ballot per word):
//   load      scatter of the voxel list into a u32 key volume (atomicMax keeps the
//             larger intensity of a repeated voxel) + max intensity
//   binarize  key -> value (float, in place of the key) and the bit volume
//   close     ball dilation then erosion on 64-voxel words: the OR over the ball's rows
//             (dy, dz) of each row's x-dilation by w(dy, dz) = floor(sqrt(r^2-dy^2-dz^2))
//             (shift-or within the word and its two neighbours); the erosion is the
//             complement of the dilation of the complement with the outside set
//   label     union-find with atomicMin links to the smaller root (every root is the
//             smallest index of its tree), then a flattening pass; labels are touched
//             only at object voxels, so each voxel ends with the smallest global index
//             of its component
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
constexpr int kMaxClose = 63;                     // x-dilations reach one word to each side

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
                                                          int W, uint32_t* __restrict__ key,
                                                          uint32_t* __restrict__ rowflag,
                                                          uint32_t* __restrict__ maxbits,
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
        rowflag[((size_t)z * P + idx) / W] = 1u;  // the row holds a non-zero key
        if (v > 0.0f) m = max(m, __float_as_uint(v));  // v >= 0: bits order as values; NaN skipped
    }
    m = wave_max_u32(m);
    if ((threadIdx.x & 63) == 0 && m) atomicMax(maxbits, m);
}

// the same from a dense complex volume: intensity = re^2 + im^2 of the non-zero voxels
__device__ __forceinline__ void dense_voxel(float re, float im, long i, int W, uint32_t* __restrict__ key,
                                            uint32_t* __restrict__ rowflag, uint32_t& m) {
#pragma clang fp contract(off)  // re^2 + im^2 with two roundings, as the file's writer computes it
    if (re != 0.0f || im != 0.0f) {  // zero keys are already in place (k_seg_clear_rows)
        const float stored = re * re + im * im;
        const float v = reloaded_intensity(stored);
        key[i] = intensity_key(v);
        if (v > 0.0f) m = max(m, __float_as_uint(v));
        rowflag[i / W] = 1u;  // the row holds a non-zero key
    }
}

// PAIRS: two voxels per 16-byte load (the volume 16-byte aligned), the last voxel of an odd
// N by thread 0 of block 0; else one voxel per load
template <bool PAIRS>
__global__ void __launch_bounds__(kThreads) k_seg_dense(const float2* __restrict__ vol, long N, int W,
                                                        uint32_t* __restrict__ key, uint32_t* __restrict__ rowflag,
                                                        uint32_t* __restrict__ maxbits) {
    uint32_t m = 0;
    const long stride = (long)gridDim.x * kThreads;
    if constexpr (PAIRS) {
        const float4* v4 = (const float4*)vol;
        for (long q = (long)blockIdx.x * kThreads + threadIdx.x; q < N / 2; q += stride) {
            const float4 c = v4[q];
            dense_voxel(c.x, c.y, 2 * q, W, key, rowflag, m);
            dense_voxel(c.z, c.w, 2 * q + 1, W, key, rowflag, m);
        }
        if ((N & 1) && blockIdx.x == 0 && threadIdx.x == 0) {
            const float2 c = vol[N - 1];
            dense_voxel(c.x, c.y, N - 1, W, key, rowflag, m);
        }
    } else {
        for (long i = (long)blockIdx.x * kThreads + threadIdx.x; i < N; i += stride) {
            const float2 c = vol[i];
            dense_voxel(c.x, c.y, i, W, key, rowflag, m);
        }
    }
    m = wave_max_u32(m);
    if ((threadIdx.x & 63) == 0 && m) atomicMax(maxbits, m);
}

// Between loads the key array is zero outside the rows the last load flagged (binarize
// rewrites only non-zero keys, and nothing else writes keys), so a load first zeroes the
// keys of the flagged rows and clears their flags, then writes only its non-zero keys.
// One wave per row at a time.
__global__ void __launch_bounds__(kThreads) k_seg_clear_rows(uint32_t* __restrict__ key,
                                                             uint32_t* __restrict__ rowflag, int W, int rows) {
    const int lane = threadIdx.x & 63;
    const int waves = gridDim.x * (kThreads / 64);
    for (int row = blockIdx.x * (kThreads / 64) + (threadIdx.x >> 6); row < rows; row += waves) {
        if (rowflag[row] == 0u) continue;  // wave-uniform
        uint32_t* kr = key + (size_t)row * W;
        for (int x = lane; x < W; x += 64) kr[x] = 0u;
        if (lane == 0) rowflag[row] = 0u;
    }
}

__device__ __forceinline__ int isqrt_floor(int n) {
    int r = (int)sqrtf((float)n);
    while (r * r > n) --r;
    while ((r + 1) * (r + 1) <= n) ++r;
    return r;
}

// ---------------------------------------------------------------------------------
// Binary volumes are bit-packed: row r = z*H + y holds R = ceil(W/64) 64-bit words,
// bit j of word k is voxel x = 64k + j; bits past x = W - 1 are zero.  A wave of 64
// lanes produces one word with a ballot.
// ---------------------------------------------------------------------------------
__device__ __forceinline__ uint64_t valid_bits(int wi, int R, int W) {
    const int tail = W & 63;
    return (wi < R - 1 || tail == 0) ? ~0ull : ((1ull << tail) - 1ull);
}

// binary = intensity > thr (float compared in double, binarize_kernel :39-50); the key
// becomes the value of the masked volume (mask_kernel :52-61): intensity or 0.  One wave
// per row at a time, kUnroll 64-bit words per step (their loads issued together).  A row
// whose flag says that all its keys are 0 only gets zero words when skip_empty (the host
// sets it when 0 > thr is false, so such a row has no voxel above the threshold and no key
// to rewrite).
constexpr int kUnroll = 8;
__global__ void __launch_bounds__(kThreads) k_seg_binarize(uint32_t* __restrict__ key_value_bits,
                                                           uint64_t* __restrict__ bits,
                                                           const uint32_t* __restrict__ rowflag,
                                                           uint32_t* __restrict__ rowbits, int W, int R,
                                                           int rows, double thr, int skip_empty) {
    const int lane = threadIdx.x & 63;
    const int waves = gridDim.x * (kThreads / 64);
    for (int row = blockIdx.x * (kThreads / 64) + (threadIdx.x >> 6); row < rows; row += waves) {
        uint32_t* kr = key_value_bits + (size_t)row * W;
        uint64_t* br = bits + (size_t)row * R;
        if (skip_empty && rowflag[row] == 0u) {  // wave-uniform: one row per wave
            for (int w = lane; w < R; w += 64) br[w] = 0ull;
            if (lane == 0) rowbits[row] = 0u;
            continue;
        }
        uint64_t any = 0;  // wave-uniform: OR of the row's words
        for (int w0 = 0; w0 < R; w0 += kUnroll) {
            uint32_t k[kUnroll];
#pragma unroll
            for (int u = 0; u < kUnroll; ++u) {
                const int x = (w0 + u) * 64 + lane;
                k[u] = (w0 + u < R && x < W) ? kr[x] : 0u;
            }
#pragma unroll
            for (int u = 0; u < kUnroll; ++u) {
                const float v = key_value(k[u]);
                const int x = (w0 + u) * 64 + lane;
                const bool b = x < W && (double)v > thr;
                if (k[u]) kr[x] = b ? __float_as_uint(v) : 0u;  // key 0 is already value 0
                const uint64_t word = __ballot(b);
                if (lane == 0 && w0 + u < R) br[w0 + u] = word;
                any |= word;
            }
        }
        if (lane == 0) rowbits[row] = any != 0ull ? 1u : 0u;
    }
}

// x-dilation of the middle word of (prev, cur, next) by w <= 63 voxels
__device__ __forceinline__ uint64_t xdilate(uint64_t prev, uint64_t cur, uint64_t next, int w) {
    uint64_t r = cur;
    for (int k = 1; k <= w; ++k) r |= (cur >> k) | (next << (64 - k)) | (cur << k) | (prev >> (64 - k));
    return r;
}

// One morphology step of the closing on 64-voxel words (one thread per word): the OR over
// the ball's rows (dy, dz) of the row's x-dilation by w(dy, dz) = floor(sqrt(r^2 - dy^2 -
// dz^2)).  Dilation (complement = 0): rows and words outside the volume are empty.
// Erosion (complement = 1) = NOT dilation of (NOT src with everything outside set): a
// voxel survives only if its whole ball lies inside the volume and inside src.  rowbits
// (dilation only, may be NULL): per row of src, 0 if all its words are 0 (such rows add
// nothing and are not read).
__global__ void __launch_bounds__(kThreads) k_seg_morph(const uint64_t* __restrict__ src,
                                                        uint64_t* __restrict__ dst,
                                                        const uint32_t* __restrict__ rowbits, int W, int R, int H,
                                                        int Nz, int r, int complement) {
    const uint32_t nwords = (uint32_t)R * H * Nz;
    const uint64_t flip = complement ? ~0ull : 0ull;  // src words are read as src ^ flip
    for (uint32_t w = blockIdx.x * kThreads + threadIdx.x; w < nwords; w += gridDim.x * kThreads) {
        const uint32_t row = w / (uint32_t)R;
        const int wi = (int)(w - row * (uint32_t)R);
        const int z = (int)(row / (uint32_t)H);
        const int y = (int)(row - (uint32_t)z * H);
        uint64_t acc = 0;
        for (int dz = -r; dz <= r && acc != ~0ull; ++dz) {
            const int zz = z + dz;
            const int rz = r * r - dz * dz;
            const int wy = isqrt_floor(rz);
            if (zz < 0 || zz >= Nz) {
                acc |= flip;  // a whole plane of the ball outside the volume
                continue;
            }
            for (int dy = -wy; dy <= wy; ++dy) {
                const int yy = y + dy;
                if (yy < 0 || yy >= H) {
                    acc |= flip;
                    continue;
                }
                if (rowbits && rowbits[(size_t)zz * H + yy] == 0u) continue;
                const uint64_t* rowp = src + ((size_t)zz * H + yy) * R;
                const uint64_t cur = rowp[wi] ^ flip;
                const uint64_t prev = wi > 0 ? rowp[wi - 1] ^ flip : flip;
                const uint64_t next = wi + 1 < R ? rowp[wi + 1] ^ flip : flip;
                if ((prev | cur | next) != 0ull) acc |= xdilate(prev, cur, next, isqrt_floor(rz - dy * dy));
            }
        }
        dst[w] = (acc ^ flip) & valid_bits(wi, R, W);
    }
}

__device__ __forceinline__ bool test_bit(const uint64_t* bits, int R, long row, int x) {
    return (bits[row * R + (x >> 6)] >> (x & 63)) & 1ull;
}

// --- ordered compaction -----------------------------------------------------
__device__ __forceinline__ uint32_t block_exclusive_scan(uint32_t v, uint32_t* lds, uint32_t* total) {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    uint32_t x = v;
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
        const uint32_t y = (uint32_t)__shfl_up((int)x, off, 64);
        if (lane >= off) x += y;
    }
    if (lane == 63) lds[wave] = x;
    __syncthreads();
    uint32_t before = 0, all = 0;
    for (int k = 0; k < kThreads / 64; ++k) {
        const uint32_t c = lds[k];
        before += k < wave ? c : 0;
        all += c;
    }
    __syncthreads();
    *total = all;
    return before + x - v;
}

// object voxels per block of 256 words (thread t: word b*256 + t)
__global__ void __launch_bounds__(kThreads) k_list_count(const uint64_t* __restrict__ bits, uint32_t nwords,
                                                         uint32_t* __restrict__ counts) {
    __shared__ uint32_t lds[kThreads / 64];
    const uint32_t w = blockIdx.x * kThreads + threadIdx.x;
    const uint64_t m = w < nwords ? bits[w] : 0ull;
    uint32_t all;
    block_exclusive_scan((uint32_t)__popcll(m), lds, &all);
    if (threadIdx.x == 0) counts[blockIdx.x] = all;
}

// exclusive scan of n block counts in one workgroup of 1024 threads; *total = the sum
__global__ void __launch_bounds__(1024) k_scan(const uint32_t* __restrict__ counts, uint32_t n,
                                               uint32_t* __restrict__ offsets, uint32_t* __restrict__ total) {
    __shared__ uint32_t sa[1024];
    const uint32_t per = (n + 1023) / 1024;
    const uint32_t b0 = threadIdx.x * per;
    const uint32_t b1 = min(n, b0 + per);
    uint32_t a = 0;
    for (uint32_t j = b0; j < b1; ++j) a += counts[j];
    sa[threadIdx.x] = a;
    __syncthreads();
    for (int off = 1; off < 1024; off <<= 1) {  // inclusive Hillis-Steele scan
        const uint32_t xa = threadIdx.x >= (unsigned)off ? sa[threadIdx.x - off] : 0u;
        __syncthreads();
        sa[threadIdx.x] += xa;
        __syncthreads();
    }
    uint32_t ra = sa[threadIdx.x] - a;
    for (uint32_t j = b0; j < b1; ++j) {
        offsets[j] = ra;
        ra += counts[j];
    }
    if (threadIdx.x == 1023) *total = sa[1023];
}

// the object voxels' global indices in raster order (thread order, then bit order)
__global__ void __launch_bounds__(kThreads) k_list_scatter(const uint64_t* __restrict__ bits, int W, int R,
                                                           uint32_t nwords, const uint32_t* __restrict__ offsets,
                                                           uint32_t* __restrict__ list) {
    __shared__ uint32_t lds[kThreads / 64];
    const uint32_t w = blockIdx.x * kThreads + threadIdx.x;
    uint64_t m = w < nwords ? bits[w] : 0ull;
    uint32_t all;
    uint32_t pos = offsets[blockIdx.x] + block_exclusive_scan((uint32_t)__popcll(m), lds, &all);
    if (m) {
        const uint32_t row = w / (uint32_t)R;
        const uint32_t base = row * (uint32_t)W + (w - row * (uint32_t)R) * 64u;
        while (m) {
            const int j = __builtin_ctzll(m);
            m &= m - 1ull;
            list[pos++] = base + (uint32_t)j;
        }
    }
}

// --- connected components ---------------------------------------------------
// One thread per object voxel of the compacted list; labels are read and written only at
// object voxels.
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

__global__ void __launch_bounds__(kThreads) k_ccl_init(const uint32_t* __restrict__ list, uint32_t n,
                                                       uint32_t* __restrict__ lab) {
    for (uint32_t i = blockIdx.x * kThreads + threadIdx.x; i < n; i += gridDim.x * kThreads) lab[list[i]] = list[i];
}

// Each object voxel unites with its 13 earlier neighbours (the other 13 unite with it):
// in-plane flat offsets dy*W + dx kept inside 0..W*H-1 (labelConnectedComponents'
// offsets, :452-480), so x = -1 is x = W-1 of the row above and x = W is x = 0 of the
// row below; on the previous plane all nine, on the same plane the four below 0.
// Sixteen lanes per voxel, lane k < 13 taking neighbour k (dz = -1: k = 0..8 for
// (dy, dx) in raster order; dz = 0: k = 9..12 for (-1,-1), (-1,0), (-1,1), (0,-1)).
__global__ void __launch_bounds__(kThreads) k_ccl_merge(const uint32_t* __restrict__ list, uint32_t n,
                                                        const uint64_t* __restrict__ bits, uint32_t* lab, int W,
                                                        int R, int H) {
    const long total = (long)n * 16;
    for (long t = (long)blockIdx.x * kThreads + threadIdx.x; t < total; t += (long)gridDim.x * kThreads) {
        const uint32_t k = (uint32_t)(t & 15);
        if (k >= 13u) continue;
        const uint32_t v = list[t >> 4];
        const uint32_t row = v / (uint32_t)W;
        const int x = (int)(v - row * (uint32_t)W);
        const int z = (int)(row / (uint32_t)H);
        const int y = (int)(row - (uint32_t)z * H);
        const int dz = k < 9u ? -1 : 0;
        const int q = k < 9u ? (int)k : (int)k - 9;
        const int dy = q / 3 - 1, dx = q % 3 - 1;
        if (z + dz < 0) continue;
        int yy = y + dy, xx = x + dx;
        if (xx < 0) {
            xx = W - 1;
            --yy;
        } else if (xx >= W) {
            xx = 0;
            ++yy;
        }
        if (yy < 0 || yy >= H) continue;  // flat index outside the plane
        const long nrow = (long)(z + dz) * H + yy;
        if (test_bit(bits, R, nrow, xx)) unite(lab, v, (uint32_t)(nrow * W + xx));
    }
}

__global__ void __launch_bounds__(kThreads) k_ccl_flatten(const uint32_t* __restrict__ list, uint32_t n,
                                                          uint32_t* lab) {
    for (uint32_t i = blockIdx.x * kThreads + threadIdx.x; i < n; i += gridDim.x * kThreads) {
        const uint32_t v = list[i];
        const uint32_t l = lab[v];
        if (l != v) lab[v] = find_root(lab, l);
    }
}

// --- objects ----------------------------------------------------------------
// roots (label == own index) per block of 256 list entries
__global__ void __launch_bounds__(kThreads) k_root_count(const uint32_t* __restrict__ list, uint32_t n,
                                                         const uint32_t* __restrict__ lab,
                                                         uint32_t* __restrict__ counts) {
    __shared__ uint32_t lds[kThreads / 64];
    const uint32_t i = blockIdx.x * kThreads + threadIdx.x;
    const bool f = i < n && lab[list[i]] == list[i];
    uint32_t all;
    block_exclusive_scan(f ? 1u : 0u, lds, &all);
    if (threadIdx.x == 0) counts[blockIdx.x] = all;
}

// rank of every root among the roots in list (raster) order: its object number
__global__ void __launch_bounds__(kThreads) k_root_rank(const uint32_t* __restrict__ list, uint32_t n,
                                                        const uint32_t* __restrict__ lab,
                                                        const uint32_t* __restrict__ offsets,
                                                        uint32_t* __restrict__ rank) {
    __shared__ uint32_t lds[kThreads / 64];
    const uint32_t i = blockIdx.x * kThreads + threadIdx.x;
    const uint32_t v = i < n ? list[i] : 0u;
    const bool f = i < n && lab[v] == v;
    uint32_t all;
    const uint32_t pos = offsets[blockIdx.x] + block_exclusive_scan(f ? 1u : 0u, lds, &all);
    if (f) rank[v] = pos;
}

// object number and value of every list entry
__global__ void __launch_bounds__(kThreads) k_object_of(const uint32_t* __restrict__ list, uint32_t n,
                                                        const uint32_t* __restrict__ lab,
                                                        const uint32_t* __restrict__ rank,
                                                        const float* __restrict__ value,
                                                        uint32_t* __restrict__ out_object,
                                                        float* __restrict__ out_value) {
    for (uint32_t i = blockIdx.x * kThreads + threadIdx.x; i < n; i += gridDim.x * kThreads) {
        const uint32_t v = list[i];
        out_object[i] = rank[lab[v]];
        out_value[i] = value[v];
    }
}

// dense copies of the state for the caller: the bit volume as bytes, the labels with
// 0xFFFFFFFF off the objects
__global__ void __launch_bounds__(kThreads) k_unpack(const uint64_t* __restrict__ bits,
                                                     const uint32_t* __restrict__ lab, int W, int R, long N,
                                                     uint8_t* __restrict__ mask_out, uint32_t* __restrict__ lab_out) {
    for (long i = (long)blockIdx.x * kThreads + threadIdx.x; i < N; i += (long)gridDim.x * kThreads) {
        const long row = i / W;
        const int x = (int)(i - row * W);
        const bool b = test_bit(bits, R, row, x);
        if (mask_out) mask_out[i] = b ? 1 : 0;
        if (lab_out) lab_out[i] = b ? lab[i] : kNone;
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
    int R = 0;                   // 64-bit words per row of a bit volume
    long nwords = 0;
    uint32_t* key = nullptr;     // keys, then values (float bits)
    uint64_t* bits = nullptr;    // binary volume, bit-packed rows
    uint64_t* bits2 = nullptr;   // the dilation during the closing
    uint32_t* rowflag = nullptr; // per row (z * H + y) of the volume: 1 if a key in it is non-zero
    uint32_t* rowbits = nullptr; // per row of the binarized bit volume: 1 if a bit in it is set
    bool rowbits_valid = false;  // rowbits describes `bits` (from binarize until the first closing)
    uint32_t* work = nullptr;    // N words: root ranks, or dense copies for the caller
    uint32_t* lab = nullptr;     // labels (object voxels only)
    uint32_t* counts = nullptr;  // per scan block
    uint32_t* offsets = nullptr;
    uint32_t* small = nullptr;   // [0] max bits, [1] bad entries, [2] object voxels, [3] objects
    long nblocks = 0;            // capacity of counts / offsets
    void* list = nullptr;        // staging of a host voxel list or volume; then the object voxels
    size_t list_bytes = 0;
    uint32_t n_list = 0;         // object voxels in the compacted list (after label)
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

hipError_t clear_rows(rihvr_segmentation* s) {
    const int rows = s->H * s->Nz;
    hipLaunchKernelGGL(k_seg_clear_rows, dim3((unsigned)std::min(rows / 4 + 1, 64 * 1024)), dim3(kThreads), 0,
                       s->stream, s->key, s->rowflag, s->W, rows);
    return hipGetLastError();
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
    s->R = (width + 63) / 64;
    s->nwords = (long)s->R * height * num_planes;
    s->nblocks = std::max(s->nwords, N) / kThreads + 1;
    rihvr_status st = RIHVR_OK;
    auto fail = [&](hipError_t e) {
        std::fprintf(stderr, "rihvr: HIP error %s in rihvr_segmentation_create\n", hipGetErrorString(e));
        st = e == hipErrorOutOfMemory ? RIHVR_E_OUT_OF_MEMORY : RIHVR_E_DEVICE;
    };
    hipError_t e = hipGetDevice(&s->device);
    if (e == hipSuccess) e = hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking);
    if (e == hipSuccess) e = hipMalloc((void**)&s->key, (size_t)N * 4);
    if (e == hipSuccess) e = hipMalloc((void**)&s->bits, (size_t)s->nwords * 8);
    if (e == hipSuccess) e = hipMalloc((void**)&s->bits2, (size_t)s->nwords * 8);
    if (e == hipSuccess) e = hipMalloc((void**)&s->rowflag, (size_t)height * num_planes * 4);
    if (e == hipSuccess) e = hipMalloc((void**)&s->rowbits, (size_t)height * num_planes * 4);
    if (e == hipSuccess) e = hipMalloc((void**)&s->work, (size_t)N * 4);
    if (e == hipSuccess) e = hipMalloc((void**)&s->lab, (size_t)N * 4);
    if (e == hipSuccess) e = hipMalloc((void**)&s->counts, (size_t)s->nblocks * 4);
    if (e == hipSuccess) e = hipMalloc((void**)&s->offsets, (size_t)s->nblocks * 4);
    if (e == hipSuccess) e = hipMalloc((void**)&s->small, 16 * sizeof(uint32_t));
    for (int t = 0; t < T_COUNT && e == hipSuccess; ++t)
        for (int k = 0; k < 2 && e == hipSuccess; ++k) e = hipEventCreate(&s->ev[t][k]);
    // the loads rely on zero keys outside the flagged rows
    if (e == hipSuccess) e = hipMemsetAsync(s->key, 0, (size_t)N * 4, s->stream);
    if (e == hipSuccess) e = hipMemsetAsync(s->rowflag, 0, (size_t)height * num_planes * 4, s->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(s->stream);
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
    for (void* p : {(void*)s->key, (void*)s->bits, (void*)s->bits2, (void*)s->rowflag, (void*)s->rowbits,
                    (void*)s->work, (void*)s->lab, (void*)s->counts, (void*)s->offsets, (void*)s->small, s->list})
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
    SEG_HIP(clear_rows(s));
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
                           (long)count, s->Nz, s->P, s->W, s->key, s->rowflag, s->small, s->small + 1);
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
    SEG_HIP(clear_rows(s));
    if ((size_t)dv % 16 == 0)
        hipLaunchKernelGGL(k_seg_dense<true>, dim3(grid_for(s->N / 2)), dim3(kThreads), 0, s->stream,
                           (const float2*)dv, s->N, s->W, s->key, s->rowflag, s->small);
    else
        hipLaunchKernelGGL(k_seg_dense<false>, dim3(grid_for(s->N)), dim3(kThreads), 0, s->stream,
                           (const float2*)dv, s->N, s->W, s->key, s->rowflag, s->small);
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
    const int rows = s->H * s->Nz;
    hipLaunchKernelGGL(k_seg_binarize, dim3((unsigned)std::min(rows / 4 + 1, 64 * 1024)), dim3(kThreads), 0,
                       s->stream, s->key, s->bits, s->rowflag, s->rowbits, s->W, s->R, rows, thr,
                       !(0.0 > thr) ? 1 : 0);
    SEG_HIP(hipGetLastError());
    st = end_timer(s, T_BINARIZE);
    if (st) return st;
    SEG_HIP(hipStreamSynchronize(s->stream));
    if (max_intensity) *max_intensity = mx;
    if (threshold) *threshold = thr;
    s->stage = BINARIZED;
    s->rowbits_valid = true;
    s->timed[T_CLOSE] = s->timed[T_LABEL] = s->timed[T_EXTRACT] = false;
    return RIHVR_OK;
}

rihvr_status rihvr_segmentation_close(rihvr_segmentation* s, int32_t close_size) {
    if (!s) return RIHVR_E_INVALID_ARGUMENT;
    if (s->stage != BINARIZED) return RIHVR_E_INVALID_STATE;
    if (close_size > kMaxClose) return RIHVR_E_INVALID_ARGUMENT;
    if (close_size <= 0) return RIHVR_OK;  // simpleMorph returns at once (:225)
    DeviceGuard g(s->device);
    rihvr_status st = begin_timer(s, T_CLOSE);
    if (st) return st;
    const unsigned grid = grid_for(s->nwords);
    // a closing can set bits in rows that were empty: the row flags hold for the first one only
    const uint32_t* rb = s->rowbits_valid ? s->rowbits : nullptr;
    s->rowbits_valid = false;
    hipLaunchKernelGGL(k_seg_morph, dim3(grid), dim3(kThreads), 0, s->stream, s->bits, s->bits2, rb, s->W, s->R,
                       s->H, s->Nz, close_size, 0);
    hipLaunchKernelGGL(k_seg_morph, dim3(grid), dim3(kThreads), 0, s->stream, s->bits2, s->bits,
                       (const uint32_t*)nullptr, s->W, s->R, s->H, s->Nz, close_size, 1);
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
    // the object voxels in raster order
    const uint32_t nw = (uint32_t)s->nwords;
    const uint32_t nbw = (nw + kThreads - 1) / kThreads;
    hipLaunchKernelGGL(k_list_count, dim3(nbw), dim3(kThreads), 0, s->stream, s->bits, nw, s->counts);
    hipLaunchKernelGGL(k_scan, dim3(1), dim3(1024), 0, s->stream, s->counts, nbw, s->offsets, s->small + 2);
    SEG_HIP(hipGetLastError());
    uint32_t n = 0;
    SEG_HIP(hipMemcpyAsync(&n, s->small + 2, sizeof(n), hipMemcpyDeviceToHost, s->stream));
    SEG_HIP(hipStreamSynchronize(s->stream));
    if (n > 0) {
        st = stage_buffer(s, (size_t)n * 12);  // list, object numbers, values
        if (st) return st;
        uint32_t* list = (uint32_t*)s->list;
        hipLaunchKernelGGL(k_list_scatter, dim3(nbw), dim3(kThreads), 0, s->stream, s->bits, s->W, s->R, nw,
                           s->offsets, list);
        const unsigned grid = grid_for(n);
        hipLaunchKernelGGL(k_ccl_init, dim3(grid), dim3(kThreads), 0, s->stream, list, n, s->lab);
        hipLaunchKernelGGL(k_ccl_merge, dim3(grid_for((long)n * 16)), dim3(kThreads), 0, s->stream, list, n, s->bits,
                           s->lab, s->W, s->R, s->H);
        hipLaunchKernelGGL(k_ccl_flatten, dim3(grid), dim3(kThreads), 0, s->stream, list, n, s->lab);
        SEG_HIP(hipGetLastError());
    }
    st = end_timer(s, T_LABEL);
    if (st) return st;
    SEG_HIP(hipStreamSynchronize(s->stream));
    s->n_list = n;
    s->stage = LABELED;
    return RIHVR_OK;
}

rihvr_status rihvr_segmentation_extract(rihvr_segmentation* s, int64_t* num_objects) {
    if (!s) return RIHVR_E_INVALID_ARGUMENT;
    if (s->stage != LABELED) return RIHVR_E_INVALID_STATE;
    DeviceGuard g(s->device);
    rihvr_status st = begin_timer(s, T_EXTRACT);
    if (st) return st;
    const uint32_t n = s->n_list;
    const size_t nvox = n;
    uint32_t n_roots = 0;
    if (n > 0) {
        const uint32_t* list = (const uint32_t*)s->list;
        uint32_t* oo = (uint32_t*)s->list + nvox;
        float* ov = (float*)(oo + nvox);
        const uint32_t nb = (n + kThreads - 1) / kThreads;
        uint32_t* rank = s->work;
        hipLaunchKernelGGL(k_root_count, dim3(nb), dim3(kThreads), 0, s->stream, list, n, s->lab, s->counts);
        hipLaunchKernelGGL(k_scan, dim3(1), dim3(1024), 0, s->stream, s->counts, nb, s->offsets, s->small + 3);
        hipLaunchKernelGGL(k_root_rank, dim3(nb), dim3(kThreads), 0, s->stream, list, n, s->lab, s->offsets, rank);
        hipLaunchKernelGGL(k_object_of, dim3(grid_for(n)), dim3(kThreads), 0, s->stream, list, n, s->lab, rank,
                           (const float*)s->key, oo, ov);
        SEG_HIP(hipGetLastError());
        SEG_HIP(hipMemcpyAsync(&n_roots, s->small + 3, sizeof(n_roots), hipMemcpyDeviceToHost, s->stream));
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
    if (s->objects.size() != n_roots) return RIHVR_E_UNKNOWN_ERROR;
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
    hipLaunchKernelGGL(k_unpack, dim3(grid_for(s->N)), dim3(kThreads), 0, s->stream, s->bits, s->lab, s->W, s->R,
                       s->N, (uint8_t*)s->work, (uint32_t*)nullptr);
    SEG_HIP(hipGetLastError());
    return copy_out(s, out, s->work, (size_t)s->N, loc);
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
    hipLaunchKernelGGL(k_unpack, dim3(grid_for(s->N)), dim3(kThreads), 0, s->stream, s->bits, s->lab, s->W, s->R,
                       s->N, (uint8_t*)nullptr, s->work);
    SEG_HIP(hipGetLastError());
    return copy_out(s, out, s->work, (size_t)s->N * 4, loc);
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
