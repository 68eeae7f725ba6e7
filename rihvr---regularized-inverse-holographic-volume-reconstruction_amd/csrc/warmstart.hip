// warmstart.hip -- the learned warm start: a 3D encoder-decoder network that maps the
// back-propagated volume A^T b of a hologram to a start volume for FISTA, with its BF16
// training and inference kernels for gfx950.
//
// Network (4 levels, 16-32-64-128 channels, 3x3x3 convolutions, zero padding):
//   x16 = in_scale * (Re, Im) of A^T b, as 16 BF16 channels (channels 2..15 zero)
//   e0 = relu(conv(x16))          16   level 0
//   e1 = relu(conv(pool(e0)))     32   level 1      pool = 2x2x2 average
//   e2 = relu(conv(pool(e1)))     64   level 2
//   e3 = relu(conv(pool(e2)))    128   level 3
//   d2 = relu(conv(e2 ++ up(e3)))  64  level 2      up = nearest 2x2x2, ++ = channel concat
//   d1 = relu(conv(e1 ++ up(d2)))  32  level 1
//   d0 = relu(conv(e0 ++ up(d1)))  16  level 0
//   out = out_scale * (Wo d0 + bo)  (1x1x1, 16 -> 2 = Re, Im of the start volume)
// Activations are BF16, channels-last [z][y][x][c]; every 3x3x3 convolution is an implicit
// GEMM on v_mfma_f32_16x16x16_bf16 with FP32 accumulation (M = 16 voxels, N = 16 output
// channels, K = 16 input channels of one tap).  The backward pass uses the same forward kernel
// for the input gradient (taps flipped, weights transposed) and an MFMA weight-gradient kernel
// (K = voxels) whose per-chunk partials are summed in a fixed order.  Loss: mean squared error
// against the solver's converged volume; optimiser: Adam on FP32 master weights.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "../../include/rihvr.h"
#include "warmstart.hpp"

namespace rihvr {
namespace net {

typedef short s4 __attribute__((ext_vector_type(4)));
typedef float f4 __attribute__((ext_vector_type(4)));

constexpr int kLevels = 4;
constexpr int kNumConv = 7;  // e0 e1 e2 e3 d2 d1 d0
constexpr int kTaps = 27;
constexpr int kFwdThreads = 256;   // 4 waves, one output row of 64 voxels each
constexpr int kMaxCB = 64;         // output channels per workgroup
constexpr int kFirstZ = 8;         // planes per workgroup of the input layer
constexpr int kEw = 256;           // elementwise block size
constexpr float kB1 = 0.9f, kB2 = 0.999f, kEps = 1e-8f;

__device__ __forceinline__ float bf2f(unsigned short h) { return __uint_as_float((unsigned)h << 16); }
__device__ __forceinline__ unsigned short f2bf(float f) {
    unsigned u = __float_as_uint(f);
    if ((u & 0x7fffffffu) > 0x7f800000u) return (unsigned short)((u >> 16) | 0x40u);  // quiet NaN
    u += 0x7fffu + ((u >> 16) & 1u);  // round to nearest even
    return (unsigned short)(u >> 16);
}

struct ConvGeo {
    int D, H, W;  // full-resolution grid of the output and of in0
    int C0, C1;   // channels of in0 (full resolution) and of in1 (half resolution, read upsampled)
    int Cout;
    int CB;       // output channels per workgroup (16, 32, 48 or 64; divides Cout)
};

// A operand: 4 input channels (c .. c+3 of the concatenated input) of voxel (z, y, x), zero
// outside the grid.  Channels below C0 come from in0, the others from in1 at (z/2, y/2, x/2).
__device__ __forceinline__ s4 load_in4(const unsigned short* __restrict__ in0, const unsigned short* __restrict__ in1,
                                       const ConvGeo& g, int z, int y, int x, int c) {
    s4 v = {0, 0, 0, 0};
    if (z < 0 || z >= g.D || y < 0 || y >= g.H || x < 0 || x >= g.W) return v;
    if (c < g.C0) {
        v = *(const s4*)(in0 + (((size_t)z * g.H + y) * g.W + x) * g.C0 + c);
    } else {
        const size_t h = (((size_t)(z >> 1) * (g.H >> 1) + (y >> 1)) * (g.W >> 1) + (x >> 1));
        v = *(const s4*)(in1 + h * g.C1 + (c - g.C0));
    }
    return v;
}

// out[v][co] = act(bias[co] + sum_tap sum_ci in[v + tap][ci] * w[tap][co][ci]).
// A workgroup computes a tile of 4 ROWS rows x 64 columns of one plane for CB output channels:
// wave w takes rows y0 + w ROWS .. + ROWS - 1, each as 4 M-tiles of 16 voxels, CB / 16 N-tiles.
// Per 16-channel chunk of the input and per dz, the (4 ROWS + 2) x 66 voxel halo of the tile in
// plane z + dz is staged in LDS (coalesced 16-byte loads, zero outside the grid; in1's channels
// read at half resolution) with the weights of that dz slab (9 taps), so the 9 taps read their A
// and B fragments from LDS.  ROWS = 2 (for CB = 16, whose stages hold little MFMA work) doubles
// the work per stage.  Grid: (ceil(W/64) * Cout/CB, ceil(H/(4 ROWS)), D).
// OUTF32: raw FP32 sums (input gradient), else BF16 with bias and optional ReLU.
// (conv_forward runs this kernel with ROWS = 1; two rows per wave run k_conv_dma below.)
constexpr int kHX = 66;  // halo columns: 64 + 2
constexpr int kHX2 = 34;  // half-resolution halo columns: 32 + 2
template <bool OUTF32, int ROWS>
__global__ void __launch_bounds__(kFwdThreads) k_conv_fwd(const unsigned short* __restrict__ in0,
                                                          const unsigned short* __restrict__ in1, ConvGeo g,
                                                          const unsigned short* __restrict__ w,
                                                          const float* __restrict__ bias, int relu,
                                                          void* __restrict__ out) {
    constexpr int TY = 4 * ROWS, HY = TY + 2, HALO = HY * kHX;
    constexpr int MT = 4 * ROWS;                       // M-tiles per wave
    constexpr int NTMAX = ROWS == 1 ? 4 : 1;           // N-tiles per wave at most
    __shared__ uint4 halo[HALO * 2];     // [row][x][16 ch] as two 16-byte halves per voxel
    constexpr int CBM = ROWS == 2 ? 16 : kMaxCB;  // ROWS = 2 runs CB = 16 only (conv_forward)
    __shared__ s4 wl[9 * CBM * 4];       // [tap of the slab][co][4 groups of 4 ci]
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int q = lane >> 4, r16 = lane & 15;
    const int nxt = (g.W + 63) >> 6;
    const int xt = blockIdx.x % nxt, cb = blockIdx.x / nxt;
    const int x0 = xt * 64, y0 = blockIdx.y * TY, z = blockIdx.z;
    const int co0 = cb * g.CB;
    const int NT = g.CB >> 4;
    const int Cin = g.C0 + g.C1;
    const int H2 = g.H >> 1, W2 = g.W >> 1;
    f4 acc[MT][NTMAX];
#pragma unroll
    for (int m = 0; m < MT; ++m)
#pragma unroll
        for (int n = 0; n < NTMAX; ++n) acc[m][n] = f4{0.f, 0.f, 0.f, 0.f};
    const int nchunks = Cin >> 4;
    const unsigned short* hs = (const unsigned short*)halo;
    // The halo and weight slots this thread stages are the same in every stage: their source
    // offsets (halo: within a plane and chunk of in0, or of in1 at half resolution, -1 outside
    // the grid, -2 no slot; weights: within a dz slab and chunk, -1 no slot) are computed once,
    // and a stage adds its base.  The weight slab is stored [tap][co][4 groups of 4 ci] with co
    // stride CB, so slot e of the slab is wl[e].  (Offsets are 32-bit: conv_forward checks that
    // they fit.)
    constexpr int kPer = (HALO * 2 + kFwdThreads - 1) / kFwdThreads;
    constexpr int kPerW = (9 * CBM * 4 + kFwdThreads - 1) / kFwdThreads;
    int hoff0[kPer], hoff1[kPer], woff[kPerW];
#pragma unroll
    for (int k = 0; k < kPer; ++k) {
        const int e = threadIdx.x + k * kFwdThreads;
        hoff0[k] = hoff1[k] = -2;
        if (e < HALO * 2) {
            const int half = e & 1, vox = e >> 1;
            const int rx = vox % kHX, ry = vox / kHX;
            const int yy = y0 + ry - 1, xx = x0 + rx - 1;
            const bool inb = yy >= 0 && yy < g.H && xx >= 0 && xx < g.W;
            hoff0[k] = inb ? (yy * g.W + xx) * g.C0 + half * 8 : -1;
            hoff1[k] = inb ? ((yy >> 1) * W2 + (xx >> 1)) * g.C1 + half * 8 : -1;
        }
    }
#pragma unroll
    for (int k = 0; k < kPerW; ++k) {
        const int e = threadIdx.x + k * kFwdThreads;
        woff[k] = -1;
        if (e < 9 * g.CB * 4) {
            const int qq = e & 3, co = (e >> 2) % g.CB, tt = (e >> 2) / g.CB;
            woff[k] = (tt * g.Cout + co0 + co) * Cin + qq * 4;
        }
    }
    for (int c = 0; c < nchunks; ++c) {
        const bool hi = c * 16 >= g.C0;
        for (int dz = -1; dz <= 1; ++dz) {
            const int zz = z + dz;
            if (zz < 0 || zz >= g.D) continue;  // a plane outside the grid adds zeros (uniform)
            __syncthreads();                    // the previous stage's LDS reads are done
            const unsigned short* base = hi ? in1 + (size_t)(zz >> 1) * H2 * W2 * g.C1 + (c * 16 - g.C0)
                                            : in0 + (size_t)zz * g.H * g.W * g.C0 + c * 16;
            uint4 v[kPer];
#pragma unroll
            for (int k = 0; k < kPer; ++k) {
                const int o = hi ? hoff1[k] : hoff0[k];
                v[k] = make_uint4(0, 0, 0, 0);
                if (o >= 0) v[k] = *(const uint4*)(base + o);
            }
#pragma unroll
            for (int k = 0; k < kPer; ++k) {
                const int e = threadIdx.x + k * kFwdThreads;
                if (e < HALO * 2) halo[e] = v[k];
            }
            const unsigned short* wbase = w + (size_t)((dz + 1) * 9) * g.Cout * Cin + c * 16;
#pragma unroll
            for (int k = 0; k < kPerW; ++k)
                if (woff[k] >= 0) wl[threadIdx.x + k * kFwdThreads] = *(const s4*)(wbase + woff[k]);
            __syncthreads();
#pragma unroll
            for (int tt = 0; tt < 9; ++tt) {
                const int dy = tt / 3 - 1, dx = tt % 3 - 1;
                s4 a[MT];
#pragma unroll
                for (int m = 0; m < MT; ++m) {
                    const int row = (wave * ROWS + (m >> 2) + dy + 1) * kHX;
                    a[m] = *(const s4*)(hs + (size_t)(row + (m & 3) * 16 + r16 + dx + 1) * 16 + q * 4);
                }
#pragma unroll
                for (int n = 0; n < NTMAX; ++n) {
                    if (n < NT) {
                        const s4 b = wl[(tt * g.CB + n * 16 + r16) * 4 + q];
#pragma unroll
                        for (int m = 0; m < MT; ++m)
                            acc[m][n] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a[m], b, acc[m][n], 0, 0, 0);
                    }
                }
            }
        }
    }
    // D[i][j]: lane (q, r16) item i holds voxel x0 + 16 (m & 3) + 4q + i of row y0 + wave ROWS + m / 4,
    // output channel co0 + 16n + r16
#pragma unroll
    for (int n = 0; n < NTMAX; ++n) {
        if (n >= NT) continue;
        const int co = co0 + n * 16 + r16;
        const float b = bias ? bias[co] : 0.0f;
#pragma unroll
        for (int m = 0; m < MT; ++m) {
            const int y = y0 + wave * ROWS + (m >> 2);
            if (y >= g.H) continue;
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                const int x = x0 + (m & 3) * 16 + q * 4 + i;
                if (x >= g.W) continue;
                const size_t o = (((size_t)z * g.H + y) * g.W + x) * g.Cout + co;
                float val = acc[m][n][i] + b;
                if (relu) val = fmaxf(val, 0.0f);
                if constexpr (OUTF32)
                    ((float*)out)[o] = val;
                else
                    ((unsigned short*)out)[o] = f2bf(val);
            }
        }
    }
}

// PAR (inference of the 16-channel decoder layer): in1's channels go through the parity-class
// form of the upsampled convolution.  An output voxel at parity (pz, py, px) reads in1 at the
// half-resolution voxels (v >> 1) + p + o - 1, o in {0, 1}^3, with wpar's combined weights
// (the taps that land on one half-resolution voxel summed: 8 taps instead of 27).  So that an
// M-tile's 16 voxels share their x parity, its voxels are x0 + 32 (m&3 >> 1) + 2 r + (m & 1)
// (r = 0..15) in every phase (the full-resolution halo's columns are stored even ones first,
// then odd ones, so such a tile reads contiguous LDS); per (chunk, oz) the (TY/2 + 2) x 34
// half-resolution halo of plane (z >> 1) + pz + oz - 1 is staged with the 16 weight slabs
// (py, px, oy, ox) of that oz.
//
// k_conv_fwd with its stages double-buffered: the halo and weight slab of stage s + 1 travel
// from global memory straight into LDS (global_load_lds_dwordx4, no registers) while stage s
// runs its MFMAs from the other buffer.  A stage is (16-channel chunk, dz) or, for PAR's
// half-resolution chunks, (chunk, oz); planes outside the grid are skipped (uniform per
// workgroup).  Each lane's 16 bytes land at the wave's LDS base + 16 lane, so every staged
// buffer is filled in slot order: a slot's source is computed per lane (elements outside the
// grid read a 16-byte zero in global memory).  Same arithmetic and order as k_conv_fwd.
__device__ const uint4 g_zero16 = {0u, 0u, 0u, 0u};

// The 1x1x1 head fused into the last decoder layer's epilogue (HEAD): the layer's BF16 outputs
// are not stored; out[v] = out_scale (bo + sum_co wo[.][co] d0[v][co]) is formed from them.
struct HeadArgs {
    const float* wo;  // [2][16]
    const float* bo;  // [2]
    float out_scale;
    float2* out;
};

__device__ __forceinline__ void dma16(const void* src, uint4* lds_wave_base) {
    __builtin_amdgcn_global_load_lds(src, (__attribute__((address_space(3))) void*)lds_wave_base, 16, 0, 0);
}

template <bool OUTF32, int ROWS, bool PAR, bool HEAD = false>
// (two waves per SIMD, which the two LDS buffers allow: without the hint the parity form's
// split-half addressing took 256 VGPRs and one wave)
__global__ void __launch_bounds__(kFwdThreads) __attribute__((amdgpu_waves_per_eu(2)))
k_conv_dma(const unsigned short* __restrict__ in0,
                                                          const unsigned short* __restrict__ in1, ConvGeo g,
                                                          const unsigned short* __restrict__ w,
                                                          const unsigned short* __restrict__ wpar,
                                                          const float* __restrict__ bias, int relu,
                                                          void* __restrict__ out, HeadArgs head) {
    constexpr int TY = 4 * ROWS, HY = TY + 2, HALO = HY * kHX;
    constexpr int HY2 = TY / 2 + 2, HALO2 = HY2 * kHX2;
    static_assert(HALO2 <= HALO, "the half-resolution halo reuses the halo buffer");
    constexpr int MT = 4 * ROWS;
    constexpr int NTMAX = ROWS == 1 ? 4 : 1;
    constexpr int SLABS = PAR ? 16 : 9;
    constexpr int CBM = ROWS == 2 ? 16 : kMaxCB;
    // two buffers as separate arrays, used in a loop unrolled by two, so that the compiler sees
    // that the MFMA stage's LDS reads and the copies in flight touch different objects and
    // does not wait for the copies before the reads
    __shared__ uint4 haloA[HALO * 2], haloB[HALO * 2];        // [row][x][2 halves of 8 ch]
    __shared__ uint4 wlsA[SLABS * CBM * 2], wlsB[SLABS * CBM * 2];  // [slab][co][2 halves of 8 ci] (co stride CB)
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int q = lane >> 4, r16 = lane & 15;
    const int nxt = (g.W + 63) >> 6;
    const int xt = blockIdx.x % nxt, cb = blockIdx.x / nxt;
    const int x0 = xt * 64, y0 = blockIdx.y * TY, z = blockIdx.z;
    const int co0 = cb * g.CB;
    const int NT = g.CB >> 4;
    const int Cin = g.C0 + g.C1;
    const int H2 = g.H >> 1, W2 = g.W >> 1, D2 = g.D >> 1;
    const int pz = z & 1;
    f4 acc[MT][NTMAX];
#pragma unroll
    for (int m = 0; m < MT; ++m)
#pragma unroll
        for (int n = 0; n < NTMAX; ++n) acc[m][n] = f4{0.f, 0.f, 0.f, 0.f};
    auto xcol = [](int m, int r) { return PAR ? 32 * ((m & 3) >> 1) + 2 * r + (m & 1) : 16 * (m & 3) + r; };
    const int S = 3 * (Cin >> 4);
    // stage s: chunk s / 3, d = s % 3 (dz = d - 1, or oz = d for PAR's half-resolution chunks)
    auto valid = [&](int st) {
        const int c = st / 3, d = st - 3 * (st / 3);
        if (PAR && c * 16 >= g.C0) {
            const int zl = (z >> 1) + pz + d - 1;
            return d < 2 && zl >= 0 && zl < D2;
        }
        const int zz = z + d - 1;
        return zz >= 0 && zz < g.D;
    };
    // PAR (inference): the halo slots this thread copies are the same in every stage, so their
    // source offsets within a plane and chunk (-1: outside the grid, a 16-byte zero is copied;
    // -2: no slot) are computed once and a stage adds its plane's and chunk's base (the
    // prediction at 1024^2 x 128 22.7 -> 20.9 ms; the kernels without PAR, which also take the
    // half-resolution channels at full-resolution slots, keep the per-stage arithmetic: the
    // registers would spill there).  Offsets are 32-bit (conv_forward checks that they fit).
    constexpr int NHS = PAR ? (HALO * 2 + kFwdThreads - 1) / kFwdThreads : 1;
    constexpr int NHS2 = PAR ? (HALO2 * 2 + kFwdThreads - 1) / kFwdThreads : 1;
    int hoff0[NHS], hoff2[NHS2];
    if constexpr (PAR) {
#pragma unroll
        for (int i = 0; i < NHS; ++i) {
            const int e = wave * 64 + i * kFwdThreads + lane;
            hoff0[i] = -2;
            if (e < HALO * 2) {
                // slot e: channel half e / HALO of voxel e % HALO (the halves in two arrays, so
                // that 16 lanes reading one 8-byte chunk of 16 voxels meet 2-way bank conflicts,
                // not the 4-way of 32-byte voxel records)
                const int vox = e % HALO, half = e / HALO;
                const int ry = vox / kHX, col = vox - ry * kHX;
                // the columns x-parity-split (even ones, then odd ones)
                const int rx = col < kHX / 2 ? 2 * col : 2 * (col - kHX / 2) + 1;
                const int yy = y0 + ry - 1, xx = x0 + rx - 1;
                const bool inb = yy >= 0 && yy < g.H && xx >= 0 && xx < g.W;
                hoff0[i] = inb ? (yy * g.W + xx) * g.C0 + half * 8 : -1;
            }
        }
#pragma unroll
        for (int i = 0; i < NHS2; ++i) {
            const int e = wave * 64 + i * kFwdThreads + lane;
            hoff2[i] = -2;
            if (e < HALO2 * 2) {
                const int vox = e % HALO2, half = e / HALO2, rx = vox % kHX2, ry = vox / kHX2;
                const int yy = (y0 >> 1) - 1 + ry, xx = (x0 >> 1) - 1 + rx;
                hoff2[i] = (yy >= 0 && yy < H2 && xx >= 0 && xx < W2) ? (yy * W2 + xx) * g.C1 + half * 8 : -1;
            }
        }
    }
    auto issue = [&](int st, uint4* hb, uint4* wb) {
        const int c = st / 3, d = st - 3 * (st / 3);
        const bool hi = c * 16 >= g.C0;
        if (PAR && hi) {
            const int zl = (z >> 1) + pz + d - 1, ci0 = c * 16 - g.C0;
            const unsigned short* base = in1 + (size_t)zl * H2 * W2 * g.C1 + ci0;
#pragma unroll
            for (int i = 0; i < NHS2; ++i)
                if (hoff2[i] != -2)
                    dma16(hoff2[i] >= 0 ? (const void*)(base + hoff2[i]) : (const void*)&g_zero16,
                          hb + wave * 64 + i * kFwdThreads);
            const int nw = 16 * g.CB * 2;
            for (int e0 = wave * 64; e0 < nw; e0 += kFwdThreads) {
                const int e = e0 + lane;
                if (e < nw) {
                    // [half][slab][co], as the halo
                    const int half = e / (16 * g.CB), rem = e - half * 16 * g.CB;
                    const int sl = rem / g.CB, co = rem - sl * g.CB;
                    const int cls = (pz * 2 + (sl >> 3)) * 2 + ((sl >> 2) & 1), tap = d * 4 + (sl & 3);
                    dma16(wpar + ((size_t)(cls * 8 + tap) * g.Cout + co0 + co) * g.C1 + ci0 + half * 8, wb + e0);
                }
            }
            return;
        }
        const int zz = z + d - 1;
        if constexpr (PAR) {  // (chunks of in0 only: in1's are the half-resolution stages above)
            const unsigned short* base = in0 + (size_t)zz * g.H * g.W * g.C0 + c * 16;
#pragma unroll
            for (int i = 0; i < NHS; ++i)
                if (hoff0[i] != -2)
                    dma16(hoff0[i] >= 0 ? (const void*)(base + hoff0[i]) : (const void*)&g_zero16,
                          hb + wave * 64 + i * kFwdThreads);
        } else {
            for (int e0 = wave * 64; e0 < HALO * 2; e0 += kFwdThreads) {
                const int e = e0 + lane;
                if (e < HALO * 2) {
                    const int vox = e >> 1;
                    const int ry = vox / kHX, rx = vox - ry * kHX;
                    const int yy = y0 + ry - 1, xx = x0 + rx - 1;
                    const void* src = &g_zero16;
                    if (yy >= 0 && yy < g.H && xx >= 0 && xx < g.W) {
                        if (!hi)
                            src = in0 + (((size_t)zz * g.H + yy) * g.W + xx) * g.C0 + c * 16 + (e & 1) * 8;
                        else
                            src = in1 + (((size_t)(zz >> 1) * H2 + (yy >> 1)) * W2 + (xx >> 1)) * g.C1 +
                                  (c * 16 - g.C0) + (e & 1) * 8;
                    }
                    dma16(src, hb + e0);
                }
            }
        }
        const int nw = 9 * g.CB * 2;
        for (int e0 = wave * 64; e0 < nw; e0 += kFwdThreads) {
            const int e = e0 + lane;
            if (e < nw) {
                int tt, co, half;
                if constexpr (PAR) {  // [half][tap][co], as the halo
                    half = e / (9 * g.CB);
                    const int rem = e - half * 9 * g.CB;
                    tt = rem / g.CB;
                    co = rem - tt * g.CB;
                } else {
                    tt = e / (g.CB * 2);
                    co = (e >> 1) % g.CB;
                    half = e & 1;
                }
                dma16(w + ((size_t)(d * 9 + tt) * g.Cout + co0 + co) * Cin + c * 16 + half * 8, wb + e0);
            }
        }
    };
    auto compute = [&](int st, const uint4* hb, const uint4* wb) {
        const int c = st / 3, d = st - 3 * (st / 3);
        const bool hi = c * 16 >= g.C0;
        const unsigned short* hs = (const unsigned short*)hb;
        const s4* wq = (const s4*)wb;
        if (PAR && hi) {
#pragma unroll
            for (int oo = 0; oo < 4; ++oo) {
                const int oy = oo >> 1, ox = oo & 1;
                s4 a[MT];
#pragma unroll
                for (int m = 0; m < MT; ++m) {
                    const int yr = wave * ROWS + (m >> 2);
                    const int hrow = (yr >> 1) + (yr & 1) + oy;
                    const int hcol = 16 * ((m & 3) >> 1) + r16 + (m & 1) + ox;
                    a[m] = *(const s4*)(hs + ((size_t)(q >> 1) * HALO2 + hrow * kHX2 + hcol) * 8 + (q & 1) * 4);
                }
#pragma unroll
                for (int n = 0; n < NTMAX; ++n) {
                    if (n < NT) {
#pragma unroll
                        for (int m = 0; m < MT; ++m) {
                            const int yr = wave * ROWS + (m >> 2);
                            const int sl = ((yr & 1) * 2 + (m & 1)) * 4 + oo;
                            const s4 bb = wq[((q >> 1) * 16 * g.CB + sl * g.CB + n * 16 + r16) * 2 + (q & 1)];
                            acc[m][n] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a[m], bb, acc[m][n], 0, 0, 0);
                        }
                    }
                }
            }
            return;
        }
        (void)d;
#pragma unroll
        for (int tt = 0; tt < 9; ++tt) {
            const int dy = tt / 3 - 1, dx = tt % 3 - 1;
            s4 a[MT];
#pragma unroll
            for (int m = 0; m < MT; ++m) {
                const int row = (wave * ROWS + (m >> 2) + dy + 1) * kHX;
                const int hc = xcol(m, r16) + dx + 1;
                const int col = PAR ? (hc & 1) * (kHX / 2) + (hc >> 1) : hc;
                a[m] = PAR ? *(const s4*)(hs + ((size_t)(q >> 1) * HALO + row + col) * 8 + (q & 1) * 4)
                           : *(const s4*)(hs + (size_t)(row + col) * 16 + q * 4);
            }
#pragma unroll
            for (int n = 0; n < NTMAX; ++n) {
                if (n < NT) {
                    const s4 bb = PAR ? wq[((q >> 1) * 9 * g.CB + tt * g.CB + n * 16 + r16) * 2 + (q & 1)]
                                      : wq[(tt * g.CB + n * 16 + r16) * 4 + q];
#pragma unroll
                    for (int m = 0; m < MT; ++m)
                        acc[m][n] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a[m], bb, acc[m][n], 0, 0, 0);
                }
            }
        }
    };
    auto next = [&](int st) {
        ++st;
        while (st < S && !valid(st)) ++st;
        return st;
    };
    int st = next(-1);
    if (st < S) issue(st, haloA, wlsA);
    for (;;) {
        // stage st is in flight to buffer A; the next goes to B
        if (st >= S) break;
        int sn = next(st);
        __builtin_amdgcn_s_waitcnt(0x0F70);  // vmcnt(0): this wave's copies of stage st are in LDS
        __syncthreads();                     // every wave's are, and the last stage's reads of B are done
        if (sn < S) issue(sn, haloB, wlsB);
        compute(st, haloA, wlsA);
        st = sn;
        if (st >= S) break;
        sn = next(st);
        __builtin_amdgcn_s_waitcnt(0x0F70);
        __syncthreads();
        if (sn < S) issue(sn, haloA, wlsA);
        compute(st, haloB, wlsB);
        st = sn;
    }
    if constexpr (HEAD) {
        // 16 output channels (NT = 1): the BF16 outputs go through LDS (the halo buffer, free
        // after the barrier), then a thread per voxel forms the head's two sums in channel order
        // (as k_head_fwd does) and stores its complex output
        __syncthreads();
        unsigned short* ot = (unsigned short*)haloA;  // [row][x][16 ch]
        const float bv = bias[r16];
#pragma unroll
        for (int m = 0; m < MT; ++m) {
            const int row = wave * ROWS + (m >> 2);
#pragma unroll
            for (int i = 0; i < 4; ++i)
                ot[(row * 64 + xcol(m, q * 4 + i)) * 16 + r16] = f2bf(fmaxf(acc[m][0][i] + bv, 0.0f));
        }
        __syncthreads();
#pragma unroll
        for (int k = 0; k < TY * 64 / kFwdThreads; ++k) {
            const int vox = threadIdx.x + k * kFwdThreads, row = vox >> 6, x = vox & 63;
            const int y = y0 + row, xx = x0 + x;
            if (y >= g.H || xx >= g.W) continue;
            float o0 = head.bo[0], o1 = head.bo[1];
#pragma unroll
            for (int hf = 0; hf < 2; ++hf) {
                const uint4 u = haloA[vox * 2 + hf];
                const unsigned wd[4] = {u.x, u.y, u.z, u.w};
#pragma unroll
                for (int j = 0; j < 8; ++j) {
                    const int co = hf * 8 + j;
                    const float h = bf2f((unsigned short)((wd[j >> 1] >> ((j & 1) * 16)) & 0xffffu));
                    o0 += head.wo[co] * h;
                    o1 += head.wo[16 + co] * h;
                }
            }
            head.out[((size_t)z * g.H + y) * g.W + xx] = make_float2(o0 * head.out_scale, o1 * head.out_scale);
        }
        return;
    }
    if constexpr (!OUTF32 && ROWS == 2) {
        // 16 output channels (CB = 16): through LDS (the halo buffer, free after the barrier),
        // so that each voxel's 32 bytes leave as two 16-byte stores of coalesced rows
        __syncthreads();
        unsigned short* ot = (unsigned short*)haloA;  // [row][x][16 ch]
        const int co = co0 + r16;
        const float bv = bias ? bias[co] : 0.0f;
#pragma unroll
        for (int m = 0; m < MT; ++m) {
            const int row = wave * ROWS + (m >> 2);
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                float val = acc[m][0][i] + bv;
                if (relu) val = fmaxf(val, 0.0f);
                ot[(row * 64 + xcol(m, q * 4 + i)) * 16 + r16] = f2bf(val);
            }
        }
        __syncthreads();
#pragma unroll
        for (int k = 0; k < TY * 64 * 2 / kFwdThreads; ++k) {
            const int e = threadIdx.x + k * kFwdThreads;  // uint4 of the tile
            const int vox = e >> 1, row = vox >> 6, x = vox & 63;
            const int y = y0 + row, xx = x0 + x;
            if (y < g.H && xx < g.W)
                *(uint4*)((unsigned short*)out + (((size_t)z * g.H + y) * g.W + xx) * g.Cout + co0 + (e & 1) * 8) =
                    haloA[e];
        }
        return;
    }
#pragma unroll
    for (int n = 0; n < NTMAX; ++n) {
        if (n >= NT) continue;
        const int co = co0 + n * 16 + r16;
        const float bv = bias ? bias[co] : 0.0f;
#pragma unroll
        for (int m = 0; m < MT; ++m) {
            const int y = y0 + wave * ROWS + (m >> 2);
            if (y >= g.H) continue;
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                const int x = x0 + xcol(m, q * 4 + i);
                if (x >= g.W) continue;
                const size_t o = (((size_t)z * g.H + y) * g.W + x) * g.Cout + co;
                float val = acc[m][n][i] + bv;
                if (relu) val = fmaxf(val, 0.0f);
                if constexpr (OUTF32)
                    ((float*)out)[o] = val;
                else
                    ((unsigned short*)out)[o] = f2bf(val);
            }
        }
    }
}

// Weight (and bias) gradient partials of one chunk of rows (z, y):
//   part[chunk][tap][co][ci] = sum_u g[u][co] * in[u + tap][ci],  part[chunk][NW + co] = sum_u g[u][co]
// (NW = 27 Cout Cin; the bias sum is written by the ci-tile 0 workgroups).  One wave per
// workgroup; grid (Cout/16, Cin/16, nchunks).  Per 16-voxel row segment the gradient tile is
// staged transposed ([co][x]) and the input window of the 9 (dz, dy) rows over x0-1 .. x0+16
// transposed too ([dzdy][ci][x]), so that a lane's B fragment (4 consecutive voxels of one
// channel) for the three dx of a row comes from two 8-byte reads; MFMA: M = co, N = ci, K = the
// 16 voxels of the segment.
constexpr int kWX = 24;  // the weight-gradient window's row: 18 voxels padded to 24 (48 bytes)
__global__ void __launch_bounds__(64) k_conv_wgrad(const unsigned short* __restrict__ in0,
                                                   const unsigned short* __restrict__ in1, ConvGeo g,
                                                   const unsigned short* __restrict__ gout, int nchunks,
                                                   float* __restrict__ part) {
    __shared__ unsigned short ga[16 * 16];         // [co][x]
    __shared__ __align__(16) unsigned short win[9 * 16 * kWX];  // [dz dy][ci][x + 1], rows padded to kWX
    const int lane = threadIdx.x;
    const int q = lane >> 4, r16 = lane & 15;
    const int co0 = blockIdx.x * 16, ci0 = blockIdx.y * 16;
    const int chunk = blockIdx.z;
    const int Cin = g.C0 + g.C1;
    const long rows = (long)g.D * g.H;
    const long r0 = rows * chunk / nchunks, r1 = rows * (chunk + 1) / nchunks;
    f4 acc[kTaps];
#pragma unroll
    for (int t = 0; t < kTaps; ++t) acc[t] = f4{0.f, 0.f, 0.f, 0.f};
    float bsum = 0.0f;
    for (long r = r0; r < r1; ++r) {
        const int z = (int)(r / g.H), y = (int)(r - (long)z * g.H);
        for (int x0 = 0; x0 < g.W; x0 += 16) {
            __syncthreads();
            {
                const int xi = lane >> 2, cg = (lane & 3) * 4;
                s4 v = {0, 0, 0, 0};
                if (x0 + xi < g.W) v = *(const s4*)(gout + (((size_t)z * g.H + y) * g.W + x0 + xi) * g.Cout + co0 + cg);
#pragma unroll
                for (int j = 0; j < 4; ++j) ga[(cg + j) * 16 + xi] = (unsigned short)v[j];
            }
            // a task: 4 voxels (x + 1 = 4 xg .. 4 xg + 3) x 4 channels of one (dz, dy) row, loaded
            // as 4 voxel vectors and written transposed as 4 channel rows of 4 voxels
            for (int e = lane; e < 9 * 5 * 4; e += 64) {
                const int combo = e / 20, rem = e - combo * 20;
                const int xg = rem >> 2, cg = (rem & 3) * 4;
                s4 v[4];
#pragma unroll
                for (int k = 0; k < 4; ++k) {
                    const int xi = xg * 4 + k;
                    v[k] = xi < 18 ? load_in4(in0, in1, g, z + combo / 3 - 1, y + combo % 3 - 1, x0 - 1 + xi, ci0 + cg)
                                   : s4{0, 0, 0, 0};
                }
#pragma unroll
                for (int j = 0; j < 4; ++j) {
                    const s4 row = {v[0][j], v[1][j], v[2][j], v[3][j]};
                    *(s4*)(win + (combo * 16 + cg + j) * kWX + xg * 4) = row;
                }
            }
            __syncthreads();
            const s4 a = *(const s4*)(ga + r16 * 16 + q * 4);
            if (blockIdx.y == 0) {
#pragma unroll
                for (int j = 0; j < 4; ++j) bsum += bf2f((unsigned short)a[j]);
            }
#pragma unroll
            for (int combo = 0; combo < 9; ++combo) {
                // B for the 3 dx of this (dz, dy): voxels 4q + j + dx + 1 of channel r16, from the
                // 8 voxels 4q .. 4q + 7 (two 8-byte reads) by byte shifts
                const unsigned short* rowp = win + (combo * 16 + r16) * kWX + q * 4;
                const unsigned long long lo = *(const unsigned long long*)rowp;
                const unsigned long long hi = *(const unsigned long long*)(rowp + 4);
                const unsigned long long bw[3] = {lo, (lo >> 16) | (hi << 48), (lo >> 32) | (hi << 32)};
#pragma unroll
                for (int d = 0; d < 3; ++d) {
                    const s4 b = __builtin_bit_cast(s4, bw[d]);
                    const int t = combo * 3 + d;
                    acc[t] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, acc[t], 0, 0, 0);
                }
            }
        }
    }
    const size_t NW = (size_t)kTaps * g.Cout * Cin;
    float* p = part + (size_t)chunk * (NW + g.Cout);
#pragma unroll
    for (int t = 0; t < kTaps; ++t)
#pragma unroll
        for (int i = 0; i < 4; ++i) p[((size_t)t * g.Cout + co0 + q * 4 + i) * Cin + ci0 + r16] = acc[t][i];
    if (blockIdx.y == 0) {
        bsum += __shfl_xor(bsum, 16);
        bsum += __shfl_xor(bsum, 32);
        if (lane < 16) p[NW + co0 + lane] = bsum;
    }
}

// out[i] = sum over chunks of part[chunk][i]: a workgroup takes 16 entries, its 16 lanes per
// entry sum chunks j, j + 16, ... in order, then a fixed tree adds the 16 lane sums
constexpr int kSumLanes = 16;
__global__ void __launch_bounds__(256) k_sum_chunks(const float* __restrict__ part, long n, int nchunks,
                                                    float* __restrict__ out) {
    __shared__ float red[kSumLanes][16 + 1];
    const int e = threadIdx.x & 15, j = threadIdx.x >> 4;
    const long i = (long)blockIdx.x * 16 + e;
    float s = 0.0f;
    if (i < n)
        for (int c = j; c < nchunks; c += kSumLanes) s += part[(size_t)c * n + i];
    red[j][e] = s;
    __syncthreads();
    for (int h = kSumLanes / 2; h > 0; h >>= 1) {
        if (j < h) red[j][e] += red[j + h][e];
        __syncthreads();
    }
    if (j == 0 && i < n) out[i] = red[0][e];
}

// out[i] = sum over chunks of part[chunk][i] for few entries and many chunks: a workgroup per
// entry, thread t sums chunks t, t + 256, ... in order, then a fixed tree over the threads
__global__ void __launch_bounds__(256) k_sum_chunks_wide(const float* __restrict__ part, long n, int nchunks,
                                                         float* __restrict__ out) {
    __shared__ float red[256];
    const long i = blockIdx.x;
    float s = 0.0f;
    for (int c = threadIdx.x; c < nchunks; c += 256) s += part[(size_t)c * n + i];
    red[threadIdx.x] = s;
    __syncthreads();
    for (int h = 128; h > 0; h >>= 1) {
        if (threadIdx.x < h) red[threadIdx.x] += red[threadIdx.x + h];
        __syncthreads();
    }
    if (threadIdx.x == 0) out[i] = red[0];
}

// x16[v] = (in_scale Re, in_scale Im, 0 x 14) in BF16
__global__ void k_input16(const float2* __restrict__ in, long V, float scale, unsigned short* __restrict__ x16) {
    const long v = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= V) return;
    const float2 a = in[v];
    uint4 lo, hi = make_uint4(0, 0, 0, 0);
    lo.x = (unsigned)f2bf(a.x * scale) | ((unsigned)f2bf(a.y * scale) << 16);
    lo.y = lo.z = lo.w = 0;
    uint4* o = (uint4*)(x16 + v * 16);
    o[0] = lo;
    o[1] = hi;
}

// The input layer e0 = relu(conv(x) + b) straight from the complex64 volume: its two real input
// channels (in_scale Re, in_scale Im in BF16) make K = 27 taps x 2 channels, packed as K = 64
// (taps 27..31 zero) into 4 MFMA K-steps: step s, lane group q holds taps 8s + 2q and 8s + 2q + 1.
// A workgroup computes 8 rows x 64 columns of kFirstZ consecutive planes (wave w: rows 2w, 2w + 1,
// 4 M-tiles each) from the 3 x 10 x 66 halo of (Re, Im) pairs staged in LDS (at one plane per
// workgroup, each loaded its three halo planes for 32 MFMAs: 2.6 ms at 1024^2 x 128, now 2.2; held
// to 6 waves per SIMD instead of 4 it spills).  wp:
// [4][16 co][16 k] BF16 with k = 2 (tap - 8s) + channel.  Grid: (ceil(W/64), ceil(H/8),
// ceil(D/kFirstZ)).
__global__ void __launch_bounds__(kFwdThreads) k_conv_first(const float2* __restrict__ in, int D, int H, int W,
                                                            float scale, const unsigned short* __restrict__ wp,
                                                            const float* __restrict__ bias,
                                                            unsigned short* __restrict__ out,
                                                            unsigned short* __restrict__ pooled) {
    // A workgroup walks kFirstZ planes of its 8 x 64 tile: a ring of three halo planes in LDS
    // (plane p in slot p mod 3), one new plane loaded per step.
    constexpr int HY = 10, HX = 66;
    __shared__ unsigned halo[3 * HY * HX];   // (Re, Im) BF16 pair per voxel, [slot][row][x]
    __shared__ uint4 otile[8 * 64 * 2];      // the tile's outputs, [row][x][16 ch] BF16
    __shared__ s4 wls[4 * 16 * 4];
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int q = lane >> 4, r16 = lane & 15;
    const int x0 = blockIdx.x * 64, y0 = blockIdx.y * 8;
    const int z0 = blockIdx.z * kFirstZ, z1 = min(D, z0 + kFirstZ);
    // plane zz's 10 halo rows into its ring slot: a wave per row, 64 + 2 voxels, one (two for
    // lanes 0, 1) coalesced 8-byte load per lane; zeros outside the grid
    auto load_plane = [&](int zz) {
        const int slot = (zz + 3) % 3;
        for (int ry = wave; ry < HY; ry += 4) {
            const int yy = y0 + ry - 1;
            const bool row_in = zz >= 0 && zz < D && yy >= 0 && yy < H;
            const float2* src = in + ((size_t)(row_in ? zz : 0) * H + (row_in ? yy : 0)) * W;
            for (int rx = lane; rx < HX; rx += 64) {
                const int xx = x0 + rx - 1;
                unsigned v = 0;
                if (row_in && xx >= 0 && xx < W) {
                    const float2 a = src[xx];
                    v = (unsigned)f2bf(a.x * scale) | ((unsigned)f2bf(a.y * scale) << 16);
                }
                halo[(slot * HY + ry) * HX + rx] = v;
            }
        }
    };
    for (int e = threadIdx.x; e < 4 * 16 * 4; e += kFwdThreads) wls[e] = *(const s4*)(wp + e * 4);
    load_plane(z0 - 1);
    load_plane(z0);
    const float bv = bias[r16];
    unsigned short* ot = (unsigned short*)otile;
    // the 2 x 2 x 2 average pool of the outputs (the next level's input), formed as k_pool does:
    // a thread holds both rows of a y pair (M-tiles m and m + 4), both voxels of an x pair (items
    // 2j, 2j + 1) and sees the z pair in consecutive steps; FP32 sums of the BF16 outputs in
    // k_pool's order (z, then y, then x), the same bits
    float psum[4][2];
    const int H2 = H >> 1, W2 = W >> 1;
    for (int z = z0; z < z1; ++z) {
        // the slot of plane z + 1 held plane z - 2, read by the last step (which ended at a
        // barrier); otile's last reads (the last step's stores) precede this step's barrier
        load_plane(z + 1);
        __syncthreads();
        f4 acc[8];
#pragma unroll
        for (int m = 0; m < 8; ++m) acc[m] = f4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
        for (int st = 0; st < 4; ++st) {
            const s4 b = wls[(st * 16 + r16) * 4 + q];
            int off[2];
            bool ok[2];
#pragma unroll
            for (int h = 0; h < 2; ++h) {
                const int t = 8 * st + 2 * q + h;
                ok[h] = t < 27;
                const int tc = ok[h] ? t : 0;
                // the tap's plane z + tc / 9 - 1 in its ring slot, at (row 0, x 0)
                off[h] = (((z + tc / 9 + 2) % 3) * HY + (tc / 3) % 3) * HX + (tc % 3);
            }
#pragma unroll
            for (int m = 0; m < 8; ++m) {
                const int base = (wave * 2 + (m >> 2)) * HX + (m & 3) * 16 + r16;
                const unsigned lo = ok[0] ? halo[base + off[0]] : 0u;
                const unsigned hi = ok[1] ? halo[base + off[1]] : 0u;
                s4 a;
                a[0] = (short)(lo & 0xffffu);
                a[1] = (short)(lo >> 16);
                a[2] = (short)(hi & 0xffffu);
                a[3] = (short)(hi >> 16);
                acc[m] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, acc[m], 0, 0, 0);
            }
        }
        // through LDS, so that the tile's rows (64 voxels x 32 bytes each) leave as 16-byte stores
        unsigned short vb[8][4];
#pragma unroll
        for (int m = 0; m < 8; ++m) {
            const int row = wave * 2 + (m >> 2);
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                const int x = (m & 3) * 16 + q * 4 + i;
                vb[m][i] = f2bf(fmaxf(acc[m][i] + bv, 0.0f));
                ot[(row * 64 + x) * 16 + r16] = vb[m][i];
            }
        }
        if (pooled) {
            if (((z - z0) & 1) == 0) {
#pragma unroll
                for (int t = 0; t < 4; ++t)
#pragma unroll
                    for (int j = 0; j < 2; ++j) psum[t][j] = 0.0f;
            }
#pragma unroll
            for (int t = 0; t < 4; ++t)
#pragma unroll
                for (int j = 0; j < 2; ++j)
#pragma unroll
                    for (int b = 0; b < 2; ++b)
#pragma unroll
                        for (int d = 0; d < 2; ++d) psum[t][j] += bf2f(vb[t + 4 * b][2 * j + d]);
            if (((z - z0) & 1) == 1) {
                const int uy = (y0 >> 1) + wave;
#pragma unroll
                for (int t = 0; t < 4; ++t)
#pragma unroll
                    for (int j = 0; j < 2; ++j) {
                        const int ux = (x0 >> 1) + t * 8 + q * 2 + j;
                        if (uy < H2 && ux < W2)
                            pooled[(((size_t)(z >> 1) * H2 + uy) * W2 + ux) * 16 + r16] = f2bf(psum[t][j] * 0.125f);
                    }
            }
        }
        __syncthreads();
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            const int e = threadIdx.x + k * kFwdThreads;  // uint4 of the tile
            const int vox = e >> 1, row = vox >> 6, x = vox & 63;
            const int y = y0 + row, xx = x0 + x;
            if (y < H && xx < W) *(uint4*)(out + (((size_t)z * H + y) * W + xx) * 16 + (e & 1) * 8) = otile[e];
        }
    }
}

// the input layer's packed weights: wp[s][co][k] = w[t = 8s + k/2][co][k % 2] (0 for t >= 27)
__global__ void k_pack_first(const float* __restrict__ p, unsigned short* __restrict__ wp) {
    const int i = threadIdx.x + blockIdx.x * blockDim.x;
    if (i >= 4 * 16 * 16) return;
    const int k = i & 15, co = (i >> 4) & 15, st = i >> 8;
    const int t = 8 * st + (k >> 1);
    wp[i] = t < 27 ? f2bf(p[((size_t)t * 16 + co) * 16 + (k & 1)]) : (unsigned short)0;
}

// p[u][c] = (sum of the 8 children e[2u + (a, b, c)][c]) / 8 (FP32 sum, BF16 out); thread per (u, 4 channels)
__global__ void k_pool(const unsigned short* __restrict__ e, int D, int H, int W, int C,
                       unsigned short* __restrict__ p) {
    const int D2 = D >> 1, H2 = H >> 1, W2 = W >> 1, C4 = C >> 2;
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    const long n = (long)D2 * H2 * W2 * C4;
    if (i >= n) return;
    const int c = (int)(i % C4) * 4;
    long u = i / C4;
    const int x = (int)(u % W2), y = (int)((u / W2) % H2), z = (int)(u / ((long)W2 * H2));
    float s[4] = {0.f, 0.f, 0.f, 0.f};
    for (int a = 0; a < 2; ++a)
        for (int b = 0; b < 2; ++b)
            for (int d = 0; d < 2; ++d) {
                const s4 v = *(const s4*)(e + (((size_t)(2 * z + a) * H + 2 * y + b) * W + 2 * x + d) * C + c);
#pragma unroll
                for (int j = 0; j < 4; ++j) s[j] += bf2f((unsigned short)v[j]);
            }
    s4 o;
#pragma unroll
    for (int j = 0; j < 4; ++j) o[j] = (short)f2bf(s[j] * 0.125f);
    *(s4*)(p + (size_t)u * C + c) = o;
}

// Pre-activation gradient of a ReLU layer at level k ([V][C], BF16):
//   post[v][c] = skip[v][c]                            (skip: [V][sk_stride], channels 0..C-1)
//              + pool[parent(v)][c] / 8                (pool: [V/8][C], the average pool's adjoint)
//              + sum over the 8 children of split[child][off + c]   (split: [8V][Ct], the upsample's adjoint)
//   pre = post * (act > 0)
struct Finalize {
    const unsigned short* act;
    const float* skip;
    int sk_stride;
    const float* pool;
    const float* split;
    int off, Ct;
    int D, H, W, C;
    unsigned short* out;
};
__global__ void k_finalize(Finalize f) {
    const int C4 = f.C >> 2;
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    const long n = (long)f.D * f.H * f.W * C4;
    if (i >= n) return;
    const int c = (int)(i % C4) * 4;
    const long v = i / C4;
    const int x = (int)(v % f.W), y = (int)((v / f.W) % f.H), z = (int)(v / ((long)f.W * f.H));
    float s[4] = {0.f, 0.f, 0.f, 0.f};
    if (f.skip) {
        const float4 a = *(const float4*)(f.skip + (size_t)v * f.sk_stride + c);
        s[0] += a.x; s[1] += a.y; s[2] += a.z; s[3] += a.w;
    }
    if (f.pool) {
        const long pu = (((long)(z >> 1) * (f.H >> 1) + (y >> 1)) * (f.W >> 1) + (x >> 1));
        const float4 a = *(const float4*)(f.pool + (size_t)pu * f.C + c);
        s[0] += a.x * 0.125f; s[1] += a.y * 0.125f; s[2] += a.z * 0.125f; s[3] += a.w * 0.125f;
    }
    if (f.split) {
        const int H2 = f.H * 2, W2 = f.W * 2;
        for (int a = 0; a < 2; ++a)
            for (int b = 0; b < 2; ++b)
                for (int d = 0; d < 2; ++d) {
                    const size_t ch = (((size_t)(2 * z + a) * H2 + 2 * y + b) * W2 + 2 * x + d);
                    const float4 t = *(const float4*)(f.split + ch * f.Ct + f.off + c);
                    s[0] += t.x; s[1] += t.y; s[2] += t.z; s[3] += t.w;
                }
    }
    const s4 act = *(const s4*)(f.act + (size_t)v * f.C + c);
    s4 o;
#pragma unroll
    for (int j = 0; j < 4; ++j) o[j] = (short)f2bf(bf2f((unsigned short)act[j]) > 0.0f ? s[j] : 0.0f);
    *(s4*)(f.out + (size_t)v * f.C + c) = o;
}

__device__ __forceinline__ float wave_sum_f(float v) {
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
    return v;
}
__device__ __forceinline__ double wave_sum_d(double v) {
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
    return v;
}

// Output head: raw o_c = bo[c] + sum_k wo[c][k] d0[v][k]; out[v] = out_scale * (o_0, o_1).
// With a target, lpart[block] = sum over the block's voxels of w |o - target / out_scale|^2 with
// w = 1 + pos_weight where the target is non-zero, else 1.
__global__ void __launch_bounds__(kEw) k_head_fwd(const unsigned short* __restrict__ d0, long V,
                                                  const float* __restrict__ wo, const float* __restrict__ bo,
                                                  float out_scale, float2* __restrict__ out,
                                                  const float2* __restrict__ target, float pos_weight,
                                                  double* __restrict__ lpart) {
    __shared__ double red[kEw / 64];
    const long v = (long)blockIdx.x * blockDim.x + threadIdx.x;
    double l = 0.0;
    if (v < V) {
        const s4* p = (const s4*)(d0 + v * 16);
        float o0 = bo[0], o1 = bo[1];
#pragma unroll
        for (int k4 = 0; k4 < 4; ++k4) {
            const s4 a = p[k4];
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                const float h = bf2f((unsigned short)a[j]);
                o0 += wo[k4 * 4 + j] * h;
                o1 += wo[16 + k4 * 4 + j] * h;
            }
        }
        if (out) out[v] = make_float2(o0 * out_scale, o1 * out_scale);
        if (target) {
            const float2 t = target[v];
            const float e0 = o0 - t.x / out_scale, e1 = o1 - t.y / out_scale;
            const float wv = (t.x != 0.0f || t.y != 0.0f) ? 1.0f + pos_weight : 1.0f;
            l = (double)wv * ((double)e0 * e0 + (double)e1 * e1);
        }
    }
    if (!lpart) return;
    l = wave_sum_d(l);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = l;
    __syncthreads();
    if (threadIdx.x == 0) {
        double s = 0.0;
        for (int k = 0; k < kEw / 64; ++k) s += red[k];
        lpart[blockIdx.x] = s;
    }
}

// Head backward for loss L = sum_v w_v |o - t/out_scale|^2 / (2 V):  g_c = w_v (o_c - t_c/out_scale) / V;
// hpart[block][0..31] = sum g_c d0[k] (dWo), [32..33] = sum g_c (dbo); gd0 = (sum_c g_c wo[c][k]) * (d0 > 0)
__global__ void __launch_bounds__(kEw) k_head_bwd(const unsigned short* __restrict__ d0, long V,
                                                  const float* __restrict__ wo, const float* __restrict__ bo,
                                                  float out_scale, const float2* __restrict__ target,
                                                  float pos_weight, unsigned short* __restrict__ gd0,
                                                  float* __restrict__ hpart) {
    __shared__ float red[34][kEw / 64];
    const long v = (long)blockIdx.x * blockDim.x + threadIdx.x;
    float h[16];
    float g0 = 0.f, g1 = 0.f;
    if (v < V) {
        const s4* p = (const s4*)(d0 + v * 16);
#pragma unroll
        for (int k4 = 0; k4 < 4; ++k4) {
            const s4 a = p[k4];
#pragma unroll
            for (int j = 0; j < 4; ++j) h[k4 * 4 + j] = bf2f((unsigned short)a[j]);
        }
        float o0 = bo[0], o1 = bo[1];
#pragma unroll
        for (int k = 0; k < 16; ++k) {
            o0 += wo[k] * h[k];
            o1 += wo[16 + k] * h[k];
        }
        const float2 t = target[v];
        const float inv = ((t.x != 0.0f || t.y != 0.0f) ? 1.0f + pos_weight : 1.0f) / (float)V;
        g0 = (o0 - t.x / out_scale) * inv;
        g1 = (o1 - t.y / out_scale) * inv;
        s4 o[4];
#pragma unroll
        for (int k = 0; k < 16; ++k) {
            const float gk = h[k] > 0.0f ? g0 * wo[k] + g1 * wo[16 + k] : 0.0f;
            o[k >> 2][k & 3] = (short)f2bf(gk);
        }
        s4* q = (s4*)(gd0 + v * 16);
#pragma unroll
        for (int k4 = 0; k4 < 4; ++k4) q[k4] = o[k4];
    } else {
#pragma unroll
        for (int k = 0; k < 16; ++k) h[k] = 0.0f;
    }
    const int w = threadIdx.x >> 6;
#pragma unroll
    for (int k = 0; k < 16; ++k) {
        const float a = wave_sum_f(g0 * h[k]);
        const float b = wave_sum_f(g1 * h[k]);
        if ((threadIdx.x & 63) == 0) {
            red[k][w] = a;
            red[16 + k][w] = b;
        }
    }
    const float s0 = wave_sum_f(g0), s1 = wave_sum_f(g1);
    if ((threadIdx.x & 63) == 0) {
        red[32][w] = s0;
        red[33][w] = s1;
    }
    __syncthreads();
    if (threadIdx.x < 34) {
        float s = 0.f;
        for (int k = 0; k < kEw / 64; ++k) s += red[threadIdx.x][k];
        hpart[(size_t)blockIdx.x * 34 + threadIdx.x] = s;
    }
}

// Adam on FP32 master parameters: m = b1 m + (1-b1) g, v = b2 v + (1-b2) g^2,
// p -= lr * (m / bc1) / (sqrt(v / bc2) + eps)
__global__ void k_adam(float* __restrict__ p, const float* __restrict__ g, float* __restrict__ m,
                       float* __restrict__ v, long n, float lr, float bc1, float bc2) {
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float gi = g[i];
    const float mi = kB1 * m[i] + (1.0f - kB1) * gi;
    const float vi = kB2 * v[i] + (1.0f - kB2) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    p[i] -= lr * (mi / bc1) / (sqrtf(vi / bc2) + kEps);
}

// BF16 copies of a layer's weights: w[t][co][ci] and the input-gradient weights wT[26-t][ci][co]
__global__ void k_prep_weights(const float* __restrict__ p, int Cout, int Cin, unsigned short* __restrict__ w,
                               unsigned short* __restrict__ wT) {
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    const long n = (long)kTaps * Cout * Cin;
    if (i >= n) return;
    const int ci = (int)(i % Cin), co = (int)((i / Cin) % Cout), t = (int)(i / ((long)Cin * Cout));
    const unsigned short b = f2bf(p[i]);
    w[i] = b;
    wT[((size_t)(kTaps - 1 - t) * Cin + ci) * Cout + co] = b;
}

// Combined BF16 weights of the upsampled channels for the parity-class form (PAR above):
// wpar[pz][py][px][oz][oy][ox][co][ci] = bf16(sum of w[t][co][C0 + ci] over the taps t whose
// offset along each axis lies in S(p, o): S(0, 0) = {-1}, S(0, 1) = {0, 1}, S(1, 0) = {-1, 0},
// S(1, 1) = {+1}), summed in double.
__global__ void k_prep_parity(const float* __restrict__ p, int Cout, int C0, int C1, unsigned short* __restrict__ wpar) {
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= 64L * Cout * C1) return;
    const int ci = (int)(i % C1), co = (int)((i / C1) % Cout), t8 = (int)((i / ((long)C1 * Cout)) % 8);
    const int cls = (int)(i / (8L * C1 * Cout));
    const int par[3] = {cls >> 2, (cls >> 1) & 1, cls & 1}, off[3] = {t8 >> 2, (t8 >> 1) & 1, t8 & 1};
    int lo[3], hi[3];
    for (int d = 0; d < 3; ++d) {
        lo[d] = (par[d] == 0) ? (off[d] == 0 ? -1 : 0) : (off[d] == 0 ? -1 : 1);
        hi[d] = (par[d] == 0) ? (off[d] == 0 ? -1 : 1) : (off[d] == 0 ? 0 : 1);
    }
    const int Cin = C0 + C1;
    double acc = 0.0;
    for (int tz = lo[0]; tz <= hi[0]; ++tz)
        for (int ty = lo[1]; ty <= hi[1]; ++ty)
            for (int tx = lo[2]; tx <= hi[2]; ++tx) {
                const int t = (tz + 1) * 9 + (ty + 1) * 3 + (tx + 1);
                acc += (double)p[((size_t)t * Cout + co) * Cin + C0 + ci];
            }
    wpar[i] = f2bf((float)acc);
}

// ----------------------------------------------------------------------------
// Host launchers
// ----------------------------------------------------------------------------
int pick_cb(int Cout) {
    for (int cb : {64, 48, 32, 16})
        if (Cout % cb == 0) return cb;
    return 0;
}

hipError_t conv_forward(const unsigned short* in0, int C0, const unsigned short* in1, int C1, int D, int H, int W,
                        const unsigned short* w, const float* bias, int Cout, int relu, void* out, bool out_f32,
                        hipStream_t s, const unsigned short* wpar = nullptr, const HeadArgs* head = nullptr) {
    if (C0 <= 0 || C0 % 16 || C1 < 0 || C1 % 16 || Cout <= 0 || Cout % 16 || D <= 0 || H <= 0 || W <= 0)
        return hipErrorInvalidValue;
    if (C1 > 0 && (D % 2 || H % 2 || W % 2 || !in1)) return hipErrorInvalidValue;
    if (D > 65535) return hipErrorInvalidValue;
    // the parity-class form (wpar, inference of decoder layers of up to 32 channels) runs 16
    // output channels per workgroup, two rows per wave
    // the kernels address a plane's elements (and a weight slab's) with 32-bit offsets
    if ((long)H * W * std::max(C0, C1) >= (1L << 31) || 27L * Cout * (C0 + C1) >= (1L << 31))
        return hipErrorInvalidValue;
    const bool par = wpar && C1 > 0 && !out_f32 && Cout <= 32;
    ConvGeo g{D, H, W, C0, C1, Cout, par ? 16 : pick_cb(Cout)};
    const long plane = (long)H * W;
    if (plane > (1L << 30)) return hipErrorInvalidValue;
    const long gx = (long)((W + 63) / 64) * (Cout / g.CB);
    const int rows = g.CB == 16 ? 2 : 1;  // rows per wave
    const int ty = 4 * rows;
    if (gx > 0x7fffffffL || (H + ty - 1) / ty > 65535) return hipErrorInvalidValue;
    dim3 grid((unsigned)gx, (unsigned)((H + ty - 1) / ty), (unsigned)D);
    const dim3 blk(kFwdThreads);
    // Two rows per wave (16 output channels) run the double-buffered kernel, the parity-class
    // form with wpar; one row per wave the register-staged one.  At 1024^2 x 128 the 16-channel
    // decoder layer took 20.5 ms register-staged with 27 taps, 13.8 with the parity form and
    // 11.4 double-buffered; the layers of 32 and 64 channels were slower double-buffered (e1
    // 2.16 against 1.70 ms, d1 8.96 against 7.33: its two buffers hold the workgroup to two per
    // CU) and slower in parity form (d1 8.05, d2 3.44 against 2.87 ms).
    const HeadArgs nohead{nullptr, nullptr, 0.0f, nullptr};
    if (head && !(par && Cout == 16 && relu)) return hipErrorInvalidValue;
    if (rows == 2) {
        if (par && head)
            hipLaunchKernelGGL((k_conv_dma<false, 2, true, true>), grid, blk, 0, s, in0, in1, g, w, wpar, bias, relu,
                               out, *head);
        else if (par)
            hipLaunchKernelGGL((k_conv_dma<false, 2, true>), grid, blk, 0, s, in0, in1, g, w, wpar, bias, relu, out,
                               nohead);
        else if (out_f32)
            hipLaunchKernelGGL((k_conv_dma<true, 2, false>), grid, blk, 0, s, in0, in1, g, w, nullptr, bias, relu, out,
                               nohead);
        else
            hipLaunchKernelGGL((k_conv_dma<false, 2, false>), grid, blk, 0, s, in0, in1, g, w, nullptr, bias, relu, out,
                               nohead);
    } else if (out_f32) {
        hipLaunchKernelGGL((k_conv_fwd<true, 1>), grid, blk, 0, s, in0, in1, g, w, bias, relu, out);
    } else {
        hipLaunchKernelGGL((k_conv_fwd<false, 1>), grid, blk, 0, s, in0, in1, g, w, bias, relu, out);
    }
    return hipGetLastError();
}

// rows per chunk so that about 8192 waves run and the partials stay within cap floats
// Rows per chunk so that about 4096 waves run and the partials stay within 16 M floats (and
// cap): the chunk sum reads nchunks x per floats, one sequential sum per weight, so many small
// chunks cost more in the sum than they gain in the MFMA pass (at 8,192 waves a 64 x 128 x 128
// training step spent 4.4 ms summing partials against 1.5 ms computing them).
int wgrad_chunks(long rows, int Cout, int Cin, long cap) {
    const long tiles = (long)(Cout / 16) * (Cin / 16);
    const long per = (long)kTaps * Cout * Cin + Cout;
    long n = std::max(1L, (4096 + tiles - 1) / tiles);
    n = std::min(n, rows);
    n = std::min(n, std::max(1L, std::min(cap, 16L << 20) / per));
    return (int)n;
}

hipError_t conv_wgrad(const unsigned short* in0, int C0, const unsigned short* in1, int C1, int D, int H, int W,
                      const unsigned short* gout, int Cout, float* part, long part_cap, float* dw_db, hipStream_t s) {
    if (C0 <= 0 || C0 % 16 || C1 < 0 || C1 % 16 || Cout <= 0 || Cout % 16 || D <= 0 || H <= 0 || W <= 0)
        return hipErrorInvalidValue;
    if (C1 > 0 && (D % 2 || H % 2 || W % 2 || !in1)) return hipErrorInvalidValue;
    const int Cin = C0 + C1;
    const long rows = (long)D * H;
    const int nch = wgrad_chunks(rows, Cout, Cin, part_cap);
    const long per = (long)kTaps * Cout * Cin + Cout;
    if ((long)nch * per > part_cap) return hipErrorInvalidValue;
    ConvGeo g{D, H, W, C0, C1, Cout, 16};
    hipLaunchKernelGGL(k_conv_wgrad, dim3(Cout / 16, Cin / 16, nch), dim3(64), 0, s, in0, in1, g, gout, nch, part);
    hipLaunchKernelGGL(k_sum_chunks, dim3((unsigned)((per + 15) / 16)), dim3(256), 0, s, (const float*)part, per,
                       nch, dw_db);
    return hipGetLastError();
}

unsigned blocks_of(long n, int bs = kEw) { return (unsigned)((n + bs - 1) / bs); }

}  // namespace net
}  // namespace rihvr

using namespace rihvr::net;

// ----------------------------------------------------------------------------
// The network object
// ----------------------------------------------------------------------------
namespace {

constexpr int kChan[kLevels] = {16, 32, 64, 128};

struct Layer {
    int level, C0, C1, Cout;  // C0: in0 channels, C1: in1 (half resolution) channels
    size_t w_off, b_off;       // offsets in the flat parameter array
    unsigned short *w = nullptr, *wT = nullptr;
    unsigned short* wpar = nullptr;  // parity-class weights of the in1 channels (decoder layers of <= 32 channels)
};

#define NET_HIP(call)                                                                              \
    do {                                                                                           \
        hipError_t e_ = (call);                                                                    \
        if (e_ != hipSuccess) {                                                                    \
            std::fprintf(stderr, "rihvr: HIP error %s at %s:%d\n", hipGetErrorString(e_), __FILE__, \
                         __LINE__);                                                                \
            return e_ == hipErrorOutOfMemory ? RIHVR_E_OUT_OF_MEMORY : RIHVR_E_DEVICE;              \
        }                                                                                          \
    } while (0)

// layer order of the flat parameter array: e0 e1 e2 e3 d2 d1 d0 (each W [27][Cout][Cin] then
// b [Cout]), then the head Wo [2][16], bo [2]
void layer_shapes(Layer* L, size_t* head_off, size_t* total) {
    const int spec[kNumConv][4] = {
        {0, 16, 0, 16},   // e0: x16 (2 channels used of 16)
        {1, 16, 0, 32},   // e1: pool(e0)
        {2, 32, 0, 64},   // e2: pool(e1)
        {3, 64, 0, 128},  // e3: pool(e2)
        {2, 64, 128, 64}, // d2: e2 ++ up(e3)
        {1, 32, 64, 32},  // d1: e1 ++ up(d2)
        {0, 16, 32, 16},  // d0: e0 ++ up(d1)
    };
    size_t off = 0;
    for (int i = 0; i < kNumConv; ++i) {
        L[i].level = spec[i][0];
        L[i].C0 = spec[i][1];
        L[i].C1 = spec[i][2];
        L[i].Cout = spec[i][3];
        L[i].w_off = off;
        off += (size_t)kTaps * L[i].Cout * (L[i].C0 + L[i].C1);
        L[i].b_off = off;
        off += L[i].Cout;
    }
    *head_off = off;
    *total = off + 34;
}

}  // namespace

struct rihvr_net {
    int D = 0, H = 0, W = 0;
    int Dk[kLevels], Hk[kLevels], Wk[kLevels];
    long Vk[kLevels];
    float in_scale = 1.0f, out_scale = 1.0f;
    float pos_weight = 0.0f;  // loss weight 1 + pos_weight where the target is non-zero
    hipStream_t stream = nullptr;
    Layer L[kNumConv];
    size_t head_off = 0, nparams = 0;
    float *params = nullptr, *grads = nullptr, *adam_m = nullptr, *adam_v = nullptr;
    int adam_t = 0;
    bool weights_ready = false;
    unsigned short* w0p = nullptr;  // the input layer's packed weights (k_conv_first)
    // activations
    unsigned short* x16 = nullptr;  // the 16-channel input (training: the input layer's weight gradient)
    unsigned short* e[kLevels] = {};
    unsigned short* p[kLevels] = {};  // p[1..3]
    unsigned short* d[3] = {};        // d0, d1, d2
    // training buffers (allocated at the first gradient call)
    bool train_ready = false;
    unsigned short* ge[kLevels] = {};
    unsigned short* gd[3] = {};
    float* dgd[3] = {};  // input gradients of the decoder convolutions d0, d1, d2
    float* dge[kLevels] = {};  // input gradients of e1, e2, e3 (= gradients of p1, p2, p3)
    float* wpart = nullptr;
    long wpart_cap = 0;
    float* hpart = nullptr;
    double* lpart = nullptr;
    double* h_lpart = nullptr;
    long head_blocks = 0;
};

namespace {

template <class T> rihvr_status nalloc(T** ptr, size_t count) {
    NET_HIP(hipMalloc((void**)ptr, std::max<size_t>(1, count) * sizeof(T)));
    return RIHVR_OK;
}

void net_release(rihvr_net* n) {
    if (!n) return;
    std::vector<void*> ptrs = {n->params, n->grads, n->adam_m, n->adam_v, n->x16, n->wpart, n->hpart, n->lpart,
                               n->w0p};
    for (int k = 0; k < kLevels; ++k) {
        ptrs.push_back(n->e[k]);
        ptrs.push_back(n->p[k]);
        ptrs.push_back(n->ge[k]);
        ptrs.push_back(n->dge[k]);
    }
    for (int k = 0; k < 3; ++k) {
        ptrs.push_back(n->d[k]);
        ptrs.push_back(n->gd[k]);
        ptrs.push_back(n->dgd[k]);
    }
    for (auto& l : n->L) {
        ptrs.push_back(l.w);
        ptrs.push_back(l.wT);
        ptrs.push_back(l.wpar);
    }
    for (void* q : ptrs) (void)hipFree(q);
    if (n->h_lpart) (void)hipHostFree(n->h_lpart);
    if (n->stream) (void)hipStreamDestroy(n->stream);
}

rihvr_status prep_weights(rihvr_net* n) {
    for (auto& l : n->L) {
        const int Cin = l.C0 + l.C1;
        const long cnt = (long)kTaps * l.Cout * Cin;
        hipLaunchKernelGGL(k_prep_weights, dim3(blocks_of(cnt)), dim3(kEw), 0, n->stream, n->params + l.w_off, l.Cout,
                           Cin, l.w, l.wT);
        if (l.wpar)
            hipLaunchKernelGGL(k_prep_parity, dim3(blocks_of(64L * l.Cout * l.C1)), dim3(kEw), 0, n->stream,
                               n->params + l.w_off, l.Cout, l.C0, l.C1, l.wpar);
    }
    hipLaunchKernelGGL(k_pack_first, dim3(4), dim3(256), 0, n->stream, n->params + n->L[0].w_off, n->w0p);
    NET_HIP(hipGetLastError());
    n->weights_ready = true;
    return RIHVR_OK;
}

// the convolution of layer i on its inputs; parity: in1's channels through the parity-class
// weights (inference), else through the 27 taps on the upsampled grid (training)
rihvr_status layer_forward(rihvr_net* n, int i, const unsigned short* in0, const unsigned short* in1,
                           unsigned short* out, bool parity = false, const HeadArgs* head = nullptr) {
    const Layer& l = n->L[i];
    const int k = l.level;
    NET_HIP(conv_forward(in0, l.C0, in1, l.C1, n->Dk[k], n->Hk[k], n->Wk[k], l.w, n->params + l.b_off, l.Cout, 1, out,
                         false, n->stream, parity ? l.wpar : nullptr, head));
    return RIHVR_OK;
}

rihvr_status pool(rihvr_net* n, int k, const unsigned short* e, unsigned short* p) {
    const long cnt = n->Vk[k + 1] * (kChan[k] / 4);
    hipLaunchKernelGGL(k_pool, dim3(blocks_of(cnt)), dim3(kEw), 0, n->stream, e, n->Dk[k], n->Hk[k], n->Wk[k],
                       kChan[k], p);
    NET_HIP(hipGetLastError());
    return RIHVR_OK;
}

// forward pass up to d0 (the head is applied by the caller); with_x16 also writes the 16-channel
// input the input layer's weight gradient reads
// with_x16: the training form (x16 kept for the backward pass, 27 taps everywhere); else the
// prediction, whose head is fused into the last decoder layer and writes pred (complex volume)
rihvr_status forward_body(rihvr_net* n, const float* in, bool with_x16, float* pred = nullptr) {
    if (!n->weights_ready) {
        const rihvr_status st = prep_weights(n);
        if (st) return st;
    }
    if (with_x16)
        hipLaunchKernelGGL(k_input16, dim3(blocks_of(n->Vk[0])), dim3(kEw), 0, n->stream, (const float2*)in,
                           n->Vk[0], n->in_scale, n->x16);
    hipLaunchKernelGGL(k_conv_first, dim3((unsigned)((n->W + 63) / 64), (unsigned)((n->H + 7) / 8),
                                          (unsigned)((n->D + kFirstZ - 1) / kFirstZ)),
                       dim3(kFwdThreads), 0, n->stream, (const float2*)in, n->D, n->H, n->W, n->in_scale, n->w0p,
                       n->params + n->L[0].b_off, n->e[0], n->p[1]);  // (its pool fused: p[1])
    NET_HIP(hipGetLastError());
    rihvr_status st;
    for (int k = 1; k < kLevels; ++k) {
        if (k > 1 && (st = pool(n, k - 1, n->e[k - 1], n->p[k]))) return st;
        if ((st = layer_forward(n, k, n->p[k], nullptr, n->e[k]))) return st;
    }
    const bool parity = !with_x16;  // prediction: the decoder's upsampled channels in parity-class form
    if ((st = layer_forward(n, 4, n->e[2], n->e[3], n->d[2], parity))) return st;
    if ((st = layer_forward(n, 5, n->e[1], n->d[2], n->d[1], parity))) return st;
    if (parity && pred) {
        const HeadArgs h{n->params + n->head_off, n->params + n->head_off + 32, n->out_scale, (float2*)pred};
        return layer_forward(n, 6, n->e[0], n->d[1], n->d[0], true, &h);
    }
    if ((st = layer_forward(n, 6, n->e[0], n->d[1], n->d[0], parity))) return st;
    return RIHVR_OK;
}

rihvr_status alloc_training(rihvr_net* n) {
    if (n->train_ready) return RIHVR_OK;
    rihvr_status st;
    if ((st = nalloc(&n->x16, (size_t)n->Vk[0] * 16))) return st;
    for (int k = 0; k < kLevels; ++k)
        if ((st = nalloc(&n->ge[k], (size_t)n->Vk[k] * kChan[k]))) return st;
    for (int k = 0; k < 3; ++k) {
        if ((st = nalloc(&n->gd[k], (size_t)n->Vk[k] * kChan[k]))) return st;
        // d_k's input is e_k ++ up(d_{k+1} or e3): C_k + C_{k+1} channels at level k
        if ((st = nalloc(&n->dgd[k], (size_t)n->Vk[k] * (kChan[k] + kChan[k + 1])))) return st;
    }
    for (int k = 1; k < kLevels; ++k)
        if ((st = nalloc(&n->dge[k], (size_t)n->Vk[k] * kChan[k - 1]))) return st;
    // weight-gradient partials: at least one chunk of the largest layer, at most 64 M floats
    long maxper = 0;
    for (auto& l : n->L) maxper = std::max(maxper, (long)kTaps * l.Cout * (l.C0 + l.C1) + l.Cout);
    n->wpart_cap = std::max(maxper, std::min(64L << 20, maxper * 1024));
    if ((st = nalloc(&n->wpart, (size_t)n->wpart_cap))) return st;
    if ((st = nalloc(&n->grads, n->nparams))) return st;
    if ((st = nalloc(&n->adam_m, n->nparams))) return st;
    if ((st = nalloc(&n->adam_v, n->nparams))) return st;
    NET_HIP(hipMemsetAsync(n->adam_m, 0, n->nparams * sizeof(float), n->stream));
    NET_HIP(hipMemsetAsync(n->adam_v, 0, n->nparams * sizeof(float), n->stream));
    n->adam_t = 0;
    if ((st = nalloc(&n->hpart, (size_t)n->head_blocks * 34))) return st;
    n->train_ready = true;
    return RIHVR_OK;
}

rihvr_status layer_backward(rihvr_net* n, int i, const unsigned short* in0, const unsigned short* in1,
                            const unsigned short* gout, float* din) {
    const Layer& l = n->L[i];
    const int k = l.level;
    NET_HIP(conv_wgrad(in0, l.C0, in1, l.C1, n->Dk[k], n->Hk[k], n->Wk[k], gout, l.Cout, n->wpart, n->wpart_cap,
                       n->grads + l.w_off, n->stream));
    if (din)
        NET_HIP(conv_forward(gout, l.Cout, nullptr, 0, n->Dk[k], n->Hk[k], n->Wk[k], l.wT, nullptr, l.C0 + l.C1, 0,
                             din, true, n->stream));
    return RIHVR_OK;
}

rihvr_status finalize(rihvr_net* n, int k, const unsigned short* act, const float* skip, int sk_stride,
                      const float* pool_g, const float* split, int off, int Ct, unsigned short* out) {
    Finalize f{act, skip, sk_stride, pool_g, split, off, Ct, n->Dk[k], n->Hk[k], n->Wk[k], kChan[k], out};
    hipLaunchKernelGGL(k_finalize, dim3(blocks_of(n->Vk[k] * (kChan[k] / 4))), dim3(kEw), 0, n->stream, f);
    NET_HIP(hipGetLastError());
    return RIHVR_OK;
}

// loss and every parameter gradient for one (input, target) pair, into n->grads
rihvr_status loss_grad(rihvr_net* n, const float* in, const float* target, double* loss) {
    rihvr_status st;
    if ((st = alloc_training(n))) return st;
    if ((st = forward_body(n, in, true))) return st;
    const float* wo = n->params + n->head_off;
    const float* bo = wo + 32;
    hipLaunchKernelGGL(k_head_fwd, dim3((unsigned)n->head_blocks), dim3(kEw), 0, n->stream, n->d[0], n->Vk[0], wo, bo,
                       n->out_scale, (float2*)nullptr, (const float2*)target, n->pos_weight, n->lpart);
    hipLaunchKernelGGL(k_head_bwd, dim3((unsigned)n->head_blocks), dim3(kEw), 0, n->stream, n->d[0], n->Vk[0], wo, bo,
                       n->out_scale, (const float2*)target, n->pos_weight, n->gd[0], n->hpart);
    hipLaunchKernelGGL(k_sum_chunks_wide, dim3(34), dim3(256), 0, n->stream, (const float*)n->hpart, 34L,
                       (int)n->head_blocks, n->grads + n->head_off);
    NET_HIP(hipGetLastError());
    // decoder
    if ((st = layer_backward(n, 6, n->e[0], n->d[1], n->gd[0], n->dgd[0]))) return st;
    if ((st = finalize(n, 1, n->d[1], nullptr, 0, nullptr, n->dgd[0], kChan[0], kChan[0] + kChan[1], n->gd[1])))
        return st;
    if ((st = layer_backward(n, 5, n->e[1], n->d[2], n->gd[1], n->dgd[1]))) return st;
    if ((st = finalize(n, 2, n->d[2], nullptr, 0, nullptr, n->dgd[1], kChan[1], kChan[1] + kChan[2], n->gd[2])))
        return st;
    if ((st = layer_backward(n, 4, n->e[2], n->e[3], n->gd[2], n->dgd[2]))) return st;
    if ((st = finalize(n, 3, n->e[3], nullptr, 0, nullptr, n->dgd[2], kChan[2], kChan[2] + kChan[3], n->ge[3])))
        return st;
    // encoder
    if ((st = layer_backward(n, 3, n->p[3], nullptr, n->ge[3], n->dge[3]))) return st;
    if ((st = finalize(n, 2, n->e[2], n->dgd[2], kChan[2] + kChan[3], n->dge[3], nullptr, 0, 0, n->ge[2]))) return st;
    if ((st = layer_backward(n, 2, n->p[2], nullptr, n->ge[2], n->dge[2]))) return st;
    if ((st = finalize(n, 1, n->e[1], n->dgd[1], kChan[1] + kChan[2], n->dge[2], nullptr, 0, 0, n->ge[1]))) return st;
    if ((st = layer_backward(n, 1, n->p[1], nullptr, n->ge[1], n->dge[1]))) return st;
    if ((st = finalize(n, 0, n->e[0], n->dgd[0], kChan[0] + kChan[1], n->dge[1], nullptr, 0, 0, n->ge[0]))) return st;
    if ((st = layer_backward(n, 0, n->x16, nullptr, n->ge[0], nullptr))) return st;
    NET_HIP(hipMemcpyAsync(n->h_lpart, n->lpart, n->head_blocks * sizeof(double), hipMemcpyDeviceToHost, n->stream));
    NET_HIP(hipStreamSynchronize(n->stream));
    if (loss) {
        double s = 0.0;
        for (long b = 0; b < n->head_blocks; ++b) s += n->h_lpart[b];
        *loss = s / (2.0 * (double)n->Vk[0]);
    }
    return RIHVR_OK;
}

rihvr_status adam_step(rihvr_net* n, float lr) {
    if (!n->train_ready) return RIHVR_E_INVALID_STATE;
    n->adam_t += 1;
    const float bc1 = (float)(1.0 - std::pow((double)kB1, n->adam_t));
    const float bc2 = (float)(1.0 - std::pow((double)kB2, n->adam_t));
    hipLaunchKernelGGL(k_adam, dim3(blocks_of((long)n->nparams)), dim3(kEw), 0, n->stream, n->params, n->grads,
                       n->adam_m, n->adam_v, (long)n->nparams, lr, bc1, bc2);
    NET_HIP(hipGetLastError());
    return prep_weights(n);
}

// splitmix64 + Box-Muller normals for the initial weights
struct Rng {
    uint64_t s;
    uint64_t next() {
        uint64_t z = (s += 0x9E3779B97F4A7C15ull);
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        return z ^ (z >> 31);
    }
    double uniform() { return ((next() >> 11) + 0.5) * (1.0 / 9007199254740992.0); }
    double normal() { return std::sqrt(-2.0 * std::log(uniform())) * std::cos(2.0 * 3.14159265358979323846 * uniform()); }
};

}  // namespace

namespace rihvr {
namespace net {

rihvr_status predict(rihvr_net* n, const float* in_dev, float* out_dev, hipStream_t wait_on) {
    if (!n || !in_dev || !out_dev) return RIHVR_E_INVALID_ARGUMENT;
    if (wait_on) {
        hipEvent_t ev;
        NET_HIP(hipEventCreateWithFlags(&ev, hipEventDisableTiming));
        NET_HIP(hipEventRecord(ev, wait_on));
        NET_HIP(hipStreamWaitEvent(n->stream, ev, 0));
        NET_HIP(hipEventDestroy(ev));
    }
    rihvr_status st;
    if ((st = forward_body(n, in_dev, false, out_dev))) return st;
    NET_HIP(hipStreamSynchronize(n->stream));
    return RIHVR_OK;
}

bool net_matches(const rihvr_net* n, int D, int H, int W) { return n && n->D == D && n->H == H && n->W == W; }

}  // namespace net
}  // namespace rihvr

extern "C" {

rihvr_status rihvr_net_param_count(int64_t* count) {
    if (!count) return RIHVR_E_INVALID_ARGUMENT;
    Layer L[kNumConv];
    size_t head = 0, total = 0;
    layer_shapes(L, &head, &total);
    *count = (int64_t)total;
    return RIHVR_OK;
}

rihvr_status rihvr_net_create(int32_t num_planes, int32_t height, int32_t width, rihvr_net** out) {
    if (!out) return RIHVR_E_INVALID_ARGUMENT;
    *out = nullptr;
    const int f = 1 << (kLevels - 1);
    if (num_planes <= 0 || height <= 0 || width <= 0 || num_planes % f || height % f || width % f)
        return RIHVR_E_INVALID_ARGUMENT;
    if (num_planes > 65535 || (long)height * width > (1L << 30)) return RIHVR_E_INVALID_ARGUMENT;
    rihvr_net* n = new rihvr_net();
    n->D = num_planes;
    n->H = height;
    n->W = width;
    for (int k = 0; k < kLevels; ++k) {
        n->Dk[k] = num_planes >> k;
        n->Hk[k] = height >> k;
        n->Wk[k] = width >> k;
        n->Vk[k] = (long)n->Dk[k] * n->Hk[k] * n->Wk[k];
    }
    layer_shapes(n->L, &n->head_off, &n->nparams);
    n->head_blocks = (n->Vk[0] + kEw - 1) / kEw;
    auto fail = [&](rihvr_status s) {
        net_release(n);
        delete n;
        return s;
    };
    if (hipStreamCreateWithFlags(&n->stream, hipStreamNonBlocking) != hipSuccess) return fail(RIHVR_E_DEVICE);
    rihvr_status st;
    if ((st = nalloc(&n->params, n->nparams))) return fail(st);
    if (hipMemset(n->params, 0, n->nparams * sizeof(float)) != hipSuccess) return fail(RIHVR_E_DEVICE);
    for (auto& l : n->L) {
        const size_t cnt = (size_t)kTaps * l.Cout * (l.C0 + l.C1);
        if ((st = nalloc(&l.w, cnt)) || (st = nalloc(&l.wT, cnt))) return fail(st);
        if (l.C1 > 0 && l.Cout <= 32 && (st = nalloc(&l.wpar, (size_t)64 * l.Cout * l.C1))) return fail(st);
    }
    if ((st = nalloc(&n->w0p, 4 * 16 * 16))) return fail(st);
    for (int k = 0; k < kLevels; ++k)
        if ((st = nalloc(&n->e[k], (size_t)n->Vk[k] * kChan[k]))) return fail(st);
    for (int k = 1; k < kLevels; ++k)
        if ((st = nalloc(&n->p[k], (size_t)n->Vk[k] * kChan[k - 1]))) return fail(st);
    for (int k = 0; k < 3; ++k)
        if ((st = nalloc(&n->d[k], (size_t)n->Vk[k] * kChan[k]))) return fail(st);
    if ((st = nalloc(&n->lpart, (size_t)n->head_blocks))) return fail(st);
    if (hipHostMalloc((void**)&n->h_lpart, n->head_blocks * sizeof(double), hipHostMallocDefault) != hipSuccess)
        return fail(RIHVR_E_OUT_OF_MEMORY);
    *out = n;
    return RIHVR_OK;
}

rihvr_status rihvr_net_destroy(rihvr_net* net) {
    if (!net) return RIHVR_OK;
    (void)hipStreamSynchronize(net->stream);
    net_release(net);
    delete net;
    return RIHVR_OK;
}

rihvr_status rihvr_net_init(rihvr_net* net, uint64_t seed) {
    if (!net) return RIHVR_E_INVALID_ARGUMENT;
    std::vector<float> p(net->nparams, 0.0f);
    Rng rng{seed};
    for (int i = 0; i < kNumConv; ++i) {
        const Layer& l = net->L[i];
        const int Cin = l.C0 + l.C1;
        // He initialisation over the used fan-in (the input layer reads 2 of its 16 channels)
        const int used = i == 0 ? 2 : Cin;
        const double sd = std::sqrt(2.0 / (double)(kTaps * used));
        for (int t = 0; t < kTaps; ++t)
            for (int co = 0; co < l.Cout; ++co)
                for (int ci = 0; ci < used; ++ci)
                    p[l.w_off + ((size_t)t * l.Cout + co) * Cin + ci] = (float)(sd * rng.normal());
    }
    const double sdo = std::sqrt(1.0 / 16.0);
    for (int k = 0; k < 32; ++k) p[net->head_off + k] = (float)(sdo * rng.normal());
    NET_HIP(hipMemcpy(net->params, p.data(), p.size() * sizeof(float), hipMemcpyHostToDevice));
    if (net->train_ready) {
        NET_HIP(hipMemset(net->adam_m, 0, net->nparams * sizeof(float)));
        NET_HIP(hipMemset(net->adam_v, 0, net->nparams * sizeof(float)));
        net->adam_t = 0;
    }
    net->weights_ready = false;
    return RIHVR_OK;
}

rihvr_status rihvr_net_set_params(rihvr_net* net, const float* params, float in_scale, float out_scale) {
    if (!net || !params || !(in_scale > 0) || !(out_scale > 0)) return RIHVR_E_INVALID_ARGUMENT;
    NET_HIP(hipMemcpy(net->params, params, net->nparams * sizeof(float), hipMemcpyHostToDevice));
    net->in_scale = in_scale;
    net->out_scale = out_scale;
    net->weights_ready = false;
    if (net->train_ready) {
        NET_HIP(hipMemset(net->adam_m, 0, net->nparams * sizeof(float)));
        NET_HIP(hipMemset(net->adam_v, 0, net->nparams * sizeof(float)));
        net->adam_t = 0;
    }
    return RIHVR_OK;
}

rihvr_status rihvr_net_set_loss_weight(rihvr_net* net, float pos_weight) {
    if (!net || !(pos_weight >= 0.0f)) return RIHVR_E_INVALID_ARGUMENT;
    net->pos_weight = pos_weight;
    return RIHVR_OK;
}

rihvr_status rihvr_net_get_params(rihvr_net* net, float* params, float* in_scale, float* out_scale) {
    if (!net) return RIHVR_E_INVALID_ARGUMENT;
    if (params) {
        NET_HIP(hipStreamSynchronize(net->stream));
        NET_HIP(hipMemcpy(params, net->params, net->nparams * sizeof(float), hipMemcpyDeviceToHost));
    }
    if (in_scale) *in_scale = net->in_scale;
    if (out_scale) *out_scale = net->out_scale;
    return RIHVR_OK;
}

rihvr_status rihvr_net_predict(rihvr_net* net, const float* volume_in, float* volume_out) {
    return rihvr::net::predict(net, volume_in, volume_out, nullptr);
}

rihvr_status rihvr_net_loss_grad(rihvr_net* net, const float* volume_in, const float* target, double* loss,
                                 float* grads) {
    if (!net || !volume_in || !target) return RIHVR_E_INVALID_ARGUMENT;
    const rihvr_status st = loss_grad(net, volume_in, target, loss);
    if (st) return st;
    if (grads) NET_HIP(hipMemcpy(grads, net->grads, net->nparams * sizeof(float), hipMemcpyDeviceToHost));
    return RIHVR_OK;
}

rihvr_status rihvr_net_adam_step(rihvr_net* net, float lr) {
    if (!net || !(lr > 0)) return RIHVR_E_INVALID_ARGUMENT;
    return adam_step(net, lr);
}

rihvr_status rihvr_net_train_step(rihvr_net* net, const float* volume_in, const float* target, float lr,
                                  double* loss) {
    if (!net || !volume_in || !target || !(lr > 0)) return RIHVR_E_INVALID_ARGUMENT;
    rihvr_status st = loss_grad(net, volume_in, target, loss);
    if (st) return st;
    return adam_step(net, lr);
}

rihvr_status rihvr_net_sync(rihvr_net* net) {
    if (!net) return RIHVR_E_INVALID_ARGUMENT;
    NET_HIP(hipStreamSynchronize(net->stream));
    return RIHVR_OK;
}

rihvr_status rihvr_conv3d_bf16_forward(const uint16_t* in0, int32_t c0, const uint16_t* in1, int32_t c1, int32_t depth,
                                       int32_t height, int32_t width, const uint16_t* weight, const float* bias,
                                       int32_t cout, int32_t relu, uint16_t* out) {
    if (!in0 || !weight || !out) return RIHVR_E_INVALID_ARGUMENT;
    const hipError_t e = conv_forward(in0, c0, in1, c1, depth, height, width, weight, bias, cout, relu, out, false, 0);
    if (e == hipErrorInvalidValue) return RIHVR_E_INVALID_ARGUMENT;
    NET_HIP(e);
    NET_HIP(hipDeviceSynchronize());
    return RIHVR_OK;
}

rihvr_status rihvr_conv3d_bf16_dgrad(const uint16_t* grad_out, int32_t cout, int32_t depth, int32_t height,
                                     int32_t width, const uint16_t* weight_t, int32_t cin, float* grad_in) {
    if (!grad_out || !weight_t || !grad_in) return RIHVR_E_INVALID_ARGUMENT;
    const hipError_t e =
        conv_forward(grad_out, cout, nullptr, 0, depth, height, width, weight_t, nullptr, cin, 0, grad_in, true, 0);
    if (e == hipErrorInvalidValue) return RIHVR_E_INVALID_ARGUMENT;
    NET_HIP(e);
    NET_HIP(hipDeviceSynchronize());
    return RIHVR_OK;
}

rihvr_status rihvr_conv3d_bf16_wgrad(const uint16_t* in0, int32_t c0, const uint16_t* in1, int32_t c1, int32_t depth,
                                     int32_t height, int32_t width, const uint16_t* grad_out, int32_t cout,
                                     float* grad_weight_bias) {
    if (!in0 || !grad_out || !grad_weight_bias) return RIHVR_E_INVALID_ARGUMENT;
    if (c0 <= 0 || c0 % 16 || c1 < 0 || c1 % 16 || cout <= 0 || cout % 16) return RIHVR_E_INVALID_ARGUMENT;
    const long per = (long)kTaps * cout * (c0 + c1) + cout;
    const long cap = std::max(per, std::min(16L << 20, per * 256));
    float* part = nullptr;
    NET_HIP(hipMalloc((void**)&part, cap * sizeof(float)));
    const hipError_t e = conv_wgrad(in0, c0, in1, c1, depth, height, width, grad_out, cout, part, cap,
                                    grad_weight_bias, 0);
    const hipError_t s = hipDeviceSynchronize();
    (void)hipFree(part);
    if (e == hipErrorInvalidValue) return RIHVR_E_INVALID_ARGUMENT;
    NET_HIP(e);
    NET_HIP(s);
    return RIHVR_OK;
}

}  // extern "C"
