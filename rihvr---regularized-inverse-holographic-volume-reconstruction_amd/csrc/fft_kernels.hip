// fft_kernels.hip -- row and column FFT kernels of the forward operator A, the
// adjoint A^T and the hologram transforms, with their prologues/epilogues fused.
//
// Layouts: planes [z][y][x] complex64, x fastest.  Row kernels transform along
// x (length W), column kernels along y (length H) on strips of 16 columns, so
// every global access is a 128-byte row segment.
#include "fft.hpp"
#include "kernels.hpp"

namespace rihvr {

constexpr int kColStrip = 16;  // columns per column-kernel workgroup

template <int L> struct RowCfg {
    static constexpr int TPT = RowPlan<L>::TPT;
    static constexpr int RPW = (256 / TPT) > 0 ? (256 / TPT) : 1;  // rows per workgroup
    static constexpr int THREADS = TPT * RPW;
    // LDS row: element j at j + j/16 (one pad slot per 16 elements breaks the
    // 16-element store stride of a radix-16 pass across the banks)
    static constexpr int LP = L + L / 16 + 1;
};
enum ColMode { COL_PLAIN = 0, COL_ADJ = 1, COL_FWD = 2, COL_REAL = 3 };
// Column plans: the dense passes use FftPlan (32 x 32 at 1024: one LDS exchange); the sparse
// forward pass keeps 16 x 8 x 8 on 64 threads per column (its workgroups mostly leave at once
// or load few rows, and more threads per column finish a strip sooner: measured 0.060 ms
// against 0.066 ms with 32 x 32 at 1024^2 x 128).
template <int L, int MODE> struct ColPlan : FftPlan<L> {};
template <> struct ColPlan<1024, COL_FWD> {
    using radices = Radices<16, 8, 8>;
    static constexpr int first = 16;
    static constexpr int TPT = 1024 / 16;
};
template <int L, int MODE> struct ColCfg {
    static constexpr int TPT = ColPlan<L, MODE>::TPT;
    static constexpr int THREADS = TPT * kColStrip;
};

template <int L>
__device__ __forceinline__ void load_twiddles(float2* tw_s, const float2* __restrict__ twg) {
    for (int m = threadIdx.x; m < L; m += blockDim.x) tw_s[m] = twg[m];
}

// ----------------------------------------------------------------------------
// Row kernels
// ----------------------------------------------------------------------------
enum RowMode { ROW_PLAIN = 0, ROW_ADJ = 1, ROW_FWDACC = 2, ROW_SUMPART = 3 };

struct RowArgs {
    const float2* in;       // PLAIN: [nrows][L]; FWD: volume; SUMPART: [nparts][H][L]
    float2* out;            // [nrows][L]
    int nrows;              // rows per plane (H) for ADJ/FWD; total rows otherwise
    const float2* twg;      // W_L^m table for the transform sign
    // ADJ
    const float2* rhat;     // [H][L] spectrum of the residual
    const double* kappa;    // [H][L]
    const float* zlist;     // plane positions
    int nz, zchunk;
    float invN;
    // (FWDACC uses kappa, zlist, nz, zchunk, invN, part_stride)
    // SUMPART
    int nparts;
    long part_stride;
    // FWDACC: 16-column strips present in the input planes ([nz][L/16]; 0 = read as zeros)
    const uint8_t* smap;
};

// The row passes' twiddles in registers across planes (TwCache) and the forward pass's
// next-plane prefetch: at L = 1024 (8.8.8.2 on 128 threads), where they were measured; at the
// other lengths, whose plans hold more elements per thread, the extra registers cost waves per
// SIMD (576, 768, 832: 2 -> 1), so those keep the plain passes.
template <int L> constexpr bool row_regs_fast() { return L == 1024; }

template <int L, int SIGN, int MODE>
__global__ void __launch_bounds__(RowCfg<L>::THREADS) k_rows(RowArgs a) {
    using C = RowCfg<L>;
    using Plan = RowPlan<L>;
    __shared__ float2 tw_s[L];
    __shared__ float2 buf[C::RPW * C::LP];
    load_twiddles<L>(tw_s, a.twg);
    const int tr = threadIdx.x / C::TPT;
    const int t = threadIdx.x - tr * C::TPT;
    float2* mybuf = buf + tr * C::LP;
    auto lds = [&](int j) -> float2& { return mybuf[j + (j >> 4)]; };
    const int row = blockIdx.x * C::RPW + tr;  // row within a plane (ADJ/FWD) or global row
    const bool valid = row < a.nrows;
    const int rrow = valid ? row : a.nrows - 1;
    __syncthreads();

    if constexpr (MODE == ROW_PLAIN) {
        const float2* src_row = a.in + (size_t)rrow * L;
        float2* dst_row = a.out + (size_t)rrow * L;
        FftRun<L, SIGN, typename Plan::radices, Plan::TPT>::run(
            tw_s, t, [&](int, int, int j) { return src_row[j]; },
            [&](int, int, int j, float2 v) { if (valid) dst_row[j] = v; }, lds);
    } else if constexpr (MODE == ROW_SUMPART) {
        const float2* src_row = a.in + (size_t)rrow * L;
        float2* dst_row = a.out + (size_t)rrow * L;
        FftRun<L, SIGN, typename Plan::radices, Plan::TPT>::run(
            tw_s, t,
            [&](int, int, int j) {
                float2 s = src_row[j];
                for (int p = 1; p < a.nparts; ++p) s = cadd(s, src_row[(size_t)p * a.part_stride + j]);
                return s;
            },
            [&](int, int, int j, float2 v) { if (valid) dst_row[j] = v; }, lds);
    } else if constexpr (MODE == ROW_FWDACC) {
        // A, second pass: forward row FFT of the column-transformed planes of a
        // chunk of planes; epilogue H_{-z} summed over the chunk (the 1 / (Nx Ny) is applied
        // to the first pass's input)
        // (sparse_rs_phase_mult_kernel + sparse_add_plane_kernel)
        using LPs = LastPass<L, Plan>;
        float2 acc[LPs::NB][LPs::R];
        double kp[LPs::NB][LPs::R];
        const double* krow_p = a.kappa + (size_t)rrow * L;
#pragma unroll
        for (int i = 0; i < LPs::NB; ++i) {
            const int b = t + i * LPs::TPT;
            const int bb = b < LPs::NBF ? b : LPs::NBF - 1;
#pragma unroll
            for (int r = 0; r < LPs::R; ++r) {
                acc[i][r] = make_float2(0.0f, 0.0f);
                kp[i][r] = krow_p[bb + r * LPs::NBF];
            }
        }
        const int z0 = blockIdx.y * a.zchunk;
        const int z1 = min(a.nz, z0 + a.zchunk);
        constexpr int NS = L / 16;  // strips per row (<= 64)
        const int lane = threadIdx.x & 63;
        // The plane's strips as a wave-uniform mask (every wave loads the same bytes); a plane
        // without any strip adds zeros to the accumulators (never -0: they start at +0), so it
        // is skipped; absent strips are read as zeros.  The same for every wave: no barrier is
        // skipped by part of the workgroup.
        auto plane_mask = [&](int z) -> uint64_t {
            const uint8_t* sm = a.smap + (size_t)z * NS;
            return __ballot(lane < NS && sm[lane < NS ? lane : 0] != 0);
        };
        if constexpr (!row_regs_fast<L>()) {
            for (int z = z0; z < z1; ++z) {
                const uint64_t smask = plane_mask(z);
                if (!smask) continue;
                const double mz = -(double)a.zlist[z];
                const float2* src_row = a.in + ((size_t)z * a.nrows + rrow) * L;
                FftRun<L, SIGN, typename Plan::radices, Plan::TPT>::run(
                    tw_s, t,
                    [&](int, int, int j) {
                        return ((smask >> (j >> 4)) & 1ull) ? src_row[j] : make_float2(0.0f, 0.0f);
                    },
                    [&](int i, int r, int, float2 v) {
                        float2 h = transfer(mz, kp[i][r]);
                        acc[i][r] = cadd(acc[i][r], cmul(h, v));  // 1/N is in the first pass's input
                    },
                    lds);
                __syncthreads();
            }
        } else {
            TwCache<L, SIGN, typename Plan::radices, Plan::TPT> twc;
            twc.load(tw_s, t);
            // The first pass's elements of the next plane with data are loaded while the current
            // plane is transformed (element b + r L/R0 of butterfly b = t + i TPT).
            constexpr int R0 = Plan::first;
            constexpr int NBF0 = L / R0;
            constexpr int NB0 = (NBF0 + Plan::TPT - 1) / Plan::TPT;
            float2 pre[NB0][R0];
            auto prefetch = [&](int z, uint64_t m) {
                const float2* src_row = a.in + ((size_t)z * a.nrows + rrow) * L;
#pragma unroll
                for (int i = 0; i < NB0; ++i) {
                    const int b = t + i * Plan::TPT;
                    const int bb = b < NBF0 ? b : NBF0 - 1;
#pragma unroll
                    for (int r = 0; r < R0; ++r) {
                        const int j = bb + r * NBF0;
                        pre[i][r] = ((m >> (j >> 4)) & 1ull) ? src_row[j] : make_float2(0.0f, 0.0f);
                    }
                }
            };
            int z = z0;
            uint64_t smask = 0;
            while (z < z1 && !(smask = plane_mask(z))) ++z;
            if (z < z1) prefetch(z, smask);
            while (z < z1) {
                int zn = z + 1;
                uint64_t nmask = 0;
                while (zn < z1 && !(nmask = plane_mask(zn))) ++zn;
                float2 cur[NB0][R0];
#pragma unroll
                for (int i = 0; i < NB0; ++i)
#pragma unroll
                    for (int r = 0; r < R0; ++r) cur[i][r] = pre[i][r];
                if (zn < z1) prefetch(zn, nmask);
                const double mz = -(double)a.zlist[z];
                twc.run(
                    tw_s, t, [&](int i, int r, int) { return cur[i][r]; },
                    [&](int i, int r, int, float2 v) {
                        const float2 h = transfer(mz, kp[i][r]);
                        acc[i][r] = cadd(acc[i][r], cmul(h, v));  // 1/N is in the first pass's input
                    },
                    lds);
                __syncthreads();
                z = zn;
            }
        }
        if (valid) {
            float2* dst_row = a.out + (size_t)blockIdx.y * a.part_stride + (size_t)rrow * L;
#pragma unroll
            for (int i = 0; i < LPs::NB; ++i) {
                const int b = t + i * LPs::TPT;
                if (b < LPs::NBF) {
#pragma unroll
                    for (int r = 0; r < LPs::R; ++r) dst_row[b + r * LPs::NBF] = acc[i][r];
                }
            }
        }
    } else {  // ROW_ADJ: H_z * R_hat / N, then inverse FFT along x, for a chunk of planes
        constexpr int NBF0 = L / Plan::first;
        constexpr int NB0 = (NBF0 + C::TPT - 1) / C::TPT;
        float2 rh[NB0][Plan::first];
        double kp[NB0][Plan::first];
        const float2* rrow_p = a.rhat + (size_t)rrow * L;
        const double* krow_p = a.kappa + (size_t)rrow * L;
#pragma unroll
        for (int i = 0; i < NB0; ++i) {
            const int b = t + i * C::TPT;
            const int bb = b < NBF0 ? b : NBF0 - 1;
#pragma unroll
            for (int r = 0; r < Plan::first; ++r) {
                rh[i][r] = rrow_p[bb + r * NBF0];
                kp[i][r] = krow_p[bb + r * NBF0];
            }
        }
        const int z0 = blockIdx.y * a.zchunk;
        const int z1 = min(a.nz, z0 + a.zchunk);
        auto plane = [&](int z, auto&& fft) {
            const double zz = (double)a.zlist[z];
            float2* dst_row = a.out + ((size_t)z * a.nrows + rrow) * L;
            fft([&](int i, int r, int) {
                    float2 h = transfer(zz, kp[i][r]);
                    return cscale(cmul(h, rh[i][r]), a.invN);
                },
                [&](int, int, int j, float2 v) { if (valid) dst_row[j] = v; });
            __syncthreads();
        };
        if constexpr (row_regs_fast<L>()) {
            TwCache<L, SIGN, typename Plan::radices, Plan::TPT> twc;
            twc.load(tw_s, t);
            for (int z = z0; z < z1; ++z)
                plane(z, [&](auto&& src, auto&& dst) { twc.run(tw_s, t, src, dst, lds); });
        } else {
            for (int z = z0; z < z1; ++z)
                plane(z, [&](auto&& src, auto&& dst) {
                    FftRun<L, SIGN, typename Plan::radices, Plan::TPT>::run(tw_s, t, src, dst, lds);
                });
        }
    }
}

// ----------------------------------------------------------------------------
// Column kernels: a workgroup transforms 16 adjacent columns of one plane
// (or of a chunk of planes, accumulating).  LDS layout [j][16] keeps the 16
// columns of one row contiguous.
// ----------------------------------------------------------------------------
struct ColArgs {
    const float2* in;       // [planes][L][W]
    float2* out;            // PLAIN/ADJ: [planes][L][W]; FWDACC: [chunks][L][W]; REAL: est [L][W]
    int W;
    const float2* twg;
    // ADJ, FWD
    const float2* phase;
    // REAL
    const float2* meas;
    float2* resid;
    double* partial;        // one double per workgroup: sum |r|^2
    // FWD: line occupancy of the input ([nz][L] row words and [nz] plane words, may be NULL)
    // and the strip map written
    const uint64_t* occ;
    const uint64_t* occ_planes;
    uint8_t* smap;
};

template <int L, int SIGN, int MODE>
__global__ void __launch_bounds__((ColCfg<L, MODE>::THREADS)) k_cols(ColArgs a) {
    using C = ColCfg<L, MODE>;
    using Plan = ColPlan<L, MODE>;
    if constexpr (MODE == COL_FWD) {
        // a strip with no line set in its plane transforms to zeros: not written, and the
        // workgroup leaves before anything else (uniform: every thread reads the same word)
        if (a.occ_planes && !((a.occ_planes[blockIdx.y] >> blockIdx.x) & 1ull)) {
            if (threadIdx.x == 0) a.smap[(size_t)blockIdx.y * (a.W / kColStrip) + blockIdx.x] = 0;
            return;
        }
    }
    __shared__ float2 tw_s[L];
    __shared__ float2 buf[L * kColStrip];
    __shared__ double red_s[C::THREADS / 64 + 1];
    load_twiddles<L>(tw_s, a.twg);
    const int ci = threadIdx.x % kColStrip;
    const int t = threadIdx.x / kColStrip;
    const int col = blockIdx.x * kColStrip + ci;
    auto lds = [&](int j) -> float2& { return buf[j * kColStrip + ci]; };
    const size_t W = (size_t)a.W;
    __syncthreads();

    if constexpr (MODE == COL_PLAIN || MODE == COL_ADJ) {
        const int z = blockIdx.y;
        const float2* src = a.in + (size_t)z * L * W + col;
        float2* dst = a.out + (size_t)z * L * W + col;
        float2 ph = make_float2(1.0f, 0.0f);
        if constexpr (MODE == COL_ADJ) ph = a.phase[z];  // e^{+i phi_z} (sparse_phase_scale_kernel2)
        FftRun<L, SIGN, typename Plan::radices, Plan::TPT>::run(
            tw_s, t, [&](int, int, int j) { return src[(size_t)j * W]; },
            [&](int, int, int j, float2 v) {
                if constexpr (MODE == COL_ADJ)
                    v = make_float2(v.x * ph.x - v.y * ph.y, v.y * ph.x + v.x * ph.y);
                dst[(size_t)j * W] = v;
            },
            lds);
    } else if constexpr (MODE == COL_FWD) {
        // A, first pass: x_z e^{-i phi_z} (sparse_phase_scale_kernel with -z), forward column FFT.
        // The workgroup's 16 columns are line blockIdx.x of every row: rows whose line is
        // empty are not read, and a strip empty in every row transforms to zeros, which are
        // not written (the strip map tells the row pass).
        const int z = blockIdx.y;
        const int line = blockIdx.x;
        const uint64_t* occ = a.occ ? a.occ + (size_t)z * L : nullptr;
        // the strip's row bits in LDS (from the map, or all set without one)
        __shared__ uint32_t rows_s[(L + 31) / 32];
        for (int k = threadIdx.x; k < (L + 31) / 32; k += C::THREADS) rows_s[k] = occ ? 0u : ~0u;
        __syncthreads();
        bool mine = false;
        if (occ) {
            for (int j = threadIdx.x; j < L; j += C::THREADS)
                if ((occ[j] >> line) & 1ull) {
                    atomicOr(&rows_s[j >> 5], 1u << (j & 31));
                    mine = true;
                }
        }
        const bool any = occ ? __syncthreads_or(mine) : true;
        if (threadIdx.x == 0) a.smap[(size_t)z * (W / kColStrip) + line] = any ? 1 : 0;
        if (!any) return;  // uniform: the whole workgroup leaves
        const float2 ph = a.phase[z];
        const float px = ph.x, py = -ph.y;
        const float2* src = a.in + (size_t)z * L * W + col;
        float2* dst = a.out + (size_t)z * L * W + col;
        FftRun<L, SIGN, typename Plan::radices, Plan::TPT>::run(
            tw_s, t,
            [&](int, int, int j) {
                if (!((rows_s[j >> 5] >> (j & 31)) & 1u)) return make_float2(0.0f, 0.0f);
                float2 d = src[(size_t)j * W];
                return make_float2(d.x * px - d.y * py, d.y * px + d.x * py);
            },
            [&](int, int, int j, float2 v) { dst[(size_t)j * W] = v; }, lds);
    } else {  // COL_REAL: est = Re(IFFT(...)), r = est - meas, partial sum |r|^2
        const float2* src = a.in + col;
        double s = 0.0;
        FftRun<L, SIGN, typename Plan::radices, Plan::TPT>::run(
            tw_s, t, [&](int, int, int j) { return src[(size_t)j * W]; },
            [&](int, int, int j, float2 v) {
                const size_t o = (size_t)j * W + col;
                float2 e = make_float2(v.x, 0.0f);  // back_propagate_finish_kernel
                a.out[o] = e;
                float2 m = a.meas[o];
                float2 r = make_float2(e.x - m.x, e.y - m.y);  // RM_EST_TRUE
                a.resid[o] = r;
                s += (double)r.x * r.x + (double)r.y * r.y;
            },
            lds);
        double tot = block_sum(s, red_s);
        if (threadIdx.x == 0) a.partial[blockIdx.x] = tot;
    }
}

// ----------------------------------------------------------------------------
// Host launchers
// ----------------------------------------------------------------------------
#define RIHVR_FOR_EACH_L(X) \
    X(16) X(32) X(48) X(64) X(80) X(96) X(128) X(192) X(256) X(320) X(384) X(448) X(512) X(576) \
    X(640) X(768) X(832) X(896) X(1024)

bool fft_length_supported(int L) {
    switch (L) {
#define X(n) case n: return true;
        RIHVR_FOR_EACH_L(X)
#undef X
        default: return false;
    }
}

template <int L, int SIGN, int MODE>
static hipError_t launch_rows_t(const RowArgs& a, int nrows_total_or_plane, int gy, hipStream_t s) {
    dim3 grid((nrows_total_or_plane + RowCfg<L>::RPW - 1) / RowCfg<L>::RPW, gy);
    hipLaunchKernelGGL((k_rows<L, SIGN, MODE>), grid, dim3(RowCfg<L>::THREADS), 0, s, a);
    return hipGetLastError();
}

template <int L, int SIGN, int MODE>
static hipError_t launch_cols_t(const ColArgs& a, int gy, hipStream_t s) {
    dim3 grid(a.W / kColStrip, gy);
    hipLaunchKernelGGL((k_cols<L, SIGN, MODE>), grid, dim3(ColCfg<L, MODE>::THREADS), 0, s, a);
    return hipGetLastError();
}

hipError_t fft_rows_plain(int L, int sign, const float2* in, float2* out, int nrows,
                          const float2* tw, hipStream_t s) {
    RowArgs a{};
    a.in = in; a.out = out; a.nrows = nrows; a.twg = tw;
    switch (L) {
#define X(n) case n: return sign < 0 ? launch_rows_t<n, -1, ROW_PLAIN>(a, nrows, 1, s) \
                                     : launch_rows_t<n, 1, ROW_PLAIN>(a, nrows, 1, s);
        RIHVR_FOR_EACH_L(X)
#undef X
        default: return hipErrorInvalidValue;
    }
}

hipError_t fft_rows_sumparts(int L, const float2* parts, int nparts, long part_stride, float2* out,
                             int nrows, const float2* tw_inv, hipStream_t s) {
    RowArgs a{};
    a.in = parts; a.out = out; a.nrows = nrows; a.twg = tw_inv; a.nparts = nparts;
    a.part_stride = part_stride;
    switch (L) {
#define X(n) case n: return launch_rows_t<n, 1, ROW_SUMPART>(a, nrows, 1, s);
        RIHVR_FOR_EACH_L(X)
#undef X
        default: return hipErrorInvalidValue;
    }
}

hipError_t fft_rows_fwdacc(int L, const float2* in, const double* kappa, const float* zlist, float2* parts,
                          int H, int nz, int zchunk, const float2* tw_fwd, const uint8_t* smap, hipStream_t s) {
    RowArgs a{};
    a.in = in; a.out = parts; a.nrows = H; a.twg = tw_fwd; a.kappa = kappa; a.zlist = zlist; a.nz = nz;
    a.zchunk = zchunk; a.part_stride = (long)H * L; a.smap = smap;
    const int gy = (nz + zchunk - 1) / zchunk;
    switch (L) {
#define X(n) case n: return launch_rows_t<n, -1, ROW_FWDACC>(a, H, gy, s);
        RIHVR_FOR_EACH_L(X)
#undef X
        default: return hipErrorInvalidValue;
    }
}

hipError_t fft_rows_adj(int L, const float2* rhat, const double* kappa, const float* zlist, float2* out,
                        int H, int nz, int zchunk, float invN, const float2* tw_inv, hipStream_t s) {
    RowArgs a{};
    a.out = out; a.nrows = H; a.twg = tw_inv; a.rhat = rhat; a.kappa = kappa; a.zlist = zlist;
    a.nz = nz; a.zchunk = zchunk; a.invN = invN;
    const int gy = (nz + zchunk - 1) / zchunk;
    switch (L) {
#define X(n) case n: return launch_rows_t<n, 1, ROW_ADJ>(a, H, gy, s);
        RIHVR_FOR_EACH_L(X)
#undef X
        default: return hipErrorInvalidValue;
    }
}

hipError_t fft_cols_plain(int L, int sign, const float2* in, float2* out, int W, int nplanes,
                          const float2* tw, hipStream_t s) {
    ColArgs a{};
    a.in = in; a.out = out; a.W = W; a.twg = tw;
    switch (L) {
#define X(n) case n: return sign < 0 ? launch_cols_t<n, -1, COL_PLAIN>(a, nplanes, s) \
                                     : launch_cols_t<n, 1, COL_PLAIN>(a, nplanes, s);
        RIHVR_FOR_EACH_L(X)
#undef X
        default: return hipErrorInvalidValue;
    }
}

hipError_t fft_cols_adj(int L, const float2* in, const float2* phase, float2* out, int W, int nz,
                        const float2* tw_inv, hipStream_t s) {
    ColArgs a{};
    a.in = in; a.out = out; a.W = W; a.twg = tw_inv; a.phase = phase;
    switch (L) {
#define X(n) case n: return launch_cols_t<n, 1, COL_ADJ>(a, nz, s);
        RIHVR_FOR_EACH_L(X)
#undef X
        default: return hipErrorInvalidValue;
    }
}

hipError_t fft_cols_fwd(int L, const float2* in, const float2* phase, float2* out, int W, int nz,
                        const float2* tw_fwd, const uint64_t* occ, const uint64_t* occ_planes, uint8_t* smap,
                        hipStream_t s) {
    ColArgs a{};
    a.in = in; a.out = out; a.W = W; a.twg = tw_fwd; a.phase = phase; a.occ = occ; a.occ_planes = occ_planes;
    a.smap = smap;
    switch (L) {
#define X(n) case n: return launch_cols_t<n, -1, COL_FWD>(a, nz, s);
        RIHVR_FOR_EACH_L(X)
#undef X
        default: return hipErrorInvalidValue;
    }
}

hipError_t fft_cols_real(int L, const float2* in, float2* est, const float2* meas, float2* resid,
                         double* partial, int W, const float2* tw_inv, hipStream_t s) {
    ColArgs a{};
    a.in = in; a.out = est; a.W = W; a.twg = tw_inv; a.meas = meas; a.resid = resid;
    a.partial = partial;
    switch (L) {
#define X(n) case n: return launch_cols_t<n, 1, COL_REAL>(a, 1, s);
        RIHVR_FOR_EACH_L(X)
#undef X
        default: return hipErrorInvalidValue;
    }
}

}  // namespace rihvr
