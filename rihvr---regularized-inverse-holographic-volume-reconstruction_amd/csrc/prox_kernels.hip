// prox_kernels.hip -- fused proximal step of the FASTA iteration and the
// elementwise/reduction kernels around it.
//
// k_prox_fl2d restates proximalFusedLasso2d (sparse_compressive_holo.cu:3652-3732):
//   x0   = y_{k-1} + s (y_{k-1} - y_{k-2})       (the FISTA momentum point, formed on the fly;
//                                                 sparse_fistaUpdate_kernel :4484-4493)
//   xhat = x0 - t g                               (sparse_get_xhat_kernel :3478-3489)
//   g_x = g_y = 0, then n_tv times:
//     yhat = xhat - mu div(g)                     (tv_denoised_estimate_2d_kernel :3491-3513)
//     g    = (g + sc grad(yhat)) / max(|.|, 1)    (tv_gradient_projection_2d_kernel :3539-3597)
//   yhat = xhat - mu div(g);  y = soft(yhat, t lambda1) per component (:3065-3110)
// with mu = t lambda_TV, sc = 1 / (12 t lambda_TV), periodic boundaries, and the
// line-search, restart and objective reductions of the same iteration fused in.
//
// Work decomposition: one wave streams down a strip of 64 columns of one plane
// (lane = column).  TV level k (k = 1..n_tv) runs one row behind level k-1, so
// every stencil neighbour in y is a register of the same lane (the row above)
// and every neighbour in x is a lane shuffle.  Rows and columns outside the
// strip's output range are recomputed halo (n_tv on each side in x, and at the
// ends of the row chunk in y).  No workgroup barriers and no LDS: the per-lane
// history of x-hat lives in registers.
#include <algorithm>
#include <utility>

#include "device_common.hpp"
#include "kernels.hpp"

namespace rihvr {

constexpr int kDefaultRowChunk = 128;  // output rows per wave
constexpr int kProxWaves = 4;    // waves per workgroup
constexpr int kMaxNtv = 10;

struct ProxGeom {
    int ow, nstrips, rc, nchunks;
};
inline ProxGeom prox_geom(int W, int H, int n, int row_chunk) {
    if (row_chunk <= 0) row_chunk = kDefaultRowChunk;
    ProxGeom g;
    g.ow = 64 - 2 * n;
    if (g.ow > W) g.ow = W;
    g.nstrips = (W + g.ow - 1) / g.ow;
    g.rc = H < row_chunk ? H : row_chunk;
    g.nchunks = (H + g.rc - 1) / g.rc;
    return g;
}

int prox_fl2d_num_waves(int W, int H, int nz, int n_tv, int row_chunk) {
    ProxGeom g = prox_geom(W, H, n_tv, row_chunk);
    return g.nstrips * g.nchunks * nz;
}
bool prox_fl2d_supported(int n_tv) { return n_tv >= 0 && n_tv <= kMaxNtv; }

// Bijection of [0, nblocks): block b, which the hardware places on XCD b % 8, gets the
// logical index of the (b / 8)-th block of a contiguous run assigned to that XCD.
constexpr int kXcds = 8;
__device__ __forceinline__ int xcd_local_block(int b, int nblocks) {
    const int q = nblocks / kXcds, r = nblocks % kXcds;
    const int x = b % kXcds, k = b / kXcds;
    return (x < r) ? x * (q + 1) + k : r * (q + 1) + (x - r) * q + k;
}

struct ProxDev {
    const float2* Y0;
    const float2* YM;
    const float2* G;
    float2* Y;
    int W, H, nz;
    float s_mom, t, mu, sc, thr;
    const float* s_mom_dev;  // when set, the momentum step is read from it (replays of captured graphs)
    int do_soft;
    int lsh;  // a map line is 2^lsh voxels of a row: 16 up to W = 1024, then 32, 64 (64 lines per row)
    ProxGeom geo;
    int nwaves;
    double* partial;
    // line occupancy ([nz][H] words, all required) of Y0, YM, Y's current content and new Y
    const uint64_t* occ_y0;
    const uint64_t* occ_ym;
    const uint64_t* occ_old;
    uint64_t* occ_new;
    uint64_t* occ_new_planes;  // [nz] plane words of the new map
};

__device__ __forceinline__ float soft1(float v, float thr) {
    // block_soft_threshold_kernel: y = max(|v| - T, 0); v * y / (y + T)
    float y = fmaxf(fabsf(v) - thr, 0.0f);
    return v * (y / (y + thr));
}

__device__ __forceinline__ int wrapi(int v, int n) {
    v %= n;
    return v < 0 ? v + n : v;
}

// Lane shifts across the whole wave with DPP (no LDS traffic): lane i receives
// lane i-1 (wave_shr:1) or lane i+1 (wave_shl:1); the edge lane receives 0 (bound
// control, so the shift folds into the consuming instruction) and is halo.
__device__ __forceinline__ float lane_from_left(float v) {
    return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), 0x138, 0xF, 0xF, true));
}
__device__ __forceinline__ float lane_from_right(float v) {
    return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), 0x130, 0xF, 0xF, true));
}

__device__ __forceinline__ float soft_fast(float v, float thr) {
    // block_soft_threshold_kernel: v * y / (y + T), y = max(|v| - T, 0); y / (y + T)
    // through the hardware reciprocal (the reference build divides with fast-math too)
    float y = fmaxf(fabsf(v) - thr, 0.0f);
    return v * (y * __builtin_amdgcn_rcpf(y + thr));
}

// One wave's state: the TV duals g_k (gx.re, gx.im, gy.re, gy.im) and y-hat_k of
// the last two rows of every level, indexed [row parity][level]; a per-lane
// history ring holds x-hat of the last NTV + 3 rows.  The row
// loop is unrolled by U (a multiple of the ring length and of 2), so every ring
// slot and row parity is a compile-time constant and the ring is registers.
template <int NTV> struct ProxWave {
    static constexpr int P = NTV + 3;
    static constexpr int U = (P % 2 == 0) ? P : 2 * P;
    float4 gd[2][NTV + 1];
    float2 yh[2][NTV + 1];
    // rows tau and tau + 1 of g, loaded two steps ahead (y_{k-1} and y_{k-2} are read at their
    // row's step, by the few rows that hold data)
    float2 pfg[2];
    // x-hat of the last P rows, indexed by compile-time slots: registers
    float2 hist[1][P];
    // Row flags, shifted in as each row is loaded (after the load of row tau + 2, bit j is
    // row tau + 2 - j): y0rows, y_{k-1} may hold a non-zero in the wave's lines (its map says
    // so, or it has no map); xrows, y_{k-1} or y_{k-2} may, so x0 may be non-zero.
    uint32_t y0rows = 0, xrows = 0;
    // occupancy words (wave-uniform, kept in scalar registers), [row parity]: of y_{k-1} and
    // y_{k-2} for the rows loaded next, and of the output's current content for the rows
    // written next
    uint64_t ow0[2] = {~0ull, ~0ull}, owm[2] = {~0ull, ~0ull}, oold[2] = {~0ull, ~0ull};
    // the lines (bit l: columns l 2^lsh .. (l + 1) 2^lsh - 1) that the wave's 64 lanes read, and those
    // its output lanes write
    uint64_t wlines = ~0ull, outlines = ~0ull;
    uint64_t plane_bits = 0;  // lines this wave set in the new map (its plane word's share)
    float acc_ls1 = 0.f, acc_ls2 = 0.f, acc_rs = 0.f, acc_l1 = 0.f, acc_tv = 0.f, acc_nnz = 0.f;
    float smom = 0.f;  // the momentum step s of x0 = y0 + s (y0 - ym)

    static __device__ __forceinline__ int wrap_row(int tau, int H) {
        const int row = tau % H;
        return row < 0 ? row + H : row;
    }
    // A wave-uniform word of a map this kernel does not write: a scalar load through the
    // constant address space (its own counter, so it never waits on the vector loads in
    // flight for the rows ahead).
    static __device__ __forceinline__ uint64_t uniform_word(const uint64_t* map, unsigned i) {
        const __attribute__((address_space(4))) uint64_t* cmap =
            (const __attribute__((address_space(4))) uint64_t*)map;
        return cmap[i];
    }
    // Plane bases (wave-uniform) and the wrapped rows of the next loads, kept as running
    // counters so that no step divides by H.
    const char* y0p;
    const char* ymp;
    const char* gp;
    char* yp;
    unsigned rowb;                // bytes per row
    int row_ld, row_wd, row_old;  // rows of the next load_row, load_words and old-map read
    int row_cur;                  // row tau of the step
    // this plane's row words of the maps of y_{k-1}, y_{k-2} and the output's current content
    const uint64_t* m0;
    const uint64_t* mm;
    const uint64_t* mold;
    static __device__ __forceinline__ int next_row(int r, int H) { return r + 1 == H ? 0 : r + 1; }

    // occupancy words of y_{k-1} and y_{k-2} for row row_wd, one step before its loads
    __device__ __forceinline__ void load_words(int slot, int H) {
        ow0[slot] = uniform_word(m0, (unsigned)row_wd);
        owm[slot] = uniform_word(mm, (unsigned)row_wd);
        row_wd = next_row(row_wd, H);
    }
    // Element at byte offset xoff (per lane) of the row that starts off bytes into the
    // plane at base (wave-uniform): the row address stays in scalar registers.
    static __device__ __forceinline__ const float2& at(const char* base, unsigned off, unsigned xoff) {
        return *(const float2*)(base + off + xoff);
    }
    // Row row_ld: g, and the row flags of y_{k-1} and y_{k-2} (a row with one of the wave's
    // lines set in either map is read at its step; a line whose bit is clear holds +0, so a
    // wave that reads the row with some of its lines clear reads zeros there).
    __device__ __forceinline__ void load_row(const ProxDev& a, int slot, unsigned xoff, bool ld) {
        const bool f0 = ld && (ow0[slot] & wlines) != 0ull;
        const bool fm = ld && (owm[slot] & wlines) != 0ull;
        y0rows = (y0rows << 1) | (f0 ? 1u : 0u);
        xrows = (xrows << 1) | ((f0 || fm) ? 1u : 0u);
        if (!ld) return;
        pfg[slot] = at(gp, (unsigned)row_ld * rowb, xoff);
        row_ld = next_row(row_ld, a.H);
    }

    // Step S of an unrolled block: row tau, ring slot S mod P, row parity S & 1.
    template <int S>
    __device__ __forceinline__ void step(const ProxDev& a, int tau, int r0, int r1,
                                         unsigned xoff, int ocol, bool out_lane, int tau_end,
                                         int zrow, int strip) {
        constexpr int SLOT = S % P;
        constexpr int PAR = S & 1;
        constexpr int Q = PAR ^ 1;
        const int W = a.W;
        const unsigned oofs = (unsigned)ocol * (unsigned)sizeof(float2);  // output lanes only
        // ---- row tau: momentum point x0 = y0 + s (y0 - ym), x-hat = x0 - t g ----
        {
            const float2 g = pfg[PAR];
            float2 x0 = make_float2(0.0f, 0.0f);  // y0 = ym = +0 give x0 = +0
            if ((xrows >> 1) & 1u) {              // row tau (its flags were shifted in 2 rows ago)
                const unsigned off = (unsigned)row_cur * rowb;
                const float2 y0 = at(y0p, off, xoff), ym = at(ymp, off, xoff);
                x0 = make_float2(fmaf(smom, y0.x - ym.x, y0.x), fmaf(smom, y0.y - ym.y, y0.y));
            }
            row_cur = next_row(row_cur, a.H);
            hist[0][SLOT] = make_float2(fmaf(-a.t, g.x, x0.x), fmaf(-a.t, g.y, x0.y));
            // the output's current map for the row written in the next step
            oold[Q] = uniform_word(mold, (unsigned)row_old);
            row_old = next_row(row_old, a.H);
            load_row(a, PAR, xoff, tau + 2 <= tau_end);
            if (tau + 3 <= tau_end) load_words(Q, a.H);
        }
        // ---- TV levels: level k works on row tau - k ----
#pragma unroll
        for (int k = 1; k <= NTV; ++k) {
            const int sk = (SLOT - k + 2 * P) % P;
            const float2 xh = hist[0][sk];
            float2 y;
            float4 go;  // g_{k-1} at row tau - k
            if (k == 1) {
                y = xh;
                go = make_float4(0.f, 0.f, 0.f, 0.f);
            } else {
                go = gd[Q][k - 1];
                const float4 gb = gd[PAR][k - 1];  // row below
                const float dvr = lane_from_right(go.x) - go.x + gb.z - go.z;
                const float dvi = lane_from_right(go.y) - go.y + gb.w - go.w;
                y = make_float2(fmaf(-a.mu, dvr, xh.x), fmaf(-a.mu, dvi, xh.y));
            }
            const float2 yu = yh[Q][k];
            const float axr = fmaf(a.sc, lane_from_left(y.x) - y.x, go.x);
            const float axi = fmaf(a.sc, lane_from_left(y.y) - y.y, go.y);
            const float ayr = fmaf(a.sc, yu.x - y.x, go.z);
            const float ayi = fmaf(a.sc, yu.y - y.y, go.w);
            // projection: divide by max(|a|, 1), i.e. by |a| where |a|^2 > 1 and by exactly 1
            // elsewhere.  Components all within 0.5 give |a|^2 <= 1 (also after rounding), so a
            // wave where every lane's are (nearly every wave of a sparse solve) neither forms
            // |a|^2 nor projects; otherwise a wave with no lane above 1 skips the reciprocal
            // square root and the products.
            gd[PAR][k] = make_float4(axr, axi, ayr, ayi);
            const float amax = fmaxf(fmaxf(fabsf(axr), fabsf(axi)), fmaxf(fabsf(ayr), fabsf(ayi)));
            if (__builtin_amdgcn_ballot_w64(amax > 0.5f)) {
                const float n2 = fmaf(axr, axr, fmaf(axi, axi, fmaf(ayr, ayr, ayi * ayi)));
                const bool proj = n2 > 1.0f;
                if (__builtin_amdgcn_ballot_w64(proj)) {
                    const float inv = proj ? __builtin_amdgcn_rsqf(n2) : 1.0f;
                    gd[PAR][k] = make_float4(axr * inv, axi * inv, ayr * inv, ayi * inv);
                }
            }
            yh[PAR][k] = y;
        }
        // ---- final estimate, threshold and reductions: row o = tau - NTV - 1 ----
        const int o = tau - NTV - 1;
        constexpr int SO = (SLOT - NTV - 1 + 2 * P) % P;
        constexpr int SU = (SO - 1 + P) % P;
        const float2 xh = hist[0][SO];
        float2 yv;
        if constexpr (NTV == 0) {
            yv = xh;
        } else {
            const float4 gu = gd[Q][NTV];
            const float4 gb = gd[PAR][NTV];
            const float dvr = lane_from_right(gu.x) - gu.x + gb.z - gu.z;
            const float dvi = lane_from_right(gu.y) - gu.y + gb.w - gu.w;
            yv = make_float2(fmaf(-a.mu, dvr, xh.x), fmaf(-a.mu, dvi, xh.y));
        }
        // Soft threshold.  A lane with |v| <= T in both parts gives zero, so a wave where every
        // lane does (most waves of a sparse solve) skips the division; NaN stays on the full
        // path.  Zeros are stored as +0 (the threshold's v * 0 is -0 for negative v), so a line
        // of zeros is a line of +0 whether it is written or skipped.
        bool live = true;  // wave-uniform: some lane may be non-zero
        if (a.do_soft) {
            const bool l = !(fabsf(yv.x) <= a.thr) || !(fabsf(yv.y) <= a.thr);
            live = __builtin_amdgcn_ballot_w64(l) != 0ull;
            if (live)
                yv = make_float2(soft_fast(yv.x, a.thr) + 0.0f, soft_fast(yv.y, a.thr) + 0.0f);
            else
                yv = make_float2(0.0f, 0.0f);
        } else {
            yv = make_float2(yv.x + 0.0f, yv.y + 0.0f);
        }
        if (o >= r0 && o < r1) {  // row o: the new map and the skipped writes
            uint64_t m = 0ull;
            if (live) {
                m = __builtin_amdgcn_ballot_w64(out_lane && (yv.x != 0.0f || yv.y != 0.0f));
                if (m) {
                    // lanes NTV .. NTV + ow - 1 hold columns strip*ow + (lane - NTV): the lines
                    // touched by non-zero lanes, then one atomic OR per row and wave
                    const int c0 = strip * a.geo.ow - NTV;  // column of lane 0
                    uint64_t bits = 0;
                    const int l_last = min(c0 + NTV + a.geo.ow - 1, W - 1) >> a.lsh;
                    for (int l = (c0 + NTV) >> a.lsh; l <= l_last; ++l) {
                        const int lo = max((l << a.lsh) - c0, NTV);
                        const int hi = min((l << a.lsh) + (1 << a.lsh) - 1 - c0, NTV + a.geo.ow - 1);
                        if (lo > hi || lo > 63) continue;
                        const uint64_t lanes =
                            (hi >= 63 ? ~0ull : ((1ull << (hi + 1)) - 1ull)) & ~((1ull << lo) - 1ull);
                        if (m & lanes) bits |= 1ull << l;
                    }
                    if ((threadIdx.x & 63) == 0) atomicOr((unsigned long long*)&a.occ_new[zrow + o], bits);
                    plane_bits |= bits;
                }
            }
            // a lane writes when its value is non-zero or its line holds an old non-zero; a
            // wave with neither writes nothing
            if (m | (oold[PAR] & outlines)) {
                const bool keep = (yv.x != 0.0f || yv.y != 0.0f) || ((oold[PAR] >> (ocol >> a.lsh)) & 1ull);
                if (out_lane && keep) *(float2*)(yp + (unsigned)o * rowb + oofs) = yv;
            }
        }
        // Every reduction term of row o is a product with (y - x0), (x0 - y), x0 or one of
        // the x0 differences, so it is zero (adding it changes no sum, which starts at +0)
        // when y, x0, y_{k-1} and the row above's x0 are zero in every lane of the wave,
        // halo lanes included (they are the left neighbours).  Late in a solve almost every
        // row is.
        // x0 of rows o and o - 1 and y_{k-1} of row o by their row flags, y by its value
        const bool xflags = ((xrows >> (NTV + 3)) & 3u) != 0u || ((y0rows >> (NTV + 3)) & 1u) != 0u;
        bool reduce = o >= r0 && o < r1 && xflags;
        if (live && o >= r0 && o < r1 && !xflags)
            reduce = __builtin_amdgcn_ballot_w64(((__float_as_uint(yv.x) | __float_as_uint(yv.y)) << 1) != 0u) != 0ull;
        if (reduce) {
            // all lanes active here.  g, y_{k-1} and x0 of row o and x0 of the row above are
            // formed again from memory (rows that reduce are few, so they are not kept in
            // registers; x0 = y0 + s (y0 - ym) exactly as at the row's load, +0 for a row
            // without data), and the left neighbours' x0 for the TV term
            const unsigned ro = (unsigned)o * rowb;
            const unsigned ru = (unsigned)(o > 0 ? o - 1 : a.H - 1) * rowb;
            const float2 g = at(gp, ro, xoff);
            const float2 y0 = at(y0p, ro, xoff);
            float2 x0o = make_float2(0.0f, 0.0f), x0u = make_float2(0.0f, 0.0f);
            if ((xrows >> (NTV + 3)) & 1u) {
                const float2 ym = at(ymp, ro, xoff);
                x0o = make_float2(fmaf(smom, y0.x - ym.x, y0.x), fmaf(smom, y0.y - ym.y, y0.y));
            }
            if ((xrows >> (NTV + 4)) & 1u) {
                const float2 y0u = at(y0p, ru, xoff), ymu = at(ymp, ru, xoff);
                x0u = make_float2(fmaf(smom, y0u.x - ymu.x, y0u.x), fmaf(smom, y0u.y - ymu.y, y0u.y));
            }
            const float x0lr = lane_from_left(x0o.x);
            const float x0li = lane_from_left(x0o.y);
            if (out_lane) {
                const float dr = yv.x - x0o.x, di = yv.y - x0o.y;
                acc_ls1 = fmaf(dr, g.x, fmaf(di, g.y, acc_ls1));
                acc_ls2 = fmaf(dr, dr, fmaf(di, di, acc_ls2));
                const float dre = (x0o.x - yv.x) * (yv.x - y0.x);
                const float dim = (x0o.y - yv.y) * (yv.y - y0.y);
                const int off = ((o * W) + ocol) & 1023;
                acc_rs += off < 512 ? fmaf(2.0f, dim, dre) : dim;
                // objective terms of x0: the hardware square root (about 1 ulp) feeds sums
                // accumulated per wave and reported in double
                const float m2 = fmaf(x0o.x, x0o.x, x0o.y * x0o.y);
                acc_l1 += __builtin_amdgcn_sqrtf(m2);
                acc_nnz += (x0o.x != 0.0f || x0o.y != 0.0f) ? 1.0f : 0.0f;
                const float gxx = x0lr - x0o.x, gyx = x0u.x - x0o.x, gxy = x0li - x0o.y, gyy = x0u.y - x0o.y;
                acc_tv += __builtin_amdgcn_sqrtf(fmaf(gxx, gxx, fmaf(gxy, gxy, fmaf(gyx, gyx, gyy * gyy))));
            }
        }
    }

    template <int... S>
    __device__ __forceinline__ void block(std::integer_sequence<int, S...>, const ProxDev& a,
                                          int tau, int r0, int r1, unsigned xoff, int ocol,
                                          bool out_lane, int tau_end, int zrow, int strip) {
        (step<S>(a, tau + S, r0, r1, xoff, ocol, out_lane, tau_end, zrow, strip), ...);
    }
    template <int... S>
    __device__ __forceinline__ void tail(std::integer_sequence<int, S...>, int rem, const ProxDev& a,
                                         int tau, int r0, int r1,
                                         unsigned xoff, int ocol, bool out_lane, int tau_end, int zrow, int strip) {
        ((S < rem ? step<S>(a, tau + S, r0, r1, xoff, ocol, out_lane, tau_end, zrow, strip) : void()),
         ...);
    }
};

// Up to 5 TV iterations the kernel fits 5 waves per SIMD (96 VGPRs; the few registers that
// do not fit are spilled only in the end-of-chunk steps, not in the main row loop); with more
// its state needs more registers than 5 waves leave.
template <int NTV> constexpr int prox_waves_per_simd() { return NTV <= 5 ? 5 : 1; }

template <int NTV>
__global__ void __launch_bounds__(64 * kProxWaves)
    __attribute__((amdgpu_waves_per_eu(prox_waves_per_simd<NTV>()))) k_prox_fl2d(ProxDev a) {
    using PW = ProxWave<NTV>;
    constexpr int P = PW::P;
    constexpr int U = PW::U;
    const int lane = threadIdx.x & 63;
    // wave index, made wave-uniform for the compiler so all row/plane
    // arithmetic below stays in scalar registers
    const int wl = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    // Workgroups are dispatched round-robin over the 8 XCDs (workgroup b to XCD b mod 8),
    // each with its own L2.  Renumber them so that consecutive strips of a row chunk --
    // which re-read each other's halo columns -- run on the same XCD.
    const int w = xcd_local_block(blockIdx.x, gridDim.x) * kProxWaves + wl;
    if (w >= a.nwaves) return;  // whole wave exits; no barriers below
    const ProxGeom geo = a.geo;
    const int per_plane = geo.nstrips * geo.nchunks;
    const int z = w / per_plane;
    const int rem = w - z * per_plane;
    const int chunk = rem / geo.nstrips;
    const int strip = rem - chunk * geo.nstrips;
    const int W = a.W;
    const int ocol = strip * geo.ow + (lane - NTV);  // output column of this lane
    const bool out_lane = (lane - NTV) >= 0 && (lane - NTV) < geo.ow && ocol < W;
    const int xcol = wrapi(strip * geo.ow - NTV + lane, W);
    const unsigned xoff = (unsigned)xcol * (unsigned)sizeof(float2);  // byte offset in a row
    const int r0 = chunk * geo.rc;
    const int r1 = min(a.H, r0 + geo.rc);
    const size_t plane_off = (size_t)z * a.H * W;
    const int zrow = z * a.H;  // first row of plane z in the occupancy maps

    PW st;
    st.smom = a.s_mom_dev ? *a.s_mom_dev : a.s_mom;
    // the wave's lines: the periodic wrap can add a line at the other end of the row
    st.wlines = 0;
    st.outlines = 0;
    for (int l = 0; l <= ((W - 1) >> a.lsh); ++l) {
        if (__builtin_amdgcn_ballot_w64((xcol >> a.lsh) == l)) st.wlines |= 1ull << l;
        if (__builtin_amdgcn_ballot_w64(out_lane && (ocol >> a.lsh) == l)) st.outlines |= 1ull << l;
    }
#pragma unroll
    for (int p = 0; p < 2; ++p)
#pragma unroll
        for (int k = 0; k <= NTV; ++k) {
            st.gd[p][k] = make_float4(0.f, 0.f, 0.f, 0.f);
            st.yh[p][k] = make_float2(0.f, 0.f);
        }
    const int tau0 = r0 - (NTV > 0 ? NTV : 1);
    const int tau1 = r1 + NTV;  // inclusive
    // the first two rows; every step then loads the row two ahead (U is even)
    st.y0p = (const char*)(a.Y0 + plane_off);
    st.ymp = (const char*)(a.YM + plane_off);
    st.gp = (const char*)(a.G + plane_off);
    st.yp = (char*)(a.Y + plane_off);
    st.rowb = (unsigned)W * (unsigned)sizeof(float2);
    st.row_ld = PW::wrap_row(tau0, a.H);
    st.row_wd = st.row_ld;
    st.row_cur = st.row_ld;
    st.m0 = a.occ_y0 + zrow;
    st.mm = a.occ_ym + zrow;
    st.mold = a.occ_old + zrow;
    st.row_old = PW::wrap_row(tau0 - NTV, a.H);
    st.load_words(0, a.H);  // tau0
    st.load_words(1, a.H);  // tau0 + 1
    st.load_row(a, 0, xoff, true);  // tau0
    st.load_row(a, 1, xoff, true);  // tau0 + 1
    st.load_words(0, a.H);  // tau0 + 2
    st.oold[0] = PW::uniform_word(st.mold, (unsigned)PW::wrap_row(tau0 - NTV - 1, a.H));
    int tau = tau0;
    for (; tau + U - 1 <= tau1; tau += U)
        st.block(std::make_integer_sequence<int, U>{}, a, tau, r0, r1, xoff, ocol, out_lane,
                 tau1, zrow, strip);
    if (tau <= tau1)
        st.tail(std::make_integer_sequence<int, U - 1>{}, tau1 - tau + 1, a, tau, r0, r1, xoff,
                ocol, out_lane, tau1, zrow, strip);

    if (st.plane_bits && lane == 0) atomicOr((unsigned long long*)&a.occ_new_planes[z], st.plane_bits);
    const double s_ls1 = wave_sum((double)st.acc_ls1);
    const double s_ls2 = wave_sum((double)st.acc_ls2);
    const double s_rs = wave_sum((double)st.acc_rs);
    const double s_l1 = wave_sum((double)st.acc_l1);
    const double s_tv = wave_sum((double)st.acc_tv);
    const double s_nnz = wave_sum((double)st.acc_nnz);
    if (lane == 0) {
        double* p = a.partial + (size_t)w * PS_COUNT;
        p[PS_LS1] = s_ls1;
        p[PS_LS2] = s_ls2;
        p[PS_RESTART] = s_rs;
        p[PS_L1X0] = s_l1;
        p[PS_TVX0] = s_tv;
        p[PS_NNZX0] = s_nnz;
        p[PS_NNZY] = 0.0;
        p[7] = 0.0;
    }
}

static ProxDev make_dev(const ProxLaunch& a) {
    ProxDev d;
    d.Y0 = a.Y0; d.YM = a.YM; d.G = a.G; d.Y = a.Y;
    d.W = a.W; d.H = a.H; d.nz = a.nz;
    d.s_mom = a.s_mom;
    d.s_mom_dev = a.s_mom_dev;
    d.t = (float)a.t;
    d.mu = (float)(a.t * a.reg_tv);
    d.sc = a.reg_tv != 0.0 ? (float)(1 / (4 * 3 * a.t * a.reg_tv)) : 0.0f;
    d.thr = (float)(a.t * a.reg);
    d.do_soft = a.reg > 0.0;
    d.lsh = 4;
    while (((a.W - 1) >> d.lsh) >= 64) ++d.lsh;
    d.geo = prox_geom(a.W, a.H, a.n_tv, a.row_chunk);
    d.nwaves = d.geo.nstrips * d.geo.nchunks * a.nz;
    d.partial = a.partial;
    d.occ_y0 = a.occ_y0;
    d.occ_ym = a.occ_ym;
    d.occ_old = a.occ_old;
    d.occ_new = a.occ_new;
    d.occ_new_planes = a.occ_new_planes;
    return d;
}

template <int NTV> static hipError_t launch_fl2d(const ProxDev& d, hipStream_t s) {
    dim3 grid((d.nwaves + kProxWaves - 1) / kProxWaves);
    hipLaunchKernelGGL(k_prox_fl2d<NTV>, grid, dim3(64 * kProxWaves), 0, s, d);
    return hipGetLastError();
}

template <int NTV> static int fl2d_blocks_per_cu() {
    int nb = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, k_prox_fl2d<NTV>, 64 * kProxWaves, 0) != hipSuccess)
        return 0;
    return nb;
}

int prox_fl2d_auto_rows(int W, int H, int nz, int n_tv) {
    int nb = 0;
    switch (n_tv) {
        case 0: nb = fl2d_blocks_per_cu<0>(); break;
        case 1: nb = fl2d_blocks_per_cu<1>(); break;
        case 2: nb = fl2d_blocks_per_cu<2>(); break;
        case 3: nb = fl2d_blocks_per_cu<3>(); break;
        case 4: nb = fl2d_blocks_per_cu<4>(); break;
        case 5: nb = fl2d_blocks_per_cu<5>(); break;
        case 6: nb = fl2d_blocks_per_cu<6>(); break;
        case 7: nb = fl2d_blocks_per_cu<7>(); break;
        case 8: nb = fl2d_blocks_per_cu<8>(); break;
        case 9: nb = fl2d_blocks_per_cu<9>(); break;
        case 10: nb = fl2d_blocks_per_cu<10>(); break;
        default: return 0;
    }
    int dev = 0, cus = 0;
    if (nb <= 0 || hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || cus <= 0)
        return 0;
    const long slots = (long)nb * kProxWaves * cus;  // resident waves
    const int nstrips = prox_geom(W, H, n_tv, H).nstrips;
    // Chunks of about 205 rows when they give at least two rounds of resident waves: waves
    // with data run longer, and the dispatcher balances them only over several rounds (on
    // MI355X at 1024^2 x 128, 205 rows was best or tied-best with 3, 4 and 5 waves per SIMD,
    // and one round of 512-row chunks 7 % slower).  Smaller problems take the chunk that
    // fills the device in the fewest rounds of the shortest waves.
    const int nch_target = std::max(1, (H + 102) / 205);
    if ((long)nstrips * nch_target * nz >= 2 * slots) return (H + nch_target - 1) / nch_target;
    long best_cost = -1;
    int best_rc = 0;
    for (int nch = 1; nch <= H; ++nch) {
        const int rc = (H + nch - 1) / nch;
        if (rc < 32 && nch > 1) break;
        const int chunks = (H + rc - 1) / rc;
        const long waves = (long)nstrips * chunks * nz;
        const long rounds = (waves + slots - 1) / slots;
        const long cost = rounds * (rc + 2 * n_tv + 1);
        if (best_cost < 0 || cost < best_cost) {  // ties keep the longer chunk (less halo)
            best_cost = cost;
            best_rc = rc;
        }
    }
    return best_rc;
}

hipError_t prox_fl2d(const ProxLaunch& a, hipStream_t s) {
    // row offsets inside a plane are 32-bit byte offsets; every map is required
    if ((size_t)a.W * a.H * sizeof(float2) > 0xffffffffull) return hipErrorInvalidValue;
    if (!a.occ_y0 || !a.occ_ym || !a.occ_old || !a.occ_new || !a.occ_new_planes) return hipErrorInvalidValue;
    ProxDev d = make_dev(a);
    switch (a.n_tv) {
        case 0: return launch_fl2d<0>(d, s);
        case 1: return launch_fl2d<1>(d, s);
        case 2: return launch_fl2d<2>(d, s);
        case 3: return launch_fl2d<3>(d, s);
        case 4: return launch_fl2d<4>(d, s);
        case 5: return launch_fl2d<5>(d, s);
        case 6: return launch_fl2d<6>(d, s);
        case 7: return launch_fl2d<7>(d, s);
        case 8: return launch_fl2d<8>(d, s);
        case 9: return launch_fl2d<9>(d, s);
        case 10: return launch_fl2d<10>(d, s);
        default: return hipErrorInvalidValue;
    }
}

// ----------------------------------------------------------------------------
// Mode 5: proximalL1 (:3274-3392), block_soft_threshold_FASTA_kernel (:3112-3157),
// elementwise, with the line-search / restart / L1 reductions.
// ----------------------------------------------------------------------------
constexpr int kEwThreads = 256;
constexpr int kEwItems = 8;  // elements per thread

int prox_l1_num_blocks(int W, int H, int nz) {
    long n = (long)W * H * nz;
    return (int)((n + (long)kEwThreads * kEwItems - 1) / ((long)kEwThreads * kEwItems));
}

__global__ void __launch_bounds__(kEwThreads) k_prox_l1(ProxDev a) {
    __shared__ double red[kEwThreads / 64];
    const long N = (long)a.W * a.H;
    const long n = N * a.nz;
    double acc[PS_COUNT] = {0, 0, 0, 0, 0, 0, 0, 0};
    const float s = a.s_mom_dev ? *a.s_mom_dev : a.s_mom;
    for (int it = 0; it < kEwItems; ++it) {
        long i = ((long)blockIdx.x * kEwItems + it) * kEwThreads + threadIdx.x;
        if (i >= n) break;
        const float2 y0 = a.Y0[i], ym = a.YM[i], g = a.G[i];
        float2 x0 = make_float2(fmaf(s, y0.x - ym.x, y0.x), fmaf(s, y0.y - ym.y, y0.y));
        float2 xh = make_float2(fmaf(-a.t, g.x, x0.x), fmaf(-a.t, g.y, x0.y));
        float2 yv = make_float2(soft1(xh.x, a.thr), soft1(xh.y, a.thr));
        a.Y[i] = yv;
        const float dr = yv.x - x0.x, di = yv.y - x0.y;
        acc[PS_LS1] += (double)(dr * g.x) + (double)(di * g.y);
        acc[PS_LS2] += (double)dr * dr + (double)di * di;
        const float dre = (x0.x - yv.x) * (yv.x - y0.x);
        const float dim = (x0.y - yv.y) * (yv.y - y0.y);
        const int off = (int)((i % N) & 1023);
        acc[PS_RESTART] += off < 512 ? (double)dre + 2.0 * (double)dim : (double)dim;
        if (x0.x != 0.0f || x0.y != 0.0f) {
            acc[PS_L1X0] += (double)sqrtf(x0.x * x0.x + x0.y * x0.y);
            acc[PS_NNZX0] += 1.0;
        }
        acc[PS_NNZY] += (yv.x != 0.0f || yv.y != 0.0f) ? 1.0 : 0.0;
    }
    double* p = a.partial + (size_t)blockIdx.x * PS_COUNT;
#pragma unroll
    for (int q = 0; q < PS_COUNT; ++q) {
        double v = block_sum(acc[q], red);
        if (threadIdx.x == 0) p[q] = v;
    }
}

hipError_t prox_l1(const ProxLaunch& a, hipStream_t s) {
    ProxDev d = make_dev(a);
    d.thr = (float)(a.t * a.reg);
    dim3 grid(prox_l1_num_blocks(a.W, a.H, a.nz));
    hipLaunchKernelGGL(k_prox_l1, grid, dim3(kEwThreads), 0, s, d);
    return hipGetLastError();
}

// ----------------------------------------------------------------------------
// Mode 6: proximalFusedLasso (:3422-3475).  Soft threshold at half the L1 weight
// (block_soft_threshold_FASTA_kernel, :3438-3442), the 3D TV stage
// (totalVariationProjection :3867-3888) and a second soft threshold at half the
// weight (:3458-3463).  The TV stage's dual update has its normaliser forced to
// 1 (tv_gradient_projection_kernel :3991), so it is linear; with
// mu * scale = t lambda_TV / (12 t lambda_TV) = 1/12 and zero initial duals it
// telescopes to y_{i+1} = y_i - M y_i / 12, M = div(grad) =
// 6 y - (sum of the six periodic neighbours), i.e. y_{i+1} = y_i / 2 + nb / 12.
// Applying that stencil n_tv times gives the reference's n_tv dual iterations in
// exact arithmetic with one volume of state instead of three dual volumes.
// ----------------------------------------------------------------------------
__global__ void __launch_bounds__(kEwThreads) k_fl3d_pre(ProxDev a) {
    __shared__ double red[kEwThreads / 64];
    const int W = a.W, H = a.H;
    const long N = (long)W * H;
    const long n = N * a.nz;
    double l1 = 0, tv = 0, nnz = 0;
    const float s = a.s_mom_dev ? *a.s_mom_dev : a.s_mom;
    auto mom = [&](long i) {
        const float2 y0 = a.Y0[i], ym = a.YM[i];
        return make_float2(fmaf(s, y0.x - ym.x, y0.x), fmaf(s, y0.y - ym.y, y0.y));
    };
    for (int it = 0; it < kEwItems; ++it) {
        long i = ((long)blockIdx.x * kEwItems + it) * kEwThreads + threadIdx.x;
        if (i >= n) break;
        const float2 x0 = mom(i), g = a.G[i];
        const float2 xh = make_float2(fmaf(-a.t, g.x, x0.x), fmaf(-a.t, g.y, x0.y));
        a.Y[i] = make_float2(soft1(xh.x, a.thr), soft1(xh.y, a.thr));
        // objective terms of the momentum point x0: L1, nnz and the z-periodic 3D TV
        // (calcL1NormComplex, calcTVNorm / tv_norm_kernel sparse_volume.cu:236-265)
        if (x0.x != 0.0f || x0.y != 0.0f) {
            l1 += (double)sqrtf(x0.x * x0.x + x0.y * x0.y);
            nnz += 1.0;
        }
        const long z = i / N;
        const long zo = z * N;
        const int p = (int)(i - zo);
        const int y = p / W, x = p - y * W;
        const int xm = x > 0 ? x - 1 : W - 1, ym = y > 0 ? y - 1 : H - 1;
        const long zpo = (z > 0 ? z - 1 : a.nz - 1) * N;
        const float2 qa = mom(zo + (long)y * W + xm), qb = mom(zo + (long)ym * W + x), qc = mom(zpo + p);
        const float gxx = qa.x - x0.x, gyx = qb.x - x0.x, gzx = qc.x - x0.x;
        const float gxy = qa.y - x0.y, gyy = qb.y - x0.y, gzy = qc.y - x0.y;
        tv += (double)sqrtf(gxx * gxx + gxy * gxy + gyx * gyx + gyy * gyy + gzx * gzx + gzy * gzy);
    }
    double* pp = a.partial + (size_t)blockIdx.x * PS_COUNT;
    double r;
    r = block_sum(l1, red);
    if (threadIdx.x == 0) pp[PS_L1X0] = r;
    r = block_sum(tv, red);
    if (threadIdx.x == 0) pp[PS_TVX0] = r;
    r = block_sum(nnz, red);
    if (threadIdx.x == 0) pp[PS_NNZX0] = r;
}

// One application of y <- y / 2 + (sum of the six periodic neighbours) / 12.
// Thread = one voxel; neighbours come through the cache.
__global__ void __launch_bounds__(256) k_fl3d_sweep(const float2* __restrict__ in, float2* __restrict__ out, int W,
                                                    int H, int nz) {
    const int x = blockIdx.x * 64 + (threadIdx.x & 63);
    const int y = blockIdx.y * 4 + (threadIdx.x >> 6);
    const int z = blockIdx.z;
    if (x >= W || y >= H) return;
    const long N = (long)W * H;
    const long zo = (long)z * N;
    const long zpo = (long)(z > 0 ? z - 1 : nz - 1) * N;
    const long zno = (long)(z < nz - 1 ? z + 1 : 0) * N;
    const int xm = x > 0 ? x - 1 : W - 1, xp = x < W - 1 ? x + 1 : 0;
    const int ym = y > 0 ? y - 1 : H - 1, yp = y < H - 1 ? y + 1 : 0;
    const long row = (long)y * W;
    const float2 c = in[zo + row + x];
    const float2 a0 = in[zo + row + xm], a1 = in[zo + row + xp];
    const float2 b0 = in[zo + (long)ym * W + x], b1 = in[zo + (long)yp * W + x];
    const float2 c0 = in[zpo + row + x], c1 = in[zno + row + x];
    const float sx = ((a0.x + a1.x) + (b0.x + b1.x)) + (c0.x + c1.x);
    const float sy = ((a0.y + a1.y) + (b0.y + b1.y)) + (c0.y + c1.y);
    const float k12 = 1.0f / 12.0f;
    out[zo + row + x] = make_float2(fmaf(sx, k12, 0.5f * c.x), fmaf(sy, k12, 0.5f * c.y));
}

// Second half threshold, with the line-search and restart reductions of the
// iteration (x0 formed again from y_{k-1}, y_{k-2}).
__global__ void __launch_bounds__(kEwThreads) k_fl3d_final(ProxDev a, const float2* ytv) {  // ytv may be a.Y
    __shared__ double red[kEwThreads / 64];
    const long N = (long)a.W * a.H;
    const long n = N * a.nz;
    double ls1 = 0, ls2 = 0, rs = 0, nnzy = 0;
    const float s = a.s_mom_dev ? *a.s_mom_dev : a.s_mom;
    for (int it = 0; it < kEwItems; ++it) {
        long i = ((long)blockIdx.x * kEwItems + it) * kEwThreads + threadIdx.x;
        if (i >= n) break;
        const float2 y0 = a.Y0[i], ym = a.YM[i], g = a.G[i], v = ytv[i];
        const float2 x0 = make_float2(fmaf(s, y0.x - ym.x, y0.x), fmaf(s, y0.y - ym.y, y0.y));
        const float2 yv = make_float2(soft1(v.x, a.thr), soft1(v.y, a.thr));
        a.Y[i] = yv;
        const float dr = yv.x - x0.x, di = yv.y - x0.y;
        ls1 += (double)(dr * g.x) + (double)(di * g.y);
        ls2 += (double)dr * dr + (double)di * di;
        const float dre = (x0.x - yv.x) * (yv.x - y0.x);
        const float dim = (x0.y - yv.y) * (yv.y - y0.y);
        const int off = (int)((i % N) & 1023);
        rs += off < 512 ? (double)dre + 2.0 * (double)dim : (double)dim;
        nnzy += (yv.x != 0.0f || yv.y != 0.0f) ? 1.0 : 0.0;
    }
    double* pp = a.partial + (size_t)blockIdx.x * PS_COUNT;
    double r;
    r = block_sum(ls1, red);
    if (threadIdx.x == 0) pp[PS_LS1] = r;
    r = block_sum(ls2, red);
    if (threadIdx.x == 0) pp[PS_LS2] = r;
    r = block_sum(rs, red);
    if (threadIdx.x == 0) pp[PS_RESTART] = r;
    r = block_sum(nnzy, red);
    if (threadIdx.x == 0) {
        pp[PS_NNZY] = r;
        pp[7] = 0.0;
    }
}

int prox_fl3d_num_blocks(int W, int H, int nz) { return prox_l1_num_blocks(W, H, nz); }

hipError_t prox_fl3d(const ProxLaunch& a, float2* scratch, hipStream_t s) {
    if (!scratch || scratch == a.Y) return hipErrorInvalidValue;
    ProxDev d = make_dev(a);
    d.thr = (float)(a.t * (a.reg / 2.0));  // both thresholds at regularization / 2 (:3441, :3461)
    const dim3 eg(prox_fl3d_num_blocks(a.W, a.H, a.nz));
    hipLaunchKernelGGL(k_fl3d_pre, eg, dim3(kEwThreads), 0, s, d);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    // ping-pong between the output volume and the scratch volume
    float2* buf[2] = {a.Y, scratch};
    int src = 0;
    const dim3 sg((a.W + 63) / 64, (a.H + 3) / 4, a.nz);
    for (int k = 0; k < a.n_tv; ++k) {
        hipLaunchKernelGGL(k_fl3d_sweep, sg, dim3(256), 0, s, buf[src], buf[src ^ 1], a.W, a.H, a.nz);
        if ((e = hipGetLastError()) != hipSuccess) return e;
        src ^= 1;
    }
    hipLaunchKernelGGL(k_fl3d_final, eg, dim3(kEwThreads), 0, s, d, buf[src]);
    return hipGetLastError();
}

// ----------------------------------------------------------------------------
// Mode 7 for any number of TV iterations (the fused k_prox_fl2d holds its per-level state in
// registers and is built for up to kMaxNtv): proximalFusedLasso2d (:3652-3732) as separate
// passes over the volume with the duals in memory, each pass the reference kernel's arithmetic
// (float storage, double where it promotes):
//   k_tv2_pre:      x0, x-hat = x0 - t g (sparse_get_xhat_kernel :3478-3489), duals cleared,
//                   L1 / 2D TV / nnz of x0
//   k_tv2_estimate: yhat = xhat - mu div(g) (tv_denoised_estimate_2d_kernel :3491-3513)
//   k_tv2_project:  g = (g + sc grad yhat) / max(|.|, 1) (tv_gradient_projection_2d_kernel
//                   :3539-3597), in place (each voxel reads only its own duals)
//   k_tv2_final:    y = soft(yhat) and the line-search / restart sums (:3065-3110, :4307-4482)
// ----------------------------------------------------------------------------
struct Tv2Buf {
    float2* xh;
    float2* yh;
    float2* gx;
    float2* gy;
};

__global__ void __launch_bounds__(kEwThreads) k_tv2_pre(ProxDev a, Tv2Buf b, double t) {
    __shared__ double red[kEwThreads / 64];
    const int W = a.W, H = a.H;
    const long N = (long)W * H;
    const long n = N * a.nz;
    double l1 = 0, tv = 0, nnz = 0;
    const float s = a.s_mom_dev ? *a.s_mom_dev : a.s_mom;
    auto mom = [&](long i) {
        const float2 y0 = a.Y0[i], ym = a.YM[i];
        return make_float2(fmaf(s, y0.x - ym.x, y0.x), fmaf(s, y0.y - ym.y, y0.y));
    };
    for (int it = 0; it < kEwItems; ++it) {
        long i = ((long)blockIdx.x * kEwItems + it) * kEwThreads + threadIdx.x;
        if (i >= n) break;
        const float2 x0 = mom(i), g = a.G[i];
        b.xh[i] = make_float2((float)((double)x0.x - t * (double)g.x), (float)((double)x0.y - t * (double)g.y));
        b.gx[i] = make_float2(0.0f, 0.0f);
        b.gy[i] = make_float2(0.0f, 0.0f);
        if (x0.x != 0.0f || x0.y != 0.0f) {
            l1 += (double)sqrtf(x0.x * x0.x + x0.y * x0.y);
            nnz += 1.0;
        }
        const long zo = (i / N) * N;
        const int p = (int)(i - zo);
        const int y = p / W, x = p - y * W;
        const int xm = x > 0 ? x - 1 : W - 1, ym = y > 0 ? y - 1 : H - 1;
        const float2 qa = mom(zo + (long)y * W + xm), qb = mom(zo + (long)ym * W + x);
        const float gxx = qa.x - x0.x, gyx = qb.x - x0.x, gxy = qa.y - x0.y, gyy = qb.y - x0.y;
        tv += (double)sqrtf(gxx * gxx + gxy * gxy + gyx * gyx + gyy * gyy);
    }
    double* pp = a.partial + (size_t)blockIdx.x * PS_COUNT;
    double r;
    r = block_sum(l1, red);
    if (threadIdx.x == 0) pp[PS_L1X0] = r;
    r = block_sum(tv, red);
    if (threadIdx.x == 0) pp[PS_TVX0] = r;
    r = block_sum(nnz, red);
    if (threadIdx.x == 0) pp[PS_NNZX0] = r;
}

// yhat = xhat - mu (gx[x+1] - gx[x] + gy[y+1] - gy[y]), periodic; thread = voxel
__global__ void __launch_bounds__(256) k_tv2_estimate(Tv2Buf b, int W, int H, double mu) {
    const int x = blockIdx.x * 64 + (threadIdx.x & 63);
    const int y = blockIdx.y * 4 + (threadIdx.x >> 6);
    if (x >= W || y >= H) return;
    const long o = (long)blockIdx.z * W * H;
    const int xp = x < W - 1 ? x + 1 : 0, yp = y < H - 1 ? y + 1 : 0;
    const long i = o + (long)y * W + x;
    const float2 gxi = b.gx[i], gyi = b.gy[i];
    const float2 gxp = b.gx[o + (long)y * W + xp], gyp = b.gy[o + (long)yp * W + x];
    const float dx = gxp.x - gxi.x + gyp.x - gyi.x;
    const float dy = gxp.y - gxi.y + gyp.y - gyi.y;
    const float2 xh = b.xh[i];
    b.yh[i] = make_float2((float)((double)xh.x - mu * (double)dx), (float)((double)xh.y - mu * (double)dy));
}

// g = (g + sc (yhat[x-1] - yhat[x], yhat[y-1] - yhat[y])) / max(|.|, 1) over (re, im) x (x, y)
__global__ void __launch_bounds__(256) k_tv2_project(Tv2Buf b, int W, int H, double sc) {
    const int x = blockIdx.x * 64 + (threadIdx.x & 63);
    const int y = blockIdx.y * 4 + (threadIdx.x >> 6);
    if (x >= W || y >= H) return;
    const long o = (long)blockIdx.z * W * H;
    const int xm = x > 0 ? x - 1 : W - 1, ym = y > 0 ? y - 1 : H - 1;
    const long i = o + (long)y * W + x;
    const float2 c = b.yh[i], a0 = b.yh[o + (long)y * W + xm], a1 = b.yh[o + (long)ym * W + x];
    float gxx = a0.x - c.x, gyx = a1.x - c.x, gxy = a0.y - c.y, gyy = a1.y - c.y;
    const float2 gxi = b.gx[i], gyi = b.gy[i];
    gxx = (float)((double)gxi.x + sc * (double)gxx);
    gyx = (float)((double)gyi.x + sc * (double)gyx);
    gxy = (float)((double)gxi.y + sc * (double)gxy);
    gyy = (float)((double)gyi.y + sc * (double)gyy);
    float nrm = gxx * gxx + gxy * gxy + gyx * gyx + gyy * gyy;
    nrm = (float)fmax(sqrt((double)nrm), 1.0);
    b.gx[i] = make_float2(gxx / nrm, gxy / nrm);
    b.gy[i] = make_float2(gyx / nrm, gyy / nrm);
}

__global__ void __launch_bounds__(kEwThreads) k_tv2_final(ProxDev a, const float2* __restrict__ yh) {
    __shared__ double red[kEwThreads / 64];
    const long N = (long)a.W * a.H;
    const long n = N * a.nz;
    double ls1 = 0, ls2 = 0, rs = 0, nnzy = 0;
    const float s = a.s_mom_dev ? *a.s_mom_dev : a.s_mom;
    for (int it = 0; it < kEwItems; ++it) {
        long i = ((long)blockIdx.x * kEwItems + it) * kEwThreads + threadIdx.x;
        if (i >= n) break;
        const float2 y0 = a.Y0[i], ym = a.YM[i], g = a.G[i], v = yh[i];
        const float2 x0 = make_float2(fmaf(s, y0.x - ym.x, y0.x), fmaf(s, y0.y - ym.y, y0.y));
        float2 yv = a.do_soft ? make_float2(soft1(v.x, a.thr), soft1(v.y, a.thr)) : v;
        if (yv.x == 0.0f) yv.x = 0.0f;  // +0 for a zero of either sign
        if (yv.y == 0.0f) yv.y = 0.0f;
        a.Y[i] = yv;
        const float dr = yv.x - x0.x, di = yv.y - x0.y;
        ls1 += (double)(dr * g.x) + (double)(di * g.y);
        ls2 += (double)dr * dr + (double)di * di;
        const float dre = (x0.x - yv.x) * (yv.x - y0.x);
        const float dim = (x0.y - yv.y) * (yv.y - y0.y);
        const int off = (int)((i % N) & 1023);
        rs += off < 512 ? (double)dre + 2.0 * (double)dim : (double)dim;
        nnzy += (yv.x != 0.0f || yv.y != 0.0f) ? 1.0 : 0.0;
    }
    double* pp = a.partial + (size_t)blockIdx.x * PS_COUNT;
    double r;
    r = block_sum(ls1, red);
    if (threadIdx.x == 0) pp[PS_LS1] = r;
    r = block_sum(ls2, red);
    if (threadIdx.x == 0) pp[PS_LS2] = r;
    r = block_sum(rs, red);
    if (threadIdx.x == 0) pp[PS_RESTART] = r;
    r = block_sum(nnzy, red);
    if (threadIdx.x == 0) {
        pp[PS_NNZY] = r;
        pp[7] = 0.0;
    }
}

int prox_tv2_num_blocks(int W, int H, int nz) { return prox_l1_num_blocks(W, H, nz); }

hipError_t prox_tv2_general(const ProxLaunch& a, float2* work, hipStream_t s) {
    if (!work) return hipErrorInvalidValue;
    ProxDev d = make_dev(a);
    const size_t V = (size_t)a.W * a.H * a.nz;
    Tv2Buf b{work, work + V, work + 2 * V, work + 3 * V};
    const dim3 eg(prox_tv2_num_blocks(a.W, a.H, a.nz));
    const dim3 sg((a.W + 63) / 64, (a.H + 3) / 4, a.nz);
    const double mu = a.t * a.reg_tv;
    const double sc = 1 / (4 * 3 * a.t * a.reg_tv);
    hipLaunchKernelGGL(k_tv2_pre, eg, dim3(kEwThreads), 0, s, d, b, a.t);
    for (int k = 0; k < a.n_tv; ++k) {
        hipLaunchKernelGGL(k_tv2_estimate, sg, dim3(256), 0, s, b, a.W, a.H, mu);
        hipLaunchKernelGGL(k_tv2_project, sg, dim3(256), 0, s, b, a.W, a.H, sc);
    }
    hipLaunchKernelGGL(k_tv2_estimate, sg, dim3(256), 0, s, b, a.W, a.H, mu);
    hipLaunchKernelGGL(k_tv2_final, eg, dim3(kEwThreads), 0, s, d, (const float2*)b.yh);
    return hipGetLastError();
}

// ----------------------------------------------------------------------------
// Returned volume: x1 = y + s (y - y0) with its L1 / TV / nnz.
// A wave takes whole rows (row r of plane z, z and r from one division per row; the lanes
// stride along x), so no element index is divided; a workgroup of kVtWaves waves writes one
// row of partials.  MAP (s = 0, no output): the volume's line-occupancy map ([nz][H] row
// words) selects the 64-column chunks whose terms can be non-zero -- a voxel's L1 and nnz
// need its line set, its TV term its own line, the line of its left neighbour, the line above
// (and the previous plane's with tv3d) -- and the others are not read.  The terms skipped are
// exactly +0, which leaves every sum unchanged, so both forms give the same bits.
// ----------------------------------------------------------------------------
constexpr int kVtWaves = 4;
inline int vt_rows_per_wave(int W) { return W >= 512 ? 1 : (512 + W - 1) / W; }
int momentum_out_num_blocks(int W, int H, int nz) {
    const long rows = (long)H * nz;
    const long per = (long)kVtWaves * vt_rows_per_wave(W);
    return (int)((rows + per - 1) / per);
}

template <bool MAP>
__global__ void __launch_bounds__(64 * kVtWaves)
k_volume_terms(const float2* __restrict__ Y, const float2* __restrict__ Y0, float s, float2* out, int W, int H,
               int nz, int tv3d, int rpw, const uint64_t* __restrict__ occ, int lsh, double* partial) {
    __shared__ double red[kVtWaves];
    const int lane = threadIdx.x & 63;
    const int wv = threadIdx.x >> 6;
    const long N = (long)W * H;
    const long rows = (long)H * nz;
    double l1 = 0, tv = 0, nnz = 0;
    // s = 0 (no momentum: a start volume's objective, or the last update restarted) reads Y only
    const bool still = s == 0.0f;
    auto mom = [&](long i) {
        const float2 y = Y[i];
        if (still) return y;
        const float2 y0 = Y0[i];
        return make_float2(fmaf(s, y.x - y0.x, y.x), fmaf(s, y.y - y0.y, y.y));
    };
    const int nl = ((W - 1) >> lsh) + 1;  // map lines per row (<= 64)
    for (int r = 0; r < rpw; ++r) {
        const long gr = ((long)blockIdx.x * kVtWaves + wv) * rpw + r;
        if (gr >= rows) break;
        const int z = (int)(gr / H);
        const int y = (int)(gr - (long)z * H);
        const int yu = y > 0 ? y - 1 : H - 1;
        const int zp = z > 0 ? z - 1 : nz - 1;
        const long ro = (long)z * N + (long)y * W;
        const long ru = (long)z * N + (long)yu * W;
        const long rp = (long)zp * N + (long)y * W;
        uint64_t mask = ~0ull;
        if constexpr (MAP) {
            const uint64_t w = occ[gr];
            // line 0's left neighbour is the row's last line (periodic x)
            mask = w | occ[(long)z * H + yu] | (w << 1) | ((w >> (nl - 1)) & 1ull);
            if (tv3d) mask |= occ[(long)zp * H + y];
            if (!mask) continue;
        }
        for (int x0 = 0; x0 < W; x0 += 64) {
            if constexpr (MAP) {
                const int la = x0 >> lsh, lb = min(x0 + 63, W - 1) >> lsh;
                const uint64_t cm = (lb >= 63 ? ~0ull : ((1ull << (lb + 1)) - 1ull)) & ~((1ull << la) - 1ull);
                if (!(mask & cm)) continue;
            }
            const int x = x0 + lane;
            if (x >= W) continue;
            const float2 v = mom(ro + x);
            if (out) out[ro + x] = v;
            if (v.x != 0.0f || v.y != 0.0f) {
                l1 += (double)sqrtf(v.x * v.x + v.y * v.y);
                nnz += 1.0;
            }
            const int xm = x > 0 ? x - 1 : W - 1;
            const float2 a = mom(ro + xm), b = mom(ru + x);
            const float gxx = a.x - v.x, gyx = b.x - v.x, gxy = a.y - v.y, gyy = b.y - v.y;
            float n2 = gxx * gxx + gxy * gxy + gyx * gyx + gyy * gyy;
            if (tv3d) {  // calcTVNorm: against the previous plane, z-periodic (sparse_volume.cu:236-265)
                const float2 c = mom(rp + x);
                const float gzx = c.x - v.x, gzy = c.y - v.y;
                n2 = n2 + gzx * gzx + gzy * gzy;
            }
            tv += (double)sqrtf(n2);
        }
    }
    double* pp = partial + (size_t)blockIdx.x * PS_COUNT;
    double rsum;
    rsum = block_sum(l1, red);
    if (threadIdx.x == 0) pp[PS_L1X0] = rsum;
    rsum = block_sum(tv, red);
    if (threadIdx.x == 0) pp[PS_TVX0] = rsum;
    rsum = block_sum(nnz, red);
    if (threadIdx.x == 0) {
        pp[PS_NNZX0] = rsum;
        pp[PS_LS1] = pp[PS_LS2] = pp[PS_RESTART] = pp[PS_NNZY] = pp[7] = 0.0;
    }
}

hipError_t momentum_out(const float2* Y, const float2* Y0, float s_mom, float2* out, int W, int H, int nz,
                        bool tv3d, double* partial, hipStream_t s) {
    dim3 grid(momentum_out_num_blocks(W, H, nz));
    hipLaunchKernelGGL(k_volume_terms<false>, grid, dim3(64 * kVtWaves), 0, s, Y, Y0, s_mom, out, W, H, nz,
                       tv3d ? 1 : 0, vt_rows_per_wave(W), (const uint64_t*)nullptr, 4, partial);
    return hipGetLastError();
}

hipError_t volume_terms_map(const float2* X, const uint64_t* occ, int W, int H, int nz, bool tv3d,
                            double* partial, hipStream_t s) {
    int lsh = 4;
    while (((W - 1) >> lsh) >= 64) ++lsh;
    dim3 grid(momentum_out_num_blocks(W, H, nz));
    hipLaunchKernelGGL(k_volume_terms<true>, grid, dim3(64 * kVtWaves), 0, s, X, X, 0.0f, (float2*)nullptr, W, H,
                       nz, tv3d ? 1 : 0, vt_rows_per_wave(W), occ, lsh, partial);
    return hipGetLastError();
}

// ----------------------------------------------------------------------------
// Hologram-sized kernels
// ----------------------------------------------------------------------------
constexpr int kHoloItems = kEwItems;  // elements per thread
int holo_num_blocks(long n) { return (int)((n + (long)kEwThreads * kHoloItems - 1) / ((long)kEwThreads * kHoloItems)); }

__global__ void __launch_bounds__(kEwThreads)
k_holo_momentum(const float2* __restrict__ cur, const float2* __restrict__ prev, double s_mom,
                const double* __restrict__ step_dev,
                const float2* __restrict__ meas, float2* est, float2* resid, long n, double* partial,
                const double* __restrict__ restart_sum) {
    __shared__ double red[kEwThreads / 64];
    // with restart_sum: the step is 0 when the iteration restarts (restart_sum > 0), as the
    // host's alpha update gives, so the step can be chosen before the host has the sum
    if (step_dev) s_mom = *step_dev;
    if (restart_sum && *restart_sum > 0) s_mom = 0.0;
    double acc = 0;
    for (int it = 0; it < kHoloItems; ++it) {
        long i = ((long)blockIdx.x * kHoloItems + it) * kEwThreads + threadIdx.x;
        if (i >= n) break;
        const float2 c = cur[i], p = prev[i], m = meas[i];
        // sparse_fistaUpdate_kernel on holograms: double step, float storage
        float2 e = make_float2((float)((double)c.x + s_mom * (double)(c.x - p.x)),
                               (float)((double)c.y + s_mom * (double)(c.y - p.y)));
        est[i] = e;
        float2 r = make_float2(e.x - m.x, e.y - m.y);
        resid[i] = r;
        acc += (double)r.x * r.x + (double)r.y * r.y;
    }
    double v = block_sum(acc, red);
    if (threadIdx.x == 0) partial[blockIdx.x] = v;
}

hipError_t holo_momentum(const float2* cur, const float2* prev, double s_mom, const float2* meas, float2* est,
                         float2* resid, long n, double* partial, hipStream_t s, const double* restart_sum,
                         const double* step_dev) {
    hipLaunchKernelGGL(k_holo_momentum, dim3(holo_num_blocks(n)), dim3(kEwThreads), 0, s, cur, prev, s_mom, step_dev,
                       meas, est, resid, n, partial, restart_sum);
    return hipGetLastError();
}

__global__ void k_fill_u64(uint64_t* __restrict__ p, long n, uint64_t v) {
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}
hipError_t fill_u64(uint64_t* p, long n, uint64_t v, hipStream_t s) {
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_fill_u64, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, p, n, v);
    return hipGetLastError();
}

// A sparse synthetic volume: voxel i is non-zero (re, im in [-0.5, 0.5)) when a hash of i falls
// in 1 / period of its range, else +0.
__global__ void k_synth_volume(float2* __restrict__ v, long n, unsigned period, unsigned seed) {
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    unsigned h = (unsigned)i * 0x9E3779B1u ^ seed;
    h ^= h >> 16;
    h *= 0x85EBCA6Bu;
    h ^= h >> 13;
    h *= 0xC2B2AE35u;
    h ^= h >> 16;
    float2 val = make_float2(0.0f, 0.0f);
    if (h % period == 0) val = make_float2((float)((h >> 8) & 0xffff) / 65536.0f - 0.5f, (float)(h & 0xff) / 256.0f - 0.5f);
    v[i] = val;
}
hipError_t synth_volume(float2* v, long n, unsigned period, unsigned seed, hipStream_t s) {
    hipLaunchKernelGGL(k_synth_volume, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, v, n, period, seed);
    return hipGetLastError();
}

__global__ void k_set_scalars(float* __restrict__ f, float fv, double* __restrict__ d, double dv) {
    if (f) *f = fv;
    if (d) *d = dv;
}
hipError_t set_scalars(float* f, float fv, double* d, double dv, hipStream_t s) {
    hipLaunchKernelGGL(k_set_scalars, dim3(1), dim3(1), 0, s, f, fv, d, dv);
    return hipGetLastError();
}

__global__ void __launch_bounds__(kEwThreads)
k_holo_residual(const float2* __restrict__ est, const float2* __restrict__ meas, float2* resid, long n,
                double* partial) {
    __shared__ double red[kEwThreads / 64];
    double acc = 0;
    for (int it = 0; it < kHoloItems; ++it) {
        long i = ((long)blockIdx.x * kHoloItems + it) * kEwThreads + threadIdx.x;
        if (i >= n) break;
        const float2 e = est[i], m = meas[i];
        float2 r = make_float2(e.x - m.x, e.y - m.y);
        resid[i] = r;
        acc += (double)r.x * r.x + (double)r.y * r.y;
    }
    double v = block_sum(acc, red);
    if (threadIdx.x == 0) partial[blockIdx.x] = v;
}

// sum |v|^2 over n complex values: nblocks partial sums (grid-stride), partial[b]
__global__ void __launch_bounds__(kEwThreads) k_norm2(const float2* __restrict__ v, long n, double* partial) {
    __shared__ double red[kEwThreads / 64];
    double acc = 0;
    for (long i = (long)blockIdx.x * kEwThreads + threadIdx.x; i < n; i += (long)gridDim.x * kEwThreads) {
        const float2 e = v[i];
        acc += (double)e.x * e.x + (double)e.y * e.y;
    }
    double r = block_sum(acc, red);
    if (threadIdx.x == 0) partial[blockIdx.x] = r;
}
hipError_t norm2_partials(const float2* v, long n, int nblocks, double* partial, hipStream_t s) {
    hipLaunchKernelGGL(k_norm2, dim3(nblocks), dim3(kEwThreads), 0, s, v, n, partial);
    return hipGetLastError();
}

// max |v| over n complex values: nblocks partial maxima (grid-stride), partial[b]
__global__ void __launch_bounds__(kEwThreads) k_maxabs(const float2* __restrict__ v, long n, double* partial) {
    __shared__ float red[kEwThreads / 64];
    float m = 0.0f;
    for (long i = (long)blockIdx.x * kEwThreads + threadIdx.x; i < n; i += (long)gridDim.x * kEwThreads) {
        const float2 e = v[i];
        m = fmaxf(m, e.x * e.x + e.y * e.y);
    }
    for (int o = 32; o > 0; o >>= 1) m = fmaxf(m, __shfl_xor(m, o));
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = m;
    __syncthreads();
    if (threadIdx.x == 0) {
        float r = red[0];
        for (int w = 1; w < kEwThreads / 64; ++w) r = fmaxf(r, red[w]);
        partial[blockIdx.x] = sqrt((double)r);
    }
}
hipError_t maxabs_partials(const float2* v, long n, int nblocks, double* partial, hipStream_t s) {
    hipLaunchKernelGGL(k_maxabs, dim3(nblocks), dim3(kEwThreads), 0, s, v, n, partial);
    return hipGetLastError();
}

// The complex soft threshold out = in * max(0, 1 - tau / |in|) (+0 where |in| <= tau) into two
// outputs at once (thresholds tau0, tau1: one read of the input), each with its line-occupancy
// map: a wave takes whole rows (as k_volume_terms), ORs its lanes' line bits over
// the row, stores the row word (the row is the wave's alone) and ORs it into the plane word
// after the Nz H row words (cleared beforehand).
__global__ void __launch_bounds__(64 * kVtWaves)
k_soft_map2(const float2* __restrict__ in, int W, int H, int nz, int rpw, int lsh, float tau0, float2* __restrict__ out0,
            uint64_t* occ0, float tau1, float2* __restrict__ out1, uint64_t* occ1) {
    const int lane = threadIdx.x & 63;
    const int wv = threadIdx.x >> 6;
    const long rows = (long)H * nz;
    auto soft = [](float2 e, float m, float tau) {
        float2 r = make_float2(0.0f, 0.0f);
        if (m > tau) {
            const float sc = (m - tau) / m;
            r = make_float2(e.x * sc, e.y * sc);
        }
        return r;
    };
    for (int rr = 0; rr < rpw; ++rr) {
        const long gr = ((long)blockIdx.x * kVtWaves + wv) * rpw + rr;
        if (gr >= rows) break;
        const long ro = gr * W;
        uint64_t b0 = 0, b1 = 0;
        for (int x = lane; x < W; x += 64) {
            const float2 e = in[ro + x];
            const float m = sqrtf(e.x * e.x + e.y * e.y);
            const float2 r0 = soft(e, m, tau0), r1 = soft(e, m, tau1);
            out0[ro + x] = r0;
            out1[ro + x] = r1;
            if (r0.x != 0.0f || r0.y != 0.0f) b0 |= 1ull << (x >> lsh);
            if (r1.x != 0.0f || r1.y != 0.0f) b1 |= 1ull << (x >> lsh);
        }
        for (int o = 32; o > 0; o >>= 1) {
            b0 |= __shfl_xor(b0, o);
            b1 |= __shfl_xor(b1, o);
        }
        if (lane == 0) {
            const long z = gr / H;
            occ0[gr] = b0;
            occ1[gr] = b1;
            if (b0) atomicOr((unsigned long long*)&occ0[rows + z], (unsigned long long)b0);
            if (b1) atomicOr((unsigned long long*)&occ1[rows + z], (unsigned long long)b1);
        }
    }
}
hipError_t soft_threshold_map2(const float2* in, int W, int H, int nz, float tau0, float2* out0, uint64_t* occ0,
                               float tau1, float2* out1, uint64_t* occ1, hipStream_t s) {
    int lsh = 4;
    while (((W - 1) >> lsh) >= 64) ++lsh;
    hipLaunchKernelGGL(k_soft_map2, dim3(momentum_out_num_blocks(W, H, nz)), dim3(64 * kVtWaves), 0, s, in, W, H, nz,
                       vt_rows_per_wave(W), lsh, tau0, out0, occ0, tau1, out1, occ1);
    return hipGetLastError();
}

hipError_t holo_residual(const float2* est, const float2* meas, float2* resid, long n, double* partial,
                         hipStream_t s) {
    hipLaunchKernelGGL(k_holo_residual, dim3(holo_num_blocks(n)), dim3(kEwThreads), 0, s, est, meas, resid, n,
                       partial);
    return hipGetLastError();
}

// The spectrum of A's real-valued estimate from the partial spectra S_p of its plane chunks:
// est = Re(IFFT(S)) with S = sum_p S_p has FFT(est) / N = (S(k) + conj S(-k)) / 2 (the real part
// of a transform is the Hermitian-symmetric half of its spectrum), written to est; with the
// spectrum of the measurement bspec (FFT(b) / N), the residual's is est - bspec, and
// partial[block] = sum |est - bspec|^2 (N times that is |A x - b|^2, by Parseval).  A thread
// takes an element k of rows 0 .. H/2 whose mirror -k is not before it and writes both
// (E(-k) = conj E(k)); n = (H/2 + 1) W elements are visited (H and W even).
// One pair per thread, so that its loads are all in flight together.
int holo_spec_num_blocks(int W, int H) { return (int)(((long)(H / 2 + 1) * W + kEwThreads - 1) / kEwThreads); }

__global__ void __launch_bounds__(kEwThreads)
k_holo_spec(const float2* __restrict__ parts, int nparts, long stride, int W, int H,
            const float2* __restrict__ bspec, float2* __restrict__ est, long n, double* partial) {
    __shared__ double red[kEwThreads / 64];
    double acc = 0;
    const long i = (long)blockIdx.x * kEwThreads + threadIdx.x;
    const int ky = (int)(i / W), kx = (int)(i - (long)ky * W);
    const long m = (long)(ky ? H - ky : 0) * W + (kx ? W - kx : 0);
    if (i < n && m >= i) {  // else: outside, or the pair belongs to element m
        float2 s1 = parts[i], s2 = parts[m];
        for (int q = 1; q < nparts; ++q) {
            s1 = cadd(s1, parts[(size_t)q * stride + i]);
            s2 = cadd(s2, parts[(size_t)q * stride + m]);
        }
        const float2 e = make_float2(0.5f * (s1.x + s2.x), 0.5f * (s1.y - s2.y));
        const float2 bi = bspec[i];
        est[i] = e;
        const float2 r = make_float2(e.x - bi.x, e.y - bi.y);
        acc += (double)r.x * r.x + (double)r.y * r.y;
        if (m != i) {
            const float2 em = make_float2(e.x, -e.y);
            const float2 bm = bspec[m];
            est[m] = em;
            const float2 rm = make_float2(em.x - bm.x, em.y - bm.y);
            acc += (double)rm.x * rm.x + (double)rm.y * rm.y;
        }
    }
    double v = block_sum(acc, red);
    if (threadIdx.x == 0) partial[blockIdx.x] = v;
}

hipError_t holo_spec(const float2* parts, int nparts, long stride, int W, int H, const float2* bspec, float2* est,
                     double* partial, hipStream_t s) {
    if ((W | H) & 1) return hipErrorInvalidValue;
    const long n = (long)(H / 2 + 1) * W;
    hipLaunchKernelGGL(k_holo_spec, dim3(holo_spec_num_blocks(W, H)), dim3(kEwThreads), 0, s, parts, nparts, stride,
                       W, H, bspec, est, n, partial);
    return hipGetLastError();
}

__global__ void __launch_bounds__(kEwThreads) k_holo_scale(const float2* __restrict__ in, float2* out, long n, float f) {
    const long i = (long)blockIdx.x * kEwThreads + threadIdx.x;
    if (i < n) out[i] = make_float2(in[i].x * f, in[i].y * f);
}

hipError_t holo_scale(const float2* in, float2* out, long n, float f, hipStream_t s) {
    hipLaunchKernelGGL(k_holo_scale, dim3((unsigned)((n + kEwThreads - 1) / kEwThreads)), dim3(kEwThreads), 0, s, in,
                       out, n, f);
    return hipGetLastError();
}

// Two fixed-order passes (deterministic): kReduceBlocks workgroups each sum a
// contiguous slice of rows of every column (row-contiguous, coalesced loads) into
// work[block][ncols]; one workgroup then sums those rows in block order.
constexpr int kReduceThreads = 256;
__global__ void __launch_bounds__(kReduceThreads) k_reduce_rows(const double* __restrict__ partial, int nrows,
                                                                int ncols, int rows_per_block, double* out) {
    __shared__ double red[kReduceThreads / 64];
    double acc[PS_COUNT];
#pragma unroll
    for (int q = 0; q < PS_COUNT; ++q) acc[q] = 0.0;
    const int r0 = blockIdx.x * rows_per_block;
    const int r1 = min(nrows, r0 + rows_per_block);
    for (int r = r0 + threadIdx.x; r < r1; r += kReduceThreads) {
        const double* row = partial + (size_t)r * ncols;
#pragma unroll
        for (int q = 0; q < PS_COUNT; ++q)
            if (q < ncols) acc[q] += row[q];
    }
#pragma unroll
    for (int q = 0; q < PS_COUNT; ++q) {
        if (q >= ncols) break;
        const double v = block_sum(acc[q], red);
        if (threadIdx.x == 0) out[(size_t)blockIdx.x * ncols + q] = v;
    }
}

hipError_t reduce_partials(const double* partial, int nrows, int ncols, double* work, double* out, hipStream_t s) {
    if (ncols < 1 || ncols > PS_COUNT || nrows < 0) return hipErrorInvalidValue;
    const int rpb = (nrows + kReduceBlocks - 1) / kReduceBlocks;
    const int nb = rpb > 0 ? (nrows + rpb - 1) / rpb : 1;
    hipLaunchKernelGGL(k_reduce_rows, dim3(nb), dim3(kReduceThreads), 0, s, partial, nrows, ncols, rpb > 0 ? rpb : 1,
                       work);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(k_reduce_rows, dim3(1), dim3(kReduceThreads), 0, s, (const double*)work, nb, ncols, nb, out);
    return hipGetLastError();
}

__global__ void __launch_bounds__(256) k_check_finite(const float2* __restrict__ a, long n, int* flag) {
    long i = (long)blockIdx.x * 256 + threadIdx.x;
    if (i < n) {
        const float2 v = a[i];
        if (!isfinite(v.x) || !isfinite(v.y)) *flag = 1;
    }
}

hipError_t check_finite(const float2* a, long n, int* flag, hipStream_t s) {
    hipLaunchKernelGGL(k_check_finite, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, a, n, flag);
    return hipGetLastError();
}

}  // namespace rihvr
