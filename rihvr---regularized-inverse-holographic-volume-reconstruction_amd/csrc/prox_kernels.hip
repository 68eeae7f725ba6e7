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
// ends of the row chunk in y).  No workgroup barriers; per-lane history rings
// of x-hat, x0, g and y_{k-1} live in LDS.
#include "device_common.hpp"
#include "kernels.hpp"

namespace rihvr {

constexpr int kRowChunk = 128;   // output rows per wave
constexpr int kProxWaves = 4;    // waves per workgroup
constexpr int kMaxNtv = 10;

struct ProxGeom {
    int ow, nstrips, rc, nchunks;
};
__host__ __device__ inline ProxGeom prox_geom(int W, int H, int n) {
    ProxGeom g;
    g.ow = 64 - 2 * n;
    if (g.ow > W) g.ow = W;
    g.nstrips = (W + g.ow - 1) / g.ow;
    g.rc = H < kRowChunk ? H : kRowChunk;
    g.nchunks = (H + g.rc - 1) / g.rc;
    return g;
}

int prox_fl2d_num_waves(int W, int H, int nz, int n_tv) {
    ProxGeom g = prox_geom(W, H, n_tv);
    return g.nstrips * g.nchunks * nz;
}
bool prox_fl2d_supported(int n_tv) { return n_tv >= 0 && n_tv <= kMaxNtv; }

struct ProxDev {
    const float2* Y0;
    const float2* YM;
    const float2* G;
    float2* Y;
    int W, H, nz;
    float s_mom, t, mu, sc, thr;
    int do_soft;
    ProxGeom geo;
    int nwaves;
    double* partial;
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

template <int NTV>
__global__ void __launch_bounds__(64 * kProxWaves) k_prox_fl2d(ProxDev a) {
    constexpr int P = NTV + 3;  // history depth: rows tau .. tau-NTV-2
    __shared__ float2 ring[kProxWaves][4][P][64];  // fields: xhat, x0, g, y_{k-1}
    const int lane = threadIdx.x & 63;
    const int wl = threadIdx.x >> 6;
    const int w = blockIdx.x * kProxWaves + wl;
    if (w >= a.nwaves) return;  // whole wave exits; no barriers below
    const ProxGeom geo = a.geo;
    const int per_plane = geo.nstrips * geo.nchunks;
    const int z = w / per_plane;
    const int rem = w - z * per_plane;
    const int chunk = rem / geo.nstrips;
    const int strip = rem - chunk * geo.nstrips;
    const int W = a.W, H = a.H;
    const int ocol = strip * geo.ow + (lane - NTV);  // output column of this lane
    const bool out_lane = (lane - NTV) >= 0 && (lane - NTV) < geo.ow && ocol < W;
    const int xcol = wrapi(strip * geo.ow - NTV + lane, W);
    const int r0 = chunk * geo.rc;
    const int r1 = min(H, r0 + geo.rc);
    const size_t plane_off = (size_t)z * H * W;
    const float s = a.s_mom, t = a.t, mu = a.mu, sc = a.sc;

    float4 gcur[NTV + 1], gprev[NTV + 1];  // (gx.re, gx.im, gy.re, gy.im) per TV level
    float2 yprev[NTV + 1];
#pragma unroll
    for (int k = 0; k <= NTV; ++k) {
        gcur[k] = make_float4(0.f, 0.f, 0.f, 0.f);
        gprev[k] = gcur[k];
        yprev[k] = make_float2(0.f, 0.f);
    }
    double acc_ls1 = 0, acc_ls2 = 0, acc_rs = 0, acc_l1 = 0, acc_tv = 0, acc_nnz0 = 0, acc_nnzy = 0;

    const int tau0 = r0 - (NTV > 0 ? NTV : 1);
    const int tau1 = r1 + NTV;  // inclusive
    int slot = 0;               // ring slot of row tau
    for (int tau = tau0; tau <= tau1; ++tau) {
        // ---- load row tau: momentum point, x-hat ----
        {
            const size_t o = plane_off + (size_t)wrapi(tau, H) * W + xcol;
            const float2 y0 = a.Y0[o], ym = a.YM[o], g = a.G[o];
            float2 x0 = make_float2(fmaf(s, y0.x - ym.x, y0.x), fmaf(s, y0.y - ym.y, y0.y));
            float2 xh = make_float2(fmaf(-t, g.x, x0.x), fmaf(-t, g.y, x0.y));
            ring[wl][0][slot][lane] = xh;
            ring[wl][1][slot][lane] = x0;
            ring[wl][2][slot][lane] = g;
            ring[wl][3][slot][lane] = y0;
        }
        // ---- TV levels: level k works on row tau - k ----
#pragma unroll
        for (int k = 1; k <= NTV; ++k) {
            int sk = slot - k;
            if (sk < 0) sk += P;
            const float2 xh = ring[wl][0][sk][lane];
            float2 yh;
            float4 gold;  // g_{k-1} at this row
            if (k == 1) {
                yh = xh;
                gold = make_float4(0.f, 0.f, 0.f, 0.f);
            } else {
                gold = gprev[k - 1];
                const float gxr_p = __shfl_down(gold.x, 1, 64);
                const float gxi_p = __shfl_down(gold.y, 1, 64);
                const float4 gdn = gcur[k - 1];  // row below (tau-k+1)
                const float dvr = gxr_p - gold.x + gdn.z - gold.z;
                const float dvi = gxi_p - gold.y + gdn.w - gold.w;
                yh = make_float2(fmaf(-mu, dvr, xh.x), fmaf(-mu, dvi, xh.y));
            }
            const float ylr = __shfl_up(yh.x, 1, 64);
            const float yli = __shfl_up(yh.y, 1, 64);
            const float2 yu = yprev[k];
            float axr = fmaf(sc, ylr - yh.x, gold.x);
            float axi = fmaf(sc, yli - yh.y, gold.y);
            float ayr = fmaf(sc, yu.x - yh.x, gold.z);
            float ayi = fmaf(sc, yu.y - yh.y, gold.w);
            const float n2 = axr * axr + axi * axi + ayr * ayr + ayi * ayi;
            const float inv = n2 > 1.0f ? __frsqrt_rn(n2) : 1.0f;
            gprev[k] = gcur[k];
            gcur[k] = make_float4(axr * inv, axi * inv, ayr * inv, ayi * inv);
            yprev[k] = yh;
        }
        // ---- final estimate and threshold: row o = tau - NTV - 1 ----
        const int o = tau - NTV - 1;
        int so = slot - NTV - 1;
        if (so < 0) so += P;
        const float2 xh = ring[wl][0][so][lane];
        float2 yv;
        if constexpr (NTV == 0) {
            yv = xh;
        } else {
            const float4 gu = gprev[NTV];
            const float4 gd = gcur[NTV];
            const float gxr_p = __shfl_down(gu.x, 1, 64);
            const float gxi_p = __shfl_down(gu.y, 1, 64);
            const float dvr = gxr_p - gu.x + gd.z - gu.z;
            const float dvi = gxi_p - gu.y + gd.w - gu.w;
            yv = make_float2(fmaf(-mu, dvr, xh.x), fmaf(-mu, dvi, xh.y));
        }
        if (a.do_soft) yv = make_float2(soft1(yv.x, a.thr), soft1(yv.y, a.thr));
        // TV of x0 needs x0 at x-1 (shuffle: every lane participates)
        const float2 x0o = ring[wl][1][so][lane];
        const float x0lr = __shfl_up(x0o.x, 1, 64);
        const float x0li = __shfl_up(x0o.y, 1, 64);
        if (o >= r0 && out_lane) {
            const size_t oo = plane_off + (size_t)o * W + ocol;
            a.Y[oo] = yv;
            int su = so - 1;
            if (su < 0) su += P;
            const float2 x0u = ring[wl][1][su][lane];
            const float2 g = ring[wl][2][so][lane];
            const float2 y0 = ring[wl][3][so][lane];
            const float dr = yv.x - x0o.x, di = yv.y - x0o.y;
            acc_ls1 += (double)(dr * g.x) + (double)(di * g.y);
            acc_ls2 += (double)dr * dr + (double)di * di;
            const float dre = (x0o.x - yv.x) * (yv.x - y0.x);
            const float dim = (x0o.y - yv.y) * (yv.y - y0.y);
            const int off = (int)(((size_t)o * W + ocol) & 1023);
            acc_rs += off < 512 ? (double)dre + 2.0 * (double)dim : (double)dim;
            const bool nz0 = (x0o.x != 0.0f) || (x0o.y != 0.0f);
            if (nz0) {
                acc_l1 += (double)sqrtf(x0o.x * x0o.x + x0o.y * x0o.y);
                acc_nnz0 += 1.0;
            }
            const float gxx = x0lr - x0o.x, gyx = x0u.x - x0o.x, gxy = x0li - x0o.y, gyy = x0u.y - x0o.y;
            acc_tv += (double)sqrtf(gxx * gxx + gxy * gxy + gyx * gyx + gyy * gyy);
            acc_nnzy += ((yv.x != 0.0f) || (yv.y != 0.0f)) ? 1.0 : 0.0;
        }
        slot = slot + 1 == P ? 0 : slot + 1;
    }
    acc_ls1 = wave_sum(acc_ls1);
    acc_ls2 = wave_sum(acc_ls2);
    acc_rs = wave_sum(acc_rs);
    acc_l1 = wave_sum(acc_l1);
    acc_tv = wave_sum(acc_tv);
    acc_nnz0 = wave_sum(acc_nnz0);
    acc_nnzy = wave_sum(acc_nnzy);
    if (lane == 0) {
        double* p = a.partial + (size_t)w * PS_COUNT;
        p[PS_LS1] = acc_ls1;
        p[PS_LS2] = acc_ls2;
        p[PS_RESTART] = acc_rs;
        p[PS_L1X0] = acc_l1;
        p[PS_TVX0] = acc_tv;
        p[PS_NNZX0] = acc_nnz0;
        p[PS_NNZY] = acc_nnzy;
        p[7] = 0.0;
    }
}

static ProxDev make_dev(const ProxLaunch& a) {
    ProxDev d;
    d.Y0 = a.Y0; d.YM = a.YM; d.G = a.G; d.Y = a.Y;
    d.W = a.W; d.H = a.H; d.nz = a.nz;
    d.s_mom = a.s_mom;
    d.t = (float)a.t;
    d.mu = (float)(a.t * a.reg_tv);
    d.sc = a.reg_tv != 0.0 ? (float)(1 / (4 * 3 * a.t * a.reg_tv)) : 0.0f;
    d.thr = (float)(a.t * a.reg);
    d.do_soft = a.reg > 0.0;
    d.geo = prox_geom(a.W, a.H, a.n_tv);
    d.nwaves = d.geo.nstrips * d.geo.nchunks * a.nz;
    d.partial = a.partial;
    return d;
}

template <int NTV> static hipError_t launch_fl2d(const ProxDev& d, hipStream_t s) {
    dim3 grid((d.nwaves + kProxWaves - 1) / kProxWaves);
    hipLaunchKernelGGL(k_prox_fl2d<NTV>, grid, dim3(64 * kProxWaves), 0, s, d);
    return hipGetLastError();
}

hipError_t prox_fl2d(const ProxLaunch& a, hipStream_t s) {
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
    const float s = a.s_mom;
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
// Returned volume: x1 = y + s (y - y0) with its L1 / TV / nnz.
// ----------------------------------------------------------------------------
int momentum_out_num_blocks(int W, int H, int nz) { return prox_l1_num_blocks(W, H, nz); }

__global__ void __launch_bounds__(kEwThreads)
k_momentum_out(const float2* __restrict__ Y, const float2* __restrict__ Y0, float s, float2* out, int W, int H,
               int nz, double* partial) {
    __shared__ double red[kEwThreads / 64];
    const long N = (long)W * H;
    const long n = N * nz;
    double l1 = 0, tv = 0, nnz = 0;
    auto mom = [&](long i) {
        const float2 y = Y[i], y0 = Y0[i];
        return make_float2(fmaf(s, y.x - y0.x, y.x), fmaf(s, y.y - y0.y, y.y));
    };
    for (int it = 0; it < kEwItems; ++it) {
        long i = ((long)blockIdx.x * kEwItems + it) * kEwThreads + threadIdx.x;
        if (i >= n) break;
        const long zo = (i / N) * N;
        const int p = (int)(i - zo);
        const int y = p / W, x = p - y * W;
        const float2 v = mom(i);
        if (out) out[i] = v;
        if (v.x != 0.0f || v.y != 0.0f) {
            l1 += (double)sqrtf(v.x * v.x + v.y * v.y);
            nnz += 1.0;
        }
        const int xm = x > 0 ? x - 1 : W - 1, ym = y > 0 ? y - 1 : H - 1;
        const float2 a = mom(zo + (long)y * W + xm), b = mom(zo + (long)ym * W + x);
        const float gxx = a.x - v.x, gyx = b.x - v.x, gxy = a.y - v.y, gyy = b.y - v.y;
        tv += (double)sqrtf(gxx * gxx + gxy * gxy + gyx * gyx + gyy * gyy);
    }
    double* pp = partial + (size_t)blockIdx.x * PS_COUNT;
    double r;
    r = block_sum(l1, red);
    if (threadIdx.x == 0) pp[PS_L1X0] = r;
    r = block_sum(tv, red);
    if (threadIdx.x == 0) pp[PS_TVX0] = r;
    r = block_sum(nnz, red);
    if (threadIdx.x == 0) {
        pp[PS_NNZX0] = r;
        pp[PS_LS1] = pp[PS_LS2] = pp[PS_RESTART] = pp[PS_NNZY] = pp[7] = 0.0;
    }
}

hipError_t momentum_out(const float2* Y, const float2* Y0, float s_mom, float2* out, int W, int H, int nz,
                        double* partial, hipStream_t s) {
    dim3 grid(momentum_out_num_blocks(W, H, nz));
    hipLaunchKernelGGL(k_momentum_out, grid, dim3(kEwThreads), 0, s, Y, Y0, s_mom, out, W, H, nz, partial);
    return hipGetLastError();
}

// ----------------------------------------------------------------------------
// Hologram-sized kernels
// ----------------------------------------------------------------------------
int holo_num_blocks(long n) { return (int)((n + (long)kEwThreads * kEwItems - 1) / ((long)kEwThreads * kEwItems)); }

__global__ void __launch_bounds__(kEwThreads)
k_holo_momentum(const float2* __restrict__ cur, const float2* __restrict__ prev, double s_mom,
                const float2* __restrict__ meas, float2* est, float2* resid, long n, double* partial) {
    __shared__ double red[kEwThreads / 64];
    double acc = 0;
    for (int it = 0; it < kEwItems; ++it) {
        long i = ((long)blockIdx.x * kEwItems + it) * kEwThreads + threadIdx.x;
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
                         float2* resid, long n, double* partial, hipStream_t s) {
    hipLaunchKernelGGL(k_holo_momentum, dim3(holo_num_blocks(n)), dim3(kEwThreads), 0, s, cur, prev, s_mom,
                       meas, est, resid, n, partial);
    return hipGetLastError();
}

__global__ void __launch_bounds__(kEwThreads)
k_holo_residual(const float2* __restrict__ est, const float2* __restrict__ meas, float2* resid, long n,
                double* partial) {
    __shared__ double red[kEwThreads / 64];
    double acc = 0;
    for (int it = 0; it < kEwItems; ++it) {
        long i = ((long)blockIdx.x * kEwItems + it) * kEwThreads + threadIdx.x;
        if (i >= n) break;
        const float2 e = est[i], m = meas[i];
        float2 r = make_float2(e.x - m.x, e.y - m.y);
        resid[i] = r;
        acc += (double)r.x * r.x + (double)r.y * r.y;
    }
    double v = block_sum(acc, red);
    if (threadIdx.x == 0) partial[blockIdx.x] = v;
}

hipError_t holo_residual(const float2* est, const float2* meas, float2* resid, long n, double* partial,
                         hipStream_t s) {
    hipLaunchKernelGGL(k_holo_residual, dim3(holo_num_blocks(n)), dim3(kEwThreads), 0, s, est, meas, resid, n,
                       partial);
    return hipGetLastError();
}

__global__ void __launch_bounds__(256) k_reduce_partials(const double* __restrict__ partial, int nrows, int ncols,
                                                         double* out) {
    __shared__ double red[4];
    const int q = blockIdx.x;
    double acc = 0;
    for (int r = threadIdx.x; r < nrows; r += 256) acc += partial[(size_t)r * ncols + q];
    double v = block_sum(acc, red);
    if (threadIdx.x == 0) out[q] = v;
}

hipError_t reduce_partials(const double* partial, int nrows, int ncols, double* out, hipStream_t s) {
    hipLaunchKernelGGL(k_reduce_partials, dim3(ncols), dim3(256), 0, s, partial, nrows, ncols, out);
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
