// fft_generic.hip -- FFTs of any length from 2 to 4096 for the forward operator A, its adjoint
// and the hologram transforms, where the compile-time plans of fft_kernels.hip do not cover the
// hologram size (the reference plans any size with cufftPlan2d(H, W),
// sparse_compressive_holo.cu:98).
//
// A row kernel transforms one row of length L per workgroup with a plan chosen at run time:
// Stockham autosort passes (the same indexing as fft.hpp) of radices 16, 8, 4, 2 (hand-written
// butterflies), 3, 5, 7 (direct DFTs with compile-time size) and a direct DFT of any other prime
// size (one output per work item, read back from LDS), ping-ponging between two LDS buffers.
// The row's prologue and epilogue carry the operator's arithmetic as in the tuned kernels:
// H_z R-hat / N for A^T, the H_{-z} z-sum into partial spectra for A, the sum of partial
// spectra for the estimate.  Transforms along y run as row transforms between tiled transposes
// (32 x 32 through LDS) that also apply the per-plane phase factors and, for the estimate, the
// real part, the residual and its norm.
#include "fft.hpp"
#include "kernels.hpp"

namespace rihvr {

namespace {

constexpr int kGenThreads = 256;
constexpr int kGenMaxL = 4096;
constexpr int kGenMaxRadices = 16;

struct GenPlan {
    int L;
    int nr;
    int rad[kGenMaxRadices];
};

// radices of L: as many of the largest power of two up to max_radix (16, 8 or 4) as divide it,
// the rest of its power of two, then its odd prime factors in increasing order
bool make_plan(int L, GenPlan* p, int max_radix = 16) {
    if (L < 2 || L > kGenMaxL) return false;
    p->L = L;
    p->nr = 0;
    int m = L;
    auto push = [&](int r) {
        if (p->nr >= kGenMaxRadices) return false;
        p->rad[p->nr++] = r;
        m /= r;
        return true;
    };
    if (max_radix >= 16)
        while (m % 16 == 0)
            if (!push(16)) return false;
    if (max_radix >= 8)
        while (m % 8 == 0)
            if (!push(8)) return false;
    while (m % 4 == 0)
        if (!push(4)) return false;
    while (m % 2 == 0)
        if (!push(2)) return false;
    for (int f = 3; m > 1; f += 2) {
        while (m % f == 0)
            if (!push(f)) return false;
        if (f * f > m && m > 1) {
            if (!push(m)) return false;
        }
    }
    return true;
}

// Direct DFT of compile-time size R with roots W_R^j = tw[j * step] (tw: W_L^m, L = R * step).
template <int R> __device__ __forceinline__ void dft_direct(float2 (&v)[R], const float2* __restrict__ tw, int step) {
    float2 o[R];
#pragma unroll
    for (int k = 0; k < R; ++k) {
        float2 acc = v[0];
#pragma unroll
        for (int r = 1; r < R; ++r) acc = cadd(acc, cmul(v[r], tw[((r * k) % R) * step]));
        o[k] = acc;
    }
#pragma unroll
    for (int k = 0; k < R; ++k) v[k] = o[k];
}

// One Stockham pass of compile-time radix R from src to dst (LDS), Ns = product of the
// previous radices.
template <int R, int SIGN>
__device__ __forceinline__ void pass_fixed(const float2* __restrict__ src, float2* __restrict__ dst, int L, int Ns,
                                          const float2* __restrict__ tw) {
    const int nbf = L / R;
    for (int b = threadIdx.x; b < nbf; b += kGenThreads) {
        const int k = b % Ns;
        float2 v[R];
#pragma unroll
        for (int r = 0; r < R; ++r) v[r] = src[b + r * nbf];
        if (Ns > 1) {
            const int step = L / (Ns * R);
#pragma unroll
            for (int r = 1; r < R; ++r) v[r] = cmul(v[r], tw[(r * k) * step]);
        }
        if constexpr (R == 2 || R == 4 || R == 8 || R == 16)
            Dft<R, SIGN, 16>::run(v, tw);
        else
            dft_direct<R>(v, tw, L / R);
        const int base = (b / Ns) * Ns * R + k;
#pragma unroll
        for (int r = 0; r < R; ++r) dst[base + r * Ns] = v[r];
    }
}

// A pass of any radix R: work item (b, kk) forms output kk of butterfly b from the R inputs.
__device__ __forceinline__ void pass_any(const float2* __restrict__ src, float2* __restrict__ dst, int L, int Ns,
                                         int R, const float2* __restrict__ tw) {
    const int nbf = L / R;
    const int step = L / (Ns * R);  // twiddle W_{Ns R}^{r k}
    const int rstep = L / R;        // root W_R^j = W_L^{j L / R}
    for (int w = threadIdx.x; w < L; w += kGenThreads) {
        const int b = w % nbf, kk = w / nbf;
        const int k = b % Ns;
        float2 acc = src[b];
        for (int r = 1; r < R; ++r) {
            float2 v = src[b + r * nbf];
            v = cmul(v, tw[((r * k) % (L / step)) * step]);
            acc = cadd(acc, cmul(v, tw[((r * kk) % R) * rstep]));
        }
        dst[(b / Ns) * Ns * R + k + kk * Ns] = acc;
    }
}

// The plan's passes on the row in buf[0]; returns the buffer holding the result.
template <int SIGN>
__device__ float2* run_plan(const GenPlan& p, float2* b0, float2* b1, const float2* __restrict__ tw) {
    float2* src = b0;
    float2* dst = b1;
    int Ns = 1;
    for (int i = 0; i < p.nr; ++i) {
        const int R = p.rad[i];
        switch (R) {
            case 16: pass_fixed<16, SIGN>(src, dst, p.L, Ns, tw); break;
            case 8: pass_fixed<8, SIGN>(src, dst, p.L, Ns, tw); break;
            case 4: pass_fixed<4, SIGN>(src, dst, p.L, Ns, tw); break;
            case 2: pass_fixed<2, SIGN>(src, dst, p.L, Ns, tw); break;
            case 3: pass_fixed<3, SIGN>(src, dst, p.L, Ns, tw); break;
            case 5: pass_fixed<5, SIGN>(src, dst, p.L, Ns, tw); break;
            case 7: pass_fixed<7, SIGN>(src, dst, p.L, Ns, tw); break;
            default: pass_any(src, dst, p.L, Ns, R, tw); break;
        }
        __syncthreads();
        Ns *= R;
        float2* t = src;
        src = dst;
        dst = t;
    }
    return src;
}

enum GenRowMode { GR_PLAIN = 0, GR_ADJ = 1, GR_FWDACC = 2, GR_SUMPART = 3 };

struct GenRowArgs {
    GenPlan plan;
    const float2* in;     // PLAIN: [rows][L]; FWDACC: [nz][nrows][L]; SUMPART: [nparts][nrows][L]
    float2* out;          // PLAIN/SUMPART: [rows][L]; ADJ: [nz][nrows][L]; FWDACC: [chunks][nrows][L]
    const float2* tw;     // W_L^m for the transform's sign
    int nrows;            // rows per plane (ADJ, FWDACC) or in total (PLAIN, SUMPART)
    const float2* rhat;   // ADJ: [nrows][L]
    const double* kappa;  // ADJ, FWDACC: [nrows][L]
    const float* zlist;
    int nz, zchunk;
    float invN;
    int nparts;
    long part_stride;
};

constexpr int kPerThread = kGenMaxL / kGenThreads;  // elements per thread in the FWDACC accumulators

template <int SIGN, int MODE> __global__ void __launch_bounds__(kGenThreads) k_gen_rows(GenRowArgs a) {
    extern __shared__ float2 gen_lds[];
    const int L = a.plan.L;
    float2* b0 = gen_lds;
    float2* b1 = gen_lds + L;
    const int row = blockIdx.x;
    if constexpr (MODE == GR_PLAIN || MODE == GR_SUMPART) {
        const float2* src = a.in + (size_t)row * L;
        for (int j = threadIdx.x; j < L; j += kGenThreads) {
            float2 v = src[j];
            if constexpr (MODE == GR_SUMPART)
                for (int p = 1; p < a.nparts; ++p) v = cadd(v, src[(size_t)p * a.part_stride + j]);
            b0[j] = v;
        }
        __syncthreads();
        const float2* res = run_plan<SIGN>(a.plan, b0, b1, a.tw);
        float2* dst = a.out + (size_t)row * L;
        for (int j = threadIdx.x; j < L; j += kGenThreads) dst[j] = res[j];
    } else if constexpr (MODE == GR_ADJ) {
        const int z0 = blockIdx.y * a.zchunk, z1 = min(a.nz, z0 + a.zchunk);
        const float2* rh = a.rhat + (size_t)row * L;
        const double* kp = a.kappa + (size_t)row * L;
        for (int z = z0; z < z1; ++z) {
            const double zz = (double)a.zlist[z];
            for (int j = threadIdx.x; j < L; j += kGenThreads) b0[j] = cscale(cmul(transfer(zz, kp[j]), rh[j]), a.invN);
            __syncthreads();
            const float2* res = run_plan<SIGN>(a.plan, b0, b1, a.tw);
            float2* dst = a.out + ((size_t)z * a.nrows + row) * L;
            for (int j = threadIdx.x; j < L; j += kGenThreads) dst[j] = res[j];
            __syncthreads();
        }
    } else {  // GR_FWDACC
        float2 acc[kPerThread];
#pragma unroll
        for (int k = 0; k < kPerThread; ++k) acc[k] = make_float2(0.0f, 0.0f);
        const double* kp = a.kappa + (size_t)row * L;
        const int z0 = blockIdx.y * a.zchunk, z1 = min(a.nz, z0 + a.zchunk);
        for (int z = z0; z < z1; ++z) {
            const double mz = -(double)a.zlist[z];
            const float2* src = a.in + ((size_t)z * a.nrows + row) * L;
            for (int j = threadIdx.x; j < L; j += kGenThreads) b0[j] = src[j];
            __syncthreads();
            const float2* res = run_plan<SIGN>(a.plan, b0, b1, a.tw);
#pragma unroll
            for (int k = 0; k < kPerThread; ++k) {
                const int j = threadIdx.x + k * kGenThreads;
                if (j < L) acc[k] = cadd(acc[k], cmul(transfer(mz, kp[j]), res[j]));
            }
            __syncthreads();
        }
        float2* dst = a.out + (size_t)blockIdx.y * a.part_stride + (size_t)row * L;
#pragma unroll
        for (int k = 0; k < kPerThread; ++k) {
            const int j = threadIdx.x + k * kGenThreads;
            if (j < L) dst[j] = acc[k];
        }
    }
}

template <int SIGN, int MODE> hipError_t launch_rows(const GenRowArgs& a, unsigned gx, unsigned gy, hipStream_t s) {
    const size_t lds = 2 * (size_t)a.plan.L * sizeof(float2);
    if (gx == 0 || gy == 0) return hipSuccess;
    hipLaunchKernelGGL((k_gen_rows<SIGN, MODE>), dim3(gx, gy), dim3(kGenThreads), lds, s, a);
    return hipGetLastError();
}

// ----------------------------------------------------------------------------
// Tiled transposes: out[z][c][r] = in[z][r][c] (times a per-plane phase factor, conjugated or
// not), and the estimate's real part / residual / norm.
// ----------------------------------------------------------------------------
constexpr int kTile = 32;

__global__ void __launch_bounds__(256) k_gen_transpose(const float2* __restrict__ in, float2* __restrict__ out,
                                                       int R, int C, const float2* __restrict__ phase, int conj) {
    __shared__ float2 tile[kTile][kTile + 1];
    const int z = blockIdx.z;
    const int r0 = blockIdx.y * kTile, c0 = blockIdx.x * kTile;
    const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;  // 32 x 8
    const float2* src = in + (size_t)z * R * C;
    float2* dst = out + (size_t)z * R * C;
    float2 ph = make_float2(1.0f, 0.0f);
    if (phase) {
        ph = phase[z];
        if (conj) ph.y = -ph.y;
    }
    for (int k = ty; k < kTile; k += 8) {
        const int r = r0 + k, c = c0 + tx;
        if (r < R && c < C) {
            float2 v = src[(size_t)r * C + c];
            if (phase) v = make_float2(v.x * ph.x - v.y * ph.y, v.y * ph.x + v.x * ph.y);
            tile[k][tx] = v;
        }
    }
    __syncthreads();
    for (int k = ty; k < kTile; k += 8) {
        const int c = c0 + k, r = r0 + tx;
        if (r < R && c < C) dst[(size_t)c * R + r] = tile[tx][k];
    }
}

// in: [C][R] (the transposed estimate); est[r][c] = (Re in[c][r], 0), resid = est - meas,
// partial[block] = sum |resid|^2 over the tile
__global__ void __launch_bounds__(256) k_gen_transpose_real(const float2* __restrict__ in, float2* __restrict__ est,
                                                            const float2* __restrict__ meas, float2* __restrict__ resid,
                                                            double* __restrict__ partial, int R, int C) {
    __shared__ float tile[kTile][kTile + 1];
    __shared__ double red[4];
    const int r0 = blockIdx.y * kTile, c0 = blockIdx.x * kTile;
    const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;
    for (int k = ty; k < kTile; k += 8) {
        const int c = c0 + k, r = r0 + tx;
        if (r < R && c < C) tile[k][tx] = in[(size_t)c * R + r].x;
    }
    __syncthreads();
    double s = 0.0;
    for (int k = ty; k < kTile; k += 8) {
        const int r = r0 + k, c = c0 + tx;
        if (r < R && c < C) {
            const size_t o = (size_t)r * C + c;
            const float2 e = make_float2(tile[tx][k], 0.0f);  // back_propagate_finish_kernel
            est[o] = e;
            const float2 m = meas[o];
            const float2 d = make_float2(e.x - m.x, e.y - m.y);  // RM_EST_TRUE
            resid[o] = d;
            s += (double)d.x * d.x + (double)d.y * d.y;
        }
    }
    const double tot = block_sum(s, red);
    if (threadIdx.x == 0) partial[(size_t)blockIdx.y * gridDim.x + blockIdx.x] = tot;
}

dim3 tgrid(int R, int C, int nz) { return dim3((C + kTile - 1) / kTile, (R + kTile - 1) / kTile, nz); }

}  // namespace

bool gen_fft_supported(int L) {
    GenPlan p;
    return make_plan(L, &p);
}

hipError_t gen_rows_plain(int L, int sign, const float2* in, float2* out, long nrows, const float2* tw, hipStream_t s,
                          int max_radix) {
    GenRowArgs a{};
    if (!make_plan(L, &a.plan, max_radix) || nrows > 0x7fffffffL) return hipErrorInvalidValue;
    a.in = in; a.out = out; a.tw = tw; a.nrows = (int)nrows;
    return sign < 0 ? launch_rows<-1, GR_PLAIN>(a, (unsigned)nrows, 1, s) : launch_rows<1, GR_PLAIN>(a, (unsigned)nrows, 1, s);
}

hipError_t gen_rows_sumparts(int L, const float2* parts, int nparts, long part_stride, float2* out, int nrows,
                             const float2* tw_inv, hipStream_t s, int max_radix) {
    GenRowArgs a{};
    if (!make_plan(L, &a.plan, max_radix)) return hipErrorInvalidValue;
    a.in = parts; a.out = out; a.tw = tw_inv; a.nrows = nrows; a.nparts = nparts; a.part_stride = part_stride;
    return launch_rows<1, GR_SUMPART>(a, (unsigned)nrows, 1, s);
}

hipError_t gen_rows_adj(int L, const float2* rhat, const double* kappa, const float* zlist, float2* out, int nrows,
                        int nz, int zchunk, float invN, const float2* tw_inv, hipStream_t s, int max_radix) {
    GenRowArgs a{};
    if (!make_plan(L, &a.plan, max_radix) || zchunk < 1) return hipErrorInvalidValue;
    a.out = out; a.tw = tw_inv; a.nrows = nrows; a.rhat = rhat; a.kappa = kappa; a.zlist = zlist; a.nz = nz;
    a.zchunk = zchunk; a.invN = invN;
    return launch_rows<1, GR_ADJ>(a, (unsigned)nrows, (unsigned)((nz + zchunk - 1) / zchunk), s);
}

hipError_t gen_rows_fwdacc(int L, const float2* in, const double* kappa, const float* zlist, float2* parts, int nrows,
                           int nz, int zchunk, const float2* tw_fwd, hipStream_t s, int max_radix) {
    GenRowArgs a{};
    if (!make_plan(L, &a.plan, max_radix) || zchunk < 1) return hipErrorInvalidValue;
    a.in = in; a.out = parts; a.tw = tw_fwd; a.nrows = nrows; a.kappa = kappa; a.zlist = zlist; a.nz = nz;
    a.zchunk = zchunk; a.part_stride = (long)nrows * L;
    return launch_rows<-1, GR_FWDACC>(a, (unsigned)nrows, (unsigned)((nz + zchunk - 1) / zchunk), s);
}

hipError_t gen_transpose(const float2* in, float2* out, int R, int C, int nz, const float2* phase, bool conj,
                         hipStream_t s) {
    if (nz > 65535 || (R + kTile - 1) / kTile > 65535) return hipErrorInvalidValue;
    hipLaunchKernelGGL(k_gen_transpose, tgrid(R, C, nz), dim3(256), 0, s, in, out, R, C, phase, conj ? 1 : 0);
    return hipGetLastError();
}

int gen_transpose_real_blocks(int R, int C) { return ((R + kTile - 1) / kTile) * ((C + kTile - 1) / kTile); }

hipError_t gen_transpose_real(const float2* in, float2* est, const float2* meas, float2* resid, double* partial, int R,
                              int C, hipStream_t s) {
    hipLaunchKernelGGL(k_gen_transpose_real, tgrid(R, C, 1), dim3(256), 0, s, in, est, meas, resid, partial, R, C);
    return hipGetLastError();
}

}  // namespace rihvr
