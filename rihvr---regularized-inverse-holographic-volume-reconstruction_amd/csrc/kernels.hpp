// kernels.hpp -- host-side launchers of the gfx950 kernels (one stream each).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace rihvr {

// Line occupancy of a prox-output volume: one 64-bit word per (plane, row), bit l set when
// the row's voxels l L .. l L + L - 1 hold a non-zero value, L = 16 voxels (one 128-byte line)
// for W <= 1024, else the least power of two giving at most 64 lines (32 up to 2048, 64 up to
// 4096; only the fixed-length FFT kernels, W <= 1024, read the maps besides the prox).  A
// volume with a valid map holds exact zeros in every line whose bit is clear, so readers
// may skip those lines and writers may skip zero lines that were already zero.  A map is
// [nz][H] row words followed by [nz] plane words, each the OR of its plane's row words.

// ---- FFT (fft_kernels.hip) -------------------------------------------------
bool fft_length_supported(int L);

// Plain row FFT of nrows contiguous rows of length L.  tw: W_L^m for the sign.
hipError_t fft_rows_plain(int L, int sign, const float2* in, float2* out, int nrows,
                          const float2* tw, hipStream_t s);
// Inverse row FFT of sum_p parts[p] (nparts partial spectra, part_stride elements apart).
hipError_t fft_rows_sumparts(int L, const float2* parts, int nparts, long part_stride, float2* out,
                             int nrows, const float2* tw_inv, hipStream_t s);
// A, first pass: forward column FFT (length L = H) of vol[z] * conj(phase[z]), nz planes of
// width W (phase[z] = e^{i phi_z} / N for A).
// occ (may be NULL): vol's line occupancy, [nz][L] row words, and occ_planes its [nz] plane
// words; lines with a clear bit are not read, and a 16-column strip with no set bit in its
// plane word is neither transformed nor written (its workgroup leaves at once).
// smap[z][W/16] (always written) = 1 for the strips written to out.
hipError_t fft_cols_fwd(int L, const float2* vol, const float2* phase, float2* out, int W, int nz,
                        const float2* tw_fwd, const uint64_t* occ, const uint64_t* occ_planes, uint8_t* smap,
                        hipStream_t s);
// A, second pass: forward row FFT (length L = W), epilogue H_{-z} summed over zchunk planes
// into parts[chunk] ([H][W] each); the 1/N of A is applied by the first pass's phase
// factors.  Strips with smap 0 are read as zeros, planes without any strip are skipped.
hipError_t fft_rows_fwdacc(int L, const float2* in, const double* kappa, const float* zlist, float2* parts,
                          int H, int nz, int zchunk, const float2* tw_fwd, const uint8_t* smap, hipStream_t s);
// A^T, first pass: inverse row FFT of H_z * rhat / N, planes processed zchunk per workgroup.
hipError_t fft_rows_adj(int L, const float2* rhat, const double* kappa, const float* zlist, float2* out,
                        int H, int nz, int zchunk, float invN, const float2* tw_inv, hipStream_t s);
// Plain column FFT (length L = H) of nplanes planes of width W.
hipError_t fft_cols_plain(int L, int sign, const float2* in, float2* out, int W, int nplanes,
                          const float2* tw, hipStream_t s);
// A^T, second pass: inverse column FFT, epilogue * e^{+i phi_z}.
hipError_t fft_cols_adj(int L, const float2* in, const float2* phase, float2* out, int W, int nz,
                        const float2* tw_inv, hipStream_t s);
// Final inverse column FFT of one plane: est = (Re, 0), resid = est - meas,
// partial[blockIdx.x] = sum |resid|^2.
hipError_t fft_cols_real(int L, const float2* in, float2* est, const float2* meas, float2* resid,
                         double* partial, int W, const float2* tw_inv, hipStream_t s);

// ---- FFTs of any length 2..4096 (fft_generic.hip) -----------------------------------
// Row transforms of length L with a plan chosen at run time; transforms along y run between
// tiled transposes ([z][R][C] -> [z][C][R]).  Same roles as the fixed-length launchers above.
bool gen_fft_supported(int L);
// max_radix: the largest radix of the plan (16, 8 or 4; an autotuner choice)
hipError_t gen_rows_plain(int L, int sign, const float2* in, float2* out, long nrows, const float2* tw, hipStream_t s,
                          int max_radix = 16);
hipError_t gen_rows_sumparts(int L, const float2* parts, int nparts, long part_stride, float2* out, int nrows,
                             const float2* tw_inv, hipStream_t s, int max_radix = 16);
hipError_t gen_rows_adj(int L, const float2* rhat, const double* kappa, const float* zlist, float2* out, int nrows,
                        int nz, int zchunk, float invN, const float2* tw_inv, hipStream_t s, int max_radix = 16);
hipError_t gen_rows_fwdacc(int L, const float2* in, const double* kappa, const float* zlist, float2* parts, int nrows,
                           int nz, int zchunk, const float2* tw_fwd, hipStream_t s, int max_radix = 16);
// out[z][c][r] = in[z][r][c] (times phase[z], conjugated with conj; phase may be NULL)
hipError_t gen_transpose(const float2* in, float2* out, int R, int C, int nz, const float2* phase, bool conj,
                         hipStream_t s);
// in [C][R] (a transposed estimate): est = (Re, 0) [R][C], resid = est - meas,
// partial[block] = sum |resid|^2 over gen_transpose_real_blocks(R, C) blocks
int gen_transpose_real_blocks(int R, int C);
hipError_t gen_transpose_real(const float2* in, float2* est, const float2* meas, float2* resid, double* partial, int R,
                              int C, hipStream_t s);

// ---- prox and elementwise (prox_kernels.hip) --------------------------------
// Reduction slots written by the prox kernels (per wave partials, summed by reduce_partials).
enum ProxSlot {
    PS_LS1 = 0,      // sum Re<y - x0, g>             (sparse_linesearch_1_kernel)
    PS_LS2 = 1,      // sum |y - x0|^2                (sparse_linesearch_2_kernel)
    PS_RESTART = 2,  // quirky restart sum            (sparse_fistaRestart_kernel)
    PS_L1X0 = 3,     // sum |x0| over x0 != 0         (calcL1NormComplex of the momentum point)
    PS_TVX0 = 4,     // sum 2D TV of x0               (calcTVNorm2d of the momentum point)
    PS_NNZX0 = 5,    // count of x0 != 0
    PS_NNZY = 6,     // count of y != 0
    PS_COUNT = 8
};

struct ProxLaunch {
    const float2* Y0;   // y_{k-1}
    const float2* YM;   // y_{k-2}
    const float2* G;    // gradient A^T r
    float2* Y;          // output y_k
    int W, H, nz;
    float s_mom;        // x0 = Y0 + s_mom * (Y0 - YM)
    const float* s_mom_dev = nullptr;  // when set, s_mom is read from device memory
    double t;           // step size
    double reg;         // L1 weight
    double reg_tv;      // TV weight
    int n_tv;
    int row_chunk;      // output rows per wave of the mode-7 prox (<= 0: default)
    double* partial;    // [nwaves][PS_COUNT]
    // mode 7 only, all required ([nz][H] words, see the line occupancy above): the maps of
    // Y0 and YM (their clear lines are not read), of Y's current content (its zero lines
    // with a clear bit are not rewritten) and the map of the new Y (OR-ed in; must start
    // cleared).  A map with every bit set stands for a volume without one.  The output's
    // zeros are +0.
    const uint64_t* occ_y0 = nullptr;
    const uint64_t* occ_ym = nullptr;
    const uint64_t* occ_old = nullptr;
    uint64_t* occ_new = nullptr;
    uint64_t* occ_new_planes = nullptr;  // [nz] plane words of the new map (OR-ed in)
};
// Number of per-wave partial rows the mode-7 prox writes for this geometry.
// row_chunk: output rows per wave (<= 0: default 128)
int prox_fl2d_num_waves(int W, int H, int nz, int n_tv, int row_chunk);
// Output rows per wave for this geometry on the current device: about 205 when that gives
// at least two rounds of resident waves, else the chunk length whose rounds of resident
// waves (rounded up) times the rows a wave steps through (its chunk plus the 2 n_tv + 1
// halo rows) is least.  0 if the occupancy query fails.
int prox_fl2d_auto_rows(int W, int H, int nz, int n_tv);
bool prox_fl2d_supported(int n_tv);
hipError_t prox_fl2d(const ProxLaunch& a, hipStream_t s);

int prox_l1_num_blocks(int W, int H, int nz);
hipError_t prox_l1(const ProxLaunch& a, hipStream_t s);

// Mode 6 (3D TV, z-periodic): three kernels, partials [nblocks][PS_COUNT];
// scratch is a volume-sized buffer distinct from a.Y.
int prox_fl3d_num_blocks(int W, int H, int nz);
// Mode 7 for any n_tv (multi-pass, the duals in memory): work holds 4 volumes
// (x-hat, y-hat, g_x, g_y); partials [nblocks][PS_COUNT].
int prox_tv2_num_blocks(int W, int H, int nz);
hipError_t prox_tv2_general(const ProxLaunch& a, float2* work, hipStream_t s);
hipError_t prox_fl3d(const ProxLaunch& a, float2* scratch, hipStream_t s);

// out = Y + s (Y - Y0) (the returned momentum point) with partials
// [nblocks][PS_COUNT]: L1, TV (2D per plane, or 3D z-periodic with tv3d) and nnz
// of out in slots PS_L1X0, PS_TVX0, PS_NNZX0.
int momentum_out_num_blocks(int W, int H, int nz);
hipError_t momentum_out(const float2* Y, const float2* Y0, float s_mom, float2* out, int W, int H,
                        int nz, bool tv3d, double* partial, hipStream_t s);
// momentum_out(X, X, 0, NULL, ...) for a volume with a valid line-occupancy map occ ([nz][H]
// row words; a clear line holds +0): the same partials, bit for bit, reading only the lines
// whose terms can be non-zero.
hipError_t volume_terms_map(const float2* X, const uint64_t* occ, int W, int H, int nz, bool tv3d,
                            double* partial, hipStream_t s);

// Hologram momentum by linearity (sparse_compressive_holo.cu:494-498):
// est = cur + s (cur - prev); resid = est - meas; partial[block] = sum |resid|^2.
// restart_sum (device, may be NULL): when > 0 the step is 0 (a restarted iteration).
int holo_num_blocks(long n);
// step_dev (device, may be NULL): when set, the step is read from it instead of s_mom.
hipError_t holo_momentum(const float2* cur, const float2* prev, double s_mom, const float2* meas,
                         float2* est, float2* resid, long n, double* partial, hipStream_t s,
                         const double* restart_sum = nullptr, const double* step_dev = nullptr);
// p[0..n) = v (a kernel, so that it can be captured in a graph)
hipError_t fill_u64(uint64_t* p, long n, uint64_t v, hipStream_t s);
// partial[b] (b < nblocks) = max |v_i| over block b's share of n values
hipError_t maxabs_partials(const float2* v, long n, int nblocks, double* partial, hipStream_t s);
// the complex soft threshold at two thresholds in one pass, out_k = in * max(0, 1 - tau_k / |in|)
// (+0 where |in| <= tau_k), each with its line-occupancy map ([nz][H] row words, written; then
// nz plane words, cleared by the caller) set for its non-zero lines
hipError_t soft_threshold_map2(const float2* in, int W, int H, int nz, float tau0, float2* out0, uint64_t* occ0,
                               float tau1, float2* out1, uint64_t* occ1, hipStream_t s);
// partial[b] (b < nblocks) = part of sum |v_i|^2 over n values
hipError_t norm2_partials(const float2* v, long n, int nblocks, double* partial, hipStream_t s);
// a sparse pseudo-random volume (about 1 / period of its voxels non-zero)
hipError_t synth_volume(float2* v, long n, unsigned period, unsigned seed, hipStream_t s);
// *f = fv and *d = dv (either may be NULL), in stream order
hipError_t set_scalars(float* f, float fv, double* d, double dv, hipStream_t s);
// resid = est - meas with partial sum |resid|^2
hipError_t holo_residual(const float2* est, const float2* meas, float2* resid, long n, double* partial,
                         hipStream_t s);
// Spectrum of A's real estimate from nparts partial spectra [nparts][H][W] (stride apart):
// est = (S(k) + conj S(-k)) / 2 with S = sum of the parts; partial[block] = sum |est - bspec|^2
// over holo_spec_num_blocks(W, H) blocks (W and H even)
int holo_spec_num_blocks(int W, int H);
hipError_t holo_spec(const float2* parts, int nparts, long stride, int W, int H, const float2* bspec, float2* est,
                     double* partial, hipStream_t s);
// out = f * in
hipError_t holo_scale(const float2* in, float2* out, long n, float f, hipStream_t s);

// sums rows of partial [nrows][ncols] into out[ncols] (fixed order: deterministic);
// work holds kReduceBlocks * ncols doubles
constexpr int kReduceBlocks = 64;
hipError_t reduce_partials(const double* partial, int nrows, int ncols, double* work, double* out, hipStream_t s);

// flags[0] = 1 if any element of a [n] complex array is not finite
hipError_t check_finite(const float2* a, long n, int* flag, hipStream_t s);

}  // namespace rihvr
