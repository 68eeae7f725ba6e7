// solver.hip -- host side of librihvr: the context, the FASTA/FISTA control
// loop of SparseCompressiveHolo::inverseReconstruct (sparse_compressive_holo.cu:173-633)
// and the C ABI of include/rihvr.h.
//
// Data stay on the GPU for the whole solve.  Per iteration the host reads back
// one small vector of reduced scalars (line-search terms, restart sum, residual
// norm, objective terms) and takes the reference's control decisions
// (backtrack, restart, momentum step) on it.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "../../include/rihvr.h"
#include "kernels.hpp"
#include "warmstart.hpp"

using namespace rihvr;

namespace {

constexpr double kPi = 3.14159265358979323846;

// scalar slots after reduction
enum Slot { SL_PROX = 0, SL_RESID_A = 8, SL_RESID_M = 9, SL_COUNT = 16 };

#define RIHVR_HIP(call)                                                   \
    do {                                                                  \
        hipError_t e_ = (call);                                           \
        if (e_ != hipSuccess) {                                           \
            std::fprintf(stderr, "rihvr: HIP error %s at %s:%d\n",       \
                         hipGetErrorString(e_), __FILE__, __LINE__);      \
            return e_ == hipErrorOutOfMemory ? RIHVR_E_OUT_OF_MEMORY      \
                                             : RIHVR_E_DEVICE;            \
        }                                                                 \
    } while (0)

#define RIHVR_TRY(expr)                       \
    do {                                      \
        rihvr_status st_ = (expr);            \
        if (st_ != RIHVR_OK) return st_;      \
    } while (0)

// Kernels timed with HIP events when timing is enabled (the reference's
// DECLARE_TIMING / SAVE_TIMING phases, src/core/umnholo.h:79-124).
enum KernelId {
    K_COLS_FWD = 0, K_ROWS_FWDACC, K_ROWS_SUMPARTS, K_COLS_REAL, K_PROX, K_HOLO_MOM, K_ROWS_PLAIN,
    K_COLS_PLAIN, K_ROWS_ADJ, K_COLS_ADJ, K_REDUCE, K_MOM_OUT, K_HOLO_SPEC, K_COUNT
};
const char* const kKernelNames[K_COUNT] = {
    "fft_cols_fwd", "fft_rows_fwdacc", "fft_rows_sumparts", "fft_cols_real", "prox", "holo_momentum",
    "fft_rows_plain", "fft_cols_plain", "fft_rows_adj", "fft_cols_adj", "reduce_partials", "momentum_out",
    "holo_spectrum"};

struct Timer {
    bool on = false;
    int only = -1;  // >= 0: record this kernel id only
    struct Pending { int id; hipEvent_t a, b; };
    std::vector<hipEvent_t> pool;
    std::vector<Pending> pending;
    double total_ms[K_COUNT] = {};
    int64_t launches[K_COUNT] = {};
};

// state carried between rihvr_begin / rihvr_iterate / rihvr_finish
struct SolveState {
    bool active = false;
    int mode = RIHVR_MODE_FASTA_FUSED_LASSO_2D;
    double t = 0;             // step size
    double alpha1 = 1.0;      // FISTA alpha
    float s_mom = 0.0f;       // momentum step of the last update
    double step = 0.0;        // the same step in double, as the hologram momentum used it
    double prev_summag = 0;   // sum |r|^2 at the current momentum point
    double f_initial = 0;
    int total_bt = 0;
    int iters = 0;
    bool g_ready = false;     // G = A^T r of the current residual (else only R-hat is)
};

}  // namespace

struct rihvr_ctx {
    rihvr_params p{};
    int W = 0, H = 0, Nz = 0;
    long N = 0, V = 0;
    int device = 0;
    hipStream_t stream = nullptr;  // every launch and copy, in order
    hipStream_t aux = nullptr;     // A^T of the next plane group while prox and A run on `stream`
    std::vector<hipEvent_t> ev_adj;  // per plane group: its G planes are written (on aux)
    hipEvent_t ev_rhat = nullptr;    // R-hat is written (on stream)
    hipEvent_t ev_scal = nullptr;    // the scalars of the last prox + A are in h_scal
    // volumes
    float2* Yb[3] = {nullptr, nullptr, nullptr};  // ring of prox outputs
    int iy = 0, iy0 = 1, iym = 2;                 // y_k, y_{k-1}, y_{k-2}
    // line occupancy of each ring volume ([Nz][H] words, kernels.hpp); a volume written by
    // anything but the mode-7 prox has no valid map (every line is then read)
    uint64_t* occ[3] = {nullptr, nullptr, nullptr};
    bool occ_valid[3] = {false, false, false};
    uint64_t* occ_tmp = nullptr;                  // the map being built by a prox
    uint64_t* occ_ones = nullptr;                 // every line set: stands for "no map"
    uint8_t* smap = nullptr;                      // [Nz][W/16] strips of A's column pass
    float2* G = nullptr;                          // A^T r
    float2* G_alt = nullptr;                      // the next iteration's A^T r, formed ahead
    float2* T = nullptr;                          // FFT scratch volume
    // FFTs of lengths outside the fixed-length kernels' set: rows with run-time plans between
    // transposes (fft_generic.hip); T2 holds the transposed planes
    bool gen = false;
    int gen_radix = 16;  // the run-time plans' largest radix
    float2* T2 = nullptr;
    float2* tvbuf = nullptr;                      // 4 volumes for mode 7 with n_tv above the fused prox's limit
    float2* out = nullptr;                        // points at the returned volume
    // holograms
    // Hologram-domain quantities of the iteration are kept as spectra FFT(.) / N (the
    // estimates are real, so their spectra are Hermitian-symmetric): hE, the estimate at the
    // momentum point; hC/hP, A y_k and A y_{k-1}; hX, A of the prox output being tested
    // (becomes hC once the iteration is accepted); bspec, the measurement's; rhat then holds
    // the residual's spectrum hE - bspec (rhat_scaled).  meas is the measurement itself;
    // hspat and resid receive the spatial estimate and residual when they are asked for.
    float2 *meas = nullptr, *hE = nullptr, *hC = nullptr, *hP = nullptr, *hX = nullptr, *resid = nullptr,
           *rhat = nullptr, *htmp = nullptr, *spart = nullptr, *bspec = nullptr, *hspat = nullptr;
    bool rhat_scaled = false;  // rhat holds FFT(r) / N (else FFT(r))
    // tables
    double* kappa = nullptr;
    float* zlist = nullptr;
    float2* phase = nullptr;
    float2* phase_fwd = nullptr;  // phase / N: A's first pass applies A's 1/N with its phase
    float2 *twfW = nullptr, *twiW = nullptr, *twfH = nullptr, *twiH = nullptr;
    // reductions
    double* partial = nullptr;
    double* partial_holo = nullptr;
    long partial_rows = 0, holo_rows = 0;
    double* scal = nullptr;
    double* red_work = nullptr;  // first-pass sums of reduce_partials
    double* h_scal = nullptr;
    int* dflag = nullptr;
    int zchunk_adj = 8, zchunk_fwd = 8, nparts = 1;
    int pipe_planes = 1;  // planes per pipeline group (a multiple of the z chunks)
    int n_groups = 1;
    bool groups_one_stream = false;  // run A^T, prox and A group by group on the main stream
    int prox_rows = 0;    // output rows per prox wave (0: default)
    // Captured iterations: the two halves of a queued-ahead iteration (prox + A, then the
    // hologram momentum + the next A^T) as HIP graphs, one per pointer and step combination
    // (the volume ring, the maps and the spectra rotate through at most 12), replayed with the
    // momentum step read from device memory
    bool graphs = true;
    float* d_smom = nullptr;
    double* d_step = nullptr;
    struct CachedGraph {
        std::vector<const void*> key;
        double t;
        hipGraphExec_t exec;
    };
    std::vector<CachedGraph> gcache[2];
    bool have_result = false;
    bool force_step = false;
    double forced_step = 0;
    // testing knob (RIHVR_TEST_BACKTRACKS, a bit mask of iteration numbers 0..63): the
    // first attempt of those iterations is rejected as if its ratio were above 1, so the
    // backtracking paths can be exercised; never set in normal use
    uint64_t test_backtracks = 0;
    Timer tm;
    SolveState ss;
};

namespace {

rihvr_status validate(const rihvr_params* p) {
    if (!p) return RIHVR_E_INVALID_ARGUMENT;
    if (p->width <= 0 || p->height <= 0 || p->num_planes <= 0) return RIHVR_E_INVALID_ARGUMENT;
    // the reference launches W/16 x H/16 blocks of 16x16 (sparse_compressive_holo.cu:125-126)
    if (p->width % 16 || p->height % 16) return RIHVR_E_INVALID_ARGUMENT;
    if (!(p->wavelength > 0) || !(p->resolution > 0)) return RIHVR_E_INVALID_ARGUMENT;
    if (p->num_inverse_iterations < 0 || p->num_tv_iterations < 0) return RIHVR_E_INVALID_ARGUMENT;
    // inverseReconstruct accepts modes 5, 6, 7 and throws HOLO_ERROR_INVALID_ARGUMENT
    // for any other (sparse_compressive_holo.cu:178-189)
    if (p->mode != RIHVR_MODE_FASTA_L1 && p->mode != RIHVR_MODE_FASTA_FUSED_LASSO &&
        p->mode != RIHVR_MODE_FASTA_FUSED_LASSO_2D)
        return RIHVR_E_INVALID_ARGUMENT;
    // any length the fixed-length FFT kernels or the run-time plans cover (up to 4096)
    if (!(fft_length_supported(p->width) && fft_length_supported(p->height)) &&
        !(gen_fft_supported(p->width) && gen_fft_supported(p->height)))
        return RIHVR_E_INVALID_ARGUMENT;
    return RIHVR_OK;
}

template <class T> rihvr_status dalloc(T** ptr, size_t count) {
    RIHVR_HIP(hipMalloc((void**)ptr, std::max<size_t>(1, count) * sizeof(T)));
    return RIHVR_OK;
}

std::vector<float2> twiddles(int L, int sign) {
    std::vector<float2> t(L);
    for (int m = 0; m < L; ++m) {
        double a = sign * 2.0 * kPi * (double)m / (double)L;
        t[m] = make_float2((float)std::cos(a), (float)std::sin(a));
    }
    return t;
}

rihvr_status upload(void* dst, const void* src, size_t bytes, hipStream_t s) {
    RIHVR_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, s));
    RIHVR_HIP(hipStreamSynchronize(s));
    return RIHVR_OK;
}

// copy count complex64 values between caller memory (loc) and device memory
rihvr_status copy_in(rihvr_ctx* c, float2* dst, const float* src, int loc, size_t count) {
    if (!src) return RIHVR_E_INVALID_ARGUMENT;
    RIHVR_HIP(hipMemcpyAsync(dst, src, count * sizeof(float2),
                             loc == RIHVR_DEVICE ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice, c->stream));
    return RIHVR_OK;
}
rihvr_status copy_out(rihvr_ctx* c, float* dst, const float2* src, int loc, size_t count) {
    if (!dst) return RIHVR_E_INVALID_ARGUMENT;
    RIHVR_HIP(hipMemcpyAsync(dst, src, count * sizeof(float2),
                             loc == RIHVR_DEVICE ? hipMemcpyDeviceToDevice : hipMemcpyDeviceToHost, c->stream));
    RIHVR_HIP(hipStreamSynchronize(c->stream));
    return RIHVR_OK;
}

// Launches through `launch` (returns hipError_t) on stream `st`, bracketed by events
// on that stream when timing is on.
template <class F> rihvr_status timed_on(rihvr_ctx* c, int id, hipStream_t st, F&& launch) {
    if (!c->tm.on || (c->tm.only >= 0 && id != c->tm.only)) {
        RIHVR_HIP(launch());
        return RIHVR_OK;
    }
    hipEvent_t ev[2];
    for (hipEvent_t& e : ev) {
        if (c->tm.pool.empty()) {
            RIHVR_HIP(hipEventCreate(&e));
        } else {
            e = c->tm.pool.back();
            c->tm.pool.pop_back();
        }
    }
    RIHVR_HIP(hipEventRecord(ev[0], st));
    RIHVR_HIP(launch());
    RIHVR_HIP(hipEventRecord(ev[1], st));
    c->tm.pending.push_back({id, ev[0], ev[1]});
    return RIHVR_OK;
}
template <class F> rihvr_status timed(rihvr_ctx* c, int id, F&& launch) {
    return timed_on(c, id, c->stream, std::forward<F>(launch));
}

// Accumulates the durations of completed event pairs (call after a stream sync).
// Pairs whose end event has not completed yet (work queued after a partial wait) stay
// pending; events complete in stream order.
rihvr_status resolve_timing(rihvr_ctx* c) {
    size_t k = 0;
    for (; k < c->tm.pending.size(); ++k) {
        auto& p = c->tm.pending[k];
        const hipError_t q = hipEventQuery(p.b);
        if (q == hipErrorNotReady) break;
        RIHVR_HIP(q);
        float ms = 0;
        RIHVR_HIP(hipEventElapsedTime(&ms, p.a, p.b));
        c->tm.total_ms[p.id] += ms;
        c->tm.launches[p.id] += 1;
        c->tm.pool.push_back(p.a);
        c->tm.pool.push_back(p.b);
    }
    c->tm.pending.erase(c->tm.pending.begin(), c->tm.pending.begin() + k);
    return RIHVR_OK;
}

// the residual norms are reduced over spectra FFT(r) / N: |r|^2 = N sum |FFT(r) / N|^2
void scale_resid_norms(rihvr_ctx* c) {
    c->h_scal[SL_RESID_A] *= (double)c->N;
    c->h_scal[SL_RESID_M] *= (double)c->N;
}

rihvr_status fetch_scalars(rihvr_ctx* c) {
    RIHVR_HIP(hipMemcpyAsync(c->h_scal, c->scal, SL_COUNT * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    RIHVR_HIP(hipStreamSynchronize(c->stream));
    scale_resid_norms(c);
    return resolve_timing(c);
}

// The same copy, waited for by an event so that work queued after it keeps the device busy.
rihvr_status fetch_scalars_async(rihvr_ctx* c) {
    RIHVR_HIP(hipMemcpyAsync(c->h_scal, c->scal, SL_COUNT * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    RIHVR_HIP(hipEventRecord(c->ev_scal, c->stream));
    return RIHVR_OK;
}
rihvr_status wait_scalars(rihvr_ctx* c) {
    RIHVR_HIP(hipEventSynchronize(c->ev_scal));
    scale_resid_norms(c);
    return resolve_timing(c);
}

// ring index of a volume pointer, or -1
int ring_index(const rihvr_ctx* c, const float2* v) {
    for (int k = 0; k < 3; ++k)
        if (c->Yb[k] == v) return k;
    return -1;
}
// the valid line-occupancy map of a ring volume, else NULL
// words of a line-occupancy map: [Nz][H] row words, then [Nz] plane words
size_t map_words(const rihvr_ctx* c) { return (size_t)c->Nz * c->H + c->Nz; }
const uint64_t* map_planes(const rihvr_ctx* c, const uint64_t* map) { return map + (size_t)c->Nz * c->H; }

const uint64_t* occ_of(const rihvr_ctx* c, const float2* v) {
    const int k = ring_index(c, v);
    return (k >= 0 && c->occ_valid[k]) ? c->occ[k] : nullptr;
}
void invalidate_occ(rihvr_ctx* c, const float2* v) {
    const int k = ring_index(c, v);
    if (k >= 0) c->occ_valid[k] = false;
}

// A for planes [g0, g0 + gz) on stream st: the column pass of vol * e^{-i phi} into the
// scratch planes T[g0..], then the row pass with the H_{-z}/N z-sum into the partial
// spectra of those planes' z chunks (backPropagate :1935-2077)
// occ: vol's line occupancy map (NULL: read every line); the strips the column pass
// writes are recorded in smap for the row pass
rihvr_status forward_group(rihvr_ctx* c, const float2* vol, int g0, int gz, hipStream_t st, const uint64_t* occ) {
    float2* Tg = c->T + (size_t)g0 * c->N;
    float2* spg = c->spart + (size_t)(g0 / c->zchunk_fwd) * c->N;
    if (c->gen) {
        // columns as rows of the transposed planes ([z][W][H]); every line is read
        float2* Tt = c->T2 + (size_t)g0 * c->N;
        RIHVR_TRY(timed_on(c, K_COLS_FWD, st, [&] {
            hipError_t e = gen_transpose(vol + (size_t)g0 * c->N, Tt, c->H, c->W, gz, c->phase_fwd + g0, true, st);
            if (e == hipSuccess) e = gen_rows_plain(c->H, -1, Tt, Tt, (long)gz * c->W, c->twfH, st, c->gen_radix);
            if (e == hipSuccess) e = gen_transpose(Tt, Tg, c->W, c->H, gz, nullptr, false, st);
            return e;
        }));
        return timed_on(c, K_ROWS_FWDACC, st, [&] {
            return gen_rows_fwdacc(c->W, Tg, c->kappa, c->zlist + g0, spg, c->H, gz, c->zchunk_fwd, c->twfW, st, c->gen_radix);
        });
    }
    uint8_t* smg = c->smap + (size_t)g0 * (c->W / 16);
    const uint64_t* og = occ ? occ + (size_t)g0 * c->H : nullptr;
    const uint64_t* opg = occ ? map_planes(c, occ) + g0 : nullptr;
    RIHVR_TRY(timed_on(c, K_COLS_FWD, st, [&] {
        return fft_cols_fwd(c->H, vol + (size_t)g0 * c->N, c->phase_fwd + g0, Tg, c->W, gz, c->twfH, og, opg, smg, st);
    }));
    RIHVR_TRY(timed_on(c, K_ROWS_FWDACC, st, [&] {
        return fft_rows_fwdacc(c->W, Tg, c->kappa, c->zlist + g0, spg, c->H, gz, c->zchunk_fwd, c->twfW, smg,
                               st);
    }));
    return RIHVR_OK;
}

// The end of A (calcResidual :2676-2723 + calcSumMagnitude): the estimate is the real part
// of the inverse transform of the summed partial spectra; it is kept as its spectrum
// (FFT(est) / N) in est, and scal[SL_RESID_A] = sum |est - bspec|^2 = |A x - b|^2 / N
// (Parseval; fetch_scalars scales it)
rihvr_status forward_finish(rihvr_ctx* c, float2* est) {
    RIHVR_TRY(timed(c, K_HOLO_SPEC, [&] {
        return holo_spec(c->spart, c->nparts, c->N, c->W, c->H, c->bspec, est, c->partial_holo, c->stream);
    }));
    RIHVR_TRY(timed(c, K_REDUCE, [&] {
        return reduce_partials(c->partial_holo, holo_spec_num_blocks(c->W, c->H), 1, c->red_work, c->scal + SL_RESID_A,
                               c->stream);
    }));
    return RIHVR_OK;
}

// The spatial estimate and residual of a spectrum est_spec (FFT(est) / N): est = Re(IFFT(.)),
// into hspat, and resid = est - meas (back_propagate_finish_kernel, RM_EST_TRUE)
rihvr_status spatial_estimate(rihvr_ctx* c, const float2* est_spec) {
    if (c->gen) {
        RIHVR_TRY(timed(c, K_ROWS_SUMPARTS, [&] {
            return gen_rows_sumparts(c->W, est_spec, 1, c->N, c->htmp, c->H, c->twiW, c->stream, c->gen_radix);
        }));
        return timed(c, K_COLS_REAL, [&] {
            hipError_t e = gen_transpose(c->htmp, c->T2, c->H, c->W, 1, nullptr, false, c->stream);
            if (e == hipSuccess) e = gen_rows_plain(c->H, 1, c->T2, c->T2, c->W, c->twiH, c->stream, c->gen_radix);
            if (e == hipSuccess)
                e = gen_transpose_real(c->T2, c->hspat, c->meas, c->resid, c->partial_holo, c->H, c->W, c->stream);
            return e;
        });
    }
    RIHVR_TRY(timed(c, K_ROWS_SUMPARTS, [&] { return fft_rows_sumparts(c->W, est_spec, 1, c->N, c->htmp, c->H, c->twiW, c->stream); }));
    RIHVR_TRY(timed(c, K_COLS_REAL, [&] { return fft_cols_real(c->H, c->htmp, c->hspat, c->meas, c->resid, c->partial_holo, c->W, c->twiH, c->stream); }));
    return RIHVR_OK;
}

rihvr_status apply_forward(rihvr_ctx* c, const float2* vol, float2* est) {
    const uint64_t* occ = occ_of(c, vol);
    for (int g0 = 0; g0 < c->Nz; g0 += c->pipe_planes)
        RIHVR_TRY(forward_group(c, vol, g0, std::min(c->pipe_planes, c->Nz - g0), c->stream, occ));
    return forward_finish(c, est);
}

// R-hat = FFT(r) of a spatial residual (SparseGradient::reconstruct, sparse_gradient.cu:47-75)
// into out; ev_rhat marks rhat
rihvr_status spectrum(rihvr_ctx* c, const float2* r, float2* out) {
    if (c->gen) {
        RIHVR_TRY(timed(c, K_ROWS_PLAIN, [&] { return gen_rows_plain(c->W, -1, r, c->htmp, c->H, c->twfW, c->stream, c->gen_radix); }));
        RIHVR_TRY(timed(c, K_COLS_PLAIN, [&] {
            hipError_t e = gen_transpose(c->htmp, c->T2, c->H, c->W, 1, nullptr, false, c->stream);
            if (e == hipSuccess) e = gen_rows_plain(c->H, -1, c->T2, c->T2, c->W, c->twfH, c->stream, c->gen_radix);
            if (e == hipSuccess) e = gen_transpose(c->T2, out, c->W, c->H, 1, nullptr, false, c->stream);
            return e;
        }));
        if (out == c->rhat) {
            c->rhat_scaled = false;
            RIHVR_HIP(hipEventRecord(c->ev_rhat, c->stream));
        }
        return RIHVR_OK;
    }
    RIHVR_TRY(timed(c, K_ROWS_PLAIN, [&] { return fft_rows_plain(c->W, -1, r, c->htmp, c->H, c->twfW, c->stream); }));
    RIHVR_TRY(timed(c, K_COLS_PLAIN, [&] { return fft_cols_plain(c->H, -1, c->htmp, out, c->W, 1, c->twfH, c->stream); }));
    if (out == c->rhat) {
        c->rhat_scaled = false;
        RIHVR_HIP(hipEventRecord(c->ev_rhat, c->stream));
    }
    return RIHVR_OK;
}
// rhat now holds the residual spectrum FFT(r) / N written by a hologram kernel
rihvr_status rhat_written(rihvr_ctx* c) {
    c->rhat_scaled = true;
    RIHVR_HIP(hipEventRecord(c->ev_rhat, c->stream));
    return RIHVR_OK;
}

// A^T from R-hat for planes [g0, g0 + gz) on stream st (SparseGradient::getPlane,
// sparse_gradient.cu:125-165): row pass of H_z R-hat / N into T[g0..], column pass
// times e^{+i phi_z} into G[g0..]
rihvr_status adjoint_group(rihvr_ctx* c, float2* G, int g0, int gz, hipStream_t st) {
    const float invN = c->rhat_scaled ? 1.0f : (float)(1.0 / (double)c->N);
    float2* Tg = c->T + (size_t)g0 * c->N;
    if (c->gen) {
        float2* Tt = c->T2 + (size_t)g0 * c->N;
        RIHVR_TRY(timed_on(c, K_ROWS_ADJ, st, [&] {
            return gen_rows_adj(c->W, c->rhat, c->kappa, c->zlist + g0, Tg, c->H, gz, std::min(c->zchunk_adj, gz), invN,
                                c->twiW, st, c->gen_radix);
        }));
        return timed_on(c, K_COLS_ADJ, st, [&] {
            hipError_t e = gen_transpose(Tg, Tt, c->H, c->W, gz, nullptr, false, st);
            if (e == hipSuccess) e = gen_rows_plain(c->H, 1, Tt, Tt, (long)gz * c->W, c->twiH, st, c->gen_radix);
            if (e == hipSuccess) e = gen_transpose(Tt, G + (size_t)g0 * c->N, c->W, c->H, gz, c->phase + g0, false, st);
            return e;
        });
    }
    RIHVR_TRY(timed_on(c, K_ROWS_ADJ, st, [&] {
        return fft_rows_adj(c->W, c->rhat, c->kappa, c->zlist + g0, Tg, c->H, gz, std::min(c->zchunk_adj, gz), invN,
                            c->twiW, st);
    }));
    RIHVR_TRY(timed_on(c, K_COLS_ADJ, st, [&] {
        return fft_cols_adj(c->H, Tg, c->phase + g0, G + (size_t)g0 * c->N, c->W, gz, c->twiH, st);
    }));
    return RIHVR_OK;
}

// G = A^T r on the main stream
rihvr_status apply_adjoint(rihvr_ctx* c, const float2* r, float2* G) {
    RIHVR_TRY(spectrum(c, r, c->rhat));
    for (int g0 = 0; g0 < c->Nz; g0 += c->pipe_planes)
        RIHVR_TRY(adjoint_group(c, G, g0, std::min(c->pipe_planes, c->Nz - g0), c->stream));
    return RIHVR_OK;
}

int effective_mode(const rihvr_ctx* c) {
    // any mode becomes mode 5 when the TV weight is zero (:192-195)
    if (c->p.regularization_TV_param == 0.0f) return RIHVR_MODE_FASTA_L1;
    return c->p.mode;
}

bool tv3d(const rihvr_ctx* c) { return c->ss.mode == RIHVR_MODE_FASTA_FUSED_LASSO; }

// y = prox(x0 - t g), x0 = Y0 + s (Y0 - YM); reductions into scal[SL_PROX..]
// s_mom_dev (may be NULL): the momentum step in device memory (a captured graph reads it there);
// enqueue = false performs only the host-side bookkeeping of the maps (a replayed graph holds
// the launches)
rihvr_status apply_prox(rihvr_ctx* c, int mode, const float2* Y0, const float2* YM, float s_mom, const float2* G,
                        double t, float2* Y, const float* s_mom_dev = nullptr, bool enqueue = true) {
    ProxLaunch a;
    a.Y0 = Y0; a.YM = YM; a.G = G; a.Y = Y;
    a.W = c->W; a.H = c->H; a.nz = c->Nz;
    a.s_mom = s_mom; a.t = t;
    a.s_mom_dev = s_mom_dev;
    a.reg = (double)c->p.regularization_param;
    a.reg_tv = (double)c->p.regularization_TV_param;
    a.n_tv = c->p.num_tv_iterations;
    a.row_chunk = c->prox_rows;
    a.partial = c->partial;
    int rows;
    const int ky = ring_index(c, Y);
    if (mode == RIHVR_MODE_FASTA_FUSED_LASSO_2D && c->tvbuf) {
        // more TV iterations than the fused kernel holds: the multi-pass form (no maps)
        if (enqueue) RIHVR_TRY(timed(c, K_PROX, [&] { return prox_tv2_general(a, c->tvbuf, c->stream); }));
        rows = prox_tv2_num_blocks(c->W, c->H, c->Nz);
        if (ky >= 0) c->occ_valid[ky] = false;
    } else if (mode == RIHVR_MODE_FASTA_FUSED_LASSO_2D) {
        // the new map is built in occ_tmp, then becomes Y's; a volume without a valid map
        // passes the all-ones map (every line read, every line of Y rewritten)
        const uint64_t* o0 = occ_of(c, Y0);
        const uint64_t* om = occ_of(c, YM);
        a.occ_y0 = o0 ? o0 : c->occ_ones;
        a.occ_ym = om ? om : c->occ_ones;
        a.occ_old = (ky >= 0 && c->occ_valid[ky]) ? c->occ[ky] : c->occ_ones;
        a.occ_new = c->occ_tmp;
        a.occ_new_planes = c->occ_tmp + (size_t)c->Nz * c->H;
        if (enqueue) {
            RIHVR_HIP(fill_u64(c->occ_tmp, (long)map_words(c), 0ull, c->stream));
            RIHVR_TRY(timed(c, K_PROX, [&] { return prox_fl2d(a, c->stream); }));
        }
        rows = prox_fl2d_num_waves(c->W, c->H, c->Nz, a.n_tv, c->prox_rows);
        if (ky >= 0) {
            std::swap(c->occ[ky], c->occ_tmp);
            c->occ_valid[ky] = true;
        }
    } else if (mode == RIHVR_MODE_FASTA_FUSED_LASSO) {
        // the FFT scratch volume is free between A^T and A
        if (enqueue) RIHVR_TRY(timed(c, K_PROX, [&] { return prox_fl3d(a, c->T, c->stream); }));
        rows = prox_fl3d_num_blocks(c->W, c->H, c->Nz);
    } else {
        if (enqueue) RIHVR_TRY(timed(c, K_PROX, [&] { return prox_l1(a, c->stream); }));
        rows = prox_l1_num_blocks(c->W, c->H, c->Nz);
    }
    if ((mode != RIHVR_MODE_FASTA_FUSED_LASSO_2D || c->tvbuf) && ky >= 0) c->occ_valid[ky] = false;
    if (enqueue)
        RIHVR_TRY(timed(c, K_REDUCE, [&] { return reduce_partials(c->partial, rows, PS_COUNT, c->red_work, c->scal + SL_PROX, c->stream); }));
    return RIHVR_OK;
}

// First pass of an iteration whose gradient is not formed yet (mode 7): A^T of plane
// group g runs on the auxiliary stream while the prox and A of earlier groups run on
// the main stream; the main stream waits for group g's G planes before its prox.
// Same kernels, same arithmetic and the same partial-row layout as the one-stream
// sequence A^T (all planes), prox (all planes), A (all planes).
rihvr_status pipelined_pass(rihvr_ctx* c, const float2* Y0, const float2* YM, float s_mom, double t, float2* Yn) {
    if (!c->groups_one_stream) {
        RIHVR_HIP(hipStreamWaitEvent(c->aux, c->ev_rhat, 0));
        for (int g = 0, g0 = 0; g < c->n_groups; ++g, g0 += c->pipe_planes) {
            RIHVR_TRY(adjoint_group(c, c->G, g0, std::min(c->pipe_planes, c->Nz - g0), c->aux));
            RIHVR_HIP(hipEventRecord(c->ev_adj[g], c->aux));
        }
    }
    ProxLaunch a;
    a.W = c->W; a.H = c->H;
    a.s_mom = s_mom; a.t = t;
    a.reg = (double)c->p.regularization_param;
    a.reg_tv = (double)c->p.regularization_TV_param;
    a.n_tv = c->p.num_tv_iterations;
    a.row_chunk = c->prox_rows;
    const int per_plane = prox_fl2d_num_waves(c->W, c->H, 1, a.n_tv, c->prox_rows);
    const uint64_t* o0 = occ_of(c, Y0);
    const uint64_t* om = occ_of(c, YM);
    const int ky = ring_index(c, Yn);
    const uint64_t* oold = (ky >= 0 && c->occ_valid[ky]) ? c->occ[ky] : c->occ_ones;
    RIHVR_HIP(hipMemsetAsync(c->occ_tmp, 0, map_words(c) * sizeof(uint64_t), c->stream));
    for (int g = 0, g0 = 0; g < c->n_groups; ++g, g0 += c->pipe_planes) {
        const int gz = std::min(c->pipe_planes, c->Nz - g0);
        const size_t off = (size_t)g0 * c->N;
        const size_t roff = (size_t)g0 * c->H;  // first map word of the group
        if (c->groups_one_stream)
            RIHVR_TRY(adjoint_group(c, c->G, g0, gz, c->stream));
        else
            RIHVR_HIP(hipStreamWaitEvent(c->stream, c->ev_adj[g], 0));
        a.Y0 = Y0 + off; a.YM = YM + off; a.G = c->G + off; a.Y = Yn + off;
        a.nz = gz;
        a.partial = c->partial + (size_t)g0 * per_plane * PS_COUNT;
        a.occ_y0 = (o0 ? o0 : c->occ_ones) + roff;
        a.occ_ym = (om ? om : c->occ_ones) + roff;
        a.occ_old = oold + roff;
        a.occ_new = c->occ_tmp + roff;
        a.occ_new_planes = c->occ_tmp + (size_t)c->Nz * c->H + g0;
        RIHVR_TRY(timed(c, K_PROX, [&] { return prox_fl2d(a, c->stream); }));
        RIHVR_TRY(forward_group(c, Yn, g0, gz, c->stream, ky >= 0 ? c->occ_tmp : nullptr));
    }
    if (ky >= 0) {
        std::swap(c->occ[ky], c->occ_tmp);
        c->occ_valid[ky] = true;
    }
    RIHVR_TRY(timed(c, K_REDUCE, [&] {
        return reduce_partials(c->partial, per_plane * c->Nz, PS_COUNT, c->red_work, c->scal + SL_PROX, c->stream);
    }));
    return forward_finish(c, c->hX);
}

double objective(const rihvr_ctx* c, int mode, double summag, double l1, double tv) {
    // calcObjectiveFunction (:2838-2899)
    double f = 0.5 * summag + (double)c->p.regularization_param * l1;
    // TV term: 2D per plane for mode 7 (calcTVNorm2d), 3D z-periodic for mode 6 (calcTVNorm)
    if (mode == RIHVR_MODE_FASTA_FUSED_LASSO_2D || mode == RIHVR_MODE_FASTA_FUSED_LASSO)
        f = f + (double)c->p.regularization_TV_param * tv;
    return f;
}

void release(rihvr_ctx* c) {
    if (!c) return;
    for (auto& b : c->Yb) (void)hipFree(b);
    for (auto& m : c->occ) (void)hipFree(m);
    (void)hipFree(c->occ_tmp);
    (void)hipFree(c->occ_ones);
    (void)hipFree(c->smap);
    void* ptrs[] = {c->G, c->G_alt, c->T, c->T2, c->tvbuf, c->meas, c->hE, c->hC, c->hP, c->hX, c->resid, c->rhat, c->htmp, c->bspec,
                    c->hspat, c->spart, c->kappa,
                    c->zlist, c->phase, c->phase_fwd, c->twfW, c->twiW, c->twfH, c->twiH, c->partial, c->partial_holo, c->scal,
                    c->red_work, c->dflag};
    for (void* q : ptrs) (void)hipFree(q);
    for (auto& p : c->tm.pending) { (void)hipEventDestroy(p.a); (void)hipEventDestroy(p.b); }
    for (hipEvent_t e : c->tm.pool) (void)hipEventDestroy(e);
    if (c->h_scal) (void)hipHostFree(c->h_scal);
    for (hipEvent_t e : c->ev_adj) (void)hipEventDestroy(e);
    if (c->ev_rhat) (void)hipEventDestroy(c->ev_rhat);
    if (c->ev_scal) (void)hipEventDestroy(c->ev_scal);
    for (auto& gc : c->gcache)
        for (auto& g : gc) (void)hipGraphExecDestroy(g.exec);
    (void)hipFree(c->d_smom);
    (void)hipFree(c->d_step);
    if (c->aux) (void)hipStreamDestroy(c->aux);
    if (c->stream) (void)hipStreamDestroy(c->stream);
}

rihvr_status check_schedule(const rihvr_schedule* s) {
    if (!s) return RIHVR_OK;
    const bool ok = (s->fft_path == 0 || s->fft_path == 1) &&
                    (s->max_radix == 0 || s->max_radix == 4 || s->max_radix == 8 || s->max_radix == 16) &&
                    s->zchunk_fwd >= 0 && s->zchunk_adj >= 0 && s->prox_rows >= 0 && (s->graphs == 0 || s->graphs == 1);
    return ok ? RIHVR_OK : RIHVR_E_INVALID_ARGUMENT;
}

// sch: the launch choices (NULL: the defaults, with the RIHVR_* environment overrides)
rihvr_status create_impl(const rihvr_params* params, const rihvr_schedule* sch, rihvr_ctx** out) {
    RIHVR_TRY(validate(params));
    RIHVR_TRY(check_schedule(sch));
    rihvr_ctx* c = new rihvr_ctx();
    *out = c;
    c->p = *params;
    c->W = params->width;
    c->H = params->height;
    c->Nz = params->num_planes;
    c->N = (long)c->W * c->H;
    c->V = c->N * c->Nz;
    RIHVR_HIP(hipGetDevice(&c->device));
    RIHVR_HIP(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
    RIHVR_HIP(hipStreamCreateWithFlags(&c->aux, hipStreamNonBlocking));
    RIHVR_HIP(hipEventCreateWithFlags(&c->ev_rhat, hipEventDisableTiming));
    RIHVR_HIP(hipEventCreateWithFlags(&c->ev_scal, hipEventDisableTiming));
    // defaults: A's partial spectra over 32 planes; A^T's row pass walks 8 planes per workgroup
    // (at 1024^2 x 128: 588 against 584 iter/s with 16 over whole solves, the same bits)
    int zc = 32, za = 8, prow = -1;
    bool force_gen = false;
    if (sch) {
        if (sch->zchunk_fwd > 0) zc = sch->zchunk_fwd;
        if (sch->zchunk_adj > 0) za = sch->zchunk_adj;
        if (sch->prox_rows > 0) prow = sch->prox_rows;
        if (sch->max_radix > 0) c->gen_radix = sch->max_radix;
        force_gen = sch->fft_path == 1;
        c->graphs = sch->graphs != 0;
    } else {
        if (const char* e = std::getenv("RIHVR_ZCHUNK")) zc = std::max(1, std::atoi(e));
        if (const char* e = std::getenv("RIHVR_ZCHUNK_ADJ")) za = std::max(1, std::atoi(e));
        if (const char* e = std::getenv("RIHVR_PROX_ROWS")) prow = std::max(0, std::atoi(e));
    }
    if (prow >= 0)
        c->prox_rows = std::min(prow, c->H);
    else if (params->num_tv_iterations >= 0 && params->num_tv_iterations <= 10)
        c->prox_rows = prox_fl2d_auto_rows(c->W, c->H, c->Nz, params->num_tv_iterations);
    c->zchunk_fwd = std::min(c->Nz, zc);
    c->nparts = (c->Nz + c->zchunk_fwd - 1) / c->zchunk_fwd;
    c->zchunk_adj = std::min(c->Nz, za);
    // pipeline groups: one group (everything on the main stream) unless RIHVR_PIPE_PLANES
    // asks for groups of that many planes (rounded up to a multiple of 16).  Measured at
    // 1024^2 x 128 on MI355X the two-stream overlap loses (groups of 32: 287 iter/s,
    // 64: 308, one stream: 316): the prox and the column FFTs each fill the CUs' LDS, so
    // they time-share CUs instead of co-issuing
    int pp = c->Nz;
    if (const char* e = std::getenv("RIHVR_PIPE_PLANES")) pp = std::atoi(e) > 0 ? std::atoi(e) : c->Nz;
    pp = ((pp + 15) / 16) * 16;
    c->pipe_planes = std::min(c->Nz, std::max(16, pp));
    if (c->pipe_planes % c->zchunk_fwd || c->pipe_planes % c->zchunk_adj) c->pipe_planes = c->Nz;
    c->gen = force_gen || !(fft_length_supported(c->W) && fft_length_supported(c->H));
    const bool tv_general = params->mode == RIHVR_MODE_FASTA_FUSED_LASSO_2D &&
                            !prox_fl2d_supported(params->num_tv_iterations);
    if (c->gen || tv_general) c->pipe_planes = c->Nz;  // the plane-group schedule needs the fused kernels
    c->n_groups = (c->Nz + c->pipe_planes - 1) / c->pipe_planes;
    if (const char* e = std::getenv("RIHVR_GROUP_ONE_STREAM")) c->groups_one_stream = std::atoi(e) != 0;
    if (const char* e = std::getenv("RIHVR_TEST_BACKTRACKS")) c->test_backtracks = std::strtoull(e, nullptr, 0);
    c->ev_adj.assign(c->n_groups, nullptr);
    for (hipEvent_t& e : c->ev_adj) RIHVR_HIP(hipEventCreateWithFlags(&e, hipEventDisableTiming));
    for (auto& b : c->Yb) RIHVR_TRY(dalloc(&b, c->V));
    for (auto& m : c->occ) RIHVR_TRY(dalloc(&m, map_words(c)));
    RIHVR_TRY(dalloc(&c->occ_tmp, map_words(c)));
    RIHVR_TRY(dalloc(&c->occ_ones, map_words(c)));
    RIHVR_HIP(hipMemset(c->occ_ones, 0xff, map_words(c) * sizeof(uint64_t)));
    RIHVR_TRY(dalloc(&c->smap, (size_t)c->Nz * (c->W / 16)));
    RIHVR_TRY(dalloc(&c->G, c->V));
    RIHVR_TRY(dalloc(&c->G_alt, c->V));
    RIHVR_TRY(dalloc(&c->T, c->V));
    if (c->gen) RIHVR_TRY(dalloc(&c->T2, c->V));
    if (tv_general) RIHVR_TRY(dalloc(&c->tvbuf, 4 * (size_t)c->V));
    float2** holos[] = {&c->meas, &c->hE, &c->hC, &c->hP, &c->hX, &c->resid, &c->rhat, &c->htmp, &c->bspec, &c->hspat};
    for (float2** h : holos) RIHVR_TRY(dalloc(h, c->N));
    RIHVR_TRY(dalloc(&c->spart, (size_t)c->N * c->nparts));
    RIHVR_TRY(dalloc(&c->kappa, c->N));
    RIHVR_TRY(dalloc(&c->zlist, c->Nz));
    RIHVR_TRY(dalloc(&c->phase, c->Nz));
    RIHVR_TRY(dalloc(&c->phase_fwd, c->Nz));
    RIHVR_TRY(dalloc(&c->twfW, c->W));
    RIHVR_TRY(dalloc(&c->twiW, c->W));
    RIHVR_TRY(dalloc(&c->twfH, c->H));
    RIHVR_TRY(dalloc(&c->twiH, c->H));
    long rows = std::max<long>(prox_fl2d_num_waves(c->W, c->H, c->Nz, std::min(params->num_tv_iterations, 10), c->prox_rows),
                               prox_l1_num_blocks(c->W, c->H, c->Nz));
    rows = std::max<long>(rows, prox_fl3d_num_blocks(c->W, c->H, c->Nz));
    rows = std::max<long>(rows, momentum_out_num_blocks(c->W, c->H, c->Nz));
    rows = std::max<long>(rows, prox_tv2_num_blocks(c->W, c->H, c->Nz));
    c->partial_rows = rows;
    RIHVR_TRY(dalloc(&c->partial, (size_t)rows * PS_COUNT));
    c->holo_rows = std::max<long>(std::max<long>(holo_num_blocks(c->N), c->W / 16), holo_spec_num_blocks(c->W, c->H));
    c->holo_rows = std::max<long>(c->holo_rows, gen_transpose_real_blocks(c->H, c->W));
    RIHVR_TRY(dalloc(&c->partial_holo, c->holo_rows));
    RIHVR_TRY(dalloc(&c->scal, SL_COUNT));
    RIHVR_TRY(dalloc(&c->red_work, (size_t)kReduceBlocks * PS_COUNT));
    RIHVR_TRY(dalloc(&c->dflag, 4));
    RIHVR_TRY(dalloc(&c->d_smom, 1));
    RIHVR_TRY(dalloc(&c->d_step, 1));
    RIHVR_HIP(hipHostMalloc((void**)&c->h_scal, SL_COUNT * sizeof(double), hipHostMallocDefault));

    // kappa table, plane list and per-plane phase factors, computed on the host
    // exactly as the reference does (sparse_compressive_holo.cu:38-56;
    // holo_ui.cpp:478-483; phase = float(2 pi z / lambda), :1849-1855)
    std::vector<double> kap(c->N);
    const double lambda = (double)params->wavelength, reso = (double)params->resolution;
    for (int y = 0; y < c->H; ++y)
        for (int x = 0; x < c->W; ++x) {
            volatile double fx = (double)(((x + c->W / 2) % c->W) - (c->W / 2)) / (double)c->W;
            volatile double fy = (double)(((y + c->H / 2) % c->H) - (c->H / 2)) / (double)c->H;
            volatile double f2 = fx * fx + fy * fy;
            volatile double q = f2 * (lambda / reso);
            volatile double q2 = q * (lambda / reso);
            double sq = 1 - q2;
            sq = (sq < 0) ? 0 : sq;
            volatile double num = -2 * kPi * std::sqrt(sq);
            kap[(size_t)y * c->W + x] = num / lambda;
        }
    std::vector<float> z(c->Nz);
    std::vector<float2> ph(c->Nz), phf(c->Nz);
    const double invN = 1.0 / (double)c->N;
    for (int i = 0; i < c->Nz; ++i) {
        float step = (float)i * params->plane_stepsize;
        z[i] = params->start_plane + step;
        float arg = (float)(2 * kPi * (double)z[i] / lambda);
        ph[i] = make_float2((float)std::cos((double)arg), (float)std::sin((double)arg));
        phf[i] = make_float2((float)((double)ph[i].x * invN), (float)((double)ph[i].y * invN));
    }
    RIHVR_TRY(upload(c->kappa, kap.data(), sizeof(double) * c->N, c->stream));
    RIHVR_TRY(upload(c->zlist, z.data(), sizeof(float) * c->Nz, c->stream));
    RIHVR_TRY(upload(c->phase, ph.data(), sizeof(float2) * c->Nz, c->stream));
    RIHVR_TRY(upload(c->phase_fwd, phf.data(), sizeof(float2) * c->Nz, c->stream));
    auto tfw = twiddles(c->W, -1), tiw = twiddles(c->W, +1), tfh = twiddles(c->H, -1), tih = twiddles(c->H, +1);
    RIHVR_TRY(upload(c->twfW, tfw.data(), sizeof(float2) * c->W, c->stream));
    RIHVR_TRY(upload(c->twiW, tiw.data(), sizeof(float2) * c->W, c->stream));
    RIHVR_TRY(upload(c->twfH, tfh.data(), sizeof(float2) * c->H, c->stream));
    RIHVR_TRY(upload(c->twiH, tih.data(), sizeof(float2) * c->H, c->stream));
    c->out = c->Yb[0];
    return RIHVR_OK;
}

// ---------------------------------------------------------------------------
// inverseReconstruct (:173-633), modes 5, 6 and 7, split at its seams:
//   begin   = the set-up before the loop (:238-296)
//   iterate = n passes of the loop body (:299-607)
//   finish  = the returned momentum point and final objective (:483-489, :608-633)
// ---------------------------------------------------------------------------
rihvr_status begin_impl(rihvr_ctx* c, const float* holo, int holo_loc, const float* warm, int warm_loc,
                        double* f_initial) {
    if (!holo) return RIHVR_E_INVALID_ARGUMENT;
    c->have_result = false;
    c->ss = SolveState{};
    c->ss.mode = effective_mode(c);
    RIHVR_TRY(copy_in(c, c->meas, holo, holo_loc, c->N));
    RIHVR_HIP(hipMemsetAsync(c->dflag, 0, sizeof(int), c->stream));
    RIHVR_HIP(check_finite(c->meas, c->N, c->dflag, c->stream));
    int hflag = 0;
    RIHVR_HIP(hipMemcpyAsync(&hflag, c->dflag, sizeof(int), hipMemcpyDeviceToHost, c->stream));
    RIHVR_HIP(hipStreamSynchronize(c->stream));
    if (hflag) return RIHVR_E_INVALID_DATA;
    // the measurement's spectrum FFT(b) / N
    RIHVR_TRY(spectrum(c, c->meas, c->bspec));
    RIHVR_HIP(holo_scale(c->bspec, c->bspec, c->N, (float)(1.0 / (double)c->N), c->stream));

    // start point x1 = x0 = y0 = warm start or zero (:258-259, sparse_volume.cu:44-48);
    // a warm start that already is one of the ring volumes (the previous frame's
    // result in a batch) becomes y_{k-1} in place
    int in_ring = -1;
    for (int k = 0; k < 3; ++k)
        if (warm && warm_loc == RIHVR_DEVICE && warm == (const float*)c->Yb[k]) in_ring = k;
    // a warm start in a ring slot with its occupancy map keeps it (a zero start has the empty
    // map; another warm start none: every line is read once)
    const bool keep_map = warm && in_ring >= 0 && c->occ_valid[in_ring];
    if (in_ring >= 0) {
        c->iy0 = in_ring; c->iy = (in_ring + 1) % 3; c->iym = (in_ring + 2) % 3;
        // with its map, y_{k-2} is a slot whose map is valid when one is (the ring's order is
        // a permutation, not a cycle)
        if (keep_map && !c->occ_valid[c->iym] && c->occ_valid[c->iy]) std::swap(c->iy, c->iym);
    } else {
        c->iy = 0; c->iy0 = 1; c->iym = 2;
    }
    float2* Y0 = c->Yb[c->iy0];
    float2* YM = c->Yb[c->iym];
    if (warm && in_ring < 0) RIHVR_TRY(copy_in(c, Y0, warm, warm_loc, c->V));
    else if (!warm) RIHVR_HIP(hipMemsetAsync(Y0, 0, c->V * sizeof(float2), c->stream));
    if (keep_map && c->occ_valid[c->iym]) {
        // y_{k-2} keeps what it holds (with its map): the first prox runs at step s = 0, where
        // x0 = y0 + 0 (y0 - ym) is y0 for any finite ym, and after it the slot is the next
        // output, written over the lines its map marks.  Its content is finite here: the ring
        // slots of a warm start with a map (the learned start) hold the network's output and
        // the threshold candidates, and no candidate is chosen when they hold a NaN.
    } else {
        RIHVR_HIP(hipMemcpyAsync(YM, Y0, c->V * sizeof(float2), hipMemcpyDeviceToDevice, c->stream));
        if (keep_map) {
            RIHVR_HIP(hipMemcpyAsync(c->occ[c->iym], c->occ[in_ring], map_words(c) * sizeof(uint64_t),
                                     hipMemcpyDeviceToDevice, c->stream));
            c->occ_valid[c->iym] = true;
        } else {
            for (int k : {c->iy0, c->iym}) {
                c->occ_valid[k] = !warm;
                if (!warm) RIHVR_HIP(hipMemsetAsync(c->occ[k], 0, map_words(c) * sizeof(uint64_t), c->stream));
            }
        }
    }
    RIHVR_HIP(hipMemsetAsync(c->hC, 0, c->N * sizeof(float2), c->stream));
    RIHVR_HIP(hipMemsetAsync(c->hP, 0, c->N * sizeof(float2), c->stream));

    // initial residual, objective and gradient (:273-286)
    RIHVR_TRY(apply_forward(c, Y0, c->hE));
    double l1 = 0, tv = 0;
    if (warm) {
        // a start with its map reads only the lines whose terms can be non-zero (the same sums)
        RIHVR_TRY(timed(c, K_MOM_OUT, [&] {
            return keep_map ? volume_terms_map(Y0, c->occ[c->iy0], c->W, c->H, c->Nz, tv3d(c), c->partial, c->stream)
                            : momentum_out(Y0, Y0, 0.0f, nullptr, c->W, c->H, c->Nz, tv3d(c), c->partial, c->stream);
        }));
        RIHVR_TRY(timed(c, K_REDUCE, [&] {
            return reduce_partials(c->partial, momentum_out_num_blocks(c->W, c->H, c->Nz), PS_COUNT,
                                   c->red_work, c->scal + SL_PROX, c->stream);
        }));
    }
    RIHVR_TRY(fetch_scalars(c));
    if (warm) {
        l1 = c->h_scal[SL_PROX + PS_L1X0];
        tv = c->h_scal[SL_PROX + PS_TVX0];
    }
    c->ss.prev_summag = c->h_scal[SL_RESID_A];
    c->ss.f_initial = objective(c, c->ss.mode, c->ss.prev_summag, l1, tv);
    if (f_initial) *f_initial = c->ss.f_initial;
    // the residual's spectrum; the gradient's planes are formed with the first prox
    RIHVR_TRY(timed(c, K_HOLO_MOM, [&] {
        return holo_residual(c->hE, c->bspec, c->rhat, c->N, c->partial_holo, c->stream);
    }));
    RIHVR_TRY(rhat_written(c));
    c->ss.t = 1.0 / (double)c->Nz;  // :247 (a forced step is overridden there too)
    c->ss.alpha1 = 1.0;
    c->ss.s_mom = 0.0f;             // x0 = Y0 for the first prox
    c->ss.active = true;
    return RIHVR_OK;
}

// reduce_partials on the context's stream into out (timed as the solver's reductions)
rihvr_status reduce_partials_on(rihvr_ctx* c, const double* partial, int rows, int cols, double* out) {
    return timed(c, K_REDUCE, [&] { return reduce_partials(partial, rows, cols, c->red_work, out, c->stream); });
}

// objective of the current momentum point Y0 + s (Y0 - YM) with residual norm summag
rihvr_status momentum_objective(rihvr_ctx* c, double summag, double* f) {
    RIHVR_TRY(timed(c, K_MOM_OUT, [&] {
        return momentum_out(c->Yb[c->iy0], c->Yb[c->iym], c->ss.s_mom, nullptr, c->W, c->H, c->Nz, tv3d(c), c->partial,
                            c->stream);
    }));
    RIHVR_TRY(timed(c, K_REDUCE, [&] {
        return reduce_partials(c->partial, momentum_out_num_blocks(c->W, c->H, c->Nz), PS_COUNT,
                               c->red_work, c->scal + SL_PROX, c->stream);
    }));
    RIHVR_TRY(fetch_scalars(c));
    *f = objective(c, c->ss.mode, summag, c->h_scal[SL_PROX + PS_L1X0], c->h_scal[SL_PROX + PS_TVX0]);
    return RIHVR_OK;
}

// The prox's weights as one word (they are baked into a captured prox's arguments).
uint64_t weight_bits(const rihvr_ctx* c) {
    uint32_t a = 0, b = 0;
    std::memcpy(&a, &c->p.regularization_param, 4);
    std::memcpy(&b, &c->p.regularization_TV_param, 4);
    return ((uint64_t)a << 32) | b;
}

// Replays the graph cached for (key, t) on the context's stream, or captures `enqueue`'s launches
// into one first.  `bookkeep` performs the host-side state changes of the launches (it runs after
// a replay; a capture runs `enqueue`, which performs them itself).
template <class F, class B>
rihvr_status graph_run(rihvr_ctx* c, int which, const std::vector<const void*>& key, double t, F&& enqueue,
                       B&& bookkeep) {
    for (auto& g : c->gcache[which])
        if (g.t == t && g.key == key) {
            RIHVR_HIP(hipGraphLaunch(g.exec, c->stream));
            return bookkeep();
        }
    RIHVR_HIP(hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
    const rihvr_status st = enqueue();
    hipGraph_t graph = nullptr;
    const hipError_t e = hipStreamEndCapture(c->stream, &graph);
    if (st != RIHVR_OK) {
        if (graph) (void)hipGraphDestroy(graph);
        return st;
    }
    RIHVR_HIP(e);
    hipGraphExec_t exec = nullptr;
    const hipError_t ie = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
    (void)hipGraphDestroy(graph);
    RIHVR_HIP(ie);
    c->gcache[which].push_back({key, t, exec});
    RIHVR_HIP(hipGraphLaunch(exec, c->stream));
    return RIHVR_OK;
}

// f_stop > 0: stop once the objective at the momentum point of a completed iteration
// is <= f_stop.  That objective comes with the next iteration's first prox (its x0
// reductions), so the test runs there and a stop abandons that iteration: the ring
// is not rotated and the estimate at the accepted point is still in hE.  *done
// receives the number of iterations completed.
rihvr_status iterate_impl(rihvr_ctx* c, int n_it, rihvr_stats* st, bool last_f_by_finish, double f_stop = 0.0,
                          int* done = nullptr) {
    if (!c->ss.active) return RIHVR_E_INVALID_STATE;
    if (n_it < 0) return RIHVR_E_INVALID_ARGUMENT;
    if (done) *done = n_it;
    if (n_it > 0) c->have_result = false;  // a snapshot lives in the ring slot the prox writes
    SolveState& s = c->ss;
    const int mode = s.mode;
    // The work of an accepted iteration (hologram momentum, R-hat and the next gradient into
    // G_alt) is queued before the host has read the scalars of the prox and A, so the device
    // does not idle while the host decides; a backtrack (rare) discards it, and so does a stop
    // at the objective target, which then restores the accepted point's hologram momentum.
    // The step it uses is the host's, chosen on the device by the restart test (a restart
    // makes it 0).
    const bool ahead = c->n_groups == 1;
    for (int i = 0; i < n_it; ++i) {
        double alpha0 = s.alpha1;
        int bt = 0;
        double ratio = 0, lim = 0;
        float2* Yn = c->Yb[c->iy];
        float2* Y0 = c->Yb[c->iy0];
        float2* YM = c->Yb[c->iym];
        bool queued = false;  // the accepted-iteration work is queued
        for (;;) {
            const bool use_graph = ahead && c->graphs && !c->tm.on;
            if (!s.g_ready && mode == RIHVR_MODE_FASTA_FUSED_LASSO_2D && c->n_groups > 1) {
                RIHVR_TRY(pipelined_pass(c, Y0, YM, s.s_mom, s.t, Yn));
            } else {
                if (!s.g_ready)
                    for (int g0 = 0; g0 < c->Nz; g0 += c->pipe_planes)
                        RIHVR_TRY(adjoint_group(c, c->G, g0, std::min(c->pipe_planes, c->Nz - g0), c->stream));
                if (use_graph) {
                    RIHVR_HIP(set_scalars(c->d_smom, s.s_mom, nullptr, 0.0, c->stream));
                    const std::vector<const void*> key = {
                        Y0, YM, Yn, c->G, c->hX, c->occ[0], c->occ[1], c->occ[2], c->occ_tmp,
                        (const void*)(intptr_t)(c->occ_valid[0] | (c->occ_valid[1] << 1) | (c->occ_valid[2] << 2)),
                        (const void*)(intptr_t)mode, (const void*)(uintptr_t)weight_bits(c)};
                    RIHVR_TRY(graph_run(
                        c, 0, key, s.t,
                        [&] {
                            RIHVR_TRY(apply_prox(c, mode, Y0, YM, s.s_mom, c->G, s.t, Yn, c->d_smom, true));
                            return apply_forward(c, Yn, c->hX);
                        },
                        [&] { return apply_prox(c, mode, Y0, YM, s.s_mom, c->G, s.t, Yn, c->d_smom, false); }));
                } else {
                    RIHVR_TRY(apply_prox(c, mode, Y0, YM, s.s_mom, c->G, s.t, Yn));
                    RIHVR_TRY(apply_forward(c, Yn, c->hX));
                }
            }
            s.g_ready = true;
            if (ahead) {
                RIHVR_TRY(fetch_scalars_async(c));
                const double a1 = (1.0 + std::sqrt(1.0 + 4.0 * alpha0 * alpha0)) / 2.0;
                std::swap(c->hP, c->hC);
                std::swap(c->hC, c->hX);
                if (use_graph) {
                    RIHVR_HIP(set_scalars(nullptr, 0.0f, c->d_step, (alpha0 - 1) / a1, c->stream));
                    c->rhat_scaled = true;  // the graph's hologram momentum writes the residual spectrum
                    const bool next = i + 1 < n_it;
                    const std::vector<const void*> key = {c->hC, c->hP, c->hE, c->rhat, c->G_alt,
                                                          (const void*)(intptr_t)next};
                    RIHVR_TRY(graph_run(
                        c, 1, key, 0.0,
                        [&] {
                            RIHVR_HIP(holo_momentum(c->hC, c->hP, 0.0, c->bspec, c->hE, c->rhat, c->N, c->partial_holo,
                                                    c->stream, c->scal + SL_PROX + PS_RESTART, c->d_step));
                            RIHVR_HIP(reduce_partials(c->partial_holo, holo_num_blocks(c->N), 1, c->red_work,
                                                      c->scal + SL_RESID_M, c->stream));
                            if (next)
                                for (int g0 = 0; g0 < c->Nz; g0 += c->pipe_planes)
                                    RIHVR_TRY(adjoint_group(c, c->G_alt, g0, std::min(c->pipe_planes, c->Nz - g0),
                                                            c->stream));
                            return RIHVR_OK;
                        },
                        [&] { return RIHVR_OK; }));
                    RIHVR_HIP(hipEventRecord(c->ev_rhat, c->stream));
                } else {
                    RIHVR_TRY(timed(c, K_HOLO_MOM, [&] {
                        return holo_momentum(c->hC, c->hP, (alpha0 - 1) / a1, c->bspec, c->hE, c->rhat, c->N,
                                             c->partial_holo, c->stream, c->scal + SL_PROX + PS_RESTART);
                    }));
                    RIHVR_TRY(timed(c, K_REDUCE, [&] {
                        return reduce_partials(c->partial_holo, holo_num_blocks(c->N), 1, c->red_work,
                                               c->scal + SL_RESID_M, c->stream);
                    }));
                    RIHVR_TRY(rhat_written(c));
                    if (i + 1 < n_it)
                        for (int g0 = 0; g0 < c->Nz; g0 += c->pipe_planes)
                            RIHVR_TRY(adjoint_group(c, c->G_alt, g0, std::min(c->pipe_planes, c->Nz - g0), c->stream));
                }
                queued = true;
                RIHVR_TRY(wait_scalars(c));
            } else {
                RIHVR_TRY(fetch_scalars(c));
            }
            const double* sc = c->h_scal;
            if (bt == 0 && s.iters > 0) {
                // residual norm at the previous momentum point (:545-549), and the
                // objective there from this prox's x0 reductions
                s.prev_summag = sc[SL_RESID_M];
                const double f_prev =
                    objective(c, mode, s.prev_summag, sc[SL_PROX + PS_L1X0], sc[SL_PROX + PS_TVX0]);
                if (i > 0 && st) {
                    if (st->sparsity_it) st->sparsity_it[i - 1] = sc[SL_PROX + PS_NNZX0] / (double)c->V;
                    if (st->f_resid) st->f_resid[i - 1] = 0.5 * s.prev_summag;
                    if (st->f) st->f[i - 1] = f_prev;
                }
                if (f_stop > 0 && f_prev <= f_stop) {
                    // keep the accepted point.  Unqueued, its estimate is still in hE (the
                    // forward pass above wrote hX).  Queued, the hologram ring is put back and
                    // the accepted point's momentum (its estimate, R-hat and residual norm) is
                    // formed again from the same spectra and step, which gives the same bits.
                    if (queued) {
                        std::swap(c->hC, c->hX);
                        std::swap(c->hP, c->hC);
                        queued = false;
                        RIHVR_HIP(holo_momentum(c->hC, c->hP, s.step, c->bspec, c->hE, c->rhat, c->N,
                                                c->partial_holo, c->stream));
                        RIHVR_HIP(reduce_partials(c->partial_holo, holo_num_blocks(c->N), 1, c->red_work,
                                                  c->scal + SL_RESID_M, c->stream));
                        RIHVR_TRY(rhat_written(c));
                        s.g_ready = false;
                    }
                    if (done) *done = i;
                    n_it = i;
                    break;
                }
            }
            lim = sc[SL_PROX + PS_LS1] + sc[SL_PROX + PS_LS2] / (2 * s.t);  // :4381
            const double f_resid = 0.5 * sc[SL_RESID_A];
            ratio = f_resid / (0.5 * s.prev_summag + lim);  // :403
            const bool forced = bt == 0 && s.iters < 64 && ((c->test_backtracks >> s.iters) & 1ull);
            if (((ratio > 1.0) || (ratio < 0) || forced) && (bt < 2)) {  // :423
                if (queued) {  // discard the queued work: the hologram ring as it was
                    std::swap(c->hC, c->hX);
                    std::swap(c->hP, c->hC);
                    queued = false;
                }
                s.t *= 0.5;
                bt++;
                s.total_bt++;
                continue;
            }
            break;
        }
        if (i == n_it) break;  // stopped at the objective target
        const double rs = c->h_scal[SL_PROX + PS_RESTART];
        const bool restart = rs > 0;  // :4481
        if (restart) alpha0 = 1;
        s.alpha1 = (1.0 + std::sqrt(1.0 + 4.0 * alpha0 * alpha0)) / 2.0;  // :476
        const double step = (alpha0 - 1) / s.alpha1;                       // :481
        if (queued) {
            // the queued work used this step (0 after a restart); G_alt is the next gradient
            if (i + 1 < n_it) {
                std::swap(c->G, c->G_alt);
                s.g_ready = true;
            } else {
                s.g_ready = false;
            }
        } else {
            // hologram momentum by linearity (:494-498): prev <- cur, cur <- A y_k, est <- cur + s (cur - prev)
            std::swap(c->hP, c->hC);
            std::swap(c->hC, c->hX);
            RIHVR_TRY(timed(c, K_HOLO_MOM, [&] {
                return holo_momentum(c->hC, c->hP, step, c->bspec, c->hE, c->rhat, c->N, c->partial_holo, c->stream);
            }));
            RIHVR_TRY(timed(c, K_REDUCE, [&] {
                return reduce_partials(c->partial_holo, holo_num_blocks(c->N), 1, c->red_work, c->scal + SL_RESID_M, c->stream);
            }));
            // the next gradient (:545-546): R-hat now, its planes with the next prox
            RIHVR_TRY(rhat_written(c));
            s.g_ready = false;
        }
        // rotate the prox-output ring: y_{k-2} <- y_{k-1} <- y_k
        const int old_m = c->iym;
        c->iym = c->iy0;
        c->iy0 = c->iy;
        c->iy = old_m;
        s.s_mom = (float)step;
        s.step = step;
        s.iters++;
        if (st) {
            if (st->ratio) st->ratio[i] = ratio;
            if (st->lim) st->lim[i] = lim;
            if (st->backtracks) st->backtracks[i] = bt;
            if (st->restarts) st->restarts[i] = restart ? 1 : 0;
            if (st->restart_sum) st->restart_sum[i] = rs;
        }
    }
    if (st) {
        st->iterations = n_it;
        st->total_backtracks = s.total_bt;
        st->stepsize = s.t;
        st->f_initial = s.f_initial;
        if (n_it > 0 && !last_f_by_finish && (st->f || st->f_resid || st->sparsity_it)) {
            RIHVR_TRY(fetch_scalars(c));
            const double summag = c->h_scal[SL_RESID_M];
            double f = 0;
            RIHVR_TRY(momentum_objective(c, summag, &f));
            if (st->f) st->f[n_it - 1] = f;
            if (st->f_resid) st->f_resid[n_it - 1] = 0.5 * summag;
            if (st->sparsity_it) st->sparsity_it[n_it - 1] = c->h_scal[SL_PROX + PS_NNZX0] / (double)c->V;
        }
    }
    return RIHVR_OK;
}

rihvr_status finish_impl(rihvr_ctx* c, rihvr_stats* st, int last_index) {
    if (!c->ss.active) return RIHVR_E_INVALID_STATE;
    SolveState& s = c->ss;
    // returned volume: x1 = y_K + s (y_K - y_{K-1})  (:483-489); with no
    // iteration s = 0 and the start point is returned
    float2* outb = c->Yb[c->iy];
    c->occ_valid[c->iy] = false;
    RIHVR_TRY(timed(c, K_MOM_OUT, [&] {
        return momentum_out(c->Yb[c->iy0], c->Yb[c->iym], s.s_mom, outb, c->W, c->H, c->Nz, tv3d(c), c->partial,
                            c->stream);
    }));
    RIHVR_TRY(timed(c, K_REDUCE, [&] {
        return reduce_partials(c->partial, momentum_out_num_blocks(c->W, c->H, c->Nz), PS_COUNT, c->red_work,
                               c->scal + SL_PROX, c->stream);
    }));
    RIHVR_TRY(fetch_scalars(c));
    const double* sc = c->h_scal;
    const double summag = s.iters > 0 ? sc[SL_RESID_M] : s.prev_summag;
    const double f_last = objective(c, s.mode, summag, sc[SL_PROX + PS_L1X0], sc[SL_PROX + PS_TVX0]);
    c->out = outb;
    c->have_result = true;
    s.active = false;
    if (st) {
        if (last_index >= 0) {
            if (st->f) st->f[last_index] = f_last;
            if (st->f_resid) st->f_resid[last_index] = 0.5 * summag;
            if (st->sparsity_it) st->sparsity_it[last_index] = sc[SL_PROX + PS_NNZX0] / (double)c->V;
        }
        st->total_backtracks = s.total_bt;
        st->stepsize = s.t;
        st->f_initial = s.f_initial;
        st->f_final = f_last;
        st->nnz = (int64_t)llround(sc[SL_PROX + PS_NNZX0]);
        st->sparsity = (double)st->nnz / (double)c->V;
    }
    return RIHVR_OK;
}

// x1 = y_K + s (y_K - y_{K-1}) into the free ring slot; the solve stays active
rihvr_status snapshot_impl(rihvr_ctx* c) {
    if (!c->ss.active) return RIHVR_E_INVALID_STATE;
    float2* outb = c->Yb[c->iy];
    c->occ_valid[c->iy] = false;
    RIHVR_TRY(timed(c, K_MOM_OUT, [&] {
        return momentum_out(c->Yb[c->iy0], c->Yb[c->iym], c->ss.s_mom, outb, c->W, c->H, c->Nz, tv3d(c), c->partial,
                            c->stream);
    }));
    RIHVR_HIP(hipStreamSynchronize(c->stream));
    c->out = outb;
    c->have_result = true;
    return RIHVR_OK;
}

// One frame run until the objective at the momentum point of a completed iteration
// reaches f_target, or max_it iterations; f_target <= 0 runs max_it iterations.  The
// objective test costs no extra pass (see iterate_impl); a start point already on
// target runs no iteration.
rihvr_status solve_frame(rihvr_ctx* c, const float* holo, int holo_loc, const float* warm, int warm_loc,
                         double f_target, int max_it, rihvr_frame_stats* fs) {
    double f0 = 0;
    RIHVR_TRY(begin_impl(c, holo, holo_loc, warm, warm_loc, &f0));
    int its = 0;
    const double f = f0;
    if (f_target > 0) {
        if (!(f <= f_target)) RIHVR_TRY(iterate_impl(c, max_it, nullptr, true, f_target, &its));
    } else {
        RIHVR_TRY(iterate_impl(c, max_it, nullptr, true));
        its = max_it;
    }
    rihvr_stats fin{};
    RIHVR_TRY(finish_impl(c, &fin, -1));
    if (fs) {
        fs->iterations = its;
        fs->total_backtracks = fin.total_backtracks;
        fs->f_initial = f0;
        fs->f_final = fin.f_final;
        fs->nnz = fin.nnz;
        // stopped by the target (its < max_it), or at the cap with the final objective on target
        fs->reached = f_target > 0 ? ((its < max_it || fin.f_final <= f_target) ? 1 : 0) : 1;
    }
    return RIHVR_OK;
}

rihvr_status batch_impl(rihvr_ctx* c, const float* holos, int holo_loc, int n_frames, const float* warm0,
                        int warm_loc, int warm_previous, const double* f_target, int32_t max_iterations,
                        float* volumes, int vol_loc, rihvr_frame_stats* fs) {
    if (!holos || n_frames < 0) return RIHVR_E_INVALID_ARGUMENT;
    const int max_it = max_iterations > 0 ? max_iterations : c->p.num_inverse_iterations;
    for (int k = 0; k < n_frames; ++k) {
        const float* h = holos + 2 * (size_t)k * c->N;
        const float* w = warm0;
        int wl = warm_loc;
        if (k > 0 && warm_previous) {
            w = (const float*)c->out;  // the previous frame's result, still on the device
            wl = RIHVR_DEVICE;
        }
        const double tgt = f_target ? f_target[k] : 0.0;
        RIHVR_TRY(solve_frame(c, h, holo_loc, w, wl, tgt, max_it, fs ? fs + k : nullptr));
        if (volumes) RIHVR_TRY(copy_out(c, volumes + 2 * (size_t)k * c->V, c->out, vol_loc, c->V));
    }
    return RIHVR_OK;
}

// The learned start: A^T b (the adjoint of the measurement, SparseGradient::reconstruct with
// r = b) -> network -> x_net in a ring slot -> soft(x_net, tau) in another, tau = q max|x_net|
// for the q of kSoftFractions whose start has the least objective (calcObjectiveFunction;
// on the 1024^2 x 128 workload q = 0.02 and 0.05 both reached the plain objective in the
// fewest iterations, 0.1 and up in more).
// That objective is compared with the objectives at the A^T b start t A^T b (t = 1/Nz; its
// data residual is zero for a real hologram since A A^T = Nz I there) and at the reference's
// zero start (0.5 ||b||^2): the thresholded start is used when it is below both, else the zero
// start.  Then FISTA to the target.
constexpr double kSoftFractions[] = {0.02, 0.05};

rihvr_status learned_impl(rihvr_ctx* c, rihvr_net* net, const float* holo, int holo_loc, double f_target,
                          int max_iterations, rihvr_learned_stats* out) {
    if (!holo || !rihvr::net::net_matches(net, c->Nz, c->H, c->W)) return RIHVR_E_INVALID_ARGUMENT;
    using clk = std::chrono::steady_clock;
    auto ms = [](clk::time_point a, clk::time_point b) {
        return std::chrono::duration<double, std::milli>(b - a).count();
    };
    rihvr_learned_stats st{};
    const int max_it = max_iterations > 0 ? max_iterations : c->p.num_inverse_iterations;
    c->have_result = false;
    c->ss = SolveState{};
    c->ss.mode = effective_mode(c);
    const int mode = c->ss.mode;
    const auto t0 = clk::now();
    RIHVR_TRY(copy_in(c, c->meas, holo, holo_loc, c->N));
    RIHVR_HIP(hipMemsetAsync(c->dflag, 0, sizeof(int), c->stream));
    RIHVR_HIP(check_finite(c->meas, c->N, c->dflag, c->stream));
    int hflag = 0;
    RIHVR_HIP(hipMemcpyAsync(&hflag, c->dflag, sizeof(int), hipMemcpyDeviceToHost, c->stream));
    RIHVR_HIP(hipStreamSynchronize(c->stream));
    if (hflag) return RIHVR_E_INVALID_DATA;
    RIHVR_TRY(apply_adjoint(c, c->meas, c->G));  // R-hat = FFT(b) (unscaled), G = A^T b
    RIHVR_HIP(hipStreamSynchronize(c->stream));
    const auto t1 = clk::now();
    float2* xw = c->Yb[1];
    RIHVR_TRY(rihvr::net::predict(net, (const float*)c->G, (float*)xw, c->stream));
    invalidate_occ(c, xw);
    // the measurement's spectrum FFT(b) / N, and 0.5 ||b||^2 = 0.5 N sum |FFT(b) / N|^2
    RIHVR_HIP(holo_scale(c->rhat, c->bspec, c->N, (float)(1.0 / (double)c->N), c->stream));
    RIHVR_HIP(hipMemsetAsync(c->hC, 0, c->N * sizeof(float2), c->stream));
    RIHVR_HIP(holo_residual(c->bspec, c->hC, c->htmp, c->N, c->partial_holo, c->stream));
    RIHVR_TRY(reduce_partials_on(c, c->partial_holo, holo_num_blocks(c->N), 1, c->scal + SL_RESID_M));
    RIHVR_TRY(fetch_scalars(c));
    st.f_zero = 0.5 * c->h_scal[SL_RESID_M];
    // objective and data residual of a start volume X
    // (map: X's valid line-occupancy map, or NULL)
    auto volume_objective = [&](const float2* X, const uint64_t* map, double* resid, double* f) -> rihvr_status {
        RIHVR_TRY(apply_forward(c, X, c->hE));
        RIHVR_TRY(timed(c, K_MOM_OUT, [&] {
            return map ? volume_terms_map(X, map, c->W, c->H, c->Nz, tv3d(c), c->partial, c->stream)
                       : momentum_out(X, X, 0.0f, nullptr, c->W, c->H, c->Nz, tv3d(c), c->partial, c->stream);
        }));
        RIHVR_TRY(reduce_partials_on(c, c->partial, momentum_out_num_blocks(c->W, c->H, c->Nz), PS_COUNT,
                                     c->scal + SL_PROX));
        RIHVR_TRY(fetch_scalars(c));
        *resid = 0.5 * c->h_scal[SL_RESID_A];
        *f = objective(c, mode, c->h_scal[SL_RESID_A], c->h_scal[SL_PROX + PS_L1X0], c->h_scal[SL_PROX + PS_TVX0]);
        return RIHVR_OK;
    };
    // The network's output is dense (every voxel holds a small value), which the objective's L1
    // and TV terms price although FISTA's first prox removes it.  The candidate starts are
    // soft(x_net, q max|x_net|), q in kSoftFractions: the one of least objective is the
    // network's start (x_net's own objective is not needed, and not evaluated).
    st.resid_net = st.f_net = std::nan("");
    const int nb = (int)std::max<long>(1, std::min<long>(4096, c->partial_rows * PS_COUNT));
    RIHVR_HIP(maxabs_partials(xw, c->V, nb, c->partial, c->stream));
    std::vector<double> mx(nb);
    RIHVR_HIP(hipMemcpyAsync(mx.data(), c->partial, nb * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    RIHVR_HIP(hipStreamSynchronize(c->stream));
    const double xmax = *std::max_element(mx.begin(), mx.end());
    // the candidates go to the two other ring slots (both formed in one pass over x_net, each with
    // its occupancy map, so A and the first proxes skip its empty lines); the best one is kept
    // where it is
    static_assert(sizeof(kSoftFractions) / sizeof(kSoftFractions[0]) == 2, "two candidates per pass");
    float2* cand[2] = {c->Yb[2], c->Yb[0]};
    float tau[2];
    for (int k = 0; k < 2; ++k) {
        const int kb = ring_index(c, cand[k]);
        tau[k] = (float)(kSoftFractions[k] * xmax);
        RIHVR_HIP(fill_u64(c->occ[kb] + (size_t)c->Nz * c->H, (long)c->Nz, 0ull, c->stream));
        c->occ_valid[kb] = true;
    }
    RIHVR_HIP(soft_threshold_map2(xw, c->W, c->H, c->Nz, tau[0], cand[0], c->occ[ring_index(c, cand[0])], tau[1],
                                  cand[1], c->occ[ring_index(c, cand[1])], c->stream));
    float2* xs = nullptr;
    st.f_thresholded = HUGE_VAL;
    for (int k = 0; k < 2; ++k) {
        double r = 0, f = 0;
        RIHVR_TRY(volume_objective(cand[k], c->occ[ring_index(c, cand[k])], &r, &f));
        if (f < st.f_thresholded) {
            st.f_thresholded = f;
            st.tau = tau[k];
            xs = cand[k];
        }
    }
    RIHVR_HIP(holo_scale(c->G, c->G_alt, c->V, (float)(1.0 / (double)c->Nz), c->stream));
    RIHVR_TRY(volume_objective(c->G_alt, nullptr, &st.resid_adjoint, &st.f_adjoint));
    st.fallback = (xs && st.f_thresholded <= st.f_adjoint && st.f_thresholded <= st.f_zero) ? 0 : 1;
    const auto t2 = clk::now();
    rihvr_frame_stats fs{};
    RIHVR_TRY(solve_frame(c, holo, holo_loc, st.fallback ? nullptr : (const float*)xs, RIHVR_DEVICE, f_target,
                          max_it, &fs));
    const auto t3 = clk::now();
    st.iterations = fs.iterations;
    st.reached = fs.reached;
    st.total_backtracks = fs.total_backtracks;
    st.f_start = fs.f_initial;
    st.f_final = fs.f_final;
    st.ms_adjoint = ms(t0, t1);
    st.ms_network = ms(t1, t2);
    st.ms_solve = ms(t2, t3);
    if (out) *out = st;
    return RIHVR_OK;
}

rihvr_status reconstruct_impl(rihvr_ctx* c, const float* holo, int holo_loc, const float* warm, int warm_loc,
                              rihvr_stats* st) {
    const int n_it = c->p.num_inverse_iterations;
    RIHVR_TRY(begin_impl(c, holo, holo_loc, warm, warm_loc, nullptr));
    RIHVR_TRY(iterate_impl(c, n_it, st, true));
    RIHVR_TRY(finish_impl(c, st, n_it - 1));
    if (st) st->iterations = n_it;
    return RIHVR_OK;
}

// sum |v|^2 of a volume into *out (host), through the prox partials' buffer
rihvr_status volume_norm2(rihvr_ctx* c, const float2* v, double* out) {
    const int nb = (int)std::max<long>(1, std::min<long>(4096, c->partial_rows * PS_COUNT));
    RIHVR_HIP(norm2_partials(v, c->V, nb, c->partial, c->stream));
    RIHVR_TRY(reduce_partials_on(c, c->partial, nb, 1, c->scal + SL_RESID_M));
    RIHVR_TRY(fetch_scalars(c));
    *out = c->h_scal[SL_RESID_M];
    return RIHVR_OK;
}

// The Lipschitz constant of the data term's gradient, ||A^T A||, by power iteration from a
// seeded pseudo-random volume.  The reference (estimateLipschitz, sparse_compressive_holo.cu:
// 2205-2464) takes the secant ||grad f(x2) - grad f(x1)|| / ||x2 - x1|| of two random volumes,
// which is ||A^T A d|| / ||d|| for d = x2 - x1 (the measurement cancels): the first ratio here.
// Each further iteration applies A^T A to the normalised last result.  Floor 1e-6 (:2460).
rihvr_status lipschitz_impl(rihvr_ctx* c, int seed, int iterations, double* L) {
    if (!L || iterations < 1) return RIHVR_E_INVALID_ARGUMENT;
    c->have_result = false;
    c->ss.active = false;
    float2* v = c->Yb[0];
    c->occ_valid[0] = false;
    RIHVR_HIP(synth_volume(v, c->V, 1u, (unsigned)seed, c->stream));  // every voxel in [-0.5, 0.5)^2
    RIHVR_HIP(hipMemsetAsync(c->meas, 0, c->N * sizeof(float2), c->stream));
    RIHVR_HIP(hipMemsetAsync(c->bspec, 0, c->N * sizeof(float2), c->stream));
    double nv = 0, ratio = 0;
    RIHVR_TRY(volume_norm2(c, v, &nv));
    for (int k = 0; k < iterations; ++k) {
        RIHVR_TRY(apply_forward(c, v, c->hE));
        RIHVR_TRY(spatial_estimate(c, c->hE));
        RIHVR_TRY(apply_adjoint(c, c->hspat, c->G));
        double ng = 0;
        RIHVR_TRY(volume_norm2(c, c->G, &ng));
        ratio = nv > 0 ? std::sqrt(ng / nv) : 0.0;
        if (k + 1 < iterations && ng > 0) {
            RIHVR_HIP(holo_scale(c->G, v, c->V, (float)(1.0 / std::sqrt(ng)), c->stream));
            RIHVR_TRY(volume_norm2(c, v, &nv));
        }
    }
    *L = std::max(ratio, 1e-6);
    return RIHVR_OK;
}

rihvr_schedule schedule_of(const rihvr_ctx* c) {
    rihvr_schedule s{};
    s.fft_path = c->gen ? 1 : 0;
    s.max_radix = c->gen_radix;
    s.zchunk_fwd = c->zchunk_fwd;
    s.zchunk_adj = c->zchunk_adj;
    s.prox_rows = c->prox_rows;
    s.graphs = c->graphs ? 1 : 0;
    return s;
}

// Owns a context made by create_impl (also a half-made one) until the scope ends.
struct CtxGuard {
    rihvr_ctx* c = nullptr;
    ~CtxGuard() {
        if (!c) return;
        if (c->stream) (void)hipStreamSynchronize(c->stream);
        if (c->aux) (void)hipStreamSynchronize(c->aux);
        release(c);
        delete c;
    }
};

// The schedule search: each candidate is a fresh context solving `holo` (device memory) from the
// zero start, `warm` untimed iterations (enough to capture every graph of the ring's pointer
// combinations) then `iterations` timed ones between stream synchronisations.
rihvr_status autotune_impl(const rihvr_params* p, const float* holo, int holo_loc, int iterations, rihvr_schedule* best,
                           rihvr_schedule* tried, int capacity, int32_t* n_tried) {
    RIHVR_TRY(validate(p));
    if (!best || iterations < 1 || capacity < 0 || (capacity > 0 && !tried)) return RIHVR_E_INVALID_ARGUMENT;
    const size_t N = (size_t)p->width * p->height;
    float2* hol = nullptr;
    RIHVR_HIP(hipMalloc((void**)&hol, N * sizeof(float2)));
    struct Free {
        float2* q;
        ~Free() { (void)hipFree(q); }
    } free_hol{hol};
    if (holo) {
        RIHVR_HIP(hipMemcpy(hol, holo, N * sizeof(float2),
                            holo_loc == RIHVR_DEVICE ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice));
    } else {
        // A x of a sparse pseudo-random x (about 1 voxel in 1000 non-zero)
        CtxGuard g;
        RIHVR_TRY(create_impl(p, nullptr, &g.c));
        rihvr_ctx* c = g.c;
        RIHVR_HIP(synth_volume(c->Yb[0], c->V, 1000u, 0x5eedu, c->stream));
        c->occ_valid[0] = false;
        RIHVR_HIP(hipMemsetAsync(c->meas, 0, c->N * sizeof(float2), c->stream));
        RIHVR_HIP(hipMemsetAsync(c->bspec, 0, c->N * sizeof(float2), c->stream));
        RIHVR_TRY(apply_forward(c, c->Yb[0], c->hE));
        RIHVR_TRY(spatial_estimate(c, c->hE));
        RIHVR_HIP(hipMemcpyAsync(hol, c->hspat, N * sizeof(float2), hipMemcpyDeviceToDevice, c->stream));
        RIHVR_HIP(hipStreamSynchronize(c->stream));
    }
    const int warm = 12;
    int count = 0;
    auto measure = [&](rihvr_schedule& s) -> rihvr_status {
        CtxGuard g;
        RIHVR_TRY(create_impl(p, &s, &g.c));
        rihvr_ctx* c = g.c;
        RIHVR_TRY(begin_impl(c, (const float*)hol, RIHVR_DEVICE, nullptr, RIHVR_DEVICE, nullptr));
        RIHVR_TRY(iterate_impl(c, warm, nullptr, false));
        RIHVR_HIP(hipStreamSynchronize(c->stream));
        const auto t0 = std::chrono::steady_clock::now();
        RIHVR_TRY(iterate_impl(c, iterations, nullptr, false));
        RIHVR_HIP(hipStreamSynchronize(c->stream));
        const auto t1 = std::chrono::steady_clock::now();
        const rihvr_schedule eff = schedule_of(c);
        s.fft_path = eff.fft_path;
        s.zchunk_fwd = eff.zchunk_fwd;
        s.zchunk_adj = eff.zchunk_adj;
        s.ms_per_iteration = std::chrono::duration<double, std::milli>(t1 - t0).count() / iterations;
        if (count < capacity) tried[count] = s;
        ++count;
        return RIHVR_OK;
    };
    auto same = [](const rihvr_schedule& a, const rihvr_schedule& b) {
        return a.fft_path == b.fft_path && a.max_radix == b.max_radix && a.zchunk_fwd == b.zchunk_fwd &&
               a.zchunk_adj == b.zchunk_adj && a.prox_rows == b.prox_rows && a.graphs == b.graphs;
    };
    // the starting point: the defaults
    rihvr_schedule cur{};
    cur.max_radix = 16;
    cur.zchunk_fwd = std::min(32, p->num_planes);
    cur.zchunk_adj = std::min(8, p->num_planes);
    cur.graphs = 1;
    const bool fixed = fft_length_supported(p->width) && fft_length_supported(p->height);
    if (!fixed) cur.fft_path = 1;
    RIHVR_TRY(measure(cur));
    // a candidate that does not fit in memory (the run-time plans' transposed planes take one more
    // volume) is skipped; any other error ends the search
    auto try_set = [&](std::vector<rihvr_schedule> cands) -> rihvr_status {
        for (rihvr_schedule s : cands) {
            if (same(s, cur)) continue;
            const rihvr_status e = measure(s);
            if (e == RIHVR_E_OUT_OF_MEMORY) continue;
            RIHVR_TRY(e);
            if (s.ms_per_iteration < cur.ms_per_iteration) cur = s;
        }
        return RIHVR_OK;
    };
    std::vector<rihvr_schedule> v;
    for (int r : {16, 8, 4}) {
        rihvr_schedule s = cur;
        s.fft_path = 1;
        s.max_radix = r;
        v.push_back(s);
    }
    RIHVR_TRY(try_set(v));
    const rihvr_schedule base = cur;
    v.clear();
    for (int z : {8, 16, 32, 64}) {
        rihvr_schedule s = base;
        s.zchunk_fwd = std::min(z, p->num_planes);
        v.push_back(s);
    }
    RIHVR_TRY(try_set(v));
    v.clear();
    for (int z : {4, 8, 16, 32}) {
        rihvr_schedule s = cur;
        s.zchunk_adj = std::min(z, p->num_planes);
        v.push_back(s);
    }
    RIHVR_TRY(try_set(v));
    if (p->mode == RIHVR_MODE_FASTA_FUSED_LASSO_2D && prox_fl2d_supported(p->num_tv_iterations)) {
        v.clear();
        for (int r : {64, 128, 256}) {
            rihvr_schedule s = cur;
            s.prox_rows = std::min(r, p->height);
            v.push_back(s);
        }
        RIHVR_TRY(try_set(v));
    }
    v.clear();
    {
        rihvr_schedule s = cur;
        s.graphs = 1 - cur.graphs;
        v.push_back(s);
    }
    RIHVR_TRY(try_set(v));
    *best = cur;
    if (n_tried) *n_tried = count;
    return RIHVR_OK;
}

// The parameters that fix a context's tables and buffers; items that agree on them share one.
bool same_geometry(const rihvr_params& a, const rihvr_params& b) {
    return a.width == b.width && a.height == b.height && a.num_planes == b.num_planes &&
           std::memcmp(&a.start_plane, &b.start_plane, 4) == 0 &&
           std::memcmp(&a.plane_stepsize, &b.plane_stepsize, 4) == 0 &&
           std::memcmp(&a.wavelength, &b.wavelength, 4) == 0 && std::memcmp(&a.resolution, &b.resolution, 4) == 0 &&
           a.num_tv_iterations == b.num_tv_iterations && a.mode == b.mode;
}

rihvr_status multi_impl(rihvr_multi_item* items, int n, int holo_loc, int vol_loc, int autotune) {
    if (n < 0 || (n > 0 && !items) || autotune < 0 || autotune > 2) return RIHVR_E_INVALID_ARGUMENT;
    for (int i = 0; i < n; ++i) {
        RIHVR_TRY(validate(&items[i].params));
        if (!items[i].holo) return RIHVR_E_INVALID_ARGUMENT;
    }
    std::vector<char> done(n, 0);
    for (int i = 0; i < n; ++i) {
        if (done[i]) continue;
        rihvr_schedule sch{};
        if (autotune == 1)
            RIHVR_TRY(autotune_impl(&items[i].params, items[i].holo, holo_loc, 10, &sch, nullptr, 0, nullptr));
        else if (autotune == 2)
            sch = items[i].schedule;
        CtxGuard g;
        RIHVR_TRY(create_impl(&items[i].params, autotune ? &sch : nullptr, &g.c));
        rihvr_ctx* c = g.c;
        for (int j = i; j < n; ++j) {
            if (done[j] || !same_geometry(items[i].params, items[j].params)) continue;
            rihvr_multi_item& it = items[j];
            // the weights and the iteration count are read per solve
            c->p.regularization_param = it.params.regularization_param;
            c->p.regularization_TV_param = it.params.regularization_TV_param;
            c->p.num_inverse_iterations = it.params.num_inverse_iterations;
            c->p.zero_padding = it.params.zero_padding;
            RIHVR_TRY(solve_frame(c, it.holo, holo_loc, nullptr, RIHVR_DEVICE, it.f_target,
                                  it.params.num_inverse_iterations, &it.stats));
            if (it.volume) RIHVR_TRY(copy_out(c, it.volume, c->out, vol_loc, c->V));
            it.schedule = schedule_of(c);
            it.schedule.ms_per_iteration = autotune ? sch.ms_per_iteration : 0.0;
            done[j] = 1;
        }
    }
    return RIHVR_OK;
}

}  // namespace

extern "C" {

int rihvr_abi_version(void) { return RIHVR_ABI_VERSION; }

const char* rihvr_status_string(rihvr_status s) {
    switch (s) {
        case RIHVR_OK: return "ok";
        case RIHVR_E_UNKNOWN_TYPE: return "unknown type";
        case RIHVR_E_BAD_FILENAME: return "bad filename";
        case RIHVR_E_MISSING_DATA: return "missing data";
        case RIHVR_E_INVALID_STATE: return "invalid state";
        case RIHVR_E_OUT_OF_MEMORY: return "out of memory";
        case RIHVR_E_UNKNOWN_MODE: return "unknown mode";
        case RIHVR_E_INVALID_ARGUMENT: return "invalid argument";
        case RIHVR_E_UNKNOWN_ERROR: return "unknown error";
        case RIHVR_E_INVALID_FILE: return "invalid file";
        case RIHVR_E_INVALID_DATA: return "invalid data";
        case RIHVR_E_CRITICAL_ASSUMPTION: return "critical assumption";
        case RIHVR_E_INFINITE_LOOP: return "infinite loop";
        case RIHVR_E_DEVICE: return "HIP device error";
    }
    return "unrecognised status";
}

rihvr_status rihvr_create(const rihvr_params* params, rihvr_ctx** out) {
    if (!out) return RIHVR_E_INVALID_ARGUMENT;
    *out = nullptr;
    rihvr_ctx* c = nullptr;
    rihvr_status s = create_impl(params, nullptr, &c);
    if (s != RIHVR_OK) {
        if (c) {
            release(c);
            delete c;
        }
        return s;
    }
    *out = c;
    return RIHVR_OK;
}

rihvr_status rihvr_estimate_lipschitz(rihvr_ctx* ctx, int32_t seed, int32_t iterations, double* lipschitz) {
    if (!ctx) return RIHVR_E_INVALID_ARGUMENT;
    return lipschitz_impl(ctx, seed, iterations, lipschitz);
}

rihvr_status rihvr_create_scheduled(const rihvr_params* params, const rihvr_schedule* schedule, rihvr_ctx** out) {
    if (!out) return RIHVR_E_INVALID_ARGUMENT;
    *out = nullptr;
    CtxGuard g;
    RIHVR_TRY(create_impl(params, schedule, &g.c));
    *out = g.c;
    g.c = nullptr;
    return RIHVR_OK;
}

rihvr_status rihvr_get_schedule(const rihvr_ctx* ctx, rihvr_schedule* out) {
    if (!ctx || !out) return RIHVR_E_INVALID_ARGUMENT;
    *out = schedule_of(ctx);
    return RIHVR_OK;
}

rihvr_status rihvr_autotune(const rihvr_params* params, const float* holo, int holo_loc, int32_t iterations,
                            rihvr_schedule* best, rihvr_schedule* tried, int32_t capacity, int32_t* n_tried) {
    if (n_tried) *n_tried = 0;
    return autotune_impl(params, holo, holo_loc, iterations, best, tried, capacity, n_tried);
}

rihvr_status rihvr_reconstruct_multi(rihvr_multi_item* items, int32_t n_items, int holo_loc, int vol_loc,
                                     int32_t autotune) {
    return multi_impl(items, n_items, holo_loc, vol_loc, autotune);
}

rihvr_status rihvr_destroy(rihvr_ctx* ctx) {
    if (!ctx) return RIHVR_OK;
    (void)hipStreamSynchronize(ctx->stream);
    if (ctx->aux) (void)hipStreamSynchronize(ctx->aux);
    release(ctx);
    delete ctx;
    return RIHVR_OK;
}

rihvr_status rihvr_reconstruct_learned(rihvr_ctx* ctx, rihvr_net* net, const float* holo, int holo_loc,
                                       double f_target, int32_t max_iterations, rihvr_learned_stats* stats) {
    if (!ctx || !net) return RIHVR_E_INVALID_ARGUMENT;
    return learned_impl(ctx, net, holo, holo_loc, f_target, max_iterations, stats);
}

rihvr_status rihvr_reconstruct(rihvr_ctx* ctx, const float* holo, int holo_loc, const float* warm, int warm_loc,
                               rihvr_stats* stats) {
    if (!ctx) return RIHVR_E_INVALID_ARGUMENT;
    return reconstruct_impl(ctx, holo, holo_loc, warm, warm_loc, stats);
}

rihvr_status rihvr_reconstruct_batch(rihvr_ctx* ctx, const float* holos, int holo_loc, int32_t n_frames,
                                     const float* warm_start, int warm_loc, int32_t warm_from_previous,
                                     const double* f_target, int32_t max_iterations, float* volumes, int vol_loc,
                                     rihvr_frame_stats* frame_stats) {
    if (!ctx) return RIHVR_E_INVALID_ARGUMENT;
    return batch_impl(ctx, holos, holo_loc, n_frames, warm_start, warm_loc, warm_from_previous, f_target,
                      max_iterations, volumes, vol_loc, frame_stats);
}

rihvr_status rihvr_begin(rihvr_ctx* ctx, const float* holo, int holo_loc, const float* warm_start, int warm_loc,
                         double* f_initial) {
    if (!ctx) return RIHVR_E_INVALID_ARGUMENT;
    return begin_impl(ctx, holo, holo_loc, warm_start, warm_loc, f_initial);
}

rihvr_status rihvr_iterate(rihvr_ctx* ctx, int32_t n_iter, rihvr_stats* stats) {
    if (!ctx) return RIHVR_E_INVALID_ARGUMENT;
    return iterate_impl(ctx, n_iter, stats, false);
}

rihvr_status rihvr_snapshot(rihvr_ctx* ctx) {
    if (!ctx) return RIHVR_E_INVALID_ARGUMENT;
    return snapshot_impl(ctx);
}

rihvr_status rihvr_finish(rihvr_ctx* ctx, rihvr_stats* stats) {
    if (!ctx) return RIHVR_E_INVALID_ARGUMENT;
    return finish_impl(ctx, stats, -1);
}

rihvr_status rihvr_set_timing(rihvr_ctx* ctx, int32_t enable) {
    if (!ctx) return RIHVR_E_INVALID_ARGUMENT;
    RIHVR_HIP(hipStreamSynchronize(ctx->stream));
    RIHVR_TRY(resolve_timing(ctx));
    for (int k = 0; k < K_COUNT; ++k) {
        ctx->tm.total_ms[k] = 0;
        ctx->tm.launches[k] = 0;
    }
    if (enable < 0 || enable >= 2 + K_COUNT) return RIHVR_E_INVALID_ARGUMENT;
    ctx->tm.on = enable != 0;
    ctx->tm.only = enable >= 2 ? enable - 2 : -1;
    return RIHVR_OK;
}

rihvr_status rihvr_get_timing(rihvr_ctx* ctx, rihvr_kernel_time* out, int32_t capacity, int32_t* count) {
    if (!ctx || !count) return RIHVR_E_INVALID_ARGUMENT;
    RIHVR_HIP(hipStreamSynchronize(ctx->stream));
    RIHVR_TRY(resolve_timing(ctx));
    *count = K_COUNT;
    if (!out) return RIHVR_OK;
    for (int k = 0; k < K_COUNT && k < capacity; ++k) {
        std::memset(out[k].name, 0, sizeof(out[k].name));
        std::strncpy(out[k].name, kKernelNames[k], sizeof(out[k].name) - 1);
        out[k].launches = ctx->tm.launches[k];
        out[k].total_ms = ctx->tm.total_ms[k];
    }
    return RIHVR_OK;
}

rihvr_status rihvr_get_params(const rihvr_ctx* ctx, rihvr_params* out) {
    if (!ctx || !out) return RIHVR_E_INVALID_ARGUMENT;
    *out = ctx->p;
    return RIHVR_OK;
}

rihvr_status rihvr_get_volume(rihvr_ctx* ctx, float* out, int out_loc) {
    if (!ctx) return RIHVR_E_INVALID_ARGUMENT;
    if (!ctx->have_result) return RIHVR_E_INVALID_STATE;
    return copy_out(ctx, out, ctx->out, out_loc, ctx->V);
}

rihvr_status rihvr_get_plane(rihvr_ctx* ctx, int32_t z, float* out, int out_loc) {
    if (!ctx) return RIHVR_E_INVALID_ARGUMENT;
    if (z < 0 || z >= ctx->Nz) return RIHVR_E_INVALID_ARGUMENT;
    if (!ctx->have_result) return RIHVR_E_INVALID_STATE;
    return copy_out(ctx, out, ctx->out + (size_t)z * ctx->N, out_loc, ctx->N);
}

rihvr_status rihvr_set_stepsize(rihvr_ctx* ctx, double step) {
    if (!ctx) return RIHVR_E_INVALID_ARGUMENT;
    ctx->force_step = true;
    ctx->forced_step = step;
    return RIHVR_OK;
}

rihvr_status rihvr_get_estimate(rihvr_ctx* ctx, float* estimate, float* residual, int out_loc) {
    if (!ctx) return RIHVR_E_INVALID_ARGUMENT;
    if (!ctx->have_result) return RIHVR_E_INVALID_STATE;
    RIHVR_TRY(spatial_estimate(ctx, ctx->hE));
    if (estimate) RIHVR_TRY(copy_out(ctx, estimate, ctx->hspat, out_loc, ctx->N));
    if (residual) RIHVR_TRY(copy_out(ctx, residual, ctx->resid, out_loc, ctx->N));
    return RIHVR_OK;
}

rihvr_status rihvr_forward(rihvr_ctx* ctx, const float* volume, int vol_loc, float* holo_out, int holo_loc) {
    if (!ctx || !volume || !holo_out) return RIHVR_E_INVALID_ARGUMENT;
    ctx->have_result = false;
    ctx->ss.active = false;
    float2* v = ctx->Yb[0];
    ctx->occ_valid[0] = false;
    RIHVR_TRY(copy_in(ctx, v, volume, vol_loc, ctx->V));
    RIHVR_HIP(hipMemsetAsync(ctx->meas, 0, ctx->N * sizeof(float2), ctx->stream));
    RIHVR_HIP(hipMemsetAsync(ctx->bspec, 0, ctx->N * sizeof(float2), ctx->stream));
    RIHVR_TRY(apply_forward(ctx, v, ctx->hE));
    RIHVR_TRY(spatial_estimate(ctx, ctx->hE));
    return copy_out(ctx, holo_out, ctx->hspat, holo_loc, ctx->N);
}

rihvr_status rihvr_adjoint(rihvr_ctx* ctx, const float* holo, int holo_loc, float* volume_out, int vol_loc) {
    if (!ctx || !holo || !volume_out) return RIHVR_E_INVALID_ARGUMENT;
    ctx->have_result = false;
    ctx->ss.active = false;
    RIHVR_TRY(copy_in(ctx, ctx->resid, holo, holo_loc, ctx->N));
    RIHVR_TRY(apply_adjoint(ctx, ctx->resid, ctx->G));
    return copy_out(ctx, volume_out, ctx->G, vol_loc, ctx->V);
}

rihvr_status rihvr_prox(rihvr_ctx* ctx, const float* x0, const float* grad, double stepsize, float* out, int loc) {
    if (!ctx || !x0 || !grad || !out) return RIHVR_E_INVALID_ARGUMENT;
    ctx->have_result = false;
    ctx->ss.active = false;
    ctx->occ_valid[1] = false;
    RIHVR_TRY(copy_in(ctx, ctx->Yb[1], x0, loc, ctx->V));
    RIHVR_TRY(copy_in(ctx, ctx->G, grad, loc, ctx->V));
    RIHVR_TRY(apply_prox(ctx, effective_mode(ctx), ctx->Yb[1], ctx->Yb[1], 0.0f, ctx->G, stepsize, ctx->Yb[0]));
    return copy_out(ctx, out, ctx->Yb[0], loc, ctx->V);
}

}  // extern "C"
