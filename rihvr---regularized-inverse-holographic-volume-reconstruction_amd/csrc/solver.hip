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
#include <cmath>
#include <cstdio>
#include <cstring>
#include <vector>

#include "../../include/rihvr.h"
#include "kernels.hpp"

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

}  // namespace

struct rihvr_ctx {
    rihvr_params p{};
    int W = 0, H = 0, Nz = 0;
    long N = 0, V = 0;
    int device = 0;
    hipStream_t stream = nullptr;
    // volumes
    float2* Yb[3] = {nullptr, nullptr, nullptr};  // ring of prox outputs
    int iy = 0, iy0 = 1, iym = 2;                 // y_k, y_{k-1}, y_{k-2}
    float2* G = nullptr;                          // A^T r
    float2* T = nullptr;                          // FFT scratch volume
    float2* out = nullptr;                        // points at the returned volume
    // holograms
    float2 *meas = nullptr, *hE = nullptr, *hC = nullptr, *hP = nullptr, *resid = nullptr, *rhat = nullptr,
           *htmp = nullptr, *spart = nullptr;
    // tables
    double* kappa = nullptr;
    float* zlist = nullptr;
    float2* phase = nullptr;
    float2 *twfW = nullptr, *twiW = nullptr, *twfH = nullptr, *twiH = nullptr;
    // reductions
    double* partial = nullptr;
    double* partial_holo = nullptr;
    long partial_rows = 0, holo_rows = 0;
    double* scal = nullptr;
    double* h_scal = nullptr;
    int* dflag = nullptr;
    int zchunk_adj = 8, zchunk_fwd = 8, nparts = 1;
    bool have_result = false;
    bool force_step = false;
    double forced_step = 0;
};

namespace {

rihvr_status validate(const rihvr_params* p) {
    if (!p) return RIHVR_E_INVALID_ARGUMENT;
    if (p->width <= 0 || p->height <= 0 || p->num_planes <= 0) return RIHVR_E_INVALID_ARGUMENT;
    // the reference launches W/16 x H/16 blocks of 16x16 (sparse_compressive_holo.cu:125-126)
    if (p->width % 16 || p->height % 16) return RIHVR_E_INVALID_ARGUMENT;
    if (!(p->wavelength > 0) || !(p->resolution > 0)) return RIHVR_E_INVALID_ARGUMENT;
    if (p->num_inverse_iterations < 0 || p->num_tv_iterations < 0) return RIHVR_E_INVALID_ARGUMENT;
    if (p->mode == RIHVR_MODE_FASTA_FUSED_LASSO) return RIHVR_E_UNKNOWN_MODE;
    if (p->mode != RIHVR_MODE_FASTA_L1 && p->mode != RIHVR_MODE_FASTA_FUSED_LASSO_2D)
        return RIHVR_E_UNKNOWN_MODE;
    if (!fft_length_supported(p->width) || !fft_length_supported(p->height)) return RIHVR_E_INVALID_ARGUMENT;
    if (p->mode == RIHVR_MODE_FASTA_FUSED_LASSO_2D && !prox_fl2d_supported(p->num_tv_iterations))
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

rihvr_status fetch_scalars(rihvr_ctx* c) {
    RIHVR_HIP(hipMemcpyAsync(c->h_scal, c->scal, SL_COUNT * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    RIHVR_HIP(hipStreamSynchronize(c->stream));
    return RIHVR_OK;
}

// est = A vol, resid = est - meas, scal[SL_RESID_A] = sum |resid|^2
// (backPropagate :1935-2077 + calcResidual :2676-2723 + calcSumMagnitude)
rihvr_status apply_forward(rihvr_ctx* c, const float2* vol, float2* est) {
    const float invN = (float)(1.0 / (double)c->N);
    RIHVR_HIP(fft_rows_fwd(c->W, vol, c->phase, c->T, c->H, c->Nz, c->twfW, c->stream));
    RIHVR_HIP(fft_cols_fwdacc(c->H, c->T, c->kappa, c->zlist, c->spart, c->W, c->Nz, c->zchunk_fwd, invN, c->twfH,
                              c->stream));
    RIHVR_HIP(fft_rows_sumparts(c->W, c->spart, c->nparts, c->N, c->htmp, c->H, c->twiW, c->stream));
    RIHVR_HIP(fft_cols_real(c->H, c->htmp, est, c->meas, c->resid, c->partial_holo, c->W, c->twiH, c->stream));
    RIHVR_HIP(reduce_partials(c->partial_holo, c->W / 16, 1, c->scal + SL_RESID_A, c->stream));
    return RIHVR_OK;
}

// G = A^T resid  (SparseGradient::reconstruct + getPlane, sparse_gradient.cu:47-165)
rihvr_status apply_adjoint(rihvr_ctx* c, const float2* r, float2* G) {
    const float invN = (float)(1.0 / (double)c->N);
    RIHVR_HIP(fft_rows_plain(c->W, -1, r, c->htmp, c->H, c->twfW, c->stream));
    RIHVR_HIP(fft_cols_plain(c->H, -1, c->htmp, c->rhat, c->W, 1, c->twfH, c->stream));
    RIHVR_HIP(fft_rows_adj(c->W, c->rhat, c->kappa, c->zlist, c->T, c->H, c->Nz, c->zchunk_adj, invN, c->twiW,
                           c->stream));
    RIHVR_HIP(fft_cols_adj(c->H, c->T, c->phase, G, c->W, c->Nz, c->twiH, c->stream));
    return RIHVR_OK;
}

int effective_mode(const rihvr_ctx* c) {
    // mode 7 becomes mode 5 when the TV weight is zero (:192-195)
    if (c->p.mode == RIHVR_MODE_FASTA_FUSED_LASSO_2D && c->p.regularization_TV_param == 0.0f)
        return RIHVR_MODE_FASTA_L1;
    return c->p.mode;
}

// y = prox(x0 - t g), x0 = Y0 + s (Y0 - YM); reductions into scal[SL_PROX..]
rihvr_status apply_prox(rihvr_ctx* c, int mode, const float2* Y0, const float2* YM, float s_mom, const float2* G,
                        double t, float2* Y) {
    ProxLaunch a;
    a.Y0 = Y0; a.YM = YM; a.G = G; a.Y = Y;
    a.W = c->W; a.H = c->H; a.nz = c->Nz;
    a.s_mom = s_mom; a.t = t;
    a.reg = (double)c->p.regularization_param;
    a.reg_tv = (double)c->p.regularization_TV_param;
    a.n_tv = c->p.num_tv_iterations;
    a.partial = c->partial;
    int rows;
    if (mode == RIHVR_MODE_FASTA_FUSED_LASSO_2D) {
        RIHVR_HIP(prox_fl2d(a, c->stream));
        rows = prox_fl2d_num_waves(c->W, c->H, c->Nz, a.n_tv);
    } else {
        RIHVR_HIP(prox_l1(a, c->stream));
        rows = prox_l1_num_blocks(c->W, c->H, c->Nz);
    }
    RIHVR_HIP(reduce_partials(c->partial, rows, PS_COUNT, c->scal + SL_PROX, c->stream));
    return RIHVR_OK;
}

double objective(const rihvr_ctx* c, int mode, double summag, double l1, double tv) {
    // calcObjectiveFunction (:2838-2899)
    double f = 0.5 * summag + (double)c->p.regularization_param * l1;
    if (mode == RIHVR_MODE_FASTA_FUSED_LASSO_2D) f = f + (double)c->p.regularization_TV_param * tv;
    return f;
}

void release(rihvr_ctx* c) {
    if (!c) return;
    for (auto& b : c->Yb) (void)hipFree(b);
    void* ptrs[] = {c->G, c->T, c->meas, c->hE, c->hC, c->hP, c->resid, c->rhat, c->htmp, c->spart, c->kappa,
                    c->zlist, c->phase, c->twfW, c->twiW, c->twfH, c->twiH, c->partial, c->partial_holo, c->scal,
                    c->dflag};
    for (void* q : ptrs) (void)hipFree(q);
    if (c->h_scal) (void)hipHostFree(c->h_scal);
    if (c->stream) (void)hipStreamDestroy(c->stream);
}

rihvr_status create_impl(const rihvr_params* params, rihvr_ctx** out) {
    RIHVR_TRY(validate(params));
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
    c->zchunk_fwd = std::min(c->Nz, 8);
    c->nparts = (c->Nz + c->zchunk_fwd - 1) / c->zchunk_fwd;
    c->zchunk_adj = std::min(c->Nz, 8);
    for (auto& b : c->Yb) RIHVR_TRY(dalloc(&b, c->V));
    RIHVR_TRY(dalloc(&c->G, c->V));
    RIHVR_TRY(dalloc(&c->T, c->V));
    float2** holos[] = {&c->meas, &c->hE, &c->hC, &c->hP, &c->resid, &c->rhat, &c->htmp};
    for (float2** h : holos) RIHVR_TRY(dalloc(h, c->N));
    RIHVR_TRY(dalloc(&c->spart, (size_t)c->N * c->nparts));
    RIHVR_TRY(dalloc(&c->kappa, c->N));
    RIHVR_TRY(dalloc(&c->zlist, c->Nz));
    RIHVR_TRY(dalloc(&c->phase, c->Nz));
    RIHVR_TRY(dalloc(&c->twfW, c->W));
    RIHVR_TRY(dalloc(&c->twiW, c->W));
    RIHVR_TRY(dalloc(&c->twfH, c->H));
    RIHVR_TRY(dalloc(&c->twiH, c->H));
    long rows = std::max<long>(prox_fl2d_num_waves(c->W, c->H, c->Nz, std::min(params->num_tv_iterations, 10)),
                               prox_l1_num_blocks(c->W, c->H, c->Nz));
    rows = std::max<long>(rows, momentum_out_num_blocks(c->W, c->H, c->Nz));
    c->partial_rows = rows;
    RIHVR_TRY(dalloc(&c->partial, (size_t)rows * PS_COUNT));
    c->holo_rows = std::max<long>(holo_num_blocks(c->N), c->W / 16);
    RIHVR_TRY(dalloc(&c->partial_holo, c->holo_rows));
    RIHVR_TRY(dalloc(&c->scal, SL_COUNT));
    RIHVR_TRY(dalloc(&c->dflag, 4));
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
    std::vector<float2> ph(c->Nz);
    for (int i = 0; i < c->Nz; ++i) {
        float step = (float)i * params->plane_stepsize;
        z[i] = params->start_plane + step;
        float arg = (float)(2 * kPi * (double)z[i] / lambda);
        ph[i] = make_float2((float)std::cos((double)arg), (float)std::sin((double)arg));
    }
    RIHVR_TRY(upload(c->kappa, kap.data(), sizeof(double) * c->N, c->stream));
    RIHVR_TRY(upload(c->zlist, z.data(), sizeof(float) * c->Nz, c->stream));
    RIHVR_TRY(upload(c->phase, ph.data(), sizeof(float2) * c->Nz, c->stream));
    auto tfw = twiddles(c->W, -1), tiw = twiddles(c->W, +1), tfh = twiddles(c->H, -1), tih = twiddles(c->H, +1);
    RIHVR_TRY(upload(c->twfW, tfw.data(), sizeof(float2) * c->W, c->stream));
    RIHVR_TRY(upload(c->twiW, tiw.data(), sizeof(float2) * c->W, c->stream));
    RIHVR_TRY(upload(c->twfH, tfh.data(), sizeof(float2) * c->H, c->stream));
    RIHVR_TRY(upload(c->twiH, tih.data(), sizeof(float2) * c->H, c->stream));
    c->out = c->Yb[0];
    return RIHVR_OK;
}

// inverseReconstruct (:173-633), modes 5 and 7
rihvr_status reconstruct_impl(rihvr_ctx* c, const float* holo, int holo_loc, const float* warm, int warm_loc,
                              rihvr_stats* st) {
    if (!holo) return RIHVR_E_INVALID_ARGUMENT;
    const int mode = effective_mode(c);
    const int n_it = c->p.num_inverse_iterations;
    RIHVR_TRY(copy_in(c, c->meas, holo, holo_loc, c->N));
    RIHVR_HIP(hipMemsetAsync(c->dflag, 0, sizeof(int), c->stream));
    RIHVR_HIP(check_finite(c->meas, c->N, c->dflag, c->stream));
    int hflag = 0;
    RIHVR_HIP(hipMemcpyAsync(&hflag, c->dflag, sizeof(int), hipMemcpyDeviceToHost, c->stream));
    RIHVR_HIP(hipStreamSynchronize(c->stream));
    if (hflag) return RIHVR_E_INVALID_DATA;

    // start point x1 = x0 = y0 = warm start or zero (:258-259)
    c->iy = 0; c->iy0 = 1; c->iym = 2;
    float2* Y0 = c->Yb[c->iy0];
    float2* YM = c->Yb[c->iym];
    if (warm) RIHVR_TRY(copy_in(c, Y0, warm, warm_loc, c->V));
    else RIHVR_HIP(hipMemsetAsync(Y0, 0, c->V * sizeof(float2), c->stream));
    RIHVR_HIP(hipMemcpyAsync(YM, Y0, c->V * sizeof(float2), hipMemcpyDeviceToDevice, c->stream));
    RIHVR_HIP(hipMemsetAsync(c->hC, 0, c->N * sizeof(float2), c->stream));
    RIHVR_HIP(hipMemsetAsync(c->hP, 0, c->N * sizeof(float2), c->stream));

    double t = 1.0 / (double)c->Nz;  // :247 (any forced step is overridden there too)
    // initial residual and gradient (:273-286)
    RIHVR_TRY(apply_forward(c, Y0, c->hE));
    RIHVR_TRY(fetch_scalars(c));
    double prev_summag = c->h_scal[SL_RESID_A];
    double prev_f_resid = 0.5 * prev_summag;
    RIHVR_TRY(apply_adjoint(c, c->resid, c->G));

    double alpha1 = 1.0;
    float s_mom = 0.0f;  // x0 = Y0 for the first prox
    int total_bt = 0;
    double f_last = 0;
    if (st) st->iterations = 0;
    for (int it = 0; it < n_it; ++it) {
        double alpha0 = alpha1;
        int bt = 0;
        double ratio = 0, lim = 0;
        float2* Yn = c->Yb[c->iy];
        Y0 = c->Yb[c->iy0];
        YM = c->Yb[c->iym];
        for (;;) {
            RIHVR_TRY(apply_prox(c, mode, Y0, YM, s_mom, c->G, t, Yn));
            RIHVR_TRY(apply_forward(c, Yn, c->hE));
            RIHVR_TRY(fetch_scalars(c));
            const double* sc = c->h_scal;
            if (bt == 0) {
                if (it > 0) {
                    // residual norm at the previous momentum point (:545-549)
                    prev_summag = sc[SL_RESID_M];
                    prev_f_resid = 0.5 * prev_summag;
                    if (st && st->f_resid) st->f_resid[it - 1] = prev_f_resid;
                }
                // objective at the previous momentum point (the x0 of this prox)
                double f = objective(c, mode, prev_summag, sc[SL_PROX + PS_L1X0], sc[SL_PROX + PS_TVX0]);
                if (it == 0) {
                    if (st) st->f_initial = f;
                } else if (st && st->f) {
                    st->f[it - 1] = f;
                }
            }
            lim = sc[SL_PROX + PS_LS1] + sc[SL_PROX + PS_LS2] / (2 * t);  // :4381
            double f_resid = 0.5 * sc[SL_RESID_A];
            ratio = f_resid / (prev_f_resid + lim);  // :403
            if (((ratio > 1.0) || (ratio < 0)) && (bt < 2)) {  // :423
                t *= 0.5;
                bt++;
                total_bt++;
                continue;
            }
            break;
        }
        const double rs = c->h_scal[SL_PROX + PS_RESTART];
        const bool restart = rs > 0;  // :4481
        if (restart) alpha0 = 1;
        alpha1 = (1.0 + std::sqrt(1.0 + 4.0 * alpha0 * alpha0)) / 2.0;  // :476
        const double step = (alpha0 - 1) / alpha1;                        // :481
        // hologram momentum by linearity (:494-498): prev <- cur, cur <- est, est <- cur + s (cur - prev)
        std::swap(c->hP, c->hC);  // hP = old cur
        std::swap(c->hC, c->hE);  // hC = A y_k, hE = scratch
        RIHVR_HIP(holo_momentum(c->hC, c->hP, step, c->meas, c->hE, c->resid, c->N, c->partial_holo, c->stream));
        RIHVR_HIP(reduce_partials(c->partial_holo, holo_num_blocks(c->N), 1, c->scal + SL_RESID_M, c->stream));
        // the next gradient (:545-546)
        RIHVR_TRY(apply_adjoint(c, c->resid, c->G));
        // rotate the prox-output ring: y_{k-2} <- y_{k-1} <- y_k
        int old_m = c->iym;
        c->iym = c->iy0;
        c->iy0 = c->iy;
        c->iy = old_m;
        s_mom = (float)step;
        if (st) {
            st->iterations = it + 1;
            if (st->ratio) st->ratio[it] = ratio;
            if (st->lim) st->lim[it] = lim;
            if (st->backtracks) st->backtracks[it] = bt;
            if (st->restarts) st->restarts[it] = restart ? 1 : 0;
            if (st->restart_sum) st->restart_sum[it] = rs;
        }
    }
    // returned volume: x1 = y_K + s (y_K - y_{K-1})  (:483-489)
    float2* outb = c->Yb[c->iy];
    if (n_it > 0) {
        RIHVR_HIP(momentum_out(c->Yb[c->iy0], c->Yb[c->iym], s_mom, outb, c->W, c->H, c->Nz, c->partial, c->stream));
        RIHVR_HIP(reduce_partials(c->partial, momentum_out_num_blocks(c->W, c->H, c->Nz), PS_COUNT,
                                  c->scal + SL_PROX, c->stream));
    } else {
        RIHVR_HIP(hipMemcpyAsync(outb, c->Yb[c->iy0], c->V * sizeof(float2), hipMemcpyDeviceToDevice, c->stream));
        RIHVR_HIP(momentum_out(c->Yb[c->iy0], c->Yb[c->iy0], 0.0f, outb, c->W, c->H, c->Nz, c->partial, c->stream));
        RIHVR_HIP(reduce_partials(c->partial, momentum_out_num_blocks(c->W, c->H, c->Nz), PS_COUNT,
                                  c->scal + SL_PROX, c->stream));
    }
    RIHVR_TRY(fetch_scalars(c));
    const double* sc = c->h_scal;
    if (n_it > 0) {
        prev_summag = sc[SL_RESID_M];
        if (st && st->f_resid) st->f_resid[n_it - 1] = 0.5 * prev_summag;
    }
    f_last = objective(c, mode, prev_summag, sc[SL_PROX + PS_L1X0], sc[SL_PROX + PS_TVX0]);
    c->out = outb;
    c->have_result = true;
    if (st) {
        if (n_it > 0 && st->f) st->f[n_it - 1] = f_last;
        if (n_it == 0) st->f_initial = f_last;
        st->total_backtracks = total_bt;
        st->stepsize = t;
        st->f_final = f_last;
        st->nnz = (int64_t)llround(sc[SL_PROX + PS_NNZX0]);
        st->sparsity = (double)st->nnz / (double)c->V;
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
    rihvr_status s = create_impl(params, &c);
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

rihvr_status rihvr_destroy(rihvr_ctx* ctx) {
    if (!ctx) return RIHVR_OK;
    (void)hipStreamSynchronize(ctx->stream);
    release(ctx);
    delete ctx;
    return RIHVR_OK;
}

rihvr_status rihvr_reconstruct(rihvr_ctx* ctx, const float* holo, int holo_loc, const float* warm, int warm_loc,
                               rihvr_stats* stats) {
    if (!ctx) return RIHVR_E_INVALID_ARGUMENT;
    ctx->have_result = false;
    return reconstruct_impl(ctx, holo, holo_loc, warm, warm_loc, stats);
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
    if (estimate) RIHVR_TRY(copy_out(ctx, estimate, ctx->hE, out_loc, ctx->N));
    if (residual) RIHVR_TRY(copy_out(ctx, residual, ctx->resid, out_loc, ctx->N));
    return RIHVR_OK;
}

rihvr_status rihvr_forward(rihvr_ctx* ctx, const float* volume, int vol_loc, float* holo_out, int holo_loc) {
    if (!ctx || !volume || !holo_out) return RIHVR_E_INVALID_ARGUMENT;
    ctx->have_result = false;
    float2* v = ctx->Yb[0];
    RIHVR_TRY(copy_in(ctx, v, volume, vol_loc, ctx->V));
    RIHVR_HIP(hipMemsetAsync(ctx->meas, 0, ctx->N * sizeof(float2), ctx->stream));
    RIHVR_TRY(apply_forward(ctx, v, ctx->hE));
    return copy_out(ctx, holo_out, ctx->hE, holo_loc, ctx->N);
}

rihvr_status rihvr_adjoint(rihvr_ctx* ctx, const float* holo, int holo_loc, float* volume_out, int vol_loc) {
    if (!ctx || !holo || !volume_out) return RIHVR_E_INVALID_ARGUMENT;
    ctx->have_result = false;
    RIHVR_TRY(copy_in(ctx, ctx->resid, holo, holo_loc, ctx->N));
    RIHVR_TRY(apply_adjoint(ctx, ctx->resid, ctx->G));
    return copy_out(ctx, volume_out, ctx->G, vol_loc, ctx->V);
}

rihvr_status rihvr_prox(rihvr_ctx* ctx, const float* x0, const float* grad, double stepsize, float* out, int loc) {
    if (!ctx || !x0 || !grad || !out) return RIHVR_E_INVALID_ARGUMENT;
    ctx->have_result = false;
    RIHVR_TRY(copy_in(ctx, ctx->Yb[1], x0, loc, ctx->V));
    RIHVR_TRY(copy_in(ctx, ctx->G, grad, loc, ctx->V));
    RIHVR_TRY(apply_prox(ctx, effective_mode(ctx), ctx->Yb[1], ctx->Yb[1], 0.0f, ctx->G, stepsize, ctx->Yb[0]));
    return copy_out(ctx, out, ctx->Yb[0], loc, ctx->V);
}

}  // extern "C"
