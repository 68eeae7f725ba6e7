// rihvr.hpp -- header-only C++ mirror of umnholo::SparseCompressiveHolo
// (src/core/sparse_compressive_holo.h:25-190) over the C ABI of rihvr.h.
//
// Same method names and argument meaning; errors are thrown as the reference
// throws them: a HoloError value (src/core/umnholo.h:132-146), or
// CUDA_ERROR_UNKNOWN's counterpart RIHVR_E_DEVICE for a device failure.
#ifndef RIHVR_HPP
#define RIHVR_HPP

#include <cstddef>
#include <vector>

#include "rihvr.h"

namespace rihvr {

// umnholo::HoloError numbering; DEVICE_ERROR has no HoloError counterpart.
enum HoloError {
    HOLO_ERROR_UNKNOWN_TYPE = 1,
    HOLO_ERROR_BAD_FILENAME = 2,
    HOLO_ERROR_MISSING_DATA = 3,
    HOLO_ERROR_INVALID_STATE = 4,
    HOLO_ERROR_OUT_OF_MEMORY = 5,
    HOLO_ERROR_UNKNOWN_MODE = 6,
    HOLO_ERROR_INVALID_ARGUMENT = 7,
    HOLO_ERROR_UNKNOWN_ERROR = 8,
    HOLO_ERROR_INVALID_FILE = 9,
    HOLO_ERROR_INVALID_DATA = 10,
    HOLO_ERROR_CRITICAL_ASSUMPTION = 11,
    HOLO_ERROR_INFINITE_LOOP = 12,
    DEVICE_ERROR = 13
};

enum CompressiveHoloMode {
    COMPRESSIVE_MODE_FASTA_L1 = RIHVR_MODE_FASTA_L1,
    COMPRESSIVE_MODE_FASTA_FUSED_LASSO = RIHVR_MODE_FASTA_FUSED_LASSO,
    COMPRESSIVE_MODE_FASTA_FUSED_LASSO_2D = RIHVR_MODE_FASTA_FUSED_LASSO_2D
};

inline void check(rihvr_status s) {
    if (s != RIHVR_OK) throw static_cast<HoloError>(s);
}

class SparseCompressiveHolo {
public:
    // SparseCompressiveHolo(Hologram) (:58-135): the hologram's size and the
    // parameter set are given by `params`; the mode used by inverseReconstruct is
    // params.mode (7 by default in the CLI).
    explicit SparseCompressiveHolo(const rihvr_params& params) : params_(params) {
        check(rihvr_create(&params_, &ctx_));
    }
    SparseCompressiveHolo(const SparseCompressiveHolo&) = delete;
    SparseCompressiveHolo& operator=(const SparseCompressiveHolo&) = delete;
    ~SparseCompressiveHolo() { destroy(); }

    void destroy() {
        if (ctx_) rihvr_destroy(ctx_);
        ctx_ = nullptr;
    }

    // inverseReconstruct(Hologram, CompressiveHoloMode) (:173-633).  holo: [H][W]
    // complex64 (CV_32FC2 data) at `loc`.  The mode must be the one the context was
    // created with (modes 5/7; 7 switches to 5 when the TV weight is 0, as in :192-195).
    rihvr_stats inverseReconstruct(const float* holo, CompressiveHoloMode mode, int loc = RIHVR_HOST,
                                   const float* warm_start = nullptr) {
        if ((int)mode != params_.mode) throw HOLO_ERROR_INVALID_ARGUMENT;
        const int n = params_.num_inverse_iterations > 0 ? params_.num_inverse_iterations : 1;
        f_.assign(n, 0.0);
        backtracks_.assign(n, 0);
        restarts_.assign(n, 0);
        rihvr_stats st{};
        st.f = f_.data();
        st.backtracks = backtracks_.data();
        st.restarts = restarts_.data();
        check(rihvr_reconstruct(ctx_, holo, loc, warm_start, loc, &st));
        return st;
    }

    // getPlane(size_t) (sparse_compressive_holo.h:92): plane z of the result, [H][W] complex64.
    std::vector<float> getPlane(size_t z) {
        std::vector<float> out(2 * (size_t)params_.width * params_.height);
        check(rihvr_get_plane(ctx_, (int32_t)z, out.data(), RIHVR_HOST));
        return out;
    }

    // The whole result volume, [Nz][H][W] complex64.
    std::vector<float> getVolume() {
        std::vector<float> out(2 * (size_t)params_.width * params_.height * params_.num_planes);
        check(rihvr_get_volume(ctx_, out.data(), RIHVR_HOST));
        return out;
    }

    // setStepsize(double) (sparse_compressive_holo.h:94)
    void setStepsize(double step) { check(rihvr_set_stepsize(ctx_, step)); }

    // per-iteration objective values of the last solve (the printed "new f", :561)
    const std::vector<double>& objectiveValues() const { return f_; }

    rihvr_ctx* handle() { return ctx_; }

private:
    rihvr_params params_;
    rihvr_ctx* ctx_ = nullptr;
    std::vector<double> f_;
    std::vector<int32_t> backtracks_, restarts_;
};

}  // namespace rihvr

#endif  // RIHVR_HPP
