// cli_main.cpp -- sparse-inverse-recon: the reconstruction executable.
//
//   sparse-inverse-recon -F <parameter file> [-S 0|1]
//
// Same flow, file names and exit behaviour as the reference's executable
// (src/core/main_sparse_inverse.cu:43-198): read the parameter file, read image
// <image_filename % start_image>, rescale / normalise or remove the background,
// crop, zero-pad, write <output_path>/enh.tif, run the fused-lasso solve (mode 7;
// mode 5 when the TV weight is 0), then write the planes (output_planes; -S 1 puts
// them in <output_path>/<start_image %04d>/), <output_path>cmb_{xy,xz,yz}.tif and
// <output_path>/inverse_voxels.csv.  Inside the solve it appends to
// <output_path>/all_metrics.csv, writes per-iteration outputs when output_steps is
// set and writes estimate_final.tif / residual_final.tif
// (sparse_compressive_holo.cu:196-205, 571-597, 621-624).  Errors print
// "Detected HoloError: <code>" and exit with -1, as the reference.
#include <getopt.h>
#include <sys/stat.h>

#include <algorithm>
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "../../include/rihvr.h"

namespace {

struct HoloError {
    rihvr_status code;
};

void check(rihvr_status s) {
    if (s != RIHVR_OK) throw HoloError{s};
}

struct Ctx {
    rihvr_ctx* c = nullptr;
    ~Ctx() {
        if (c) rihvr_destroy(c);
    }
};

void min_max_real(const std::vector<float>& holo, double* lo, double* hi) {
    const size_t n = holo.size() / 2;
    *lo = *hi = holo[0];
    for (size_t i = 1; i < n; ++i) {
        *lo = std::min(*lo, (double)holo[2 * i]);
        *hi = std::max(*hi, (double)holo[2 * i]);
    }
}

int run(int argc, char* argv[]) {
    std::printf("Begining sparse inverse reconstruction\n");
    const auto start = std::chrono::steady_clock::now();
    char params_filename[RIHVR_PATH_MAX] = {0};
    bool params_set = false;
    bool save_planes_in_subfolder = false;
    std::printf("Parsing command line inputs\n");
    int opt;
    while ((opt = getopt(argc, argv, "F:S:")) != -1) {
        switch (opt) {
            case 'F':
                std::snprintf(params_filename, sizeof(params_filename), "%s", optarg);
                params_set = true;
                break;
            case 'S':
                save_planes_in_subfolder = std::atoi(optarg) != 0;
                std::printf("Save planes in subfolder flag set to %d\n", save_planes_in_subfolder ? 1 : 0);
                break;
            default:
                std::printf("Unknown input argument\n");
                return -1;
        }
    }
    if (!params_set) {
        std::printf("User must give input parameter file using -F flag\n");
        return -1;
    }
    std::printf("Parameters have been set, begining processing\n");

    static rihvr_file_params fp;
    // the reference ignores the parser's status: it keeps whatever was read
    const rihvr_status ps = rihvr_read_param_file(params_filename, &fp);
    if (ps == RIHVR_E_BAD_FILENAME) std::printf("Unable to read file <%s>\n", params_filename);

    char read_filename[RIHVR_PATH_MAX + 64];
    std::snprintf(read_filename, sizeof(read_filename), fp.image_filename, fp.start_image);
    int32_t w = 0, h = 0;
    if (rihvr_read_image(read_filename, nullptr, &w, &h) != RIHVR_OK) {
        std::printf("Unable to read file: %s\n", read_filename);
        throw HoloError{RIHVR_E_BAD_FILENAME};
    }
    std::vector<float> image((size_t)w * h), background;
    check(rihvr_read_image(read_filename, image.data(), &w, &h));
    if (fp.remove_background) {
        std::printf("Removing background\n");
        int32_t bw = 0, bh = 0;
        if (rihvr_read_image(fp.background_filename, nullptr, &bw, &bh) != RIHVR_OK) {
            std::printf("Unable to read file: %s\n", fp.background_filename);
            throw HoloError{RIHVR_E_BAD_FILENAME};
        }
        if ((long)bw * bh != (long)w * h) throw HoloError{RIHVR_E_INVALID_ARGUMENT};
        background.resize((size_t)bw * bh);
        check(rihvr_read_image(fp.background_filename, background.data(), &bw, &bh));
    } else {
        std::printf("Rescaling hologram\n");
    }
    int32_t W = 0, H = 0;
    const float* bg = fp.remove_background ? background.data() : nullptr;
    check(rihvr_prepare_hologram(&fp, image.data(), bg, w, h, nullptr, &W, &H));
    std::vector<float> holo(2 * (size_t)W * H);
    check(rihvr_prepare_hologram(&fp, image.data(), bg, w, h, holo.data(), &W, &H));

    char fn[RIHVR_PATH_MAX + 64];
    // Hologram::write does not check cv::imwrite's result: neither do these writes
    std::snprintf(fn, sizeof(fn), "%s/enh.tif", fp.output_path);
    rihvr_write_hologram(fn, holo.data(), W, H, 1, 0.0, 0.0);

    rihvr_params p = fp.solver;
    p.width = W;
    p.height = H;
    p.mode = RIHVR_MODE_FASTA_FUSED_LASSO_2D;
    Ctx ctx;
    check(rihvr_create(&p, &ctx.c));

    // --- inverseReconstruct (sparse_compressive_holo.cu:183-628) ---
    double min_measured = 0, max_measured = 0;
    min_max_real(holo, &min_measured, &max_measured);
    std::snprintf(fn, sizeof(fn), "%s/all_metrics.csv", fp.output_path);
    FILE* metrics = std::fopen(fn, "a");
    if (!metrics) {
        std::printf("Unable to open file: %s\n", fn);
        throw HoloError{RIHVR_E_INVALID_FILE};
    }
    std::fprintf(metrics, "it, f, sparsity, stepsize, total_backtracks\n");
    std::printf("before recon.inverseReconstruct\n");
    const int n_it = p.num_inverse_iterations;
    std::vector<float> est(2 * (size_t)W * H), res(2 * (size_t)W * H);
    rihvr_status st = RIHVR_OK;
    rihvr_stats fin{};
    if (!fp.output_steps) {
        // one call for the whole solve; the per-iteration values come back in arrays
        const size_t n = (size_t)std::max(1, n_it);
        std::vector<double> f(n), fr(n), ratio(n), lim(n), rsum(n), sp(n);
        std::vector<int32_t> bt(n), rs(n);
        fin.f = f.data(); fin.f_resid = fr.data(); fin.ratio = ratio.data(); fin.lim = lim.data();
        fin.backtracks = bt.data(); fin.restarts = rs.data(); fin.restart_sum = rsum.data();
        fin.sparsity_it = sp.data();
        st = rihvr_reconstruct(ctx.c, holo.data(), RIHVR_HOST, nullptr, RIHVR_HOST, &fin);
        int total_bt = 0;
        for (int it = 0; it < n_it && st == RIHVR_OK; ++it) {
            total_bt += bt[it];
            std::printf("it %d: new f = %e, backtracks %d, sparsity %g\n", it, f[it], total_bt, sp[it]);
        }
        fin.f = fin.f_resid = fin.ratio = fin.lim = fin.restart_sum = fin.sparsity_it = nullptr;
        fin.backtracks = fin.restarts = nullptr;
    } else {
        // stepwise, writing the per-iteration outputs of the momentum point x1
        double f1 = 0, fr = 0, ratio = 0, lim = 0, rsum = 0, sp = 0;
        int32_t bt = 0, rs = 0;
        rihvr_stats s{};
        s.f = &f1; s.f_resid = &fr; s.ratio = &ratio; s.lim = &lim; s.backtracks = &bt; s.restarts = &rs;
        s.restart_sum = &rsum; s.sparsity_it = &sp;
        st = rihvr_begin(ctx.c, holo.data(), RIHVR_HOST, nullptr, RIHVR_HOST, nullptr);
        for (int it = 0; it < n_it && st == RIHVR_OK; ++it) {
            if ((st = rihvr_iterate(ctx.c, 1, &s))) break;
            std::printf("it %d: new f = %e, backtracks %d, stepsize %g, sparsity %g\n", it, f1, s.total_backtracks,
                        s.stepsize, sp);
            if ((st = rihvr_snapshot(ctx.c))) break;
            char name[64];
            std::snprintf(name, sizeof(name), "cmb_iter%03d_end", it);
            if ((st = rihvr_save_projections(ctx.c, fp.output_path, name, 1))) break;
            if ((st = rihvr_get_estimate(ctx.c, est.data(), res.data(), RIHVR_HOST))) break;
            std::snprintf(fn, sizeof(fn), "%s/estimate_iter%03d.tif", fp.output_path, it);
            rihvr_write_hologram(fn, est.data(), W, H, 0, min_measured, max_measured);
            std::snprintf(fn, sizeof(fn), "%s/residual_iter%03d.tif", fp.output_path, it);
            rihvr_write_hologram(fn, res.data(), W, H, 0, min_measured, max_measured);
            std::fprintf(metrics, "%d, %f, %f, %f, %d\n", it, f1, sp, s.stepsize, s.total_backtracks);
        }
        if (st == RIHVR_OK) st = rihvr_finish(ctx.c, &fin);
    }
    if (st == RIHVR_OK) st = rihvr_get_estimate(ctx.c, est.data(), res.data(), RIHVR_HOST);
    if (st == RIHVR_OK) {
        std::snprintf(fn, sizeof(fn), "%s/estimate_final.tif", fp.output_path);
        rihvr_write_hologram(fn, est.data(), W, H, 0, min_measured, max_measured);
        std::snprintf(fn, sizeof(fn), "%s/residual_final.tif", fp.output_path);
        rihvr_write_hologram(fn, res.data(), W, H, 0, min_measured, max_measured);
    }
    std::fclose(metrics);
    check(st);
    std::printf("after recon.inverseReconstruct: f = %e, sparsity %g, nnz %lld\n", fin.f_final, fin.sparsity,
                (long long)fin.nnz);

    if (fp.output_planes) {
        char prefix[RIHVR_PATH_MAX + 64];
        if (save_planes_in_subfolder) {
            char sub[RIHVR_PATH_MAX + 32];
            std::snprintf(sub, sizeof(sub), "%s/%04d/", fp.output_path, fp.start_image);
            mkdir(sub, 0755);
            std::snprintf(prefix, sizeof(prefix), "%s%s", sub, "inverse");
        } else {
            std::snprintf(prefix, sizeof(prefix), "%s/%s", fp.output_path, "inverse");
        }
        check(rihvr_save_planes(ctx.c, prefix));
    }
    check(rihvr_save_projections(ctx.c, fp.output_path, "cmb", 0));
    std::snprintf(fn, sizeof(fn), "%s/%s.csv", fp.output_path, "inverse_voxels");
    check(rihvr_save_sparse(ctx.c, fn));

    const double ms =
        std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - start).count();
    std::printf("Done. Total time: %g ms\n", ms);
    return 0;
}

}  // namespace

int main(int argc, char* argv[]) {
    try {
        return run(argc, argv);
    } catch (const HoloError& e) {
        std::printf("Detected HoloError: %d\n", (int)e.code);
        return -1;
    }
}
