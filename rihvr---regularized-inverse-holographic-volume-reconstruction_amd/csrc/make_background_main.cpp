// make_background_main.cpp -- make-background: background images for the
// reconstruction's background removal.
//
//   make-background -F <parameter file>
//
// The reference's executable (src/core/main_create_background.cu:196-274) with the
// file names, messages and formats it writes:
//   use_median_background (7006) set: the median of a window of five frames
//     (median_background_calc, :104-193) written to <output_path>/background_%04d.png,
//     with the reference's window order (frames start..start+4 for the first three
//     outputs; then frame start+i replaces stack slot (i-3) % 5 for output i);
//   otherwise: the mean background of HoloSequence::useBackgroundSubtraction
//     (holo_sequence.cu:309-437) over frames start + k*step (step = count/200 when
//     more than 500 frames), written to <output_path>/background.png, and its statistics
//     (alpha, beta, minimum, average; "%f" lines) to <output_path>/background_stats.txt.
// The reference lets a HoloError escape main (an abort); here it is printed as
// "Detected HoloError: N" and the exit code is -1, as the reconstruction executable does.
#include <getopt.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <vector>

#include "../../include/rihvr.h"

namespace {

struct HoloError {
    rihvr_status code;
};

void check(rihvr_status s) {
    if (s != RIHVR_OK) throw HoloError{s};
}

// Hologram::read: the image at path, raw grey values; width/height must match when set
std::vector<float> read_frame(const char* path, int32_t* w, int32_t* h) {
    int32_t fw = 0, fh = 0;
    if (rihvr_read_image(path, nullptr, &fw, &fh) != RIHVR_OK) {
        std::printf("Unable to read file: %s\n", path);
        throw HoloError{RIHVR_E_BAD_FILENAME};
    }
    if (*w > 0 && (fw != *w || fh != *h)) throw HoloError{RIHVR_E_INVALID_ARGUMENT};
    std::vector<float> px((size_t)fw * fh);
    check(rihvr_read_image(path, px.data(), &fw, &fh));
    *w = fw;
    *h = fh;
    return px;
}

// Hologram::write of a LOADED-state hologram: convertTo(CV_8U) of the real part
// (round to nearest even, saturate), then the image file
void write_png(const char* path, const std::vector<float>& v, int32_t w, int32_t h) {
    std::vector<uint8_t> px(v.size());
    for (size_t i = 0; i < v.size(); ++i) {
        const double r = std::nearbyint((double)v[i]);
        px[i] = (uint8_t)(r < 0 ? 0 : (r > 255 ? 255 : r));
    }
    check(rihvr_write_png8(path, px.data(), w, h));
}

int median_background(const rihvr_file_params& fp) {
    char fn[RIHVR_PATH_MAX + 64];
    int32_t w = 0, h = 0;
    std::snprintf(fn, sizeof(fn), fp.image_filename, fp.start_image);
    read_frame(fn, &w, &h);  // the reference reads the start image first (:107-110)
    const long N = (long)w * h;
    const int num = fp.background_end - fp.background_start;
    if (num < 5) {
        std::printf("Median background calculation requires at least 5 frames\n");
        throw HoloError{RIHVR_E_CRITICAL_ASSUMPTION};
    }
    std::vector<float> stack(5 * (size_t)N), med((size_t)N);
    for (int i = 0; i < 5; ++i) {
        std::snprintf(fn, sizeof(fn), fp.image_filename, fp.background_start + i);
        std::vector<float> f = read_frame(fn, &w, &h);
        std::copy(f.begin(), f.end(), stack.begin() + (size_t)i * N);
    }
    check(rihvr_background_median5(stack.data(), w, h, RIHVR_HOST, med.data()));
    for (int i = 0; i < 3; ++i) {
        std::snprintf(fn, sizeof(fn), "%s/background_%04d.png", fp.output_path, fp.background_start + i);
        write_png(fn, med, w, h);
    }
    std::printf("Computing background for image %04d of %04d\n", 3, num);
    for (int i = 3; i < num - 2; ++i) {
        std::printf("Computing background for image %04d of %04d\n", i, num);
        std::snprintf(fn, sizeof(fn), fp.image_filename, fp.background_start + i);
        std::vector<float> f = read_frame(fn, &w, &h);
        std::copy(f.begin(), f.end(), stack.begin() + (size_t)((i - 3) % 5) * N);
        check(rihvr_background_median5(stack.data(), w, h, RIHVR_HOST, med.data()));
        std::snprintf(fn, sizeof(fn), "%s/background_%04d.png", fp.output_path, fp.background_start + i);
        write_png(fn, med, w, h);
    }
    return 0;
}

int mean_background(const rihvr_file_params& fp) {
    char fn[RIHVR_PATH_MAX + 64];
    int32_t w = 0, h = 0;
    // HoloSequence::findLargestCount reads the start image for its size (holo_sequence.cu:58-61)
    std::snprintf(fn, sizeof(fn), fp.image_filename, fp.start_image);
    read_frame(fn, &w, &h);
    const long N = (long)w * h;
    // frames of useBackgroundSubtraction (:332-352) and loadNextChunk (:138-160)
    int start = fp.start_image, total = fp.image_count;
    if (fp.background_start >= 0 && fp.background_end >= 0) {
        start = fp.background_start;
        total = fp.background_end - fp.background_start - 1;
    }
    const int step = total > 500 ? total / 200 : 1;
    const int count = total > 0 ? total / step : 0;
    std::vector<float> frames((size_t)std::max(count, 1) * N);
    for (int k = 0; k < count; ++k) {
        std::snprintf(fn, sizeof(fn), fp.image_filename, start + k * step);
        std::vector<float> f = read_frame(fn, &w, &h);
        std::copy(f.begin(), f.end(), frames.begin() + (size_t)k * N);
    }
    std::vector<float> bg((size_t)N);
    rihvr_background_stats st{};
    check(rihvr_background_mean(frames.data(), count, w, h, RIHVR_HOST, bg.data(), &st));
    std::printf("Used %d images for background calculation\n", count);
    std::printf("  stepsize = %d\n", step);
    std::snprintf(fn, sizeof(fn), "%s/background.png", fp.output_path);
    write_png(fn, bg, w, h);
    std::snprintf(fn, sizeof(fn), "%s/background_stats.txt", fp.output_path);
    FILE* f = std::fopen(fn, "w");
    if (!f) {
        std::printf("Unable to open output file: %s\n", fn);
        return -1;
    }
    std::fprintf(f, "%f\n", st.alpha);
    std::fprintf(f, "%f\n", st.beta);
    std::fprintf(f, "%f\n", st.minimum);
    std::fprintf(f, "%f\n", st.average);
    std::fclose(f);
    return 0;
}

int run(int argc, char* argv[]) {
    std::printf("Begining background calculation\n");
    char params_filename[RIHVR_PATH_MAX] = {0};
    bool params_set = false;
    std::printf("Parsing command line inputs\n");
    int opt;
    while ((opt = getopt(argc, argv, "F:")) != -1) {
        switch (opt) {
            case 'F':
                std::snprintf(params_filename, sizeof(params_filename), "%s", optarg);
                params_set = true;
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
    if (rihvr_read_param_file(params_filename, &fp) == RIHVR_E_BAD_FILENAME)
        std::printf("Unable to read file <%s>\n", params_filename);
    return fp.use_median_background ? median_background(fp) : mean_background(fp);
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
