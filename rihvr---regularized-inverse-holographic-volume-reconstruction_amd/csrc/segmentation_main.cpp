// segmentation_main.cpp -- segmentation: objects and centroids of a reconstructed volume.
//
//   segmentation -F <parameter file>
//
// The reference's executable (src/core/main_segmentation.cu:36-138): the volume size is
// that of the start image after crop and zero padding; the voxels are read from
// <output_path>/inverse_voxels_%04d.csv (start_image), binarized with
// segment_min_intensity (5010), closed with segment_close_size (5011), labelled, turned
// into objects, filtered by segment_min_area (5009) and written to
// <output_path>/centroids_%04d.csv, centroids_weighted_%04d.csv and
// centroids_verbose_%04d.csv.  The messages are the reference's; a HoloError is printed
// as "Detected HoloError: N" and the exit code is -1.
#include <getopt.h>

#include <chrono>
#include <cstdio>
#include <vector>

#include "../../include/rihvr.h"

namespace {

struct HoloError {
    rihvr_status code;
};

void check(rihvr_status s) {
    if (s != RIHVR_OK) throw HoloError{s};
}

struct Segmentation {
    rihvr_segmentation* p = nullptr;
    ~Segmentation() { rihvr_segmentation_destroy(p); }
};

int run(int argc, char* argv[]) {
    std::printf("Begining sparse inverse reconstruction\n");  // the reference's banner (:42)
    const auto t0 = std::chrono::steady_clock::now();
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

    // Hologram::read of the start image, crop, zeroPad: the volume's width and height
    char fn[RIHVR_PATH_MAX + 64];
    std::snprintf(fn, sizeof(fn), fp.image_filename, fp.start_image);
    int32_t iw = 0, ih = 0;
    if (rihvr_read_image(fn, nullptr, &iw, &ih) != RIHVR_OK) {
        std::printf("Unable to read file: %s\n", fn);
        throw HoloError{RIHVR_E_BAD_FILENAME};
    }
    std::vector<float> image((size_t)iw * ih);
    check(rihvr_read_image(fn, image.data(), &iw, &ih));
    int32_t W = 0, H = 0;
    check(rihvr_prepare_hologram(&fp, image.data(), nullptr, iw, ih, nullptr, &W, &H));
    const int32_t Nz = fp.solver.num_planes;

    std::snprintf(fn, sizeof(fn), "%s/inverse_voxels_%04d.csv", fp.output_path, fp.start_image);
    int64_t count = 0;
    rihvr_status s = rihvr_read_sparse_voxels(fn, W, H, Nz, &count, nullptr, nullptr, nullptr);
    if (s == RIHVR_E_BAD_FILENAME) std::printf("Unable to open file: %s\n", fn);
    if (s == RIHVR_E_INVALID_FILE) std::printf("SparseVolume::loadData Error:\nFile <%s> is not correctly formatted\n", fn);
    check(s);
    std::printf("Read first line data: x, y, z, value\n");
    std::vector<int32_t> plane((size_t)count), index((size_t)count);
    std::vector<float> intensity((size_t)count);
    check(rihvr_read_sparse_voxels(fn, W, H, Nz, &count, plane.data(), index.data(), intensity.data()));

    Segmentation seg;
    check(rihvr_segmentation_create(W, H, Nz, &seg.p));
    check(rihvr_segmentation_load(seg.p, count, plane.data(), index.data(), intensity.data(), RIHVR_HOST));
    std::printf("binarizing volume...\n");
    double mx = 0, thr = 0;
    check(rihvr_segmentation_binarize(seg.p, fp.segment_min_intensity, &mx, &thr));
    std::printf("max intensity = %f, thr = %f\n", mx, thr);
    std::printf("dilate...\n");
    std::printf("erode...\n");
    check(rihvr_segmentation_close(seg.p, fp.segment_close_size));
    std::printf("labeling connected components...\n");
    check(rihvr_segmentation_label(seg.p));
    std::printf("extracting and saving object centroids\n");
    int64_t n_objects = 0, removed = 0;
    check(rihvr_segmentation_extract(seg.p, &n_objects));
    std::printf("Number of particles: %d\n", (int)n_objects);
    check(rihvr_segmentation_size_filter(seg.p, fp.segment_min_area, &removed));
    std::printf("blobSizeFilter removed %d objects\n", (int)removed);

    const char* names[3] = {"centroids", "centroids_weighted", "centroids_verbose"};
    for (int k = 0; k < 3; ++k) {
        std::snprintf(fn, sizeof(fn), "%s/%s_%04d.csv", fp.output_path, names[k], fp.start_image);
        s = rihvr_segmentation_write_centroids(seg.p, fn, k == 1, k == 2);
        if (s == RIHVR_E_BAD_FILENAME)
            std::printf("ObjectCloud::writeCentroids: Error: Unable to open file: %s\n", fn);
        check(s);
    }
    const double ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
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
