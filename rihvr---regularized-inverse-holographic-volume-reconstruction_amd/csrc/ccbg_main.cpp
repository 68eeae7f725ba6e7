// ccbg_main.cpp -- ccbg: correlation-based backgrounds, one per frame.
//
//   ccbg [params file]      (default params.txt)
//
// The reference's program (src/core/ccbg.cpp): a parameter file of "key = value" lines ('#'
// comments; keys inputFolder ./input, outputFolder ./backgrounds, filePattern im_%04d.tif,
// startIdx 1, endIdx 10, minFrames 5, maxFrames 100, resizePct 0.25, saveEnhanced false),
// frames <inputFolder>/<filePattern % idx> for idx = startIdx .. endIdx, and for each frame
// <outputFolder>/background_%04d.png (and enhanced_%04d.png with saveEnhanced), computed by
// rihvr_background_correlation on the GPU.  resizePct must be 1/k for an integer k (the box
// average INTER_AREA computes at an integer factor).  Errors print "Error: ..." and exit 1.
#include <sys/stat.h>

#include <algorithm>
#include <cctype>
#include <cmath>
#include <cstdio>
#include <fstream>
#include <map>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/rihvr.h"

namespace {

std::string trim(const std::string& s) {
    const size_t a = s.find_first_not_of(" \t\r\n");
    if (a == std::string::npos) return "";
    const size_t b = s.find_last_not_of(" \t\r\n");
    return s.substr(a, b - a + 1);
}

std::map<std::string, std::string> read_params(const std::string& path) {
    std::ifstream f(path);
    if (!f) throw std::runtime_error("Cannot open params file: " + path);
    std::map<std::string, std::string> p;
    std::string line;
    while (std::getline(f, line)) {
        line = trim(line);
        if (line.empty() || line[0] == '#') continue;
        const size_t eq = line.find('=');
        if (eq == std::string::npos) continue;
        p[trim(line.substr(0, eq))] = trim(line.substr(eq + 1));
    }
    return p;
}

void make_dirs(const std::string& path) {
    std::string cur;
    for (size_t i = 0; i <= path.size(); ++i) {
        if (i == path.size() || path[i] == '/') {
            if (!cur.empty()) mkdir(cur.c_str(), 0755);
        }
        if (i < path.size()) cur += path[i];
    }
}

std::string format(const std::string& pattern, int idx) {
    std::vector<char> buf(pattern.size() + 64);
    std::snprintf(buf.data(), buf.size(), pattern.c_str(), idx);
    return std::string(buf.data());
}

}  // namespace

int main(int argc, char** argv) {
    try {
        const std::string pfile = argc > 1 ? argv[1] : "params.txt";
        auto p = read_params(pfile);
        auto get = [&](const char* k, const char* d) { return p.count(k) ? p[k] : std::string(d); };
        const std::string in_dir = get("inputFolder", "./input");
        const std::string out_dir = get("outputFolder", "./backgrounds");
        const std::string pattern = get("filePattern", "im_%04d.tif");
        const int start = std::stoi(get("startIdx", "1"));
        const int end = std::stoi(get("endIdx", "10"));
        const int min_frames = std::stoi(get("minFrames", "5"));
        const int max_frames = std::stoi(get("maxFrames", "100"));
        const float pct = std::stof(get("resizePct", "0.25"));
        std::string se = get("saveEnhanced", "false");
        std::transform(se.begin(), se.end(), se.begin(), ::tolower);
        const bool enhanced = se == "1" || se == "true" || se == "yes";
        const int k = pct >= 1.0f ? 1 : (int)std::lround(1.0 / pct);
        if (k < 1 || std::fabs(1.0 / k - pct) > 1e-6) throw std::runtime_error("resizePct must be 1/k for an integer k");
        make_dirs(out_dir);
        const int n = end - start + 1;
        if (n < 2) throw std::runtime_error("at least two frames are needed");
        int32_t w = 0, h = 0;
        std::vector<float> frames;
        for (int i = 0; i < n; ++i) {
            const std::string path = format(in_dir + "/" + pattern, start + i);
            int32_t fw = 0, fh = 0;
            if (rihvr_read_image(path.c_str(), nullptr, &fw, &fh) != RIHVR_OK)
                throw std::runtime_error("Failed to load image: " + path);
            if (i == 0) {
                w = fw;
                h = fh;
                frames.resize((size_t)n * w * h);
            } else if (fw != w || fh != h) {
                throw std::runtime_error("Size mismatch");
            }
            if (rihvr_read_image(path.c_str(), frames.data() + (size_t)i * w * h, &fw, &fh) != RIHVR_OK)
                throw std::runtime_error("Failed to load image: " + path);
        }
        std::vector<uint8_t> bg((size_t)n * w * h), enh(enhanced ? (size_t)n * w * h : 0);
        std::vector<int32_t> used(n);
        const rihvr_status st = rihvr_background_correlation(frames.data(), n, w, h, RIHVR_HOST, min_frames, max_frames,
                                                             k, bg.data(), enhanced ? enh.data() : nullptr, used.data());
        if (st != RIHVR_OK) throw std::runtime_error(std::string("background: ") + rihvr_status_string(st));
        for (int i = 0; i < n; ++i) {
            const std::string bp = format(out_dir + "/background_%04d.png", start + i);
            if (rihvr_write_png8(bp.c_str(), bg.data() + (size_t)i * w * h, w, h) != RIHVR_OK)
                throw std::runtime_error("Cannot write " + bp);
            if (enhanced) {
                const std::string ep = format(out_dir + "/enhanced_%04d.png", start + i);
                if (rihvr_write_png8(ep.c_str(), enh.data() + (size_t)i * w * h, w, h) != RIHVR_OK)
                    throw std::runtime_error("Cannot write " + ep);
            }
            std::printf("Saved background for image %d using %d frames%s\n", start + i, used[i],
                        enhanced ? " + enhanced image" : "");
        }
        return 0;
    } catch (const std::exception& e) {
        std::fprintf(stderr, "Error: %s\n", e.what());
        return 1;
    }
}
