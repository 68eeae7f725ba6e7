// io.cpp -- host-side input and output of the reconstruction: the parameter
// file, image decoding, the hologram input contract and the result writers.
// Built on the public C ABI only; no device code.
#include <algorithm>
#include <cctype>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/rihvr.h"
#include "io_internal.hpp"

namespace {

constexpr double kPi = 3.14159265358979323846;

// HoloParameter ids (src/core/holo_ui.h:22-78)
enum Id {
    NUM_IMAGES = 1001, START_IMAGE = 1002, BG_START = 1003, BG_END = 1004,
    INPUT_FN = 2001, INPUT_PATHN = 2002, INPUT_FORMATFN = 2003, OUTPUT_PATHN = 2004, INPUT_BACKGROUND = 2005,
    EXE_PATHN = 2006,
    START_PLANE = 3001, PLANE_STEP = 3002, NUM_PLANES = 3003, WAVELENGTH = 3004, RESOLUTION = 3005, ROI_X = 3006,
    ROI_Y = 3007, ROI_Z = 3008, ROI_SIZE = 3009, ZERO_PADDING = 3010, ROI_RECT_X = 3011, ROI_RECT_Y = 3012,
    ROI_RECT_W = 3013, ROI_RECT_H = 3014,
    BETA = 4001, IIPE_ITERATIONS = 4002,
    PARTICLE_DIAMETER = 5001, SIM_CHAMBER_RADIUS = 5002, SIM_CONCENTRATION = 5003, SIM_LENGTH = 5004,
    SIM_RESOLUTION = 5005, SIM_IMAGE_SIZE = 5006, SIM_ROTATION_STEP = 5007, SIM_OVERSAMPLE = 5008,
    SEGMENT_MIN_AREA = 5009, SEGMENT_MIN_INTENSITY = 5010, SEGMENT_CLOSE_SIZE = 5011,
    NUM_INVERSE_ITS = 6001, INVERSE_REG_TAU = 6002, INVERSE_MAX_SVD = 6003, INVERSE_REG_TV = 6004,
    NUM_TV_ITS = 6005,
    FLAG_OUTPUT_PLANES = 7001, FLAG_STANDARD_RECON = 7002, FLAG_INVERSE_RECON = 7003, FLAG_DECONV_RECON = 7004,
    FLAG_OUTPUT_STEPS = 7005, FLAG_MEDIAN_BACKGROUND = 7006
};

void copy_str(char* dst, const char* src) {
    std::strncpy(dst, src, RIHVR_PATH_MAX - 1);
    dst[RIHVR_PATH_MAX - 1] = 0;
}

// a flag line: the float value, overridden by the words true/false (holo_ui.cpp:437-481)
int flag_value(float value, const char* text) {
    int v = value != 0.0f ? 1 : 0;
    if (std::strcmp(text, "true") == 0) v = 1;
    if (std::strcmp(text, "false") == 0) v = 0;
    return v;
}

// ---------------------------------------------------------------------------
// TIFF / PGM decoding
// ---------------------------------------------------------------------------
struct Reader {
    std::vector<uint8_t> b;
    bool le = true;
    bool ok(size_t off, size_t n) const { return off + n <= b.size(); }
    uint16_t u16(size_t o) const {
        return le ? (uint16_t)(b[o] | (b[o + 1] << 8)) : (uint16_t)((b[o] << 8) | b[o + 1]);
    }
    uint32_t u32(size_t o) const {
        return le ? (uint32_t)b[o] | ((uint32_t)b[o + 1] << 8) | ((uint32_t)b[o + 2] << 16) | ((uint32_t)b[o + 3] << 24)
                  : ((uint32_t)b[o] << 24) | ((uint32_t)b[o + 1] << 16) | ((uint32_t)b[o + 2] << 8) | (uint32_t)b[o + 3];
    }
};

bool read_file(const char* path, std::vector<uint8_t>& out) {
    FILE* f = std::fopen(path, "rb");
    if (!f) return false;
    std::fseek(f, 0, SEEK_END);
    long n = std::ftell(f);
    std::fseek(f, 0, SEEK_SET);
    if (n < 0) {
        std::fclose(f);
        return false;
    }
    out.resize((size_t)n);
    size_t got = n ? std::fread(out.data(), 1, (size_t)n, f) : 0;
    std::fclose(f);
    return got == (size_t)n;
}

// BGR -> gray of 8-bit data (cv::cvtColor COLOR_BGR2GRAY in Hologram::read):
// 0.299 R + 0.587 G + 0.114 B in 14-bit fixed point, rounded
float bgr_to_gray(uint8_t b, uint8_t g, uint8_t r) {
    return (float)((b * 1868 + g * 9617 + r * 4899 + (1 << 13)) >> 14);
}

// The image sizes cv::imread accepts by default (OpenCV's validateInputImageSize:
// CV_IO_MAX_IMAGE_WIDTH = CV_IO_MAX_IMAGE_HEIGHT = 2^20, CV_IO_MAX_IMAGE_PIXELS = 2^30);
// a header describing a larger image is refused, so a reported size always fits int32 and
// its pixel count fits the caller's buffer arithmetic.
constexpr uint64_t kMaxImageSide = 1ull << 20, kMaxImagePixels = 1ull << 30;
bool image_size_ok(uint64_t W, uint64_t H) {
    return W >= 1 && H >= 1 && W <= kMaxImageSide && H <= kMaxImageSide && W * H <= kMaxImagePixels;
}

rihvr_status decode_tiff(Reader& r, float* pixels, int32_t* width, int32_t* height) {
    if (!r.ok(0, 8)) return RIHVR_E_UNKNOWN_TYPE;
    r.le = r.b[0] == 'I';
    if (r.u16(2) != 42) return RIHVR_E_UNKNOWN_TYPE;
    size_t ifd = r.u32(4);
    if (!r.ok(ifd, 2)) return RIHVR_E_UNKNOWN_TYPE;
    const int n = r.u16(ifd);
    uint32_t W = 0, H = 0, bps = 8, spp = 1, comp = 1, rows_per_strip = 0xffffffffu, photometric = 1, planar = 1;
    std::vector<uint32_t> offsets, counts;
    auto values = [&](size_t e, std::vector<uint32_t>& out) {
        const uint16_t type = r.u16(e + 2);
        const uint32_t cnt = r.u32(e + 4);
        const size_t sz = type == 3 ? 2 : 4;
        size_t p = cnt * sz <= 4 ? e + 8 : r.u32(e + 8);
        out.clear();
        for (uint32_t i = 0; i < cnt && r.ok(p + i * sz, sz); ++i) out.push_back(sz == 2 ? r.u16(p + i * sz) : r.u32(p + i * sz));
    };
    for (int i = 0; i < n; ++i) {
        const size_t e = ifd + 2 + 12 * (size_t)i;
        if (!r.ok(e, 12)) return RIHVR_E_UNKNOWN_TYPE;
        const uint16_t tag = r.u16(e);
        std::vector<uint32_t> v;
        values(e, v);
        if (v.empty()) continue;
        switch (tag) {
            case 256: W = v[0]; break;
            case 257: H = v[0]; break;
            case 258: bps = v[0]; break;
            case 259: comp = v[0]; break;
            case 262: photometric = v[0]; break;
            case 273: offsets = v; break;
            case 277: spp = v[0]; break;
            case 278: rows_per_strip = v[0]; break;
            case 279: counts = v; break;
            case 284: planar = v[0]; break;
            default: break;
        }
    }
    if (!image_size_ok(W, H) || comp != 1 || planar != 1 || offsets.empty()) return RIHVR_E_UNKNOWN_TYPE;
    if (!((spp == 1 && (bps == 8 || bps == 16)) || (spp == 3 && bps == 8))) return RIHVR_E_UNKNOWN_TYPE;
    *width = (int32_t)W;
    *height = (int32_t)H;
    if (!pixels) return RIHVR_OK;
    const size_t bpp = (size_t)spp * (bps / 8);
    const size_t row_bytes = bpp * W;
    if (rows_per_strip == 0 || rows_per_strip > H) rows_per_strip = H;
    size_t y = 0;
    for (size_t s = 0; s < offsets.size() && y < H; ++s) {
        const size_t rows = std::min<size_t>(rows_per_strip, H - y);
        const size_t off = offsets[s];
        if (!r.ok(off, rows * row_bytes)) return RIHVR_E_INVALID_FILE;
        for (size_t yy = 0; yy < rows; ++yy, ++y) {
            const size_t base = off + yy * row_bytes;
            for (size_t x = 0; x < W; ++x) {
                float v;
                if (spp == 3) {
                    const uint8_t* p = &r.b[base + 3 * x];  // RGB order on disk
                    v = bgr_to_gray(p[2], p[1], p[0]);
                } else if (bps == 8) {
                    v = r.b[base + x];
                } else {
                    v = r.u16(base + 2 * x);
                }
                if (photometric == 0) v = (bps == 16 ? 65535.0f : 255.0f) - v;  // WhiteIsZero
                pixels[y * W + x] = v;
            }
        }
    }
    return y == H ? RIHVR_OK : RIHVR_E_INVALID_FILE;
}

rihvr_status decode_pgm(const std::vector<uint8_t>& b, float* pixels, int32_t* width, int32_t* height) {
    size_t p = 2;
    auto skip = [&]() {
        while (p < b.size()) {
            if (b[p] == '#') {
                while (p < b.size() && b[p] != '\n') ++p;
            } else if (b[p] == ' ' || b[p] == '\n' || b[p] == '\r' || b[p] == '\t') {
                ++p;
            } else {
                break;
            }
        }
    };
    // a decimal field; values beyond 2^40 stop growing (every such field is refused below)
    auto num = [&]() {
        skip();
        long v = 0;
        while (p < b.size() && b[p] >= '0' && b[p] <= '9') {
            const long d = b[p++] - '0';
            if (v <= (1L << 40)) v = v * 10 + d;
        }
        return v;
    };
    const long W = num(), H = num(), maxv = num();
    ++p;  // one whitespace byte before the raster
    if (!image_size_ok((uint64_t)std::max(W, 0L), (uint64_t)std::max(H, 0L)) || maxv <= 0 || maxv > 65535)
        return RIHVR_E_UNKNOWN_TYPE;
    *width = (int32_t)W;
    *height = (int32_t)H;
    if (!pixels) return RIHVR_OK;
    const size_t bpp = maxv > 255 ? 2 : 1;
    if (p + (size_t)W * H * bpp > b.size()) return RIHVR_E_INVALID_FILE;
    for (size_t i = 0; i < (size_t)W * H; ++i)
        pixels[i] = bpp == 1 ? (float)b[p + i] : (float)((b[p + 2 * i] << 8) | b[p + 2 * i + 1]);
    return RIHVR_OK;
}

// convertTo(CV_8U): round to nearest (ties to even), saturate
uint8_t to_u8(double v) {
    if (!(v == v)) return 0;
    double r = std::nearbyint(v);
    return (uint8_t)std::min(255.0, std::max(0.0, r));
}
// convertTo(CV_8U, 255) of a float image: the product is formed in float
uint8_t unit_to_u8(float v) { return to_u8((double)(v * 255.0f)); }

// The writers take the sizes the readers accept (image_size_ok), which also keeps every
// byte count within the 32-bit fields of the TIFF header and of a PNG chunk.
rihvr_status write_tiff(const char* path, const uint8_t* px, int32_t W, int32_t H) {
    if (!path || !px || W <= 0 || H <= 0 || !image_size_ok((uint64_t)W, (uint64_t)H)) return RIHVR_E_INVALID_ARGUMENT;
    FILE* f = std::fopen(path, "wb");
    if (!f) return RIHVR_E_BAD_FILENAME;
    const uint32_t n = (uint32_t)W * (uint32_t)H;
    std::vector<uint8_t> b;
    auto u16 = [&](uint16_t v) { b.push_back(v & 0xff); b.push_back(v >> 8); };
    auto u32 = [&](uint32_t v) { for (int i = 0; i < 4; ++i) b.push_back((v >> (8 * i)) & 0xff); };
    b.push_back('I'); b.push_back('I'); u16(42); u32(8);
    const uint16_t ntags = 9;
    const uint32_t data_off = 8 + 2 + ntags * 12 + 4;
    u16(ntags);
    auto tag = [&](uint16_t t, uint16_t type, uint32_t cnt, uint32_t val) { u16(t); u16(type); u32(cnt); u32(val); };
    tag(256, 4, 1, (uint32_t)W);
    tag(257, 4, 1, (uint32_t)H);
    tag(258, 3, 1, 8);
    tag(259, 3, 1, 1);
    tag(262, 3, 1, 1);
    tag(273, 4, 1, data_off);
    tag(277, 3, 1, 1);
    tag(278, 4, 1, (uint32_t)H);
    tag(279, 4, 1, n);
    u32(0);
    bool ok = std::fwrite(b.data(), 1, b.size(), f) == b.size() && std::fwrite(px, 1, n, f) == n;
    ok = (std::fclose(f) == 0) && ok;
    return ok ? RIHVR_OK : RIHVR_E_BAD_FILENAME;
}

// ---------------------------------------------------------------------------
// PNG (ISO/IEC 15948): signature, IHDR, one IDAT holding a zlib stream of stored
// (uncompressed) deflate blocks (RFC 1950/1951), IEND; CRC-32 over each chunk's type
// and data, Adler-32 over the filtered scanlines (filter byte 0 per row).
// ---------------------------------------------------------------------------
struct Crc32Table {
    uint32_t t[256];
    Crc32Table() {
        for (uint32_t i = 0; i < 256; ++i) {
            uint32_t v = i;
            for (int k = 0; k < 8; ++k) v = (v & 1) ? 0xEDB88320u ^ (v >> 1) : v >> 1;
            t[i] = v;
        }
    }
};

uint32_t crc32_update(uint32_t c, const uint8_t* p, size_t n) {
    static const Crc32Table table;  // initialised once, also when threads write PNGs at once
    for (size_t i = 0; i < n; ++i) c = table.t[(c ^ p[i]) & 0xff] ^ (c >> 8);
    return c;
}

void put_be32(std::vector<uint8_t>& b, uint32_t v) {
    for (int i = 3; i >= 0; --i) b.push_back((uint8_t)(v >> (8 * i)));
}

void png_chunk(std::vector<uint8_t>& out, const char* type, const std::vector<uint8_t>& data) {
    put_be32(out, (uint32_t)data.size());
    const size_t at = out.size();
    out.insert(out.end(), type, type + 4);
    out.insert(out.end(), data.begin(), data.end());
    const uint32_t crc = crc32_update(0xFFFFFFFFu, out.data() + at, 4 + data.size()) ^ 0xFFFFFFFFu;
    put_be32(out, crc);
}

rihvr_status write_png(const char* path, const uint8_t* px, int32_t W, int32_t H) {
    if (!path || !px || W <= 0 || H <= 0 || !image_size_ok((uint64_t)W, (uint64_t)H)) return RIHVR_E_INVALID_ARGUMENT;
    std::vector<uint8_t> raw;
    raw.reserve((size_t)(W + 1) * H);
    for (int y = 0; y < H; ++y) {
        raw.push_back(0);  // filter: none
        raw.insert(raw.end(), px + (size_t)y * W, px + (size_t)(y + 1) * W);
    }
    std::vector<uint8_t> z = {0x78, 0x01};  // zlib: deflate, 32 KiB window, no dictionary
    for (size_t at = 0; at < raw.size() || at == 0;) {
        const size_t len = std::min<size_t>(65535, raw.size() - at);
        const bool last = at + len >= raw.size();
        z.push_back(last ? 1 : 0);  // BFINAL, BTYPE = 00 (stored)
        z.push_back((uint8_t)(len & 0xff));
        z.push_back((uint8_t)(len >> 8));
        z.push_back((uint8_t)(~len & 0xff));
        z.push_back((uint8_t)((~len >> 8) & 0xff));
        z.insert(z.end(), raw.begin() + at, raw.begin() + at + len);
        at += len;
        if (last) break;
    }
    uint32_t a = 1, b = 0;  // Adler-32
    for (uint8_t v : raw) {
        a = (a + v) % 65521u;
        b = (b + a) % 65521u;
    }
    put_be32(z, (b << 16) | a);
    std::vector<uint8_t> out = {0x89, 'P', 'N', 'G', '\r', '\n', 0x1a, '\n'};
    std::vector<uint8_t> ihdr;
    put_be32(ihdr, (uint32_t)W);
    put_be32(ihdr, (uint32_t)H);
    ihdr.push_back(8);  // bit depth
    ihdr.push_back(0);  // colour type: greyscale
    ihdr.push_back(0);  // compression
    ihdr.push_back(0);  // filter method
    ihdr.push_back(0);  // no interlace
    png_chunk(out, "IHDR", ihdr);
    png_chunk(out, "IDAT", z);
    png_chunk(out, "IEND", {});
    FILE* f = std::fopen(path, "wb");
    if (!f) return RIHVR_E_BAD_FILENAME;
    bool ok = std::fwrite(out.data(), 1, out.size(), f) == out.size();
    ok = (std::fclose(f) == 0) && ok;
    return ok ? RIHVR_OK : RIHVR_E_BAD_FILENAME;
}

struct Volume {
    rihvr_params p{};
    std::vector<float> v;  // interleaved complex64
    long N = 0;
};

rihvr_status fetch_volume(rihvr_ctx* ctx, Volume& vol) {
    rihvr_status s = rihvr_get_params(ctx, &vol.p);
    if (s) return s;
    vol.N = (long)vol.p.width * vol.p.height;
    vol.v.resize(2 * (size_t)vol.N * vol.p.num_planes);
    return rihvr_get_volume(ctx, vol.v.data(), RIHVR_HOST);
}

// |v| as cv::magnitude (float)
inline float mag(const float* c) { return std::sqrt(c[0] * c[0] + c[1] * c[1]); }

void global_mag_range(const Volume& vol, double* lo, double* hi) {
    // minMaxLoc per plane, folded plane by plane from plane 0 (savePlanes :1048-1064)
    const long N = vol.N;
    for (int z = 0; z < vol.p.num_planes; ++z) {
        double mn = 0, mx = 0;
        for (long i = 0; i < N; ++i) {
            double m = mag(&vol.v[2 * ((size_t)z * N + i)]);
            if (i == 0 || m < mn) mn = m;
            if (i == 0 || m > mx) mx = m;
        }
        if (z == 0) { *lo = mn; *hi = mx; }
        *lo = std::min(*lo, mn);
        *hi = std::max(*hi, mx);
    }
}

}  // namespace

namespace rihvr {

rihvr_status crop_geometry(const rihvr_file_params* fp, int width, int height, CropGeometry* g) {
    // crop (hologram.cu:768-880)
    int rx, ry, rw, rh;
    if (fp->roi_w * fp->roi_h > 0) {
        rx = fp->roi_x; ry = fp->roi_y; rw = fp->roi_w; rh = fp->roi_h;
    } else {
        const int cx = (int)fp->center_x, cy = (int)fp->center_y, size = fp->roi_size;
        int sx = cx - size / 2, sy = cy - size / 2;
        rw = size; rh = size;
        if (sx < 0) sx = 0;
        if (sy < 0) sy = 0;
        if (sx + size > width) sx = width - size;
        if (sy + size > height) sy = height - size;
        if (width < size) { sx = 0; rw = width; }
        if (height < size) { sy = 0; rh = height; }
        if (size < 1) { sx = 0; sy = 0; rw = width; rh = height; }
        rx = sx; ry = sy;
    }
    // crop(Rect): zero borders where the ROI leaves the image (the right border is
    // computed from the ROI height, hologram.cu:848-852)
    int top = 0, bottom = 0, left = 0, right = 0, bw = width, bh = height;
    if (ry < 0) { top = -ry; ry = 0; bh += top; }
    if (ry + rh > bh) { bottom = ry + rh - bh; bh += bottom; }
    if (rx < 0) { left = -rx; rx = 0; bw += left; }
    if (rx + rw > bw) { right = rx + rh - bw; bw += right; }
    if (rw <= 0 || rh <= 0 || rx + rw > bw || ry + rh > bh) return RIHVR_E_INVALID_ARGUMENT;
    // zeroPad (hologram.cu:882-945)
    const int pad = fp->solver.zero_padding;
    if (pad < 0) return RIHVR_E_INVALID_ARGUMENT;
    if (rw != rh) return RIHVR_E_INVALID_ARGUMENT;  // the reference's padded Mat has rows and cols swapped
    g->src_x0 = rx - left;
    g->src_y0 = ry - top;
    g->rw = rw;
    g->rh = rh;
    g->pad = pad;
    g->ow = rw + 2 * pad;
    g->oh = rh + 2 * pad;
    return RIHVR_OK;
}

}  // namespace rihvr

extern "C" {

rihvr_status rihvr_read_param_file(const char* path, rihvr_file_params* out) {
    if (!path || !out) return RIHVR_E_INVALID_ARGUMENT;
    std::memset(out, 0, sizeof(*out));
    // HoloUi::HoloUi defaults (holo_ui.cpp:10-62)
    rihvr_params& p = out->solver;
    p.start_plane = 0;
    p.plane_stepsize = 1.0f;
    p.num_planes = 256;
    p.wavelength = 632;
    p.resolution = 1000;
    p.zero_padding = 256;
    p.num_inverse_iterations = 500;
    p.num_tv_iterations = 5;
    p.regularization_param = 0.01f;
    p.regularization_TV_param = 0.0f;
    p.mode = RIHVR_MODE_FASTA_FUSED_LASSO_2D;
    out->center_x = out->center_y = out->center_z = -1;
    out->roi_size = -1;
    out->background_start = out->background_end = -1;
    out->segment_min_area = 2;
    out->segment_min_intensity = 1;
    out->segment_close_size = 3;
    FILE* fid = std::fopen(path, "r");
    if (!fid) return RIHVR_E_BAD_FILENAME;
    char buff[RIHVR_PATH_MAX] = {0};
    char description[RIHVR_PATH_MAX] = {0};
    char char_value[RIHVR_PATH_MAX] = {0};
    int id = 0;
    float value = 0;
    bool set_full_filename = false;
    // while (!feof) { fgets; parse }: the failed fgets at the end of a file that ends
    // in a newline leaves the last line in the buffer, which is parsed again, and a
    // line that sscanf cannot read re-applies the previous id and values; both are
    // idempotent
    while (!std::feof(fid)) {
        if (!std::fgets(buff, sizeof(buff), fid) && buff[0] == 0) break;
        if (buff[0] == '%' || buff[0] == '#') continue;
        std::sscanf(buff, "%4095s %d %f", description, &id, &value);
        std::sscanf(buff, "%4095s %d %4095[^\t\n]", description, &id, char_value);
        switch (id) {
            case NUM_IMAGES: out->image_count = (int32_t)value; break;
            case START_IMAGE: out->start_image = (int32_t)value; break;
            case BG_START: out->background_start = (int32_t)value; break;
            case BG_END: out->background_end = (int32_t)value; break;
            case INPUT_FN: copy_str(out->image_filename, char_value); set_full_filename = true; break;
            case INPUT_PATHN: copy_str(out->image_path, char_value); break;
            case INPUT_FORMATFN: copy_str(out->image_formatted_name, char_value); break;
            case OUTPUT_PATHN: copy_str(out->output_path, char_value); break;
            case INPUT_BACKGROUND: copy_str(out->background_filename, char_value); out->remove_background = 1; break;
            case EXE_PATHN: break;
            case START_PLANE: p.start_plane = value; break;
            case PLANE_STEP: p.plane_stepsize = value; break;
            case NUM_PLANES: p.num_planes = (int32_t)value; break;
            case RESOLUTION: p.resolution = value; break;
            case WAVELENGTH: p.wavelength = value; break;
            case ROI_X: out->center_x = value; break;
            case ROI_Y: out->center_y = value; break;
            case ROI_Z: out->center_z = value; break;
            case ROI_SIZE: out->roi_size = (int32_t)value; break;
            case ZERO_PADDING: p.zero_padding = (int32_t)value; break;
            case ROI_RECT_X: out->roi_x = (int32_t)value; break;
            case ROI_RECT_Y: out->roi_y = (int32_t)value; break;
            case ROI_RECT_W: out->roi_w = (int32_t)value; break;
            case ROI_RECT_H: out->roi_h = (int32_t)value; break;
            case BETA: case IIPE_ITERATIONS: case PARTICLE_DIAMETER: case SIM_CHAMBER_RADIUS:
            case SIM_CONCENTRATION: case SIM_LENGTH: case SIM_RESOLUTION: case SIM_IMAGE_SIZE:
            case SIM_ROTATION_STEP: case SIM_OVERSAMPLE: case INVERSE_MAX_SVD:
                break;  // parameters of other executables: accepted, not used here
            case SEGMENT_MIN_AREA: out->segment_min_area = (int32_t)value; break;
            case SEGMENT_MIN_INTENSITY: out->segment_min_intensity = value; break;
            case SEGMENT_CLOSE_SIZE: out->segment_close_size = (int32_t)value; break;
            case NUM_INVERSE_ITS: p.num_inverse_iterations = (int32_t)value; break;
            case INVERSE_REG_TAU: p.regularization_param = value; break;
            case INVERSE_REG_TV: p.regularization_TV_param = value; break;
            case NUM_TV_ITS: p.num_tv_iterations = (int32_t)value; break;
            case FLAG_OUTPUT_PLANES: out->output_planes = flag_value(value, char_value); break;
            case FLAG_OUTPUT_STEPS: out->output_steps = flag_value(value, char_value); break;
            case FLAG_MEDIAN_BACKGROUND:
                out->use_median_background = flag_value(value, char_value);
                break;
            case FLAG_STANDARD_RECON: case FLAG_INVERSE_RECON: case FLAG_DECONV_RECON:
                break;
            default:
                // holo_ui.cpp:480-484: stop at the first unknown id and return, keeping
                // what was read (image_filename is then not composed)
                std::fprintf(stderr, "Unknown or unsupported input argument: %s\n", buff);
                out->unknown_id = id;
                std::fclose(fid);
                return RIHVR_E_INVALID_FILE;
        }
    }
    std::fclose(fid);
    if (!set_full_filename)
        std::snprintf(out->image_filename, RIHVR_PATH_MAX, "%s%s", out->image_path, out->image_formatted_name);
    return RIHVR_OK;
}

rihvr_status rihvr_read_image(const char* path, float* pixels, int32_t* width, int32_t* height) {
    if (!path || !width || !height) return RIHVR_E_INVALID_ARGUMENT;
    Reader r;
    if (!read_file(path, r.b)) return RIHVR_E_BAD_FILENAME;
    if (r.b.size() >= 4 && ((r.b[0] == 'I' && r.b[1] == 'I') || (r.b[0] == 'M' && r.b[1] == 'M')))
        return decode_tiff(r, pixels, width, height);
    if (r.b.size() >= 2 && r.b[0] == 'P' && r.b[1] == '5') return decode_pgm(r.b, pixels, width, height);
    return RIHVR_E_UNKNOWN_TYPE;
}

rihvr_status rihvr_prepare_hologram(const rihvr_file_params* fp, const float* image, const float* background,
                                    int32_t width, int32_t height, float* out, int32_t* out_width,
                                    int32_t* out_height) {
    if (!fp || !image || width <= 0 || height <= 0 || !out_width || !out_height) return RIHVR_E_INVALID_ARGUMENT;
    if (!image_size_ok((uint64_t)width, (uint64_t)height)) return RIHVR_E_INVALID_ARGUMENT;  // the readers' limits
    if (fp->remove_background && !background) return RIHVR_E_MISSING_DATA;
    const long N = (long)width * height;
    std::vector<float> h(N);
    // rescale(HOLOGRAM_SCALE_0_1): / 255 (hologram.cu:401-439), a float scale as cv::Mat / s
    const float inv255 = (float)(1.0 / 255.0);
    for (long i = 0; i < N; ++i) h[i] = image[i] * inv255;
    if (fp->remove_background) {
        // removeBackground (hologram.cu:1365-1432): (I - B) / max(sqrt(B), 0.1)
        for (long i = 0; i < N; ++i) {
            const float b = background[i] * inv255;
            const float sb = std::max(std::sqrt(b), 0.1f);
            h[i] = (h[i] - b) / sb;
        }
    } else {
        // normalize_mean0 (hologram.cu:332-358): I / mean - 1
        double sum = 0;
        for (long i = 0; i < N; ++i) sum += h[i];
        const float inv_mean = (float)(1.0 / (sum / (double)N));
        for (long i = 0; i < N; ++i) h[i] = h[i] * inv_mean - 1.0f;
    }
    rihvr::CropGeometry g;
    const rihvr_status gs = rihvr::crop_geometry(fp, width, height, &g);
    if (gs) return gs;
    const int ow = g.ow, oh = g.oh, pad = g.pad, rw = g.rw, rh = g.rh;
    *out_width = ow;
    *out_height = oh;
    if (!out) return RIHVR_OK;
    std::fill(out, out + 2 * (size_t)ow * oh, 0.0f);
    for (int y = 0; y < rh; ++y)
        for (int x = 0; x < rw; ++x) {
            const int sy = g.src_y0 + y, sx = g.src_x0 + x;  // coordinates in the unbordered image
            float v = 0.0f;
            if (sy >= 0 && sy < height && sx >= 0 && sx < width) v = h[(size_t)sy * width + sx];
            out[2 * ((size_t)(y + pad) * ow + (x + pad))] = v;
        }
    return RIHVR_OK;
}

rihvr_status rihvr_write_tiff8(const char* path, const uint8_t* pixels, int32_t width, int32_t height) {
    return write_tiff(path, pixels, width, height);
}

rihvr_status rihvr_write_png8(const char* path, const uint8_t* pixels, int32_t width, int32_t height) {
    return write_png(path, pixels, width, height);
}

rihvr_status rihvr_write_hologram(const char* path, const float* holo, int32_t width, int32_t height,
                                  int32_t scale_to_minmax, double lo, double hi) {
    if (!holo || width <= 0 || height <= 0 || !image_size_ok((uint64_t)width, (uint64_t)height))
        return RIHVR_E_INVALID_ARGUMENT;
    const long N = (long)width * height;
    std::vector<uint8_t> px(N);
    if (scale_to_minmax) {
        double mn = holo[0], mx = holo[0];
        for (long i = 1; i < N; ++i) {
            mn = std::min(mn, (double)holo[2 * i]);
            mx = std::max(mx, (double)holo[2 * i]);
        }
        // cv::normalize(NORM_MINMAX, 0, 255) then convertTo(CV_8U)
        const double sc = mx > mn ? 255.0 / (mx - mn) : 0.0;
        for (long i = 0; i < N; ++i) px[i] = to_u8((float)((holo[2 * i] - mn) * sc));
    } else {
        if (!(hi > lo)) return RIHVR_E_INVALID_ARGUMENT;
        for (long i = 0; i < N; ++i) px[i] = to_u8((float)(255 * (holo[2 * i] - lo) / (hi - lo)));
    }
    return write_tiff(path, px.data(), width, height);
}

rihvr_status rihvr_read_sparse_voxels(const char* path, int32_t width, int32_t height, int32_t num_planes,
                                      int64_t* count, int32_t* plane, int32_t* index, float* intensity) {
    if (!path || !count || width <= 0 || height <= 0 || num_planes <= 0) return RIHVR_E_INVALID_ARGUMENT;
    if (plane && (!index || !intensity)) return RIHVR_E_INVALID_ARGUMENT;
    FILE* f = std::fopen(path, "rb");
    if (!f) return RIHVR_E_BAD_FILENAME;
    std::string text;
    char chunk[1 << 16];
    size_t got;
    while ((got = std::fread(chunk, 1, sizeof(chunk), f)) > 0) text.append(chunk, got);
    std::fclose(f);
    // the header check reads 14 characters (fgets(read_str, 15), sparse_volume.cu:577-597)
    static const char kHeader[] = "x, y, z, value";
    if (text.compare(0, 14, kHeader) != 0) return RIHVR_E_INVALID_FILE;
    size_t pos = 14;
    while (pos < text.size() && text[pos] != '\n') {
        if (!std::isspace((unsigned char)text[pos])) return RIHVR_E_INVALID_FILE;
        ++pos;
    }
    struct Row {
        int32_t x, y, z;
        float v;
    };
    std::vector<Row> rows;
    const char* s = text.c_str();
    const char* end = s + text.size();
    const char* q = s + pos;
    // rows "%d, %d, %d, %e\n" (:607-609): literal commas, any white space around
    auto skip_ws = [&]() {
        while (q < end && std::isspace((unsigned char)*q)) ++q;
    };
    auto read_int = [&](int32_t* out) -> bool {
        skip_ws();
        char* e = nullptr;
        const long v = std::strtol(q, &e, 10);
        if (e == q || v < INT32_MIN || v > INT32_MAX) return false;
        *out = (int32_t)v;
        q = e;
        return true;
    };
    auto comma = [&]() -> bool {
        if (q >= end || *q != ',') return false;
        ++q;
        return true;
    };
    const long W = width, H = height, D = num_planes;
    for (;;) {
        skip_ws();
        if (q >= end) break;
        Row r;
        if (!read_int(&r.x) || !comma() || !read_int(&r.y) || !comma() || !read_int(&r.z) || !comma())
            return RIHVR_E_INVALID_FILE;
        skip_ws();
        char* e = nullptr;
        r.v = std::strtof(q, &e);
        if (e == q) return RIHVR_E_INVALID_FILE;
        q = e;
        // the reference's test (:613-623), then what it would store out of range
        if (r.x > W || r.y > H || r.z > D) return RIHVR_E_INVALID_DATA;
        if (r.x < 0 || r.y < 0 || r.z < 0 || r.z >= D || (long)r.y * W + r.x >= W * H) return RIHVR_E_INVALID_DATA;
        rows.push_back(r);
    }
    *count = (int64_t)rows.size();
    if (!plane) return RIHVR_OK;
    // plane z takes the next count[z] rows of the file (:629-654)
    std::vector<int64_t> per(num_planes, 0);
    for (const Row& r : rows) ++per[r.z];
    size_t k = 0;
    for (int32_t z = 0; z < num_planes; ++z)
        for (int64_t j = 0; j < per[z]; ++j, ++k) {
            const Row& r = rows[k];
            plane[k] = z;
            index[k] = (int32_t)(r.y * W + r.x);
            intensity[k] = r.v;
        }
    return RIHVR_OK;
}

rihvr_status rihvr_save_sparse(rihvr_ctx* ctx, const char* csv_path) {
    if (!ctx || !csv_path) return RIHVR_E_INVALID_ARGUMENT;
    Volume vol;
    rihvr_status s = fetch_volume(ctx, vol);
    if (s) return s;
    FILE* f = std::fopen(csv_path, "w");
    if (!f) return RIHVR_E_BAD_FILENAME;
    std::fprintf(f, "x, y, z, value\n");
    const int W = vol.p.width;
    for (int z = 0; z < vol.p.num_planes; ++z)
        for (long i = 0; i < vol.N; ++i) {
            const float* c = &vol.v[2 * ((size_t)z * vol.N + i)];
            if (c[0] != 0.0f || c[1] != 0.0f) {
                const float intensity = c[0] * c[0] + c[1] * c[1];
                std::fprintf(f, "%ld, %ld, %d, %e\n", i % W, i / W, z, intensity);
            }
        }
    return std::fclose(f) == 0 ? RIHVR_OK : RIHVR_E_BAD_FILENAME;
}

rihvr_status rihvr_save_projections(rihvr_ctx* ctx, const char* output_path, const char* prefix,
                                    int32_t prefix_as_suffix) {
    if (!ctx || !output_path || !prefix) return RIHVR_E_INVALID_ARGUMENT;
    Volume vol;
    rihvr_status s = fetch_volume(ctx, vol);
    if (s) return s;
    const int W = vol.p.width, H = vol.p.height, D = vol.p.num_planes;
    double gmin = 0, gmax = 0;
    global_mag_range(vol, &gmin, &gmax);
    std::vector<float> xy((size_t)W * H, 0.0f), xz((size_t)D * W, 0.0f), yz((size_t)H * D, 0.0f);
    const float fmin = (float)gmin, frange = (float)(gmax - gmin);
    for (int z = 0; z < D; ++z) {
        std::vector<float> xzv(W, -INFINITY), yzv(H, -INFINITY);
        for (int y = 0; y < H; ++y)
            for (int x = 0; x < W; ++x) {
                const float m = (mag(&vol.v[2 * ((size_t)z * vol.N + (size_t)y * W + x)]) - fmin) / frange;
                xzv[x] = std::max(xzv[x], m);
                yzv[y] = std::max(yzv[y], m);
                if (z > 0) xy[(size_t)y * W + x] = std::max(xy[(size_t)y * W + x], m);  // plane 0 left out
            }
        for (int x = 0; x < W; ++x) xz[(size_t)z * W + x] = xzv[x];
        for (int y = 0; y < H; ++y) yz[(size_t)y * D + z] = yzv[y];
    }
    auto save = [&](const std::vector<float>& img, int w, int h, const char* axes) {
        std::vector<uint8_t> px(img.size());
        for (size_t i = 0; i < img.size(); ++i) px[i] = unit_to_u8(img[i]);
        std::string fn = prefix_as_suffix ? std::string(output_path) + axes + "_" + prefix + ".tif"
                                          : std::string(output_path) + prefix + "_" + axes + ".tif";
        // a failed imwrite throws HOLO_ERROR_UNKNOWN_ERROR there (sparse_volume.cu:795-799)
        return write_tiff(fn.c_str(), px.data(), w, h) ? RIHVR_E_UNKNOWN_ERROR : RIHVR_OK;
    };
    if ((s = save(xy, W, H, "xy"))) return s;
    if ((s = save(xz, W, D, "xz"))) return s;
    return save(yz, D, H, "yz");
}

rihvr_status rihvr_save_planes(rihvr_ctx* ctx, const char* path_prefix) {
    if (!ctx || !path_prefix) return RIHVR_E_INVALID_ARGUMENT;
    Volume vol;
    rihvr_status s = fetch_volume(ctx, vol);
    if (s) return s;
    const int W = vol.p.width, H = vol.p.height;
    double gmin = 0, gmax = 0;
    global_mag_range(vol, &gmin, &gmax);
    const double ri_max = std::sqrt(gmax) / 10, ri_min = -ri_max;
    std::vector<uint8_t> px((size_t)W * H);
    char fn[RIHVR_PATH_MAX + 64];
    for (int kind = 0; kind < 4; ++kind) {
        static const char* const fmt[4] = {"%s_%04d.tif", "%s_phase_%04d.tif", "%s_real_%04d.tif", "%s_imag_%04d.tif"};
        for (int z = 0; z < vol.p.num_planes; ++z) {
            for (long i = 0; i < vol.N; ++i) {
                const float* c = &vol.v[2 * ((size_t)z * vol.N + i)];
                float v;
                if (kind == 0) {
                    v = (float)((mag(c) - gmin) / (gmax - gmin));
                } else if (kind == 1) {
                    double a = std::atan2((double)c[1], (double)c[0]);  // cv::phase: [0, 2 pi)
                    if (a < 0) a += 2 * kPi;
                    v = (float)(a / (2 * kPi));
                } else {
                    v = (float)((c[kind == 2 ? 0 : 1] - ri_min) / (ri_max - ri_min));
                }
                px[i] = unit_to_u8(v);
            }
            std::snprintf(fn, sizeof(fn), fmt[kind], path_prefix, z);
            // a failed imwrite throws HOLO_ERROR_UNKNOWN_ERROR there (:1073-1077)
            if (write_tiff(fn, px.data(), W, H)) return RIHVR_E_UNKNOWN_ERROR;
        }
    }
    return RIHVR_OK;
}

}  // extern "C"
