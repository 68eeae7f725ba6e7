// io_internal.hpp -- host helpers shared by io.cpp and the device preprocessing.
#pragma once
#include "../../include/rihvr.h"

namespace rihvr {

// The crop and zero-pad of the input contract (Hologram::crop, hologram.cu:768-880;
// zeroPad, :882-945): output pixel (x, y) of the rw x rh crop, placed at (x + pad,
// y + pad) of the ow x oh output, reads source pixel (src_x0 + x, src_y0 + y) of the
// image, zero where that lies outside it.
struct CropGeometry {
    int src_x0, src_y0;  // source of the crop's first pixel (may be negative)
    int rw, rh;          // crop size
    int pad;             // zero padding on every side
    int ow, oh;          // output size
};

// RIHVR_E_INVALID_ARGUMENT for an empty or non-square crop or a negative padding.
rihvr_status crop_geometry(const rihvr_file_params* fp, int width, int height, CropGeometry* g);

}  // namespace rihvr
