// warmstart.hpp -- the learned warm start's entry points used by the solver.
#pragma once
#include <hip/hip_runtime.h>

#include "../../include/rihvr.h"

namespace rihvr {
namespace net {

// out_dev = the network's start volume for in_dev (both [D][H][W] complex64 on the device);
// the network's stream first waits for the work queued on wait_on (may be null).  Blocks until done.
rihvr_status predict(rihvr_net* n, const float* in_dev, float* out_dev, hipStream_t wait_on);
// true when the network was created for a D x H x W volume
bool net_matches(const rihvr_net* n, int D, int H, int W);

}  // namespace net
}  // namespace rihvr
