// Host-side launch entry points of the kernel translation units.  Each
// returns the HIP error of its launches; the C-ABI layer turns it into a
// status code and a message.
#pragma once

#include <hip/hip_runtime.h>

#include "chaos_lyap.h"
#include "lyap_device.hpp"

namespace lyap {

// 3-D flows, one lane per point (lyap_sweep.hip)
hipError_t launch_lane_sweep(const lyap_sweep_desc& d, const SweepIO& io, hipStream_t s,
                             int num_cus, int max_blocks);

}  // namespace lyap
