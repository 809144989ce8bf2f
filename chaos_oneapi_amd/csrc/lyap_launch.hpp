// Host-side launch entry points of the kernel translation units.  Each
// returns the HIP error of its launches; the C-ABI layer turns it into a
// status code and a message.
#pragma once

#include <hip/hip_runtime.h>

#include "chaos_lyap.h"
#include "lyap_device.hpp"

namespace lyap {

// records the message returned by lyap_last_error() and returns `code`
int set_error(int code, const char* fmt, ...);

// 3-D flows, one lane per point (lyap_sweep.hip)
hipError_t launch_lane_sweep(const lyap_sweep_desc& d, const SweepIO& io, hipStream_t s,
                             int num_cus, int max_blocks);

// Lorenz-96, one point per block of `waves` wave64s (lyap_l96.hip);
// waves <= 0 picks the default for the dimension
hipError_t launch_l96_sweep(const lyap_sweep_desc& d, const SweepIO& io, hipStream_t s, int num_cus,
                            int waves);
bool l96_supported(int n, int waves);
int l96_default_waves(int n);

}  // namespace lyap
