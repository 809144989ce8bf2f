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

// 3-D flows and maps, one lane per point (lyap_sweep.hip); quad: the sweep's
// method on the 3-D flows with four lanes per point; watch: progress signals
// and skips of the sweep's method with one lane per point (null: none)
hipError_t launch_lane_sweep(const lyap_sweep_desc& d, const SweepIO& io, hipStream_t s,
                             int num_cus, int max_blocks, bool quad = false,
                             const LaneWatch* watch = nullptr);

// Lorenz-96 (lyap_l96.hip): ppw > 0 runs the wave-level drain with ppw
// points per wave64; otherwise one point per block of `waves` wave64s
// (waves <= 0 picks the default for the dimension); max_blocks > 0 caps the
// persistent grid
hipError_t launch_l96_sweep(const lyap_sweep_desc& d, const SweepIO& io, hipStream_t s, int num_cus,
                            int waves, int ppw, int max_blocks);
bool l96_supported(int n, int waves);
bool l96_wave_supported(int n, int ppw);
int l96_default_waves(int n);

// heavy-first order (lyap_order.hip): keys n_acc + n_rej of a probe run over
// points [first, first + count) of a grid with n2 points per row, with
// radius > 0 the largest over the (2 radius + 1)^2 grid neighbourhood, sorted
// descending with the point indices (vals_out: the index list of the full
// run); temp: order_sort_bytes(count) bytes of sort scratch (0 when the query
// fails)
size_t order_sort_bytes(int64_t count);
hipError_t order_heavy_first(const uint32_t* n_acc, const uint32_t* n_rej, int64_t first, int64_t count,
                             int32_t n2, int32_t radius, uint32_t* keys_in, uint32_t* keys_out,
                             int64_t* vals_in, int64_t* vals_out, void* temp, size_t temp_bytes,
                             hipStream_t s);
// outputs of a run in list order (idx: the list) back to point order; null
// st/na/nr outputs are skipped
hipError_t order_scatter(const int64_t* idx, int64_t first, int64_t count, int n_exps, const double* lam_s,
                         const uint32_t* st_s, const uint32_t* na_s, const uint32_t* nr_s, double* lam,
                         uint32_t* st, uint32_t* na, uint32_t* nr, hipStream_t s);

// ---------------------------------------------------------------------------
// batched mixed-model launch (lyap_batch.hip)
// ---------------------------------------------------------------------------
enum BatchKind : int32_t {
    BATCH_SHIMIZU = 0,
    BATCH_LORENZ = 1,
    BATCH_ROSSLER = 2,
    BATCH_L96_8 = 3,
    BATCH_L96_16 = 4,
    BATCH_L96_32 = 5,
    // the 3-D flows with the trace closure (N - 1 vectors, lambda_N from tr J)
    BATCH_SHIMIZU_TRACE = 6,
    BATCH_LORENZ_TRACE = 7,
    BATCH_ROSSLER_TRACE = 8,
};
constexpr bool batch_kind_is_l96(int kind) {
    return kind == BATCH_L96_8 || kind == BATCH_L96_16 || kind == BATCH_L96_32;
}
constexpr int kMaxBatchClasses = 16;
constexpr int kBatchLayouts = 2;

// one class of a batch as the kernel reads it (device memory)
struct BatchClass {
    lyap_sweep_desc d;  // finalised
    SweepIO io;         // io.counter: this class's own work counter
    int32_t kind;       // BatchKind
    int32_t pad;
};

bool batch_variant_built(int integ, bool adapt, bool mixed, int layout, bool w3);
int batch_points_per_wave(int layout, int dim);
hipError_t batch_max_waves_per_cu(int integ, bool adapt, bool mixed, int layout, bool w3, int* out);
// reset of the class counters + the persistent kernel on `waves` one-wave workgroups
hipError_t launch_batch(const BatchClass* d_cls, int n_cls, int integ, bool adapt, bool mixed,
                        int layout, bool w3, int64_t waves, hipStream_t s);

}  // namespace lyap
