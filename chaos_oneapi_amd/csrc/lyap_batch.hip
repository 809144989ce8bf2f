// Batched mixed-model sweep in one persistent launch.
//
// A batch is a list of classes (one model and dimension each, with its own
// descriptor, points and outputs).  Every workgroup is one wave64 that walks
// the classes in the order the host put them in and drains each one through
// that class's own work counter:
//   - 3-D flows (shimizu, Lorenz, Rössler): one point per lane, lanes refill
//     from the counter as their points finish (lane_drain), with N or (trace
//     closure) N - 1 deviation vectors;
//   - Lorenz-96 N = 8, 16, 32: P points per wave, each on 64/P lanes, slots
//     refill independently (l96_wave_drain).
// A wave moves to the next class once the current class has no items left and
// all of its lanes are done, so waves spread over the classes by themselves and
// the queue needs no global phase barrier.  The host orders the classes longest
// task first (per-lane 3-D points are the longest, then L96 by N descending),
// so the last class drained is the one with the shortest items.
// The per-point arithmetic is exactly that of the single-class kernels, so a
// batch reproduces their results bit for bit.
#include "lyap_device.hpp"
#include "lyap_l96_wave.hpp"
#include "lyap_lane.hpp"
#include "lyap_launch.hpp"

namespace lyap {

namespace {

// Lorenz-96 points per wave for N = 8, 16, 32 of each batch layout
struct BatchLayout {
    int p8, p16, p32;
};
constexpr BatchLayout kLayouts[kBatchLayouts] = {
    {8, 4, 1},  // one point per N-lane group (no replicated base state) except N = 32
    {4, 2, 1},  // two vector groups per point for N = 8, 16
};

constexpr int imax(int a, int b) { return a > b ? a : b; }

LYAP_D uint32_t uni(uint32_t v) { return (uint32_t)__builtin_amdgcn_readfirstlane((int)v); }
LYAP_D int64_t uni(int64_t v) {
    const uint64_t b = (uint64_t)v;
    return (int64_t)(((uint64_t)uni((uint32_t)(b >> 32)) << 32) | (uint64_t)uni((uint32_t)b));
}
LYAP_D double uni(double v) { return __builtin_bit_cast(double, uni(__builtin_bit_cast(int64_t, v))); }
template <class Ptr>
LYAP_D Ptr* uni(Ptr* v) { return (Ptr*)uni((int64_t)(uintptr_t)v); }

// The fields of a class the step loops read, loaded once and made wave-uniform
// (the class records sit in global memory, where the compiler cannot see that
// they are the same for every lane).  Only these fields of the returned
// descriptor are set; per-point set-up reads the record itself.
LYAP_D lyap_sweep_desc hot_fields(const lyap_sweep_desc& g) {
    lyap_sweep_desc d;
    d.n_renorm_transient = uni(g.n_renorm_transient);
    d.n_renorm_average = uni(g.n_renorm_average);
    d.max_steps = uni(g.max_steps);
    d.steps_per_renorm = uni(g.steps_per_renorm);
    d.rtol = uni(g.rtol);
    d.atol = uni(g.atol);
    d.h_min = uni(g.h_min);
    d.h_max = uni(g.h_max);
    d.t_renorm = uni(g.t_renorm);
    d.escape_radius = uni(g.escape_radius);
    d.h = uni(g.h);
    d.n_exps = (int32_t)uni((uint32_t)g.n_exps);
    return d;
}

LYAP_D SweepIO uniform_io(const SweepIO& g) {
    SweepIO io;
    io.first = uni(g.first);
    io.count = uni(g.count);
    io.idx_list = uni(g.idx_list);
    io.params = uni(g.params);
    io.inits = uni(g.inits);
    io.lambda = uni(g.lambda);
    io.status = uni(g.status);
    io.n_acc = uni(g.n_acc);
    io.n_rej = uni(g.n_rej);
    io.counter = uni(g.counter);
    io.stride = uni(g.stride);
    io.gather = (int32_t)uni((uint32_t)g.gather);
    io.pad = 0;
    return io;
}

template <int INTEG, bool ADAPT, bool MIXED, int LAYOUT>
LYAP_D void batch_body(const BatchClass* __restrict__ cls, int n_cls, double* lds) {
    constexpr BatchLayout L = kLayouts[LAYOUT];
    for (int c = 0; c < n_cls; ++c) {
        const BatchClass* bc = cls + c;
        const lyap_sweep_desc d = hot_fields(bc->d);
        const SweepIO io = uniform_io(bc->io);
        auto getd = [bc] { return &bc->d; };
        auto getio = [bc] { return &bc->io; };
        switch ((int)uni((uint32_t)bc->kind)) {
            case BATCH_SHIMIZU: lane_drain<Shimizu, INTEG, ADAPT, MIXED>(d, getd, io, getio); break;
            case BATCH_LORENZ: lane_drain<Lorenz, INTEG, ADAPT, MIXED>(d, getd, io, getio); break;
            case BATCH_ROSSLER: lane_drain<Rossler, INTEG, ADAPT, MIXED>(d, getd, io, getio); break;
            case BATCH_SHIMIZU_TRACE: lane_drain<Shimizu, INTEG, ADAPT, MIXED, true>(d, getd, io, getio); break;
            case BATCH_LORENZ_TRACE: lane_drain<Lorenz, INTEG, ADAPT, MIXED, true>(d, getd, io, getio); break;
            case BATCH_ROSSLER_TRACE: lane_drain<Rossler, INTEG, ADAPT, MIXED, true>(d, getd, io, getio); break;
            case BATCH_L96_8: l96_wave_drain<8, L.p8, INTEG, ADAPT, MIXED>(d, getd, io, lds); break;
            case BATCH_L96_16: l96_wave_drain<16, L.p16, INTEG, ADAPT, MIXED>(d, getd, io, lds); break;
            case BATCH_L96_32: l96_wave_drain<32, L.p32, INTEG, ADAPT, MIXED>(d, getd, io, lds); break;
            default: break;
        }
    }
}

template <int LAYOUT>
constexpr int batch_lds_doubles() {
    return imax(l96_wave_lds_doubles<8, kLayouts[LAYOUT].p8>(),
                imax(l96_wave_lds_doubles<16, kLayouts[LAYOUT].p16>(),
                     l96_wave_lds_doubles<32, kLayouts[LAYOUT].p32>()));
}

}  // namespace

// compiler-chosen register budget
template <int INTEG, bool ADAPT, bool MIXED, int LAYOUT>
__global__ __launch_bounds__(64) void k_batch(const BatchClass* __restrict__ cls, int n_cls) {
    __shared__ double lds[batch_lds_doubles<LAYOUT>()];
    batch_body<INTEG, ADAPT, MIXED, LAYOUT>(cls, n_cls, lds);
}

// register budget for 3 waves per SIMD
template <int INTEG, bool ADAPT, bool MIXED, int LAYOUT>
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(3))) void k_batch_w3(
    const BatchClass* __restrict__ cls, int n_cls) {
    __shared__ double lds[batch_lds_doubles<LAYOUT>()];
    batch_body<INTEG, ADAPT, MIXED, LAYOUT>(cls, n_cls, lds);
}

__global__ void k_batch_reset(const BatchClass* __restrict__ cls, int n_cls) {
    const int c = (int)threadIdx.x;
    if (blockIdx.x == 0 && c < n_cls) *cls[c].io.counter = 0ULL;
}

namespace {

using BatchKernel = void (*)(const BatchClass*, int);

template <int INTEG, bool ADAPT, bool MIXED>
BatchKernel pick_tuned(int layout, bool w3) {
    if (layout == 0) return w3 ? k_batch_w3<INTEG, ADAPT, MIXED, 0> : k_batch<INTEG, ADAPT, MIXED, 0>;
    return w3 ? k_batch_w3<INTEG, ADAPT, MIXED, 1> : k_batch<INTEG, ADAPT, MIXED, 1>;
}

// Every (integrator, adaptive, precision) has the default layout; the adaptive
// Dopri5 variants (uneven step counts, the case batching is for) also have the
// tuned ones.
BatchKernel pick_kernel(int integ, bool adapt, bool mixed, int layout, bool w3) {
    const bool tuned = layout != 0 || w3;
    if (integ == LYAP_INT_DOPRI5 && adapt)
        return mixed ? pick_tuned<LYAP_INT_DOPRI5, true, true>(layout, w3)
                     : pick_tuned<LYAP_INT_DOPRI5, true, false>(layout, w3);
    if (tuned) return nullptr;
    const int key = (integ << 2) | (adapt ? 2 : 0) | (mixed ? 1 : 0);
    switch (key) {
        case (LYAP_INT_DVERK << 2) | 0: return k_batch<LYAP_INT_DVERK, false, false, 0>;
        case (LYAP_INT_DVERK << 2) | 1: return k_batch<LYAP_INT_DVERK, false, true, 0>;
        case (LYAP_INT_DVERK << 2) | 2: return k_batch<LYAP_INT_DVERK, true, false, 0>;
        case (LYAP_INT_DVERK << 2) | 3: return k_batch<LYAP_INT_DVERK, true, true, 0>;
        case (LYAP_INT_DOPRI5 << 2) | 0: return k_batch<LYAP_INT_DOPRI5, false, false, 0>;
        case (LYAP_INT_DOPRI5 << 2) | 1: return k_batch<LYAP_INT_DOPRI5, false, true, 0>;
        default: return nullptr;
    }
}

}  // namespace

bool batch_variant_built(int integ, bool adapt, bool mixed, int layout, bool w3) {
    if (layout < 0 || layout >= kBatchLayouts) return false;
    return pick_kernel(integ, adapt, mixed, layout, w3) != nullptr;
}

int batch_points_per_wave(int layout, int dim) {
    if (layout < 0 || layout >= kBatchLayouts) return 0;
    return dim == 8 ? kLayouts[layout].p8 : (dim == 16 ? kLayouts[layout].p16 : (dim == 32 ? kLayouts[layout].p32 : 64));
}

hipError_t batch_max_waves_per_cu(int integ, bool adapt, bool mixed, int layout, bool w3, int* out) {
    BatchKernel k = pick_kernel(integ, adapt, mixed, layout, w3);
    if (!k) return hipErrorInvalidValue;
    return hipOccupancyMaxActiveBlocksPerMultiprocessor(out, k, 64, 0);
}

hipError_t launch_batch(const BatchClass* d_cls, int n_cls, int integ, bool adapt, bool mixed,
                        int layout, bool w3, int64_t waves, hipStream_t s) {
    BatchKernel k = pick_kernel(integ, adapt, mixed, layout, w3);
    if (!k || n_cls < 1 || n_cls > kMaxBatchClasses || waves < 1) return hipErrorInvalidValue;
    hipLaunchKernelGGL(k_batch_reset, dim3(1), dim3(64), 0, s, d_cls, n_cls);
    hipLaunchKernelGGL(k, dim3((unsigned)waves), dim3(64), 0, s, d_cls, n_cls);
    return hipGetLastError();
}

}  // namespace lyap
