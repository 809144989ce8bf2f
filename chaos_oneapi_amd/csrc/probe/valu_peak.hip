// Vector-ALU FMA issue-rate probe for the roofline's peak: binary64 FMAs,
// unpacked binary32 FMAs (this library is built without packed binary32
// instructions) and a 1:1 interleave of the two.  Every lane runs CH
// independent fma chains for `iters` iterations on 16 waves per CU; a launch
// does 2 * CH * iters flops per lane.  Prints one JSON line with the best of
// `reps` launches per kind (HIP events on the launch stream).
//
//     valu_peak [iters] [reps]
#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstdlib>

namespace {

constexpr int CH = 8;

template <typename T>
__global__ __launch_bounds__(256) void k_fma(T* out, T a, T b, int iters) {
    T x[CH];
#pragma unroll
    for (int c = 0; c < CH; ++c) x[c] = (T)(threadIdx.x * 1e-3) + (T)c;
    for (int i = 0; i < iters; ++i) {
#pragma unroll
        for (int c = 0; c < CH; ++c) x[c] = fma(x[c], a, b);
    }
    T s = 0;
#pragma unroll
    for (int c = 0; c < CH; ++c) s += x[c];
    out[blockIdx.x * blockDim.x + threadIdx.x] = s;
}

__global__ __launch_bounds__(256) void k_mix(double* out, double a, double b, float af, float bf, int iters) {
    double x[CH];
    float y[CH];
#pragma unroll
    for (int c = 0; c < CH; ++c) {
        x[c] = threadIdx.x * 1e-3 + c;
        y[c] = (float)x[c];
    }
    for (int i = 0; i < iters; ++i) {
#pragma unroll
        for (int c = 0; c < CH; ++c) {
            x[c] = fma(x[c], a, b);
            y[c] = fmaf(y[c], af, bf);
        }
    }
    double s = 0;
#pragma unroll
    for (int c = 0; c < CH; ++c) s += x[c] + (double)y[c];
    out[blockIdx.x * blockDim.x + threadIdx.x] = s;
}

int check(hipError_t e, const char* what) {
    if (e != hipSuccess) {
        std::fprintf(stderr, "%s: %s\n", what, hipGetErrorString(e));
        std::exit(1);
    }
    return 0;
}

}  // namespace

int main(int argc, char** argv) {
    const int iters = argc > 1 ? std::atoi(argv[1]) : 200000;
    const int reps = argc > 2 ? std::atoi(argv[2]) : 5;
    if (iters < 1 || reps < 1) {
        std::fprintf(stderr, "usage: valu_peak [iters >= 1] [reps >= 1]\n");
        return 2;
    }
    int cus = 0, clk_khz = 0;
    check(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, 0), "CU count");
    check(hipDeviceGetAttribute(&clk_khz, hipDeviceAttributeClockRate, 0), "clock");
    const int threads = 256;
    const int blocks = cus * 4;  // 16 waves per CU, 4 per SIMD
    double* out = nullptr;
    check(hipMalloc(&out, sizeof(double) * (size_t)blocks * threads), "hipMalloc");
    hipEvent_t e0, e1;
    check(hipEventCreate(&e0), "event");
    check(hipEventCreate(&e1), "event");
    const double flops = 2.0 * CH * (double)iters * blocks * threads;
    double best[3] = {0, 0, 0};
    for (int r = 0; r < reps; ++r) {
        for (int kind = 0; kind < 3; ++kind) {
            check(hipEventRecord(e0, 0), "record");
            if (kind == 0)
                hipLaunchKernelGGL(k_fma<double>, dim3(blocks), dim3(threads), 0, 0, out, 0.9999999, 1e-9, iters);
            else if (kind == 1)
                hipLaunchKernelGGL(k_fma<float>, dim3(blocks), dim3(threads), 0, 0, (float*)out, 0.9999999f,
                                   1e-9f, iters);
            else
                hipLaunchKernelGGL(k_mix, dim3(blocks), dim3(threads), 0, 0, out, 0.9999999, 1e-9, 0.9999999f,
                                   1e-9f, iters);
            check(hipGetLastError(), "launch");
            check(hipEventRecord(e1, 0), "record");
            check(hipEventSynchronize(e1), "sync");
            float ms = 0;
            check(hipEventElapsedTime(&ms, e0, e1), "elapsed");
            const double tf = (kind == 2 ? 2.0 : 1.0) * flops / (ms * 1e-3) / 1e12;
            if (tf > best[kind]) best[kind] = tf;
        }
    }
    const double clk_peak = cus * 128.0 * clk_khz * 1e3 / 1e12;  // 64 lanes x FMA per CU and clock
    std::printf("{\"cus\": %d, \"clock_khz\": %d, \"iters\": %d, \"reps\": %d, "
                "\"fp64_fma_tflops\": %.3f, \"fp32_unpacked_fma_tflops\": %.3f, "
                "\"mix_1to1_tflops\": %.3f, \"one_fma_per_lane_per_clock_tflops\": %.3f}\n",
                cus, clk_khz, iters, reps, best[0], best[1], best[2], clk_peak);
    check(hipFree(out), "hipFree");
    return 0;
}
