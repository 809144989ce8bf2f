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

// one kind of operation per kernel: 0 add, 1 mul (x = x op a), on T
template <typename T, int OP>
__global__ __launch_bounds__(256) void k_op(T* out, T a, int iters) {
    T x[CH];
#pragma unroll
    for (int c = 0; c < CH; ++c) x[c] = (T)(threadIdx.x + 1) + (T)c;
    for (int i = 0; i < iters; ++i) {
#pragma unroll
        for (int c = 0; c < CH; ++c) {
            // integer adds chain through the next slot (a constant add
            // would fold into one multiply); floating ops cannot reassociate
            if constexpr (sizeof(T) == 4 && T(1) / T(2) == T(0)) x[c] = x[c] + x[(c + 1) % CH];
            else x[c] = OP == 0 ? x[c] + a : x[c] * a;
        }
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
    // wave instructions of one launch: CH per iteration per wave (x2 for the mix)
    const double wave_instr = (double)CH * iters * blocks * threads / 64.0;
    constexpr int kKinds = 8;
    const char* names[kKinds] = {"fp64_fma", "fp32_fma", "mix_fp64_fp32_fma", "fp64_add",
                                 "fp64_mul", "fp32_add", "fp32_mul", "int32_add"};
    double best_ms[kKinds];
    for (int k = 0; k < kKinds; ++k) best_ms[k] = 1e30;
    for (int r = 0; r < reps; ++r) {
        for (int kind = 0; kind < kKinds; ++kind) {
            check(hipEventRecord(e0, 0), "record");
            float* of = reinterpret_cast<float*>(out);
            int* oi = reinterpret_cast<int*>(out);
            switch (kind) {
                case 0: hipLaunchKernelGGL(k_fma<double>, dim3(blocks), dim3(threads), 0, 0, out, 0.9999999, 1e-9, iters); break;
                case 1: hipLaunchKernelGGL(k_fma<float>, dim3(blocks), dim3(threads), 0, 0, of, 0.9999999f, 1e-9f, iters); break;
                case 2: hipLaunchKernelGGL(k_mix, dim3(blocks), dim3(threads), 0, 0, out, 0.9999999, 1e-9, 0.9999999f, 1e-9f, iters); break;
                case 3: hipLaunchKernelGGL((k_op<double, 0>), dim3(blocks), dim3(threads), 0, 0, out, 1e-9, iters); break;
                case 4: hipLaunchKernelGGL((k_op<double, 1>), dim3(blocks), dim3(threads), 0, 0, out, 0.9999999, iters); break;
                case 5: hipLaunchKernelGGL((k_op<float, 0>), dim3(blocks), dim3(threads), 0, 0, of, 1e-9f, iters); break;
                case 6: hipLaunchKernelGGL((k_op<float, 1>), dim3(blocks), dim3(threads), 0, 0, of, 0.9999999f, iters); break;
                default: hipLaunchKernelGGL((k_op<int, 0>), dim3(blocks), dim3(threads), 0, 0, oi, 3, iters); break;
            }
            check(hipGetLastError(), "launch");
            check(hipEventRecord(e1, 0), "record");
            check(hipEventSynchronize(e1), "sync");
            float ms = 0;
            check(hipEventElapsedTime(&ms, e0, e1), "elapsed");
            if (ms < best_ms[kind]) best_ms[kind] = ms;
        }
    }
    const double flops = 2.0 * wave_instr * 64.0;
    const double clk_peak = cus * 128.0 * clk_khz * 1e3 / 1e12;  // 64 lanes x FMA per CU and clock
    std::printf("{\"cus\": %d, \"clock_khz\": %d, \"iters\": %d, \"reps\": %d, "
                "\"fp64_fma_tflops\": %.3f, \"fp32_unpacked_fma_tflops\": %.3f, "
                "\"mix_1to1_tflops\": %.3f, \"one_fma_per_lane_per_clock_tflops\": %.3f, "
                "\"simd_cycles_per_wave_instr_at_nominal_clock\": {",
                cus, clk_khz, iters, reps, flops / (best_ms[0] * 1e-3) / 1e12, flops / (best_ms[1] * 1e-3) / 1e12,
                2 * flops / (best_ms[2] * 1e-3) / 1e12, clk_peak);
    // SIMD cycles a wave instruction occupies: launch cycles x SIMDs / wave instructions
    for (int k = 0; k < kKinds; ++k) {
        const double n = wave_instr * (k == 2 ? 2.0 : 1.0);
        const double cyc = best_ms[k] * 1e-3 * clk_khz * 1e3 * cus * 4.0 / n;
        std::printf("%s\"%s\": %.3f", k ? ", " : "", names[k], cyc);
    }
    std::printf("}}\n");
    check(hipFree(out), "hipFree");
    return 0;
}
