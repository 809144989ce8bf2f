// Heavy-first order of an adaptive sweep's points.
//
// A persistent launch whose grid holds one to a few points per lane lasts as
// long as the lane that ends last: a heavy point taken in the second round
// finishes long after the rest.  A short probe run of every point (the same
// integration over a small fraction of its times) gives each point's step
// count; the points are then sorted by it (or by the largest count of their
// grid neighbourhood), heaviest first, and the full run takes them in that
// order.  Its outputs land in list order in staging
// buffers and are scattered back to point order.  Every point's integration
// is the same whatever the order, so the results are the same bits.
#include <hip/hip_runtime.h>

#include <hipcub/device/device_radix_sort.hpp>

#include "lyap_launch.hpp"

namespace lyap {
namespace {

__global__ void __launch_bounds__(256) k_order_keys(const uint32_t* __restrict__ n_acc,
                                                    const uint32_t* __restrict__ n_rej, int64_t first,
                                                    int64_t count, uint32_t* __restrict__ keys,
                                                    int64_t* __restrict__ vals) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= count) return;
    const uint64_t k = (uint64_t)n_acc[i] + (uint64_t)n_rej[i];
    keys[i] = k > 0xffffffffULL ? 0xffffffffu : (uint32_t)k;
    vals[i] = first + i;
}

__global__ void __launch_bounds__(256) k_order_scatter(const int64_t* __restrict__ idx, int64_t first,
                                                       int64_t count, int n_exps,
                                                       const double* __restrict__ lam_s,
                                                       const uint32_t* __restrict__ st_s,
                                                       const uint32_t* __restrict__ na_s,
                                                       const uint32_t* __restrict__ nr_s,
                                                       double* __restrict__ lam, uint32_t* __restrict__ st,
                                                       uint32_t* __restrict__ na, uint32_t* __restrict__ nr) {
    const int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= count) return;
    const int64_t p = idx[j] - first;
    for (int e = 0; e < n_exps; ++e) lam[p * n_exps + e] = lam_s[j * n_exps + e];
    if (st) st[p] = st_s[j];
    if (na) na[p] = na_s[j];
    if (nr) nr[p] = nr_s[j];
}

// keys of grid points: the largest raw key over the (2r+1)^2 grid points
// around each point that lie in the run's range [first, first + count) of the
// n1 x n2 grid (index p1 * n2 + p2).  Points whose neighbourhood holds a heavy
// point count as heavy, so that a point the short probe takes for light is
// not left for a late round when its neighbours are heavy.
__global__ void __launch_bounds__(256) k_order_nbr_max(const uint32_t* __restrict__ raw, int64_t first,
                                                       int64_t count, int32_t n2, int32_t r,
                                                       uint32_t* __restrict__ keys) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= count) return;
    const int64_t g = first + i;
    const int64_t g1 = g / n2, g2 = g % n2;
    uint32_t m = raw[i];
    for (int64_t a = g1 - r; a <= g1 + r; ++a) {
        for (int64_t b = g2 - r; b <= g2 + r; ++b) {
            if (b < 0 || b >= n2) continue;
            const int64_t j = a * n2 + b - first;
            if (a < 0 || j < 0 || j >= count) continue;
            const uint32_t v = raw[j];
            m = v > m ? v : m;
        }
    }
    keys[i] = m;
}

unsigned grid_of(int64_t count) { return (unsigned)((count + 255) / 256); }

}  // namespace

size_t order_sort_bytes(int64_t count) {
    size_t bytes = 0;
    if (hipcub::DeviceRadixSort::SortPairsDescending(nullptr, bytes, (const uint32_t*)nullptr,
                                                     (uint32_t*)nullptr, (const int64_t*)nullptr,
                                                     (int64_t*)nullptr, (int)count) != hipSuccess)
        return 0;
    return bytes;
}

hipError_t order_heavy_first(const uint32_t* n_acc, const uint32_t* n_rej, int64_t first, int64_t count,
                             int32_t n2, int32_t radius, uint32_t* keys_in, uint32_t* keys_out,
                             int64_t* vals_in, int64_t* vals_out, void* temp, size_t temp_bytes,
                             hipStream_t s) {
    // raw keys into keys_out (the sort's output, free until the sort)
    hipLaunchKernelGGL(k_order_keys, dim3(grid_of(count)), dim3(256), 0, s, n_acc, n_rej, first, count,
                       radius > 0 ? keys_out : keys_in, vals_in);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    if (radius > 0) {
        hipLaunchKernelGGL(k_order_nbr_max, dim3(grid_of(count)), dim3(256), 0, s, keys_out, first, count, n2,
                           radius, keys_in);
        if ((e = hipGetLastError()) != hipSuccess) return e;
    }
    return hipcub::DeviceRadixSort::SortPairsDescending(temp, temp_bytes, keys_in, keys_out, vals_in,
                                                        vals_out, (int)count, 0, 32, s);
}

hipError_t order_scatter(const int64_t* idx, int64_t first, int64_t count, int n_exps, const double* lam_s,
                         const uint32_t* st_s, const uint32_t* na_s, const uint32_t* nr_s, double* lam,
                         uint32_t* st, uint32_t* na, uint32_t* nr, hipStream_t s) {
    hipLaunchKernelGGL(k_order_scatter, dim3(grid_of(count)), dim3(256), 0, s, idx, first, count, n_exps,
                       lam_s, st_s, na_s, nr_s, lam, st, na, nr);
    return hipGetLastError();
}

}  // namespace lyap
