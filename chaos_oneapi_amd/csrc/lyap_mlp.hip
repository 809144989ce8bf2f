// Surrogate network of the sweep: (p1, p2) -> 2*F Fourier features -> three
// hidden layers of H GELU units -> [mu1, mu2, mu3, log-variance, escape logit].
//
// Weights are kept as an FP32 master copy plus two BF16 copies (W^T for the
// forward products, W for the input-gradient products); every matrix product
// runs on gfx950's v_mfma_f32_16x16x32_bf16 with FP32 accumulation in one NT
// GEMM kernel, C[M,N] = A[M,K] * B[N,K]^T (+ bias), both operands row-major so
// each lane fetches eight contiguous bf16 (16 bytes) per operand and k-step.  Activations are stored in
// both row- and column-major BF16 so weight gradients are NT products too.
// Adam runs on the FP32 master weights and re-issues the BF16 copies.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "chaos_lyap.h"
#include "lyap_device.hpp"
#include "lyap_launch.hpp"

typedef short v4s __attribute__((ext_vector_type(4)));
typedef __bf16 v8bf __attribute__((ext_vector_type(8)));
typedef float v4f __attribute__((ext_vector_type(4)));

namespace lyap {
namespace mlp {

constexpr int kOut = 5;       // mu1, mu2, mu3, log-variance, escape logit
constexpr int kOutPad = 16;   // head rows padded to one MFMA tile
constexpr int kLayers = 4;
constexpr int kMaxSplits = 16;  // split-K of the weight-gradient products
constexpr int kColsumChunks = 64;  // row chunks of the bias-gradient sums

LYAP_HD uint16_t f2bf(float f) {
    uint32_t u = __builtin_bit_cast(uint32_t, f);
    if ((u & 0x7fffffffu) > 0x7f800000u) return 0x7fc0;  // NaN
    u += 0x7fffu + ((u >> 16) & 1u);                     // round to nearest even
    return (uint16_t)(u >> 16);
}
LYAP_HD float bf2f(uint16_t b) { return __builtin_bit_cast(float, (uint32_t)b << 16); }

// ---------------------------------------------------------------------------
// NT GEMM on MFMA.  One wave computes a 16 x (16*NB) tile of C with gfx950's
// double-K BF16 product, v_mfma_f32_16x16x32_bf16: per 32-wide k-step, lane l
// holds A[m0 + l%16][k0 + 8(l/16) .. +7] and B[n0 + 16j + l%16][k0 + 8(l/16) .. +7]
// (one 16-byte load each); a k range whose length is an odd multiple of 16
// ends with one v_mfma_f32_16x16x16bf16_1k step (four bf16 per lane).  The
// accumulator element i of lane l is C[m0 + 4(l/16) + i][n0 + 16j + l%16].
// M, N multiples of 16 (of 16*NB for N), K and the k range multiples of 16.
// ---------------------------------------------------------------------------
// acc[j] += A-tile . B-tile_j over k in [k_lo, k_hi).  a_row / b_row are the
// lane's rows (A row m0 + l%16, B row n0 + l%16) without the k offset; q = l/16.
// Four k-steps' loads are issued before their MFMAs (one step's loads would
// otherwise wait out the L2 latency alone), then a step at a time for the
// rest; each accumulator takes the k-steps in increasing order.
template <int NB>
__device__ __forceinline__ void mfma_k_range(const uint16_t* __restrict__ a_row,
                                             const uint16_t* __restrict__ b_row, int q, int K, int k_lo,
                                             int k_hi, v4f (&acc)[NB]) {
    const uint16_t* a8 = a_row + 8 * q;
    const uint16_t* b8 = b_row + 8 * q;
    int k0 = k_lo;
    for (; k0 + 128 <= k_hi; k0 += 128) {
        v8bf a[4], b[4][NB];
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            a[u] = *reinterpret_cast<const v8bf*>(a8 + k0 + 32 * u);
#pragma unroll
            for (int j = 0; j < NB; ++j)
                b[u][j] = *reinterpret_cast<const v8bf*>(b8 + (size_t)(16 * j) * K + k0 + 32 * u);
        }
#pragma unroll
        for (int u = 0; u < 4; ++u)
#pragma unroll
            for (int j = 0; j < NB; ++j) acc[j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[u], b[u][j], acc[j], 0, 0, 0);
    }
    for (; k0 + 32 <= k_hi; k0 += 32) {
        const v8bf a = *reinterpret_cast<const v8bf*>(a8 + k0);
#pragma unroll
        for (int j = 0; j < NB; ++j) {
            const v8bf b = *reinterpret_cast<const v8bf*>(b8 + (size_t)(16 * j) * K + k0);
            acc[j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, acc[j], 0, 0, 0);
        }
    }
    if (k0 < k_hi) {  // a last 16-wide step
        const v4s a = *reinterpret_cast<const v4s*>(a_row + 4 * q + k0);
#pragma unroll
        for (int j = 0; j < NB; ++j) {
            const v4s b = *reinterpret_cast<const v4s*>(b_row + 4 * q + (size_t)(16 * j) * K + k0);
            acc[j] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, acc[j], 0, 0, 0);
        }
    }
}

__device__ __forceinline__ float gelu(float x) { return 0.5f * x * (1.f + erff(x * 0.70710678118654752f)); }

// With Hr (forward products of the hidden layers) the epilogue also stores
// H = gelu(C) in row-major Hr[M][N] and column-major Ht[N][M] BF16.
template <int NB>
__global__ __launch_bounds__(256) void k_gemm_nt(const uint16_t* __restrict__ A,
                                                 const uint16_t* __restrict__ B, float* __restrict__ C,
                                                 const float* __restrict__ bias, int M, int N, int K,
                                                 float beta, uint16_t* __restrict__ Hr,
                                                 uint16_t* __restrict__ Ht) {
    const int wave = (int)(blockIdx.x * 4 + (threadIdx.x >> 6));
    const int lane = (int)(threadIdx.x & 63);
    const int tiles_n = N / (16 * NB);
    const int m0 = (wave / tiles_n) * 16;
    const int n0 = (wave % tiles_n) * (16 * NB);
    if (m0 >= M) return;
    const int r = lane & 15, q = lane >> 4;
    v4f acc[NB];
#pragma unroll
    for (int j = 0; j < NB; ++j) acc[j] = v4f{0.f, 0.f, 0.f, 0.f};
    mfma_k_range<NB>(A + (size_t)(m0 + r) * K, B + (size_t)(n0 + r) * K, q, K, 0, K, acc);
#pragma unroll
    for (int j = 0; j < NB; ++j) {
        const int col = n0 + 16 * j + r;
        const float bv = bias ? bias[col] : 0.f;
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            float* c = C + (size_t)(m0 + 4 * q + i) * N + col;
            const float z = (beta != 0.f ? beta * *c : 0.f) + acc[j][i] + bv;
            *c = z;
            if (Hr) {
                const int row = m0 + 4 * q + i;
                const uint16_t h = f2bf(gelu(z));
                Hr[(size_t)row * N + col] = h;
                Ht[(size_t)col * M + row] = h;
            }
        }
    }
}

// Weight gradient (small M x N, K = the batch) with the split-K inside one
// workgroup: wave s of block t computes tile t over k in [s*kc, min(K,
// (s+1)*kc)); the partial tiles meet in LDS and are added in split order, the
// same sums as separate split-K partial products added in order.  With bias_part
// ([chunks][M] row-chunk sums of the layer's dZ), the blocks of the first tile
// column also add the bias gradient of their 16 rows in chunk order.
// wgrad_tile is the body for tile t, on a block of 64 * splits threads with
// `part` its LDS ([splits][16][16 * NB], then [kColsumChunks][16]).
template <int NB>
__device__ __forceinline__ void wgrad_tile(int t, float* part, const uint16_t* __restrict__ A,
                                           const uint16_t* __restrict__ B, float* __restrict__ C, int M,
                                           int N, int K, int kc, int splits,
                                           const float* __restrict__ bias_part, int chunks,
                                           float* __restrict__ db) {
    constexpr int TW = 16 * NB;
    const int split = (int)(threadIdx.x >> 6);
    const int lane = (int)(threadIdx.x & 63);
    const int tiles_n = N / TW;
    const int m0 = (t / tiles_n) * 16;
    const int n0 = (t % tiles_n) * TW;
    const int r = lane & 15, q = lane >> 4;
    // the bias chunk sums of rows m0 .. m0+15 are fetched before the products
    // (their latency hides behind the MFMA loop) and parked in LDS after them
    const bool do_bias = bias_part != nullptr && n0 == 0;
    const int nb_vals = 16 * chunks;
    float bv[2] = {0.f, 0.f};
    if (do_bias) {
#pragma unroll
        for (int u = 0; u < 2; ++u) {
            const int e = (int)threadIdx.x + u * (int)blockDim.x;
            if (e < nb_vals) bv[u] = bias_part[(size_t)(e >> 4) * M + m0 + (e & 15)];
        }
    }
    v4f acc[NB];
#pragma unroll
    for (int j = 0; j < NB; ++j) acc[j] = v4f{0.f, 0.f, 0.f, 0.f};
    const int k_lo = split * kc, k_hi = min(K, k_lo + kc);
    mfma_k_range<NB>(A + (size_t)(m0 + r) * K, B + (size_t)(n0 + r) * K, q, K, k_lo, k_hi, acc);
    float* P = part + split * 16 * TW;
#pragma unroll
    for (int j = 0; j < NB; ++j)
#pragma unroll
        for (int i = 0; i < 4; ++i) P[(4 * q + i) * TW + 16 * j + r] = acc[j][i];
    float* bl = part + splits * 16 * TW;
    if (do_bias) {
#pragma unroll
        for (int u = 0; u < 2; ++u) {
            const int e = (int)threadIdx.x + u * (int)blockDim.x;
            if (e < nb_vals) bl[e] = bv[u];
        }
        // more chunk sums than two per thread (few splits): the rest directly
        for (int e = (int)threadIdx.x + 2 * (int)blockDim.x; e < nb_vals; e += (int)blockDim.x)
            bl[e] = bias_part[(size_t)(e >> 4) * M + m0 + (e & 15)];
    }
    __syncthreads();
    for (int e = (int)threadIdx.x; e < 16 * TW; e += (int)blockDim.x) {
        float s = part[e];
        for (int k = 1; k < splits; ++k) s += part[k * 16 * TW + e];
        C[(size_t)(m0 + e / TW) * N + n0 + e % TW] = s;
    }
    if (do_bias && threadIdx.x < 16) {
        float s = bl[threadIdx.x];
        for (int k = 1; k < chunks; ++k) s += bl[k * 16 + threadIdx.x];
        db[m0 + threadIdx.x] = s;
    }
}

template <int NB>
__global__ __launch_bounds__(1024) void k_wgrad(const uint16_t* __restrict__ A,
                                                const uint16_t* __restrict__ B, float* __restrict__ C,
                                                int M, int N, int K, int kc, int splits,
                                                const float* __restrict__ bias_part, int chunks,
                                                float* __restrict__ db) {
    extern __shared__ float part[];
    wgrad_tile<NB>((int)blockIdx.x, part, A, B, C, M, N, K, kc, splits, bias_part, chunks, db);
}


// ---------------------------------------------------------------------------
// elementwise kernels
// ---------------------------------------------------------------------------
__device__ __forceinline__ float gelu_grad(float x) {
    return 0.5f * (1.f + erff(x * 0.70710678118654752f)) +
           x * 0.39894228040143268f * expf(-0.5f * x * x);
}

// features[b][j] = sin(2 pi (B_j . u)), features[b][F+j] = cos(...); row- and column-major bf16.
// One thread per (feature j, row b), rows fastest (the column-major stores
// coalesce).  (loss, when given, is cleared for the loss reduction of the same pass)
__global__ void k_fourier(const double* __restrict__ p, int n, int n_valid, const float* __restrict__ fb,
                          int F, double lo1, double inv1, double lo2, double inv2,
                          uint16_t* __restrict__ X, uint16_t* __restrict__ Xt, float* __restrict__ loss) {
    const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (loss && idx == 0) *loss = 0.f;
    if (idx >= (int64_t)n * F) return;
    const int j = (int)(idx / n), b = (int)(idx % n);
    float u0 = 0.f, u1 = 0.f;
    if (b < n_valid) {
        u0 = (float)((p[2 * b] - lo1) * inv1);
        u1 = (float)((p[2 * b + 1] - lo2) * inv2);
    }
    const float f = 6.283185307179586f * (fb[2 * j] * u0 + fb[2 * j + 1] * u1);
    float s, c;
    sincosf(f, &s, &c);
    const uint16_t sb = f2bf(s), cb = f2bf(c);
    X[(size_t)b * 2 * F + j] = sb;
    X[(size_t)b * 2 * F + F + j] = cb;
    Xt[(size_t)j * n + b] = sb;
    Xt[(size_t)(F + j) * n + b] = cb;
}

// dZ = dH * gelu'(Z) in row- and column-major bf16, and the partial column
// sums of dZ (as rounded to bf16) for the bias gradient: block (x, r) covers
// columns 64x .. 64x+63 and rows [r*rc, min(n, (r+1)*rc)); thread (c, g) takes
// rows g, g+4, ... of the chunk in order, the four row groups are added in a
// fixed tree and the chunk's sum goes to part[r][j] (k_wgrad of the layer adds
// the chunks in order, so the bias gradient is deterministic).
__global__ __launch_bounds__(256) void k_gelu_bwd_colsum(const float* __restrict__ dH,
                                                         const float* __restrict__ Z, int n, int H, int rc,
                                                         uint16_t* __restrict__ dZr, uint16_t* __restrict__ dZt,
                                                         float* __restrict__ part) {
    __shared__ float red[4][64];
    const int c = (int)(threadIdx.x & 63), g = (int)(threadIdx.x >> 6);
    const int j = (int)blockIdx.x * 64 + c;
    const int r = (int)blockIdx.y;
    const int lo = r * rc, hi = min(n, lo + rc);
    float s = 0.f;
    if (j < H)
        for (int b = lo + g; b < hi; b += 4) {
            const size_t idx = (size_t)b * H + j;
            const uint16_t v = f2bf(dH[idx] * gelu_grad(Z[idx]));
            dZr[idx] = v;
            dZt[(size_t)j * n + b] = v;
            s += bf2f(v);
        }
    red[g][c] = s;
    __syncthreads();
    if (g < 2) red[g][c] += red[g + 2][c];
    __syncthreads();
    if (g == 0 && j < H) part[(size_t)r * H + j] = red[0][c] + red[1][c];
}

// Input gradient through the GELU of the layer below, in one launch:
// dH[B][N] = dZ[B][K] . W[N][K]^T on 16x16 MFMA tiles, a block's four waves
// covering one 64-row chunk of one 16-column group, then dZ' = bf16(dH *
// gelu'(Z)) to row- and column-major dZr / dZt and the chunk's bias-gradient
// column sums to part[chunk][col] -- the same operations, in the same order,
// as the product followed by k_gelu_bwd_colsum with 64-row chunks.  dZr must
// not alias dZ (other blocks are still reading it).
// the body of k_gemm_dgelu for virtual block vb, run by 256 threads (tid
// 0..255) with their own LDS tile and reduction rows; every thread of the
// workgroup reaches its three barriers
__device__ __forceinline__ void dgelu_chunk(int vb, int tid, float (*tile)[17], float (*red)[16],
                                            const uint16_t* __restrict__ dZ, const uint16_t* __restrict__ W,
                                            const float* __restrict__ Z, int M, int N, int K,
                                            uint16_t* __restrict__ dZr, uint16_t* __restrict__ dZt,
                                            float* __restrict__ part) {
    const int w = tid >> 6, lane = tid & 63;
    const int groups_n = N / 16;
    const int chunk = vb / groups_n;
    const int n0 = (vb % groups_n) * 16;
    const int m0 = chunk * 64 + w * 16;
    const int r = lane & 15, q = lane >> 4;
    v4f acc[1] = {v4f{0.f, 0.f, 0.f, 0.f}};
    mfma_k_range<1>(dZ + (size_t)(m0 + r) * K, W + (size_t)(n0 + r) * K, q, K, 0, K, acc);
    const int col = n0 + r;
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        const int row = m0 + 4 * q + i;
        const float dh = 0.f + acc[0][i] + 0.f;  // the stored dH of the two-kernel form
        const size_t idx = (size_t)row * N + col;
        const uint16_t v = f2bf(dh * gelu_grad(Z[idx]));
        dZr[idx] = v;
        dZt[(size_t)col * M + row] = v;
        tile[w * 16 + 4 * q + i][r] = bf2f(v);
    }
    __syncthreads();
    if (tid < 64) {
        const int c = tid & 15, g = tid >> 4;
        float s = 0.f;
        for (int b = g; b < 64; b += 4) s += tile[b][c];
        red[g][c] = s;
    }
    __syncthreads();
    if (tid < 32) {
        const int c = tid & 15, g = tid >> 4;
        red[g][c] += red[g + 2][c];
    }
    __syncthreads();
    if (tid < 16) part[(size_t)chunk * N + n0 + tid] = red[0][tid] + red[1][tid];
}

__global__ __launch_bounds__(256) void k_gemm_dgelu(const uint16_t* __restrict__ dZ,
                                                    const uint16_t* __restrict__ W,
                                                    const float* __restrict__ Z, int M, int N, int K,
                                                    uint16_t* __restrict__ dZr, uint16_t* __restrict__ dZt,
                                                    float* __restrict__ part) {
    __shared__ float tile[64][17];
    __shared__ float red[4][16];
    dgelu_chunk((int)blockIdx.x, (int)threadIdx.x, tile, red, dZ, W, Z, M, N, K, dZr, dZt, part);
}

// One launch for a layer's weight gradient (k_wgrad, 16x16 tiles, eight
// splits: 512-thread blocks) and the input gradient of the layer below
// (k_gemm_dgelu, two 64-row chunks a block): the two products are independent
// and each is latency-bound, so side by side they take about as long as the
// longer one.  Blocks [0, wgrad_blocks) are weight-gradient tiles.
struct WgradArgs {
    const uint16_t* A;
    const uint16_t* B;
    float* C;
    int M, N, K, kc, splits;
    const float* bias_part;
    int chunks;
    float* db;
};
struct DgeluArgs {
    const uint16_t* dZ;
    const uint16_t* W;
    const float* Z;
    int M, N, K;
    uint16_t* dZr;
    uint16_t* dZt;
    float* part;
};
__global__ __launch_bounds__(512) void k_wgrad_dgelu(WgradArgs wa, DgeluArgs da, int wgrad_blocks) {
    extern __shared__ float lds[];
    if ((int)blockIdx.x < wgrad_blocks) {
        wgrad_tile<1>((int)blockIdx.x, lds, wa.A, wa.B, wa.C, wa.M, wa.N, wa.K, wa.kc, wa.splits, wa.bias_part,
                      wa.chunks, wa.db);
        return;
    }
    const int half = (int)(threadIdx.x >> 8);
    float(*tile)[17] = reinterpret_cast<float(*)[17]>(lds + half * (64 * 17 + 64));
    float(*red)[16] = reinterpret_cast<float(*)[16]>(lds + half * (64 * 17 + 64) + 64 * 17);
    dgelu_chunk(((int)blockIdx.x - wgrad_blocks) * 2 + half, (int)(threadIdx.x & 255), tile, red, da.dZ, da.W,
                da.Z, da.M, da.N, da.K, da.dZr, da.dZt, da.part);
}

// loss and head gradient: Gaussian NLL on mu1 with the learned log-variance,
// squared error on mu2, mu3 (targets normalised), logistic loss on escape.
// Rows b >= n_valid are padding (zero gradient).
// the loss of row b and its gradient g[kOutPad] (zero for padding rows)
__device__ __forceinline__ float head_row(int b, const float* __restrict__ Z, const double* __restrict__ lam,
                                          const uint32_t* __restrict__ status, int n_valid,
                                          const double* __restrict__ ym, const double* __restrict__ ys,
                                          float inv_n, float (&g)[kOutPad]) {
#pragma unroll
    for (int j = 0; j < kOutPad; ++j) g[j] = 0.f;
    float l = 0.f;
    if (b < n_valid) {
        const float* z = Z + (size_t)b * kOutPad;
        const float bad = status[b] != 0 ? 1.f : 0.f;
        const float zl = z[4];
        const float sp = zl > 0.f ? zl + log1pf(expf(-zl)) : log1pf(expf(zl));
        const float sig = 1.f / (1.f + expf(-zl));
        l += sp - bad * zl;
        g[4] = (sig - bad) * inv_n;
        if (bad == 0.f) {
            const float y1 = (float)((lam[3 * b] - ym[0]) / ys[0]);
            const float y2 = (float)((lam[3 * b + 1] - ym[1]) / ys[1]);
            const float y3 = (float)((lam[3 * b + 2] - ym[2]) / ys[2]);
            const float s = z[3], e = expf(-s);
            const float d1 = z[0] - y1, d2 = z[1] - y2, d3 = z[2] - y3;
            l += 0.5f * (d1 * d1 * e + s) + 0.5f * (d2 * d2 + d3 * d3);
            g[0] = d1 * e * inv_n;
            g[1] = d2 * inv_n;
            g[2] = d3 * inv_n;
            g[3] = 0.5f * (1.f - d1 * d1 * e) * inv_n;
        }
    }
    return l;
}

__global__ void k_head_loss(const float* __restrict__ Z, const double* __restrict__ lam,
                            const uint32_t* __restrict__ status, int n, int n_valid,
                            const double* __restrict__ ym, const double* __restrict__ ys, float inv_n,
                            uint16_t* __restrict__ dZr, uint16_t* __restrict__ dZt,
                            float* __restrict__ dZf, float* __restrict__ loss) {
    const int b = (int)(blockIdx.x * blockDim.x + threadIdx.x);
    if (b >= n) return;
    float g[kOutPad];
    const float l = head_row(b, Z, lam, status, n_valid, ym, ys, inv_n, g);
#pragma unroll
    for (int j = 0; j < kOutPad; ++j) {
        const uint16_t gb = f2bf(g[j]);
        dZr[(size_t)b * kOutPad + j] = gb;
        dZt[(size_t)j * n + b] = gb;
        dZf[(size_t)b * kOutPad + j] = g[j];
    }
    if (loss && b < n_valid) atomicAdd(loss, l * inv_n);
}

// The head in one launch (training): Z = H . W^T + bias on 16x16 MFMA tiles,
// a block's four waves covering one 64-row chunk, then each row's loss and
// gradient (as k_head_loss) and the chunk's bias-gradient column sums (as
// k_colsum_part with 64-row chunks) -- the same operations in the same order.
__global__ __launch_bounds__(256) void k_gemm_head_loss(const uint16_t* __restrict__ A,
                                                        const uint16_t* __restrict__ W,
                                                        const float* __restrict__ bias, int n, int K,
                                                        float* __restrict__ Z, const double* __restrict__ lam,
                                                        const uint32_t* __restrict__ status, int n_valid,
                                                        const double* __restrict__ ym,
                                                        const double* __restrict__ ys, float inv_n,
                                                        uint16_t* __restrict__ dZr, uint16_t* __restrict__ dZt,
                                                        float* __restrict__ dZf, float* __restrict__ loss,
                                                        float* __restrict__ part) {
    __shared__ float gt[64][kOutPad + 1];
    __shared__ float red[4][kOutPad];
    const int w = (int)(threadIdx.x >> 6), lane = (int)(threadIdx.x & 63);
    const int m0 = (int)blockIdx.x * 64 + w * 16;
    const int r = lane & 15, q = lane >> 4;
    v4f acc[1] = {v4f{0.f, 0.f, 0.f, 0.f}};
    mfma_k_range<1>(A + (size_t)(m0 + r) * K, W + (size_t)r * K, q, K, 0, K, acc);
    const float bv = bias[r];
#pragma unroll
    for (int i = 0; i < 4; ++i) Z[(size_t)(m0 + 4 * q + i) * kOutPad + r] = 0.f + acc[0][i] + bv;
    __syncthreads();  // the block's Z rows are read back below (same workgroup)
    if (threadIdx.x < 64) {
        const int b = (int)blockIdx.x * 64 + (int)threadIdx.x;
        float g[kOutPad];
        const float l = head_row(b, Z, lam, status, n_valid, ym, ys, inv_n, g);
#pragma unroll
        for (int j = 0; j < kOutPad; ++j) {
            const uint16_t gb = f2bf(g[j]);
            dZr[(size_t)b * kOutPad + j] = gb;
            dZt[(size_t)j * n + b] = gb;
            dZf[(size_t)b * kOutPad + j] = g[j];
            gt[threadIdx.x][j] = g[j];
        }
        if (loss && b < n_valid) atomicAdd(loss, l * inv_n);
    }
    __syncthreads();
    if (threadIdx.x < 64) {
        const int c = (int)(threadIdx.x & 15), g = (int)(threadIdx.x >> 4);
        float s = 0.f;
        for (int b = g; b < 64; b += 4) s += gt[b][c];
        red[g][c] = s;
    }
    __syncthreads();
    if (threadIdx.x < 32) {
        const int c = (int)(threadIdx.x & 15), g = (int)(threadIdx.x >> 4);
        red[g][c] += red[g + 2][c];
    }
    __syncthreads();
    if (threadIdx.x < kOutPad) part[(size_t)blockIdx.x * kOutPad + threadIdx.x] = red[0][threadIdx.x] + red[1][threadIdx.x];
}

// Partial column sums of dZ [n][H] over row chunks: block (x, r) covers
// columns 64x .. 64x+63 and rows [r*rc, min(n, (r+1)*rc)); thread (c, g) sums
// rows g, g+4, ... of the chunk in order (coalesced across c), the four row
// groups are added in a fixed tree and the chunk's sum goes to part[r][j].
// k_wgrad of the head then adds the chunks in order, so db is deterministic.
__global__ __launch_bounds__(256) void k_colsum_part(const float* __restrict__ dZ, int n, int H, int rc,
                                                     float* __restrict__ part) {
    __shared__ float red[4][64];
    const int c = (int)(threadIdx.x & 63), g = (int)(threadIdx.x >> 6);
    const int j = (int)blockIdx.x * 64 + c;
    const int r = (int)blockIdx.y;
    const int lo = r * rc, hi = min(n, lo + rc);
    float s = 0.f;
    if (j < H)
        for (int b = lo + g; b < hi; b += 4) s += dZ[(size_t)b * H + j];
    red[g][c] = s;
    __syncthreads();
    if (g < 2) red[g][c] += red[g + 2][c];
    __syncthreads();
    if (g == 0 && j < H) part[(size_t)r * H + j] = red[0][c] + red[1][c];
}

// where each layer's weight matrix [O][I] sits in the parameter vector, and
// its BF16 copies (Wt [O][I], Wn [I][O])
struct WeightMap {
    int64_t off[kLayers];
    int O[kLayers], I[kLayers];
    uint16_t* wt[kLayers];
    uint16_t* wn[kLayers];
};

// Adam on every trainable parameter, then the new value of every weight entry
// issued to its BF16 copies (one pass over the parameters)
__global__ void k_adam_issue(float* __restrict__ w, const float* __restrict__ g, float* __restrict__ m,
                             float* __restrict__ v, const uint8_t* __restrict__ trainable, int64_t n,
                             float lr, float b1, float b2, float eps, float c1, float c2, WeightMap wm) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    float wi = w[i];
    if (trainable[i]) {
        const float gi = g[i];
        const float mi = b1 * m[i] + (1.f - b1) * gi;
        const float vi = b2 * v[i] + (1.f - b2) * gi * gi;
        m[i] = mi;
        v[i] = vi;
        wi -= lr * (mi / c1) / (sqrtf(vi / c2) + eps);
        w[i] = wi;
    }
#pragma unroll
    for (int l = 0; l < kLayers; ++l) {
        const int64_t k = i - wm.off[l];
        if (k >= 0 && k < (int64_t)wm.O[l] * wm.I[l]) {
            const int o = (int)(k / wm.I[l]), c = (int)(k % wm.I[l]);
            const uint16_t bv = f2bf(wi);
            wm.wt[l][k] = bv;
            wm.wn[l][(size_t)c * wm.O[l] + o] = bv;
        }
    }
}

// master [O][I] fp32 -> Wt [O][I] bf16 and Wn [I][O] bf16
__global__ void k_issue_bf16(const float* __restrict__ w, int O, int I, uint16_t* __restrict__ wt,
                             uint16_t* __restrict__ wn) {
    const int idx = (int)(blockIdx.x * blockDim.x + threadIdx.x);
    if (idx >= O * I) return;
    const int o = idx / I, i = idx % I;
    const uint16_t b = f2bf(w[idx]);
    wt[idx] = b;
    wn[(size_t)i * O + o] = b;
}

// inference output: de-normalised means, std of lambda1, escape probability
__global__ void k_head_out(const float* __restrict__ Z, int n_valid, const double* __restrict__ ym,
                           const double* __restrict__ ys, float* __restrict__ out) {
    const int b = (int)(blockIdx.x * blockDim.x + threadIdx.x);
    if (b >= n_valid) return;
    const float* z = Z + (size_t)b * kOutPad;
    float* o = out + (size_t)b * kOut;
    o[0] = (float)(z[0] * ys[0] + ym[0]);
    o[1] = (float)(z[1] * ys[1] + ym[1]);
    o[2] = (float)(z[2] * ys[2] + ym[2]);
    o[3] = (float)(expf(0.5f * z[3]) * ys[0]);
    o[4] = 1.f / (1.f + expf(-z[4]));
}

}  // namespace mlp
}  // namespace lyap

// ---------------------------------------------------------------------------
// C-ABI object
// ---------------------------------------------------------------------------
using namespace lyap::mlp;

namespace {

template <class... A>
int mfail(int code, const char* fmt, A... args) {
    return lyap::set_error(code, fmt, args...);
}

#define MLP_HIP(call)                                                               \
    do {                                                                            \
        hipError_t e_ = (call);                                                     \
        if (e_ != hipSuccess) return mfail(LYAP_ERR_HIP, "%s: %s", #call, hipGetErrorString(e_)); \
    } while (0)

}  // namespace

struct lyap_mlp {
    int device = 0;
    lyap_mlp_config cfg{};
    int dims[kLayers + 1];  // in, hidden, hidden, hidden, out_pad
    int64_t off_w[kLayers], off_b[kLayers], n_params = 0;
    float* master = nullptr;
    float* grad = nullptr;
    float* adam_m = nullptr;
    float* adam_v = nullptr;
    uint8_t* trainable = nullptr;
    uint16_t* wt[kLayers] = {};
    uint16_t* wn[kLayers] = {};
    float* fourier = nullptr;
    double* norm = nullptr;  // y_mean[3], y_std[3]
    // workspace for max_batch rows
    uint16_t *act[kLayers] = {}, *act_t[kLayers] = {};  // inputs of layer l (row, column major)
    float* z[kLayers] = {};                              // pre-activations
    uint16_t *dz = nullptr, *dz_t = nullptr;
    uint16_t *dz2 = nullptr, *dz_t2 = nullptr;  // the other dZ buffers (the input-gradient pass reads one pair, writes the other)
    float* dh = nullptr;
    float* dzf = nullptr;
    float* loss = nullptr;
    float* split_ws = nullptr;  // the head's bias-gradient chunk sums
    float *col_ws = nullptr, *col_ws2 = nullptr;  // bias-gradient chunk sums of the hidden layers (a pair)
    int64_t step = 0;
    std::vector<void*> allocs;
};

extern "C" {

static int mlp_alloc(lyap_mlp* m, void** p, size_t bytes) {
    hipError_t e = hipMalloc(p, bytes);
    if (e != hipSuccess) return mfail(LYAP_ERR_HIP, "hipMalloc(%zu): %s", bytes, hipGetErrorString(e));
    m->allocs.push_back(*p);
    return LYAP_SUCCESS;
}

int lyap_mlp_destroy(lyap_mlp* m) {
    if (!m) return LYAP_SUCCESS;
    (void)hipSetDevice(m->device);
    for (void* p : m->allocs) (void)hipFree(p);
    delete m;
    return LYAP_SUCCESS;
}

int lyap_mlp_create(int device, const lyap_mlp_config* cfg, lyap_mlp** out) {
    if (!cfg || !out) return mfail(LYAP_ERR_INVALID, "lyap_mlp_create: NULL argument");
    *out = nullptr;
    if (cfg->max_batch <= 0 || cfg->max_batch % 64 != 0)
        return mfail(LYAP_ERR_INVALID, "max_batch must be a positive multiple of 64");
    if (cfg->hidden <= 0 || cfg->hidden % 64 != 0)
        return mfail(LYAP_ERR_INVALID, "hidden must be a positive multiple of 64");
    if (cfg->n_fourier <= 0 || (2 * cfg->n_fourier) % 16 != 0)
        return mfail(LYAP_ERR_INVALID, "2*n_fourier must be a multiple of 16");
    if (!(cfg->hi1 > cfg->lo1) || !(cfg->hi2 > cfg->lo2))
        return mfail(LYAP_ERR_INVALID, "input ranges must be non-empty");
    MLP_HIP(hipSetDevice(device));
    lyap_mlp* m = new lyap_mlp();
    m->device = device;
    m->cfg = *cfg;
    const int F2 = 2 * cfg->n_fourier, H = cfg->hidden, B = cfg->max_batch;
    int dims[kLayers + 1] = {F2, H, H, H, kOutPad};
    std::memcpy(m->dims, dims, sizeof(dims));
    for (int l = 0; l < kLayers; ++l) {
        m->off_w[l] = m->n_params;
        m->n_params += (int64_t)dims[l + 1] * dims[l];
        m->off_b[l] = m->n_params;
        m->n_params += dims[l + 1];
    }
    int rc = LYAP_SUCCESS;
    const size_t P = (size_t)m->n_params;
    rc |= mlp_alloc(m, (void**)&m->master, P * 4);
    rc |= mlp_alloc(m, (void**)&m->grad, P * 4);
    rc |= mlp_alloc(m, (void**)&m->adam_m, P * 4);
    rc |= mlp_alloc(m, (void**)&m->adam_v, P * 4);
    rc |= mlp_alloc(m, (void**)&m->trainable, P);
    rc |= mlp_alloc(m, (void**)&m->fourier, (size_t)cfg->n_fourier * 2 * 4);
    rc |= mlp_alloc(m, (void**)&m->norm, 6 * 8);
    for (int l = 0; l < kLayers; ++l) {
        const size_t w = (size_t)dims[l + 1] * dims[l];
        rc |= mlp_alloc(m, (void**)&m->wt[l], w * 2);
        rc |= mlp_alloc(m, (void**)&m->wn[l], w * 2);
        rc |= mlp_alloc(m, (void**)&m->act[l], (size_t)B * dims[l] * 2);
        rc |= mlp_alloc(m, (void**)&m->act_t[l], (size_t)B * dims[l] * 2);
        rc |= mlp_alloc(m, (void**)&m->z[l], (size_t)B * dims[l + 1] * 4);
    }
    rc |= mlp_alloc(m, (void**)&m->dz, (size_t)B * H * 2);
    rc |= mlp_alloc(m, (void**)&m->dz2, (size_t)B * H * 2);
    rc |= mlp_alloc(m, (void**)&m->dz_t2, (size_t)B * H * 2);
    rc |= mlp_alloc(m, (void**)&m->dz_t, (size_t)B * H * 2);
    rc |= mlp_alloc(m, (void**)&m->dh, (size_t)B * H * 4);
    rc |= mlp_alloc(m, (void**)&m->dzf, (size_t)B * kOutPad * 4);  // the head's dZ in fp32
    rc |= mlp_alloc(m, (void**)&m->loss, 4);
    size_t max_d = 0;
    for (int l = 0; l < kLayers; ++l) max_d = std::max(max_d, (size_t)dims[l + 1]);
    rc |= mlp_alloc(m, (void**)&m->split_ws, (size_t)kColsumChunks * max_d * 4);
    rc |= mlp_alloc(m, (void**)&m->col_ws, (size_t)kColsumChunks * max_d * 4);
    rc |= mlp_alloc(m, (void**)&m->col_ws2, (size_t)kColsumChunks * max_d * 4);
    if (rc != LYAP_SUCCESS) {
        lyap_mlp_destroy(m);
        return LYAP_ERR_HIP;
    }
    // seeded initialisation on the host: U(-sqrt(3/fan_in), sqrt(3/fan_in)) weights,
    // zero biases, head rows beyond the 5 outputs frozen at zero
    std::vector<float> w(P, 0.f);
    std::vector<uint8_t> tr(P, 1);
    for (int l = 0; l < kLayers; ++l) {
        const float a = std::sqrt(3.0f / (float)dims[l]) * (l == kLayers - 1 ? 0.5f : 1.0f);
        for (int o = 0; o < dims[l + 1]; ++o)
            for (int i = 0; i < dims[l]; ++i) {
                const int64_t k = m->off_w[l] + (int64_t)o * dims[l] + i;
                const double u = lyap::uniform01(cfg->seed, k, 7);
                w[k] = (float)((2.0 * u - 1.0) * a);
                if (l == kLayers - 1 && o >= kOut) {
                    w[k] = 0.f;
                    tr[k] = 0;
                }
            }
        if (l == kLayers - 1)
            for (int o = kOut; o < kOutPad; ++o) tr[m->off_b[l] + o] = 0;
    }
    std::vector<float> fb((size_t)cfg->n_fourier * 2);
    for (size_t j = 0; j < fb.size(); ++j) {
        double z = 0.0;
        for (int r = 0; r < 12; ++r) z += lyap::uniform01(cfg->seed, (int64_t)j, 100 + r);
        fb[j] = (float)((z - 6.0) * cfg->fourier_sigma);
    }
    double norm[6] = {cfg->y_mean[0], cfg->y_mean[1], cfg->y_mean[2],
                      cfg->y_std[0] > 0 ? cfg->y_std[0] : 1.0, cfg->y_std[1] > 0 ? cfg->y_std[1] : 1.0,
                      cfg->y_std[2] > 0 ? cfg->y_std[2] : 1.0};
    hipError_t e = hipMemcpy(m->master, w.data(), P * 4, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(m->trainable, tr.data(), P, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(m->fourier, fb.data(), fb.size() * 4, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(m->norm, norm, sizeof(norm), hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemset(m->adam_m, 0, P * 4);
    if (e == hipSuccess) e = hipMemset(m->adam_v, 0, P * 4);
    if (e == hipSuccess) e = hipMemset(m->grad, 0, P * 4);
    if (e != hipSuccess) {
        lyap_mlp_destroy(m);
        return mfail(LYAP_ERR_HIP, "mlp init copy: %s", hipGetErrorString(e));
    }
    for (int l = 0; l < kLayers; ++l) {
        const int n = dims[l + 1] * dims[l];
        hipLaunchKernelGGL(k_issue_bf16, dim3((n + 255) / 256), dim3(256), 0, 0, m->master + m->off_w[l],
                           dims[l + 1], dims[l], m->wt[l], m->wn[l]);
    }
    // the null stream only: work of the caller's other streams is not waited for
    e = hipStreamSynchronize(nullptr);
    if (e != hipSuccess) {
        lyap_mlp_destroy(m);
        return mfail(LYAP_ERR_HIP, "mlp init: %s", hipGetErrorString(e));
    }
    *out = m;
    return LYAP_SUCCESS;
}

// C = A . B^T (+ bias); with Hr also H = gelu(C) (row- and column-major bf16)
static void gemm(hipStream_t s, const uint16_t* A, const uint16_t* B, float* C, const float* bias, int M,
                 int N, int K, float beta, uint16_t* Hr = nullptr, uint16_t* Ht = nullptr) {
    // 16 x 64 tiles per wave only when they still give four waves per CU (the
    // inference batches); the training batches' products are latency-bound and
    // run faster on four times as many 16 x 16 tiles (same sums per element)
    const int nb = (N % 64 == 0 && (M / 16) * (N / 64) >= 1024) ? 4 : 1;
    const int waves = (M / 16) * (N / (16 * nb));
    const dim3 grid((unsigned)((waves + 3) / 4)), block(256);
    if (nb == 4) hipLaunchKernelGGL(k_gemm_nt<4>, grid, block, 0, s, A, B, C, bias, M, N, K, beta, Hr, Ht);
    else hipLaunchKernelGGL(k_gemm_nt<1>, grid, block, 0, s, A, B, C, bias, M, N, K, beta, Hr, Ht);
}

// weight-gradient product dW^T[M][N] (small M x N, K = the batch): K split
// into up to kMaxSplits chunks of a multiple of 32, one wave each, inside one
// workgroup per output tile; with bias_part the same launch adds the layer's
// bias-gradient chunk sums ([chunks][M]) into db
static void wgrad_shape(int M, int N, int K, int* splits_out, int* kc_out, int* nb_out) {
    int splits = K / 256;
    if (splits > kMaxSplits) splits = kMaxSplits;
    if (splits < 1) splits = 1;
    const int kc = ((K + splits - 1) / splits + 31) / 32 * 32;  // whole double-K steps
    *splits_out = (K + kc - 1) / kc;
    *kc_out = kc;
    // 16 x 16 tiles unless 16 x 64 ones still give a block per CU: these
    // products are latency-bound, so more blocks beat more reuse per wave
    *nb_out = (N % 64 == 0 && (M / 16) * (N / 64) >= 256) ? 4 : 1;
}

static void gemm_wgrad(hipStream_t s, const uint16_t* A, const uint16_t* B, float* C, int M, int N, int K,
                       const float* bias_part, int chunks, float* db) {
    int splits, kc, nb;
    wgrad_shape(M, N, K, &splits, &kc, &nb);
    const int tiles = (M / 16) * (N / (16 * nb));
    const size_t lds = ((size_t)splits * 16 * 16 * nb + (size_t)kColsumChunks * 16) * sizeof(float);
    const dim3 grid((unsigned)tiles), block((unsigned)(64 * splits));
    if (nb == 4)
        hipLaunchKernelGGL(k_wgrad<4>, grid, block, lds, s, A, B, C, M, N, K, kc, splits, bias_part, chunks, db);
    else
        hipLaunchKernelGGL(k_wgrad<1>, grid, block, lds, s, A, B, C, M, N, K, kc, splits, bias_part, chunks, db);
}

// row chunks of the bias-gradient sums db[j] = sum_b dZ[b][j]: up to
// kColsumChunks chunks summed in parallel, then added in chunk order
static void colsum_chunks(int n, int* chunks, int* rc) {
    int c = n / 64;  // 64-row chunks: 16 rows per thread
    if (c > kColsumChunks) c = kColsumChunks;
    if (c < 1) c = 1;
    *rc = ((n + c - 1) / c + 3) / 4 * 4;
    *chunks = (n + *rc - 1) / *rc;
}

// clear_loss: the loss accumulator is cleared by the first kernel (training)
// head = false: the hidden layers only (training runs the head with its loss)
static int forward(lyap_mlp* m, const double* d_p, int n_pad, int n_valid, hipStream_t s, bool clear_loss,
                   bool head = true) {
    const auto& c = m->cfg;
    const int* D = m->dims;
    hipLaunchKernelGGL(k_fourier, dim3((unsigned)(((int64_t)n_pad * c.n_fourier + 255) / 256)), dim3(256), 0, s,
                       d_p, n_pad, n_valid, m->fourier,
                       c.n_fourier, c.lo1, 1.0 / (c.hi1 - c.lo1), c.lo2, 1.0 / (c.hi2 - c.lo2), m->act[0],
                       m->act_t[0], clear_loss ? m->loss : nullptr);
    for (int l = 0; l < (head ? kLayers : kLayers - 1); ++l) {
        const bool hidden = l + 1 < kLayers;
        gemm(s, m->act[l], m->wt[l], m->z[l], m->master + m->off_b[l], n_pad, D[l + 1], D[l], 0.f,
             hidden ? m->act[l + 1] : nullptr, hidden ? m->act_t[l + 1] : nullptr);
    }
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return mfail(LYAP_ERR_HIP, "mlp forward: %s", hipGetErrorString(e));
    return LYAP_SUCCESS;
}

// the head runs in one launch with its loss when the bias chunks are 64 rows
static bool head_fused(int n_pad) {
    int chunks, rc;
    colsum_chunks(n_pad, &chunks, &rc);
    return rc == 64 && n_pad == chunks * 64;
}

static int backward(lyap_mlp* m, const double* d_lam, const uint32_t* d_status, int n_pad, int n_valid,
                    hipStream_t s) {
    const int* D = m->dims;
    int chunks, rc;
    colsum_chunks(n_pad, &chunks, &rc);
    const bool fused_head = head_fused(n_pad);
    if (fused_head) {
        const int l = kLayers - 1;
        hipLaunchKernelGGL(k_gemm_head_loss, dim3((unsigned)chunks), dim3(256), 0, s, m->act[l], m->wt[l],
                           m->master + m->off_b[l], n_pad, D[l], m->z[l], d_lam, d_status, n_valid, m->norm,
                           m->norm + 3, 1.0f / (float)n_valid, m->dz, m->dz_t, m->dzf, m->loss, m->split_ws);
    } else {
        hipLaunchKernelGGL(k_head_loss, dim3((n_pad + 255) / 256), dim3(256), 0, s, m->z[kLayers - 1], d_lam,
                           d_status, n_pad, n_valid, m->norm, m->norm + 3, 1.0f / (float)n_valid, m->dz, m->dz_t,
                           m->dzf, m->loss);
    }
    // dZ of the layer being differentiated (row- and column-major) and its
    // bias-gradient chunk sums; the input-gradient pass writes the layer
    // below's into the other buffer of each pair (a paired launch still reads
    // the current ones)
    uint16_t *dz_in = m->dz, *dz_out = m->dz2;
    uint16_t *dzt_in = m->dz_t, *dzt_out = m->dz_t2;
    const float* bias_in = m->split_ws;  // the head's, from its fp32 dZ
    float* cw_out = m->col_ws;
    if (!fused_head)
        hipLaunchKernelGGL(k_colsum_part, dim3((unsigned)((D[kLayers] + 63) / 64), (unsigned)chunks), dim3(256), 0,
                           s, m->dzf, n_pad, D[kLayers], rc, m->split_ws);
    const bool dgelu_ok = rc == 64 && n_pad == chunks * 64;
    for (int l = kLayers - 1; l >= 0; --l) {
        // weight gradient dW^T[out][in] = dZ^T[out][B] . act^T[in][B]^T with the
        // bias gradient from the chunk sums; input gradient dH[B][in] = dZ[B][out]
        // . W[in][out]^T through the GELU below (l >= 1)
        float* const dW = m->grad + m->off_w[l];
        float* const db = m->grad + m->off_b[l];
        int splits, kc, nb;
        wgrad_shape(D[l + 1], D[l], n_pad, &splits, &kc, &nb);
        if (l >= 1 && dgelu_ok && D[l] % 16 == 0 && nb == 1 && splits == 8 && chunks % 2 == 0) {
            const int wb = (D[l + 1] / 16) * (D[l] / 16);
            const WgradArgs wa{dzt_in, m->act_t[l], dW, D[l + 1], D[l], n_pad, kc, splits, bias_in, chunks, db};
            const DgeluArgs da{dz_in, m->wn[l], m->z[l - 1], n_pad, D[l], D[l + 1], dz_out, dzt_out, cw_out};
            const size_t lds = std::max((size_t)(splits * 16 * 16 + kColsumChunks * 16), (size_t)(2 * (64 * 17 + 64))) *
                               sizeof(float);
            hipLaunchKernelGGL(k_wgrad_dgelu, dim3((unsigned)(wb + chunks * (D[l] / 16) / 2)), dim3(512), lds, s, wa,
                               da, wb);
        } else {
            gemm_wgrad(s, dzt_in, m->act_t[l], dW, D[l + 1], D[l], n_pad, bias_in, chunks, db);
            if (l == 0) break;
            if (dgelu_ok && D[l] % 16 == 0) {
                hipLaunchKernelGGL(k_gemm_dgelu, dim3((unsigned)(chunks * (D[l] / 16))), dim3(256), 0, s, dz_in,
                                   m->wn[l], m->z[l - 1], n_pad, D[l], D[l + 1], dz_out, dzt_out, cw_out);
            } else {
                gemm(s, dz_in, m->wn[l], m->dh, nullptr, n_pad, D[l], D[l + 1], 0.f);
                hipLaunchKernelGGL(k_gelu_bwd_colsum, dim3((unsigned)((D[l] + 63) / 64), (unsigned)chunks), dim3(256),
                                   0, s, m->dh, m->z[l - 1], n_pad, D[l], rc, dz_out, dzt_out, cw_out);
            }
        }
        std::swap(dz_in, dz_out);
        std::swap(dzt_in, dzt_out);
        bias_in = cw_out;
        cw_out = cw_out == m->col_ws ? m->col_ws2 : m->col_ws;
    }
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return mfail(LYAP_ERR_HIP, "mlp backward: %s", hipGetErrorString(e));
    return LYAP_SUCCESS;
}

static int check_batch(lyap_mlp* m, int64_t n) {
    if (!m) return mfail(LYAP_ERR_INVALID, "NULL network");
    if (n <= 0 || n > m->cfg.max_batch)
        return mfail(LYAP_ERR_INVALID, "batch %lld outside (0, %d]", (long long)n, m->cfg.max_batch);
    return LYAP_SUCCESS;
}

int lyap_mlp_backward(lyap_mlp* m, const double* d_p, const double* d_lambda, const uint32_t* d_status,
                      int64_t n, float* d_loss, void* stream) {
    int rc = check_batch(m, n);
    if (rc) return rc;
    if (!d_p || !d_lambda || !d_status) return mfail(LYAP_ERR_INVALID, "NULL batch pointer");
    hipStream_t s = (hipStream_t)stream;
    MLP_HIP(hipSetDevice(m->device));
    const int n_pad = (int)((n + 63) / 64 * 64);
    if ((rc = forward(m, d_p, n_pad, (int)n, s, true, !head_fused(n_pad)))) return rc;
    if ((rc = backward(m, d_lambda, d_status, n_pad, (int)n, s))) return rc;
    if (d_loss) MLP_HIP(hipMemcpyAsync(d_loss, m->loss, 4, hipMemcpyDeviceToDevice, s));
    return LYAP_SUCCESS;
}

int lyap_mlp_train_step(lyap_mlp* m, const double* d_p, const double* d_lambda, const uint32_t* d_status,
                        int64_t n, float* d_loss, void* stream) {
    int rc = lyap_mlp_backward(m, d_p, d_lambda, d_status, n, d_loss, stream);
    if (rc) return rc;
    hipStream_t s = (hipStream_t)stream;
    const auto& c = m->cfg;
    m->step += 1;
    const float c1 = 1.f - std::pow(c.beta1, (float)m->step), c2 = 1.f - std::pow(c.beta2, (float)m->step);
    WeightMap wm;
    for (int l = 0; l < kLayers; ++l) {
        wm.off[l] = m->off_w[l];
        wm.O[l] = m->dims[l + 1];
        wm.I[l] = m->dims[l];
        wm.wt[l] = m->wt[l];
        wm.wn[l] = m->wn[l];
    }
    hipLaunchKernelGGL(k_adam_issue, dim3((unsigned)((m->n_params + 255) / 256)), dim3(256), 0, s, m->master,
                       m->grad, m->adam_m, m->adam_v, m->trainable, m->n_params, c.lr, c.beta1, c.beta2, c.eps, c1,
                       c2, wm);
    MLP_HIP(hipGetLastError());
    return LYAP_SUCCESS;
}

int lyap_mlp_forward(lyap_mlp* m, const double* d_p, int64_t n, float* d_out, void* stream) {
    int rc = check_batch(m, n);
    if (rc) return rc;
    if (!d_p || !d_out) return mfail(LYAP_ERR_INVALID, "NULL pointer");
    hipStream_t s = (hipStream_t)stream;
    MLP_HIP(hipSetDevice(m->device));
    const int n_pad = (int)((n + 63) / 64 * 64);
    if ((rc = forward(m, d_p, n_pad, (int)n, s, false))) return rc;
    hipLaunchKernelGGL(k_head_out, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, m->z[kLayers - 1], (int)n,
                       m->norm, m->norm + 3, d_out);
    MLP_HIP(hipGetLastError());
    return LYAP_SUCCESS;
}

int lyap_mlp_param_count(lyap_mlp* m, int64_t* n) {
    if (!m || !n) return mfail(LYAP_ERR_INVALID, "NULL argument");
    *n = m->n_params;
    return LYAP_SUCCESS;
}

// which = 0: parameters, 1: gradients, 2: Fourier matrix [n_fourier][2]; host copies
int lyap_mlp_get(lyap_mlp* m, int which, float* host, int64_t n) {
    if (!m || !host) return mfail(LYAP_ERR_INVALID, "NULL argument");
    MLP_HIP(hipSetDevice(m->device));
    MLP_HIP(hipDeviceSynchronize());
    const float* src = which == 0 ? m->master : (which == 1 ? m->grad : m->fourier);
    const int64_t cnt = which == 2 ? (int64_t)m->cfg.n_fourier * 2 : m->n_params;
    if (which < 0 || which > 2 || n != cnt) return mfail(LYAP_ERR_INVALID, "lyap_mlp_get: size %lld", (long long)cnt);
    MLP_HIP(hipMemcpy(host, src, (size_t)cnt * 4, hipMemcpyDeviceToHost));
    return LYAP_SUCCESS;
}

int lyap_mlp_set_params(lyap_mlp* m, const float* host, int64_t n) {
    if (!m || !host || n != m->n_params) return mfail(LYAP_ERR_INVALID, "lyap_mlp_set_params: bad size");
    MLP_HIP(hipSetDevice(m->device));
    MLP_HIP(hipMemcpy(m->master, host, (size_t)n * 4, hipMemcpyHostToDevice));
    for (int l = 0; l < kLayers; ++l) {
        const int nw = m->dims[l + 1] * m->dims[l];
        hipLaunchKernelGGL(k_issue_bf16, dim3((nw + 255) / 256), dim3(256), 0, 0, m->master + m->off_w[l],
                           m->dims[l + 1], m->dims[l], m->wt[l], m->wn[l]);
    }
    MLP_HIP(hipStreamSynchronize(nullptr));
    return LYAP_SUCCESS;
}

}  // extern "C"
