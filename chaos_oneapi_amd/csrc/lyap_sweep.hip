// Per-lane sweep kernels for the 3-D flows (shimizu, Lorenz, Rössler).
//
// One wave64 lane owns one parameter point; the whole state (base point,
// deviation matrix, step control, log sums) lives in VGPRs.  The launch is
// persistent: when a lane's point finishes (adaptive step counts differ per
// point), the wave refills every idle lane with one wave-aggregated atomic on
// the work counter, so lanes never wait for the slowest point of their wave
// to finish before taking new work.
//
// Kernels:
//   k_var_lane   coupled variational RK (Jacobian at every stage state), MGS
//                on exact time boundaries, FP64 or FP32-stage tangent,
//                DVERK 6(5) / Dopri5, fixed or adaptive   (the sweep's method);
//                with TRACE N - 1 vectors and the last exponent from the
//                integrated Jacobian trace
//   k_varref_lane  reference VAR mode     cuda/lyapunov.dp.cpp:133-226
//   k_qrref_lane   reference QR mode      cuda/lyapunov.dp.cpp:229-259
//   k_fdref_lane   reference FD mode      cuda/lyapunov.dp.cpp:5-129
#include "lyap_device.hpp"
#include "lyap_lane.hpp"
#include "lyap_launch.hpp"

namespace lyap {

// ---------------------------------------------------------------------------
// the sweep's method: one lane per point, persistent
// ---------------------------------------------------------------------------
// TRACE: N - 1 deviation vectors, the last exponent from the integrated trace
template <class M, int INTEG, bool ADAPT, bool MIXED, bool TRACE = false>
__global__ __launch_bounds__(256) void k_var_lane(const lyap_sweep_desc d, const SweepIO io) {
    lane_drain<M, INTEG, ADAPT, MIXED, TRACE>(d, [] { return kernarg_desc(); }, io, [] { return kernarg_io(); });
}

// the same with progress signals and skips (w, read from the kernarg segment
// when a lane takes or finishes a point); adaptive steps only.  Compiled for
// three waves a SIMD (168 VGPRs), the occupancy of the plain kernel it stands
// in for: the signal and skip code would otherwise push some variants past it
template <class M, int INTEG, bool MIXED, bool TRACE>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(3))) void k_var_lane_watch(const lyap_sweep_desc d, const SweepIO io, const LaneWatch w) {
    lane_drain<M, INTEG, true, MIXED, TRACE>(d, [] { return kernarg_desc(); }, io, [] { return kernarg_io(); },
                                             [] { return kernarg_watch(); });
}

// the same method with four lanes per point (latency-bound launches)
template <class M, int INTEG, bool ADAPT, bool MIXED, bool TRACE = false>
__global__ __launch_bounds__(256) void k_var_quad(const lyap_sweep_desc d, const SweepIO io) {
    quad_drain<M, INTEG, ADAPT, MIXED, TRACE>(d, [] { return kernarg_desc(); }, io, [] { return kernarg_io(); });
}

// ---------------------------------------------------------------------------
// reference modes (fixed-step DVERK, one lane per point, grid-stride)
// ---------------------------------------------------------------------------
// dverkStepVarMat: all stages use J at the same base point
template <class M>
LYAP_D void frozen_tangent_step(const double* xf, const double* p, double h, double* v) {
    using TB = Tab<LYAP_INT_DVERK>;
    constexpr int S = TB::S, N = M::N;
    const typename M::template J<double> jf = M::template jac<double>(xf, p);
    double kv[S][N];
#pragma unroll
    for (int s = 0; s < S; ++s) {
        if (!TB::for_solution(s)) {
#pragma unroll
            for (int i = 0; i < N; ++i) kv[s][i] = 0.0;
            continue;
        }
        double arg[N];
        const RowA<TB, double, S> row(s);
#pragma unroll
        for (int i = 0; i < N; ++i) {
            bool any;
            const double acc = coef_sum_a(s, row, kv, i, any);
            arg[i] = any ? fma(h, acc, v[i]) : v[i];
        }
        M::template jvp<double>(jf, p, arg, kv[s]);
    }
    const RowB<TB, double, S> rb;
#pragma unroll
    for (int i = 0; i < N; ++i) v[i] = fma(h, coef_sum_b(rb, kv, i), v[i]);
}

template <class M>
LYAP_D void plain_step(double* x, const double* p, double h) {
    double xn[M::N];
    typename M::template J<double> X[Tab<LYAP_INT_DVERK>::S];
    base_step<M, LYAP_INT_DVERK, false, double>(x, p, h, xn, nullptr, X, nullptr);
#pragma unroll
    for (int i = 0; i < M::N; ++i) x[i] = xn[i];
}

// cuda/tools/linalg.dp.cpp:5-18
template <int N>
LYAP_D double dot_ref(const double* a, const double* b) {
    double r = 0.0;
#pragma unroll
    for (int i = 0; i < N; ++i) r = r + a[i] * b[i];
    return r;
}

// cuda/tools/linalg.dp.cpp:36-51 (classical GS) and :20-28 (normalise to eps)
template <int K, int N>
LYAP_D void cgs_ref(double (&V)[K][N]) {
#pragma unroll
    for (int i = 1; i < K; ++i) {
        double proj[N];
#pragma unroll
        for (int q = 0; q < N; ++q) proj[q] = 0.0;
#pragma unroll
        for (int j = 0; j < i; ++j) {
            const double c = dot_ref<N>(V[j], V[i]) / dot_ref<N>(V[j], V[j]);
#pragma unroll
            for (int q = 0; q < N; ++q) proj[q] = proj[q] + c * V[j][q];
        }
#pragma unroll
        for (int q = 0; q < N; ++q) V[i][q] = V[i][q] - proj[q];
    }
}

template <int K, int N>
LYAP_D void normalize_ref(double (&V)[K][N], double eps) {
#pragma unroll
    for (int j = 0; j < K; ++j) {
        const double nrm = sqrt(dot_ref<N>(V[j], V[j]));
#pragma unroll
        for (int i = 0; i < N; ++i) V[j][i] = (V[j][i] / nrm) * eps;
    }
}

template <class M>
__global__ __launch_bounds__(256) void k_varref_lane(const lyap_sweep_desc d, const SweepIO io) {
    constexpr int N = M::N, P = M::P, K = M::N;
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t item = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; item < io.count;
         item += stride) {
        const int64_t idx = io.idx_list ? io.idx_list[item] : io.first + item;
        double p[P], x[N], V[K][N], L[K];
        point_params<P>(d, idx, io.params, p);
#pragma unroll
        for (int i = 0; i < N; ++i)
            x[i] = io.inits ? io.inits[idx * N + i] : init_component(d, idx, i, p);
        for (int64_t s = 0; s < d.ref_steps_transient; ++s) plain_step<M>(x, p, d.h);
#pragma unroll
        for (int r = 0; r < K; ++r) {
            L[r] = 0.0;
#pragma unroll
            for (int i = 0; i < N; ++i) V[r][i] = 0.0;
            V[r][r] = V[r][r] + d.eps;
        }
        for (int64_t s = 0; s < d.ref_steps_slave; ++s) {
#pragma unroll
            for (int r = 0; r < K; ++r) frozen_tangent_step<M>(x, p, d.h, V[r]);
            plain_step<M>(x, p, d.h);
            cgs_ref<K, N>(V);
            normalize_ref<K, N>(V, d.eps);
        }
        int32_t skip = 0;
        for (int64_t s = 0; s < d.ref_steps_average; ++s) {
            ++skip;
#pragma unroll
            for (int r = 0; r < K; ++r) frozen_tangent_step<M>(x, p, d.h, V[r]);
            plain_step<M>(x, p, d.h);
            if ((double)skip == d.add_skip_steps) {
                cgs_ref<K, N>(V);
#pragma unroll
                for (int r = 0; r < K; ++r)
                    L[r] = L[r] + det_log(sqrt(dot_ref<N>(V[r], V[r])) / d.eps);
                normalize_ref<K, N>(V, d.eps);
                skip = 0;
            }
        }
        bool ok = true;
        for (int r = 0; r < d.n_exps; ++r) {
            double lr = 0.0;
#pragma unroll
            for (int q = 0; q < K; ++q)
                if (q == r) lr = L[q];
            const double lam = lr / d.t_average;
            ok = ok && isfinite(lam);
            io.lambda[item * d.n_exps + r] = lam;
        }
        if (io.status) io.status[item] = ok ? LYAP_PT_OK : LYAP_PT_NONFINITE;
        if (io.n_acc)
            io.n_acc[item] =
                (uint32_t)(d.ref_steps_transient + d.ref_steps_slave + d.ref_steps_average);
        if (io.n_rej) io.n_rej[item] = 0;
    }
}

// cuda/tools/linalg.dp.cpp:78-128 for a square N x N matrix whose columns are
// the vectors: CGS on the columns, normalise to 1, diag(R) = diag(Q^T A)
template <int N>
LYAP_D void qr_ref(double (&Q)[N][N], double* rd) {
    double A[N][N];
#pragma unroll
    for (int r = 0; r < N; ++r)
#pragma unroll
        for (int c = 0; c < N; ++c) A[r][c] = Q[r][c];
#pragma unroll
    for (int i = 1; i < N; ++i) {
        double proj[N];
#pragma unroll
        for (int q = 0; q < N; ++q) proj[q] = 0.0;
#pragma unroll
        for (int j = 0; j < i; ++j) {
            double dji = 0.0, djj = 0.0;
#pragma unroll
            for (int q = 0; q < N; ++q) dji = dji + Q[q][j] * Q[q][i];
#pragma unroll
            for (int q = 0; q < N; ++q) djj = djj + Q[q][j] * Q[q][j];
            const double c = dji / djj;
#pragma unroll
            for (int q = 0; q < N; ++q) proj[q] = proj[q] + c * Q[q][j];
        }
#pragma unroll
        for (int q = 0; q < N; ++q) Q[q][i] = Q[q][i] - proj[q];
    }
#pragma unroll
    for (int i = 0; i < N; ++i) {
        double s = 0.0;
#pragma unroll
        for (int q = 0; q < N; ++q) s = s + Q[q][i] * Q[q][i];
        const double nrm = sqrt(s);
#pragma unroll
        for (int q = 0; q < N; ++q) Q[q][i] = (Q[q][i] / nrm) * 1.0;
    }
#pragma unroll
    for (int i = 0; i < N; ++i) {
        double r = 0.0;
#pragma unroll
        for (int q = 0; q < N; ++q) r = r + Q[q][i] * A[q][i];
        rd[i] = r;
    }
}

template <class M>
__global__ __launch_bounds__(256) void k_qrref_lane(const lyap_sweep_desc d, const SweepIO io) {
    constexpr int N = M::N, P = M::P;
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t item = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; item < io.count;
         item += stride) {
        const int64_t idx = io.idx_list ? io.idx_list[item] : io.first + item;
        double p[P], x[N], Q[N][N], L[N], rd[N];
        point_params<P>(d, idx, io.params, p);
#pragma unroll
        for (int i = 0; i < N; ++i)
            x[i] = io.inits ? io.inits[idx * N + i] : init_component(d, idx, i, p);
        for (int64_t s = 0; s < d.ref_steps_transient; ++s) plain_step<M>(x, p, d.h);
#pragma unroll
        for (int r = 0; r < N; ++r) {
            L[r] = 0.0;
#pragma unroll
            for (int c = 0; c < N; ++c) Q[r][c] = (r == c) ? 1.0 : 0.0;
        }
        for (int64_t s = 0; s < d.ref_steps_average; ++s) {
            plain_step<M>(x, p, d.h);
#pragma unroll
            for (int c = 0; c < N; ++c) {
                double col[N];
#pragma unroll
                for (int q = 0; q < N; ++q) col[q] = Q[q][c];
                frozen_tangent_step<M>(x, p, d.h, col);
#pragma unroll
                for (int q = 0; q < N; ++q) Q[q][c] = col[q];
            }
            qr_ref<N>(Q, rd);
#pragma unroll
            for (int i = 0; i < N; ++i)
                if (i < d.n_exps) L[i] = L[i] + det_log(fabs(rd[i]));
        }
        bool ok = true;
        for (int r = 0; r < d.n_exps; ++r) {
            double lr = 0.0;
#pragma unroll
            for (int q = 0; q < N; ++q)
                if (q == r) lr = L[q];
            const double lam = lr / d.t_average;
            ok = ok && isfinite(lam);
            io.lambda[item * d.n_exps + r] = lam;
        }
        if (io.status) io.status[item] = ok ? LYAP_PT_OK : LYAP_PT_NONFINITE;
        if (io.n_acc) io.n_acc[item] = (uint32_t)(d.ref_steps_transient + d.ref_steps_average);
        if (io.n_rej) io.n_rej[item] = 0;
    }
}

template <class M>
__global__ __launch_bounds__(256) void k_fdref_lane(const lyap_sweep_desc d, const SweepIO io) {
    constexpr int N = M::N, P = M::P, K = M::N;
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t item = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; item < io.count;
         item += stride) {
        const int64_t idx = io.idx_list ? io.idx_list[item] : io.first + item;
        double p[P], x[N], S[K][N], L[K];
        point_params<P>(d, idx, io.params, p);
#pragma unroll
        for (int i = 0; i < N; ++i)
            x[i] = io.inits ? io.inits[idx * N + i] : init_component(d, idx, i, p);
        for (int64_t s = 0; s < d.ref_steps_transient; ++s) plain_step<M>(x, p, d.h);
#pragma unroll
        for (int r = 0; r < K; ++r) {
            L[r] = 0.0;
#pragma unroll
            for (int i = 0; i < N; ++i) S[r][i] = x[i];
            S[r][r] = S[r][r] + d.eps;
        }
        for (int64_t s = 0; s < d.ref_steps_slave; ++s) {
            plain_step<M>(x, p, d.h);
#pragma unroll
            for (int r = 0; r < K; ++r) {
                plain_step<M>(S[r], p, d.h);
#pragma unroll
                for (int i = 0; i < N; ++i) S[r][i] = S[r][i] - x[i];
            }
            cgs_ref<K, N>(S);
            normalize_ref<K, N>(S, d.eps);
#pragma unroll
            for (int r = 0; r < K; ++r)
#pragma unroll
                for (int i = 0; i < N; ++i) S[r][i] = S[r][i] + x[i];
        }
        int32_t skip = 0;
        for (int64_t s = 0; s < d.ref_steps_average; ++s) {
            ++skip;
            plain_step<M>(x, p, d.h);
#pragma unroll
            for (int r = 0; r < K; ++r) plain_step<M>(S[r], p, d.h);
            if ((double)skip == d.add_skip_steps) {
#pragma unroll
                for (int r = 0; r < K; ++r)
#pragma unroll
                    for (int i = 0; i < N; ++i) S[r][i] = S[r][i] - x[i];
                cgs_ref<K, N>(S);
#pragma unroll
                for (int r = 0; r < K; ++r)
                    L[r] = L[r] + det_log(sqrt(dot_ref<N>(S[r], S[r])) / d.eps);
                normalize_ref<K, N>(S, d.eps);
#pragma unroll
                for (int r = 0; r < K; ++r)
#pragma unroll
                    for (int i = 0; i < N; ++i) S[r][i] = S[r][i] + x[i];
                skip = 0;
            }
        }
        bool ok = true;
        for (int r = 0; r < d.n_exps; ++r) {
            double lr = 0.0;
#pragma unroll
            for (int q = 0; q < K; ++q)
                if (q == r) lr = L[q];
            const double lam = lr / d.t_average;
            ok = ok && isfinite(lam);
            io.lambda[item * d.n_exps + r] = lam;
        }
        if (io.status) io.status[item] = ok ? LYAP_PT_OK : LYAP_PT_NONFINITE;
        if (io.n_acc)
            io.n_acc[item] =
                (uint32_t)(d.ref_steps_transient + d.ref_steps_slave + d.ref_steps_average);
        if (io.n_rej) io.n_rej[item] = 0;
    }
}

// ---------------------------------------------------------------------------
// maps
// ---------------------------------------------------------------------------
template <class M, bool MIXED>
__global__ __launch_bounds__(256) void k_map_lane(const lyap_sweep_desc d, const SweepIO io) {
    map_drain<M, MIXED>(d, [] { return kernarg_desc(); }, io);
}

template <class M>
LYAP_D void map_step(double* x, const double* p) {
    double xn[M::N];
    M::rhs(x, p, xn);
#pragma unroll
    for (int i = 0; i < M::N; ++i) x[i] = xn[i];
}

// StepMAPVAR for the k row vectors, J at the current base point
template <class M, int K>
LYAP_D void map_tangent(const double* x, const double* p, double (&V)[K][M::N]) {
    const typename M::template J<double> jx = M::template jac<double>(x, p);
#pragma unroll
    for (int r = 0; r < K; ++r) {
        double o[M::N];
        M::template jvp<double>(jx, p, V[r], o);
#pragma unroll
        for (int i = 0; i < M::N; ++i) V[r][i] = o[i];
    }
}

// reference VAR mode, map branch: cuda/lyapunov.dp.cpp:182-222,224-225
template <class M>
__global__ __launch_bounds__(256) void k_mapvarref_lane(const lyap_sweep_desc d, const SweepIO io) {
    constexpr int N = M::N, P = M::P, K = M::N;
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t item = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; item < io.count;
         item += stride) {
        const int64_t idx = io.idx_list ? io.idx_list[item] : io.first + item;
        double p[P], x[N], V[K][N], L[K];
        point_params<P>(d, idx, io.params, p);
#pragma unroll
        for (int i = 0; i < N; ++i)
            x[i] = io.inits ? io.inits[idx * N + i] : init_component(d, idx, i, p);
        for (int64_t s = 0; s < d.ref_steps_transient; ++s) map_step<M>(x, p);
#pragma unroll
        for (int r = 0; r < K; ++r) {
            L[r] = 0.0;
#pragma unroll
            for (int i = 0; i < N; ++i) V[r][i] = 0.0;
            V[r][r] = V[r][r] + d.eps;
        }
        for (int64_t s = 0; s < d.ref_steps_slave; ++s) {
            map_tangent<M, K>(x, p, V);
            map_step<M>(x, p);
            cgs_ref<K, N>(V);
            normalize_ref<K, N>(V, d.eps);
        }
        for (int64_t s = 0; s < d.ref_steps_average; ++s) {
            map_tangent<M, K>(x, p, V);
            map_step<M>(x, p);
            cgs_ref<K, N>(V);
#pragma unroll
            for (int r = 0; r < K; ++r) L[r] = L[r] + det_log(sqrt(dot_ref<N>(V[r], V[r])) / d.eps);
            normalize_ref<K, N>(V, d.eps);
        }
        bool ok = true;
        for (int r = 0; r < d.n_exps; ++r) {
            double lr = 0.0;
#pragma unroll
            for (int q = 0; q < K; ++q)
                if (q == r) lr = L[q];
            const double lam = lr / d.t_average;
            ok = ok && isfinite(lam);
            io.lambda[item * d.n_exps + r] = lam;
        }
        if (io.status) io.status[item] = ok ? LYAP_PT_OK : LYAP_PT_NONFINITE;
        if (io.n_acc)
            io.n_acc[item] =
                (uint32_t)(d.ref_steps_transient + d.ref_steps_slave + d.ref_steps_average);
        if (io.n_rej) io.n_rej[item] = 0;
    }
}

// reference FD (two-trajectory) mode, map branch: cuda/lyapunov.dp.cpp:70-122,127-128
template <class M>
__global__ __launch_bounds__(256) void k_mapfdref_lane(const lyap_sweep_desc d, const SweepIO io) {
    constexpr int N = M::N, P = M::P, K = M::N;
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t item = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; item < io.count;
         item += stride) {
        const int64_t idx = io.idx_list ? io.idx_list[item] : io.first + item;
        double p[P], x[N], S[K][N], L[K];
        point_params<P>(d, idx, io.params, p);
#pragma unroll
        for (int i = 0; i < N; ++i)
            x[i] = io.inits ? io.inits[idx * N + i] : init_component(d, idx, i, p);
        for (int64_t s = 0; s < d.ref_steps_transient; ++s) map_step<M>(x, p);
#pragma unroll
        for (int r = 0; r < K; ++r) {
            L[r] = 0.0;
#pragma unroll
            for (int i = 0; i < N; ++i) S[r][i] = x[i];
            S[r][r] = S[r][r] + d.eps;
        }
        for (int64_t s = 0; s < d.ref_steps_slave; ++s) {
            map_step<M>(x, p);
#pragma unroll
            for (int r = 0; r < K; ++r) {
                map_step<M>(S[r], p);
#pragma unroll
                for (int i = 0; i < N; ++i) S[r][i] = S[r][i] - x[i];
            }
            cgs_ref<K, N>(S);
            normalize_ref<K, N>(S, d.eps);
#pragma unroll
            for (int r = 0; r < K; ++r)
#pragma unroll
                for (int i = 0; i < N; ++i) S[r][i] = S[r][i] + x[i];
        }
        for (int64_t s = 0; s < d.ref_steps_average; ++s) {
            map_step<M>(x, p);
#pragma unroll
            for (int r = 0; r < K; ++r) map_step<M>(S[r], p);
#pragma unroll
            for (int r = 0; r < K; ++r)
#pragma unroll
                for (int i = 0; i < N; ++i) S[r][i] = S[r][i] - x[i];
            cgs_ref<K, N>(S);
#pragma unroll
            for (int r = 0; r < K; ++r) L[r] = L[r] + det_log(sqrt(dot_ref<N>(S[r], S[r])) / d.eps);
            normalize_ref<K, N>(S, d.eps);
#pragma unroll
            for (int r = 0; r < K; ++r)
#pragma unroll
                for (int i = 0; i < N; ++i) S[r][i] = S[r][i] + x[i];
        }
        bool ok = true;
        for (int r = 0; r < d.n_exps; ++r) {
            double lr = 0.0;
#pragma unroll
            for (int q = 0; q < K; ++q)
                if (q == r) lr = L[q];
            const double lam = lr / d.t_average;
            ok = ok && isfinite(lam);
            io.lambda[item * d.n_exps + r] = lam;
        }
        if (io.status) io.status[item] = ok ? LYAP_PT_OK : LYAP_PT_NONFINITE;
        if (io.n_acc)
            io.n_acc[item] =
                (uint32_t)(d.ref_steps_transient + d.ref_steps_slave + d.ref_steps_average);
        if (io.n_rej) io.n_rej[item] = 0;
    }
}

__global__ void k_reset_counter(unsigned long long* c) {
    if (threadIdx.x == 0 && blockIdx.x == 0) *c = 0ULL;
}

// the two milestone counters of a watched launch
__global__ void k_reset_watch(unsigned long long* done) {
    if (blockIdx.x == 0 && threadIdx.x < 2) done[threadIdx.x] = 0ULL;
}

// ---------------------------------------------------------------------------
// launch table
// ---------------------------------------------------------------------------
// lanes_per_point: 1 (k_var_lane and the reference kernels) or 4 (k_var_quad);
// max_blocks_hint > 0 caps the grid, < 0 leaves that many blocks of the
// occupancy-sized grid free for kernels on other streams (at least one block runs)
template <class K, class... Extra>
static hipError_t launch_persistent(K kern, const lyap_sweep_desc& d, const SweepIO& io,
                                    hipStream_t stream, int num_cus, int max_blocks_hint,
                                    int lanes_per_point, const Extra&... extra) {
    int per_cu = 0;
    hipError_t e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kern, 256, 0);
    if (e != hipSuccess) return e;
    if (per_cu < 1) per_cu = 1;
    int64_t blocks = (int64_t)per_cu * num_cus;
    if (max_blocks_hint < 0) blocks += max_blocks_hint;
    const int64_t need = (io.count * lanes_per_point + 255) / 256;
    if (blocks > need) blocks = need;
    if (max_blocks_hint > 0 && blocks > max_blocks_hint) blocks = max_blocks_hint;
    if (blocks < 1) blocks = 1;
    hipLaunchKernelGGL(k_reset_counter, dim3(1), dim3(64), 0, stream, io.counter);
    hipLaunchKernelGGL(kern, dim3((unsigned)blocks), dim3(256), 0, stream, d, io, extra...);
    return hipGetLastError();
}

// watched launches: the sweep's method, adaptive, one lane per point
template <class M>
static hipError_t dispatch_watch(const lyap_sweep_desc& d, const SweepIO& io, hipStream_t s,
                                 int num_cus, int max_blocks, const LaneWatch& w) {
    if (!d.adaptive) return hipErrorInvalidValue;
    const bool trace = d.method == LYAP_METHOD_VAR_TRACE;
    const int key = (d.integrator << 2) | (trace ? 2 : 0) | (d.precision ? 1 : 0);
    switch (key) {
#define LYAP_CASE(I, T, X)                             \
    case ((I) << 2) | ((T) ? 2 : 0) | ((X) ? 1 : 0): \
        return launch_persistent(k_var_lane_watch<M, I, X, T>, d, io, s, num_cus, max_blocks, 1, w);
        LYAP_CASE(LYAP_INT_DVERK, false, false)
        LYAP_CASE(LYAP_INT_DVERK, false, true)
        LYAP_CASE(LYAP_INT_DVERK, true, false)
        LYAP_CASE(LYAP_INT_DVERK, true, true)
        LYAP_CASE(LYAP_INT_DOPRI5, false, false)
        LYAP_CASE(LYAP_INT_DOPRI5, false, true)
        LYAP_CASE(LYAP_INT_DOPRI5, true, false)
        LYAP_CASE(LYAP_INT_DOPRI5, true, true)
#undef LYAP_CASE
        default:
            return hipErrorInvalidValue;
    }
}

template <class M>
static hipError_t dispatch_model(const lyap_sweep_desc& d, const SweepIO& io, hipStream_t s,
                                 int num_cus, int max_blocks, bool quad, const LaneWatch* w) {
    if (w) return dispatch_watch<M>(d, io, s, num_cus, max_blocks, *w);
    if (d.method == LYAP_METHOD_VAR) {
        const int key = (d.integrator << 2) | (d.adaptive ? 2 : 0) | (d.precision ? 1 : 0);
        switch (key) {
#define LYAP_CASE(I, A, X)                                                                  \
    case ((I) << 2) | ((A) ? 2 : 0) | ((X) ? 1 : 0):                                        \
        return quad ? launch_persistent(k_var_quad<M, I, A, X>, d, io, s, num_cus, max_blocks, 4) \
                    : launch_persistent(k_var_lane<M, I, A, X>, d, io, s, num_cus, max_blocks, 1);
            LYAP_CASE(LYAP_INT_DVERK, false, false)
            LYAP_CASE(LYAP_INT_DVERK, false, true)
            LYAP_CASE(LYAP_INT_DVERK, true, false)
            LYAP_CASE(LYAP_INT_DVERK, true, true)
            LYAP_CASE(LYAP_INT_DOPRI5, false, false)
            LYAP_CASE(LYAP_INT_DOPRI5, false, true)
            LYAP_CASE(LYAP_INT_DOPRI5, true, false)
            LYAP_CASE(LYAP_INT_DOPRI5, true, true)
#undef LYAP_CASE
            default:
                return hipErrorInvalidValue;
        }
    }
    if (d.method == LYAP_METHOD_VAR_TRACE) {
        const int key = (d.integrator << 2) | (d.adaptive ? 2 : 0) | (d.precision ? 1 : 0);
        switch (key) {
#define LYAP_CASE(I, A, X)                                                                          \
    case ((I) << 2) | ((A) ? 2 : 0) | ((X) ? 1 : 0):                                                \
        return quad ? launch_persistent(k_var_quad<M, I, A, X, true>, d, io, s, num_cus, max_blocks, 4) \
                    : launch_persistent(k_var_lane<M, I, A, X, true>, d, io, s, num_cus, max_blocks, 1);
            LYAP_CASE(LYAP_INT_DVERK, false, false)
            LYAP_CASE(LYAP_INT_DVERK, false, true)
            LYAP_CASE(LYAP_INT_DVERK, true, false)
            LYAP_CASE(LYAP_INT_DVERK, true, true)
            LYAP_CASE(LYAP_INT_DOPRI5, false, false)
            LYAP_CASE(LYAP_INT_DOPRI5, false, true)
            LYAP_CASE(LYAP_INT_DOPRI5, true, false)
            LYAP_CASE(LYAP_INT_DOPRI5, true, true)
#undef LYAP_CASE
            default:
                return hipErrorInvalidValue;
        }
    }
    const int64_t grid = (io.count + 255) / 256 > 4096 ? 4096 : (io.count + 255) / 256;
    const dim3 g((unsigned)(grid < 1 ? 1 : grid)), b(256);
    switch (d.method) {
        case LYAP_METHOD_VAR_REF:
            hipLaunchKernelGGL(k_varref_lane<M>, g, b, 0, s, d, io);
            break;
        case LYAP_METHOD_QR_REF:
            hipLaunchKernelGGL(k_qrref_lane<M>, g, b, 0, s, d, io);
            break;
        case LYAP_METHOD_FD_REF:
            hipLaunchKernelGGL(k_fdref_lane<M>, g, b, 0, s, d, io);
            break;
        default:
            return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

template <class M>
static hipError_t dispatch_map(const lyap_sweep_desc& d, const SweepIO& io, hipStream_t s, int num_cus,
                               int max_blocks) {
    if (d.method == LYAP_METHOD_VAR) {
        if (d.adaptive) return hipErrorInvalidValue;
        if (d.precision == LYAP_PREC_MIXED)
            return launch_persistent(k_map_lane<M, true>, d, io, s, num_cus, max_blocks, 1);
        return launch_persistent(k_map_lane<M, false>, d, io, s, num_cus, max_blocks, 1);
    }
    const int64_t grid = (io.count + 255) / 256 > 4096 ? 4096 : (io.count + 255) / 256;
    const dim3 g((unsigned)(grid < 1 ? 1 : grid)), b(256);
    switch (d.method) {
        case LYAP_METHOD_VAR_REF:
            hipLaunchKernelGGL(k_mapvarref_lane<M>, g, b, 0, s, d, io);
            break;
        case LYAP_METHOD_FD_REF:
            hipLaunchKernelGGL(k_mapfdref_lane<M>, g, b, 0, s, d, io);
            break;
        default:
            return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

hipError_t launch_lane_sweep(const lyap_sweep_desc& d, const SweepIO& io, hipStream_t s,
                             int num_cus, int max_blocks, bool quad, const LaneWatch* watch) {
    if (watch && watch->done) hipLaunchKernelGGL(k_reset_watch, dim3(1), dim3(64), 0, s, watch->done);
    switch (d.model) {
        case LYAP_MODEL_HENON:
            return dispatch_map<Henon>(d, io, s, num_cus, max_blocks);
        case LYAP_MODEL_LOGISTIC:
            return dispatch_map<Logistic>(d, io, s, num_cus, max_blocks);
        case LYAP_MODEL_SHIMIZU:
            return dispatch_model<Shimizu>(d, io, s, num_cus, max_blocks, quad, watch);
        case LYAP_MODEL_LORENZ:
            return dispatch_model<Lorenz>(d, io, s, num_cus, max_blocks, quad, watch);
        case LYAP_MODEL_ROSSLER:
            return dispatch_model<Rossler>(d, io, s, num_cus, max_blocks, quad, watch);
        default:
            return hipErrorInvalidValue;
    }
}

}  // namespace lyap
