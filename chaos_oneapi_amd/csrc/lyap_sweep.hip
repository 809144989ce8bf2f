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
//                DVERK 6(5) / Dopri5, fixed or adaptive   (the sweep's method)
//   k_varref_lane  reference VAR mode     cuda/lyapunov.dp.cpp:133-226
//   k_qrref_lane   reference QR mode      cuda/lyapunov.dp.cpp:229-259
//   k_fdref_lane   reference FD mode      cuda/lyapunov.dp.cpp:5-129
#include "lyap_device.hpp"
#include "lyap_launch.hpp"

namespace lyap {

// ---------------------------------------------------------------------------
// base RK step: x -> xn; stage states X[s] (for the coupled tangent); error
// vector when ERR.  With FSAL (Dopri5 + ERR) k1 comes in from the previous
// accepted step and the new k1 (= f(xn)) goes out.
// ---------------------------------------------------------------------------
template <class M, int INTEG, bool ERR, class TX>
LYAP_D void base_step(const double* x, const double* p, double h, double* xn, double* ev,
                      typename M::template J<TX> (&X)[Tab<INTEG>::S], double* k1_io) {
    using TB = Tab<INTEG>;
    constexpr int S = TB::S, N = M::N;
    constexpr bool FSAL = (INTEG == LYAP_INT_DOPRI5) && ERR;
    double k[S][N];
#pragma unroll
    for (int s = 0; s < S; ++s) {
        if (!ERR && !TB::for_solution(s)) {
#pragma unroll
            for (int i = 0; i < N; ++i) k[s][i] = 0.0;
            continue;
        }
        double arg[N];
        const RowA<TB, double, S> row(s);
#pragma unroll
        for (int i = 0; i < N; ++i) {
            bool any;
            const double acc = coef_sum_a(s, row, k, i, any);
            arg[i] = any ? fma(h, acc, x[i]) : x[i];
        }
        if (ERR || TB::for_solution(s)) X[s] = M::template jac<TX>(arg, p);
        if (s == 0 && FSAL) {
#pragma unroll
            for (int i = 0; i < N; ++i) k[0][i] = k1_io[i];
        } else {
            M::rhs(arg, p, k[s]);
        }
    }
    {
        const RowB<TB, double, S> rb;
#pragma unroll
        for (int i = 0; i < N; ++i) xn[i] = fma(h, coef_sum_b(rb, k, i), x[i]);
    }
    if (ERR) {
        const RowE<TB, double, S> re;
#pragma unroll
        for (int i = 0; i < N; ++i) ev[i] = h * coef_sum_e(re, k, i);
    }
    if (FSAL) {
#pragma unroll
        for (int i = 0; i < N; ++i) k1_io[i] = k[S - 1][i];
    }
}

// coupled tangent step for one deviation vector v (FP64 storage).
// FP64 form: stages in double.  Mixed form: stages in binary32, increment
// h*sum(b_j K_j) formed in binary32 and added to the FP64 vector.
template <class M, int INTEG, bool MIXED, class TX>
LYAP_D void tangent_step(const typename M::template J<TX> (&X)[Tab<INTEG>::S], const double* p,
                         const float* pf, double h, double* v) {
    using TB = Tab<INTEG>;
    constexpr int S = TB::S, N = M::N;
    using T = typename std::conditional<MIXED, float, double>::type;
    const T hT = (T)h;
    T vt[N];
#pragma unroll
    for (int i = 0; i < N; ++i) vt[i] = (T)v[i];
    T kv[S][N];
#pragma unroll
    for (int s = 0; s < S; ++s) {
        if (!TB::for_solution(s)) {
#pragma unroll
            for (int i = 0; i < N; ++i) kv[s][i] = T(0);
            continue;
        }
        T arg[N];
        const RowA<TB, T, S> row(s);
#pragma unroll
        for (int i = 0; i < N; ++i) {
            bool any;
            const T acc = coef_sum_a(s, row, kv, i, any);
            arg[i] = any ? fma(hT, acc, vt[i]) : vt[i];
        }
        if constexpr (MIXED) M::template jvp<float>(X[s], pf, arg, kv[s]);
        else M::template jvp<double>(X[s], p, arg, kv[s]);
    }
    const RowB<TB, T, S> rb;
#pragma unroll
    for (int i = 0; i < N; ++i) {
        if constexpr (MIXED) {
            const float inc = hT * coef_sum_b(rb, kv, i);
            v[i] = v[i] + (double)inc;
        } else {
            v[i] = fma(h, coef_sum_b(rb, kv, i), v[i]);
        }
    }
}

// modified Gram-Schmidt on K rows of length N (sequential fma dot products)
template <int K, int N>
LYAP_D void mgs_rows(double (&V)[K][N], double* ln) {
#pragma unroll
    for (int j = 0; j < K; ++j) {
#pragma unroll
        for (int r = 0; r < j; ++r) {
            double c = V[r][0] * V[j][0];
#pragma unroll
            for (int i = 1; i < N; ++i) c = fma(V[r][i], V[j][i], c);
#pragma unroll
            for (int i = 0; i < N; ++i) V[j][i] = fma(-c, V[r][i], V[j][i]);
        }
        double s = V[j][0] * V[j][0];
#pragma unroll
        for (int i = 1; i < N; ++i) s = fma(V[j][i], V[j][i], s);
        const double nrm = sqrt(s);
        ln[j] = det_log(nrm);
        const double inv = 1.0 / nrm;
#pragma unroll
        for (int i = 0; i < N; ++i) V[j][i] = V[j][i] * inv;
    }
}

template <int N>
LYAP_D double err_norm_seq(const double* ev, const double* x, const double* xn, double rtol,
                           double atol) {
    double q[N];
#pragma unroll
    for (int i = 0; i < N; ++i) {
        const double ax = fabs(x[i]), an = fabs(xn[i]);
        const double sk = fma(rtol, ax > an ? ax : an, atol);
        q[i] = ev[i] / sk;
    }
    double s = q[0] * q[0];
#pragma unroll
    for (int i = 1; i < N; ++i) s = fma(q[i], q[i], s);
    return sqrt(s / (double)N);
}

template <int N>
LYAP_D bool escaped(const double* x, double radius) {
    bool e = false;
#pragma unroll
    for (int i = 0; i < N; ++i) e = e || !(fabs(x[i]) <= radius);
    return e;
}

// wave-aggregated refill: lanes with want=true get consecutive work items
LYAP_D int64_t grab_items(bool want, unsigned long long* counter) {
    const uint64_t mask = __ballot(want);
    if (mask == 0) return -1;
    const int lane = (int)__lane_id();
    const int leader = __ffsll((unsigned long long)mask) - 1;
    unsigned long long base = 0;
    if (lane == leader) base = atomicAdd(counter, (unsigned long long)__popcll(mask));
    base = __shfl(base, leader);
    const uint64_t below = mask & ((1ULL << lane) - 1ULL);
    return want ? (int64_t)(base + (unsigned long long)__popcll(below)) : -1;
}

// ---------------------------------------------------------------------------
// the sweep's method: one lane per point, persistent
// ---------------------------------------------------------------------------
template <class M, int INTEG, bool ADAPT, bool MIXED>
__global__ __launch_bounds__(256) void k_var_lane(const lyap_sweep_desc d, const SweepIO io) {
    constexpr int N = M::N, P = M::P, K = M::N, S = Tab<INTEG>::S;
    using TX = typename std::conditional<MIXED, float, double>::type;
    // scalars the step loop reads; everything else of the descriptor is
    // re-read through an opaque pointer when a lane takes a new point, so the
    // compiler does not keep (and spill) the whole descriptor in SGPRs
    const int64_t n_tr = d.n_renorm_transient;
    const int64_t m_total = d.n_renorm_transient + d.n_renorm_average;
    const int64_t max_steps = d.max_steps, spr = d.steps_per_renorm;
    const double rtol = d.rtol, atol = d.atol, h_min = d.h_min, h_max = d.h_max;
    const double t_renorm = d.t_renorm, esc = d.escape_radius, h_fixed = d.h;
    const double t_avg = ADAPT ? (double)d.n_renorm_average * d.t_renorm
                               : (double)d.n_renorm_average * (double)d.steps_per_renorm * d.h;
    const int n_exps = d.n_exps;
    const double log_floor = det_log(1e-4);

    double x[N], V[K][N], sum[K], comp[K], p[P], k1[N];
    float pf[P];
    double t = 0.0, h = 0.0;
    int64_t m = 0, step_in = 0, item = -1;
    uint32_t nacc = 0, nrej = 0;
    Ctrl ctrl{1e-4, log_floor, false};
    bool active = false, exhausted = false;

    for (;;) {
        const int64_t got = grab_items(!active && !exhausted, io.counter);
        if (!active && !exhausted) {
            if (got < io.count) {
                KernargDesc* dp = kernarg_desc();
                item = item_position(io, got);
                const int64_t idx = io.idx_list ? io.idx_list[item] : io.first + item;
                point_params<P>(*dp, idx, io.params, p);
#pragma unroll
                for (int j = 0; j < P; ++j) pf[j] = (float)p[j];
#pragma unroll
                for (int i = 0; i < N; ++i)
                    x[i] = io.inits ? io.inits[idx * N + i] : init_component(*dp, idx, i, p);
#pragma unroll
                for (int r = 0; r < K; ++r) {
                    sum[r] = 0.0;
                    comp[r] = 0.0;
#pragma unroll
                    for (int i = 0; i < N; ++i) V[r][i] = (r == i) ? 1.0 : 0.0;
                }
                t = 0.0;
                h = dp->h;
                m = 0;
                step_in = 0;
                nacc = nrej = 0;
                ctrl.fac_old = 1e-4;
                ctrl.log_fac_old = log_floor;
                ctrl.rejected = false;
                if (INTEG == LYAP_INT_DOPRI5 && ADAPT) M::rhs(x, p, k1);
                active = true;
            } else {
                exhausted = true;
            }
        }
        if (!__any(active)) break;
        if (!active) continue;

        uint32_t st = LYAP_PT_OK;
        bool finished = false;
        double xn[N], ln[K];
        typename M::template J<TX> X[S];
        if constexpr (ADAPT) {
            if ((int64_t)nacc + (int64_t)nrej >= max_steps) {
                st = LYAP_PT_MAXSTEPS;
                finished = true;
            } else {
                const double t_next = (double)(m + 1) * t_renorm;
                double h_try = h;
                bool clipped = false;
                const double jn = M::jnorm(x, p);
                if (h_try * jn > 1.0) h_try = 1.0 / jn;
                // land on the boundary when it is within 1/16 of the step, so no
                // round-off sliver is left before it
                if (t_next - t <= 1.0625 * h_try) {
                    h_try = t_next - t;
                    clipped = true;
                }
                if (h_try < h_min) {
                    st = LYAP_PT_STEP_UNDERFLOW;
                    finished = true;
                } else {
                    double ev[N], k1_next[N];
#pragma unroll
                    for (int i = 0; i < N; ++i) k1_next[i] = k1[i];
                    base_step<M, INTEG, true, TX>(x, p, h_try, xn, ev, X, k1_next);
                    const double err = err_norm_seq<N>(ev, x, xn, rtol, atol);
                    if (!(err < kInf)) {
                        st = LYAP_PT_NONFINITE;
                        finished = true;
                    } else {
                        double h_next;
                        const double lerr = det_log(err);
                        const bool acc =
                            ctrl_update<INTEG>(ctrl, err, lerr, log_floor, h_try, h_max, h_next);
                        h = h_next;
                        if (!acc) {
                            ++nrej;
                        } else {
#pragma unroll
                            for (int r = 0; r < K; ++r)
                                tangent_step<M, INTEG, MIXED, TX>(X, p, pf, h_try, V[r]);
#pragma unroll
                            for (int i = 0; i < N; ++i) {
                                x[i] = xn[i];
                                k1[i] = k1_next[i];
                            }
                            t = clipped ? t_next : t + h_try;
                            ++nacc;
                            if (escaped<N>(x, esc)) {
                                st = LYAP_PT_ESCAPED;
                                finished = true;
                            } else if (clipped) {
                                mgs_rows<K, N>(V, ln);
                                ++m;
                                if (m > n_tr) {
#pragma unroll
                                    for (int r = 0; r < K; ++r) kahan_add(sum[r], comp[r], ln[r]);
                                }
                                if (m == m_total) finished = true;
                            }
                        }
                    }
                }
            }
        } else {
            base_step<M, INTEG, false, TX>(x, p, h_fixed, xn, nullptr, X, nullptr);
#pragma unroll
            for (int r = 0; r < K; ++r) tangent_step<M, INTEG, MIXED, TX>(X, p, pf, h_fixed, V[r]);
#pragma unroll
            for (int i = 0; i < N; ++i) x[i] = xn[i];
            ++nacc;
            if (++step_in == spr) {
                step_in = 0;
                bool bad = false;
#pragma unroll
                for (int i = 0; i < N; ++i) bad = bad || !isfinite(x[i]);
                if (bad) {
                    st = LYAP_PT_NONFINITE;
                    finished = true;
                } else if (escaped<N>(x, esc)) {
                    st = LYAP_PT_ESCAPED;
                    finished = true;
                } else {
                    mgs_rows<K, N>(V, ln);
                    ++m;
                    if (m > n_tr) {
#pragma unroll
                        for (int r = 0; r < K; ++r) kahan_add(sum[r], comp[r], ln[r]);
                    }
                    if (m == m_total) finished = true;
                }
            }
        }
        if (finished) {
            for (int r = 0; r < n_exps; ++r) {
                double s = 0.0;
#pragma unroll
                for (int q = 0; q < K; ++q)
                    if (q == r) s = sum[q];
                io.lambda[item * n_exps + r] = (st == LYAP_PT_OK) ? s / t_avg : __builtin_nan("");
            }
            if (io.status) io.status[item] = st;
            if (io.n_acc) io.n_acc[item] = nacc;
            if (io.n_rej) io.n_rej[item] = nrej;
            active = false;
        }
    }
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

__global__ void k_reset_counter(unsigned long long* c) {
    if (threadIdx.x == 0 && blockIdx.x == 0) *c = 0ULL;
}

// ---------------------------------------------------------------------------
// launch table
// ---------------------------------------------------------------------------
template <class K>
static hipError_t launch_persistent(K kern, const lyap_sweep_desc& d, const SweepIO& io,
                                    hipStream_t stream, int num_cus, int max_blocks_hint) {
    int per_cu = 0;
    hipError_t e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kern, 256, 0);
    if (e != hipSuccess) return e;
    if (per_cu < 1) per_cu = 1;
    int64_t blocks = (int64_t)per_cu * num_cus;
    const int64_t need = (io.count + 255) / 256;
    if (blocks > need) blocks = need;
    if (max_blocks_hint > 0 && blocks > max_blocks_hint) blocks = max_blocks_hint;
    if (blocks < 1) blocks = 1;
    hipLaunchKernelGGL(k_reset_counter, dim3(1), dim3(64), 0, stream, io.counter);
    hipLaunchKernelGGL(kern, dim3((unsigned)blocks), dim3(256), 0, stream, d, io);
    return hipGetLastError();
}

template <class M>
static hipError_t dispatch_model(const lyap_sweep_desc& d, const SweepIO& io, hipStream_t s,
                                 int num_cus, int max_blocks) {
    if (d.method == LYAP_METHOD_VAR) {
        const int key = (d.integrator << 2) | (d.adaptive ? 2 : 0) | (d.precision ? 1 : 0);
        switch (key) {
#define LYAP_CASE(I, A, X)                                                                  \
    case ((I) << 2) | ((A) ? 2 : 0) | ((X) ? 1 : 0):                                        \
        return launch_persistent(k_var_lane<M, I, A, X>, d, io, s, num_cus, max_blocks);
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

hipError_t launch_lane_sweep(const lyap_sweep_desc& d, const SweepIO& io, hipStream_t s,
                             int num_cus, int max_blocks) {
    switch (d.model) {
        case LYAP_MODEL_SHIMIZU:
            return dispatch_model<Shimizu>(d, io, s, num_cus, max_blocks);
        case LYAP_MODEL_LORENZ:
            return dispatch_model<Lorenz>(d, io, s, num_cus, max_blocks);
        case LYAP_MODEL_ROSSLER:
            return dispatch_model<Rossler>(d, io, s, num_cus, max_blocks);
        default:
            return hipErrorInvalidValue;
    }
}

}  // namespace lyap
