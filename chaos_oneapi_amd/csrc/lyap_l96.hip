// Lorenz-96 sweep kernels, N in {8, 16, 32}, full N-exponent spectrum.
//
// Layout: one point per block of W wave64s.  Lane l of wave w owns component
// i = l % N of the base state and of E = N*N/(64*W) deviation vectors, the
// vectors r = gid + (W*64/N)*m (gid = w*64/N + l/N, m < E).  The base state is
// replicated in every N-lane group (the replicas cost no extra instructions),
// so the Jacobian at each stage state is available to every vector's lanes.
//   - neighbour couplings x_{i+1}, x_{i-1}, x_{i-2} (and of each vector) are
//     moves inside the N-lane group (DPP row rotations for N = 16; for N = 32
//     row shifts plus v_permlane16_swap across the row pair; lane shuffles
//     otherwise);
//   - the error norm and the Jacobian-norm step cap are butterfly reductions
//     over the group (Ring::sum / Ring::max: the xor butterfly with offsets
//     N/2 .. 1, on the VALU for N = 16 and 32), identical in every lane;
//   - renormalisation stages the N x N deviation matrix in LDS, transposed so
//     lane r owns vector r, and runs right-looking MGS with one barrier per
//     vector (double-buffered q_j).
// Every lane computes the same control decisions from the same bits, so
// all branches are uniform; they are also made wave-uniform explicitly.
// The launch is persistent: a block takes its next point from the counter.
#include "lyap_device.hpp"
#include "lyap_l96_wave.hpp"
#include "lyap_launch.hpp"

namespace lyap {

__global__ void k_reset_counter_l96(unsigned long long* c) {
    if (threadIdx.x == 0 && blockIdx.x == 0) *c = 0ULL;
}

template <int N, int W, int INTEG, bool ADAPT, bool MIXED>
__global__ __launch_bounds__(64 * W) void k_l96(const lyap_sweep_desc d, const SweepIO io) {
    using TB = Tab<INTEG>;
    using T = typename std::conditional<MIXED, float, double>::type;
    constexpr int G = 64 / N;   // vector groups per wave
    constexpr int NG = W * G;   // vector groups per block
    constexpr int K = N;        // exponents
    constexpr int E = K / NG;   // vectors per lane
    constexpr int S = TB::S;
    static_assert(K % NG == 0 && E >= 1, "layout");

    __shared__ double sV[K * N];
    __shared__ double sQ[2][N];
    __shared__ unsigned long long s_item;

    const int tid = (int)threadIdx.x;
    const int lane = tid & 63;
    const int i = lane % N;
    const int gid = (tid >> 6) * G + lane / N;
    const Ring<N> ring(lane);

    const int64_t n_tr = d.n_renorm_transient;
    const int64_t m_total = d.n_renorm_transient + d.n_renorm_average;
    const int64_t max_steps = d.max_steps, spr = d.steps_per_renorm;
    const double rtol = d.rtol, atol = d.atol, h_min = d.h_min, h_max = d.h_max;
    const double t_renorm = d.t_renorm, esc = d.escape_radius, h_fixed = d.h;
    const double t_avg = ADAPT ? (double)d.n_renorm_average * d.t_renorm
                               : (double)d.n_renorm_average * (double)d.steps_per_renorm * d.h;
    const int n_exps = d.n_exps;
    const double log_floor = ctrl_log(1e-4);

    for (;;) {
        if (tid == 0) s_item = atomicAdd(io.counter, 1ULL);
        __syncthreads();
        const int64_t got = (int64_t)s_item;
        __syncthreads();
        if (got >= io.count) break;
        const int64_t pos = item_position(io, got);
        KernargDesc* dp = kernarg_desc();
        const int64_t idx = io.idx_list ? io.idx_list[pos] : io.first + pos;
        double p[2];
        point_params<2>(*dp, idx, io.params, p);
        const double F = p[0], dd = p[1];
        const T FT = (T)F, dT = (T)dd;
        double x = io.inits ? io.inits[idx * N + i] : init_component(*dp, idx, i, p);
        double V[E];
#pragma unroll
        for (int m = 0; m < E; ++m) V[m] = (gid + NG * m == i) ? 1.0 : 0.0;
        double sum = 0.0, comp = 0.0;
        double lmant = 1.0;  // open log window of row tid's norms (tid < K)
        int32_t lex = 0;
        int lcnt = 0;
        double t = 0.0, h = dp->h;
        int64_t mr = 0, step_in = 0;
        uint32_t nacc = 0, nrej = 0;
        Ctrl ctrl{log_floor, false};
        uint32_t st = LYAP_PT_OK;
        double k1 = 0.0;
        if (INTEG == LYAP_INT_DOPRI5 && ADAPT) {
            double xp1, xm1, xm2;
            ring.moves(x, xp1, xm1, xm2);
            k1 = fma(xp1 - xm2, xm1, fma(-dd, x, F));
        }

        for (;;) {
            double h_try = h;
            bool clipped = false;
            double t_next = 0.0;
            if constexpr (ADAPT) {
                if (wave_uniform((int64_t)nacc + (int64_t)nrej >= max_steps)) {
                    st = LYAP_PT_MAXSTEPS;
                    break;
                }
                t_next = (double)(mr + 1) * t_renorm;
                {
                    double xp1, xm1, xm2;
                    ring.moves(x, xp1, xm1, xm2);
                    const double row = (2.0 * fabs(xm1) + fabs(xp1 - xm2)) + fabs(dd);
                    const double jn = ring.max(row);
                    if (h_try * jn > 1.0) h_try = 1.0 / jn;
                }
                if (t_next - t <= 1.0625 * h_try) {
                    h_try = t_next - t;
                    clipped = true;
                }
                if (wave_uniform(h_try < h_min)) {
                    st = LYAP_PT_STEP_UNDERFLOW;
                    break;
                }
            } else {
                h_try = h_fixed;
            }

            // base step, own component; J data of every stage kept for the tangent
            double k[S];
            T jm1[S], jd[S];
            {
                const RowB<TB, double, S> rb;
#pragma unroll
                for (int s = 0; s < S; ++s) {
                    if (!ADAPT && !TB::for_solution(s)) {
                        k[s] = 0.0;
                        jm1[s] = jd[s] = T(0);
                        continue;
                    }
                    const RowA<TB, double, S> row(s);
                    double acc = 0.0;
                    bool any = false;
#pragma unroll
                    for (int j = 0; j < S; ++j) {
                        if (j >= s || TB::a(s, j) == 0.0) continue;
                        acc = any ? fma(row.c[j], k[j], acc) : row.c[j] * k[j];
                        any = true;
                    }
                    const double arg = any ? fma(h_try, acc, x) : x;
                    double xp1, xm1, xm2;
                    ring.moves(arg, xp1, xm1, xm2);
                    if constexpr (MIXED) {
                        // the binary32 tangent folds h into the Jacobian data
                        jm1[s] = (T)(h_try * xm1);
                        jd[s] = (T)(h_try * (xp1 - xm2));
                    } else {
                        jm1[s] = (T)xm1;
                        jd[s] = (T)(xp1 - xm2);
                    }
                    if (s == 0 && INTEG == LYAP_INT_DOPRI5 && ADAPT) k[0] = k1;
                    else k[s] = fma(xp1 - xm2, xm1, fma(-dd, arg, F));
                }
                double accb = 0.0;
                bool anyb = false;
#pragma unroll
                for (int j = 0; j < S; ++j) {
                    if (TB::b(j) == 0.0) continue;
                    accb = anyb ? fma(rb.c[j], k[j], accb) : rb.c[j] * k[j];
                    anyb = true;
                }
                const double xn = fma(h_try, accb, x);
                bool accept = true;
                if constexpr (ADAPT) {
                    const RowE<TB, double, S> re;
                    double acce = 0.0;
                    bool anye = false;
#pragma unroll
                    for (int j = 0; j < S; ++j) {
                        if (TB::e(j) == 0.0) continue;
                        acce = anye ? fma(re.c[j], k[j], acce) : re.c[j] * k[j];
                        anye = true;
                    }
                    const double ev = h_try * acce;
                    const double ax = fabs(x), an = fabs(xn);
                    const double sk = fma(rtol, fmax(ax, an), atol);
                    const double q = ev / sk;
                    const double err = sqrt(ring.sum(q * q) / (double)N);
                    if (wave_uniform(!(err < kInf))) {
                        st = LYAP_PT_NONFINITE;
                        break;
                    }
                    double h_next;
                    accept = ctrl_update<INTEG>(ctrl, err <= 1.0, err > 1e-4, ctrl_log(err), log_floor,
                                                h_try, h_max, h_next);
                    accept = wave_uniform(accept);
                    h = h_next;
                    if (!accept) {
                        ++nrej;
                        continue;
                    }
                }
                // tangent step of the lane's E vectors, J at every stage state
                // mixed form: h folded into the Jacobian data (jm1, jd and
                // dh = h d), so the stage arguments are v + sum_j a_sj K'_j and
                // the products K'_s = h K_s; the increment is sum_j b_j K'_j
                const T hT = (T)h_try;
                const T dh = MIXED ? (T)(h_try * dd) : dT;
                const RowB<TB, T, S> rbt;
#pragma unroll
                for (int m = 0; m < E; ++m) {
                    const T v0 = (T)V[m];
                    T kv[S];
#pragma unroll
                    for (int s = 0; s < S; ++s) {
                        if (!TB::for_solution(s)) {
                            kv[s] = T(0);
                            continue;
                        }
                        const RowA<TB, T, S> row(s);
                        T acc = MIXED ? v0 : T(0);
                        bool any = false;
#pragma unroll
                        for (int j = 0; j < S; ++j) {
                            if (j >= s || TB::a(s, j) == 0.0) continue;
                            acc = (any || MIXED) ? fma(row.c[j], kv[j], acc) : row.c[j] * kv[j];
                            any = true;
                        }
                        const T varg = MIXED ? acc : (any ? fma(hT, acc, v0) : v0);
                        T vp1, vm1, vm2;
                        ring.moves(varg, vp1, vm1, vm2);
                        if constexpr (MIXED)
                            kv[s] = fma(vp1 - vm2, jm1[s], fma(-dh, varg, jd[s] * vm1));
                        else
                            kv[s] = fma(vp1 - vm2, jm1[s], fma(jd[s], vm1, -dT * varg));
                    }
                    T accv = T(0);
                    bool anyv = false;
#pragma unroll
                    for (int j = 0; j < S; ++j) {
                        if (TB::b(j) == 0.0) continue;
                        accv = anyv ? fma(rbt.c[j], kv[j], accv) : rbt.c[j] * kv[j];
                        anyv = true;
                    }
                    if constexpr (MIXED) V[m] = V[m] + (double)accv;
                    else V[m] = fma(h_try, accv, V[m]);
                }
                x = xn;
                if (INTEG == LYAP_INT_DOPRI5 && ADAPT) k1 = k[S - 1];
            }
            ++nacc;
            bool boundary;
            if constexpr (ADAPT) {
                t = clipped ? t_next : t + h_try;
                boundary = clipped;
            } else {
                boundary = (++step_in == spr);
                if (boundary) step_in = 0;
            }
            if constexpr (!ADAPT) {
                if (boundary && wave_uniform(__any(!isfinite(x)))) {
                    st = LYAP_PT_NONFINITE;
                    break;
                }
            }
            if (ADAPT || boundary) {
                if (wave_uniform(__any(!(fabs(x) <= esc)))) {
                    st = LYAP_PT_ESCAPED;
                    break;
                }
            }
            if (!wave_uniform(boundary)) continue;

            // renormalisation: MGS over the deviation matrix staged in LDS
            __syncthreads();
#pragma unroll
            for (int m = 0; m < E; ++m) sV[(gid + NG * m) * N + i] = V[m];
            __syncthreads();
            double vr[N];
            double vn = 0.0;
            if (tid < K) {
#pragma unroll
                for (int c = 0; c < N; ++c) vr[c] = sV[tid * N + c];
            }
            for (int j = 0; j < K; ++j) {
                if (tid == j) {
                    double s2 = vr[0] * vr[0];
#pragma unroll
                    for (int c = 1; c < N; ++c) s2 = fma(vr[c], vr[c], s2);
                    const double inv = inv_sqrt_det(s2);
                    vn = s2 * inv;  // accumulated after the loop, by every row lane at once
#pragma unroll
                    for (int c = 0; c < N; ++c) {
                        vr[c] = vr[c] * inv;
                        sQ[j & 1][c] = vr[c];
                    }
                }
                __syncthreads();
                if (tid > j && tid < K) {
                    const double* q = sQ[j & 1];
                    double cdot = q[0] * vr[0];
#pragma unroll
                    for (int c = 1; c < N; ++c) cdot = fma(q[c], vr[c], cdot);
#pragma unroll
                    for (int c = 0; c < N; ++c) vr[c] = fma(-cdot, q[c], vr[c]);
                }
            }
            if (tid < K) {
#pragma unroll
                for (int c = 0; c < N; ++c) sV[tid * N + c] = vr[c];
            }
            __syncthreads();
#pragma unroll
            for (int m = 0; m < E; ++m) V[m] = sV[(gid + NG * m) * N + i];
            ++mr;
            if (mr > n_tr && tid < K) {
                lognorm_push(lmant, lex, vn);
                if (++lcnt == kLogWindow || mr == m_total) {
                    lognorm_flush(lmant, lex, sum, comp);
                    lcnt = 0;
                }
            }
            if (mr == m_total) break;
        }

        if (tid < n_exps) io.lambda[pos * n_exps + tid] = (st == LYAP_PT_OK) ? sum / t_avg : __builtin_nan("");
        if (tid == 0) {
            if (io.status) io.status[pos] = st;
            if (io.n_acc) io.n_acc[pos] = nacc;
            if (io.n_rej) io.n_rej[pos] = nrej;
        }
    }
}

// wave-level drain: P points per wave64, one wave per workgroup
template <int N, int P, int INTEG, bool ADAPT, bool MIXED>
__global__ __launch_bounds__(64) void k_l96_wave(const lyap_sweep_desc d, const SweepIO io) {
    __shared__ double lds[l96_wave_lds_doubles<N, P>()];
    l96_wave_drain<N, P, INTEG, ADAPT, MIXED>(d, [] { return kernarg_desc(); }, io, lds);
}

// the same with a register budget for 3 waves per SIMD: the layouts whose
// compiler-chosen budget runs 2 (N = 16 with 4 points per wave: 179-187 VGPRs;
// N = 32 with 1: 183-197) fit 168 with a few spilled registers and run 4-8 %
// faster; the other layouts already fit
template <int N, int P, int INTEG, bool ADAPT, bool MIXED>
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(3))) void k_l96_wave_w3(
    const lyap_sweep_desc d, const SweepIO io) {
    __shared__ double lds[l96_wave_lds_doubles<N, P>()];
    l96_wave_drain<N, P, INTEG, ADAPT, MIXED>(d, [] { return kernarg_desc(); }, io, lds);
}

namespace {

template <class Kern>
hipError_t launch_persistent_l96(Kern kern, int threads, const lyap_sweep_desc& d, const SweepIO& io,
                                 int64_t points_per_block, hipStream_t s, int num_cus, int max_blocks) {
    int per_cu = 0;
    hipError_t e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kern, threads, 0);
    if (e != hipSuccess) return e;
    if (per_cu < 1) per_cu = 1;
    int64_t blocks = (int64_t)per_cu * num_cus;
    const int64_t need = (io.count + points_per_block - 1) / points_per_block;
    if (blocks > need) blocks = need;
    if (max_blocks > 0 && blocks > max_blocks) blocks = max_blocks;
    if (blocks < 1) blocks = 1;
    hipLaunchKernelGGL(k_reset_counter_l96, dim3(1), dim3(64), 0, s, io.counter);
    hipLaunchKernelGGL(kern, dim3((unsigned)blocks), dim3(threads), 0, s, d, io);
    return hipGetLastError();
}

template <int N, int W, int INTEG, bool ADAPT, bool MIXED>
hipError_t launch_one(const lyap_sweep_desc& d, const SweepIO& io, hipStream_t s, int num_cus,
                      int max_blocks) {
    return launch_persistent_l96(k_l96<N, W, INTEG, ADAPT, MIXED>, 64 * W, d, io, 1, s, num_cus,
                                 max_blocks);
}

template <int N, int P, int INTEG, bool ADAPT, bool MIXED>
hipError_t launch_wave_one(const lyap_sweep_desc& d, const SweepIO& io, hipStream_t s, int num_cus,
                           int max_blocks) {
    if constexpr ((N == 16 && P == 4) || (N == 32 && P == 1))
        return launch_persistent_l96(k_l96_wave_w3<N, P, INTEG, ADAPT, MIXED>, 64, d, io, P, s, num_cus,
                                     max_blocks);
    else
        return launch_persistent_l96(k_l96_wave<N, P, INTEG, ADAPT, MIXED>, 64, d, io, P, s, num_cus,
                                     max_blocks);
}

#define LYAP_L96_VARIANTS(LAUNCH, ...)                                                    \
    switch ((d.integrator << 2) | (d.adaptive ? 2 : 0) | (d.precision ? 1 : 0)) {           \
        case (LYAP_INT_DVERK << 2) | 0: return LAUNCH<__VA_ARGS__, LYAP_INT_DVERK, false, false>(d, io, s, num_cus, max_blocks);   \
        case (LYAP_INT_DVERK << 2) | 1: return LAUNCH<__VA_ARGS__, LYAP_INT_DVERK, false, true>(d, io, s, num_cus, max_blocks);    \
        case (LYAP_INT_DVERK << 2) | 2: return LAUNCH<__VA_ARGS__, LYAP_INT_DVERK, true, false>(d, io, s, num_cus, max_blocks);    \
        case (LYAP_INT_DVERK << 2) | 3: return LAUNCH<__VA_ARGS__, LYAP_INT_DVERK, true, true>(d, io, s, num_cus, max_blocks);     \
        case (LYAP_INT_DOPRI5 << 2) | 0: return LAUNCH<__VA_ARGS__, LYAP_INT_DOPRI5, false, false>(d, io, s, num_cus, max_blocks); \
        case (LYAP_INT_DOPRI5 << 2) | 1: return LAUNCH<__VA_ARGS__, LYAP_INT_DOPRI5, false, true>(d, io, s, num_cus, max_blocks);  \
        case (LYAP_INT_DOPRI5 << 2) | 2: return LAUNCH<__VA_ARGS__, LYAP_INT_DOPRI5, true, false>(d, io, s, num_cus, max_blocks);  \
        case (LYAP_INT_DOPRI5 << 2) | 3: return LAUNCH<__VA_ARGS__, LYAP_INT_DOPRI5, true, true>(d, io, s, num_cus, max_blocks);   \
        default: return hipErrorInvalidValue;                                                  \
    }

template <int N, int W>
hipError_t launch_nw(const lyap_sweep_desc& d, const SweepIO& io, hipStream_t s, int num_cus,
                     int max_blocks) {
    LYAP_L96_VARIANTS(launch_one, N, W)
}

template <int N, int P>
hipError_t launch_np(const lyap_sweep_desc& d, const SweepIO& io, hipStream_t s, int num_cus,
                     int max_blocks) {
    LYAP_L96_VARIANTS(launch_wave_one, N, P)
}

#undef LYAP_L96_VARIANTS

}  // namespace

int l96_default_waves(int n) { return n == 8 ? 1 : (n == 16 ? 2 : 4); }

bool l96_supported(int n, int waves) {
    return (n == 8 && waves == 1) || (n == 16 && (waves == 1 || waves == 2 || waves == 4)) ||
           (n == 32 && (waves == 2 || waves == 4));
}

bool l96_wave_supported(int n, int ppw) {
    return (n == 8 && (ppw == 1 || ppw == 2 || ppw == 4 || ppw == 8)) ||
           (n == 16 && (ppw == 1 || ppw == 2 || ppw == 4)) || (n == 32 && (ppw == 1 || ppw == 2));
}

hipError_t launch_l96_sweep(const lyap_sweep_desc& d, const SweepIO& io, hipStream_t s, int num_cus,
                            int waves, int ppw, int max_blocks) {
    if (d.method != LYAP_METHOD_VAR) return hipErrorInvalidValue;
    if (ppw > 0) {
        if (!l96_wave_supported(d.dim, ppw)) return hipErrorInvalidValue;
        switch (d.dim * 16 + ppw) {
            case 8 * 16 + 1: return launch_np<8, 1>(d, io, s, num_cus, max_blocks);
            case 8 * 16 + 2: return launch_np<8, 2>(d, io, s, num_cus, max_blocks);
            case 8 * 16 + 4: return launch_np<8, 4>(d, io, s, num_cus, max_blocks);
            case 8 * 16 + 8: return launch_np<8, 8>(d, io, s, num_cus, max_blocks);
            case 16 * 16 + 1: return launch_np<16, 1>(d, io, s, num_cus, max_blocks);
            case 16 * 16 + 2: return launch_np<16, 2>(d, io, s, num_cus, max_blocks);
            case 16 * 16 + 4: return launch_np<16, 4>(d, io, s, num_cus, max_blocks);
            case 32 * 16 + 1: return launch_np<32, 1>(d, io, s, num_cus, max_blocks);
            case 32 * 16 + 2: return launch_np<32, 2>(d, io, s, num_cus, max_blocks);
            default: return hipErrorInvalidValue;
        }
    }
    if (waves <= 0) waves = l96_default_waves(d.dim);
    if (!l96_supported(d.dim, waves)) return hipErrorInvalidValue;
    switch (d.dim * 16 + waves) {
        case 8 * 16 + 1: return launch_nw<8, 1>(d, io, s, num_cus, max_blocks);
        case 16 * 16 + 1: return launch_nw<16, 1>(d, io, s, num_cus, max_blocks);
        case 16 * 16 + 2: return launch_nw<16, 2>(d, io, s, num_cus, max_blocks);
        case 16 * 16 + 4: return launch_nw<16, 4>(d, io, s, num_cus, max_blocks);
        case 32 * 16 + 2: return launch_nw<32, 2>(d, io, s, num_cus, max_blocks);
        case 32 * 16 + 4: return launch_nw<32, 4>(d, io, s, num_cus, max_blocks);
        default: return hipErrorInvalidValue;
    }
}

}  // namespace lyap
