// Lorenz-96 device code shared by the block-per-point sweep kernel
// (lyap_l96.hip) and the wave-level drain used by the batched launch.
//
// Wave-level drain layout: a wave holds P points ("slots"), LP = 64/P lanes
// each.  Inside a slot, lane u owns component i = u % N of the base state and
// of E = N/G deviation vectors, G = LP/N being the vector groups per point:
// lane u holds the vectors r = u/N + G*m, m < E.  The slot's base state is
// replicated in each of its G groups (G = 1 at P = 64/N: no replication).
//   - neighbour couplings are shuffles inside the N-lane group;
//   - the error norm and the Jacobian-norm cap are xor butterflies over the
//     group, bit-identical in every lane of the group;
//   - renormalisation stages the slot's N x N deviation matrix in the wave's
//     LDS region, transposed so lane u < N owns vector u, and runs the same
//     right-looking MGS as the block kernel (sequential dot products) on the
//     rows in place, so the rows take no registers.
// Slots advance independently (their adaptive steps differ), so the wave's
// lanes diverge by slot; every control decision is taken from the slot's first
// lane so that a slot's lanes always agree, and the LDS exchange inside a slot
// orders the wave's own accesses (no workgroup barrier).
#pragma once

#include "lyap_device.hpp"
#include "lyap_lane.hpp"

namespace lyap {

template <int N>
LYAP_D double group_sum(double v) {
#pragma unroll
    for (int o = N / 2; o >= 1; o >>= 1) v = v + __shfl_xor(v, o);
    return v;
}

template <int N>
LYAP_D double group_max(double v) {
#pragma unroll
    for (int o = N / 2; o >= 1; o >>= 1) {
        const double w = __shfl_xor(v, o);
        v = v > w ? v : w;
    }
    return v;
}

LYAP_D bool wave_uniform(bool b) { return __builtin_amdgcn_readfirstlane((int)b) != 0; }

// orders this wave's LDS writes before its later LDS reads (cross-lane)
LYAP_D void wave_lds_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// doubles of LDS one wave of the drain uses for P points of L96-N: per point an
// N x N matrix with rows padded to N+1 doubles (lanes reading the same column
// of different rows then hit different banks)
template <int N, int P>
constexpr int l96_wave_lds_doubles() { return P * N * (N + 1); }

template <int N, int P, int INTEG, bool ADAPT, bool MIXED, class GetDesc>
LYAP_D void l96_wave_drain(const lyap_sweep_desc& d, GetDesc getd, const SweepIO& io, double* lds) {
    using TB = Tab<INTEG>;
    using T = typename std::conditional<MIXED, float, double>::type;
    constexpr int LP = 64 / P;  // lanes per point
    constexpr int G = LP / N;   // vector groups per point
    constexpr int K = N;        // exponents
    constexpr int E = K / G;    // vectors per lane
    constexpr int S = TB::S;
    static_assert(64 % P == 0 && LP % N == 0 && K % G == 0 && E >= 1, "layout");
    constexpr uint64_t SLOT_MASK = LP == 64 ? ~0ULL : ((1ULL << LP) - 1ULL);

    const int lane = (int)__lane_id();
    const int lead = lane - lane % LP;  // the slot's first lane
    const int u = lane % LP;
    const int i = u % N;
    const int vg = u / N;
    const int base_lane = lane - i;
    const int src_p1 = base_lane + (i + 1) % N;
    const int src_m1 = base_lane + (i + N - 1) % N;
    const int src_m2 = base_lane + (i + N - 2) % N;
    constexpr int RS = N + 1;  // LDS row stride (doubles)
    double* sV = lds + (lane / LP) * (N * RS);

    // a decision of the slot: the value its first lane computed
    auto slot_uniform = [lead](bool b) { return __shfl((int)b, lead) != 0; };
    auto slot_any = [lead](bool b) { return ((__ballot(b) >> lead) & SLOT_MASK) != 0ULL; };

    const int64_t n_tr = d.n_renorm_transient;
    const int64_t m_total = d.n_renorm_transient + d.n_renorm_average;
    const int64_t max_steps = d.max_steps, spr = d.steps_per_renorm;
    const double rtol = d.rtol, atol = d.atol, h_min = d.h_min, h_max = d.h_max;
    const double t_renorm = d.t_renorm, esc = d.escape_radius, h_fixed = d.h;
    const double t_avg = ADAPT ? (double)d.n_renorm_average * d.t_renorm
                               : (double)d.n_renorm_average * (double)d.steps_per_renorm * d.h;
    const int n_exps = d.n_exps;
    const double log_floor = det_log(1e-4);

    double x = 0.0, V[E], sum = 0.0, comp = 0.0, t = 0.0, h = 0.0, k1 = 0.0, F = 0.0, dd = 0.0;
    T dT = T(0);
    int64_t mr = 0, step_in = 0, pos = 0;
    uint32_t nacc = 0, nrej = 0;
    Ctrl ctrl{1e-4, log_floor, false};
    bool active = false, exhausted = false;
#pragma unroll
    for (int m = 0; m < E; ++m) V[m] = 0.0;

    for (;;) {
        const bool want = !active && !exhausted;
        int64_t got = grab_items(want && u == 0, io.counter);
        got = __shfl(got, lead);
        if (want) {
            if (got < io.count) {
                const auto dp = getd();
                pos = item_position(io, got);
                const int64_t idx = io.idx_list ? io.idx_list[pos] : io.first + pos;
                double p[2];
                point_params<2>(*dp, idx, io.params, p);
                F = p[0];
                dd = p[1];
                dT = (T)dd;
                x = io.inits ? io.inits[idx * N + i] : init_component(*dp, idx, i, p);
#pragma unroll
                for (int m = 0; m < E; ++m) V[m] = (vg + G * m == i) ? 1.0 : 0.0;
                sum = comp = 0.0;
                t = 0.0;
                h = dp->h;
                mr = step_in = 0;
                nacc = nrej = 0;
                ctrl.fac_old = 1e-4;
                ctrl.log_fac_old = log_floor;
                ctrl.rejected = false;
                if (INTEG == LYAP_INT_DOPRI5 && ADAPT) {
                    const double xp1 = __shfl(x, src_p1), xm1 = __shfl(x, src_m1),
                                 xm2 = __shfl(x, src_m2);
                    k1 = fma(xp1 - xm2, xm1, fma(-dd, x, F));
                }
                active = true;
            } else {
                exhausted = true;
            }
        }
        if (!__any(active)) break;
        if (!active) continue;

        uint32_t st = LYAP_PT_OK;
        bool finished = false, go = true;
        double h_try = h, t_next = 0.0;
        bool clipped = false;
        if constexpr (ADAPT) {
            if (slot_uniform((int64_t)nacc + (int64_t)nrej >= max_steps)) {
                st = LYAP_PT_MAXSTEPS;
                finished = true;
                go = false;
            } else {
                t_next = (double)(mr + 1) * t_renorm;
                const double xp1 = __shfl(x, src_p1), xm1 = __shfl(x, src_m1), xm2 = __shfl(x, src_m2);
                const double row = (2.0 * fabs(xm1) + fabs(xp1 - xm2)) + fabs(dd);
                const double jn = group_max<N>(row);
                if (h_try * jn > 1.0) h_try = 1.0 / jn;
                if (slot_uniform(t_next - t <= 1.0625 * h_try)) {
                    h_try = t_next - t;
                    clipped = true;
                }
                if (slot_uniform(h_try < h_min)) {
                    st = LYAP_PT_STEP_UNDERFLOW;
                    finished = true;
                    go = false;
                }
            }
        } else {
            h_try = h_fixed;
        }

        if (go) {
            // base step, own component; J data of every stage kept for the tangent
            double k[S];
            T jm1[S], jd[S];
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
                const double xp1 = __shfl(arg, src_p1), xm1 = __shfl(arg, src_m1),
                             xm2 = __shfl(arg, src_m2);
                jm1[s] = (T)xm1;
                jd[s] = (T)(xp1 - xm2);
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
                const double sk = fma(rtol, ax > an ? ax : an, atol);
                const double q = ev / sk;
                const double err = sqrt(group_sum<N>(q * q) / (double)N);
                if (slot_uniform(!(err < kInf))) {
                    st = LYAP_PT_NONFINITE;
                    finished = true;
                    accept = false;
                } else {
                    double h_next;
                    accept = ctrl_update<INTEG>(ctrl, err, det_log(err), log_floor, h_try, h_max,
                                                h_next);
                    accept = slot_uniform(accept);
                    h = h_next;
                    if (!accept) ++nrej;
                }
            }
            if (accept) {
                // tangent step of the lane's E vectors, J at every stage state
                const T hT = (T)h_try;
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
                        T acc = T(0);
                        bool any = false;
#pragma unroll
                        for (int j = 0; j < S; ++j) {
                            if (j >= s || TB::a(s, j) == 0.0) continue;
                            acc = any ? fma(row.c[j], kv[j], acc) : row.c[j] * kv[j];
                            any = true;
                        }
                        const T varg = any ? fma(hT, acc, v0) : v0;
                        const T vp1 = __shfl(varg, src_p1), vm1 = __shfl(varg, src_m1),
                                vm2 = __shfl(varg, src_m2);
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
                    if constexpr (MIXED) V[m] = V[m] + (double)(hT * accv);
                    else V[m] = fma(h_try, accv, V[m]);
                }
                x = xn;
                if (INTEG == LYAP_INT_DOPRI5 && ADAPT) k1 = k[S - 1];
                ++nacc;

                bool boundary;
                if constexpr (ADAPT) {
                    t = clipped ? t_next : t + h_try;
                    boundary = clipped;
                } else {
                    boundary = (++step_in == spr);
                    if (boundary) step_in = 0;
                }
                if (!ADAPT && boundary && slot_any(!isfinite(x))) {
                    st = LYAP_PT_NONFINITE;
                    finished = true;
                } else if ((ADAPT || boundary) && slot_any(!(fabs(x) <= esc))) {
                    st = LYAP_PT_ESCAPED;
                    finished = true;
                } else if (boundary) {
                    // renormalisation: MGS over the slot's matrix staged in LDS,
                    // row r (vector r) at sV + r*RS, worked on in place by lane r
                    wave_lds_sync();
#pragma unroll
                    for (int m = 0; m < E; ++m) sV[(vg + G * m) * RS + i] = V[m];
                    wave_lds_sync();
                    double* row = sV + u * RS;
                    double ln = 0.0;
                    for (int j = 0; j < K; ++j) {
                        if (u == j) {
                            double s2 = row[0] * row[0];
#pragma unroll 8
                            for (int c = 1; c < N; ++c) s2 = fma(row[c], row[c], s2);
                            const double nrm = sqrt(s2);
                            ln = det_log(nrm);
                            const double inv = 1.0 / nrm;
#pragma unroll 8
                            for (int c = 0; c < N; ++c) row[c] = row[c] * inv;
                        }
                        wave_lds_sync();
                        if (u > j && u < K) {
                            const double* q = sV + j * RS;
                            double cdot = q[0] * row[0];
#pragma unroll 8
                            for (int c = 1; c < N; ++c) cdot = fma(q[c], row[c], cdot);
#pragma unroll 8
                            for (int c = 0; c < N; ++c) row[c] = fma(-cdot, q[c], row[c]);
                        }
                    }
                    wave_lds_sync();
#pragma unroll
                    for (int m = 0; m < E; ++m) V[m] = sV[(vg + G * m) * RS + i];
                    ++mr;
                    if (mr > n_tr && u < K) kahan_add(sum, comp, ln);
                    if (mr == m_total) finished = true;
                }
            }
        }
        if (finished) {
            if (u < n_exps) io.lambda[pos * n_exps + u] = (st == LYAP_PT_OK) ? sum / t_avg : __builtin_nan("");
            if (u == 0) {
                if (io.status) io.status[pos] = st;
                if (io.n_acc) io.n_acc[pos] = nacc;
                if (io.n_rej) io.n_rej[pos] = nrej;
            }
            active = false;
        }
    }
}

}  // namespace lyap
