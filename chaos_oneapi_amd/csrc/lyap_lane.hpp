// Per-lane device code of the 3-D flows (shimizu, Lorenz, Rössler): the base
// RK step, the coupled tangent step, MGS, the error norm and the persistent
// per-lane loop.  Used by the standalone sweep kernels (lyap_sweep.hip) and
// the batched mixed-model launch (lyap_batch.hip).
#pragma once

#include "lyap_device.hpp"

namespace lyap {

// ---------------------------------------------------------------------------
// base RK step: x -> xn; stage states X[s] (for the coupled tangent); error
// vector when ERR.  With FSAL (Dopri5 + ERR) k1 comes in from the previous
// accepted step and the new k1 (= f(xn)) goes out.  With TRACE, *tr_step is
// the step's integral of tr J by the solution's quadrature:
// h * (c + sum_j b_j g(X_j)) for tr J = c(p) + g(x) (h * c when g is absent).
// ---------------------------------------------------------------------------
template <class M, int INTEG, bool ERR, class TX, bool TRACE = false>
LYAP_D void base_step(const double* x, const double* p, double h, double* xn, double* ev,
                      typename M::template J<TX> (&X)[Tab<INTEG>::S], double* k1_io,
                      double* tr_step = nullptr) {
    using TB = Tab<INTEG>;
    constexpr int S = TB::S, N = M::N;
    constexpr bool FSAL = (INTEG == LYAP_INT_DOPRI5) && ERR;
    double k[S][N];
    double trq = 0.0;
    bool tr_any = false;
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
        if constexpr (TRACE && M::kTraceVar) {
            if (TB::b(s) != 0.0) {
                const double bs = sconst(TB::b(s));
                trq = tr_any ? fma(bs, M::trace_var(arg), trq) : bs * M::trace_var(arg);
                tr_any = true;
            }
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
    if constexpr (TRACE) *tr_step = M::kTraceVar ? h * (M::trace_const(p) + trq) : h * M::trace_const(p);
}

// stages of the embedded fifth-order weights (DVERK); other tableaux have none
template <class TB>
struct LowStages {
    static constexpr int value = TB::S;
};
template <>
struct LowStages<Tab<LYAP_INT_DVERK>> {
    static constexpr int value = Tab<LYAP_INT_DVERK>::S_LOW;
};

// adaptive steps take the low-order tangent in the mixed DVERK form
template <int INTEG, bool MIXED>
inline constexpr bool kLowTangent = MIXED && INTEG == LYAP_INT_DVERK;

// coupled tangent step for one deviation vector v (FP64 storage).
// FP64 form: stages in double.  Mixed form: stages in binary32, increment
// h*sum(b_j K_j) formed in binary32 and added to the FP64 vector.  LOW (the
// mixed form with adaptive DVERK): the embedded fifth-order weights b - e on
// stages 1-6 (the base step computes all eight stage states for its error
// estimate), one stage fewer than the sixth-order solution's seven.
template <class M, int INTEG, bool MIXED, class TX, bool LOW = false>
LYAP_D void tangent_step(const typename M::template J<TX> (&X)[Tab<INTEG>::S], const double* p,
                         const float* pf, double h, double* v) {
    using TB = Tab<INTEG>;
    constexpr int S = TB::S, N = M::N;
    static_assert(!LOW || (MIXED && INTEG == LYAP_INT_DVERK), "the low-order tangent is the mixed DVERK form's");
    using T = typename std::conditional<MIXED, float, double>::type;
    const T hT = (T)h;
    T vt[N];
#pragma unroll
    for (int i = 0; i < N; ++i) vt[i] = (T)v[i];
    T kv[S][N];
#pragma unroll
    for (int s = 0; s < S; ++s) {
        if (LOW ? s >= LowStages<TB>::value : !TB::for_solution(s)) {
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
    if constexpr (LOW) {
        const RowBh<TB, T, S> rbh;
#pragma unroll
        for (int i = 0; i < N; ++i) {
            const float inc = hT * coef_sum_bh(rbh, kv, i);
            v[i] = v[i] + (double)inc;
        }
    } else {
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
}

// modified Gram-Schmidt on K rows of length N (sequential fma dot products);
// nrm[j] = row j's norm before its normalisation
template <int K, int N>
LYAP_D void mgs_rows(double (&V)[K][N], double* nrm_out) {
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
        const double inv = inv_sqrt_det(s);
        nrm_out[j] = s * inv;
#pragma unroll
        for (int i = 0; i < N; ++i) V[j][i] = V[j][i] * inv;
    }
}

// The error norm squared with one division:
//   err^2 = (1/N) sum_i (ev_i / sk_i)^2 = sum_i ev_i^2 w_i / (N prod_j A_j),
// A_j = sk_j^2, w_i = prod_{j != i} A_j as prefix x suffix products (a factor 1
// left out).  Three divisions, a square root and a division by N become one
// division; the 3-D flows only (products of N - 1 squared scales in range).
template <int N>
// fmax(|x|, |x_new|): one v_max_f64 with abs modifiers.  It differs from a
// compare-and-select only when |x_new| is NaN, and then e is not finite either
// (every stage that feeds x_new also feeds the error estimate), so the norm is
// not finite and the point ends as non-finite all the same.
LYAP_D double err2_seq(const double* ev, const double* x, const double* xn, double rtol, double atol) {
    double A[N], pre[N], suf[N];
#pragma unroll
    for (int i = 0; i < N; ++i) {
        const double ax = fabs(x[i]), an = fabs(xn[i]);
        const double sk = fma(rtol, fmax(ax, an), atol);
        A[i] = sk * sk;
    }
    pre[0] = 1.0;
#pragma unroll
    for (int i = 1; i < N; ++i) pre[i] = (i == 1) ? A[0] : pre[i - 1] * A[i - 1];
    suf[N - 1] = 1.0;
#pragma unroll
    for (int i = N - 2; i >= 0; --i) suf[i] = (i == N - 2) ? A[N - 1] : A[i + 1] * suf[i + 1];
    double num = 0.0;
#pragma unroll
    for (int i = 0; i < N; ++i) {
        const double w = (i == 0) ? suf[0] : ((i == N - 1) ? pre[N - 1] : pre[i] * suf[i]);
        const double e2 = ev[i] * ev[i];
        num = (i == 0) ? e2 * w : fma(e2, w, num);
    }
    const double den = (double)N * ((N == 1) ? A[0] : pre[N - 1] * A[N - 1]);
    return num / den;
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

// no progress signals or skips (the batched launch)
struct NoWatch {
    LYAP_D KernargWatch* operator()() const { return nullptr; }
};

// A lane finished (or skipped) watched item `item`: its outputs are released
// to the device scope before the count, and the lane that completes a
// milestone stores the launch's sequence number to the host-visible flag.
template <class W>
LYAP_D void watch_count(W* w, int64_t item) {
    if (!w->done || item >= w->end[1]) return;
    const int k = item < w->end[0] ? 0 : 1;
    const int64_t size = k ? w->end[1] - w->end[0] : w->end[0];
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
    const unsigned long long c =
        __hip_atomic_fetch_add(w->done + k, 1ULL, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    if ((int64_t)c + 1 == size) {
        // the other lanes' release fences synchronise with this acquire fence
        // through the counter's increments, so their outputs are ordered before
        // the flag the host reads
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
        __hip_atomic_store(w->flag + k, w->seq, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_SYSTEM);
    }
}

// Item `item` of a watched launch is skipped: its outputs say so (status
// LYAP_PT_SKIPPED, NaN exponents, no steps) and it counts for its milestone.
template <class GetWatch, class GetIO, class GetDesc>
LYAP_D bool skip_item(GetWatch getw, GetIO getio, GetDesc getd, int64_t item) {
    const auto w = getw();
    if (!w->skip) return false;
    if (__hip_atomic_load(w->skip + item, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM) == 0u) return false;
    const auto dp = getd();
    const auto iop = getio();
    for (int r = 0; r < dp->n_exps; ++r) iop->lambda[item * dp->n_exps + r] = __builtin_nan("");
    if (iop->status) iop->status[item] = LYAP_PT_SKIPPED;
    if (iop->n_acc) iop->n_acc[item] = 0u;
    if (iop->n_rej) iop->n_rej[item] = 0u;
    watch_count(w, item);
    return true;
}

// Persistent per-lane loop over the points of one sweep: lanes that finish a
// point take the next item from io.counter until the items are exhausted and
// every lane of the wave is done.  getd() and getio() return the descriptor and
// the IO record to read the per-point set-up and output fields from (opaque
// pointers, so those fields are not kept live in SGPRs across the step loop);
// io supplies only the fields every iteration reads (counter, count, gather).
// getw() returns the launch's LaneWatch record (progress signals, skips) or
// null.
template <class M, int INTEG, bool ADAPT, bool MIXED, bool TRACE = false, class GetDesc, class GetIO,
          class GetWatch = NoWatch>
LYAP_D void lane_drain(const lyap_sweep_desc& d, GetDesc getd, const SweepIO& io, GetIO getio,
                       GetWatch getw = GetWatch()) {
    // TRACE: N - 1 deviation vectors; the last exponent is the averaged
    // integral of tr J (Liouville) minus the others
    constexpr int N = M::N, P = M::P, K = TRACE ? M::N - 1 : M::N, S = Tab<INTEG>::S;
    using TX = typename std::conditional<MIXED, float, double>::type;
    constexpr bool kWatch = !std::is_same<GetWatch, NoWatch>::value;
    // scalars the step loop reads; everything else of the descriptor is
    // re-read through an opaque pointer when a lane takes a new point, so the
    // compiler does not keep (and spill) the whole descriptor in SGPRs
    const int64_t n_tr = d.n_renorm_transient;
    const int64_t m_total = d.n_renorm_transient + d.n_renorm_average;
    const int64_t max_steps = d.max_steps, spr = d.steps_per_renorm;
    const double rtol = d.rtol, atol = d.atol, h_min = d.h_min, h_max = d.h_max;
    const double t_renorm = d.t_renorm, esc = d.escape_radius, h_fixed = d.h;
    const double log_floor = ctrl_log(1e-4);
    // io gives the fields read every iteration (work counter, count, gather);
    // the others are re-read through getio() when a lane takes or finishes a
    // point
    unsigned long long* const counter = io.counter;
    const int64_t count = io.count;

    double x[N], V[K][N], sum[K], comp[K], lmant[K], p[P], k1[N];
    int32_t lex[K];
    int lcnt = 0;  // renormalisations in the open log window
    double tri = 0.0, trs = 0.0, trc = 0.0;  // TRACE: tr J integral (interval, Kahan sum)
    float pf[P];
    double t = 0.0, h = 0.0;
    double t_next = 0.0;  // (m + 1) * t_renorm: the next renormalisation time
    int64_t m = 0, step_in = 0, item = -1;
    uint32_t nacc = 0, nrej = 0;
    Ctrl ctrl{log_floor, false};
    bool active = false, exhausted = false;
    // a lane that reached a renormalisation boundary waits (pending) until
    // io.gather lanes of the wave wait, or every active lane does; then they
    // renormalise together.  Each lane's arithmetic is unchanged, only the
    // wave executes the MGS block less often with most lanes masked off.
    bool pending = false;
    const int gather = io.gather;

    auto write_out = [&](uint32_t st) {
        const auto dp = getd();
        const auto iop = getio();
        const double t_avg = ADAPT ? (double)dp->n_renorm_average * dp->t_renorm
                                   : (double)dp->n_renorm_average * (double)dp->steps_per_renorm * dp->h;
        const int n_exps = dp->n_exps;
        double* const lam = iop->lambda;
        double closing = trs;
#pragma unroll
        for (int q = 0; q < K; ++q) closing = closing - sum[q];
        for (int r = 0; r < n_exps; ++r) {
            double s = 0.0;
#pragma unroll
            for (int q = 0; q < K; ++q)
                if (q == r) s = sum[q];
            if (TRACE && r == K) s = closing;
            lam[item * n_exps + r] = (st == LYAP_PT_OK) ? s / t_avg : __builtin_nan("");
        }
        if (iop->status) iop->status[item] = st;
        if (iop->n_acc) iop->n_acc[item] = nacc;
        if (iop->n_rej) iop->n_rej[item] = nrej;
        if constexpr (kWatch) watch_count(getw(), item);
    };

    for (;;) {
        const int64_t got = grab_items(!active && !exhausted, counter);
        // taking a point and finishing one are rare next to a step: the hints
        // keep the register allocator's spill code out of the step loop
        if (__builtin_expect(!active && !exhausted, 0)) {
            bool skipped = false;
            if constexpr (kWatch) {
                if (got < count) skipped = skip_item(getw, getio, getd, item_position(got, getio()->stride, count));
            }
            if (skipped) {
                // not integrated (outputs written); the lane takes the next item
            } else if (got < count) {
                const auto dp = getd();
                const auto iop = getio();
                item = item_position(got, iop->stride, count);
                const int64_t idx = iop->idx_list ? iop->idx_list[item] : iop->first + item;
                point_params<P>(*dp, idx, iop->params, p);
#pragma unroll
                for (int j = 0; j < P; ++j) pf[j] = (float)p[j];
                const double* const inits = iop->inits;
#pragma unroll
                for (int i = 0; i < N; ++i)
                    x[i] = inits ? inits[idx * N + i] : init_component(*dp, idx, i, p);
#pragma unroll
                for (int r = 0; r < K; ++r) {
                    sum[r] = 0.0;
                    comp[r] = 0.0;
                    lmant[r] = 1.0;
                    lex[r] = 0;
#pragma unroll
                    for (int i = 0; i < N; ++i) V[r][i] = (r == i) ? 1.0 : 0.0;
                }
                lcnt = 0;
                tri = trs = trc = 0.0;
                t = 0.0;
                h = dp->h;
                m = 0;
                t_next = t_renorm;
                step_in = 0;
                nacc = nrej = 0;
                ctrl.log_fac_old = log_floor;
                ctrl.rejected = false;
                if (INTEG == LYAP_INT_DOPRI5 && ADAPT) M::rhs(x, p, k1);
                active = true;
            } else {
                exhausted = true;
            }
        }
        if constexpr (kWatch) {
            // a lane whose item was skipped is neither active nor exhausted
            if (!__any(active || !exhausted)) break;
        } else {
            if (!__any(active)) break;
        }
        {
            const uint64_t pm = __ballot(active && pending);
            if (pm) {
                const int np = __popcll(pm);
                if (np >= gather || np == __popcll(__ballot(active))) {
                    if (pending) {
                        double nrm[K];
                        mgs_rows<K, N>(V, nrm);
                        ++m;
                        t_next = (double)(m + 1) * t_renorm;
                        if (m > n_tr) {
#pragma unroll
                            for (int r = 0; r < K; ++r) lognorm_push(lmant[r], lex[r], nrm[r]);
                            if (++lcnt == kLogWindow || m == m_total) {
#pragma unroll
                                for (int r = 0; r < K; ++r) lognorm_flush(lmant[r], lex[r], sum[r], comp[r]);
                                lcnt = 0;
                            }
                            if constexpr (TRACE) kahan_add(trs, trc, tri);
                        }
                        if constexpr (TRACE) tri = 0.0;
                        pending = false;
                        if (m == m_total) {
                            write_out(LYAP_PT_OK);
                            active = false;
                        }
                    }
                }
            }
        }
        if (!active || pending) continue;

        uint32_t st = LYAP_PT_OK;
        bool finished = false;
        double xn[N];
        typename M::template J<TX> X[S];
        if constexpr (ADAPT) {
            if ((int64_t)nacc + (int64_t)nrej >= max_steps) {
                st = LYAP_PT_MAXSTEPS;
                finished = true;
            } else {
                double h_try = h;
                bool clipped = false;
                const double jn = M::jnorm(x, p);
                // the division only when a lane of the wave needs the cap (the
                // empty asm keeps it from being hoisted above the branch)
                if (__any(h_try * jn > 1.0)) {
                    double jv = jn;
                    asm volatile("" : "+v"(jv));
                    if (h_try * jn > 1.0) h_try = 1.0 / jv;
                }
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
                    double ev[N], k1_next[N], tr_step = 0.0;
#pragma unroll
                    for (int i = 0; i < N; ++i) k1_next[i] = k1[i];
                    base_step<M, INTEG, true, TX, TRACE>(x, p, h_try, xn, ev, X, k1_next, &tr_step);
                    const double e2 = err2_seq<N>(ev, x, xn, rtol, atol);
                    if (!(e2 < kInf)) {
                        st = LYAP_PT_NONFINITE;
                        finished = true;
                    } else {
                        double h_next;
                        const double lerr = 0.5 * ctrl_log(e2);
                        const bool acc = ctrl_update<INTEG>(ctrl, e2 <= 1.0, e2 > 1e-8, lerr, log_floor,
                                                            h_try, h_max, h_next);
                        h = h_next;
                        if (!acc) {
                            ++nrej;
                        } else {
#pragma unroll
                            for (int r = 0; r < K; ++r)
                                tangent_step<M, INTEG, MIXED, TX, kLowTangent<INTEG, MIXED>>(X, p, pf, h_try,
                                                                                             V[r]);
                            if constexpr (TRACE) tri = tri + tr_step;
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
                                pending = true;
                            }
                        }
                    }
                }
            }
        } else {
            double tr_step = 0.0;
            base_step<M, INTEG, false, TX, TRACE>(x, p, h_fixed, xn, nullptr, X, nullptr, &tr_step);
#pragma unroll
            for (int r = 0; r < K; ++r) tangent_step<M, INTEG, MIXED, TX>(X, p, pf, h_fixed, V[r]);
            if constexpr (TRACE) tri = tri + tr_step;
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
                    pending = true;
                }
            }
        }
        if (__builtin_expect(finished, 0)) {
            write_out(st);
            active = false;
        }
    }
}

// lane r of a quad receives lane q of the same quad (DPP quad_perm broadcast)
template <int Q>
LYAP_D double quad_bcast(double v) {
    constexpr int ctrl = Q | (Q << 2) | (Q << 4) | (Q << 6);
    const uint64_t b = __builtin_bit_cast(uint64_t, v);
    const uint32_t lo = (uint32_t)__builtin_amdgcn_mov_dpp((int)(uint32_t)b, ctrl, 0xF, 0xF, true);
    const uint32_t hi = (uint32_t)__builtin_amdgcn_mov_dpp((int)(uint32_t)(b >> 32), ctrl, 0xF, 0xF, true);
    return __builtin_bit_cast(double, ((uint64_t)hi << 32) | (uint64_t)lo);
}

// Persistent loop of the sweep's method with four lanes per point (a quad),
// for launches that are about one point per lane and so last as long as their
// slowest point.  Every lane of the quad runs the point's base step, error
// norm and step control (identical operands, identical results, so the quad's
// control flow agrees without communication); lane r < K advances deviation
// vector r alone, so a step costs one tangent vector instead of K.  At a
// renormalisation the quad's vectors are broadcast to every lane (DPP
// quad_perm), every lane runs the same MGS on the full matrix and keeps its
// own row; lane r accumulates exponent r.  The arithmetic of each point is
// that of lane_drain, so both give the same bits.
template <class M, int INTEG, bool ADAPT, bool MIXED, bool TRACE = false, class GetDesc, class GetIO>
LYAP_D void quad_drain(const lyap_sweep_desc& d, GetDesc getd, const SweepIO& io, GetIO getio) {
    // TRACE: N - 1 vectors (lanes 0 .. N - 2); every lane integrates tr J with
    // the replicated base step, and lane N - 1 writes the closing exponent
    constexpr int N = M::N, P = M::P, K = TRACE ? M::N - 1 : M::N, S = Tab<INTEG>::S;
    static_assert(K <= 4, "four lanes per point");
    using TX = typename std::conditional<MIXED, float, double>::type;
    const int64_t n_tr = d.n_renorm_transient;
    const int64_t m_total = d.n_renorm_transient + d.n_renorm_average;
    const int64_t max_steps = d.max_steps, spr = d.steps_per_renorm;
    const double rtol = d.rtol, atol = d.atol, h_min = d.h_min, h_max = d.h_max;
    const double t_renorm = d.t_renorm, esc = d.escape_radius, h_fixed = d.h;
    const double log_floor = ctrl_log(1e-4);
    unsigned long long* const counter = io.counter;
    const int64_t count = io.count;
    const int ql = (int)(__lane_id() & 3);
    const int lead = (int)(__lane_id() & ~3);

    double x[N], v[N], p[P], k1[N];
    double sum = 0.0, comp = 0.0, lmant = 1.0;
    double tri = 0.0, trs = 0.0, trc = 0.0;  // TRACE: tr J integral (interval, Kahan sum)
    int32_t lex = 0;
    int lcnt = 0;
    float pf[P];
    double t = 0.0, h = 0.0, t_next = 0.0;
    int64_t m = 0, step_in = 0, item = -1;
    uint32_t nacc = 0, nrej = 0;
    Ctrl ctrl{log_floor, false};
    bool active = false, exhausted = false, pending = false;

    auto write_out = [&](uint32_t st) {
        const auto dp = getd();
        const auto iop = getio();
        const double t_avg = ADAPT ? (double)dp->n_renorm_average * dp->t_renorm
                                   : (double)dp->n_renorm_average * (double)dp->steps_per_renorm * dp->h;
        double s = sum;
        if constexpr (TRACE) {
            // the closing exponent: the trace integral minus lanes 0 .. K-1's sums
            double closing = trs;
            closing = closing - quad_bcast<0>(sum);
            if constexpr (K > 1) closing = closing - quad_bcast<1>(sum);
            if constexpr (K > 2) closing = closing - quad_bcast<2>(sum);
            if (ql == K) s = closing;
        }
        if (ql < dp->n_exps)
            iop->lambda[item * dp->n_exps + ql] = (st == LYAP_PT_OK) ? s / t_avg : __builtin_nan("");
        if (ql == 0) {
            if (iop->status) iop->status[item] = st;
            if (iop->n_acc) iop->n_acc[item] = nacc;
            if (iop->n_rej) iop->n_rej[item] = nrej;
        }
    };

    for (;;) {
        const bool want = !active && !exhausted;
        int64_t got = grab_items(want && ql == 0, counter);
        got = __shfl(got, lead);
        if (__builtin_expect(want, 0)) {
            if (got < count) {
                const auto dp = getd();
                const auto iop = getio();
                item = item_position(got, iop->stride, count);
                const int64_t idx = iop->idx_list ? iop->idx_list[item] : iop->first + item;
                point_params<P>(*dp, idx, iop->params, p);
#pragma unroll
                for (int j = 0; j < P; ++j) pf[j] = (float)p[j];
                const double* const inits = iop->inits;
#pragma unroll
                for (int i = 0; i < N; ++i) {
                    x[i] = inits ? inits[idx * N + i] : init_component(*dp, idx, i, p);
                    v[i] = (i == ql) ? 1.0 : 0.0;
                }
                sum = comp = 0.0;
                tri = trs = trc = 0.0;
                lmant = 1.0;
                lex = 0;
                lcnt = 0;
                t = 0.0;
                h = dp->h;
                m = 0;
                t_next = t_renorm;
                step_in = 0;
                nacc = nrej = 0;
                ctrl.log_fac_old = log_floor;
                ctrl.rejected = false;
                pending = false;
                if (INTEG == LYAP_INT_DOPRI5 && ADAPT) M::rhs(x, p, k1);
                active = true;
            } else {
                exhausted = true;
            }
        }
        if (!__any(active)) break;
        if (__any(active && pending)) {
            if (active && pending) {
                double V[K][N], nrm[K];
#pragma unroll
                for (int i = 0; i < N; ++i) {
                    V[0][i] = quad_bcast<0>(v[i]);
                    if constexpr (K > 1) V[1][i] = quad_bcast<1>(v[i]);
                    if constexpr (K > 2) V[2][i] = quad_bcast<2>(v[i]);
                    if constexpr (K > 3) V[3][i] = quad_bcast<3>(v[i]);
                }
                mgs_rows<K, N>(V, nrm);
                double own = nrm[0];
#pragma unroll
                for (int r = 1; r < K; ++r) own = (ql == r) ? nrm[r] : own;
#pragma unroll
                for (int i = 0; i < N; ++i) {
                    double w = V[0][i];
#pragma unroll
                    for (int r = 1; r < K; ++r) w = (ql == r) ? V[r][i] : w;
                    v[i] = w;
                }
                ++m;
                t_next = (double)(m + 1) * t_renorm;
                if (m > n_tr) {
                    lognorm_push(lmant, lex, own);
                    if (++lcnt == kLogWindow || m == m_total) {
                        lognorm_flush(lmant, lex, sum, comp);
                        lcnt = 0;
                    }
                    if constexpr (TRACE) kahan_add(trs, trc, tri);
                }
                if constexpr (TRACE) tri = 0.0;
                pending = false;
                if (m == m_total) {
                    write_out(LYAP_PT_OK);
                    active = false;
                }
            }
        }
        if (!active || pending) continue;

        uint32_t st = LYAP_PT_OK;
        bool finished = false;
        double xn[N];
        typename M::template J<TX> X[S];
        if constexpr (ADAPT) {
            if ((int64_t)nacc + (int64_t)nrej >= max_steps) {
                st = LYAP_PT_MAXSTEPS;
                finished = true;
            } else {
                double h_try = h;
                bool clipped = false;
                const double jn = M::jnorm(x, p);
                if (__any(h_try * jn > 1.0)) {
                    double jv = jn;
                    asm volatile("" : "+v"(jv));
                    if (h_try * jn > 1.0) h_try = 1.0 / jv;
                }
                if (t_next - t <= 1.0625 * h_try) {
                    h_try = t_next - t;
                    clipped = true;
                }
                if (h_try < h_min) {
                    st = LYAP_PT_STEP_UNDERFLOW;
                    finished = true;
                } else {
                    double ev[N], k1_next[N], tr_step = 0.0;
#pragma unroll
                    for (int i = 0; i < N; ++i) k1_next[i] = k1[i];
                    base_step<M, INTEG, true, TX, TRACE>(x, p, h_try, xn, ev, X, k1_next, &tr_step);
                    const double e2 = err2_seq<N>(ev, x, xn, rtol, atol);
                    if (!(e2 < kInf)) {
                        st = LYAP_PT_NONFINITE;
                        finished = true;
                    } else {
                        double h_next;
                        const double lerr = 0.5 * ctrl_log(e2);
                        const bool acc = ctrl_update<INTEG>(ctrl, e2 <= 1.0, e2 > 1e-8, lerr, log_floor,
                                                            h_try, h_max, h_next);
                        h = h_next;
                        if (!acc) {
                            ++nrej;
                        } else {
                            if (ql < K) tangent_step<M, INTEG, MIXED, TX, kLowTangent<INTEG, MIXED>>(X, p, pf, h_try, v);
                            if constexpr (TRACE) tri = tri + tr_step;
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
                                pending = true;
                            }
                        }
                    }
                }
            }
        } else {
            double tr_step = 0.0;
            base_step<M, INTEG, false, TX, TRACE>(x, p, h_fixed, xn, nullptr, X, nullptr, &tr_step);
            if (ql < K) tangent_step<M, INTEG, MIXED, TX>(X, p, pf, h_fixed, v);
            if constexpr (TRACE) tri = tri + tr_step;
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
                    pending = true;
                }
            }
        }
        if (__builtin_expect(finished, 0)) {
            write_out(st);
            active = false;
        }
    }
}

// Persistent per-lane loop of the sweep's method for maps: per iteration
// V <- J(x) V (binary32 products in the mixed form), x <- f(x); MGS and the
// windowed log-norm accumulation every steps_per_renorm iterations; the escape test after
// every iteration (a diverging map overflows within a few iterations).
template <class M, bool MIXED, class GetDesc>
LYAP_D void map_drain(const lyap_sweep_desc& d, GetDesc getd, const SweepIO& io) {
    constexpr int N = M::N, P = M::P, K = M::N;
    using TX = typename std::conditional<MIXED, float, double>::type;
    const int64_t n_tr = d.n_renorm_transient;
    const int64_t m_total = d.n_renorm_transient + d.n_renorm_average;
    const int64_t spr = d.steps_per_renorm;
    const double esc = d.escape_radius;
    const double t_avg = (double)d.n_renorm_average * (double)d.steps_per_renorm * d.h;
    const int n_exps = d.n_exps;

    double x[N], V[K][N], sum[K], comp[K], lmant[K], p[P];
    int32_t lex[K];
    int lcnt = 0;
    float pf[P];
    int64_t m = 0, step_in = 0, item = -1;
    uint32_t nacc = 0;
    bool active = false, exhausted = false;

    for (;;) {
        const int64_t got = grab_items(!active && !exhausted, io.counter);
        if (__builtin_expect(!active && !exhausted, 0)) {
            if (got < io.count) {
                const auto dp = getd();
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
                    sum[r] = comp[r] = 0.0;
                    lmant[r] = 1.0;
                    lex[r] = 0;
#pragma unroll
                    for (int i = 0; i < N; ++i) V[r][i] = (r == i) ? 1.0 : 0.0;
                }
                lcnt = 0;
                m = step_in = 0;
                nacc = 0;
                active = true;
            } else {
                exhausted = true;
            }
        }
        if (!__any(active)) break;
        if (!active) continue;

        uint32_t st = LYAP_PT_OK;
        bool finished = false;
        {
            const typename M::template J<TX> jx = M::template jac<TX>(x, p);
#pragma unroll
            for (int r = 0; r < K; ++r) {
                TX v[N], o[N];
#pragma unroll
                for (int i = 0; i < N; ++i) v[i] = (TX)V[r][i];
                if constexpr (MIXED) M::template jvp<float>(jx, pf, v, o);
                else M::template jvp<double>(jx, p, v, o);
#pragma unroll
                for (int i = 0; i < N; ++i) V[r][i] = (double)o[i];
            }
            double xn[N];
            M::rhs(x, p, xn);
#pragma unroll
            for (int i = 0; i < N; ++i) x[i] = xn[i];
            ++nacc;
        }
        if (escaped<N>(x, esc)) {
            st = LYAP_PT_ESCAPED;
            finished = true;
        } else if (++step_in == spr) {
            step_in = 0;
            double nrm[K];
            mgs_rows<K, N>(V, nrm);
            ++m;
            if (m > n_tr) {
#pragma unroll
                for (int r = 0; r < K; ++r) lognorm_push(lmant[r], lex[r], nrm[r]);
                if (++lcnt == kLogWindow || m == m_total) {
#pragma unroll
                    for (int r = 0; r < K; ++r) lognorm_flush(lmant[r], lex[r], sum[r], comp[r]);
                    lcnt = 0;
                }
            }
            if (m == m_total) finished = true;
        }
        if (__builtin_expect(finished, 0)) {
            for (int r = 0; r < n_exps; ++r) {
                double s = 0.0;
#pragma unroll
                for (int q = 0; q < K; ++q)
                    if (q == r) s = sum[q];
                io.lambda[item * n_exps + r] = (st == LYAP_PT_OK) ? s / t_avg : __builtin_nan("");
            }
            if (io.status) io.status[item] = st;
            if (io.n_acc) io.n_acc[item] = nacc;
            if (io.n_rej) io.n_rej[item] = 0;
            active = false;
        }
    }
}

}  // namespace lyap
