// Lorenz-96 device code shared by the block-per-point sweep kernel
// (lyap_l96.hip) and the wave-level drain used by the batched launch.
//
// Wave-level drain layout: a wave holds P points ("slots") on LP = 64/P lanes
// each, as G = LP/N groups of N lanes (WaveMap: consecutive lanes for N = 16,
// two groups interleaved per 16-lane row for N = 8, and for N = 32 a row pair
// whose first row holds the even components and second row the odd ones).  Lane (group vg of the
// slot, component i) owns component i of the base state and of E = N/G
// deviation vectors, r = vg + G*m, m < E; u = vg*N + i is its position in the
// slot.  The slot's base state is replicated in each of its G groups (G = 1 at
// P = 64/N: no replication).
//   - neighbour couplings are moves inside the N-lane group (Ring: DPP row
//     rotations for N = 8 and 16; for N = 32, the other row of the pair
//     through v_permlane16_swap and row-masked DPP rotations);
//   - the error norm and the Jacobian-norm cap are butterflies over the group
//     (Ring::sum / Ring::max), bit-identical in every lane of the group;
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

LYAP_D bool wave_uniform(bool b) { return __builtin_amdgcn_readfirstlane((int)b) != 0; }

// row rotation by DPP: lane i of each 16-lane row receives lane (i - R) mod 16.
// Every lane has a valid source, so the move needs no "old" value (mov_dpp).
template <int R>
LYAP_D float dpp_ror(float v) {
    return __builtin_bit_cast(float, __builtin_amdgcn_mov_dpp(__builtin_bit_cast(int, v), 0x120 + R, 0xF, 0xF,
                                                              true));
}
template <int R>
LYAP_D double dpp_ror(double v) {
    const uint64_t b = __builtin_bit_cast(uint64_t, v);
    const uint32_t lo = (uint32_t)__builtin_amdgcn_mov_dpp((int)(uint32_t)b, 0x120 + R, 0xF, 0xF, true);
    const uint32_t hi = (uint32_t)__builtin_amdgcn_mov_dpp((int)(uint32_t)(b >> 32), 0x120 + R, 0xF, 0xF, true);
    return __builtin_bit_cast(double, ((uint64_t)hi << 32) | (uint64_t)lo);
}

// the value of lane l ^ 16 (the other 16-lane row of each row pair), by
// gfx950's v_permlane16_swap: with both operands v it returns (A, B) where row
// r of A is row r & ~1 of v and row r of B is row r | 1 of v, so an odd row
// takes A and an even row B.  Every lane of the wave's row pairs takes part.
LYAP_D uint32_t row_partner_u32(uint32_t v, bool odd_row) {
    const auto r = __builtin_amdgcn_permlane16_swap(v, v, false, false);
    return odd_row ? r[0] : r[1];
}
LYAP_D float row_partner(float v, bool odd_row) {
    return __builtin_bit_cast(float, row_partner_u32(__builtin_bit_cast(uint32_t, v), odd_row));
}
LYAP_D double row_partner(double v, bool odd_row) {
    const uint64_t b = __builtin_bit_cast(uint64_t, v);
    const uint32_t lo = row_partner_u32((uint32_t)b, odd_row);
    const uint32_t hi = row_partner_u32((uint32_t)(b >> 32), odd_row);
    return __builtin_bit_cast(double, ((uint64_t)hi << 32) | (uint64_t)lo);
}

// row shift by DPP that keeps `old` where the source lane falls off the row:
// lane j of each row receives lane j - R (R > 0: row_shr) or j + |R| (R < 0:
// row_shl) of the same row, and the lanes with no such source keep old
template <int R>
LYAP_D uint32_t dpp_shift_u32(uint32_t old, uint32_t v) {
    constexpr int ctrl = R > 0 ? 0x110 + R : 0x100 - R;
    return (uint32_t)__builtin_amdgcn_update_dpp((int)old, (int)v, ctrl, 0xF, 0xF, false);
}
template <int R>
LYAP_D float dpp_shift(float old, float v) {
    return __builtin_bit_cast(float, dpp_shift_u32<R>(__builtin_bit_cast(uint32_t, old), __builtin_bit_cast(uint32_t, v)));
}
template <int R>
LYAP_D double dpp_shift(double old, double v) {
    const uint64_t o = __builtin_bit_cast(uint64_t, old), b = __builtin_bit_cast(uint64_t, v);
    const uint32_t lo = dpp_shift_u32<R>((uint32_t)o, (uint32_t)b);
    const uint32_t hi = dpp_shift_u32<R>((uint32_t)(o >> 32), (uint32_t)(b >> 32));
    return __builtin_bit_cast(double, ((uint64_t)hi << 32) | (uint64_t)lo);
}

// acc += (lane i - R of v within its 16-lane row) * w: one v_fmac_f32 with the
// row rotation as its DPP source (the compiler folds a rotation into a VOP2
// multiply but not into an fma).  The s_nop gives the two wait states a DPP
// read of a VGPR written by the previous VALU instruction needs.
template <int R>
LYAP_D void fmac_ror_f32(float& acc, float v, float w) {
    static_assert(R >= 1 && R <= 15, "row rotation");
    __asm__("s_nop 1\n\tv_fmac_f32_dpp %0, %1, %2 row_ror:%3 row_mask:0xf bank_mask:0xf bound_ctrl:1"
            : "+v"(acc)
            : "v"(v), "v"(w), "i"(R));
}

// Rotation by R lanes within the 16-lane rows selected by ROWS (bit r: row r);
// the other rows keep `old`.
template <int R, int ROWS>
LYAP_D uint32_t dpp_ror_rows_u32(uint32_t old, uint32_t v) {
    return (uint32_t)__builtin_amdgcn_update_dpp((int)old, (int)v, 0x120 + R, ROWS, 0xF, false);
}
// Neighbours of component i = 2m + r on the interleaved row pair (row r of the
// pair holds the components of parity r, lane m of the row component 2m + r):
// one v_permlane16_swap of v gives the pair's even and odd components in both
// rows, and i+1 is the odd row's value (even rows) or the even row's rotated by
// one lane (odd rows), i-1 the odd row's rotated (even rows) or the even row's
// value (odd rows), each one row-masked DPP move; i-2 is the own row rotated.
LYAP_D void pair_il_moves_u32(uint32_t v, uint32_t& p1, uint32_t& m1, uint32_t& m2) {
    const auto r = __builtin_amdgcn_permlane16_swap(v, v, false, false);
    const uint32_t ev = r[0], od = r[1];
    p1 = dpp_ror_rows_u32<15, 0xA>(od, ev);
    m1 = dpp_ror_rows_u32<1, 0x5>(ev, od);
    m2 = (uint32_t)__builtin_amdgcn_mov_dpp((int)v, 0x121, 0xF, 0xF, true);
}

// Couplings x_{i+1}, x_{i-1}, x_{i-2} of component i inside its N-lane group
// (periodic).  CS is the lane stride between consecutive components: 1 for
// groups of consecutive lanes, 2 for two N = 8 groups interleaved in one
// 16-lane row.  Groups that are whole DPP rows (N = 16, CS = 1; N = 8, CS = 2)
// move by row rotations on the VALU.  Groups of 32 consecutive lanes (N = 32)
// span a row pair: a component's neighbours are the row rotation of its own
// row, except at the row's edge lanes, which take the rotation of the other
// row of the pair (v_permlane16_swap) — all on the VALU.  Other groups move by
// lane shuffles (ds_bpermute).  IL (N = 32): the group's components are
// interleaved over its row pair, even components in the first row and odd
// ones in the second (WaveMap), so that the neighbours are the other row's
// value or a rotation of it (pair_il_moves_u32): four moves where the
// consecutive layout takes eight.  Every lane of a group is active whenever
// one is (the callers' control is group-uniform).
template <int N, int CS = 1, bool IL = false>
struct Ring {
    static constexpr bool kIL = IL && N == 32 && CS == 1;
    static constexpr bool kDpp = N * CS == 16;
    static constexpr bool kPair = N == 32 && CS == 1 && !kIL;
    int src_p1, src_m1, src_m2;
    bool odd_row;  // rows 1 and 3 of the wave
    LYAP_D explicit Ring(int lane) {
        static_assert(CS == 1 || (CS == 2 && N == 8), "component stride");
        const int i = CS == 1 ? lane % N : (lane & 15) / CS;
        const int base = lane - i * CS;
        src_p1 = base + ((i + 1) % N) * CS;
        src_m1 = base + ((i + N - 1) % N) * CS;
        src_m2 = base + ((i + N - 2) % N) * CS;
        odd_row = ((lane >> 4) & 1) != 0;
    }
    // the three neighbours at once (a row pair's edge lanes share one
    // v_permlane16_swap of v): the row shift fills the inner lanes and the
    // edge lanes keep the other row's value rotated into their place
    template <class T>
    LYAP_D void moves(T v, T& vp1, T& vm1, T& vm2) const {
        if constexpr (kIL) {
            if constexpr (sizeof(T) == 4) {
                uint32_t a, b, c;
                pair_il_moves_u32(__builtin_bit_cast(uint32_t, v), a, b, c);
                vp1 = __builtin_bit_cast(T, a);
                vm1 = __builtin_bit_cast(T, b);
                vm2 = __builtin_bit_cast(T, c);
            } else {
                const uint64_t w = __builtin_bit_cast(uint64_t, v);
                uint32_t a0, b0, c0, a1, b1, c1;
                pair_il_moves_u32((uint32_t)w, a0, b0, c0);
                pair_il_moves_u32((uint32_t)(w >> 32), a1, b1, c1);
                vp1 = __builtin_bit_cast(T, ((uint64_t)a1 << 32) | (uint64_t)a0);
                vm1 = __builtin_bit_cast(T, ((uint64_t)b1 << 32) | (uint64_t)b0);
                vm2 = __builtin_bit_cast(T, ((uint64_t)c1 << 32) | (uint64_t)c0);
            }
        } else if constexpr (kPair) {
            const T pv = row_partner(v, odd_row);
            vp1 = dpp_shift<-1>(dpp_ror<15>(pv), v);
            vm1 = dpp_shift<1>(dpp_ror<1>(pv), v);
            vm2 = dpp_shift<2>(dpp_ror<2>(pv), v);
        } else {
            vp1 = p1(v);
            vm1 = m1(v);
            vm2 = m2(v);
        }
    }
    template <class T>
    LYAP_D T p1(T v) const {
        if constexpr (kIL) {
            T a, b, c;
            moves(v, a, b, c);
            return a;
        } else if constexpr (kDpp) return dpp_ror<16 - CS>(v);
        else if constexpr (kPair) return dpp_shift<-1>(dpp_ror<15>(row_partner(v, odd_row)), v);
        else return __shfl(v, src_p1);
    }
    template <class T>
    LYAP_D T m1(T v) const {
        if constexpr (kIL) {
            T a, b, c;
            moves(v, a, b, c);
            return b;
        } else if constexpr (kDpp) return dpp_ror<CS>(v);
        else if constexpr (kPair) return dpp_shift<1>(dpp_ror<1>(row_partner(v, odd_row)), v);
        else return __shfl(v, src_m1);
    }
    template <class T>
    LYAP_D T m2(T v) const {
        if constexpr (kIL) return dpp_ror<1>(v);
        else if constexpr (kDpp) return dpp_ror<2 * CS>(v);
        else if constexpr (kPair) return dpp_shift<2>(dpp_ror<2>(row_partner(v, odd_row)), v);
        else return __shfl(v, src_m2);
    }
    // Butterfly sum / max over the group, identical in every lane and equal
    // bit for bit to the xor butterfly with component offsets N/2, ..., 1:
    // after the level with offset o the partial results repeat with period o
    // (in components), so the value o components away by rotation is the xor
    // partner's value, and a + b == b + a.
    LYAP_D double sum(double v) const {
        if constexpr (kIL) {
            // component offsets 16, 8, 4, 2 are lane offsets 8, 4, 2, 1 of the
            // row; offset 1 is the other row of the pair
            v = v + dpp_ror<8>(v);
            v = v + dpp_ror<4>(v);
            v = v + dpp_ror<2>(v);
            v = v + dpp_ror<1>(v);
            v = v + row_partner(v, odd_row);
            return v;
        } else if constexpr (kDpp || kPair) {
            if constexpr (kPair) v = v + row_partner(v, odd_row);  // offset 16: the other row
            if constexpr (N >= 16) v = v + dpp_ror<8 * CS>(v);
            v = v + dpp_ror<4 * CS>(v);
            v = v + dpp_ror<2 * CS>(v);
            v = v + dpp_ror<1 * CS>(v);
            return v;
        } else {
#pragma unroll
            for (int o = N / 2; o >= 1; o >>= 1) v = v + __shfl_xor(v, o * CS);
            return v;
        }
    }
    LYAP_D double max(double v) const {
        auto mx = [](double a, double b) { return a > b ? a : b; };
        if constexpr (kIL) {
            v = mx(v, dpp_ror<8>(v));
            v = mx(v, dpp_ror<4>(v));
            v = mx(v, dpp_ror<2>(v));
            v = mx(v, dpp_ror<1>(v));
            v = mx(v, row_partner(v, odd_row));
            return v;
        } else if constexpr (kDpp || kPair) {
            if constexpr (kPair) v = mx(v, row_partner(v, odd_row));
            if constexpr (N >= 16) v = mx(v, dpp_ror<8 * CS>(v));
            v = mx(v, dpp_ror<4 * CS>(v));
            v = mx(v, dpp_ror<2 * CS>(v));
            v = mx(v, dpp_ror<1 * CS>(v));
            return v;
        } else {
#pragma unroll
            for (int o = N / 2; o >= 1; o >>= 1) v = mx(v, __shfl_xor(v, o * CS));
            return v;
        }
    }
};

// Lanes of the wave-level drain: lane -> (group q, component c).  N >= 16:
// group q holds lanes q*N .. q*N+N-1; with IL and N = 32 component c sits at
// lane q*32 + 16*(c & 1) + c/2 (even components in the first row of the pair,
// odd ones in the second).  N = 8: the two groups of each 16-lane row are
// interleaved (lane = 16*row + 2*c + h, q = 2*row + h), so that the component
// rotations above are row rotations.
template <int N, bool IL = false>
struct WaveMap {
    static constexpr int CS = N >= 16 ? 1 : 16 / N;
    static constexpr bool kIL = IL && N == 32;
    LYAP_D static int comp(int lane) {
        if constexpr (kIL) return 2 * (lane & 15) + ((lane >> 4) & 1);
        else return CS == 1 ? lane % N : (lane & 15) / CS;
    }
    LYAP_D static int group(int lane) { return CS == 1 ? lane / N : (lane >> 4) * CS + (lane & 15) % CS; }
    LYAP_D static int lane_of(int q, int c) {
        if constexpr (kIL) return q * N + (c & 1) * 16 + (c >> 1);
        else return CS == 1 ? q * N + c : 16 * (q / CS) + c * CS + q % CS;
    }
    // ballot bits of the lanes of group q
    LYAP_D static uint64_t group_bits(int q) {
        if constexpr (CS == 1) return (N == 64 ? ~0ULL : ((1ULL << N) - 1ULL)) << (q * N);
        else return 0x5555ULL << (16 * (q / CS) + q % CS);
    }
};

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

    using Map = WaveMap<N, true>;  // N = 32: components interleaved over each row pair
    constexpr int CS = Map::CS;
    const int lane = (int)__lane_id();
    const int q = Map::group(lane);   // N-lane group of the wave
    const int i = Map::comp(lane);    // component
    const int slot = q / G;           // point slot
    const int vg = q % G;             // vector group within the slot
    const int u = vg * N + i;         // position within the slot
    const int lead = Map::lane_of(slot * G, 0);  // the slot's first lane (u = 0)
    uint64_t slot_bits = 0;
#pragma unroll
    for (int g = 0; g < G; ++g) slot_bits |= Map::group_bits(slot * G + g);
    const Ring<N, CS, true> ring(lane);
    constexpr int RS = N + 1;  // LDS row stride (doubles)
    double* sV = lds + slot * (N * RS);

    // a decision of the slot: the value its first lane computed
    auto slot_uniform = [lead](bool b) { return __shfl((int)b, lead) != 0; };
    auto slot_any = [slot_bits](bool b) { return (__ballot(b) & slot_bits) != 0ULL; };

    const int64_t n_tr = d.n_renorm_transient;
    const int64_t m_total = d.n_renorm_transient + d.n_renorm_average;
    const int64_t max_steps = d.max_steps, spr = d.steps_per_renorm;
    const double rtol = d.rtol, atol = d.atol, h_min = d.h_min, h_max = d.h_max;
    const double t_renorm = d.t_renorm, esc = d.escape_radius, h_fixed = d.h;
    const double t_avg = ADAPT ? (double)d.n_renorm_average * d.t_renorm
                               : (double)d.n_renorm_average * (double)d.steps_per_renorm * d.h;
    const int n_exps = d.n_exps;
    const double log_floor = ctrl_log(1e-4);

    double x = 0.0, V[E], sum = 0.0, comp = 0.0, t = 0.0, h = 0.0, k1 = 0.0, F = 0.0, dd = 0.0;
    double lmant = 1.0;  // open log window of row u's norms (lanes u < K)
    int32_t lex = 0;
    int lcnt = 0;
    T dT = T(0);
    int64_t mr = 0, step_in = 0, pos = 0;
    uint32_t nacc = 0, nrej = 0;
    Ctrl ctrl{log_floor, false};
    bool active = false, exhausted = false;
#pragma unroll
    for (int m = 0; m < E; ++m) V[m] = 0.0;

    for (;;) {
        const bool want = !active && !exhausted;
        int64_t got = grab_items(want && u == 0, io.counter);
        got = __shfl(got, lead);
        // taking and finishing a point are rare next to a step (branch hints
        // keep the register allocator's spill code out of the step loop)
        if (__builtin_expect(want, 0)) {
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
                lmant = 1.0;
                lex = 0;
                lcnt = 0;
                t = 0.0;
                h = dp->h;
                mr = step_in = 0;
                nacc = nrej = 0;
                ctrl.log_fac_old = log_floor;
                ctrl.rejected = false;
                if (INTEG == LYAP_INT_DOPRI5 && ADAPT) {
                    double xp1, xm1, xm2;
                    ring.moves(x, xp1, xm1, xm2);
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
                double xp1, xm1, xm2;
                ring.moves(x, xp1, xm1, xm2);
                const double row = (2.0 * fabs(xm1) + fabs(xp1 - xm2)) + fabs(dd);
                const double jn = ring.max(row);
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
                const double qe = ev / sk;
                const double err = sqrt(ring.sum(qe * qe) / (double)N);
                if (slot_uniform(!(err < kInf))) {
                    st = LYAP_PT_NONFINITE;
                    finished = true;
                    accept = false;
                } else {
                    double h_next;
                    accept = ctrl_update<INTEG>(ctrl, err <= 1.0, err > 1e-4, ctrl_log(err), log_floor,
                                                h_try, h_max, h_next);
                    accept = slot_uniform(accept);
                    h = h_next;
                    if (!accept) ++nrej;
                }
            }
            if (accept) {
                // tangent step of the lane's E vectors, J at every stage state;
                // NV vectors advance through the stages together so that one
                // vector's lane moves fill the others' wait states
                // mixed form: h folded into the Jacobian data (jm1, jd and
                // dh = h d), so the stage arguments are v + sum_j a_sj K'_j and
                // the products K'_s = h K_s; the increment is sum_j b_j K'_j
                const T hT = (T)h_try;
                const T dh = MIXED ? (T)(h_try * dd) : dT;
                const RowB<TB, T, S> rbt;
                constexpr int NV = E % 2 == 0 ? 2 : 1;
#pragma unroll
                for (int m0 = 0; m0 < E; m0 += NV) {
                    T v0[NV], kv[NV][S];
#pragma unroll
                    for (int q2 = 0; q2 < NV; ++q2) v0[q2] = (T)V[m0 + q2];
#pragma unroll
                    for (int s = 0; s < S; ++s) {
                        if (!TB::for_solution(s)) {
#pragma unroll
                            for (int q2 = 0; q2 < NV; ++q2) kv[q2][s] = T(0);
                            continue;
                        }
                        const RowA<TB, T, S> row(s);
                        T varg[NV];
#pragma unroll
                        for (int q2 = 0; q2 < NV; ++q2) {
                            if constexpr (MIXED) {
                                T acc = v0[q2];
#pragma unroll
                                for (int j = 0; j < S; ++j) {
                                    if (j >= s || TB::a(s, j) == 0.0) continue;
                                    acc = fma(row.c[j], kv[q2][j], acc);
                                }
                                varg[q2] = acc;
                            } else {
                                T acc = T(0);
                                bool any = false;
#pragma unroll
                                for (int j = 0; j < S; ++j) {
                                    if (j >= s || TB::a(s, j) == 0.0) continue;
                                    acc = any ? fma(row.c[j], kv[q2][j], acc) : row.c[j] * kv[q2][j];
                                    any = true;
                                }
                                varg[q2] = any ? fma(hT, acc, v0[q2]) : v0[q2];
                            }
                        }
#pragma unroll
                        for (int q2 = 0; q2 < NV; ++q2) {
                            T vp1, vm1, vm2;
                            if constexpr (MIXED && Ring<N, CS, true>::kDpp) {
                                // v(i+1) - v(i-2) is d(i+1) for d(j) = v(j) - v(j-3): one DPP
                                // subtraction, and its move to lane i rides on the
                                // multiply-add (the same operations as the general form)
                                const T dlt = varg[q2] - dpp_ror<3 * CS>(varg[q2]);
                                vm1 = dpp_ror<CS>(varg[q2]);
                                T acc = fma(-dh, varg[q2], jd[s] * vm1);
                                fmac_ror_f32<16 - CS>(acc, dlt, jm1[s]);
                                kv[q2][s] = acc;
                                continue;
                            }
                            ring.moves(varg[q2], vp1, vm1, vm2);
                            // binary32: the jd*v(i-1) product first, so its lane move folds
                            // into a DPP multiply (the FP64 form keeps the fused order)
                            if constexpr (MIXED)
                                kv[q2][s] = fma(vp1 - vm2, jm1[s], fma(-dh, varg[q2], jd[s] * vm1));
                            else
                                kv[q2][s] = fma(vp1 - vm2, jm1[s], fma(jd[s], vm1, -dT * varg[q2]));
                        }
                    }
#pragma unroll
                    for (int q2 = 0; q2 < NV; ++q2) {
                        T accv = T(0);
                        bool anyv = false;
#pragma unroll
                        for (int j = 0; j < S; ++j) {
                            if (TB::b(j) == 0.0) continue;
                            accv = anyv ? fma(rbt.c[j], kv[q2][j], accv) : rbt.c[j] * kv[q2][j];
                            anyv = true;
                        }
                        if constexpr (MIXED) V[m0 + q2] = V[m0 + q2] + (double)accv;
                        else V[m0 + q2] = fma(h_try, accv, V[m0 + q2]);
                    }
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
                    double vn = 0.0;
                    if constexpr (G == 1 && N <= 16) {
                        // one vector group per point (E = N <= 16): lane r keeps row r
                        // in the registers its E vectors leave free while the
                        // matrix is staged, and writes it to LDS once,
                        // normalised, for the later rows
                        double rw[N];
                        if (u < K) {
#pragma unroll
                            for (int c = 0; c < N; ++c) rw[c] = row[c];
                        }
                        for (int j = 0; j < K; ++j) {
                            if (u == j) {
                                double s2 = rw[0] * rw[0];
#pragma unroll
                                for (int c = 1; c < N; ++c) s2 = fma(rw[c], rw[c], s2);
                                const double inv = inv_sqrt_det(s2);
                                vn = s2 * inv;  // accumulated after the loop, by every row lane at once
#pragma unroll
                                for (int c = 0; c < N; ++c) {
                                    rw[c] = rw[c] * inv;
                                    row[c] = rw[c];
                                }
                            }
                            wave_lds_sync();
                            if (u > j && u < K) {
                                const double* qj = sV + j * RS;
                                double cdot = qj[0] * rw[0];
#pragma unroll
                                for (int c = 1; c < N; ++c) cdot = fma(qj[c], rw[c], cdot);
#pragma unroll
                                for (int c = 0; c < N; ++c) rw[c] = fma(-cdot, qj[c], rw[c]);
                            }
                        }
                    } else {
                        for (int j = 0; j < K; ++j) {
                            if (u == j) {
                                double s2 = row[0] * row[0];
#pragma unroll 8
                                for (int c = 1; c < N; ++c) s2 = fma(row[c], row[c], s2);
                                const double inv = inv_sqrt_det(s2);
                                vn = s2 * inv;  // accumulated after the loop, by every row lane at once
#pragma unroll 8
                                for (int c = 0; c < N; ++c) row[c] = row[c] * inv;
                            }
                            wave_lds_sync();
                            if (u > j && u < K) {
                                const double* qj = sV + j * RS;
                                double cdot = qj[0] * row[0];
#pragma unroll 8
                                for (int c = 1; c < N; ++c) cdot = fma(qj[c], row[c], cdot);
#pragma unroll 8
                                for (int c = 0; c < N; ++c) row[c] = fma(-cdot, qj[c], row[c]);
                            }
                        }
                    }
                    wave_lds_sync();
#pragma unroll
                    for (int m = 0; m < E; ++m) V[m] = sV[(vg + G * m) * RS + i];
                    ++mr;
                    if (mr > n_tr && u < K) {
                        lognorm_push(lmant, lex, vn);
                        if (++lcnt == kLogWindow || mr == m_total) {
                            lognorm_flush(lmant, lex, sum, comp);
                            lcnt = 0;
                        }
                    }
                    if (mr == m_total) finished = true;
                }
            }
        }
        if (__builtin_expect(finished, 0)) {
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
