// Device building blocks of the gfx950 Lyapunov sweep: deterministic
// elementary functions, the counter-based generator, the DVERK 6(5) and
// Dormand-Prince 5(4) tableaux, the model right-hand sides and Jacobian
// products, the step-size controller and the per-lane orthogonalisation.
//
// Arithmetic contract: compiled with -ffp-contract=off; every fused
// multiply-add is written as fma().  Stage sums run in increasing stage order
// (first term a product, then fma(a_j, k_j, acc), then fma(h, acc, y)), so
// the FP64 path is bit-reproducible against a CPU build of the same
// operations.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <type_traits>

#include "chaos_lyap.h"

#define LYAP_HD __host__ __device__ __forceinline__
#define LYAP_D __device__ __forceinline__

namespace lyap {

// ---------------------------------------------------------------------------
// deterministic log / exp / pow (no libm; identical bits on host and device)
// ---------------------------------------------------------------------------
LYAP_HD uint64_t bits_of(double x) { return __builtin_bit_cast(uint64_t, x); }
LYAP_HD double double_of(uint64_t u) { return __builtin_bit_cast(double, u); }

// An FP64 coefficient materialised where it is used.  gfx9 VALU encodings
// take no 64-bit literal, so FP64 constants live in SGPR pairs; left alone,
// the compiler hoists all of them out of the step loop, overflows the SGPR
// budget and spills them into VGPR lanes (a v_readlane per use).  Passing the
// value through an empty asm keeps it a fresh s_mov pair at its use site.  The
// two halves go through separate 32-bit asm operands: with one 64-bit operand
// the register coalescer merges halves of different constants and spills the
// pairs anyway (about 5x more v_readlane/v_writelane in the lane kernels).
LYAP_HD double sconst(double v) {
#if defined(__HIP_DEVICE_COMPILE__)
    const uint64_t b = __builtin_bit_cast(uint64_t, v);
    uint32_t lo = (uint32_t)b, hi = (uint32_t)(b >> 32);
    __asm__ volatile("" : "+s"(lo));
    __asm__ volatile("" : "+s"(hi));
    v = __builtin_bit_cast(double, ((uint64_t)hi << 32) | lo);
#endif
    return v;
}
LYAP_HD float sconst(float v) { return v; }

// fma(a, b, c) for a binary64 constant c: one v_fma_f64 with c in an SGPR pair
// materialised at the use site.  Left to itself the compiler copies c into
// the destination VGPRs and issues v_fmac_f64: three VALU instructions for
// one.  The same rounding as fma(), so host and device agree bit for bit.
LYAP_HD double fma_c(double a, double b, double c) {
#if defined(__HIP_DEVICE_COMPILE__)
    const double cs = sconst(c);
    double r;
    __asm__("v_fma_f64 %0, %1, %2, %3" : "=v"(r) : "v"(a), "v"(b), "s"(cs));
    return r;
#else
    return fma(a, b, c);
#endif
}

constexpr double kLn2Hi = 6.93147180369123816490e-01;
constexpr double kLn2Lo = 1.90821492927058770002e-10;
constexpr double kInf = __builtin_huge_val();

// log(1+f) = f - s*(f - R), s = f/(2+f), R = 2*sum_{j>=1} s^(2j)/(2j+1)
LYAP_HD double det_log(double x) {
    if (x != x) return x;
    if (x < 0.0) return __builtin_nan("");
    if (x == 0.0) return -kInf;
    if (x == kInf) return kInf;
    uint64_t b = bits_of(x);
    int e = (int)((b >> 52) & 0x7ff);
    int k;
    if (e == 0) {
        x = x * 18014398509481984.0;
        b = bits_of(x);
        e = (int)((b >> 52) & 0x7ff);
        k = e - 1023 - 54;
    } else {
        k = e - 1023;
    }
    double m = double_of((b & 0x000fffffffffffffULL) | 0x3ff0000000000000ULL);
    if (m > 1.4142135623730951) {
        m = m * 0.5;
        k = k + 1;
    }
    const double f = m - 1.0;
    const double s = f / (2.0 + f);
    const double z = s * s;
    double r = sconst(2.0 / 25.0);
    r = fma_c(r, z, 2.0 / 23.0);
    r = fma_c(r, z, 2.0 / 21.0);
    r = fma_c(r, z, 2.0 / 19.0);
    r = fma_c(r, z, 2.0 / 17.0);
    r = fma_c(r, z, 2.0 / 15.0);
    r = fma_c(r, z, 2.0 / 13.0);
    r = fma_c(r, z, 2.0 / 11.0);
    r = fma_c(r, z, 2.0 / 9.0);
    r = fma_c(r, z, 2.0 / 7.0);
    r = fma_c(r, z, 2.0 / 5.0);
    r = fma_c(r, z, 2.0 / 3.0);
    r = r * z;
    const double lg = f - s * (f - r);
    const double dk = (double)k;
    return fma(dk, sconst(kLn2Hi), fma(dk, sconst(kLn2Lo), lg));
}

// exp: nearest-integer reduction by ln2, degree-13 Taylor polynomial
LYAP_HD double det_exp(double x) {
    if (x != x) return x;
    if (x > 709.782712893384) return kInf;
    if (x < -745.1332191019412) return 0.0;
    double kd = x * 1.4426950408889634;
    kd = (kd >= 0.0) ? floor(kd + 0.5) : ceil(kd - 0.5);
    const double r = fma(-kd, sconst(kLn2Lo), fma(-kd, sconst(kLn2Hi), x));
    double p = sconst(1.0 / 6227020800.0);
    p = fma_c(p, r, 1.0 / 479001600.0);
    p = fma_c(p, r, 1.0 / 39916800.0);
    p = fma_c(p, r, 1.0 / 3628800.0);
    p = fma_c(p, r, 1.0 / 362880.0);
    p = fma_c(p, r, 1.0 / 40320.0);
    p = fma_c(p, r, 1.0 / 5040.0);
    p = fma_c(p, r, 1.0 / 720.0);
    p = fma_c(p, r, 1.0 / 120.0);
    p = fma_c(p, r, 1.0 / 24.0);
    p = fma_c(p, r, 1.0 / 6.0);
    p = fma_c(p, r, 0.5);
    p = fma_c(p, r, 1.0);
    p = fma_c(p, r, 1.0);
    int k = (int)kd;
    if (k > 1023) {
        p = p * double_of((uint64_t)(1023 + 1023) << 52);
        k -= 1023;
    }
    if (k < -1022) {
        p = p * double_of((uint64_t)1 << 52);
        k += 1022;
    }
    return p * double_of((uint64_t)(k + 1023) << 52);
}

LYAP_HD double det_pow(double x, double y) {
    if (x == 0.0) return 0.0;
    return det_exp(y * det_log(x));
}

// ---------------------------------------------------------------------------
// counter-based generator: splitmix64 finaliser keyed on (seed, point, stream)
// ---------------------------------------------------------------------------
LYAP_HD uint64_t mix64(uint64_t z) {
    z += 0x9E3779B97F4A7C15ULL;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
}

LYAP_HD double uniform01(uint64_t seed, int64_t point, int64_t stream) {
    const uint64_t z = mix64(seed ^ mix64((uint64_t)point * 1024ULL + (uint64_t)stream));
    return (double)(z >> 11) * (1.0 / 9007199254740992.0);
}

// ---------------------------------------------------------------------------
// tableaux
// ---------------------------------------------------------------------------
template <int INTEG>
struct Tab;

// DVERK 6(5) (Verner's pair as published by Hull, Enright & Jackson); the
// reference's fixed-step form is cuda/tools/dverk.dp.cpp:11-42.
template <>
struct Tab<LYAP_INT_DVERK> {
    static constexpr int S = 8;
    static constexpr double expo = 1.0 / 6.0 - 0.75 * 0.04;
    static constexpr double a(int s, int j) {
        constexpr double A[8][8] = {
            {0},
            {1.0 / 6.0},
            {4.0 / 75.0, 16.0 / 75.0},
            {5.0 / 6.0, -8.0 / 3.0, 5.0 / 2.0},
            {-165.0 / 64.0, 55.0 / 6.0, -425.0 / 64.0, 85.0 / 96.0},
            {12.0 / 5.0, -8.0, 4015.0 / 612.0, -11.0 / 36.0, 88.0 / 255.0},
            {-8263.0 / 15000.0, 124.0 / 75.0, -643.0 / 680.0, -81.0 / 250.0, 2484.0 / 10625.0, 0.0},
            {3501.0 / 1720.0, -300.0 / 43.0, 297275.0 / 52632.0, -319.0 / 2322.0,
             24068.0 / 84065.0, 0.0, 3850.0 / 26703.0}};
        return A[s][j];
    }
    static constexpr double b(int j) {
        constexpr double B[8] = {3.0 / 40.0,      0.0, 875.0 / 2244.0,    23.0 / 72.0,
                                 264.0 / 1955.0, 0.0, 125.0 / 11592.0, 43.0 / 616.0};
        return B[j];
    }
    static constexpr double e(int j) {
        constexpr double E[8] = {-1.0 / 160.0,    0.0,          -125.0 / 17952.0, 1.0 / 144.0,
                                 -12.0 / 1955.0, -3.0 / 44.0, 125.0 / 11592.0,  43.0 / 616.0};
        return E[j];
    }
    // stage 6 feeds only the error estimate
    static constexpr bool for_solution(int s) { return s != 5; }
    // the embedded fifth-order weights b - e, which use stages 1-6 only: the
    // binary32 tangent of the mixed form with adaptive steps takes these (its
    // rounding, ~6e-8 a step, is far above the fifth-order truncation at steps
    // whose sixth-order error estimate is held near rtol)
    static constexpr int S_LOW = 6;
    static constexpr double bh(int j) {
        constexpr double BH[8] = {13.0 / 160.0, 0.0, 2375.0 / 5984.0, 5.0 / 16.0,
                                  12.0 / 85.0,  3.0 / 44.0, 0.0,    0.0};
        return BH[j];
    }
};

// Dormand-Prince 5(4), cuda/tools/dopri.dp.cpp:14-45; stage 7 is FSAL
template <>
struct Tab<LYAP_INT_DOPRI5> {
    static constexpr int S = 7;
    static constexpr double expo = 1.0 / 5.0 - 0.75 * 0.04;
    static constexpr double a(int s, int j) {
        constexpr double A[7][7] = {
            {0},
            {1.0 / 5.0},
            {3.0 / 40.0, 9.0 / 40.0},
            {44.0 / 45.0, -56.0 / 15.0, 32.0 / 9.0},
            {19372.0 / 6561.0, -25360.0 / 2187.0, 64448.0 / 6561.0, -212.0 / 729.0},
            {9017.0 / 3168.0, -355.0 / 33.0, 46732.0 / 5247.0, 49.0 / 176.0, -5103.0 / 18656.0},
            {35.0 / 384.0, 0.0, 500.0 / 1113.0, 125.0 / 192.0, -2187.0 / 6784.0, 11.0 / 84.0}};
        return A[s][j];
    }
    static constexpr double b(int j) {
        constexpr double B[7] = {35.0 / 384.0,      0.0,          500.0 / 1113.0, 125.0 / 192.0,
                                 -2187.0 / 6784.0, 11.0 / 84.0, 0.0};
        return B[j];
    }
    static constexpr double e(int j) {
        constexpr double E[7] = {71.0 / 57600.0,       0.0,           -71.0 / 16695.0, 71.0 / 1920.0,
                                 -17253.0 / 339200.0, 22.0 / 525.0, -1.0 / 40.0};
        return E[j];
    }
    static constexpr bool for_solution(int s) { return s != 6; }
};



// One tableau row (stage s) of coefficients as opaque SGPR values, shared by
// every component and vector that uses the row.
template <class TB, class T, int LD>
struct RowA {
    T c[LD];
    LYAP_D explicit RowA(int s) {
#pragma unroll
        for (int j = 0; j < LD; ++j) c[j] = (j < s && TB::a(s, j) != 0.0) ? sconst((T)TB::a(s, j)) : T(0);
    }
};
template <class TB, class T, int LD>
struct RowB {
    T c[LD];
    LYAP_D RowB() {
#pragma unroll
        for (int j = 0; j < LD; ++j) c[j] = TB::b(j) != 0.0 ? sconst((T)TB::b(j)) : T(0);
    }
};
template <class TB, class T, int LD>
struct RowBh {
    T c[LD];
    LYAP_D RowBh() {
#pragma unroll
        for (int j = 0; j < LD; ++j) c[j] = TB::bh(j) != 0.0 ? sconst((T)TB::bh(j)) : T(0);
    }
};
template <class TB, class T, int LD>
struct RowE {
    T c[LD];
    LYAP_D RowE() {
#pragma unroll
        for (int j = 0; j < LD; ++j) c[j] = TB::e(j) != 0.0 ? sconst((T)TB::e(j)) : T(0);
    }
};

// Stage sums: coefficients in increasing stage order, first nonzero term a
// product, later terms fma(c_j, k_j, acc).  The zero tests use the constexpr
// tableau, so they resolve at compile time after unrolling.
template <class TB, class T, int LD, int N>
LYAP_D T coef_sum_a(int s, const RowA<TB, T, LD>& row, const T (&k)[LD][N], int i, bool& any) {
    T acc = T(0);
    any = false;
#pragma unroll
    for (int j = 0; j < LD; ++j) {
        if (j >= s || TB::a(s, j) == 0.0) continue;
        if (!any) {
            acc = row.c[j] * k[j][i];
            any = true;
        } else {
            acc = fma(row.c[j], k[j][i], acc);
        }
    }
    return acc;
}

template <class TB, class T, int LD, int N>
LYAP_D T coef_sum_b(const RowB<TB, T, LD>& row, const T (&k)[LD][N], int i) {
    T acc = T(0);
    bool any = false;
#pragma unroll
    for (int j = 0; j < LD; ++j) {
        if (TB::b(j) == 0.0) continue;
        if (!any) {
            acc = row.c[j] * k[j][i];
            any = true;
        } else {
            acc = fma(row.c[j], k[j][i], acc);
        }
    }
    return acc;
}

template <class TB, class T, int LD, int N>
LYAP_D T coef_sum_bh(const RowBh<TB, T, LD>& row, const T (&k)[LD][N], int i) {
    T acc = T(0);
    bool any = false;
#pragma unroll
    for (int j = 0; j < LD; ++j) {
        if (TB::bh(j) == 0.0) continue;
        if (!any) {
            acc = row.c[j] * k[j][i];
            any = true;
        } else {
            acc = fma(row.c[j], k[j][i], acc);
        }
    }
    return acc;
}

template <class TB, class T, int LD, int N>
LYAP_D T coef_sum_e(const RowE<TB, T, LD>& row, const T (&k)[LD][N], int i) {
    T acc = T(0);
    bool any = false;
#pragma unroll
    for (int j = 0; j < LD; ++j) {
        if (TB::e(j) == 0.0) continue;
        if (!any) {
            acc = row.c[j] * k[j][i];
            any = true;
        } else {
            acc = fma(row.c[j], k[j][i], acc);
        }
    }
    return acc;
}

// ---------------------------------------------------------------------------
// models: f(x) and J(x) v.  T is double, or float for the mixed tangent (the
// caller rounds the state and parameters to float first).
// ---------------------------------------------------------------------------
// Each model gives f(x), its Jacobian data at a stage state and J v from that
// data.  The Jacobian data are the entries that depend on the state; entries
// that are differences of state and parameter are formed in FP64 and rounded
// once to the tangent type T (float in the mixed form), so they do not lose
// digits to cancellation in binary32.  With T = double this is exactly the
// FP64 Jacobian.  jnorm(x) = max row sum of |J(x)| caps adaptive steps at
// h*jnorm <= 1, which keeps the tangent integration accurate where the base
// error estimate vanishes (stable fixed points).
struct Shimizu {  // reference cuda/models/flows/shimizu.dp.cpp:15-27; (alpha, lambda)
    static constexpr int ID = LYAP_MODEL_SHIMIZU, N = 3, P = 2;
    static constexpr bool kMap = false;
    template <class T>
    struct J {
        T x, oz;  // x, 1 - z
    };
    LYAP_D static void rhs(const double* x, const double* p, double* o) {
        o[0] = x[1];
        o[1] = fma(-p[0], x[1], fma(-x[0], x[2], x[0]));
        o[2] = fma(x[0], x[0], -p[1] * x[2]);
    }
    template <class T>
    LYAP_D static J<T> jac(const double* x, const double* p) {
        return J<T>{(T)x[0], (T)(1.0 - x[2])};
    }
    template <class T>
    LYAP_D static void jvp(const J<T>& j, const T* p, const T* v, T* o) {
        o[0] = v[1];
        o[1] = fma(j.oz, v[0], fma(-p[0], v[1], -j.x * v[2]));
        o[2] = fma(T(2) * j.x, v[0], -p[1] * v[2]);
    }
    LYAP_D static double jnorm(const double* x, const double* p) {
        const double r0 = 1.0;
        const double r1 = (fabs(1.0 - x[2]) + fabs(p[0])) + fabs(x[0]);
        const double r2 = 2.0 * fabs(x[0]) + fabs(p[1]);
        const double m = r0 > r1 ? r0 : r1;
        return m > r2 ? m : r2;
    }
    // tr J = c(p) + g(x) (the divergence of f): -alpha - lambda
    static constexpr bool kTraceVar = false;
    LYAP_D static double trace_const(const double* p) { return -(p[0] + p[1]); }
    LYAP_D static double trace_var(const double*) { return 0.0; }
};

struct Lorenz {  // (sigma, rho, beta)
    static constexpr int ID = LYAP_MODEL_LORENZ, N = 3, P = 3;
    static constexpr bool kMap = false;
    template <class T>
    struct J {
        T x, y, rz;  // x, y, rho - z
    };
    LYAP_D static void rhs(const double* x, const double* p, double* o) {
        o[0] = p[0] * (x[1] - x[0]);
        o[1] = fma(x[0], p[1] - x[2], -x[1]);
        o[2] = fma(x[0], x[1], -p[2] * x[2]);
    }
    template <class T>
    LYAP_D static J<T> jac(const double* x, const double* p) {
        return J<T>{(T)x[0], (T)x[1], (T)(p[1] - x[2])};
    }
    template <class T>
    LYAP_D static void jvp(const J<T>& j, const T* p, const T* v, T* o) {
        o[0] = p[0] * (v[1] - v[0]);
        o[1] = fma(j.rz, v[0], -fma(j.x, v[2], v[1]));
        o[2] = fma(j.y, v[0], fma(j.x, v[1], -p[2] * v[2]));
    }
    LYAP_D static double jnorm(const double* x, const double* p) {
        const double r0 = 2.0 * fabs(p[0]);
        const double r1 = (fabs(p[1] - x[2]) + 1.0) + fabs(x[0]);
        const double r2 = (fabs(x[1]) + fabs(x[0])) + fabs(p[2]);
        const double m = r0 > r1 ? r0 : r1;
        return m > r2 ? m : r2;
    }
    // tr J = -sigma - 1 - beta
    static constexpr bool kTraceVar = false;
    LYAP_D static double trace_const(const double* p) { return -((p[0] + 1.0) + p[2]); }
    LYAP_D static double trace_var(const double*) { return 0.0; }
};

struct Rossler {  // (a, c, b)
    static constexpr int ID = LYAP_MODEL_ROSSLER, N = 3, P = 3;
    static constexpr bool kMap = false;
    template <class T>
    struct J {
        T z, xc;  // z, x - c
    };
    LYAP_D static void rhs(const double* x, const double* p, double* o) {
        o[0] = -(x[1] + x[2]);
        o[1] = fma(p[0], x[1], x[0]);
        o[2] = fma(x[2], x[0] - p[1], p[2]);
    }
    template <class T>
    LYAP_D static J<T> jac(const double* x, const double* p) {
        return J<T>{(T)x[2], (T)(x[0] - p[1])};
    }
    template <class T>
    LYAP_D static void jvp(const J<T>& j, const T* p, const T* v, T* o) {
        o[0] = -(v[1] + v[2]);
        o[1] = fma(p[0], v[1], v[0]);
        o[2] = fma(j.z, v[0], j.xc * v[2]);
    }
    LYAP_D static double jnorm(const double* x, const double* p) {
        const double r0 = 2.0;
        const double r1 = 1.0 + fabs(p[0]);
        const double r2 = fabs(x[2]) + fabs(x[0] - p[1]);
        const double m = r0 > r1 ? r0 : r1;
        return m > r2 ? m : r2;
    }
    // tr J = (a - c) + x
    static constexpr bool kTraceVar = true;
    LYAP_D static double trace_const(const double* p) { return p[0] - p[1]; }
    LYAP_D static double trace_var(const double* x) { return x[0]; }
};

// Map-type models (ModelType::map, model_factory.h:22-26): rhs() is the next
// state f(x), jvp() the product J(x) v (StepMAP / StepMAPVAR,
// cuda/tools/mapstep.dp.cpp:5-18).  Jacobian entries that combine state and
// parameters are formed in FP64 and rounded once for the mixed form.
struct Henon {  // (a, b): x' = 1 - a x^2 + y, y' = b x
    static constexpr int ID = LYAP_MODEL_HENON, N = 2, P = 2;
    static constexpr bool kMap = true;
    template <class T>
    struct J {
        T j00;  // -2 a x
    };
    LYAP_D static void rhs(const double* x, const double* p, double* o) {
        o[0] = fma(-p[0], x[0] * x[0], 1.0 + x[1]);
        o[1] = p[1] * x[0];
    }
    template <class T>
    LYAP_D static J<T> jac(const double* x, const double* p) {
        return J<T>{(T)((-2.0 * p[0]) * x[0])};
    }
    template <class T>
    LYAP_D static void jvp(const J<T>& j, const T* p, const T* v, T* o) {
        o[0] = fma(j.j00, v[0], v[1]);
        o[1] = p[1] * v[0];
    }
};

struct Logistic {  // (r): x' = r x (1 - x)
    static constexpr int ID = LYAP_MODEL_LOGISTIC, N = 1, P = 1;
    static constexpr bool kMap = true;
    template <class T>
    struct J {
        T j;  // r (1 - 2x)
    };
    LYAP_D static void rhs(const double* x, const double* p, double* o) { o[0] = (p[0] * x[0]) * (1.0 - x[0]); }
    template <class T>
    LYAP_D static J<T> jac(const double* x, const double* p) {
        return J<T>{(T)(p[0] * fma(-2.0, x[0], 1.0))};
    }
    template <class T>
    LYAP_D static void jvp(const J<T>& j, const T*, const T* v, T* o) {
        o[0] = j.j * v[0];
    }
};

// ---------------------------------------------------------------------------
// step-size control (constants and update rule of cuda/tools/dopri.dp.cpp:9-12,
// 100-139; the exponent is 1/q - 0.75*beta with q the embedded order + 1)
// ---------------------------------------------------------------------------
// Step-control log and exp: the det_log / det_exp reductions with shorter
// series (about 1e-10 relative, ample for a step-size factor).  ctrl_exp is
// only asked for the range that survives the factor limits of ctrl_update, so
// its argument is clamped to [-1.6, 2.5] first.
LYAP_HD double ctrl_log(double x) {
    if (x != x) return x;
    if (x < 0.0) return __builtin_nan("");
    if (x == 0.0) return -kInf;
    if (x == kInf) return kInf;
    uint64_t b = bits_of(x);
    int e = (int)((b >> 52) & 0x7ff);
    int k;
    if (e == 0) {
        x = x * 18014398509481984.0;
        b = bits_of(x);
        e = (int)((b >> 52) & 0x7ff);
        k = e - 1023 - 54;
    } else {
        k = e - 1023;
    }
    double m = double_of((b & 0x000fffffffffffffULL) | 0x3ff0000000000000ULL);
    if (m > 1.4142135623730951) {
        m = m * 0.5;
        k = k + 1;
    }
    const double f = m - 1.0;
    const double s = f / (2.0 + f);
    const double z = s * s;
    double r = sconst(2.0 / 11.0);
    r = fma_c(r, z, 2.0 / 9.0);
    r = fma_c(r, z, 2.0 / 7.0);
    r = fma_c(r, z, 2.0 / 5.0);
    r = fma_c(r, z, 2.0 / 3.0);
    r = r * z;
    const double lg = f - s * (f - r);
    const double dk = (double)k;
    return fma(dk, sconst(kLn2Hi), fma(dk, sconst(kLn2Lo), lg));
}

LYAP_HD double ctrl_exp(double x) {
    x = fmin(fmax(x, -1.6), 2.5);  // never NaN here (see ctrl_update)
    const double kd = floor(x * 1.4426950408889634 + 0.5);
    const double r = fma(-kd, sconst(kLn2Lo), fma(-kd, sconst(kLn2Hi), x));
    double p = sconst(1.0 / 40320.0);
    p = fma_c(p, r, 1.0 / 5040.0);
    p = fma_c(p, r, 1.0 / 720.0);
    p = fma_c(p, r, 1.0 / 120.0);
    p = fma_c(p, r, 1.0 / 24.0);
    p = fma_c(p, r, 1.0 / 6.0);
    p = fma(p, r, 0.5);
    p = fma(p, r, 1.0);
    p = fma(p, r, 1.0);
    return p * double_of((uint64_t)(1023 + (int)kd) << 52);
}

struct Ctrl {
    double log_fac_old;  // ctrl_log of facOld (floored at 1e-4)
    bool rejected;
};

// dopriStep's PI rule (cuda/tools/dopri.dp.cpp:109-139): fac = err^expo /
// facOld^beta limited to [1/10, 1/0.2] after division by the safety factor,
// h_new = h / fac; after a rejection h_new = h / min(5, err^expo / 0.9).  Here
// in the equivalent form h_new = h * r, r = clamp(0.9 * exp(-(expo*lerr -
// beta*log facOld)), 0.2, 10): one exponential and no division.
// ok: err <= 1; above_floor: err > 1e-4; lerr = log(err) (ctrl_log);
// log_floor = ctrl_log(1e-4).  Both outcomes share one ctrl_exp call on the
// selected argument, so a wave whose lanes disagree evaluates it once.
template <int INTEG>
LYAP_D bool ctrl_update(Ctrl& c, bool ok, bool above_floor, double lerr, double log_floor, double h_try,
                        double h_max, double& h_next) {
    constexpr double kSafe = 0.9, kBeta = 0.04;
    const double e_rej = Tab<INTEG>::expo * lerr;
    const double e = ok ? fma(-kBeta, c.log_fac_old, e_rej) : e_rej;
    // e is finite or -inf (err = 0), never NaN; r, h_try and hn are finite
    // and positive, so min/max (one instruction each) give the clamps' bits
    double r = kSafe * ctrl_exp(-e);
    r = fmax(r, 0.2);
    if (ok) {
        r = fmin(r, 10.0);
        double hn = h_try * r;
        c.log_fac_old = above_floor ? lerr : log_floor;
        hn = fmin(hn, h_max);
        if (c.rejected) hn = fmin(hn, h_try);
        c.rejected = false;
        h_next = hn;
        return true;
    }
    c.rejected = true;
    h_next = h_try * r;
    return false;
}

// ---------------------------------------------------------------------------
// per-point set-up shared by every kernel
// ---------------------------------------------------------------------------
// parameters: tasklyapunov.cpp:45-63 grid (p1 outer) unless given explicitly
template <int P, class D>
LYAP_D void point_params(const D& d, int64_t idx, const double* params_in, double* p) {
    if (params_in) {
#pragma unroll
        for (int j = 0; j < P; ++j) p[j] = params_in[idx * P + j];
        return;
    }
    const int64_t i1 = idx / d.n2, i2 = idx % d.n2;
    const double v1 = d.start1 + (double)i1 * d.delta1;
    const double v2 = d.start2 + (double)i2 * d.delta2;
#pragma unroll
    for (int j = 0; j < P; ++j)
        p[j] = (j == d.sweep_idx1) ? v1 : ((j == d.sweep_idx2) ? v2 : d.base_params[j]);
}

// initial state: given, or init_state (+ p[offset]) + seeded perturbation
// (uniform on [-scale, scale], or scale * Irwin-Hall(12) - 6 normal deviate)
template <class D>
LYAP_D double init_component(const D& d, int64_t idx, int i, const double* p) {
    double x0 = d.init_state[i];
    if (d.init_param_offset >= 0) x0 = x0 + p[d.init_param_offset];
    double dx = 0.0;
    if (d.init_kind == LYAP_INIT_UNIFORM) {
        const double u = uniform01(d.seed, idx, i);
        dx = d.init_scale * (2.0 * u - 1.0);
    } else if (d.init_kind == LYAP_INIT_NORMAL) {
        double z = 0.0;
        for (int j = 0; j < 12; ++j) z = z + uniform01(d.seed, idx, 32 + i * 12 + j);
        dx = d.init_scale * (z - 6.0);
    }
    return x0 + dx;
}

// 1/sqrt(s) for the renormalisations of the sweep's method: the binary64
// bit-pattern seed of the classic inverse-square-root trick (relative error
// below 3.5 %) and four Newton steps y <- y (3/2 - (s/2) y^2), which end within
// ~1.2 ulp of 1/sqrt(s) (1.8e-3, 4.6e-6, 3.2e-11, 3.3e-16 after 1..4 steps).
// IEEE multiplies and fmas only, so the CPU restatement gives the same bits;
// 16 instructions where sqrt and a division take ~28.  The norm is s * y.
LYAP_HD double inv_sqrt_det(double s) {
    double y = double_of(0x5FE6EB50C7B537A9ULL - (bits_of(s) >> 1));
    const double hs = 0.5 * s;
#pragma unroll
    for (int it = 0; it < 4; ++it) y = y * fma_c(-hs, y * y, 1.5);
    return y;
}

LYAP_D void kahan_add(double& s, double& c, double v) {
    const double y = v - c;
    const double t = s + y;
    c = (t - s) - y;
    s = t;
}

// Log-norm accumulation of the sweep's method.  The norms of up to
// kLogWindow consecutive renormalisations are multiplied into a product kept
// as a mantissa in [1, 2) and an exact binary exponent; the log of each
// window's product, log(mant) + ex*ln2, is Kahan-added to the sum.  One log
// per window instead of one per renormalisation (the product rounds once per
// factor, far below the log's own error).  A mantissa that is zero,
// subnormal, infinite or NaN keeps its exponent field, so its log is what the
// log of the norm itself would have been.
constexpr int kLogWindow = 64;

LYAP_HD double split_exponent(double m, int32_t& ex) {
    const uint64_t b = bits_of(m);
    const int e = (int)((b >> 52) & 0x7ff);
    if (e == 0 || e == 0x7ff) return m;
    ex += e - 1023;
    return double_of((b & 0x800fffffffffffffULL) | 0x3ff0000000000000ULL);
}

LYAP_HD double log_scaled(double mant, int32_t ex) {
    const double dk = (double)ex;
    return fma(dk, sconst(kLn2Hi), fma(dk, sconst(kLn2Lo), det_log(mant)));
}

// one vector's share: push a norm; flush adds the window's log to (s, c)
LYAP_HD void lognorm_push(double& mant, int32_t& ex, double nrm) { mant = split_exponent(mant * nrm, ex); }
LYAP_HD void lognorm_flush(double& mant, int32_t& ex, double& s, double& c) {
    kahan_add(s, c, log_scaled(mant, ex));
    mant = 1.0;
    ex = 0;
}

// The descriptor as it sits in the kernarg segment (the first kernel argument,
// offset 0), read through an opaque constant-address-space pointer: fields
// needed only when a lane or block takes a new point are then scalar loads at
// that moment, instead of values the compiler keeps live (and spills) across
// the whole step loop.  Taking the address of the by-value parameter instead
// would copy the descriptor to scratch.
using KernargDesc = const __attribute__((address_space(4))) lyap_sweep_desc;
LYAP_D KernargDesc* kernarg_desc() {
    KernargDesc* p = (KernargDesc*)__builtin_amdgcn_kernarg_segment_ptr();
    __asm__ volatile("" : "+s"(p));
    return p;
}

// outputs shared by all sweep kernels
struct SweepIO {
    int64_t first, count;
    const int64_t* idx_list;
    const double* params;
    const double* inits;
    double* lambda;
    uint32_t* status;
    uint32_t* n_acc;
    uint32_t* n_rej;
    unsigned long long* counter;  // work counter (persistent kernels)
    int64_t stride;               // work item i takes position (i * stride) mod count
    int32_t gather;               // per-lane kernels: lanes at a renormalisation wait until
                                  // this many lanes of the wave wait (<= 1: renormalise at once)
    int32_t pad;
};

// position of work item `item`: a multiplicative permutation of [0, count)
// (stride coprime to count) that spreads consecutive items over the grid, so
// the last items taken are not one expensive region of the parameter plane
LYAP_D int64_t item_position(int64_t item, int64_t stride, int64_t count) {
    return (int64_t)(((unsigned __int128)item * (unsigned __int128)stride) % (unsigned __int128)count);
}
LYAP_D int64_t item_position(const SweepIO& io, int64_t item) { return item_position(item, io.stride, io.count); }

// The IO record of a kernel whose arguments are (lyap_sweep_desc, SweepIO, ...):
// it follows the descriptor in the kernarg segment; read through an opaque
// pointer like kernarg_desc().
using KernargIO = const __attribute__((address_space(4))) SweepIO;
static_assert(sizeof(lyap_sweep_desc) % alignof(SweepIO) == 0, "SweepIO follows the descriptor unpadded");
LYAP_D KernargIO* kernarg_io() {
    const __attribute__((address_space(4))) char* b =
        (const __attribute__((address_space(4))) char*)__builtin_amdgcn_kernarg_segment_ptr();
    KernargIO* p = (KernargIO*)(b + sizeof(lyap_sweep_desc));
    __asm__ volatile("" : "+s"(p));
    return p;
}

// Progress signals and per-item skips of one launch of the sweep's method
// (k_var_lane's third argument; all pointers null: neither).  Items are list
// positions.  Items [0, end[0]) form milestone 0 and [end[0], end[1])
// milestone 1: a lane that finishes such an item releases its outputs to the
// device scope and counts it in done[k]; the lane that finishes the last one
// stores `seq` to flag[k], memory the host reads while the launch runs.  An
// item whose skip word is nonzero when a lane takes it is not integrated: its
// status is LYAP_PT_SKIPPED and its exponents NaN.  The skip words are read
// past the caches, so words written while the launch runs are seen.
struct LaneWatch {
    const uint32_t* skip;
    unsigned long long* done;  // [2], zero at launch
    uint32_t* flag;            // [2], host-visible
    int64_t end[2];
    uint32_t seq;
    uint32_t pad;
};
using KernargWatch = const __attribute__((address_space(4))) LaneWatch;
static_assert(sizeof(SweepIO) % alignof(LaneWatch) == 0, "LaneWatch follows SweepIO unpadded");
LYAP_D KernargWatch* kernarg_watch() {
    const __attribute__((address_space(4))) char* b =
        (const __attribute__((address_space(4))) char*)__builtin_amdgcn_kernarg_segment_ptr();
    KernargWatch* p = (KernargWatch*)(b + sizeof(lyap_sweep_desc) + sizeof(SweepIO));
    __asm__ volatile("" : "+s"(p));
    return p;
}

}  // namespace lyap
