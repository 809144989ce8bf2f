/*
 * CPU FP64 oracle for the Lyapunov-spectrum sweep — TEST INFRASTRUCTURE ONLY.
 * See lyap_oracle.h for what it restates and who may call it.
 *
 * Arithmetic contract (shared with the GPU path, which is written separately):
 *   - built with -ffp-contract=off; every fused multiply-add is an explicit fma();
 *   - RK stage sums are accumulated in increasing stage order, first term a
 *     product, later terms fma(a_j, k_j, acc), then fma(h, acc, y);
 *   - dot products are sequential fma chains; the error norm of a step is a
 *     sequential fma chain for the 3-D models and a xor-butterfly (offsets
 *     n/2, n/4, ..., 1) over plain products for Lorenz-96, whose components
 *     live in different GPU lanes;
 *   - log/exp/pow are the deterministic routines below (no libm), so every
 *     branch the adaptive controller takes is reproducible bit for bit.
 *
 * Parity status of each part (how it is pinned) is recorded with the tests
 * in tests/test_oracle_*.py.
 */
#include "lyap_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* ------------------------------------------------------------------------ */
/* deterministic elementary functions                                        */
/* ------------------------------------------------------------------------ */
static inline uint64_t as_u64(double x) { uint64_t u; memcpy(&u, &x, 8); return u; }
static inline double as_f64(uint64_t u) { double x; memcpy(&x, &u, 8); return x; }

static const double LN2_HI = 6.93147180369123816490e-01; /* 0x3FE62E42FEE00000 */
static const double LN2_LO = 1.90821492927058770002e-10; /* ln2 - LN2_HI */

/* log via the atanh series: log(1+f) = f - s*(f - R), s = f/(2+f),
   R = 2*sum_{j>=1} s^{2j}/(2j+1); mantissa reduced to [sqrt(1/2), sqrt(2)]. */
double orc_log(double x) {
    if (x != x) return x;
    if (x < 0.0) return NAN;
    if (x == 0.0) return -INFINITY;
    if (x == INFINITY) return INFINITY;
    uint64_t b = as_u64(x);
    int e = (int)((b >> 52) & 0x7ff);
    int k;
    if (e == 0) { /* subnormal */
        x = x * 18014398509481984.0; /* 2^54 */
        b = as_u64(x);
        e = (int)((b >> 52) & 0x7ff);
        k = e - 1023 - 54;
    } else {
        k = e - 1023;
    }
    double m = as_f64((b & 0x000fffffffffffffULL) | 0x3ff0000000000000ULL);
    if (m > 1.4142135623730951) { m = m * 0.5; k = k + 1; }
    double f = m - 1.0;
    double s = f / (2.0 + f);
    double z = s * s;
    double r = 2.0 / 25.0;
    r = fma(r, z, 2.0 / 23.0);
    r = fma(r, z, 2.0 / 21.0);
    r = fma(r, z, 2.0 / 19.0);
    r = fma(r, z, 2.0 / 17.0);
    r = fma(r, z, 2.0 / 15.0);
    r = fma(r, z, 2.0 / 13.0);
    r = fma(r, z, 2.0 / 11.0);
    r = fma(r, z, 2.0 / 9.0);
    r = fma(r, z, 2.0 / 7.0);
    r = fma(r, z, 2.0 / 5.0);
    r = fma(r, z, 2.0 / 3.0);
    r = r * z;
    double lg = f - s * (f - r);
    double dk = (double)k;
    return fma(dk, LN2_HI, fma(dk, LN2_LO, lg));
}

/* exp via round-to-nearest reduction by ln2 and a degree-13 Taylor polynomial */
double orc_exp(double x) {
    if (x != x) return x;
    if (x > 709.782712893384) return INFINITY;
    if (x < -745.1332191019412) return 0.0;
    double kd = x * 1.4426950408889634;
    kd = (kd >= 0.0) ? floor(kd + 0.5) : ceil(kd - 0.5);
    double r = fma(-kd, LN2_LO, fma(-kd, LN2_HI, x));
    double p = 1.0 / 6227020800.0; /* 1/13! */
    p = fma(p, r, 1.0 / 479001600.0);
    p = fma(p, r, 1.0 / 39916800.0);
    p = fma(p, r, 1.0 / 3628800.0);
    p = fma(p, r, 1.0 / 362880.0);
    p = fma(p, r, 1.0 / 40320.0);
    p = fma(p, r, 1.0 / 5040.0);
    p = fma(p, r, 1.0 / 720.0);
    p = fma(p, r, 1.0 / 120.0);
    p = fma(p, r, 1.0 / 24.0);
    p = fma(p, r, 1.0 / 6.0);
    p = fma(p, r, 0.5);
    p = fma(p, r, 1.0);
    p = fma(p, r, 1.0);
    int k = (int)kd;
    if (k > 1023) { p = p * as_f64((uint64_t)(1023 + 1023) << 52); k -= 1023; }
    if (k < -1022) { p = p * as_f64((uint64_t)1 << 52); k += 1022; } /* times 2^-1022 */
    return p * as_f64((uint64_t)(k + 1023) << 52);
}

double orc_pow(double x, double y) {
    if (x == 0.0) return 0.0;
    return orc_exp(y * orc_log(x));
}

/* ------------------------------------------------------------------------ */
/* counter-based generator (splitmix64 finaliser), loop counts               */
/* ------------------------------------------------------------------------ */
static inline uint64_t mix64(uint64_t z) {
    z += 0x9E3779B97F4A7C15ULL;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
}

double orc_uniform(uint64_t seed, int64_t point, int64_t stream) {
    uint64_t z = mix64(seed ^ mix64((uint64_t)point * 1024ULL + (uint64_t)stream));
    return (double)(z >> 11) * (1.0 / 9007199254740992.0);
}

/* tasklyapunov.cpp / cuda/lyapunov.dp.cpp:11,139,158,244 iterate
   `for (double t = t0; t < t1; t += h)`; the count is not (t1-t0)/h. */
int64_t orc_loop_count(double t0, double t1, double h) {
    int64_t n = 0;
    for (double t = t0; t < t1; t += h) ++n;
    return n;
}

/* ------------------------------------------------------------------------ */
/* models                                                                    */
/* ------------------------------------------------------------------------ */
/* shimizu: cuda/models/flows/shimizu.dp.cpp:15-27 (x'=y, y'=-a*y-x*z+x,
   z'=-l*z+x^2; params (alpha, lambda)).  lorenz (sigma, rho, beta),
   rossler (a, c, b), Lorenz-96 (F, d): x_i' = (x_{i+1}-x_{i-2}) x_{i-1} - d x_i + F. */
void orc_rhs(int model, int n, const double* x, const double* p, double* out) {
    switch (model) {
    case ORC_MODEL_SHIMIZU:
        out[0] = x[1];
        out[1] = fma(-p[0], x[1], fma(-x[0], x[2], x[0]));
        out[2] = fma(x[0], x[0], -p[1] * x[2]);
        break;
    case ORC_MODEL_LORENZ:
        out[0] = p[0] * (x[1] - x[0]);
        out[1] = fma(x[0], p[1] - x[2], -x[1]);
        out[2] = fma(x[0], x[1], -p[2] * x[2]);
        break;
    case ORC_MODEL_ROSSLER:
        out[0] = -(x[1] + x[2]);
        out[1] = fma(p[0], x[1], x[0]);
        out[2] = fma(x[2], x[0] - p[1], p[2]);
        break;
    case ORC_MODEL_L96:
        for (int i = 0; i < n; ++i) {
            int ip1 = (i + 1) % n, im1 = (i + n - 1) % n, im2 = (i + n - 2) % n;
            out[i] = fma(x[ip1] - x[im2], x[im1], fma(-p[1], x[i], p[0]));
        }
        break;
    case ORC_MODEL_HENON: /* (a, b): x' = 1 - a x^2 + y, y' = b x */
        out[0] = fma(-p[0], x[0] * x[0], 1.0 + x[1]);
        out[1] = p[1] * x[0];
        break;
    case ORC_MODEL_LOGISTIC: /* (r): x' = r x (1 - x) */
        out[0] = (p[0] * x[0]) * (1.0 - x[0]);
        break;
    default:
        for (int i = 0; i < n; ++i) out[i] = NAN;
    }
}

int orc_is_map(int model) { return model == ORC_MODEL_HENON || model == ORC_MODEL_LOGISTIC; }

void orc_jvp(int model, int n, const double* x, const double* p, const double* v, double* out) {
    switch (model) {
    case ORC_MODEL_SHIMIZU:
        out[0] = v[1];
        out[1] = fma(1.0 - x[2], v[0], fma(-p[0], v[1], -x[0] * v[2]));
        out[2] = fma(2.0 * x[0], v[0], -p[1] * v[2]);
        break;
    case ORC_MODEL_LORENZ:
        out[0] = p[0] * (v[1] - v[0]);
        out[1] = fma(p[1] - x[2], v[0], -fma(x[0], v[2], v[1]));
        out[2] = fma(x[1], v[0], fma(x[0], v[1], -p[2] * v[2]));
        break;
    case ORC_MODEL_ROSSLER:
        out[0] = -(v[1] + v[2]);
        out[1] = fma(p[0], v[1], v[0]);
        out[2] = fma(x[2], v[0], (x[0] - p[1]) * v[2]);
        break;
    case ORC_MODEL_L96:
        for (int i = 0; i < n; ++i) {
            int ip1 = (i + 1) % n, im1 = (i + n - 1) % n, im2 = (i + n - 2) % n;
            out[i] = fma(v[ip1] - v[im2], x[im1], fma(x[ip1] - x[im2], v[im1], -p[1] * v[i]));
        }
        break;
    case ORC_MODEL_HENON: /* J = [[-2 a x, 1], [b, 0]] */
        out[0] = fma((-2.0 * p[0]) * x[0], v[0], v[1]);
        out[1] = p[1] * v[0];
        break;
    case ORC_MODEL_LOGISTIC: /* J = r (1 - 2x) */
        out[0] = (p[0] * fma(-2.0, x[0], 1.0)) * v[0];
        break;
    default:
        for (int i = 0; i < n; ++i) out[i] = NAN;
    }
}

/* max row sum of |J(x)|, used to cap adaptive steps: h*||J||_inf <= 1 keeps the
   tangent integration accurate where the base error estimate vanishes
   (stable fixed points) */
double orc_jac_norm(int model, int n, const double* x, const double* p) {
    double r0, r1, r2, m;
    switch (model) {
    case ORC_MODEL_SHIMIZU:
        r0 = 1.0;
        r1 = (fabs(1.0 - x[2]) + fabs(p[0])) + fabs(x[0]);
        r2 = 2.0 * fabs(x[0]) + fabs(p[1]);
        break;
    case ORC_MODEL_LORENZ:
        r0 = 2.0 * fabs(p[0]);
        r1 = (fabs(p[1] - x[2]) + 1.0) + fabs(x[0]);
        r2 = (fabs(x[1]) + fabs(x[0])) + fabs(p[2]);
        break;
    case ORC_MODEL_ROSSLER:
        r0 = 2.0;
        r1 = 1.0 + fabs(p[0]);
        r2 = fabs(x[2]) + fabs(x[0] - p[1]);
        break;
    case ORC_MODEL_L96:
        m = 0.0;
        for (int i = 0; i < n; ++i) {
            int ip1 = (i + 1) % n, im1 = (i + n - 1) % n, im2 = (i + n - 2) % n;
            double r = (2.0 * fabs(x[im1]) + fabs(x[ip1] - x[im2])) + fabs(p[1]);
            m = r > m ? r : m;
        }
        return m;
    default:
        return NAN;
    }
    m = r0 > r1 ? r0 : r1;
    return m > r2 ? m : r2;
}

/* tr J = c(p) + g(x) of the 3-D flows (the divergence of f): shimizu
   -alpha - lambda, Lorenz -sigma - 1 - beta, Rossler (a - c) + x */
static double trace_const(int model, const double* p) {
    switch (model) {
    case ORC_MODEL_SHIMIZU: return -(p[0] + p[1]);
    case ORC_MODEL_LORENZ: return -((p[0] + 1.0) + p[2]);
    case ORC_MODEL_ROSSLER: return p[0] - p[1];
    default: return NAN;
    }
}
static int trace_has_var(int model) { return model == ORC_MODEL_ROSSLER; }
static double trace_var(int model, const double* x) { return model == ORC_MODEL_ROSSLER ? x[0] : 0.0; }

double orc_trace(int model, int n, const double* x, const double* p) {
    (void)n;
    return trace_has_var(model) ? trace_const(model, p) + trace_var(model, x) : trace_const(model, p);
}

/* the same products with binary32 arithmetic (the mixed form): Jacobian
   entries that are differences of state and parameter (rho - z, x - c,
   1 - z, x_{i+1} - x_{i-2}) are formed in FP64 from the FP64 stage state and
   rounded once; every other operand is rounded to float first and every
   operation is a float operation */
void orc_jvp_f32(int model, int n, const double* xd, const double* pd, const float* v, float* out) {
    float x[ORC_MAX_DIM], p[ORC_MAX_PARAMS];
    for (int i = 0; i < n; ++i) x[i] = (float)xd[i];
    for (int i = 0; i < ORC_MAX_PARAMS; ++i) p[i] = (float)pd[i];
    switch (model) {
    case ORC_MODEL_SHIMIZU: {
        const float oz = (float)(1.0 - xd[2]);
        out[0] = v[1];
        out[1] = fmaf(oz, v[0], fmaf(-p[0], v[1], -x[0] * v[2]));
        out[2] = fmaf(2.0f * x[0], v[0], -p[1] * v[2]);
        break;
    }
    case ORC_MODEL_LORENZ: {
        const float rz = (float)(pd[1] - xd[2]);
        out[0] = p[0] * (v[1] - v[0]);
        out[1] = fmaf(rz, v[0], -fmaf(x[0], v[2], v[1]));
        out[2] = fmaf(x[1], v[0], fmaf(x[0], v[1], -p[2] * v[2]));
        break;
    }
    case ORC_MODEL_ROSSLER: {
        const float xc = (float)(xd[0] - pd[1]);
        out[0] = -(v[1] + v[2]);
        out[1] = fmaf(p[0], v[1], v[0]);
        out[2] = fmaf(x[2], v[0], xc * v[2]);
        break;
    }
    case ORC_MODEL_L96:
        for (int i = 0; i < n; ++i) {
            int ip1 = (i + 1) % n, im1 = (i + n - 1) % n, im2 = (i + n - 2) % n;
            const float dx = (float)(xd[ip1] - xd[im2]);
            /* the dx*v(i-1) product rounded first (the kernels fold its lane
               move into a DPP multiply) */
            out[i] = fmaf(v[ip1] - v[im2], x[im1], fmaf(-p[1], v[i], dx * v[im1]));
        }
        break;
    case ORC_MODEL_HENON: {
        const float j00 = (float)((-2.0 * pd[0]) * xd[0]);
        out[0] = fmaf(j00, v[0], v[1]);
        out[1] = p[1] * v[0];
        break;
    }
    case ORC_MODEL_LOGISTIC: {
        const float j = (float)(pd[0] * fma(-2.0, xd[0], 1.0));
        out[0] = j * v[0];
        break;
    }
    default:
        for (int i = 0; i < n; ++i) out[i] = NAN;
    }
}

/* ------------------------------------------------------------------------ */
/* tableaux                                                                  */
/* ------------------------------------------------------------------------ */
/* DVERK 6(5) (Verner; Hull, Enright & Jackson 1976), as used by
   cuda/tools/dverk.dp.cpp:11-42.  b: 6th order, e = b - bhat (5th order). */
static const double DV_A[8][8] = {
    {0},
    {1.0 / 6.0},
    {4.0 / 75.0, 16.0 / 75.0},
    {5.0 / 6.0, -8.0 / 3.0, 5.0 / 2.0},
    {-165.0 / 64.0, 55.0 / 6.0, -425.0 / 64.0, 85.0 / 96.0},
    {12.0 / 5.0, -8.0, 4015.0 / 612.0, -11.0 / 36.0, 88.0 / 255.0},
    {-8263.0 / 15000.0, 124.0 / 75.0, -643.0 / 680.0, -81.0 / 250.0, 2484.0 / 10625.0, 0.0},
    {3501.0 / 1720.0, -300.0 / 43.0, 297275.0 / 52632.0, -319.0 / 2322.0, 24068.0 / 84065.0, 0.0,
     3850.0 / 26703.0}};
static const double DV_B[8] = {3.0 / 40.0, 0.0, 875.0 / 2244.0, 23.0 / 72.0, 264.0 / 1955.0, 0.0,
                               125.0 / 11592.0, 43.0 / 616.0};
static const double DV_E[8] = {-1.0 / 160.0, 0.0, -125.0 / 17952.0, 1.0 / 144.0, -12.0 / 1955.0,
                               -3.0 / 44.0, 125.0 / 11592.0, 43.0 / 616.0};
/* the embedded fifth-order weights b - e (stages 1-6 only): the binary32
   tangent of the mixed form with adaptive DVERK steps takes these
   (lyap_lane.hpp tangent_step LOW) */
static const double DV_BH[8] = {13.0 / 160.0, 0.0, 2375.0 / 5984.0, 5.0 / 16.0, 12.0 / 85.0,
                                3.0 / 44.0, 0.0, 0.0};

/* Dormand-Prince 5(4), cuda/tools/dopri.dp.cpp:14-45 */
static const double DP_A[7][7] = {
    {0},
    {1.0 / 5.0},
    {3.0 / 40.0, 9.0 / 40.0},
    {44.0 / 45.0, -56.0 / 15.0, 32.0 / 9.0},
    {19372.0 / 6561.0, -25360.0 / 2187.0, 64448.0 / 6561.0, -212.0 / 729.0},
    {9017.0 / 3168.0, -355.0 / 33.0, 46732.0 / 5247.0, 49.0 / 176.0, -5103.0 / 18656.0},
    {35.0 / 384.0, 0.0, 500.0 / 1113.0, 125.0 / 192.0, -2187.0 / 6784.0, 11.0 / 84.0}};
static const double DP_B[7] = {35.0 / 384.0, 0.0, 500.0 / 1113.0, 125.0 / 192.0, -2187.0 / 6784.0,
                               11.0 / 84.0, 0.0};
static const double DP_E[7] = {71.0 / 57600.0, 0.0, -71.0 / 16695.0, 71.0 / 1920.0,
                               -17253.0 / 339200.0, 22.0 / 525.0, -1.0 / 40.0};

/* stage sum: acc = sum_{j<s, a_j != 0} a_j k_j, in increasing j, first term a
   product; returns fma(h, acc, y) (or y itself when no coefficient). */
static inline double comb_d(const double* a, int s, const double* k, int stride, int i, double h,
                            double y) {
    double acc = 0.0;
    int first = 1;
    for (int j = 0; j < s; ++j) {
        if (a[j] == 0.0) continue;
        double kj = k[j * stride + i];
        if (first) { acc = a[j] * kj; first = 0; }
        else acc = fma(a[j], kj, acc);
    }
    return first ? y : fma(h, acc, y);
}

static inline float comb_f(const double* a, int s, const float* k, int stride, int i, float h,
                           float y) {
    float acc = 0.0f;
    int first = 1;
    for (int j = 0; j < s; ++j) {
        if (a[j] == 0.0) continue;
        float aj = (float)a[j];
        float kj = k[j * stride + i];
        if (first) { acc = aj * kj; first = 0; }
        else acc = fmaf(aj, kj, acc);
    }
    return first ? y : fmaf(h, acc, y);
}

static inline float sum_f(const double* a, int s, const float* k, int stride, int i) {
    float acc = 0.0f;
    int first = 1;
    for (int j = 0; j < s; ++j) {
        if (a[j] == 0.0) continue;
        float aj = (float)a[j];
        float kj = k[j * stride + i];
        if (first) { acc = aj * kj; first = 0; }
        else acc = fmaf(aj, kj, acc);
    }
    return acc;
}

static int n_stages(int integ) { return integ == ORC_INT_DVERK ? 8 : 7; }
static const double* tab_a(int integ, int s) { return integ == ORC_INT_DVERK ? DV_A[s] : DP_A[s]; }
static const double* tab_b(int integ) { return integ == ORC_INT_DVERK ? DV_B : DP_B; }
static const double* tab_e(int integ) { return integ == ORC_INT_DVERK ? DV_E : DP_E; }
/* the step's integral of tr J by the solution's quadrature (the RK method on
   the added equation y' = tr J(x)): h * (c + sum_j b_j g(X_j)) */
double orc_trace_step(int model, int integ, int n, const double* stage_x, const double* p, double h) {
    const double c = trace_const(model, p);
    if (!trace_has_var(model)) return h * c;
    const double* b = tab_b(integ);
    double q = 0.0;
    int first = 1;
    for (int j = 0; j < n_stages(integ); ++j) {
        if (b[j] == 0.0) continue;
        const double g = trace_var(model, stage_x + j * n);
        if (first) { q = b[j] * g; first = 0; }
        else q = fma(b[j], g, q);
    }
    return h * (c + q);
}

/* stages the solution needs: DVERK stage 6 only feeds the error estimate,
   Dopri5 stage 7 only feeds the error estimate (and is the next k1). */
static int stage_needed_for_solution(int integ, int s) {
    return integ == ORC_INT_DVERK ? (s != 5) : (s != 6);
}

int orc_rk_step(int model, int integ, int n, const double* x, const double* p, double h,
                double* x_new, double* err_vec, double* stage_x) {
    double k[8 * ORC_MAX_DIM], arg[ORC_MAX_DIM];
    const int S = n_stages(integ);
    const int want_err = err_vec != NULL;
    int evals = 0;
    for (int s = 0; s < S; ++s) {
        if (!want_err && !stage_needed_for_solution(integ, s)) {
            for (int i = 0; i < n; ++i) k[s * ORC_MAX_DIM + i] = 0.0;
            continue;
        }
        const double* a = tab_a(integ, s);
        for (int i = 0; i < n; ++i) arg[i] = comb_d(a, s, k, ORC_MAX_DIM, i, h, x[i]);
        if (stage_x) memcpy(stage_x + s * n, arg, sizeof(double) * n);
        orc_rhs(model, n, arg, p, k + s * ORC_MAX_DIM);
        ++evals;
    }
    const double* b = tab_b(integ);
    for (int i = 0; i < n; ++i) x_new[i] = comb_d(b, S, k, ORC_MAX_DIM, i, h, x[i]);
    if (want_err) {
        const double* e = tab_e(integ);
        for (int i = 0; i < n; ++i) {
            double acc = 0.0;
            int first = 1;
            for (int j = 0; j < S; ++j) {
                if (e[j] == 0.0) continue;
                if (first) { acc = e[j] * k[j * ORC_MAX_DIM + i]; first = 0; }
                else acc = fma(e[j], k[j * ORC_MAX_DIM + i], acc);
            }
            err_vec[i] = h * acc;
        }
    }
    /* Dopri5 is first-same-as-last: stage 7 is the next step's stage 1 */
    if (integ == ORC_INT_DOPRI5 && evals > 1) --evals;
    return evals;
}

/* coupled tangent step: stage s uses J(X_s), X_s from orc_rk_step */
void orc_tangent_step(int model, int integ, int n, const double* stage_x, const double* p,
                      double h, double* v) {
    double kv[8 * ORC_MAX_DIM], arg[ORC_MAX_DIM];
    const int S = n_stages(integ);
    for (int s = 0; s < S; ++s) {
        if (!stage_needed_for_solution(integ, s)) {
            for (int i = 0; i < n; ++i) kv[s * ORC_MAX_DIM + i] = 0.0;
            continue;
        }
        const double* a = tab_a(integ, s);
        for (int i = 0; i < n; ++i) arg[i] = comb_d(a, s, kv, ORC_MAX_DIM, i, h, v[i]);
        orc_jvp(model, n, stage_x + s * n, p, arg, kv + s * ORC_MAX_DIM);
    }
    const double* b = tab_b(integ);
    double out[ORC_MAX_DIM];
    for (int i = 0; i < n; ++i) out[i] = comb_d(b, S, kv, ORC_MAX_DIM, i, h, v[i]);
    memcpy(v, out, sizeof(double) * n);
}

/* mixed form: the deviation vector stays FP64; the stage arguments, the
   Jacobian products and the stage sums are binary32; the step increment
   h*sum(b_j K_j) is formed in binary32 and added to the FP64 vector.  low
   (adaptive DVERK steps): the embedded fifth-order weights DV_BH on stages
   1-6, whose states the error-controlled base step has computed. */
void orc_tangent_step_f32(int model, int integ, int n, const double* stage_x, const double* p,
                          double h, double* v, int low) {
    float kv[8 * ORC_MAX_DIM], arg[ORC_MAX_DIM], vf[ORC_MAX_DIM];
    const int S = n_stages(integ);
    const float hf = (float)h;
    low = low && integ == ORC_INT_DVERK;
    for (int i = 0; i < n; ++i) vf[i] = (float)v[i];
    for (int s = 0; s < S; ++s) {
        if (low ? s >= 6 : !stage_needed_for_solution(integ, s)) {
            for (int i = 0; i < n; ++i) kv[s * ORC_MAX_DIM + i] = 0.0f;
            continue;
        }
        const double* a = tab_a(integ, s);
        for (int i = 0; i < n; ++i) arg[i] = comb_f(a, s, kv, ORC_MAX_DIM, i, hf, vf[i]);
        orc_jvp_f32(model, n, stage_x + s * n, p, arg, kv + s * ORC_MAX_DIM);
    }
    const double* b = low ? DV_BH : tab_b(integ);
    for (int i = 0; i < n; ++i) {
        float inc = hf * sum_f(b, S, kv, ORC_MAX_DIM, i);
        v[i] = v[i] + (double)inc;
    }
}

/* Lorenz-96 mixed form: as tangent_step_f32 with the step h folded into the
   binary32 Jacobian data (the tangent system is linear), which the kernels
   form once per stage and share between the vectors of a lane:
   J'_s = ((float)(h x_{i-1}), (float)(h (x_{i+1} - x_{i-2})), (float)(h d)),
   stage arguments v + sum_j a_sj K'_j as one fma chain started on v, products
   K'_s = J'_s arg (h times the stage derivatives), and the increment
   sum(b_j K'_j) added to the FP64 vector. */
static void l96_tangent_step_f32(int integ, int n, const double* stage_x, const double* p,
                                 double h, double* v) {
    float kv[8 * ORC_MAX_DIM], arg[ORC_MAX_DIM], vf[ORC_MAX_DIM];
    const int S = n_stages(integ);
    const float dh = (float)(h * p[1]);
    for (int i = 0; i < n; ++i) vf[i] = (float)v[i];
    for (int s = 0; s < S; ++s) {
        float* ks = kv + s * ORC_MAX_DIM;
        if (!stage_needed_for_solution(integ, s)) {
            for (int i = 0; i < n; ++i) ks[i] = 0.0f;
            continue;
        }
        const double* a = tab_a(integ, s);
        for (int i = 0; i < n; ++i) {
            float acc = vf[i];
            for (int j = 0; j < s; ++j)
                if (a[j] != 0.0) acc = fmaf((float)a[j], kv[j * ORC_MAX_DIM + i], acc);
            arg[i] = acc;
        }
        const double* xs = stage_x + s * n;
        for (int i = 0; i < n; ++i) {
            const int ip1 = (i + 1) % n, im1 = (i + n - 1) % n, im2 = (i + n - 2) % n;
            const float jm1 = (float)(h * xs[im1]);
            const float jd = (float)(h * (xs[ip1] - xs[im2]));
            ks[i] = fmaf(arg[ip1] - arg[im2], jm1, fmaf(-dh, arg[i], jd * arg[im1]));
        }
    }
    const double* b = tab_b(integ);
    for (int i = 0; i < n; ++i) v[i] = v[i] + (double)sum_f(b, S, kv, ORC_MAX_DIM, i);
}

/* cuda/tools/dverk.dp.cpp:46-80: every stage uses J at the same base point */
void orc_frozen_tangent_step(int model, int n, const double* x_frozen, const double* p, double h,
                             double* v) {
    double kv[8 * ORC_MAX_DIM], arg[ORC_MAX_DIM];
    for (int s = 0; s < 8; ++s) {
        if (!stage_needed_for_solution(ORC_INT_DVERK, s)) {
            for (int i = 0; i < n; ++i) kv[s * ORC_MAX_DIM + i] = 0.0;
            continue;
        }
        for (int i = 0; i < n; ++i) arg[i] = comb_d(DV_A[s], s, kv, ORC_MAX_DIM, i, h, v[i]);
        orc_jvp(model, n, x_frozen, p, arg, kv + s * ORC_MAX_DIM);
    }
    double out[ORC_MAX_DIM];
    for (int i = 0; i < n; ++i) out[i] = comb_d(DV_B, 8, kv, ORC_MAX_DIM, i, h, v[i]);
    memcpy(v, out, sizeof(double) * n);
}

/* ------------------------------------------------------------------------ */
/* reductions and orthogonalisation                                          */
/* ------------------------------------------------------------------------ */
/* xor butterfly over plain products: every lane of a GPU group ends with the
   same bits, equal to this value */
static double tree_sum(int n, double* a /* clobbered */) {
    for (int o = n / 2; o >= 1; o /= 2) {
        double t[ORC_MAX_DIM];
        for (int i = 0; i < n; ++i) t[i] = a[i] + a[i ^ o];
        memcpy(a, t, sizeof(double) * n);
    }
    return a[0];
}

static double dot_ns(int n, int tree, const double* a, const double* b) {
    if (tree) {
        double t[ORC_MAX_DIM];
        for (int i = 0; i < n; ++i) t[i] = a[i] * b[i];
        return tree_sum(n, t);
    }
    double s = a[0] * b[0];
    for (int i = 1; i < n; ++i) s = fma(a[i], b[i], s);
    return s;
}

/* 1/sqrt(s) of the sweep's renormalisations: the binary64 bit-pattern seed
   of the classic inverse-square-root trick (relative error below 3.5 %) and
   four Newton steps y <- y*(3/2 - (s/2)*y^2); within ~1.2 ulp of 1/sqrt(s)
   (pinned against sqrt in tests/test_oracle_pins.py) and made of IEEE
   multiplies and fmas only, so the GPU gives the same bits. */
double orc_inv_sqrt(double s) {
    double y = as_f64(0x5FE6EB50C7B537A9ULL - (as_u64(s) >> 1));
    const double hs = 0.5 * s;
    for (int it = 0; it < 4; ++it) y = y * fma(-hs, y * y, 1.5);
    return y;
}

/* modified Gram-Schmidt on k row vectors of length n; rows end unit length
   (scaled by orc_inv_sqrt of the squared norm); norms[j] = row j's norm
   before its normalisation, s * orc_inv_sqrt(s) */
static void mgs_norms(int n, int k, int tree, double* V, double* norms) {
    for (int j = 0; j < k; ++j) {
        double* vj = V + j * n;
        for (int r = 0; r < j; ++r) {
            const double* vr = V + r * n;
            double c = dot_ns(n, tree, vr, vj);
            for (int i = 0; i < n; ++i) vj[i] = fma(-c, vr[i], vj[i]);
        }
        const double s2 = dot_ns(n, tree, vj, vj);
        const double inv = orc_inv_sqrt(s2);
        norms[j] = s2 * inv;
        for (int i = 0; i < n; ++i) vj[i] = vj[i] * inv;
    }
}

/* the same, with log_norms[j] = log of row j's norm */
void orc_mgs(int n, int k, int tree, double* V, double* log_norms) {
    mgs_norms(n, k, tree, V, log_norms);
    for (int j = 0; j < k; ++j) log_norms[j] = orc_log(log_norms[j]);
}

/* Log-norm accumulation of the sweep's method.  The norms of up to
   ORC_LOG_WINDOW consecutive renormalisations are multiplied into a product
   kept as a mantissa in [1, 2) and an exact binary exponent; the log of each
   window's product, log(mant) + ex*ln2, is Kahan-added to the sum.  A
   mantissa that is zero, subnormal, infinite or NaN keeps its exponent field
   (its log is then -inf, a finite value, +inf or NaN, as the log of the
   norm itself would have been). */
double orc_split_exponent(double m, int32_t* ex) {
    uint64_t b;
    memcpy(&b, &m, 8);
    const int e = (int)((b >> 52) & 0x7ff);
    if (e == 0 || e == 0x7ff) return m;
    *ex += e - 1023;
    b = (b & 0x800fffffffffffffULL) | 0x3ff0000000000000ULL;
    memcpy(&m, &b, 8);
    return m;
}

double orc_log_scaled(double mant, int32_t ex) {
    const double dk = (double)ex;
    return fma(dk, LN2_HI, fma(dk, LN2_LO, orc_log(mant)));
}

/* cuda/tools/linalg.dp.cpp:5-10 */
static double dot_ref(int n, const double* a, const double* b) {
    double r = 0.0;
    for (int i = 0; i < n; ++i) r = r + a[i] * b[i];
    return r;
}

/* cuda/tools/linalg.dp.cpp:36-51: classical GS, coefficients against the
   original vector i, unnormalised basis, subtraction after the j loop */
void orc_cgs_ref(int n, int k, double* V) {
    double proj[ORC_MAX_DIM];
    for (int i = 1; i < k; ++i) {
        for (int m = 0; m < n; ++m) proj[m] = 0.0;
        double* vi = V + i * n;
        for (int j = 0; j < i; ++j) {
            const double* vj = V + j * n;
            double c = dot_ref(n, vj, vi) / dot_ref(n, vj, vj);
            for (int m = 0; m < n; ++m) proj[m] = proj[m] + c * vj[m];
        }
        for (int m = 0; m < n; ++m) vi[m] = vi[m] - proj[m];
    }
}

/* cuda/tools/linalg.dp.cpp:12-28 */
static double norm_ref(int n, const double* v) { return sqrt(dot_ref(n, v, v)); }

void orc_normalize_ref(int n, int k, double* V, double eps) {
    for (int j = 0; j < k; ++j) {
        double nrm = norm_ref(n, V + j * n);
        for (int i = 0; i < n; ++i) V[j * n + i] = (V[j * n + i] / nrm) * eps;
    }
}

/* cuda/tools/linalg.dp.cpp:78-128 (square case): CGS on the columns of a copy,
   normalise columns to 1, R = Q^T A; returns diag(R); Q is replaced by the
   orthonormal factor */
void orc_qr_ref(int n, double* Q, double* Rdiag) {
    double A[ORC_MAX_DIM * ORC_MAX_DIM], col[ORC_MAX_DIM], proj[ORC_MAX_DIM];
    memcpy(A, Q, sizeof(double) * n * n);
    for (int i = 1; i < n; ++i) {
        for (int m = 0; m < n; ++m) proj[m] = 0.0;
        for (int j = 0; j < i; ++j) {
            double djj = 0.0, dji = 0.0;
            for (int m = 0; m < n; ++m) dji = dji + Q[m * n + j] * Q[m * n + i];
            for (int m = 0; m < n; ++m) djj = djj + Q[m * n + j] * Q[m * n + j];
            double c = dji / djj;
            for (int m = 0; m < n; ++m) proj[m] = proj[m] + c * Q[m * n + j];
        }
        for (int m = 0; m < n; ++m) Q[m * n + i] = Q[m * n + i] - proj[m];
    }
    for (int i = 0; i < n; ++i) {
        double s = 0.0;
        for (int m = 0; m < n; ++m) s = s + Q[m * n + i] * Q[m * n + i];
        double nrm = sqrt(s);
        for (int m = 0; m < n; ++m) Q[m * n + i] = (Q[m * n + i] / nrm) * 1.0;
    }
    for (int i = 0; i < n; ++i) {
        double r = 0.0;
        for (int m = 0; m < n; ++m) r = r + Q[m * n + i] * A[m * n + i];
        Rdiag[i] = r;
    }
    (void)col;
}

/* ------------------------------------------------------------------------ */
/* step-size control: cuda/tools/dopri.dp.cpp:9-12,100-139                   */
/* ------------------------------------------------------------------------ */
double orc_err_norm(int n, int tree, const double* ev, const double* x, const double* xn,
                    double rtol, double atol) {
    double q[ORC_MAX_DIM];
    if (n <= 0) return 0.0;
    for (int i = 0; i < n; ++i) {
        double ax = fabs(x[i]), an = fabs(xn[i]);
        double sk = fma(rtol, fmax(ax, an), atol);
        q[i] = ev[i] / sk;
    }
    double s;
    if (tree) {
        for (int i = 0; i < n; ++i) q[i] = q[i] * q[i];
        s = tree_sum(n, q);
    } else {
        s = q[0] * q[0];
        for (int i = 1; i < n; ++i) s = fma(q[i], q[i], s);
    }
    return sqrt(s / (double)n);
}

/* The 3-D flows' error norm, squared, with one division:
     err^2 = (1/n) sum_i (ev_i / sk_i)^2 = sum_i ev_i^2 w_i / (n prod_j A_j),
   A_j = sk_j^2, w_i = prod_{j != i} A_j as prefix x suffix products (a factor
   1 is left out, which is exact).  For small n only (the products of n - 1
   squared scales must stay in range). */
double orc_err2_seq(int n, const double* ev, const double* x, const double* xn, double rtol,
                    double atol) {
    double A[ORC_MAX_DIM], pre[ORC_MAX_DIM], suf[ORC_MAX_DIM];
    if (n <= 0) return 0.0;
    for (int i = 0; i < n; ++i) {
        double ax = fabs(x[i]), an = fabs(xn[i]);
        double sk = fma(rtol, fmax(ax, an), atol);
        A[i] = sk * sk;
    }
    pre[0] = 1.0;
    for (int i = 1; i < n; ++i) pre[i] = (i == 1) ? A[0] : pre[i - 1] * A[i - 1];
    suf[n - 1] = 1.0;
    for (int i = n - 2; i >= 0; --i) suf[i] = (i == n - 2) ? A[n - 1] : A[i + 1] * suf[i + 1];
    double num = 0.0;
    for (int i = 0; i < n; ++i) {
        const double w = (i == 0) ? suf[0] : ((i == n - 1) ? pre[n - 1] : pre[i] * suf[i]);
        const double e2 = ev[i] * ev[i];
        num = (i == 0) ? e2 * w : fma(e2, w, num);
    }
    const double den = (double)n * ((n == 1) ? A[0] : pre[n - 1] * A[n - 1]);
    return num / den;
}

/* Step-control log and exp: the reductions of orc_log / orc_exp with shorter
   series (about 1e-10 relative, ample for a step-size factor).  ctrl_exp is
   only asked for the range that survives the factor limits below, so its
   argument is clamped to [-1.6, 2.5] first. */
double orc_ctrl_log(double x) {
    if (x != x) return x;
    if (x < 0.0) return NAN;
    if (x == 0.0) return -INFINITY;
    if (x == INFINITY) return INFINITY;
    uint64_t b = as_u64(x);
    int e = (int)((b >> 52) & 0x7ff), k;
    if (e == 0) {
        x = x * 18014398509481984.0;
        b = as_u64(x);
        e = (int)((b >> 52) & 0x7ff);
        k = e - 1023 - 54;
    } else {
        k = e - 1023;
    }
    double m = as_f64((b & 0x000fffffffffffffULL) | 0x3ff0000000000000ULL);
    if (m > 1.4142135623730951) {
        m = m * 0.5;
        k = k + 1;
    }
    const double f = m - 1.0, s = f / (2.0 + f), z = s * s;
    double r = 2.0 / 11.0;
    r = fma(r, z, 2.0 / 9.0);
    r = fma(r, z, 2.0 / 7.0);
    r = fma(r, z, 2.0 / 5.0);
    r = fma(r, z, 2.0 / 3.0);
    r = r * z;
    const double lg = f - s * (f - r), dk = (double)k;
    return fma(dk, LN2_HI, fma(dk, LN2_LO, lg));
}

double orc_ctrl_exp(double x) {
    x = x < -1.6 ? -1.6 : x;
    x = x > 2.5 ? 2.5 : x;
    const double kd = floor(x * 1.4426950408889634 + 0.5);
    const double r = fma(-kd, LN2_LO, fma(-kd, LN2_HI, x));
    double p = 1.0 / 40320.0;
    p = fma(p, r, 1.0 / 5040.0);
    p = fma(p, r, 1.0 / 720.0);
    p = fma(p, r, 1.0 / 120.0);
    p = fma(p, r, 1.0 / 24.0);
    p = fma(p, r, 1.0 / 6.0);
    p = fma(p, r, 0.5);
    p = fma(p, r, 1.0);
    p = fma(p, r, 1.0);
    return p * as_f64((uint64_t)(1023 + (int)kd) << 52);
}

typedef struct {
    double log_fac_old; /* log of facOld, floored at 1e-4 */
    int rejected;
} ctrl_t;

static const double CTRL_SAFE = 0.9, CTRL_BETA = 0.04;

static double ctrl_expo(int integ) {
    return integ == ORC_INT_DVERK ? (1.0 / 6.0 - 0.75 * 0.04) : (1.0 / 5.0 - 0.75 * 0.04);
}

/* dopriStep's PI rule (cuda/tools/dopri.dp.cpp:109-139): fac = err^expo /
   facOld^beta, limited to [1/10, 1/0.2] after division by the safety factor,
   h_new = h / fac; after a rejection h_new = h / min(5, err^expo / 0.9).  Here
   in the equivalent form h_new = h * r, r = clamp(0.9 * exp(-(expo*lerr -
   beta*log facOld)), 0.2, 10): one exponential and no division.  lerr is
   orc_ctrl_log(err).  Returns 1 on accept; *h_next is the next attempt. */
/* ok: err <= 1; above_floor: err > 1e-4; lerr = log(err) */
static int ctrl_update(ctrl_t* c, int integ, int ok, int above_floor, double lerr, double log_floor,
                       double h_try, double h_max, double* h_next) {
    const double expo = ctrl_expo(integ);
    if (ok) {
        const double e = fma(-CTRL_BETA, c->log_fac_old, expo * lerr);
        double r = CTRL_SAFE * orc_ctrl_exp(-e);
        r = r < 0.2 ? 0.2 : r;
        r = r > 10.0 ? 10.0 : r;
        double hn = h_try * r;
        c->log_fac_old = above_floor ? lerr : log_floor;
        if (fabs(hn) > h_max) hn = h_max;
        if (c->rejected) hn = fabs(hn) < fabs(h_try) ? fabs(hn) : fabs(h_try);
        c->rejected = 0;
        *h_next = hn;
        return 1;
    }
    double r = CTRL_SAFE * orc_ctrl_exp(-(expo * lerr));
    r = r < 0.2 ? 0.2 : r;
    c->rejected = 1;
    *h_next = h_try * r;
    return 0;
}

/* One decision of the step controller on an error norm `err` (not its
   log), exported so that tests can hold it against the reference's pow-form
   rule (cuda/tools/dopri.dp.cpp:9-12,27-29,109-139): the controller state is
   facOld as its log and the rejected-last-step flag, both updated in place.
   Returns 1 on accept; *h_next is the next attempt. */
int orc_ctrl_update(int integ, double err, double* log_fac_old, int* rejected, double h_try,
                    double h_max, double* h_next) {
    ctrl_t c = {*log_fac_old, *rejected};
    const int acc = ctrl_update(&c, integ, err <= 1.0, err > 1e-4, orc_ctrl_log(err),
                                orc_ctrl_log(1e-4), h_try, h_max, h_next);
    *log_fac_old = c.log_fac_old;
    *rejected = c.rejected;
    return acc;
}

/* ------------------------------------------------------------------------ */
/* points                                                                    */
/* ------------------------------------------------------------------------ */
/* tasklyapunov.cpp:45-63: p1 outer, p = start + i*delta, other parameters from
   initial_model_params */
void orc_point_params(const orc_desc* d, int64_t idx, const double* params_in, double* p) {
    for (int j = 0; j < ORC_MAX_PARAMS; ++j) p[j] = 0.0;
    if (params_in) {
        for (int j = 0; j < d->n_params; ++j) p[j] = params_in[idx * d->n_params + j];
        return;
    }
    int64_t i1 = idx / d->n2, i2 = idx % d->n2;
    double v1 = d->start1 + (double)i1 * d->delta1;
    double v2 = d->start2 + (double)i2 * d->delta2;
    for (int j = 0; j < d->n_params; ++j) {
        if (j == d->sweep_idx1) p[j] = v1;
        else if (j == d->sweep_idx2) p[j] = v2;
        else p[j] = d->base_params[j];
    }
}

void orc_point_init(const orc_desc* d, int64_t idx, const double* p, const double* inits_in,
                    double* x) {
    const int n = d->dim;
    if (inits_in) {
        for (int i = 0; i < n; ++i) x[i] = inits_in[idx * n + i];
        return;
    }
    for (int i = 0; i < n; ++i) {
        double x0 = d->init_state[i];
        if (d->init_param_offset >= 0) x0 = x0 + p[d->init_param_offset];
        double dx = 0.0;
        if (d->init_kind == ORC_INIT_UNIFORM) {
            double u = orc_uniform(d->seed, idx, i);
            dx = d->init_scale * (2.0 * u - 1.0);
        } else if (d->init_kind == ORC_INIT_NORMAL) {
            /* Irwin-Hall: sum of 12 uniforms minus 6 has mean 0, variance 1 */
            double z = 0.0;
            for (int j = 0; j < 12; ++j) z = z + orc_uniform(d->seed, idx, 32 + i * 12 + j);
            dx = d->init_scale * (z - 6.0);
        }
        x[i] = x0 + dx;
    }
}

static int is_tree_model(int model) { return model == ORC_MODEL_L96; }

static int escaped(int n, const double* x, double radius) {
    for (int i = 0; i < n; ++i)
        if (!(fabs(x[i]) <= radius)) return 1;
    return 0;
}

static void kahan_add(double* s, double* c, double v) {
    double y = v - *c;
    double t = *s + y;
    *c = (t - *s) - y;
    *s = t;
}

typedef struct {
    double sum[ORC_MAX_DIM], comp[ORC_MAX_DIM], mant[ORC_MAX_DIM];
    int32_t ex[ORC_MAX_DIM];
    int cnt;
} lnacc_t;

static void lnacc_init(lnacc_t* a, int k) {
    for (int r = 0; r < k; ++r) {
        a->sum[r] = a->comp[r] = 0.0;
        a->mant[r] = 1.0;
        a->ex[r] = 0;
    }
    a->cnt = 0;
}

/* one renormalisation of the averaging phase; `last` closes the window */
static void lnacc_push(lnacc_t* a, int k, const double* nrm, int last) {
    for (int r = 0; r < k; ++r) a->mant[r] = orc_split_exponent(a->mant[r] * nrm[r], &a->ex[r]);
    if (++a->cnt == ORC_LOG_WINDOW || last) {
        for (int r = 0; r < k; ++r) {
            kahan_add(&a->sum[r], &a->comp[r], orc_log_scaled(a->mant[r], a->ex[r]));
            a->mant[r] = 1.0;
            a->ex[r] = 0;
        }
        a->cnt = 0;
    }
}

/* added method: coupled variational RK, MGS on exact time boundaries */
static void point_var(const orc_desc* d, const double* p, double* x, double* lam, uint32_t* status,
                      uint32_t* n_acc_out, uint32_t* n_rej_out) {
    /* ORC_METHOD_VAR_TRACE: n - 1 vectors; the last exponent is the averaged
       integral of tr J over the averaging phase minus the others (Liouville) */
    const int trace = d->method == ORC_METHOD_VAR_TRACE;
    const int n = d->dim, k = trace ? d->dim - 1 : d->n_exps, integ = d->integrator;
    double tri = 0.0, trs = 0.0, trc = 0.0;
    const int tree = is_tree_model(d->model);
    const int mixed = d->precision == ORC_PREC_MIXED;
    const int64_t n_tr = llround((d->t_transient + d->t_slave) / d->t_renorm);
    const int64_t n_av = llround(d->t_average / d->t_renorm);
    const int64_t m_total = n_tr + n_av;
    double V[ORC_MAX_DIM * ORC_MAX_DIM];
    double nrm[ORC_MAX_DIM];
    lnacc_t lna;
    double xn[ORC_MAX_DIM], ev[ORC_MAX_DIM], sx[8 * ORC_MAX_DIM];
    for (int r = 0; r < k; ++r)
        for (int i = 0; i < n; ++i) {
            V[r * n + i] = (r == i) ? 1.0 : 0.0;
        }
    lnacc_init(&lna, k);
    uint32_t st = ORC_OK;
    uint64_t n_acc = 0, n_rej = 0;
    double t_avg_exact;

/* MGS dot products are sequential for every model (the GPU stages the
   deviation matrix in LDS and gives each vector to one lane); the step's
   error norm uses the butterfly for Lorenz-96 */
#define ORC_RENORM() mgs_norms(n, k, 0, V, nrm)

    if (!d->adaptive) {
        const int64_t R = llround(d->t_renorm / d->h);
        const double h = d->h;
        t_avg_exact = (double)n_av * (double)R * h;
        for (int64_t m = 0; m < m_total && st == ORC_OK; ++m) {
            for (int64_t s = 0; s < R; ++s) {
                orc_rk_step(d->model, integ, n, x, p, h, xn, NULL, sx);
                for (int r = 0; r < k; ++r) {
                    if (mixed && d->model == ORC_MODEL_L96) l96_tangent_step_f32(integ, n, sx, p, h, V + r * n);
                    else if (mixed) orc_tangent_step_f32(d->model, integ, n, sx, p, h, V + r * n, 0);
                    else orc_tangent_step(d->model, integ, n, sx, p, h, V + r * n);
                }
                if (trace) tri = tri + orc_trace_step(d->model, integ, n, sx, p, h);
                memcpy(x, xn, sizeof(double) * n);
                ++n_acc;
            }
            for (int i = 0; i < n; ++i)
                if (!isfinite(x[i])) st = ORC_NONFINITE;
            if (st == ORC_OK && escaped(n, x, d->escape_radius)) st = ORC_ESCAPED;
            if (st != ORC_OK) break;
            ORC_RENORM();
            if (m + 1 > n_tr) {
                lnacc_push(&lna, k, nrm, m + 1 == m_total);
                if (trace) kahan_add(&trs, &trc, tri);
            }
            tri = 0.0;
        }
    } else {
        t_avg_exact = (double)n_av * d->t_renorm;
        const double log_floor = orc_ctrl_log(1e-4);
        ctrl_t c = {log_floor, 0};
        double t = 0.0, h = d->h;
        int64_t m = 0;
        double t_next = (double)(m + 1) * d->t_renorm;
        for (;;) {
            if ((int64_t)(n_acc + n_rej) >= d->max_steps) { st = ORC_MAXSTEPS; break; }
            double h_try = h;
            int clipped = 0;
            const double jn = orc_jac_norm(d->model, n, x, p);
            if (h_try * jn > 1.0) h_try = 1.0 / jn;
            /* land on the boundary when it is within 1/16 of the step, so no
               round-off sliver is left before it */
            if (t_next - t <= 1.0625 * h_try) { h_try = t_next - t; clipped = 1; }
            if (h_try < d->h_min) { st = ORC_STEP_UNDERFLOW; break; }
            orc_rk_step(d->model, integ, n, x, p, h_try, xn, ev, sx);
            /* Lorenz-96: the norm (butterfly sum); the 3-D flows: its square
               (one division), compared with 1 and 1e-8, log halved */
            int ok, above;
            double lerr;
            if (tree) {
                const double err = orc_err_norm(n, tree, ev, x, xn, d->rtol, d->atol);
                if (!(err < INFINITY)) { st = ORC_NONFINITE; break; }
                ok = err <= 1.0;
                above = err > 1e-4;
                lerr = orc_ctrl_log(err);
            } else {
                const double e2 = orc_err2_seq(n, ev, x, xn, d->rtol, d->atol);
                if (!(e2 < INFINITY)) { st = ORC_NONFINITE; break; }
                ok = e2 <= 1.0;
                above = e2 > 1e-8;
                lerr = 0.5 * orc_ctrl_log(e2);
            }
            double h_next;
            int acc = ctrl_update(&c, integ, ok, above, lerr, log_floor, h_try, d->h_max, &h_next);
            h = h_next;
            if (!acc) { ++n_rej; continue; }
            for (int r = 0; r < k; ++r) {
                if (mixed && d->model == ORC_MODEL_L96) l96_tangent_step_f32(integ, n, sx, p, h_try, V + r * n);
                else if (mixed) orc_tangent_step_f32(d->model, integ, n, sx, p, h_try, V + r * n, 1);
                else orc_tangent_step(d->model, integ, n, sx, p, h_try, V + r * n);
            }
            if (trace) tri = tri + orc_trace_step(d->model, integ, n, sx, p, h_try);
            memcpy(x, xn, sizeof(double) * n);
            t = clipped ? t_next : t + h_try;
            ++n_acc;
            if (escaped(n, x, d->escape_radius)) { st = ORC_ESCAPED; break; }
            if (clipped) {
                ORC_RENORM();
                ++m;
                if (m > n_tr) {
                    lnacc_push(&lna, k, nrm, m == m_total);
                    if (trace) kahan_add(&trs, &trc, tri);
                }
                tri = 0.0;
                if (m == m_total) break;
                t_next = (double)(m + 1) * d->t_renorm;
            }
        }
    }
#undef ORC_RENORM
    for (int r = 0; r < k; ++r) lam[r] = (st == ORC_OK) ? lna.sum[r] / t_avg_exact : NAN;
    if (trace) {
        double closing = trs;
        for (int r = 0; r < k; ++r) closing = closing - lna.sum[r];
        lam[k] = (st == ORC_OK) ? closing / t_avg_exact : NAN;
    }
    *status = st;
    *n_acc_out = (uint32_t)n_acc;
    *n_rej_out = (uint32_t)n_rej;
}

/* reference VAR mode: cuda/lyapunov.dp.cpp:138-180,224-225 */
static void point_var_ref(const orc_desc* d, const double* p, double* x, double* lam,
                          uint64_t* steps) {
    const int n = d->dim, k = d->n_exps;
    const double h = d->h, eps = d->eps;
    const double add_skip = d->t_renorm / d->h; /* tasklyapunov.cpp:65 */
    const int64_t n_tr = orc_loop_count(0.0, d->t_transient, h);
    const int64_t n_sl = orc_loop_count(0.0, d->t_slave, h);
    const int64_t n_av = orc_loop_count(0.0, d->t_average, h);
    double V[ORC_MAX_DIM * ORC_MAX_DIM], xn[ORC_MAX_DIM], L[ORC_MAX_DIM];
    for (int r = 0; r < k; ++r) L[r] = 0.0;
    for (int64_t s = 0; s < n_tr; ++s) {
        orc_rk_step(d->model, ORC_INT_DVERK, n, x, p, h, xn, NULL, NULL);
        memcpy(x, xn, sizeof(double) * n);
    }
    for (int q = 0; q < k * n; ++q) V[q] = 0.0;
    for (int r = 0; r < k; ++r) V[r * n + r] = V[r * n + r] + eps;
    for (int64_t s = 0; s < n_sl; ++s) {
        for (int r = 0; r < k; ++r) orc_frozen_tangent_step(d->model, n, x, p, h, V + r * n);
        orc_rk_step(d->model, ORC_INT_DVERK, n, x, p, h, xn, NULL, NULL);
        memcpy(x, xn, sizeof(double) * n);
        orc_cgs_ref(n, k, V);
        orc_normalize_ref(n, k, V, eps);
    }
    int32_t skip = 0;
    for (int64_t s = 0; s < n_av; ++s) {
        ++skip;
        for (int r = 0; r < k; ++r) orc_frozen_tangent_step(d->model, n, x, p, h, V + r * n);
        orc_rk_step(d->model, ORC_INT_DVERK, n, x, p, h, xn, NULL, NULL);
        memcpy(x, xn, sizeof(double) * n);
        if ((double)skip == add_skip) {
            orc_cgs_ref(n, k, V);
            for (int r = 0; r < k; ++r) L[r] = L[r] + orc_log(norm_ref(n, V + r * n) / eps);
            orc_normalize_ref(n, k, V, eps);
            skip = 0;
        }
    }
    for (int r = 0; r < k; ++r) lam[r] = L[r] / d->t_average;
    *steps = (uint64_t)(n_tr + n_sl + n_av);
}

/* reference QR mode: cuda/lyapunov.dp.cpp:229-259 (k == n) */
static void point_qr_ref(const orc_desc* d, const double* p, double* x, double* lam,
                         uint64_t* steps) {
    const int n = d->dim;
    const double h = d->h;
    const int64_t n_tr = orc_loop_count(0.0, d->t_transient, h);
    const int64_t n_av = orc_loop_count(d->t_transient, d->t_transient + d->t_average, h);
    double Q[ORC_MAX_DIM * ORC_MAX_DIM], col[ORC_MAX_DIM], xn[ORC_MAX_DIM], L[ORC_MAX_DIM],
        Rd[ORC_MAX_DIM];
    for (int r = 0; r < n; ++r) L[r] = 0.0;
    for (int64_t s = 0; s < n_tr; ++s) {
        orc_rk_step(d->model, ORC_INT_DVERK, n, x, p, h, xn, NULL, NULL);
        memcpy(x, xn, sizeof(double) * n);
    }
    for (int q = 0; q < n * n; ++q) Q[q] = 0.0;
    for (int i = 0; i < n; ++i) Q[i * n + i] = 1.0;
    for (int64_t s = 0; s < n_av; ++s) {
        orc_rk_step(d->model, ORC_INT_DVERK, n, x, p, h, xn, NULL, NULL);
        memcpy(x, xn, sizeof(double) * n);
        for (int c = 0; c < n; ++c) {
            for (int m = 0; m < n; ++m) col[m] = Q[m * n + c];
            orc_frozen_tangent_step(d->model, n, x, p, h, col);
            for (int m = 0; m < n; ++m) Q[m * n + c] = col[m];
        }
        orc_qr_ref(n, Q, Rd);
        for (int i = 0; i < d->n_exps; ++i) L[i] = L[i] + orc_log(fabs(Rd[i]));
    }
    for (int i = 0; i < d->n_exps; ++i) lam[i] = L[i] / d->t_average;
    *steps = (uint64_t)(n_tr + n_av);
}

/* reference FD (two-trajectory) mode: cuda/lyapunov.dp.cpp:10-68,127-128 */
static void point_fd_ref(const orc_desc* d, const double* p, double* x, double* lam,
                         uint64_t* steps) {
    const int n = d->dim, k = d->n_exps;
    const double h = d->h, eps = d->eps;
    const double add_skip = d->t_renorm / d->h;
    const int64_t n_tr = orc_loop_count(0.0, d->t_transient, h);
    const int64_t n_sl = orc_loop_count(0.0, d->t_slave, h);
    const int64_t n_av = orc_loop_count(0.0, d->t_average, h);
    double S[ORC_MAX_DIM * ORC_MAX_DIM], xn[ORC_MAX_DIM], L[ORC_MAX_DIM];
    for (int r = 0; r < k; ++r) L[r] = 0.0;
    for (int64_t s = 0; s < n_tr; ++s) {
        orc_rk_step(d->model, ORC_INT_DVERK, n, x, p, h, xn, NULL, NULL);
        memcpy(x, xn, sizeof(double) * n);
    }
    for (int r = 0; r < k; ++r) {
        memcpy(S + r * n, x, sizeof(double) * n);
        S[r * n + r] = S[r * n + r] + eps;
    }
    for (int64_t s = 0; s < n_sl; ++s) {
        orc_rk_step(d->model, ORC_INT_DVERK, n, x, p, h, xn, NULL, NULL);
        memcpy(x, xn, sizeof(double) * n);
        for (int r = 0; r < k; ++r) {
            orc_rk_step(d->model, ORC_INT_DVERK, n, S + r * n, p, h, xn, NULL, NULL);
            for (int i = 0; i < n; ++i) S[r * n + i] = xn[i] - x[i];
        }
        orc_cgs_ref(n, k, S);
        orc_normalize_ref(n, k, S, eps);
        for (int r = 0; r < k; ++r)
            for (int i = 0; i < n; ++i) S[r * n + i] = S[r * n + i] + x[i];
    }
    int32_t skip = 0;
    for (int64_t s = 0; s < n_av; ++s) {
        ++skip;
        orc_rk_step(d->model, ORC_INT_DVERK, n, x, p, h, xn, NULL, NULL);
        memcpy(x, xn, sizeof(double) * n);
        for (int r = 0; r < k; ++r) {
            orc_rk_step(d->model, ORC_INT_DVERK, n, S + r * n, p, h, xn, NULL, NULL);
            memcpy(S + r * n, xn, sizeof(double) * n);
        }
        if ((double)skip == add_skip) {
            for (int r = 0; r < k; ++r)
                for (int i = 0; i < n; ++i) S[r * n + i] = S[r * n + i] - x[i];
            orc_cgs_ref(n, k, S);
            for (int r = 0; r < k; ++r) L[r] = L[r] + orc_log(norm_ref(n, S + r * n) / eps);
            orc_normalize_ref(n, k, S, eps);
            for (int r = 0; r < k; ++r)
                for (int i = 0; i < n; ++i) S[r * n + i] = S[r * n + i] + x[i];
            skip = 0;
        }
    }
    for (int r = 0; r < k; ++r) lam[r] = L[r] / d->t_average;
    *steps = (uint64_t)(n_tr + n_sl + n_av);
}

/* ------------------------------------------------------------------------ */
/* maps (ModelType::map): one iteration is one "step", time counts iterations */
/* ------------------------------------------------------------------------ */
static void map_step(int model, int n, double* x, const double* p) {
    double xn[ORC_MAX_DIM];
    orc_rhs(model, n, x, p, xn);
    memcpy(x, xn, sizeof(double) * n);
}

/* V <- J(x) V for k row vectors (StepMAPVAR, cuda/tools/mapstep.dp.cpp:12-18);
   the mixed form takes the product in binary32 */
static void map_tangent(int model, int n, int k, const double* x, const double* p, int mixed,
                        double* V) {
    double out[ORC_MAX_DIM];
    float vf[ORC_MAX_DIM], of[ORC_MAX_DIM];
    for (int r = 0; r < k; ++r) {
        if (mixed) {
            for (int i = 0; i < n; ++i) vf[i] = (float)V[r * n + i];
            orc_jvp_f32(model, n, x, p, vf, of);
            for (int i = 0; i < n; ++i) V[r * n + i] = (double)of[i];
        } else {
            orc_jvp(model, n, x, p, V + r * n, out);
            memcpy(V + r * n, out, sizeof(double) * n);
        }
    }
}

/* the sweep's method for maps: R = t_renorm iterations per renormalisation
   (MGS, windowed log-norm accumulation), the transient from the start */
static void point_map_var(const orc_desc* d, const double* p, double* x, double* lam,
                          uint32_t* status, uint32_t* n_acc_out) {
    const int n = d->dim, k = d->n_exps;
    const int mixed = d->precision == ORC_PREC_MIXED;
    const int64_t R = llround(d->t_renorm / d->h);
    const int64_t n_tr = llround((d->t_transient + d->t_slave) / d->t_renorm);
    const int64_t n_av = llround(d->t_average / d->t_renorm);
    double V[ORC_MAX_DIM * ORC_MAX_DIM], nrm[ORC_MAX_DIM];
    lnacc_t lna;
    lnacc_init(&lna, k);
    for (int r = 0; r < k; ++r)
        for (int i = 0; i < n; ++i) V[r * n + i] = (r == i) ? 1.0 : 0.0;
    uint32_t st = ORC_OK;
    uint64_t n_acc = 0;
    /* a diverging map orbit overflows within a few iterations (x ~ x^2), so
       the escape test runs after every iteration */
    for (int64_t m = 0; m < n_tr + n_av && st == ORC_OK; ++m) {
        for (int64_t s = 0; s < R; ++s) {
            map_tangent(d->model, n, k, x, p, mixed, V);
            map_step(d->model, n, x, p);
            ++n_acc;
            if (escaped(n, x, d->escape_radius)) { st = ORC_ESCAPED; break; }
        }
        if (st != ORC_OK) break;
        mgs_norms(n, k, 0, V, nrm);
        if (m + 1 > n_tr) lnacc_push(&lna, k, nrm, m + 1 == n_tr + n_av);
    }
    const double t_avg = (double)n_av * (double)R * d->h;
    for (int r = 0; r < k; ++r) lam[r] = (st == ORC_OK) ? lna.sum[r] / t_avg : NAN;
    *status = st;
    *n_acc_out = (uint32_t)n_acc;
}

/* reference VAR mode for maps: cuda/lyapunov.dp.cpp:182-222,224-225 */
static void point_map_var_ref(const orc_desc* d, const double* p, double* x, double* lam,
                              uint64_t* steps) {
    const int n = d->dim, k = d->n_exps;
    const double eps = d->eps;
    const int64_t n_tr = orc_loop_count(0.0, d->t_transient, 1.0);
    const int64_t n_sl = orc_loop_count(0.0, d->t_slave, 1.0);
    const int64_t n_av = orc_loop_count(0.0, d->t_average, 1.0);
    double V[ORC_MAX_DIM * ORC_MAX_DIM], L[ORC_MAX_DIM];
    for (int r = 0; r < k; ++r) L[r] = 0.0;
    for (int64_t s = 0; s < n_tr; ++s) map_step(d->model, n, x, p);
    for (int q = 0; q < k * n; ++q) V[q] = 0.0;
    for (int r = 0; r < k; ++r) V[r * n + r] = V[r * n + r] + eps;
    for (int64_t s = 0; s < n_sl; ++s) {
        map_tangent(d->model, n, k, x, p, 0, V);
        map_step(d->model, n, x, p);
        orc_cgs_ref(n, k, V);
        orc_normalize_ref(n, k, V, eps);
    }
    for (int64_t s = 0; s < n_av; ++s) {
        map_tangent(d->model, n, k, x, p, 0, V);
        map_step(d->model, n, x, p);
        orc_cgs_ref(n, k, V);
        for (int r = 0; r < k; ++r) L[r] = L[r] + orc_log(norm_ref(n, V + r * n) / eps);
        orc_normalize_ref(n, k, V, eps);
    }
    for (int r = 0; r < k; ++r) lam[r] = L[r] / d->t_average;
    *steps = (uint64_t)(n_tr + n_sl + n_av);
}

/* reference FD (two-trajectory) mode for maps: cuda/lyapunov.dp.cpp:70-122,127-128 */
static void point_map_fd_ref(const orc_desc* d, const double* p, double* x, double* lam,
                             uint64_t* steps) {
    const int n = d->dim, k = d->n_exps;
    const double eps = d->eps;
    const int64_t n_tr = orc_loop_count(0.0, d->t_transient, 1.0);
    const int64_t n_sl = orc_loop_count(0.0, d->t_slave, 1.0);
    const int64_t n_av = orc_loop_count(0.0, d->t_average, 1.0);
    double S[ORC_MAX_DIM * ORC_MAX_DIM], L[ORC_MAX_DIM];
    for (int r = 0; r < k; ++r) L[r] = 0.0;
    for (int64_t s = 0; s < n_tr; ++s) map_step(d->model, n, x, p);
    for (int r = 0; r < k; ++r) {
        memcpy(S + r * n, x, sizeof(double) * n);
        S[r * n + r] = S[r * n + r] + eps;
    }
    for (int64_t s = 0; s < n_sl; ++s) {
        map_step(d->model, n, x, p);
        for (int r = 0; r < k; ++r) {
            map_step(d->model, n, S + r * n, p);
            for (int i = 0; i < n; ++i) S[r * n + i] = S[r * n + i] - x[i];
        }
        orc_cgs_ref(n, k, S);
        orc_normalize_ref(n, k, S, eps);
        for (int r = 0; r < k; ++r)
            for (int i = 0; i < n; ++i) S[r * n + i] = S[r * n + i] + x[i];
    }
    for (int64_t s = 0; s < n_av; ++s) {
        map_step(d->model, n, x, p);
        for (int r = 0; r < k; ++r) map_step(d->model, n, S + r * n, p);
        for (int r = 0; r < k; ++r)
            for (int i = 0; i < n; ++i) S[r * n + i] = S[r * n + i] - x[i];
        orc_cgs_ref(n, k, S);
        for (int r = 0; r < k; ++r) L[r] = L[r] + orc_log(norm_ref(n, S + r * n) / eps);
        orc_normalize_ref(n, k, S, eps);
        for (int r = 0; r < k; ++r)
            for (int i = 0; i < n; ++i) S[r * n + i] = S[r * n + i] + x[i];
    }
    for (int r = 0; r < k; ++r) lam[r] = L[r] / d->t_average;
    *steps = (uint64_t)(n_tr + n_sl + n_av);
}

void orc_point(const orc_desc* d, int64_t idx, const double* params_in, const double* inits_in,
               double* lam, uint32_t* status, uint32_t* n_acc, uint32_t* n_rej) {
    double p[ORC_MAX_PARAMS], x[ORC_MAX_DIM];
    orc_point_params(d, idx, params_in, p);
    orc_point_init(d, idx, p, inits_in, x);
    uint64_t steps = 0;
    *n_rej = 0;
    if (orc_is_map(d->model)) {
        switch (d->method) {
        case ORC_METHOD_VAR:
            point_map_var(d, p, x, lam, status, n_acc);
            return;
        case ORC_METHOD_VAR_REF: point_map_var_ref(d, p, x, lam, &steps); break;
        case ORC_METHOD_FD_REF: point_map_fd_ref(d, p, x, lam, &steps); break;
        default:
            for (int r = 0; r < d->n_exps; ++r) lam[r] = NAN;
            *status = ORC_NONFINITE;
            *n_acc = 0;
            return;
        }
        uint32_t st = ORC_OK;
        for (int r = 0; r < d->n_exps; ++r)
            if (!isfinite(lam[r])) st = ORC_NONFINITE;
        *status = st;
        *n_acc = (uint32_t)steps;
        return;
    }
    switch (d->method) {
    case ORC_METHOD_VAR:
    case ORC_METHOD_VAR_TRACE:
        point_var(d, p, x, lam, status, n_acc, n_rej);
        return;
    case ORC_METHOD_VAR_REF: point_var_ref(d, p, x, lam, &steps); break;
    case ORC_METHOD_QR_REF: point_qr_ref(d, p, x, lam, &steps); break;
    case ORC_METHOD_FD_REF: point_fd_ref(d, p, x, lam, &steps); break;
    default:
        for (int r = 0; r < d->n_exps; ++r) lam[r] = NAN;
        *status = ORC_NONFINITE;
        *n_acc = 0;
        return;
    }
    uint32_t st = ORC_OK;
    for (int r = 0; r < d->n_exps; ++r)
        if (!isfinite(lam[r])) st = ORC_NONFINITE;
    *status = st;
    *n_acc = (uint32_t)steps;
}

void orc_sweep(const orc_desc* d, int64_t first, int64_t count, const int64_t* idx_list,
               const double* params_in, const double* inits_in, double* lam, uint32_t* status,
               uint32_t* n_acc, uint32_t* n_rej, int n_threads) {
#ifdef _OPENMP
    if (n_threads > 0) omp_set_num_threads(n_threads);
#pragma omp parallel for schedule(dynamic, 1)
#endif
    for (int64_t i = 0; i < count; ++i) {
        int64_t idx = idx_list ? idx_list[i] : first + i;
        uint32_t st, na, nr;
        orc_point(d, idx, params_in, inits_in, lam + i * d->n_exps, &st, &na, &nr);
        if (status) status[i] = st;
        if (n_acc) n_acc[i] = na;
        if (n_rej) n_rej[i] = nr;
    }
    (void)n_threads;
}
