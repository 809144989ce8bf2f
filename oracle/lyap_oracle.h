/*
 * CPU FP64 oracle for the Lyapunov-spectrum sweep.  TEST INFRASTRUCTURE ONLY:
 * only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may load
 * it, and only as the checker.  The product (chaos_oneapi_amd) never calls it.
 *
 * It restates, in plain C, the per-point algorithms of the reference:
 *   - reference VAR mode  : cuda/lyapunov.dp.cpp:133-226 (deviceLESpectrumVARDverk_)
 *   - reference QR mode   : cuda/lyapunov.dp.cpp:229-259 (deviceLESpectrumQR_)
 *   - reference FD mode   : cuda/lyapunov.dp.cpp:5-129   (deviceLESpectrumDverk_)
 *   - the map branches of the VAR and FD modes: cuda/lyapunov.dp.cpp:182-222 and
 *     70-122, StepMAP/StepMAPVAR (cuda/tools/mapstep.dp.cpp:5-18); the
 *     reference registers no map model, so the two maps here (Henon,
 *     logistic) are this build's
 *   - DVERK 6(5) tableau  : cuda/tools/dverk.dp.cpp:7-80
 *   - Dopri5 + controller : cuda/tools/dopri.dp.cpp:7-140
 *   - CGS / normalise / QR: cuda/tools/linalg.dp.cpp:5-128
 *   - shimizu RHS/var     : cuda/models/flows/shimizu.dp.cpp:15-52
 *   - grid rule           : tasklyapunov.cpp:45-63
 * plus the method the sweep adds on top of the reference (coupled variational
 * RK with the Jacobian at every stage state, modified Gram-Schmidt on exact
 * time boundaries, adaptive DVERK/Dopri5, FP32 tangent stages with FP64
 * renormalisation, norm products over windows of 64 renormalisations
 * (ORC_LOG_WINDOW) and Kahan-compensated sums of the windows' logs; for the
 * 3-D flows optionally N - 1 vectors with the last exponent from the
 * integrated Jacobian trace, ORC_METHOD_VAR_TRACE).
 *
 * The descriptor below has the same memory layout as lyap_sweep_desc in
 * include/chaos_lyap.h so one Python dictionary can fill both.
 */
#ifndef LYAP_ORACLE_H
#define LYAP_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define ORC_MAX_DIM 32
#define ORC_MAX_PARAMS 8

enum { ORC_MODEL_SHIMIZU = 0, ORC_MODEL_LORENZ = 1, ORC_MODEL_ROSSLER = 2, ORC_MODEL_L96 = 3,
       ORC_MODEL_HENON = 4, ORC_MODEL_LOGISTIC = 5 };
/* maps: orc_rhs returns the next state f(x) and orc_jvp the product J(x) v */
int orc_is_map(int model);
/* the step controller's log and exp (shorter series; exp clamps its argument
   to [-1.6, 2.5], the range that survives the step-factor limits) */
double orc_ctrl_log(double x);
double orc_ctrl_exp(double x);
/* one step-controller decision on error norm err (dopriStep's PI rule,
   cuda/tools/dopri.dp.cpp:109-139, in the one-exponential form); the state
   (log facOld, rejected flag) is updated in place; returns 1 on accept */
int orc_ctrl_update(int integ, double err, double* log_fac_old, int* rejected, double h_try,
                    double h_max, double* h_next);
enum { ORC_METHOD_VAR = 0, ORC_METHOD_VAR_REF = 1, ORC_METHOD_QR_REF = 2, ORC_METHOD_FD_REF = 3,
       ORC_METHOD_VAR_TRACE = 4 };
enum { ORC_INT_DVERK = 0, ORC_INT_DOPRI5 = 1 };
enum { ORC_PREC_FP64 = 0, ORC_PREC_MIXED = 1 };
enum { ORC_INIT_NONE = 0, ORC_INIT_UNIFORM = 1, ORC_INIT_NORMAL = 2 };
enum { ORC_OK = 0, ORC_NONFINITE = 1, ORC_STEP_UNDERFLOW = 2, ORC_MAXSTEPS = 3, ORC_ESCAPED = 4 };

typedef struct orc_desc {
    int32_t model, dim, n_exps, n_params;
    int32_t method, integrator, adaptive, precision;
    double h, rtol, atol, h_min, h_max, eps;
    double t_transient, t_slave, t_average, t_renorm;
    /* reference modes: loop counts of `for (double t = t0; t < T; t += h)` */
    int64_t ref_steps_transient, ref_steps_slave, ref_steps_average;
    double add_skip_steps;
    /* added method: integer schedule */
    int64_t steps_per_renorm;    /* fixed step only */
    int64_t n_renorm_transient, n_renorm_average;
    int64_t max_steps;
    double escape_radius;
    /* grid (p1 outer): p = start + i * delta */
    int32_t sweep_idx1, sweep_idx2, n1, n2;
    double start1, delta1, start2, delta2;
    double base_params[ORC_MAX_PARAMS];
    /* initial state */
    double init_state[ORC_MAX_DIM];
    int32_t init_kind, init_param_offset;
    double init_scale;
    uint64_t seed;
} orc_desc;

/* deterministic elementary functions (same bits on every IEEE machine) */
double orc_log(double x);
double orc_exp(double x);
double orc_pow(double x, double y);

/* counter-based generator: uniform in [0,1) from (seed, point, stream) */
double orc_uniform(uint64_t seed, int64_t point, int64_t stream);
/* `for (double t = t0; t < t1; t += h) ++n;` */
int64_t orc_loop_count(double t0, double t1, double h);

/* model right-hand side and Jacobian-vector product (FP64 and FP32 forms) */
void orc_rhs(int model, int n, const double* x, const double* p, double* out);
void orc_jvp(int model, int n, const double* x, const double* p, const double* v, double* out);
double orc_jac_norm(int model, int n, const double* x, const double* p);
/* trace of the Jacobian (divergence of f) of a 3-D flow at x, and its
   integral over one RK step from the step's stage states (stride n):
   h * (c + sum_j b_j g(X_j)) with tr J = c(p) + g(x), first term a product,
   later terms fma, stages in order (h * c for the flows whose trace is
   constant) */
double orc_trace(int model, int n, const double* x, const double* p);
double orc_trace_step(int model, int integ, int n, const double* stage_x, const double* p, double h);
void orc_jvp_f32(int model, int n, const double* x, const double* p, const float* v, float* out);

/* one base RK step; err_vec (may be NULL) receives h*sum(e_j k_j); returns
   the number of RHS evaluations.  stage_x (may be NULL) receives the stage
   states, stride n, as needed by the coupled tangent step. */
int orc_rk_step(int model, int integ, int n, const double* x, const double* p, double h,
                double* x_new, double* err_vec, double* stage_x);
/* coupled tangent step (Jacobian at the stage states returned by orc_rk_step) */
void orc_tangent_step(int model, int integ, int n, const double* stage_x, const double* p,
                      double h, double* v);
/* the mixed form's coupled tangent step (binary32 stages, FP64 vector); low:
   adaptive DVERK's fifth-order weights b - e on stages 1-6 */
void orc_tangent_step_f32(int model, int integ, int n, const double* stage_x, const double* p,
                          double h, double* v, int low);
/* reference frozen-Jacobian tangent step (dverkStepVarMat) */
void orc_frozen_tangent_step(int model, int n, const double* x_frozen, const double* p,
                             double h, double* v);

/* orthogonalisation kernels */
void orc_mgs(int n, int k, int tree, double* V /* k rows of n */, double* log_norms);
/* 1/sqrt(s) of the renormalisations (seed + four Newton steps, IEEE ops only) */
double orc_inv_sqrt(double s);
/* log-norm accumulation of the sweep's method: norms of up to ORC_LOG_WINDOW
   consecutive renormalisations multiply into mant * 2^ex (mant in [1, 2),
   orc_split_exponent moves the binary exponent of a positive normal value
   into *ex), and log(mant * 2^ex) = orc_log_scaled(mant, ex) of each window
   is Kahan-summed */
#define ORC_LOG_WINDOW 64
double orc_split_exponent(double m, int32_t* ex);
double orc_log_scaled(double mant, int32_t ex);
void orc_cgs_ref(int n, int k, double* V /* k rows of n */);
void orc_normalize_ref(int n, int k, double* V, double eps);
void orc_qr_ref(int n, double* Q /* n x n row-major, columns are vectors */, double* Rdiag);

/* controller (reference dopriStep constants) */
double orc_err_norm(int n, int tree, const double* err_vec, const double* x, const double* x_new,
                    double rtol, double atol);
/* the 3-D flows' form: the error norm squared with one division (small n) */
double orc_err2_seq(int n, const double* err_vec, const double* x, const double* x_new,
                    double rtol, double atol);

/* point parameters / initial state */
void orc_point_params(const orc_desc* d, int64_t idx, const double* params_in, double* p);
void orc_point_init(const orc_desc* d, int64_t idx, const double* p, const double* inits_in, double* x);

/* per-point and whole-sweep drivers */
void orc_point(const orc_desc* d, int64_t idx, const double* params_in, const double* inits_in,
               double* lam, uint32_t* status, uint32_t* n_acc, uint32_t* n_rej);
/* idx = first + i for i in [0, count); outputs are indexed by i */
void orc_sweep(const orc_desc* d, int64_t first, int64_t count, const int64_t* idx_list,
               const double* params_in, const double* inits_in, double* lam,
               uint32_t* status, uint32_t* n_acc, uint32_t* n_rej, int n_threads);

#ifdef __cplusplus
}
#endif
#endif
