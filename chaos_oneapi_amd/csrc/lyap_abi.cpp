// C-ABI of the sweep (include/chaos_lyap.h): contexts, the model table, the
// descriptor finalisation that TaskLyapunov::init/allocate performed in the
// reference (tasklyapunov.cpp:17-31,45-46,65), and the launches.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <string>

#include "chaos_lyap.h"
#include "lyap_launch.hpp"

namespace {

thread_local std::string g_last_error;

int fail(int code, const char* fmt, ...) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    g_last_error = buf;
    return code;
}

}  // namespace

int lyap::set_error(int code, const char* fmt, ...) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    g_last_error = buf;
    return code;
}

namespace {

int hip_fail(hipError_t e, const char* what) {
    return fail(LYAP_ERR_HIP, "%s: %s", what, hipGetErrorString(e));
}

struct ModelRow {
    int32_t model, dim, n_params, type;
    const char *name, *params, *phase, *equations;
};

// ModelFactoryRegistry's table (model_factory.h:43-63), compiled in
const ModelRow kModels[] = {
    {LYAP_MODEL_SHIMIZU, 3, 2, LYAP_TYPE_FLOW, "shimizu", "alpha lambda", "x y z",
     "x'=y; y'=-alpha*y-x*z+x; z'=-lambda*z+x^2"},
    {LYAP_MODEL_LORENZ, 3, 3, LYAP_TYPE_FLOW, "lorenz", "sigma rho beta", "x y z",
     "x'=sigma*(y-x); y'=x*(rho-z)-y; z'=x*y-beta*z"},
    {LYAP_MODEL_ROSSLER, 3, 3, LYAP_TYPE_FLOW, "rossler", "a c b", "x y z",
     "x'=-y-z; y'=x+a*y; z'=b+z*(x-c)"},
    {LYAP_MODEL_L96, 8, 2, LYAP_TYPE_FLOW, "lorenz96_8", "F d", "x0..x7",
     "x_i'=(x_{i+1}-x_{i-2})*x_{i-1}-d*x_i+F (periodic)"},
    {LYAP_MODEL_L96, 16, 2, LYAP_TYPE_FLOW, "lorenz96_16", "F d", "x0..x15",
     "x_i'=(x_{i+1}-x_{i-2})*x_{i-1}-d*x_i+F (periodic)"},
    {LYAP_MODEL_L96, 32, 2, LYAP_TYPE_FLOW, "lorenz96_32", "F d", "x0..x31",
     "x_i'=(x_{i+1}-x_{i-2})*x_{i-1}-d*x_i+F (periodic)"},
};
constexpr int kNumModels = sizeof(kModels) / sizeof(kModels[0]);

void fill_info(const ModelRow& r, lyap_model_info* out) {
    out->model = r.model;
    out->dim = r.dim;
    out->n_params = r.n_params;
    out->type = r.type;
    out->name = r.name;
    out->param_names = r.params;
    out->phase_names = r.phase;
    out->equations = r.equations;
}

int64_t gcd64(int64_t a, int64_t b) {
    while (b) {
        const int64_t t = a % b;
        a = b;
        b = t;
    }
    return a;
}

// a stride near count/phi that is coprime to count
int64_t spread_stride(int64_t count) {
    if (count <= 2) return 1;
    int64_t s = (int64_t)((double)count * 0.6180339887498949);
    if (s < 1) s = 1;
    while (gcd64(s, count) != 1) ++s;
    return s;
}

bool lane_model(int32_t model) {
    return model == LYAP_MODEL_SHIMIZU || model == LYAP_MODEL_LORENZ ||
           model == LYAP_MODEL_ROSSLER;
}

}  // namespace

struct lyap_ctx {
    int device = 0;
    int num_cus = 0;
    int l96_waves[3] = {0, 0, 0};  // waves per point for N = 8, 16, 32 (0: default)
    int lane_max_blocks = 0;
    unsigned long long* counter = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    bool timed = false;
};

extern "C" {

const char* lyap_last_error(void) { return g_last_error.c_str(); }

int lyap_ctx_create(int device, lyap_ctx** out) {
    if (!out) return fail(LYAP_ERR_INVALID, "lyap_ctx_create: out is NULL");
    *out = nullptr;
    int count = 0;
    hipError_t e = hipGetDeviceCount(&count);
    if (e != hipSuccess) return hip_fail(e, "hipGetDeviceCount");
    if (device < 0 || device >= count)
        return fail(LYAP_ERR_INVALID, "lyap_ctx_create: device %d of %d", device, count);
    if ((e = hipSetDevice(device)) != hipSuccess) return hip_fail(e, "hipSetDevice");
    lyap_ctx* c = new lyap_ctx();
    c->device = device;
    if ((e = hipDeviceGetAttribute(&c->num_cus, hipDeviceAttributeMultiprocessorCount, device)) !=
        hipSuccess) {
        delete c;
        return hip_fail(e, "hipDeviceGetAttribute");
    }
    if ((e = hipMalloc(&c->counter, 256)) != hipSuccess) {
        delete c;
        return hip_fail(e, "hipMalloc(counter)");
    }
    if ((e = hipEventCreate(&c->ev0)) != hipSuccess || (e = hipEventCreate(&c->ev1)) != hipSuccess) {
        hipFree(c->counter);
        delete c;
        return hip_fail(e, "hipEventCreate");
    }
    *out = c;
    return LYAP_SUCCESS;
}

int lyap_ctx_destroy(lyap_ctx* c) {
    if (!c) return LYAP_SUCCESS;
    hipSetDevice(c->device);
    if (c->ev0) hipEventDestroy(c->ev0);
    if (c->ev1) hipEventDestroy(c->ev1);
    if (c->counter) hipFree(c->counter);
    delete c;
    return LYAP_SUCCESS;
}

int lyap_model_count(void) { return kNumModels; }

int lyap_model_at(int i, lyap_model_info* out) {
    if (i < 0 || i >= kNumModels || !out) return fail(LYAP_ERR_INVALID, "lyap_model_at: %d", i);
    fill_info(kModels[i], out);
    return LYAP_SUCCESS;
}

int lyap_model_lookup(const char* name, lyap_model_info* out) {
    if (!name || !out) return fail(LYAP_ERR_INVALID, "lyap_model_lookup: NULL argument");
    for (int i = 0; i < kNumModels; ++i)
        if (std::strcmp(kModels[i].name, name) == 0) {
            fill_info(kModels[i], out);
            return LYAP_SUCCESS;
        }
    return fail(LYAP_ERR_NOT_FOUND, "model '%s' is not registered", name);
}

// the reference's `for (double t = t0; t < t1; t += h)` loops
// (cuda/lyapunov.dp.cpp:11,19,37,139,147,158,236,244)
int64_t lyap_loop_count(double t0, double t1, double h) {
    if (!(h > 0.0)) return -1;
    int64_t n = 0;
    for (double t = t0; t < t1; t += h) ++n;
    return n;
}

int lyap_desc_finalize(lyap_sweep_desc* d) {
    if (!d) return fail(LYAP_ERR_INVALID, "lyap_desc_finalize: NULL descriptor");
    bool known = false;
    for (int i = 0; i < kNumModels; ++i)
        if (kModels[i].model == d->model && kModels[i].dim == d->dim) {
            known = true;
            if (d->n_params != kModels[i].n_params)
                return fail(LYAP_ERR_INVALID, "model %s takes %d parameters, got %d",
                            kModels[i].name, kModels[i].n_params, d->n_params);
        }
    if (!known) return fail(LYAP_ERR_INVALID, "unknown model %d with dimension %d", d->model, d->dim);
    if (d->n_exps < 1 || d->n_exps > d->dim)
        return fail(LYAP_ERR_INVALID, "lyapunov_exponent_num %d outside [1, %d]", d->n_exps, d->dim);
    if (!(d->h > 0.0)) return fail(LYAP_ERR_INVALID, "integration_step must be > 0");
    if (!(d->t_average > 0.0)) return fail(LYAP_ERR_INVALID, "total_time must be > 0");
    if (d->t_transient < 0.0 || d->t_slave < 0.0)
        return fail(LYAP_ERR_INVALID, "skip times must be >= 0");
    if (d->sweep_idx1 < 0 || d->sweep_idx1 >= d->n_params || d->sweep_idx2 < 0 ||
        d->sweep_idx2 >= d->n_params)
        return fail(LYAP_ERR_INVALID, "sweep parameter index outside [0, %d)", d->n_params);
    if (d->n1 < 1 || d->n2 < 1) return fail(LYAP_ERR_INVALID, "sweep counts must be >= 1");
    if (d->init_param_offset >= d->n_params)
        return fail(LYAP_ERR_INVALID, "init_param_offset %d outside the parameters",
                    d->init_param_offset);
    if (d->max_steps <= 0) d->max_steps = 20000000LL;
    if (!(d->escape_radius > 0.0)) d->escape_radius = 1e6;
    switch (d->method) {
        case LYAP_METHOD_VAR: {
            if (!(d->t_renorm > 0.0))
                return fail(LYAP_ERR_INVALID, "skip_time_normalization must be > 0");
            d->n_renorm_transient = llround((d->t_transient + d->t_slave) / d->t_renorm);
            d->n_renorm_average = llround(d->t_average / d->t_renorm);
            if (d->n_renorm_average < 1)
                return fail(LYAP_ERR_INVALID, "total_time shorter than one renormalisation");
            if (d->integrator != LYAP_INT_DVERK && d->integrator != LYAP_INT_DOPRI5)
                return fail(LYAP_ERR_INVALID, "unknown integrator %d", d->integrator);
            if (d->adaptive) {
                if (!(d->rtol > 0.0) || !(d->atol > 0.0))
                    return fail(LYAP_ERR_INVALID, "adaptive stepping needs rtol, atol > 0");
                if (!(d->h_min > 0.0)) d->h_min = 1e-12;
                if (!(d->h_max > 0.0)) d->h_max = d->t_renorm;
                d->steps_per_renorm = 0;
            } else {
                d->steps_per_renorm = llround(d->t_renorm / d->h);
                if (d->steps_per_renorm < 1)
                    return fail(LYAP_ERR_INVALID, "skip_time_normalization shorter than a step");
            }
            if (d->precision != LYAP_PREC_FP64 && d->precision != LYAP_PREC_MIXED)
                return fail(LYAP_ERR_INVALID, "unknown precision %d", d->precision);
            break;
        }
        case LYAP_METHOD_VAR_REF:
        case LYAP_METHOD_QR_REF:
        case LYAP_METHOD_FD_REF: {
            if (!lane_model(d->model))
                return fail(LYAP_ERR_UNSUPPORTED,
                            "the reference modes run the 3-D models only (model %d)", d->model);
            if (d->integrator != LYAP_INT_DVERK || d->adaptive || d->precision != LYAP_PREC_FP64)
                return fail(LYAP_ERR_INVALID,
                            "the reference modes are fixed-step FP64 DVERK (cuda/tools/dverk.dp.cpp)");
            if (d->method == LYAP_METHOD_QR_REF && d->n_exps != d->dim)
                return fail(LYAP_ERR_INVALID, "the reference QR mode needs n_exps == dimension");
            if (d->method != LYAP_METHOD_QR_REF && !(d->eps > 0.0))
                return fail(LYAP_ERR_INVALID, "epsilon must be > 0");
            d->ref_steps_transient = lyap_loop_count(0.0, d->t_transient, d->h);
            d->ref_steps_slave = lyap_loop_count(0.0, d->t_slave, d->h);
            d->ref_steps_average =
                d->method == LYAP_METHOD_QR_REF
                    ? lyap_loop_count(d->t_transient, d->t_transient + d->t_average, d->h)
                    : lyap_loop_count(0.0, d->t_average, d->h);
            d->add_skip_steps = d->t_renorm / d->h;
            break;
        }
        default:
            return fail(LYAP_ERR_INVALID, "unknown method %d", d->method);
    }
    return LYAP_SUCCESS;
}

}  // extern "C"

namespace {

int run_finalized(lyap_ctx* c, const lyap_sweep_desc& fd, int64_t first, int64_t count,
                  const int64_t* d_idx_list, const double* d_params, const double* d_inits,
                  double* d_lambda, uint32_t* d_status, uint32_t* d_n_acc, uint32_t* d_n_rej,
                  void* stream) {
    hipError_t e = hipSetDevice(c->device);
    if (e != hipSuccess) return hip_fail(e, "hipSetDevice");
    hipStream_t s = (hipStream_t)stream;
    lyap::SweepIO io{first,    count,    d_idx_list, d_params, d_inits,  d_lambda,
                     d_status, d_n_acc,  d_n_rej,    c->counter, spread_stride(count)};
    if ((e = hipEventRecord(c->ev0, s)) != hipSuccess) return hip_fail(e, "hipEventRecord");
    if (lane_model(fd.model)) {
        e = lyap::launch_lane_sweep(fd, io, s, c->num_cus, c->lane_max_blocks);
    } else if (fd.model == LYAP_MODEL_L96) {
        if (fd.method != LYAP_METHOD_VAR)
            return fail(LYAP_ERR_UNSUPPORTED, "Lorenz-96 runs the coupled variational method only");
        e = lyap::launch_l96_sweep(fd, io, s, c->num_cus, c->l96_waves[fd.dim == 8 ? 0 : (fd.dim == 16 ? 1 : 2)]);
    } else {
        return fail(LYAP_ERR_UNSUPPORTED, "no kernel for model %d (dimension %d)", fd.model, fd.dim);
    }
    if (e != hipSuccess) return hip_fail(e, "sweep launch");
    if ((e = hipEventRecord(c->ev1, s)) != hipSuccess) return hip_fail(e, "hipEventRecord");
    c->timed = true;
    return LYAP_SUCCESS;
}

}  // namespace

extern "C" {

int lyap_sweep_run(lyap_ctx* c, const lyap_sweep_desc* d, int64_t first, int64_t count,
                   const int64_t* d_idx_list, const double* d_params, const double* d_inits,
                   double* d_lambda, uint32_t* d_status, uint32_t* d_n_acc, uint32_t* d_n_rej,
                   void* stream) {
    if (!c || !d) return fail(LYAP_ERR_INVALID, "lyap_sweep_run: NULL context or descriptor");
    if (count < 0 || first < 0) return fail(LYAP_ERR_INVALID, "lyap_sweep_run: negative range");
    if (count == 0) return LYAP_SUCCESS;
    if (!d_lambda) return fail(LYAP_ERR_INVALID, "lyap_sweep_run: d_lambda is NULL");
    if (!d_idx_list && !d_params && first + count > (int64_t)d->n1 * (int64_t)d->n2)
        return fail(LYAP_ERR_INVALID, "range [%lld, %lld) beyond the %d x %d grid",
                    (long long)first, (long long)(first + count), d->n1, d->n2);
    lyap_sweep_desc fd = *d;
    int rc = lyap_desc_finalize(&fd);
    if (rc != LYAP_SUCCESS) return rc;
    return run_finalized(c, fd, first, count, d_idx_list, d_params, d_inits, d_lambda, d_status,
                         d_n_acc, d_n_rej, stream);
}

int lyap_ctx_set_option(lyap_ctx* c, const char* key, int64_t value) {
    if (!c || !key) return fail(LYAP_ERR_INVALID, "lyap_ctx_set_option: NULL argument");
    const std::string k(key);
    const int dims[3] = {8, 16, 32};
    for (int j = 0; j < 3; ++j)
        if (k == "l96_waves_" + std::to_string(dims[j])) {
            if (value != 0 && !lyap::l96_supported(dims[j], (int)value))
                return fail(LYAP_ERR_INVALID, "%s = %lld is not a built layout", key, (long long)value);
            c->l96_waves[j] = (int)value;
            return LYAP_SUCCESS;
        }
    if (k == "lane_max_blocks") {
        if (value < 0) return fail(LYAP_ERR_INVALID, "lane_max_blocks must be >= 0");
        c->lane_max_blocks = (int)value;
        return LYAP_SUCCESS;
    }
    return fail(LYAP_ERR_INVALID, "unknown option '%s'", key);
}

int lyap_last_kernel_ms(lyap_ctx* c, float* ms) {
    if (!c || !ms) return fail(LYAP_ERR_INVALID, "lyap_last_kernel_ms: NULL argument");
    if (!c->timed) return fail(LYAP_ERR_INVALID, "no sweep has been launched on this context");
    hipError_t e = hipEventSynchronize(c->ev1);
    if (e != hipSuccess) return hip_fail(e, "hipEventSynchronize");
    if ((e = hipEventElapsedTime(ms, c->ev0, c->ev1)) != hipSuccess)
        return hip_fail(e, "hipEventElapsedTime");
    return LYAP_SUCCESS;
}

int lyap_cuda_lyapunov(lyap_ctx* c, int block_num, int block_size, size_t task_num, int is_qr,
                       int is_var, int model_type, const double* d_inits, int32_t dimension,
                       int32_t model, const double* d_params, int32_t params_dim, double eps,
                       double integration_step, double* d_lyap_exps, int32_t exps_num,
                       double time_skip, double time_skip_clp, double calc_time,
                       double add_skip_steps, void* stream) {
    (void)block_num;
    (void)block_size;
    if (model_type != LYAP_TYPE_FLOW)
        return fail(LYAP_ERR_UNSUPPORTED, "map-type models have no kernel in this build");
    if (!d_inits || !d_params)
        return fail(LYAP_ERR_INVALID, "lyap_cuda_lyapunov: inits and params are required");
    lyap_sweep_desc d;
    std::memset(&d, 0, sizeof(d));
    d.model = model;
    d.dim = dimension;
    d.n_exps = exps_num;
    d.n_params = params_dim;
    d.method = is_qr ? LYAP_METHOD_QR_REF : (is_var ? LYAP_METHOD_VAR_REF : LYAP_METHOD_FD_REF);
    d.integrator = LYAP_INT_DVERK;
    d.h = integration_step;
    d.eps = eps;
    d.t_transient = time_skip;
    d.t_slave = time_skip_clp;
    d.t_average = calc_time;
    d.t_renorm = add_skip_steps * integration_step;
    d.n1 = 1;
    d.n2 = (int32_t)task_num;
    d.sweep_idx1 = 0;
    d.sweep_idx2 = params_dim > 1 ? 1 : 0;
    d.init_kind = LYAP_INIT_NONE;
    d.init_param_offset = -1;
    int rc = lyap_desc_finalize(&d);
    if (rc != LYAP_SUCCESS) return rc;
    d.add_skip_steps = add_skip_steps;  // the caller's value, compared as the reference does
    if (task_num == 0) return LYAP_SUCCESS;
    if (!d_lyap_exps) return fail(LYAP_ERR_INVALID, "lyap_cuda_lyapunov: lyapExps is NULL");
    return run_finalized(c, d, 0, (int64_t)task_num, nullptr, d_params, d_inits, d_lyap_exps,
                         nullptr, nullptr, nullptr, stream);
}

}  // extern "C"
