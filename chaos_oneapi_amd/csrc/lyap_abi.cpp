// C-ABI of the sweep (include/chaos_lyap.h): contexts, the model table, the
// descriptor finalisation that TaskLyapunov::init/allocate performed in the
// reference (tasklyapunov.cpp:17-31,45-46,65), and the launches.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <algorithm>
#include <atomic>
#include <chrono>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

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
    {LYAP_MODEL_HENON, 2, 2, LYAP_TYPE_MAP, "henon", "a b", "x y", "x_{n+1}=1-a*x_n^2+y_n; y_{n+1}=b*x_n"},
    {LYAP_MODEL_LOGISTIC, 1, 1, LYAP_TYPE_MAP, "logistic", "r", "x", "x_{n+1}=r*x_n*(1-x_n)"},
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

bool map_model(int32_t model) { return model == LYAP_MODEL_HENON || model == LYAP_MODEL_LOGISTIC; }

// models whose kernels run one point per lane
bool lane_model(int32_t model) {
    return model == LYAP_MODEL_SHIMIZU || model == LYAP_MODEL_LORENZ ||
           model == LYAP_MODEL_ROSSLER || map_model(model);
}

}  // namespace

struct lyap_ctx {
    int device = 0;
    int num_cus = 0;
    int l96_waves[3] = {0, 0, 0};  // waves per point for N = 8, 16, 32 (0: default)
    int l96_ppw[3] = {0, 0, 0};    // points per wave (wave-level drain) for N = 8, 16, 32; 0: off
    int lane_max_blocks = 0;
    int l96_max_blocks = 0;
    int lane_gather = 0;
    int item_order = 0;  // 0: items spread over the grid; 1: in list order
    int lane_quad = 0;   // 1: the sweep's method on the 3-D flows with four lanes per point
    int order_probe = 0; // > 0: adaptive sweeps run heavy-first, ordered by a probe run over
                         // this many thousandths of the integration times
    int order_radius = 2;  // grid neighbourhood whose heaviest probe count orders a point
    unsigned long long* counter = nullptr;
    void* order_buf = nullptr;  // staging, keys, index lists and sort scratch of ordered runs
    size_t order_cap = 0;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    bool timed = false;
    // progress signals and skips of the next run (lyap_ctx_set_watch): device
    // counters, host-visible flags (pinned, allocated on first use), the
    // milestone ends, the caller's skip words, and the run's sequence number
    bool watch_next = false;
    int64_t watch_end[2] = {0, 0};
    const uint32_t* watch_skip = nullptr;
    unsigned long long* watch_done = nullptr;
    uint32_t* watch_flag = nullptr;
    uint32_t watch_seq = 0;
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
    if (c->watch_done) hipFree(c->watch_done);
    if (c->watch_flag) hipHostFree(c->watch_flag);
    if (c->order_buf) hipFree(c->order_buf);
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
    // a map advances one iteration per unit of time (the reference's map loops
    // count t += 1, cuda/lyapunov.dp.cpp:70-122,182-222)
    if (map_model(d->model)) d->h = 1.0;
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
    if (map_model(d->model)) {
        if (d->adaptive) return fail(LYAP_ERR_INVALID, "a map has no adaptive step");
        if (d->method == LYAP_METHOD_QR_REF)
            return fail(LYAP_ERR_UNSUPPORTED, "the reference QR mode has no map branch");
    }
    if (d->method == LYAP_METHOD_VAR_TRACE) {
        if (!lane_model(d->model) || map_model(d->model))
            return fail(LYAP_ERR_UNSUPPORTED, "the trace closure runs the 3-D flows only (model %d)",
                        d->model);
        if (d->n_exps != d->dim)
            return fail(LYAP_ERR_INVALID, "the trace closure needs n_exps == dimension");
    }
    switch (d->method) {
        case LYAP_METHOD_VAR:
        case LYAP_METHOD_VAR_TRACE: {
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
                            "the reference modes run the 3-D flows and the maps only (model %d)",
                            d->model);
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

// device buffers of a heavy-first run, carved from the context's order_buf
struct OrderBufs {
    double* lam;
    uint32_t *st, *na, *nr, *keys_in, *keys_out;
    int64_t *vals_in, *vals_out;
    void* temp;
    size_t temp_bytes;
};

size_t align256(size_t n) { return (n + 255) & ~(size_t)255; }

int order_buffers(lyap_ctx* c, int64_t count, int n_exps, OrderBufs* ob) {
    if (count > (int64_t)0x7fffffff)
        return fail(LYAP_ERR_INVALID, "heavy-first order takes at most 2^31 - 1 points, got %lld",
                    (long long)count);
    const size_t n = (size_t)count;
    const size_t temp_bytes = lyap::order_sort_bytes(count);
    if (temp_bytes == 0) return fail(LYAP_ERR_HIP, "heavy-first order: sort scratch query failed");
    const size_t sizes[9] = {n * (size_t)n_exps * sizeof(double), n * 4, n * 4, n * 4, n * 4, n * 4,
                             n * 8, n * 8, temp_bytes};
    size_t need = 0;
    for (size_t z : sizes) need += align256(z);
    if (need > c->order_cap) {
        if (c->order_buf) hipFree(c->order_buf);
        c->order_buf = nullptr;
        c->order_cap = 0;
        hipError_t e = hipMalloc(&c->order_buf, need);
        if (e != hipSuccess) return hip_fail(e, "hipMalloc(order buffers)");
        c->order_cap = need;
    }
    char* p = (char*)c->order_buf;
    void* parts[9];
    for (int i = 0; i < 9; ++i) {
        parts[i] = p;
        p += align256(sizes[i]);
    }
    *ob = OrderBufs{(double*)parts[0],   (uint32_t*)parts[1], (uint32_t*)parts[2], (uint32_t*)parts[3],
                    (uint32_t*)parts[4], (uint32_t*)parts[5], (int64_t*)parts[6],  (int64_t*)parts[7],
                    parts[8],            temp_bytes};
    return LYAP_SUCCESS;
}

int run_finalized(lyap_ctx* c, const lyap_sweep_desc& fd, int64_t first, int64_t count,
                  const int64_t* d_idx_list, const double* d_params, const double* d_inits,
                  double* d_lambda, uint32_t* d_status, uint32_t* d_n_acc, uint32_t* d_n_rej,
                  void* stream) {
    hipError_t e = hipSetDevice(c->device);
    if (e != hipSuccess) return hip_fail(e, "hipSetDevice");
    hipStream_t s = (hipStream_t)stream;
    lyap::SweepIO io{first,    count,    d_idx_list, d_params,   d_inits,
                     d_lambda, d_status, d_n_acc,    d_n_rej,    c->counter,
                     c->item_order ? 1 : spread_stride(count), c->lane_gather, 0};
    const bool quad = c->lane_quad != 0 &&
                      (fd.method == LYAP_METHOD_VAR || fd.method == LYAP_METHOD_VAR_TRACE) && fd.dim == 3;
    lyap::LaneWatch watch{};
    const bool watched = c->watch_next;
    c->watch_next = false;  // one run only
    if (watched) {
        if (!lane_model(fd.model) || map_model(fd.model) || quad || !fd.adaptive ||
            (fd.method != LYAP_METHOD_VAR && fd.method != LYAP_METHOD_VAR_TRACE))
            return fail(LYAP_ERR_UNSUPPORTED,
                        "progress signals and skips need the sweep's method with adaptive steps on a 3-D "
                        "flow, one lane per point");
        if (c->watch_end[1] > count)
            return fail(LYAP_ERR_INVALID, "watched items [0, %lld) beyond the %lld items of the run",
                        (long long)c->watch_end[1], (long long)count);
        watch = lyap::LaneWatch{c->watch_skip, c->watch_done, c->watch_flag,
                                {c->watch_end[0], c->watch_end[1]}, ++c->watch_seq, 0u};
    }
    if (!lane_model(fd.model)) {
        if (fd.model != LYAP_MODEL_L96)
            return fail(LYAP_ERR_UNSUPPORTED, "no kernel for model %d (dimension %d)", fd.model, fd.dim);
        if (fd.method != LYAP_METHOD_VAR)
            return fail(LYAP_ERR_UNSUPPORTED, "Lorenz-96 runs the coupled variational method only");
    }
    auto launch = [&](const lyap_sweep_desc& d, const lyap::SweepIO& x) -> hipError_t {
        if (lane_model(d.model))
            return lyap::launch_lane_sweep(d, x, s, c->num_cus, c->lane_max_blocks, quad,
                                           watched ? &watch : nullptr);
        const int j = d.dim == 8 ? 0 : (d.dim == 16 ? 1 : 2);
        return lyap::launch_l96_sweep(d, x, s, c->num_cus, c->l96_waves[j], c->l96_ppw[j], c->l96_max_blocks);
    };
    // heavy-first order: adaptive runs of the grid (no caller's list, not watched)
    const bool ordered = c->order_probe > 0 && !d_idx_list && !watched && fd.adaptive &&
                         (fd.method == LYAP_METHOD_VAR || fd.method == LYAP_METHOD_VAR_TRACE);
    OrderBufs ob{};
    if (ordered) {
        const int rc = order_buffers(c, count, fd.n_exps, &ob);
        if (rc != LYAP_SUCCESS) return rc;
    }
    // ev0/ev1 bracket the run's own launch (not an ordered run's probe, sort
    // and scatter)
    if (ordered) {
        // the probe: the same integration over order_probe thousandths of the
        // times (at least one renormalisation interval of averaging)
        lyap_sweep_desc pd = fd;
        const double f = (double)c->order_probe / 1000.0;
        pd.t_transient *= f;
        pd.t_slave *= f;
        pd.t_average = std::max(fd.t_average * f, fd.t_renorm);
        int rc = lyap_desc_finalize(&pd);
        if (rc != LYAP_SUCCESS) return rc;
        lyap::SweepIO pio = io;
        pio.idx_list = nullptr;
        pio.lambda = ob.lam;
        pio.status = ob.st;
        pio.n_acc = ob.na;
        pio.n_rej = ob.nr;
        pio.stride = spread_stride(count);
        if ((e = launch(pd, pio)) != hipSuccess) return hip_fail(e, "probe launch");
        // the grid neighbourhood is the parameter plane's only where the
        // points are the grid's own (no caller's parameter rows)
        const int32_t radius = d_params ? 0 : c->order_radius;
        if ((e = lyap::order_heavy_first(ob.na, ob.nr, first, count, fd.n2, radius, ob.keys_in, ob.keys_out,
                                         ob.vals_in, ob.vals_out, ob.temp, ob.temp_bytes, s)) != hipSuccess)
            return hip_fail(e, "heavy-first sort");
        lyap::SweepIO mio = pio;
        mio.idx_list = ob.vals_out;
        mio.stride = 1;  // list order
        if ((e = hipEventRecord(c->ev0, s)) != hipSuccess) return hip_fail(e, "hipEventRecord");
        if ((e = launch(fd, mio)) != hipSuccess) return hip_fail(e, "sweep launch");
        if ((e = hipEventRecord(c->ev1, s)) != hipSuccess) return hip_fail(e, "hipEventRecord");
        if ((e = lyap::order_scatter(ob.vals_out, first, count, fd.n_exps, ob.lam, ob.st, ob.na, ob.nr,
                                     d_lambda, d_status, d_n_acc, d_n_rej, s)) != hipSuccess)
            return hip_fail(e, "output scatter");
    } else {
        if ((e = hipEventRecord(c->ev0, s)) != hipSuccess) return hip_fail(e, "hipEventRecord");
        if ((e = launch(fd, io)) != hipSuccess) return hip_fail(e, "sweep launch");
        if ((e = hipEventRecord(c->ev1, s)) != hipSuccess) return hip_fail(e, "hipEventRecord");
    }
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
    for (int j = 0; j < 3; ++j)
        if (k == "l96_ppw_" + std::to_string(dims[j])) {
            if (value != 0 && !lyap::l96_wave_supported(dims[j], (int)value))
                return fail(LYAP_ERR_INVALID, "%s = %lld is not a built layout", key, (long long)value);
            c->l96_ppw[j] = (int)value;
            return LYAP_SUCCESS;
        }
    if (k == "lane_max_blocks") {
        if (value < -65536 || value > INT32_MAX)
            return fail(LYAP_ERR_INVALID, "lane_max_blocks must be in [-65536, 2^31 - 1]");
        c->lane_max_blocks = (int)value;
        return LYAP_SUCCESS;
    }
    if (k == "lane_gather") {
        if (value < 0 || value > 64) return fail(LYAP_ERR_INVALID, "lane_gather must be in [0, 64]");
        c->lane_gather = (int)value;
        return LYAP_SUCCESS;
    }
    if (k == "lane_quad") {
        if (value != 0 && value != 1) return fail(LYAP_ERR_INVALID, "lane_quad must be 0 or 1");
        c->lane_quad = (int)value;
        return LYAP_SUCCESS;
    }
    if (k == "order_probe") {
        if (value < 0 || value > 500) return fail(LYAP_ERR_INVALID, "order_probe must be in [0, 500]");
        c->order_probe = (int)value;
        return LYAP_SUCCESS;
    }
    if (k == "order_radius") {
        if (value < 0 || value > 8) return fail(LYAP_ERR_INVALID, "order_radius must be in [0, 8]");
        c->order_radius = (int)value;
        return LYAP_SUCCESS;
    }
    if (k == "item_order") {
        if (value != 0 && value != 1) return fail(LYAP_ERR_INVALID, "item_order must be 0 or 1");
        c->item_order = (int)value;
        return LYAP_SUCCESS;
    }
    if (k == "l96_max_blocks") {
        if (value < 0) return fail(LYAP_ERR_INVALID, "l96_max_blocks must be >= 0");
        c->l96_max_blocks = (int)value;
        return LYAP_SUCCESS;
    }
    return fail(LYAP_ERR_INVALID, "unknown option '%s'", key);
}

int lyap_ctx_set_watch(lyap_ctx* c, int64_t end0, int64_t end1, const uint32_t* d_skip) {
    if (!c) return fail(LYAP_ERR_INVALID, "lyap_ctx_set_watch: NULL context");
    if (end0 < 0 || end1 < end0)
        return fail(LYAP_ERR_INVALID, "lyap_ctx_set_watch: milestones [0, %lld) and [%lld, %lld) are not ranges",
                    (long long)end0, (long long)end0, (long long)end1);
    hipError_t e = hipSetDevice(c->device);
    if (e != hipSuccess) return hip_fail(e, "hipSetDevice");
    if (!c->watch_done) {
        if ((e = hipMalloc(&c->watch_done, 2 * sizeof(unsigned long long))) != hipSuccess) {
            c->watch_done = nullptr;
            return hip_fail(e, "hipMalloc(watch counters)");
        }
    }
    if (!c->watch_flag) {
        void* h = nullptr;
        if ((e = hipHostMalloc(&h, 2 * sizeof(uint32_t), hipHostMallocCoherent | hipHostMallocMapped)) !=
            hipSuccess)
            return hip_fail(e, "hipHostMalloc(watch flags)");
        c->watch_flag = static_cast<uint32_t*>(h);
        c->watch_flag[0] = c->watch_flag[1] = 0u;
    }
    c->watch_next = true;
    c->watch_end[0] = end0;
    c->watch_end[1] = end1;
    c->watch_skip = d_skip;
    return LYAP_SUCCESS;
}

int lyap_ctx_watch_wait(lyap_ctx* c, int milestone, double timeout_s, int* reached) {
    if (!c || !reached || milestone < 0 || milestone > 1)
        return fail(LYAP_ERR_INVALID, "lyap_ctx_watch_wait: bad argument");
    *reached = 0;
    if (!c->watch_flag || c->watch_seq == 0)
        return fail(LYAP_ERR_INVALID, "lyap_ctx_watch_wait: no watched run on this context");
    const int64_t size = milestone ? c->watch_end[1] - c->watch_end[0] : c->watch_end[0];
    if (size == 0) {
        *reached = 1;
        return LYAP_SUCCESS;
    }
    const volatile uint32_t* f = c->watch_flag + milestone;
    const auto t0 = std::chrono::steady_clock::now();
    for (;;) {
        if (*f == c->watch_seq) {
            std::atomic_thread_fence(std::memory_order_acquire);
            *reached = 1;
            return LYAP_SUCCESS;
        }
        const double dt = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
        if (dt >= timeout_s) return LYAP_SUCCESS;
        std::this_thread::sleep_for(std::chrono::microseconds(50));
    }
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
    if (!c) return fail(LYAP_ERR_INVALID, "lyap_cuda_lyapunov: NULL context");
    if (model_type != LYAP_TYPE_FLOW && model_type != LYAP_TYPE_MAP)
        return fail(LYAP_ERR_INVALID, "lyap_cuda_lyapunov: unknown model type %d", model_type);
    if (task_num > (size_t)INT32_MAX)
        return fail(LYAP_ERR_INVALID, "lyap_cuda_lyapunov: task_num %zu above %d", task_num, INT32_MAX);
    if (!d_inits || !d_params)
        return fail(LYAP_ERR_INVALID, "lyap_cuda_lyapunov: inits and params are required");
    // the model type given must be the model's (cuda/lyapunov.dp.cpp:274-317
    // branches on it); a map advances one iteration per unit of time, so the
    // descriptor's finalisation sets h = 1 and rejects the QR mode for maps as
    // the reference does (cuda/lyapunov.dp.cpp:302-305)
    bool type_ok = false;
    for (int i = 0; i < kNumModels; ++i)
        if (kModels[i].model == model && kModels[i].dim == dimension) type_ok = kModels[i].type == model_type;
    if (!type_ok)
        return fail(LYAP_ERR_INVALID, "lyap_cuda_lyapunov: model %d (dimension %d) is not of type %d", model,
                    dimension, model_type);
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
    d.t_renorm = add_skip_steps * (model_type == LYAP_TYPE_MAP ? 1.0 : integration_step);
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

// ---------------------------------------------------------------------------
// batched mixed-model launch
// ---------------------------------------------------------------------------
struct lyap_batch {
    int device = 0;
    int n = 0;                              // classes on the device (empty ones dropped)
    lyap::BatchClass* d_cls = nullptr;      // n records
    unsigned long long* d_counters = nullptr;
    hipStream_t capture = nullptr;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    bool timed = false;
    int64_t waves = 0;
    lyap_batch_tuning tuning{};
};

namespace {

int batch_kind(const lyap_sweep_desc& d) {
    const bool trace = d.method == LYAP_METHOD_VAR_TRACE;
    switch (d.model) {
        case LYAP_MODEL_SHIMIZU: return trace ? lyap::BATCH_SHIMIZU_TRACE : lyap::BATCH_SHIMIZU;
        case LYAP_MODEL_LORENZ: return trace ? lyap::BATCH_LORENZ_TRACE : lyap::BATCH_LORENZ;
        case LYAP_MODEL_ROSSLER: return trace ? lyap::BATCH_ROSSLER_TRACE : lyap::BATCH_ROSSLER;
        case LYAP_MODEL_L96:
            return d.dim == 8 ? lyap::BATCH_L96_8
                              : (d.dim == 16 ? lyap::BATCH_L96_16 : (d.dim == 32 ? lyap::BATCH_L96_32 : -1));
        default: return -1;
    }
}

// queue position of a class kind: per-lane 3-D points are the longest tasks
// (one lane runs a whole point), then Lorenz-96 by N descending
int longest_first_rank(int kind) {
    switch (kind) {
        case lyap::BATCH_LORENZ:
        case lyap::BATCH_LORENZ_TRACE: return 0;
        case lyap::BATCH_ROSSLER:
        case lyap::BATCH_ROSSLER_TRACE: return 1;
        case lyap::BATCH_SHIMIZU:
        case lyap::BATCH_SHIMIZU_TRACE: return 2;
        case lyap::BATCH_L96_32: return 3;
        case lyap::BATCH_L96_16: return 4;
        default: return 5;
    }
}

// work counters of a batch: one per class, 8 words apart (separate cache lines)
constexpr int kCounterStride = 8;

// A batch's class records, validated and in queue order, built on the host
// without allocating (fixed arrays), shared by lyap_batch_create and
// lyap_batch_run_noalloc.  io.counter is left for the caller to point.
struct BatchPlan {
    lyap::BatchClass recs[lyap::kMaxBatchClasses];
    int n = 0;  // non-empty classes
    int integ = -1, adapt = -1, prec = -1;
    int64_t wave_tasks = 0;
    lyap_batch_tuning tuning{};
};

int plan_batch(const lyap_batch_class* classes, int n_classes, const lyap_batch_tuning* tuning,
               BatchPlan* plan) {
    if (n_classes < 1 || n_classes > lyap::kMaxBatchClasses)
        return fail(LYAP_ERR_INVALID, "a batch holds 1 to %d classes, got %d", lyap::kMaxBatchClasses,
                    n_classes);
    lyap_batch_tuning tu{};
    if (tuning) tu = *tuning;
    if (tu.layout < 0 || tu.layout >= lyap::kBatchLayouts)
        return fail(LYAP_ERR_INVALID, "batch layout %d is not built", tu.layout);
    if (tu.waves_per_cu < 0 || tu.order < 0 || tu.order > 2 || tu.gather < 0 || tu.gather > 64)
        return fail(LYAP_ERR_INVALID, "batch tuning: waves_per_cu >= 0, order in 0..2, gather in 0..64");
    plan->tuning = tu;
    lyap::BatchClass recs[lyap::kMaxBatchClasses];
    int rank[lyap::kMaxBatchClasses];
    int n = 0;
    for (int j = 0; j < n_classes; ++j) {
        const lyap_batch_class& k = classes[j];
        lyap_sweep_desc fd = k.desc;
        int rc = lyap_desc_finalize(&fd);
        if (rc != LYAP_SUCCESS) return fail(rc, "batch class %d: %s", j, g_last_error.c_str());
        if (fd.method != LYAP_METHOD_VAR && fd.method != LYAP_METHOD_VAR_TRACE)
            return fail(LYAP_ERR_UNSUPPORTED,
                        "batch class %d: a batch runs the variational method (and its trace closure) only", j);
        const int kind = batch_kind(fd);
        if (kind < 0)
            return fail(LYAP_ERR_UNSUPPORTED, "batch class %d: no batch kernel for model %d dimension %d",
                        j, fd.model, fd.dim);
        if (j == 0) {
            plan->integ = fd.integrator;
            plan->adapt = fd.adaptive ? 1 : 0;
            plan->prec = fd.precision;
        } else if (fd.integrator != plan->integ || (fd.adaptive ? 1 : 0) != plan->adapt ||
                   fd.precision != plan->prec) {
            return fail(LYAP_ERR_INVALID,
                        "batch class %d: integrator, adaptive flag and precision must match class 0", j);
        }
        if (k.count < 0) return fail(LYAP_ERR_INVALID, "batch class %d: negative count", j);
        if (k.count == 0) continue;
        if (!k.d_lambda) return fail(LYAP_ERR_INVALID, "batch class %d: d_lambda is NULL", j);
        if (!k.d_params && k.count > (int64_t)fd.n1 * (int64_t)fd.n2)
            return fail(LYAP_ERR_INVALID, "batch class %d: %lld points beyond the %d x %d grid", j,
                        (long long)k.count, fd.n1, fd.n2);
        lyap::BatchClass& r = recs[n];
        std::memset(&r, 0, sizeof(r));
        r.d = fd;
        r.io = lyap::SweepIO{0,          k.count,    nullptr,   k.d_params, k.d_inits,
                             k.d_lambda, k.d_status, k.d_n_acc, k.d_n_rej,  nullptr,
                             spread_stride(k.count), tu.gather, 0};
        r.kind = kind;
        rank[n] = tu.order == 0 ? longest_first_rank(kind)
                                : (tu.order == 2 ? (lyap::batch_kind_is_l96(kind) ? 0 : 1) : 0);
        const int ppw = lyap::batch_points_per_wave(tu.layout, fd.dim);
        plan->wave_tasks += (k.count + ppw - 1) / ppw;
        ++n;
    }
    if (!lyap::batch_variant_built(plan->integ, plan->adapt != 0, plan->prec == LYAP_PREC_MIXED, tu.layout,
                                   tu.occupancy3 != 0))
        return fail(LYAP_ERR_UNSUPPORTED, "batch kernel layout %d%s is built for adaptive Dopri5 only",
                    tu.layout, tu.occupancy3 ? " (3 waves per SIMD)" : "");
    // stable order by rank (insertion sort: at most 16 records)
    int ord[lyap::kMaxBatchClasses];
    for (int j = 0; j < n; ++j) {
        int q = j;
        while (q > 0 && rank[ord[q - 1]] > rank[j]) {
            ord[q] = ord[q - 1];
            --q;
        }
        ord[q] = j;
    }
    for (int j = 0; j < n; ++j) plan->recs[j] = recs[ord[j]];
    plan->n = n;
    return LYAP_SUCCESS;
}

// persistent grid of a planned batch on a device with num_cus CUs
hipError_t plan_waves(const BatchPlan& plan, int num_cus, int64_t* waves) {
    int per_cu = 0;
    hipError_t e = lyap::batch_max_waves_per_cu(plan.integ, plan.adapt != 0, plan.prec == LYAP_PREC_MIXED,
                                                plan.tuning.layout, plan.tuning.occupancy3 != 0, &per_cu);
    if (e != hipSuccess) return e;
    if (per_cu < 1) per_cu = 1;
    if (plan.tuning.waves_per_cu > 0 && plan.tuning.waves_per_cu < per_cu) per_cu = plan.tuning.waves_per_cu;
    int64_t w = (int64_t)per_cu * num_cus;
    if (w > plan.wave_tasks) w = plan.wave_tasks;
    *waves = w < 1 ? 1 : w;
    return hipSuccess;
}

void batch_free(lyap_batch* b) {
    if (!b) return;
    hipSetDevice(b->device);
    if (b->exec) hipGraphExecDestroy(b->exec);
    if (b->graph) hipGraphDestroy(b->graph);
    if (b->capture) hipStreamDestroy(b->capture);
    if (b->ev0) hipEventDestroy(b->ev0);
    if (b->ev1) hipEventDestroy(b->ev1);
    if (b->d_cls) hipFree(b->d_cls);
    if (b->d_counters) hipFree(b->d_counters);
    delete b;
}

}  // namespace

extern "C" {

int lyap_batch_create(lyap_ctx* c, const lyap_batch_class* classes, int n_classes,
                      const lyap_batch_tuning* tuning, lyap_batch** out) {
    if (!c || !classes || !out) return fail(LYAP_ERR_INVALID, "lyap_batch_create: NULL argument");
    *out = nullptr;
    BatchPlan plan;
    int rc = plan_batch(classes, n_classes, tuning, &plan);
    if (rc != LYAP_SUCCESS) return rc;
    const lyap_batch_tuning& tu = plan.tuning;

    hipError_t e = hipSetDevice(c->device);
    if (e != hipSuccess) return hip_fail(e, "hipSetDevice");
    lyap_batch* b = new lyap_batch();
    b->device = c->device;
    b->tuning = tu;
    b->n = plan.n;
    if (b->n == 0) {
        *out = b;
        return LYAP_SUCCESS;
    }
    if ((e = plan_waves(plan, c->num_cus, &b->waves)) != hipSuccess) {
        batch_free(b);
        return hip_fail(e, "batch occupancy");
    }
    if ((e = hipMalloc(&b->d_counters, lyap::kMaxBatchClasses * kCounterStride * sizeof(unsigned long long))) !=
            hipSuccess ||
        (e = hipMalloc(&b->d_cls, plan.n * sizeof(lyap::BatchClass))) != hipSuccess) {
        batch_free(b);
        return hip_fail(e, "hipMalloc(batch)");
    }
    for (int j = 0; j < plan.n; ++j) plan.recs[j].io.counter = b->d_counters + kCounterStride * j;
    if ((e = hipMemcpy(b->d_cls, plan.recs, plan.n * sizeof(lyap::BatchClass), hipMemcpyHostToDevice)) !=
        hipSuccess) {
        batch_free(b);
        return hip_fail(e, "hipMemcpy(batch classes)");
    }
    // capture: counter reset (a kernel, not a memset node) + the persistent kernel
    if ((e = hipStreamCreateWithFlags(&b->capture, hipStreamNonBlocking)) != hipSuccess ||
        (e = hipEventCreate(&b->ev0)) != hipSuccess || (e = hipEventCreate(&b->ev1)) != hipSuccess) {
        batch_free(b);
        return hip_fail(e, "batch stream/events");
    }
    if ((e = hipStreamBeginCapture(b->capture, hipStreamCaptureModeThreadLocal)) != hipSuccess) {
        batch_free(b);
        return hip_fail(e, "hipStreamBeginCapture");
    }
    hipError_t el = lyap::launch_batch(b->d_cls, b->n, plan.integ, plan.adapt != 0, plan.prec == LYAP_PREC_MIXED,
                                       tu.layout, tu.occupancy3 != 0, b->waves, b->capture);
    e = hipStreamEndCapture(b->capture, &b->graph);
    if (el != hipSuccess || e != hipSuccess) {
        batch_free(b);
        return hip_fail(el != hipSuccess ? el : e, "batch graph capture");
    }
    if ((e = hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0)) != hipSuccess) {
        batch_free(b);
        return hip_fail(e, "hipGraphInstantiate");
    }
    *out = b;
    return LYAP_SUCCESS;
}

int lyap_batch_launch(lyap_batch* b, void* stream) {
    if (!b) return fail(LYAP_ERR_INVALID, "lyap_batch_launch: NULL batch");
    if (b->n == 0) return LYAP_SUCCESS;
    hipError_t e = hipSetDevice(b->device);
    if (e != hipSuccess) return hip_fail(e, "hipSetDevice");
    hipStream_t s = (hipStream_t)stream;
    if ((e = hipEventRecord(b->ev0, s)) != hipSuccess) return hip_fail(e, "hipEventRecord");
    if ((e = hipGraphLaunch(b->exec, s)) != hipSuccess) return hip_fail(e, "hipGraphLaunch");
    if ((e = hipEventRecord(b->ev1, s)) != hipSuccess) return hip_fail(e, "hipEventRecord");
    b->timed = true;
    return LYAP_SUCCESS;
}

int lyap_batch_last_ms(lyap_batch* b, float* ms) {
    if (!b || !ms) return fail(LYAP_ERR_INVALID, "lyap_batch_last_ms: NULL argument");
    if (!b->timed) return fail(LYAP_ERR_INVALID, "the batch has not been launched");
    hipError_t e = hipEventSynchronize(b->ev1);
    if (e != hipSuccess) return hip_fail(e, "hipEventSynchronize");
    if ((e = hipEventElapsedTime(ms, b->ev0, b->ev1)) != hipSuccess)
        return hip_fail(e, "hipEventElapsedTime");
    return LYAP_SUCCESS;
}

int lyap_batch_info(lyap_batch* b, int64_t* waves, int* n_classes) {
    if (!b) return fail(LYAP_ERR_INVALID, "lyap_batch_info: NULL batch");
    if (waves) *waves = b->waves;
    if (n_classes) *n_classes = b->n;
    return LYAP_SUCCESS;
}

int lyap_batch_destroy(lyap_batch* b) {
    batch_free(b);
    return LYAP_SUCCESS;
}

int lyap_batch_layout_ppw(int layout, int dim, int* ppw) {
    if (!ppw) return fail(LYAP_ERR_INVALID, "lyap_batch_layout_ppw: NULL output");
    const int v = lyap::batch_points_per_wave(layout, dim);
    if (v <= 0) return fail(LYAP_ERR_INVALID, "no batch layout %d", layout);
    *ppw = v;
    return LYAP_SUCCESS;
}

}  // extern "C"

// ---------------------------------------------------------------------------
// batched launch that allocates nothing
// ---------------------------------------------------------------------------
// The class records and work counters of lyap_batch_run_noalloc live in a
// file-scope device array: records contiguous from offset 0, counters from
// kNoallocCounterOffset.  Each device holds its own copy; host code reaches it
// only through the address hipGetSymbolAddress returns on that device.
namespace {

constexpr size_t kNoallocCounterOffset =
    ((lyap::kMaxBatchClasses * sizeof(lyap::BatchClass) + 255) / 256) * 256;
constexpr size_t kNoallocScratchBytes =
    kNoallocCounterOffset + lyap::kMaxBatchClasses * kCounterStride * sizeof(unsigned long long);
constexpr int kNoallocMaxDevices = 64;

}  // namespace

static __device__ __align__(256) unsigned char scratch[kNoallocScratchBytes];

// one class record, passed by value (kernel arguments: no host-to-device copy
// and no staging buffer), written into the scratch array
__global__ void k_stage_batch_class(lyap::BatchClass rec, lyap::BatchClass* dst) {
    if (blockIdx.x == 0 && threadIdx.x == 0) *dst = rec;
}

namespace {

std::mutex g_scratch_mu;
unsigned char* g_scratch_ptr[kNoallocMaxDevices] = {};
int g_scratch_cus[kNoallocMaxDevices] = {};

// the device's scratch address and CU count, looked up once per device with
// that device current (the caller has made it current)
int noalloc_scratch(int device, unsigned char** ptr, int* cus) {
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_ptr[device]) {
        void* p = nullptr;
        hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return hip_fail(e, "hipGetSymbolAddress(scratch)");
        int n = 0;
        if ((e = hipDeviceGetAttribute(&n, hipDeviceAttributeMultiprocessorCount, device)) != hipSuccess)
            return hip_fail(e, "hipDeviceGetAttribute");
        g_scratch_ptr[device] = static_cast<unsigned char*>(p);
        g_scratch_cus[device] = n;
    }
    *ptr = g_scratch_ptr[device];
    *cus = g_scratch_cus[device];
    return LYAP_SUCCESS;
}

}  // namespace

extern "C" int lyap_batch_run_noalloc(int device, const lyap_batch_class* classes, int n_classes,
                                      const lyap_batch_tuning* tuning, void* stream) {
    if (!classes) return fail(LYAP_ERR_INVALID, "lyap_batch_run_noalloc: NULL classes");
    if (n_classes > lyap::kMaxBatchClasses)
        return fail(LYAP_ERR_INVALID,
                    "lyap_batch_run_noalloc: %d classes exceed the scratch array's %d records (%zu bytes)",
                    n_classes, lyap::kMaxBatchClasses, kNoallocScratchBytes);
    int count = 0;
    hipError_t e = hipGetDeviceCount(&count);
    if (e != hipSuccess) return hip_fail(e, "hipGetDeviceCount");
    if (device < 0 || device >= count || device >= kNoallocMaxDevices)
        return fail(LYAP_ERR_INVALID, "lyap_batch_run_noalloc: device %d of %d", device, count);
    BatchPlan plan;
    int rc = plan_batch(classes, n_classes, tuning, &plan);
    if (rc != LYAP_SUCCESS) return rc;
    if (plan.n == 0) return LYAP_SUCCESS;
    if ((e = hipSetDevice(device)) != hipSuccess) return hip_fail(e, "hipSetDevice");
    unsigned char* base = nullptr;
    int cus = 0;
    if ((rc = noalloc_scratch(device, &base, &cus)) != LYAP_SUCCESS) return rc;
    int64_t waves = 0;
    if ((e = plan_waves(plan, cus, &waves)) != hipSuccess) return hip_fail(e, "batch occupancy");
    lyap::BatchClass* d_cls = reinterpret_cast<lyap::BatchClass*>(base);
    unsigned long long* d_counters = reinterpret_cast<unsigned long long*>(base + kNoallocCounterOffset);
    hipStream_t s = (hipStream_t)stream;
    for (int j = 0; j < plan.n; ++j) {
        plan.recs[j].io.counter = d_counters + kCounterStride * j;
        hipLaunchKernelGGL(k_stage_batch_class, dim3(1), dim3(64), 0, s, plan.recs[j], d_cls + j);
    }
    if ((e = hipGetLastError()) != hipSuccess) return hip_fail(e, "batch class staging");
    // the counters are zeroed by the launch's own reset kernel
    e = lyap::launch_batch(d_cls, plan.n, plan.integ, plan.adapt != 0, plan.prec == LYAP_PREC_MIXED,
                           plan.tuning.layout, plan.tuning.occupancy3 != 0, waves, s);
    if (e != hipSuccess) return hip_fail(e, "batch launch");
    return LYAP_SUCCESS;
}
