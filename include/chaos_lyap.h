/*
 * chaos_lyap.h — C-ABI of the MI355X (gfx950) Lyapunov-spectrum sweep.
 *
 * This is the drop-in boundary under the reference's TaskLyapunov: every entry
 * point takes plain pointers and sizes, reports errors as status codes (the
 * message is in lyap_last_error()), never throws, and launches on the HIP
 * stream it is given (NULL = the null stream).  Device buffers are owned by
 * the caller; scratch (work counters) is owned by the context.
 *
 * Reference interfaces replaced (paths relative to the reference root):
 *   lyap_cuda_lyapunov    <- cudaLyapunov,            cuda/lyapunov.dp.hpp:10-17,
 *                                                      cuda/lyapunov.dp.cpp:321-367
 *   lyap_sweep_run        <- kernLEDverk_ launch,      cuda/lyapunov.dp.cpp:262-318
 *                            (plus the host grid of TaskLyapunov::allocate,
 *                             tasklyapunov.cpp:45-71, computed on the device)
 *   lyap_model_lookup     <- ModelFactoryRegistry::Storage lookup,
 *                            model_factory.h:28-69, tasklyapunov.cpp:15-16
 *   lyap_loop_count       <- the `for (double t = t0; t < T; t += h)` step
 *                            count of cuda/lyapunov.dp.cpp:11,139,158,244
 *   lyap_desc_finalize    <- TaskLyapunov::init/allocate's derived scalars,
 *                            tasklyapunov.cpp:17-31,45-46,65
 */
#ifndef CHAOS_LYAP_H
#define CHAOS_LYAP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define LYAP_MAX_DIM 32
#define LYAP_MAX_PARAMS 8

/* status codes of the entry points */
enum {
    LYAP_SUCCESS = 0,
    LYAP_ERR_INVALID = -1,     /* bad argument or descriptor */
    LYAP_ERR_HIP = -2,         /* a HIP runtime call failed */
    LYAP_ERR_UNSUPPORTED = -3, /* valid request this build has no kernel for */
    LYAP_ERR_NOT_FOUND = -4    /* unknown model name */
};

/* models compiled into the library (ModelFactoryRegistry replacement) */
enum {
    LYAP_MODEL_SHIMIZU = 0, /* reference cuda/models/flows/shimizu.dp.cpp; params (alpha, lambda) */
    LYAP_MODEL_LORENZ = 1,  /* params (sigma, rho, beta) */
    LYAP_MODEL_ROSSLER = 2, /* params (a, c, b) */
    LYAP_MODEL_L96 = 3,     /* Lorenz-96, N in {8,16,32}; params (F, d) */
    LYAP_MODEL_HENON = 4,   /* Henon map, N = 2; params (a, b) */
    LYAP_MODEL_LOGISTIC = 5 /* logistic map, N = 1; params (r) */
};
/* ModelType of model_factory.h:22-26 */
enum { LYAP_TYPE_FLOW = 0, LYAP_TYPE_MAP = 1 };

/* how the spectrum is computed */
enum {
    LYAP_METHOD_VAR = 0,     /* coupled variational RK, MGS on exact time boundaries */
    LYAP_METHOD_VAR_REF = 1, /* reference VAR mode, cuda/lyapunov.dp.cpp:133-226 */
    LYAP_METHOD_QR_REF = 2,  /* reference QR mode,  cuda/lyapunov.dp.cpp:229-259 */
    LYAP_METHOD_FD_REF = 3,  /* reference FD mode,  cuda/lyapunov.dp.cpp:5-129 */
    /* LYAP_METHOD_VAR with N - 1 deviation vectors; the last exponent closes
       the spectrum through Liouville's formula, sum_i lambda_i = <tr J>, the
       trace integrated along the base trajectory by the step's own RK
       quadrature.  The 3-D flows only, n_exps == dimension; the first N - 1
       exponents are those of LYAP_METHOD_VAR bit for bit.  No counterpart in
       the reference (which integrates every vector). */
    LYAP_METHOD_VAR_TRACE = 4
};
enum { LYAP_INT_DVERK = 0, LYAP_INT_DOPRI5 = 1 };
enum { LYAP_PREC_FP64 = 0, LYAP_PREC_MIXED = 1 };
enum { LYAP_INIT_NONE = 0, LYAP_INIT_UNIFORM = 1, LYAP_INIT_NORMAL = 2 };

/* per-point status word */
enum {
    LYAP_PT_OK = 0,
    LYAP_PT_NONFINITE = 1,
    LYAP_PT_STEP_UNDERFLOW = 2,
    LYAP_PT_MAXSTEPS = 3,
    LYAP_PT_ESCAPED = 4,
    LYAP_PT_SKIPPED = 5 /* not integrated: its skip word was set (lyap_ctx_set_watch) */
};

/* One sweep.  Field meanings follow LyapunovParameters (tool_params.h:29-43);
   the derived fields are filled by lyap_desc_finalize(). */
typedef struct lyap_sweep_desc {
    int32_t model, dim, n_exps, n_params;
    int32_t method, integrator, adaptive, precision;
    double h;           /* integration_step (initial step when adaptive) */
    double rtol, atol;  /* adaptive tolerances */
    double h_min, h_max;
    double eps;         /* epsilon of the reference modes */
    double t_transient; /* skip_time_attractor */
    double t_slave;     /* skip_time_slave_trajectory */
    double t_average;   /* total_time */
    double t_renorm;    /* skip_time_normalization */
    /* derived, reference modes: step counts of the float-accumulated loops */
    int64_t ref_steps_transient, ref_steps_slave, ref_steps_average;
    double add_skip_steps; /* t_renorm / h, compared with an int (tasklyapunov.cpp:65) */
    /* derived, coupled-variational method */
    int64_t steps_per_renorm; /* fixed step */
    int64_t n_renorm_transient, n_renorm_average;
    int64_t max_steps;    /* per point; exceeded -> LYAP_PT_MAXSTEPS */
    double escape_radius; /* |x_i| above it -> LYAP_PT_ESCAPED */
    /* sweep grid, p1 outer (tasklyapunov.cpp:45-63): p = start + i*delta */
    int32_t sweep_idx1, sweep_idx2, n1, n2;
    double start1, delta1, start2, delta2;
    double base_params[LYAP_MAX_PARAMS]; /* initial_model_params */
    /* initial state: init_state (+ p[init_param_offset]) + seeded perturbation */
    double init_state[LYAP_MAX_DIM];
    int32_t init_kind, init_param_offset;
    double init_scale;
    uint64_t seed;
} lyap_sweep_desc;

typedef struct lyap_model_info {
    int32_t model;    /* LYAP_MODEL_* */
    int32_t dim;      /* _dimension */
    int32_t n_params; /* _param_number */
    int32_t type;     /* LYAP_TYPE_* */
    const char* name;
    const char* param_names;
    const char* phase_names;
    const char* equations;
} lyap_model_info;

typedef struct lyap_ctx lyap_ctx;

/* message of the last failing call on this thread ("" if none) */
const char* lyap_last_error(void);

/* a context per (host thread, device); owns the work counters, the staging
   buffers of heavy-first runs and the watch counters, so runs on one context
   follow each other on a stream: two runs of one context in flight at once on
   different streams share them and race (use a context per concurrent run) */
int lyap_ctx_create(int device, lyap_ctx** out);
int lyap_ctx_destroy(lyap_ctx* ctx);

/* model table */
int lyap_model_count(void);
int lyap_model_at(int i, lyap_model_info* out);
int lyap_model_lookup(const char* name, lyap_model_info* out);

/* host helpers */
int64_t lyap_loop_count(double t0, double t1, double h);
int lyap_desc_finalize(lyap_sweep_desc* d);

/*
 * Run points [first, first+count) of the sweep (or the points listed in
 * d_idx_list[0..count) when it is not NULL).  Every pointer is device memory:
 *   d_params  : optional, [n_total * n_params] AoS parameters (else the grid)
 *   d_inits   : optional, [n_total * dim] AoS initial states (else the seeded rule)
 *   d_lambda  : [count * n_exps] output, row i belongs to work item i
 *   d_status, d_n_acc, d_n_rej : optional [count] outputs
 * The call is stream-ordered and returns before the kernels finish.
 */
int lyap_sweep_run(lyap_ctx* ctx, const lyap_sweep_desc* d, int64_t first, int64_t count,
                   const int64_t* d_idx_list, const double* d_params, const double* d_inits,
                   double* d_lambda, uint32_t* d_status, uint32_t* d_n_acc, uint32_t* d_n_rej,
                   void* stream);

/*
 * Launch-shape options of a context (no counterpart in the reference, whose
 * launch shape was fixed at cuda/lyapunov.dp.cpp:331-332); set by callers or
 * by the autotuner.  Keys:
 *   "l96_waves_8", "l96_waves_16", "l96_waves_32": wave64s per point of the
 *       Lorenz-96 kernel (N=8: 1; N=16: 1, 2, 4; N=32: 2, 4; 0 = default)
 *   "l96_ppw_8", "l96_ppw_16", "l96_ppw_32": points per wave64 of the
 *       wave-level Lorenz-96 drain (N=8: 1, 2, 4, 8; N=16: 1, 2, 4; N=32: 1, 2);
 *       0 (default) = the block-per-point kernel above
 *   "lane_max_blocks": cap on the persistent grid of the 3-D kernels (0 = none;
 *       -k = the occupancy-sized grid less k blocks, whose slots stay free for
 *       kernels on other streams; at least one block runs; -65536 and above)
 *   "l96_max_blocks": cap on the persistent grid of the Lorenz-96 kernels (0 = none)
 *   "lane_gather": 3-D kernels: lanes at a renormalisation wait until this many
 *       lanes of the wave wait, then renormalise together (0 or 1 = at once;
 *       results are the same bits either way)
 *   "item_order": order in which the persistent kernels take the points of a
 *       sweep: 0 (default) spread over the grid (item i is point i*s mod count,
 *       s ~ count/phi coprime to count), 1 in order (point i, i.e. the order of
 *       idx_list: a list sorted by decreasing expected cost schedules the
 *       longest points first); results are the same bits either way
 *   "order_probe": 1..500 runs adaptive sweeps of the grid (the sweep's
 *       method, no idx_list, not watched) heavy-first: a probe run of every
 *       point over this many thousandths of the integration times counts its
 *       steps, the points are sorted by the largest count in their grid
 *       neighbourhood ("order_radius"), most first, and the full
 *       run takes them in that order (its outputs scattered back to point
 *       order); 0 (default) off.  Pays where the grid holds more points than
 *       the launch has lanes, so that points start in later rounds; the run
 *       allocates staging and sort buffers on the context (kept for later
 *       runs).  lyap_last_kernel_ms times the full run's launch alone (not
 *       the probe, sort or scatter);
 *       results are the same bits either way
 *   "order_radius": 0..8 (default 2): with order_probe, a point of the grid
 *       is ordered by the largest probe count over the grid points within
 *       this many rows and columns of it (0: its own count; runs with caller's
 *       parameter rows use their own counts): a short probe misjudges single
 *       points, not the regions of the plane that are heavy
 *   "lane_quad": 1 runs the sweep's method on the 3-D flows with four lanes
 *       per point (each lane one deviation vector; base step and control
 *       repeated in the four lanes), for launches of about one point per lane
 *       whose length is their slowest point; 0 (default) one lane per point;
 *       results are the same bits either way
 */
int lyap_ctx_set_option(lyap_ctx* ctx, const char* key, int64_t value);

/* duration of the last sweep kernel launched by lyap_sweep_run on ctx
   (HIP events on the launch stream; valid once that stream has drained) */
int lyap_last_kernel_ms(lyap_ctx* ctx, float* ms);

/*
 * Progress signals and skips for the next lyap_sweep_run on ctx (that run
 * only).  No counterpart in the reference, whose launches were synchronous
 * (tasklyapunov.cpp:200-215); it lets a caller consume the first results of a
 * long launch while it runs, and drop points it no longer needs.  The run must
 * be the sweep's method (LYAP_METHOD_VAR or _VAR_TRACE) with adaptive steps on
 * a 3-D flow, one lane per point (not lane_quad).  Items are work items (list positions;
 * item_order 1 takes them in that order):
 *   items [0, end0) form milestone 0 and [end0, end1) milestone 1; once every
 *       item of a milestone has finished, its outputs are in device memory and
 *       lyap_ctx_watch_wait reports the milestone reached, while the launch
 *       goes on (end1 <= the run's count);
 *   d_skip (optional, [count] words, device memory): an item whose word is
 *       nonzero when a lane takes it is not integrated (status
 *       LYAP_PT_SKIPPED, exponents NaN, step counts 0); words written while
 *       the launch runs (by a kernel on another stream) are seen by the items
 *       taken after they land.
 */
int lyap_ctx_set_watch(lyap_ctx* ctx, int64_t end0, int64_t end1, const uint32_t* d_skip);

/* Waits on the host, up to timeout_s seconds, until milestone 0 or 1 of the
   last watched run on ctx is reached; *reached = 1 if it was (an empty
   milestone is reached at once), 0 on time-out. */
int lyap_ctx_watch_wait(lyap_ctx* ctx, int milestone, double timeout_s, int* reached);

/*
 * Reference-shaped launch: the arguments of cudaLyapunov
 * (cuda/lyapunov.dp.hpp:10-17) minus the device function pointers (the model
 * is named by id) and minus the per-work-item scratch (the kernels keep their
 * state in registers).  d_inits/d_params are AoS per point as in
 * tasklyapunov.cpp:47-71; d_lyap_exps receives task_num * exps_num values.
 * is_qr/is_var select the reference QR/VAR/FD mode; block_num/block_size are
 * accepted and ignored, like the reference's own override
 * (cuda/lyapunov.dp.cpp:331-332).
 */
int lyap_cuda_lyapunov(lyap_ctx* ctx, int block_num, int block_size, size_t task_num, int is_qr,
                       int is_var, int model_type, const double* d_inits, int32_t dimension,
                       int32_t model, const double* d_params, int32_t params_dim, double eps,
                       double integration_step, double* d_lyap_exps, int32_t exps_num,
                       double time_skip, double time_skip_clp, double calc_time,
                       double add_skip_steps, void* stream);

/*
 * Batched mixed-model sweep (no counterpart in the reference, which launches
 * one model per cudaLyapunov call, cuda/lyapunov.dp.cpp:321-367).  Up to 16
 * classes, each one model and dimension with its own descriptor, points and
 * outputs, run by ONE persistent launch: one-wave workgroups walk the classes
 * in queue order and drain each through its own work counter (3-D flows one
 * point per lane, Lorenz-96 N = 8/16/32 several points per wave).  The counter
 * reset (a kernel) and the launch are captured once into a hipGraph at create
 * time; lyap_batch_launch replays it on the caller's stream.  All classes share
 * the integrator, the adaptive flag and the precision; the method is VAR, or
 * VAR_TRACE for a 3-D class (classes may differ in that).
 * Buffers are device memory owned by the caller and must stay valid for the
 * life of the batch (the graph holds their addresses).
 */
typedef struct lyap_batch_class {
    lyap_sweep_desc desc;    /* finalised by lyap_batch_create */
    int64_t count;           /* points of the class (0: class skipped) */
    const double* d_params;  /* optional [count * n_params] AoS, else grid points 0..count */
    const double* d_inits;   /* optional [count * dim] AoS, else the seeded rule */
    double* d_lambda;        /* [count * n_exps] */
    uint32_t* d_status;      /* optional [count] */
    uint32_t* d_n_acc;       /* optional [count] */
    uint32_t* d_n_rej;       /* optional [count] */
} lyap_batch_class;

typedef struct lyap_batch_tuning {
    int32_t layout;        /* Lorenz-96 points per wave for N = 8/16/32: 0 -> 8/4/1, 1 -> 4/2/1 */
    int32_t occupancy3;    /* 1: kernel built for 3 waves per SIMD (168 VGPRs, may spill) */
    int32_t waves_per_cu;  /* persistent grid = min(waves_per_cu, fit) * CUs; 0 = as many as fit */
    int32_t order;         /* queue order: 0 longest task first (3-D lanes, then L96 32/16/8),
                              1 as given, 2 Lorenz-96 classes first */
    int32_t gather;        /* 3-D lanes at a renormalisation wait until this many lanes of the
                              wave wait, then renormalise together (0, 1: at once) */
} lyap_batch_tuning;

typedef struct lyap_batch lyap_batch;

/* tuning NULL = all zero; layouts other than 0 and occupancy3 are built for
   adaptive Dopri5 only */
int lyap_batch_create(lyap_ctx* ctx, const lyap_batch_class* classes, int n_classes,
                      const lyap_batch_tuning* tuning, lyap_batch** out);
/* replays the captured graph on `stream` (stream-ordered, returns at once) */
int lyap_batch_launch(lyap_batch* b, void* stream);
/* duration of the last replay (events on the launch stream; waits for it) */
int lyap_batch_last_ms(lyap_batch* b, float* ms);
/* persistent grid size and the number of non-empty classes */
int lyap_batch_info(lyap_batch* b, int64_t* waves, int* n_classes);
int lyap_batch_destroy(lyap_batch* b);
/* points per wave64 of batch layout `layout` for a class of dimension dim
   (Lorenz-96 N = 8, 16, 32; 64 for the per-lane 3-D classes); host only */
int lyap_batch_layout_ppw(int layout, int dim, int* ppw);

/*
 * The batched launch without any allocation (no counterpart in the
 * reference): for callers that cannot allocate device memory or create
 * runtime objects while they run — code under HIP graph capture, tight
 * serving loops, harnesses that forbid hipMalloc.  It runs the same kernels
 * as lyap_batch_create + lyap_batch_launch, with the same results, but keeps
 * the class records and work counters in a fixed device array of the library
 * (at most 16 classes; 12,032 bytes per device, held until the process ends),
 * staged by kernels on `stream`, and launches directly on `stream` (no graph
 * of its own; the caller may capture the call).  The caller still owns every
 * input and output buffer.  Every call on a device shares that device's
 * array: calls that overlap on one device (other streams, other threads)
 * race; order them on one stream.  More than 16 classes is an error
 * (LYAP_ERR_INVALID).  The first call on a device looks the array up
 * (hipGetSymbolAddress) and reads the CU count; no call allocates.
 */
int lyap_batch_run_noalloc(int device, const lyap_batch_class* classes, int n_classes,
                           const lyap_batch_tuning* tuning, void* stream);

/*
 * Surrogate network (no counterpart in the reference): (p1, p2) -> 2*n_fourier
 * Fourier features -> three hidden layers of `hidden` GELU units ->
 * [mu1, mu2, mu3, log-variance of lambda1, escape logit].  BF16 weights with
 * an FP32 master copy, FP32 accumulation on MFMA, Adam.  Inputs are device
 * arrays: p as [n][2] doubles, lambda as [n][3] doubles, status as [n] words
 * (non-zero: the point escaped or failed; its lambdas are not used).
 */
typedef struct lyap_mlp_config {
    int32_t max_batch;    /* rows per call, a multiple of 64 */
    int32_t hidden;       /* a multiple of 64 */
    int32_t n_fourier;    /* 2*n_fourier a multiple of 16 */
    int32_t reserved0;
    float fourier_sigma;  /* std of the Fourier frequencies (inputs scaled to [0,1]) */
    float lr, beta1, beta2, eps;
    float reserved1;
    uint64_t seed;
    double lo1, hi1, lo2, hi2;       /* input ranges */
    double y_mean[3], y_std[3];      /* target normalisation */
} lyap_mlp_config;

typedef struct lyap_mlp lyap_mlp;

int lyap_mlp_create(int device, const lyap_mlp_config* cfg, lyap_mlp** out);
int lyap_mlp_destroy(lyap_mlp* m);
/* forward + backward into the gradient buffer (no update); mean loss to *d_loss */
int lyap_mlp_backward(lyap_mlp* m, const double* d_p, const double* d_lambda,
                      const uint32_t* d_status, int64_t n, float* d_loss, void* stream);
/* lyap_mlp_backward followed by one Adam step */
int lyap_mlp_train_step(lyap_mlp* m, const double* d_p, const double* d_lambda,
                        const uint32_t* d_status, int64_t n, float* d_loss, void* stream);
/* d_out [n][5]: mu1, mu2, mu3 (lambda units), std of lambda1, escape probability */
int lyap_mlp_forward(lyap_mlp* m, const double* d_p, int64_t n, float* d_out, void* stream);
int lyap_mlp_param_count(lyap_mlp* m, int64_t* n);
/* which: 0 parameters, 1 gradients, 2 Fourier frequencies [n_fourier][2] (host copy) */
int lyap_mlp_get(lyap_mlp* m, int which, float* host, int64_t n);
int lyap_mlp_set_params(lyap_mlp* m, const float* host, int64_t n);

#ifdef __cplusplus
}
#endif
#endif /* CHAOS_LYAP_H */
