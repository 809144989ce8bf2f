"""ctypes bindings of liblyap.so (include/chaos_lyap.h).

The library is the product: the HIP kernels and the C-ABI.  This module only
marshals descriptors and device pointers; device memory comes from torch
(plumbing).  Loading fails loudly when the library has not been built — there
is no CPU fallback.
"""
from __future__ import annotations

import ctypes
import os

PKG = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(PKG, "liblyap.so")

MAX_DIM = 32
MAX_PARAMS = 8

SUCCESS, ERR_INVALID, ERR_HIP, ERR_UNSUPPORTED, ERR_NOT_FOUND = 0, -1, -2, -3, -4
MODEL_SHIMIZU, MODEL_LORENZ, MODEL_ROSSLER, MODEL_L96 = 0, 1, 2, 3
MODEL_HENON, MODEL_LOGISTIC = 4, 5
TYPE_FLOW, TYPE_MAP = 0, 1
METHOD_VAR, METHOD_VAR_REF, METHOD_QR_REF, METHOD_FD_REF, METHOD_VAR_TRACE = 0, 1, 2, 3, 4
INT_DVERK, INT_DOPRI5 = 0, 1
PREC_FP64, PREC_MIXED = 0, 1
INIT_NONE, INIT_UNIFORM, INIT_NORMAL = 0, 1, 2
PT_OK, PT_NONFINITE, PT_STEP_UNDERFLOW, PT_MAXSTEPS, PT_ESCAPED, PT_SKIPPED = 0, 1, 2, 3, 4, 5

# every entry point include/chaos_lyap.h declares
EXPORTS = (
    "lyap_last_error", "lyap_ctx_create", "lyap_ctx_destroy", "lyap_model_count",
    "lyap_model_at", "lyap_model_lookup", "lyap_loop_count", "lyap_desc_finalize",
    "lyap_sweep_run", "lyap_last_kernel_ms", "lyap_cuda_lyapunov", "lyap_ctx_set_option",
    "lyap_mlp_create", "lyap_mlp_destroy", "lyap_mlp_backward", "lyap_mlp_train_step",
    "lyap_mlp_forward", "lyap_mlp_param_count", "lyap_mlp_get", "lyap_mlp_set_params",
    "lyap_batch_create", "lyap_batch_launch", "lyap_batch_last_ms", "lyap_batch_info",
    "lyap_batch_layout_ppw",
    "lyap_batch_destroy", "lyap_batch_run_noalloc", "lyap_ctx_set_watch", "lyap_ctx_watch_wait",
)


class SweepDesc(ctypes.Structure):
    _fields_ = [
        ("model", ctypes.c_int32), ("dim", ctypes.c_int32),
        ("n_exps", ctypes.c_int32), ("n_params", ctypes.c_int32),
        ("method", ctypes.c_int32), ("integrator", ctypes.c_int32),
        ("adaptive", ctypes.c_int32), ("precision", ctypes.c_int32),
        ("h", ctypes.c_double), ("rtol", ctypes.c_double), ("atol", ctypes.c_double),
        ("h_min", ctypes.c_double), ("h_max", ctypes.c_double), ("eps", ctypes.c_double),
        ("t_transient", ctypes.c_double), ("t_slave", ctypes.c_double),
        ("t_average", ctypes.c_double), ("t_renorm", ctypes.c_double),
        ("ref_steps_transient", ctypes.c_int64), ("ref_steps_slave", ctypes.c_int64),
        ("ref_steps_average", ctypes.c_int64), ("add_skip_steps", ctypes.c_double),
        ("steps_per_renorm", ctypes.c_int64),
        ("n_renorm_transient", ctypes.c_int64), ("n_renorm_average", ctypes.c_int64),
        ("max_steps", ctypes.c_int64), ("escape_radius", ctypes.c_double),
        ("sweep_idx1", ctypes.c_int32), ("sweep_idx2", ctypes.c_int32),
        ("n1", ctypes.c_int32), ("n2", ctypes.c_int32),
        ("start1", ctypes.c_double), ("delta1", ctypes.c_double),
        ("start2", ctypes.c_double), ("delta2", ctypes.c_double),
        ("base_params", ctypes.c_double * MAX_PARAMS),
        ("init_state", ctypes.c_double * MAX_DIM),
        ("init_kind", ctypes.c_int32), ("init_param_offset", ctypes.c_int32),
        ("init_scale", ctypes.c_double),
        ("seed", ctypes.c_uint64),
    ]

    def to_dict(self) -> dict:
        out = {}
        for name, _ in self._fields_:
            v = getattr(self, name)
            out[name] = list(v) if name in ("base_params", "init_state") else v
        return out


class BatchClassRec(ctypes.Structure):
    """lyap_batch_class"""
    _fields_ = [("desc", SweepDesc), ("count", ctypes.c_int64), ("d_params", ctypes.c_void_p),
                ("d_inits", ctypes.c_void_p), ("d_lambda", ctypes.c_void_p),
                ("d_status", ctypes.c_void_p), ("d_n_acc", ctypes.c_void_p),
                ("d_n_rej", ctypes.c_void_p)]


class BatchTuning(ctypes.Structure):
    """lyap_batch_tuning"""
    _fields_ = [("layout", ctypes.c_int32), ("occupancy3", ctypes.c_int32),
                ("waves_per_cu", ctypes.c_int32), ("order", ctypes.c_int32),
                ("gather", ctypes.c_int32)]


class ModelInfo(ctypes.Structure):
    _fields_ = [
        ("model", ctypes.c_int32), ("dim", ctypes.c_int32), ("n_params", ctypes.c_int32),
        ("type", ctypes.c_int32), ("name", ctypes.c_char_p), ("param_names", ctypes.c_char_p),
        ("phase_names", ctypes.c_char_p), ("equations", ctypes.c_char_p),
    ]


class MlpConfig(ctypes.Structure):
    _fields_ = [
        ("max_batch", ctypes.c_int32), ("hidden", ctypes.c_int32), ("n_fourier", ctypes.c_int32),
        ("reserved0", ctypes.c_int32), ("fourier_sigma", ctypes.c_float), ("lr", ctypes.c_float),
        ("beta1", ctypes.c_float), ("beta2", ctypes.c_float), ("eps", ctypes.c_float),
        ("reserved1", ctypes.c_float), ("seed", ctypes.c_uint64),
        ("lo1", ctypes.c_double), ("hi1", ctypes.c_double), ("lo2", ctypes.c_double),
        ("hi2", ctypes.c_double), ("y_mean", ctypes.c_double * 3), ("y_std", ctypes.c_double * 3),
    ]


class LyapError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"[{code}] {msg}")
        self.code = code


_lib = None


def load(path: str = LIB_PATH):
    """Load liblyap.so; raises if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise ImportError(f"{path} is missing: run chaos_oneapi_amd.build.build() first "
                          "(the sweep has no CPU fallback)")
    # torch ships its own libamdhip64.so.7; importing it first makes liblyap's
    # NEEDED entry bind to that already-loaded runtime, so both share one HIP
    # runtime (device buffers and streams pass between them).  Loading liblyap
    # first would bring in a second runtime, after which torch sees no GPU.
    import torch  # noqa: F401
    L = ctypes.CDLL(path)
    P, i32, i64, d, vp = ctypes.POINTER, ctypes.c_int32, ctypes.c_int64, ctypes.c_double, ctypes.c_void_p
    L.lyap_last_error.restype = ctypes.c_char_p
    L.lyap_ctx_create.argtypes = [ctypes.c_int, P(vp)]
    L.lyap_ctx_destroy.argtypes = [vp]
    L.lyap_model_count.restype = ctypes.c_int
    L.lyap_model_at.argtypes = [ctypes.c_int, P(ModelInfo)]
    L.lyap_model_lookup.argtypes = [ctypes.c_char_p, P(ModelInfo)]
    L.lyap_loop_count.restype = i64
    L.lyap_loop_count.argtypes = [d, d, d]
    L.lyap_desc_finalize.argtypes = [P(SweepDesc)]
    L.lyap_sweep_run.argtypes = [vp, P(SweepDesc), i64, i64, vp, vp, vp, vp, vp, vp, vp, vp]
    L.lyap_last_kernel_ms.argtypes = [vp, P(ctypes.c_float)]
    L.lyap_ctx_set_option.argtypes = [vp, ctypes.c_char_p, i64]
    L.lyap_ctx_set_watch.argtypes = [vp, i64, i64, vp]
    L.lyap_ctx_watch_wait.argtypes = [vp, ctypes.c_int, d, P(ctypes.c_int)]
    L.lyap_mlp_create.argtypes = [ctypes.c_int, P(MlpConfig), P(vp)]
    L.lyap_mlp_destroy.argtypes = [vp]
    L.lyap_mlp_backward.argtypes = [vp, vp, vp, vp, i64, vp, vp]
    L.lyap_mlp_train_step.argtypes = [vp, vp, vp, vp, i64, vp, vp]
    L.lyap_mlp_forward.argtypes = [vp, vp, i64, vp, vp]
    L.lyap_mlp_param_count.argtypes = [vp, P(i64)]
    L.lyap_mlp_get.argtypes = [vp, ctypes.c_int, vp, i64]
    L.lyap_mlp_set_params.argtypes = [vp, vp, i64]
    L.lyap_batch_create.argtypes = [vp, P(BatchClassRec), ctypes.c_int, P(BatchTuning), P(vp)]
    L.lyap_batch_launch.argtypes = [vp, vp]
    L.lyap_batch_last_ms.argtypes = [vp, P(ctypes.c_float)]
    L.lyap_batch_info.argtypes = [vp, P(i64), P(ctypes.c_int)]
    L.lyap_batch_layout_ppw.argtypes = [ctypes.c_int, ctypes.c_int, P(ctypes.c_int)]
    L.lyap_batch_destroy.argtypes = [vp]
    L.lyap_batch_run_noalloc.argtypes = [ctypes.c_int, P(BatchClassRec), ctypes.c_int, P(BatchTuning), vp]
    L.lyap_cuda_lyapunov.argtypes = [
        vp, ctypes.c_int, ctypes.c_int, ctypes.c_size_t, ctypes.c_int, ctypes.c_int, ctypes.c_int,
        vp, i32, i32, vp, i32, d, d, vp, i32, d, d, d, d, vp]
    _lib = L
    return L


def check(rc: int) -> None:
    if rc != SUCCESS:
        raise LyapError(rc, load().lyap_last_error().decode())


def make_desc(spec: dict) -> SweepDesc:
    d = SweepDesc()
    for name, _ in SweepDesc._fields_:
        if name not in spec:
            continue
        v = spec[name]
        if name in ("base_params", "init_state"):
            arr = getattr(d, name)
            for j, x in enumerate(v):
                arr[j] = float(x)
        else:
            setattr(d, name, v)
    return d


def finalize(desc: SweepDesc) -> SweepDesc:
    """Validate and fill the derived fields (host only, no GPU needed)."""
    out = SweepDesc.from_buffer_copy(desc)
    check(load().lyap_desc_finalize(ctypes.byref(out)))
    return out


def loop_count(t0: float, t1: float, h: float) -> int:
    return int(load().lyap_loop_count(float(t0), float(t1), float(h)))


def model_lookup(name: str) -> ModelInfo:
    info = ModelInfo()
    check(load().lyap_model_lookup(name.encode(), ctypes.byref(info)))
    return info


def models() -> list:
    L = load()
    out = []
    for i in range(L.lyap_model_count()):
        info = ModelInfo()
        check(L.lyap_model_at(i, ctypes.byref(info)))
        out.append(info)
    return out


def _ptr(t):
    return None if t is None else ctypes.c_void_p(t.data_ptr())


class Context:
    """A lyap_ctx on one device; run() launches on torch's current stream."""

    def __init__(self, device: int = 0):
        self._h = ctypes.c_void_p()
        check(load().lyap_ctx_create(int(device), ctypes.byref(self._h)))
        self.device = device

    def close(self):
        if self._h:
            load().lyap_ctx_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def run(self, desc: SweepDesc, first: int = 0, count: int | None = None, idx_list=None,
            params=None, inits=None, out=None, stream=None):
        """Launch a sweep; returns (lam, status, n_acc, n_rej) device tensors (not synced)."""
        import torch

        if count is None:
            count = len(idx_list) if idx_list is not None else desc.n1 * desc.n2 - first
        dev = torch.device("cuda", self.device)
        if out is None:
            lam = torch.empty((count, desc.n_exps), dtype=torch.float64, device=dev)
            st = torch.empty(count, dtype=torch.int32, device=dev)
            na = torch.empty(count, dtype=torch.int32, device=dev)
            nr = torch.empty(count, dtype=torch.int32, device=dev)
        else:
            lam, st, na, nr = out
        for t, name, dt in ((idx_list, "idx_list", torch.int64), (params, "params", torch.float64),
                            (inits, "inits", torch.float64)):
            if t is not None and not (t.is_cuda and t.is_contiguous() and t.dtype == dt):
                raise ValueError(f"{name} must be a contiguous {dt} CUDA tensor")
        if idx_list is not None and count > idx_list.numel():
            raise ValueError(f"count {count} exceeds the {idx_list.numel()} entries of idx_list")
        watch, self._watch = getattr(self, "_watch", None), None
        if watch is not None:
            if watch[1] > count:
                raise ValueError(f"watched items [0, {watch[1]}) beyond the {count} items of the run")
            if watch[2] is not None and watch[2].numel() < count:
                raise ValueError(f"skip holds {watch[2].numel()} words, {count} needed")
        n_rows = first + count
        if params is not None and params.numel() < n_rows * desc.n_params and idx_list is None:
            raise ValueError(f"params holds {params.numel()} values, {n_rows * desc.n_params} needed")
        if inits is not None and inits.numel() < n_rows * desc.dim and idx_list is None:
            raise ValueError(f"inits holds {inits.numel()} values, {n_rows * desc.dim} needed")
        if out is not None:
            want = ((count, desc.n_exps), (count,), (count,), (count,))
            for t, w in zip(out, want):
                if t is not None and (tuple(t.shape[:len(w)]) != w or not t.is_contiguous()):
                    raise ValueError(f"output tensor of shape {tuple(t.shape)}, {w} expected")
        s = stream if stream is not None else torch.cuda.current_stream(dev)
        if watch is not None:
            check(load().lyap_ctx_set_watch(self._h, watch[0], watch[1], _ptr(watch[2])))
        check(load().lyap_sweep_run(self._h, ctypes.byref(desc), int(first), int(count),
                                    _ptr(idx_list), _ptr(params), _ptr(inits), _ptr(lam),
                                    _ptr(st), _ptr(na), _ptr(nr), ctypes.c_void_p(s.cuda_stream)))
        return lam, st, na, nr

    def set_option(self, key: str, value: int) -> None:
        check(load().lyap_ctx_set_option(self._h, key.encode(), int(value)))

    def set_watch(self, end0: int, end1: int, skip=None) -> None:
        """Progress signals and skips for the next run() on this context:
        milestones items [0, end0) and [end0, end1); skip, an optional
        contiguous int32/uint32 CUDA tensor of one word per item, read while
        the launch runs (nonzero: the item is not integrated, status
        PT_SKIPPED)."""
        import torch

        if skip is not None and not (skip.is_cuda and skip.is_contiguous() and skip.dtype in (torch.int32, torch.uint32)):
            raise ValueError("skip must be a contiguous 32-bit CUDA tensor")
        if not 0 <= end0 <= end1:
            raise ValueError(f"milestones [0, {end0}) and [{end0}, {end1}) are not ranges")
        # armed in the C context by the next run(), which checks the sizes
        self._watch = (int(end0), int(end1), skip)

    def watch_wait(self, milestone: int, timeout_s: float = 120.0) -> bool:
        """Waits on the host until milestone 0 or 1 of the last watched run is
        reached (True) or timeout_s passes (False)."""
        r = ctypes.c_int(0)
        check(load().lyap_ctx_watch_wait(self._h, int(milestone), float(timeout_s), ctypes.byref(r)))
        return bool(r.value)

    def kernel_ms(self) -> float:
        ms = ctypes.c_float()
        check(load().lyap_last_kernel_ms(self._h, ctypes.byref(ms)))
        return float(ms.value)

    def cuda_lyapunov(self, task_num, is_qr, is_var, inits, dimension, model, params, params_dim,
                      eps, h, lyap_exps, exps_num, time_skip, time_skip_clp, calc_time,
                      add_skip_steps, model_type=TYPE_FLOW, stream=None):
        """The reference-shaped launch (cudaLyapunov's arguments); inits,
        params and lyap_exps are float64 CUDA tensors, AoS per point."""
        import torch

        for t, name, need in ((inits, "inits", task_num * dimension), (params, "params", task_num * params_dim),
                              (lyap_exps, "lyap_exps", task_num * exps_num)):
            if not (t.is_cuda and t.is_contiguous() and t.dtype == torch.float64):
                raise ValueError(f"{name} must be a contiguous float64 CUDA tensor")
            if t.numel() < need:
                raise ValueError(f"{name} holds {t.numel()} values, {need} needed")
        s = stream if stream is not None else torch.cuda.current_stream(torch.device("cuda", self.device))
        check(load().lyap_cuda_lyapunov(
            self._h, 32, 32, int(task_num), int(is_qr), int(is_var), int(model_type), _ptr(inits),
            int(dimension), int(model), _ptr(params), int(params_dim), float(eps), float(h),
            _ptr(lyap_exps), int(exps_num), float(time_skip), float(time_skip_clp),
            float(calc_time), float(add_skip_steps), ctypes.c_void_p(s.cuda_stream)))


def batch_layout_ppw(layout: int, dim: int) -> int:
    """Points per wave64 of a batch layout for a class of dimension dim."""
    v = ctypes.c_int()
    check(load().lyap_batch_layout_ppw(int(layout), int(dim), ctypes.byref(v)))
    return int(v.value)


class Batch:
    """A captured batched launch (lyap_batch_*) over classes of points.

    classes: list of dicts with keys desc (SweepDesc), count, and optional
    params / inits (device tensors, AoS).  Output tensors are allocated here
    and stay owned by this object (the captured graph holds their addresses).
    """

    def __init__(self, ctx: Context, classes, layout: int = 0, occupancy3: bool = False,
                 waves_per_cu: int = 0, order: int = 0, gather: int = 0):
        import torch

        dev = torch.device("cuda", ctx.device)
        self.ctx = ctx
        self.outputs = []
        self._keep = []
        recs = (BatchClassRec * len(classes))()
        for j, c in enumerate(classes):
            desc, n = c["desc"], int(c["count"])
            lam = torch.empty((n, desc.n_exps), dtype=torch.float64, device=dev)
            st = torch.empty(n, dtype=torch.int32, device=dev)
            na = torch.empty(n, dtype=torch.int32, device=dev)
            nr = torch.empty(n, dtype=torch.int32, device=dev)
            self.outputs.append((lam, st, na, nr))
            for key in ("params", "inits"):
                t = c.get(key)
                if t is not None:
                    assert t.is_cuda and t.is_contiguous() and t.dtype == torch.float64
                    self._keep.append(t)
            r = recs[j]
            r.desc = desc
            r.count = n
            r.d_params = c["params"].data_ptr() if c.get("params") is not None else None
            r.d_inits = c["inits"].data_ptr() if c.get("inits") is not None else None
            r.d_lambda, r.d_status = lam.data_ptr(), st.data_ptr()
            r.d_n_acc, r.d_n_rej = na.data_ptr(), nr.data_ptr()
        tu = BatchTuning(int(layout), int(bool(occupancy3)), int(waves_per_cu), int(order), int(gather))
        self._h = ctypes.c_void_p()
        check(load().lyap_batch_create(ctx._h, recs, len(classes), ctypes.byref(tu),
                                       ctypes.byref(self._h)))
        waves, ncls = ctypes.c_int64(), ctypes.c_int()
        check(load().lyap_batch_info(self._h, ctypes.byref(waves), ctypes.byref(ncls)))
        self.waves, self.n_classes = int(waves.value), int(ncls.value)

    def launch(self, stream=None):
        import torch

        s = stream if stream is not None else torch.cuda.current_stream(torch.device("cuda", self.ctx.device))
        check(load().lyap_batch_launch(self._h, ctypes.c_void_p(s.cuda_stream)))
        return self.outputs

    def last_ms(self) -> float:
        ms = ctypes.c_float()
        check(load().lyap_batch_last_ms(self._h, ctypes.byref(ms)))
        return float(ms.value)

    def close(self):
        if self._h:
            load().lyap_batch_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def batch_run_noalloc(device: int, classes, layout: int = 0, occupancy3: bool = False,
                      waves_per_cu: int = 0, order: int = 0, gather: int = 0, stream=None):
    """lyap_batch_run_noalloc: the batched launch on the caller's stream with
    the library's fixed scratch array (no allocation in the library).
    classes: dicts with desc, count, optional params / inits, and outputs
    (lam, status, n_acc, n_rej) as device tensors the caller allocated."""
    import torch

    recs = (BatchClassRec * len(classes))()
    for j, c in enumerate(classes):
        lam, st, na, nr = c["outputs"]
        n = int(c["count"])
        if lam.shape != (n, c["desc"].n_exps) or lam.dtype != torch.float64 or not lam.is_contiguous():
            raise ValueError("lambda output must be a contiguous float64 [count, n_exps] tensor")
        for t in (st, na, nr):
            if t is not None and (t.numel() < n or t.dtype != torch.int32):
                raise ValueError("status/step outputs must be int32 tensors of count entries")
        for key in ("params", "inits"):
            t = c.get(key)
            if t is not None and not (t.is_cuda and t.is_contiguous() and t.dtype == torch.float64):
                raise ValueError(f"{key} must be a contiguous float64 CUDA tensor")
        r = recs[j]
        r.desc = c["desc"]
        r.count = n
        r.d_params = c["params"].data_ptr() if c.get("params") is not None else None
        r.d_inits = c["inits"].data_ptr() if c.get("inits") is not None else None
        r.d_lambda, r.d_status = lam.data_ptr(), _ptr(st)
        r.d_n_acc, r.d_n_rej = _ptr(na), _ptr(nr)
    tu = BatchTuning(int(layout), int(bool(occupancy3)), int(waves_per_cu), int(order), int(gather))
    s = stream if stream is not None else torch.cuda.current_stream(torch.device("cuda", device))
    check(load().lyap_batch_run_noalloc(int(device), recs, len(classes), ctypes.byref(tu),
                                        ctypes.c_void_p(s.cuda_stream)))


class Mlp:
    """The surrogate network object of liblyap.so (lyap_mlp_*)."""

    def __init__(self, device: int = 0, max_batch: int = 4096, hidden: int = 256,
                 n_fourier: int = 32, fourier_sigma: float = 6.0, lr: float = 1e-3,
                 beta1: float = 0.9, beta2: float = 0.999, eps: float = 1e-8, seed: int = 1,
                 ranges=(0.0, 1.0, 0.0, 1.0), y_mean=(0.0, 0.0, 0.0), y_std=(1.0, 1.0, 1.0)):
        c = MlpConfig()
        c.max_batch, c.hidden, c.n_fourier = max_batch, hidden, n_fourier
        c.fourier_sigma, c.lr, c.beta1, c.beta2, c.eps = fourier_sigma, lr, beta1, beta2, eps
        c.seed = seed
        c.lo1, c.hi1, c.lo2, c.hi2 = ranges
        for j in range(3):
            c.y_mean[j], c.y_std[j] = y_mean[j], y_std[j]
        self.cfg = c
        self.device = device
        self._h = ctypes.c_void_p()
        check(load().lyap_mlp_create(int(device), ctypes.byref(c), ctypes.byref(self._h)))
        n = ctypes.c_int64()
        check(load().lyap_mlp_param_count(self._h, ctypes.byref(n)))
        self.n_params = int(n.value)

    def close(self):
        if self._h:
            load().lyap_mlp_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _stream(self, stream):
        import torch
        return stream if stream is not None else torch.cuda.current_stream(torch.device("cuda", self.device))

    def backward(self, p, lam, status, loss=None, stream=None):
        check(load().lyap_mlp_backward(self._h, _ptr(p), _ptr(lam), _ptr(status), int(p.shape[0]),
                                       _ptr(loss), ctypes.c_void_p(self._stream(stream).cuda_stream)))

    def train_step(self, p, lam, status, loss=None, stream=None):
        check(load().lyap_mlp_train_step(self._h, _ptr(p), _ptr(lam), _ptr(status), int(p.shape[0]),
                                         _ptr(loss), ctypes.c_void_p(self._stream(stream).cuda_stream)))

    def forward(self, p, out, stream=None):
        check(load().lyap_mlp_forward(self._h, _ptr(p), int(p.shape[0]), _ptr(out),
                                      ctypes.c_void_p(self._stream(stream).cuda_stream)))

    def get(self, which: int):
        import numpy as np
        n = self.cfg.n_fourier * 2 if which == 2 else self.n_params
        out = np.empty(n, dtype=np.float32)
        check(load().lyap_mlp_get(self._h, int(which), out.ctypes.data_as(ctypes.c_void_p), n))
        return out

    def set_params(self, host):
        import numpy as np
        host = np.ascontiguousarray(host, dtype=np.float32)
        check(load().lyap_mlp_set_params(self._h, host.ctypes.data_as(ctypes.c_void_p), host.size))
