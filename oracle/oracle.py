"""ctypes front end of the CPU FP64 oracle — TEST INFRASTRUCTURE ONLY.

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg import
this module, and only to check or time the reference algorithm on the CPU.
The product package (chaos_oneapi_amd) never imports it.

The oracle library is built from oracle/lyap_oracle.c by oracle/Makefile.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liblyap_oracle.so")

MAX_DIM = 32
MAX_PARAMS = 8

MODEL_SHIMIZU, MODEL_LORENZ, MODEL_ROSSLER, MODEL_L96 = 0, 1, 2, 3
MODEL_HENON, MODEL_LOGISTIC = 4, 5
METHOD_VAR, METHOD_VAR_REF, METHOD_QR_REF, METHOD_FD_REF, METHOD_VAR_TRACE = 0, 1, 2, 3, 4
INT_DVERK, INT_DOPRI5 = 0, 1
PREC_FP64, PREC_MIXED = 0, 1
INIT_NONE, INIT_UNIFORM, INIT_NORMAL = 0, 1, 2
OK, NONFINITE, STEP_UNDERFLOW, MAXSTEPS, ESCAPED = 0, 1, 2, 3, 4


class OrcDesc(ctypes.Structure):
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


def build(force: bool = False) -> str:
    if force or not os.path.exists(LIB_PATH):
        subprocess.run(["make", "-s", "-C", HERE], check=True)
    return LIB_PATH


_lib = None


def lib():
    global _lib
    if _lib is None:
        build()
        L = ctypes.CDLL(LIB_PATH)
        d, i, i64, u64 = ctypes.c_double, ctypes.c_int, ctypes.c_int64, ctypes.c_uint64
        P = ctypes.c_void_p
        L.orc_log.restype = d; L.orc_log.argtypes = [d]
        L.orc_exp.restype = d; L.orc_exp.argtypes = [d]
        L.orc_pow.restype = d; L.orc_pow.argtypes = [d, d]
        L.orc_ctrl_log.restype = d; L.orc_ctrl_log.argtypes = [d]
        L.orc_ctrl_exp.restype = d; L.orc_ctrl_exp.argtypes = [d]
        L.orc_ctrl_update.restype = i
        L.orc_ctrl_update.argtypes = [i, d, ctypes.POINTER(d), ctypes.POINTER(ctypes.c_int), d, d,
                                      ctypes.POINTER(d)]
        L.orc_jac_norm.restype = d
        L.orc_jac_norm.argtypes = [i, i, P, P]
        L.orc_trace.restype = d
        L.orc_trace.argtypes = [i, i, P, P]
        L.orc_trace_step.restype = d
        L.orc_trace_step.argtypes = [i, i, i, P, P, d]
        L.orc_uniform.restype = d; L.orc_uniform.argtypes = [u64, i64, i64]
        L.orc_loop_count.restype = i64; L.orc_loop_count.argtypes = [d, d, d]
        L.orc_rhs.argtypes = [i, i, P, P, P]
        L.orc_jvp.argtypes = [i, i, P, P, P, P]
        L.orc_jvp_f32.argtypes = [i, i, P, P, P, P]
        L.orc_rk_step.restype = i
        L.orc_rk_step.argtypes = [i, i, i, P, P, d, P, P, P]
        L.orc_tangent_step.argtypes = [i, i, i, P, P, d, P]
        L.orc_tangent_step_f32.argtypes = [i, i, i, P, P, d, P, i]
        L.orc_frozen_tangent_step.argtypes = [i, i, P, P, d, P]
        L.orc_mgs.argtypes = [i, i, i, P, P]
        L.orc_split_exponent.restype = d
        L.orc_split_exponent.argtypes = [d, ctypes.POINTER(ctypes.c_int32)]
        L.orc_log_scaled.restype = d
        L.orc_log_scaled.argtypes = [d, ctypes.c_int32]
        L.orc_inv_sqrt.restype = d
        L.orc_inv_sqrt.argtypes = [d]
        L.orc_cgs_ref.argtypes = [i, i, P]
        L.orc_normalize_ref.argtypes = [i, i, P, d]
        L.orc_qr_ref.argtypes = [i, P, P]
        L.orc_err_norm.restype = d
        L.orc_err_norm.argtypes = [i, i, P, P, P, d, d]
        L.orc_err2_seq.restype = d
        L.orc_err2_seq.argtypes = [i, P, P, P, d, d]
        L.orc_point_params.argtypes = [P, i64, P, P]
        L.orc_point_init.argtypes = [P, i64, P, P, P]
        L.orc_point.argtypes = [P, i64, P, P, P, P, P, P]
        L.orc_sweep.argtypes = [P, i64, i64, P, P, P, P, P, P, P, i]
        _lib = L
    return _lib


def _ptr(a):
    return None if a is None else a.ctypes.data_as(ctypes.c_void_p)


def make_desc(spec: dict) -> OrcDesc:
    """Fill an OrcDesc from a plain dictionary (same keys as the C struct)."""
    d = OrcDesc()
    for name, ctype in OrcDesc._fields_:
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


def f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def rhs(model, x, p):
    x, p = f64(x), f64(np.pad(p, (0, MAX_PARAMS - len(p))))
    out = np.empty_like(x)
    lib().orc_rhs(model, len(x), _ptr(x), _ptr(p), _ptr(out))
    return out


def jvp(model, x, p, v):
    x, v, p = f64(x), f64(v), f64(np.pad(p, (0, MAX_PARAMS - len(p))))
    out = np.empty_like(x)
    lib().orc_jvp(model, len(x), _ptr(x), _ptr(p), _ptr(v), _ptr(out))
    return out


def ctrl_update(integ, err, log_fac_old, rejected, h_try, h_max):
    """One step-controller decision: (accepted, h_next, log_fac_old, rejected)."""
    lf = ctypes.c_double(float(log_fac_old))
    rj = ctypes.c_int(int(rejected))
    hn = ctypes.c_double()
    acc = lib().orc_ctrl_update(int(integ), float(err), ctypes.byref(lf), ctypes.byref(rj),
                                float(h_try), float(h_max), ctypes.byref(hn))
    return bool(acc), hn.value, lf.value, bool(rj.value)


def jac_norm(model, x, p):
    """max_i sum_j |J_ij(x)|: the adaptive step cap's norm."""
    x, p = f64(x), f64(np.pad(p, (0, MAX_PARAMS - len(p))))
    return lib().orc_jac_norm(model, len(x), _ptr(x), _ptr(p))


def trace(model, x, p):
    """tr J(x) of a 3-D flow (its divergence)."""
    x, p = f64(x), f64(np.pad(p, (0, MAX_PARAMS - len(p))))
    return lib().orc_trace(model, len(x), _ptr(x), _ptr(p))


def trace_step(model, integ, stage_x, p, h):
    """The step's integral of tr J from its stage states (rows of stage_x)."""
    sx = f64(np.asarray(stage_x).reshape(-1))
    p = f64(np.pad(p, (0, MAX_PARAMS - len(p))))
    n = np.asarray(stage_x).shape[-1]
    return lib().orc_trace_step(model, integ, n, _ptr(sx), _ptr(p), float(h))


def rk_step(model, integ, x, p, h, want_err=True):
    x, p = f64(x), f64(np.pad(p, (0, MAX_PARAMS - len(p))))
    n = len(x)
    xn = np.empty(n)
    ev = np.empty(n) if want_err else None
    sx = np.empty(8 * n)
    lib().orc_rk_step(model, integ, n, _ptr(x), _ptr(p), float(h), _ptr(xn), _ptr(ev), _ptr(sx))
    return xn, ev, sx.reshape(8, n)


def tangent_step(model, integ, stage_x, p, h, v):
    p = f64(np.pad(p, (0, MAX_PARAMS - len(p))))
    v = f64(v).copy()
    sx = f64(stage_x)
    lib().orc_tangent_step(model, integ, len(v), _ptr(sx), _ptr(p), float(h), _ptr(v))
    return v


def tangent_step_f32(model, integ, stage_x, p, h, v, low=False):
    """The mixed form's tangent step; low: adaptive DVERK's fifth-order weights."""
    p = f64(np.pad(p, (0, MAX_PARAMS - len(p))))
    v = f64(v).copy()
    sx = f64(stage_x)
    lib().orc_tangent_step_f32(model, integ, len(v), _ptr(sx), _ptr(p), float(h), _ptr(v), int(low))
    return v


def frozen_tangent_step(model, x, p, h, v):
    p = f64(np.pad(p, (0, MAX_PARAMS - len(p))))
    v = f64(v).copy()
    x = f64(x)
    lib().orc_frozen_tangent_step(model, len(v), _ptr(x), _ptr(p), float(h), _ptr(v))
    return v


def mgs(V, tree=False):
    V = f64(V).copy()
    k, n = V.shape
    ln = np.empty(k)
    lib().orc_mgs(n, k, int(tree), _ptr(V), _ptr(ln))
    return V, ln


LOG_WINDOW = 64  # ORC_LOG_WINDOW


def split_exponent(m, ex=0):
    """(mant, ex): a positive normal m as mant * 2^(ex - ex_in), mant in [1, 2);
    other values come back unchanged with ex unchanged."""
    e = ctypes.c_int32(int(ex))
    mant = lib().orc_split_exponent(float(m), ctypes.byref(e))
    return mant, e.value


def log_scaled(mant, ex):
    """log(mant * 2^ex) as the sweep's method takes it per log window."""
    return lib().orc_log_scaled(float(mant), int(ex))


def cgs_ref(V):
    V = f64(V).copy()
    k, n = V.shape
    lib().orc_cgs_ref(n, k, _ptr(V))
    return V


def normalize_ref(V, eps):
    V = f64(V).copy()
    k, n = V.shape
    lib().orc_normalize_ref(n, k, _ptr(V), float(eps))
    return V


def qr_ref(Q):
    Q = f64(Q).copy()
    n = Q.shape[0]
    rd = np.empty(n)
    lib().orc_qr_ref(n, _ptr(Q), _ptr(rd))
    return Q, rd


def point_params(desc: OrcDesc, idx: int, params_in=None):
    p = np.empty(MAX_PARAMS)
    lib().orc_point_params(ctypes.byref(desc), idx, _ptr(params_in), _ptr(p))
    return p


def point_init(desc: OrcDesc, idx: int, params_in=None, inits_in=None):
    p = point_params(desc, idx, params_in)
    x = np.empty(desc.dim)
    lib().orc_point_init(ctypes.byref(desc), idx, _ptr(p), _ptr(inits_in), _ptr(x))
    return x


def sweep(desc: OrcDesc, first=0, count=None, idx_list=None, params_in=None, inits_in=None,
          n_threads=0):
    """Run the oracle over points; returns (lam[count,k], status, n_acc, n_rej)."""
    if idx_list is not None:
        idx_list = np.ascontiguousarray(idx_list, dtype=np.int64)
        count = len(idx_list)
    if count is None:
        count = desc.n1 * desc.n2 - first
    k = desc.n_exps
    lam = np.empty((count, k))
    st = np.empty(count, dtype=np.uint32)
    na = np.empty(count, dtype=np.uint32)
    nr = np.empty(count, dtype=np.uint32)
    if params_in is not None:
        params_in = f64(params_in)
    if inits_in is not None:
        inits_in = f64(inits_in)
    lib().orc_sweep(ctypes.byref(desc), first, count, _ptr(idx_list), _ptr(params_in),
                    _ptr(inits_in), _ptr(lam), _ptr(st), _ptr(na), _ptr(nr), int(n_threads))
    return lam, st, na, nr
