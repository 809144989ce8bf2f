"""Host-side checks of liblyap.so that need no GPU: every symbol the public
header declares is exported, the model table, and lyap_desc_finalize's
derivation of the reference's scalars (tasklyapunov.cpp:17-31,45-46,65) and
its validation (status codes instead of the reference's exit()/throw)."""
import ctypes
import os
import re

import numpy as np
import pytest

from chaos_oneapi_amd import lyap as L
from chaos_oneapi_amd import workloads as W
from oracle import oracle as O

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def header_symbols():
    text = open(os.path.join(ROOT, "include", "chaos_lyap.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(lyap_[a-z_0-9]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    lib = L.load()
    syms = header_symbols()
    assert len(syms) >= 10
    for s in syms:
        assert hasattr(lib, s), f"{s} declared in include/chaos_lyap.h but not exported"
    assert set(syms) == set(L.EXPORTS)


def test_model_table_and_lookup():
    names = [m.name.decode() for m in L.models()]
    assert names == ["shimizu", "lorenz", "rossler", "lorenz96_8", "lorenz96_16", "lorenz96_32",
                     "henon", "logistic"]
    info = L.model_lookup("shimizu")
    assert (info.model, info.dim, info.n_params, info.type) == (L.MODEL_SHIMIZU, 3, 2, L.TYPE_FLOW)
    assert L.model_lookup("lorenz96_16").dim == 16
    info = L.model_lookup("henon")
    assert (info.model, info.dim, info.n_params, info.type) == (L.MODEL_HENON, 2, 2, L.TYPE_MAP)
    assert L.model_lookup("logistic").type == L.TYPE_MAP
    with pytest.raises(L.LyapError) as e:
        L.model_lookup("lorenz4d")
    assert e.value.code == L.ERR_NOT_FOUND


def test_loop_count_matches_oracle():
    for t0, t1, h in [(0, 1000, 0.01), (0, 100, 0.01), (1000, 11000, 0.01), (0, 3.3, 1e-3)]:
        assert L.loop_count(t0, t1, h) == O.lib().orc_loop_count(float(t0), float(t1), h)


def ref_spec(method, **kw):
    spec = W.lorenz_fixed(n=4)
    spec.update(method=method, integrator=L.INT_DVERK, eps=1e-7, t_transient=1000.0,
                t_slave=1000.0, t_average=10000.0, t_renorm=0.01)
    spec.update(kw)
    return spec


@pytest.mark.parametrize("method", [L.METHOD_VAR_REF, L.METHOD_FD_REF, L.METHOD_QR_REF])
def test_finalize_reference_counts(method):
    d = L.finalize(L.make_desc(ref_spec(method)))
    assert d.ref_steps_transient == O.lib().orc_loop_count(0.0, 1000.0, 0.01)
    assert d.ref_steps_slave == O.lib().orc_loop_count(0.0, 1000.0, 0.01)
    if method == L.METHOD_QR_REF:  # loop runs from t_attr (cuda/lyapunov.dp.cpp:244)
        assert d.ref_steps_average == O.lib().orc_loop_count(1000.0, 11000.0, 0.01)
    else:
        assert d.ref_steps_average == O.lib().orc_loop_count(0.0, 10000.0, 0.01)
    assert d.add_skip_steps == 0.01 / 0.01


def test_finalize_var_schedule():
    d = L.finalize(L.make_desc(W.rossler_dverk(8)))
    assert (d.n_renorm_transient, d.n_renorm_average) == (400, 4000)
    assert d.h_max == 0.5 and d.h_min > 0 and d.steps_per_renorm == 0
    d = L.finalize(L.make_desc(W.lorenz_fixed(8)))
    assert (d.n_renorm_transient, d.n_renorm_average, d.steps_per_renorm) == (100, 1000, 100)


@pytest.mark.parametrize("bad,code", [
    (dict(model=7), L.ERR_INVALID),
    (dict(n_exps=4), L.ERR_INVALID),
    (dict(n_params=2), L.ERR_INVALID),
    (dict(h=0.0), L.ERR_INVALID),
    (dict(t_average=0.0), L.ERR_INVALID),
    (dict(sweep_idx1=3), L.ERR_INVALID),
    (dict(method=9), L.ERR_INVALID),
    (dict(adaptive=1, rtol=0.0), L.ERR_INVALID),
    (dict(method=L.METHOD_VAR_REF, integrator=L.INT_DOPRI5), L.ERR_INVALID),
    (dict(method=L.METHOD_QR_REF, n_exps=2), L.ERR_INVALID),
    (dict(method=L.METHOD_FD_REF, eps=0.0), L.ERR_INVALID),
])
def test_finalize_rejects(bad, code):
    spec = W.lorenz_fixed(n=4)
    spec.update(eps=1e-7)
    spec.update(bad)
    with pytest.raises(L.LyapError) as e:
        L.finalize(L.make_desc(spec))
    assert e.value.code == code
    assert L.load().lyap_last_error().decode()


def test_reference_modes_need_3d_models():
    spec = W.lorenz96_dopri(n=4, dim=8)
    spec.update(method=L.METHOD_VAR_REF, integrator=L.INT_DVERK, adaptive=0, eps=1e-7)
    with pytest.raises(L.LyapError) as e:
        L.finalize(L.make_desc(spec))
    assert e.value.code == L.ERR_UNSUPPORTED


def test_descriptor_layout_matches_oracle():
    assert ctypes.sizeof(L.SweepDesc) == ctypes.sizeof(O.OrcDesc)
    for (a, ta), (b, tb) in zip(L.SweepDesc._fields_, O.OrcDesc._fields_):
        assert a == b and getattr(L.SweepDesc, a).offset == getattr(O.OrcDesc, b).offset


def test_set_option_validation():
    lib = L.load()
    assert lib.lyap_ctx_set_option(None, b"l96_waves_16", 2) == L.ERR_INVALID


def test_null_arguments_return_codes():
    lib = L.load()
    assert lib.lyap_ctx_create(0, None) == L.ERR_INVALID
    assert lib.lyap_sweep_run(None, None, 0, 1, None, None, None, None, None, None, None,
                              None) == L.ERR_INVALID
    assert lib.lyap_desc_finalize(None) == L.ERR_INVALID
    assert lib.lyap_model_lookup(None, None) == L.ERR_INVALID


def test_batch_entry_points_reject_null():
    lib = L.load()
    out = ctypes.c_void_p()
    assert lib.lyap_batch_create(None, None, 1, None, ctypes.byref(out)) == L.ERR_INVALID
    assert lib.lyap_batch_launch(None, None) == L.ERR_INVALID
    assert lib.lyap_batch_last_ms(None, None) == L.ERR_INVALID
    assert lib.lyap_batch_info(None, None, None) == L.ERR_INVALID
    assert lib.lyap_batch_destroy(None) == L.SUCCESS
    assert b"NULL" in lib.lyap_last_error()


def test_batch_layout_points_per_wave():
    # layout 0: one point per N-lane group (N = 32: one per wave); layout 1: two
    # vector groups per point for N = 8, 16; the 3-D classes run 64 per wave
    want = {(0, 8): 8, (0, 16): 4, (0, 32): 1, (1, 8): 4, (1, 16): 2, (1, 32): 1, (0, 3): 64, (1, 3): 64}
    for (layout, dim), ppw in want.items():
        assert L.batch_layout_ppw(layout, dim) == ppw
    with pytest.raises(L.LyapError):
        L.batch_layout_ppw(2, 16)
    assert L.load().lyap_batch_layout_ppw(0, 16, None) == L.ERR_INVALID


def test_batch_record_layout():
    # lyap_batch_class: the descriptor, then count and six pointers
    assert L.BatchClassRec.count.offset == ctypes.sizeof(L.SweepDesc)
    assert ctypes.sizeof(L.BatchClassRec) == ctypes.sizeof(L.SweepDesc) + 7 * 8
    assert ctypes.sizeof(L.BatchTuning) == 20


def test_finalize_maps():
    from chaos_oneapi_amd import workloads as W

    d = L.finalize(L.make_desc(W.henon(n=4)))
    assert d.h == 1.0 and d.steps_per_renorm == 10
    assert d.n_renorm_transient == 100 and d.n_renorm_average == 10000
    spec = W.henon(n=4)
    spec.update(method=L.METHOD_VAR_REF, eps=1e-8, t_transient=100.5, t_average=1000.0)
    d = L.finalize(L.make_desc(spec))
    assert (d.ref_steps_transient, d.ref_steps_average) == (101, 1000)
    bad = W.henon(n=4)
    bad.update(adaptive=1, rtol=1e-9, atol=1e-9)
    with pytest.raises(L.LyapError, match="no adaptive step"):
        L.finalize(L.make_desc(bad))
    qr = W.logistic(n=4)
    qr.update(method=L.METHOD_QR_REF)
    with pytest.raises(L.LyapError, match="no map branch") as e:
        L.finalize(L.make_desc(qr))
    assert e.value.code == L.ERR_UNSUPPORTED


def test_noalloc_and_reference_entry_reject_before_touching_a_device():
    """The allocation-free batched launch refuses a NULL class list and more
    classes than its fixed array holds, and the cudaLyapunov-shaped entry a
    NULL context, before any device call (these run without a GPU)."""
    lib = L.load()
    assert lib.lyap_batch_run_noalloc(0, None, 1, None, None) == L.ERR_INVALID
    recs = (L.BatchClassRec * 17)()
    assert lib.lyap_batch_run_noalloc(0, recs, 17, None, None) == L.ERR_INVALID
    assert b"scratch" in lib.lyap_last_error()
    assert lib.lyap_cuda_lyapunov(None, 32, 32, 4, 0, 1, L.TYPE_FLOW, None, 3, L.MODEL_SHIMIZU, None, 2,
                                  1e-7, 0.01, None, 3, 1.0, 1.0, 1.0, 1.0, None) == L.ERR_INVALID


def test_ref_option_names_are_checked():
    lib = L.load()
    assert lib.lyap_ctx_set_option(None, b"no_such_option", 1) == L.ERR_INVALID


def test_watch_entry_points_reject_bad_arguments():
    """lyap_ctx_set_watch and lyap_ctx_watch_wait refuse a NULL context, bad
    milestone ranges and bad milestone numbers before any device call; the
    Python wrapper checks the ranges too (no GPU needed)."""
    lib = L.load()
    assert lib.lyap_ctx_set_watch(None, 1, 2, None) == L.ERR_INVALID
    r = ctypes.c_int(7)
    assert lib.lyap_ctx_watch_wait(None, 0, 1.0, ctypes.byref(r)) == L.ERR_INVALID
    assert lib.lyap_ctx_watch_wait(None, 2, 1.0, ctypes.byref(r)) == L.ERR_INVALID
    assert lib.lyap_ctx_watch_wait(None, 0, 1.0, None) == L.ERR_INVALID
    ctx = L.Context.__new__(L.Context)  # no device: only the argument checks run
    ctx._h = ctypes.c_void_p()
    with pytest.raises(ValueError, match="not ranges"):
        ctx.set_watch(5, 3)
    with pytest.raises(ValueError, match="not ranges"):
        ctx.set_watch(-1, 3)
    assert L.PT_SKIPPED == 5
