"""The drop-in boundary on the GPU: the reference-shaped entry point
lyap_cuda_lyapunov (cudaLyapunov's arguments, cuda/lyapunov.dp.hpp:10-17) with
AoS device buffers, and the chaos_lyap driver (main.cpp:11-80's flow:
config.json -> TaskLyapunov init/allocate/execute -> saveCalculatedData ->
clear) on the reference's shimizu entry with its shipped default mode (FD,
ChaosMigration/config.json:8-9,23-36), grid counts and times reduced.  Both
are checked bit for bit against the CPU oracle."""
import json
import os
import subprocess
import zlib

import numpy as np
import pytest
import torch

from chaos_oneapi_amd import lyap as L
from oracle import oracle as O

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CLI = os.path.join(ROOT, "chaos_oneapi_amd", "chaos_lyap")


def ref_desc(model, dim, n_params, method, task_num, h, eps, t_skip, t_clp, t_calc, add_skip, exps):
    """The descriptor lyap_cuda_lyapunov builds from cudaLyapunov's arguments."""
    spec = dict(model=model, dim=dim, n_exps=exps, n_params=n_params, method=method,
                integrator=L.INT_DVERK, adaptive=0, precision=L.PREC_FP64, h=h, eps=eps,
                t_transient=t_skip, t_slave=t_clp, t_average=t_calc,
                t_renorm=add_skip * (1.0 if model in (L.MODEL_HENON, L.MODEL_LOGISTIC) else h),
                n1=1, n2=task_num, sweep_idx1=0, sweep_idx2=1 if n_params > 1 else 0,
                init_kind=L.INIT_NONE, init_param_offset=-1)
    return L.finalize(L.make_desc(spec))


CASES = [
    # shimizu (alpha, lambda) over the reference config's ranges, every mode
    ("shimizu-fd", L.MODEL_SHIMIZU, 3, 2, L.TYPE_FLOW, False, False),
    ("shimizu-var", L.MODEL_SHIMIZU, 3, 2, L.TYPE_FLOW, False, True),
    ("shimizu-qr", L.MODEL_SHIMIZU, 3, 2, L.TYPE_FLOW, True, False),
    # Henon (a, b): the reference's map VAR and FD branches
    ("henon-var", L.MODEL_HENON, 2, 2, L.TYPE_MAP, False, True),
    ("henon-fd", L.MODEL_HENON, 2, 2, L.TYPE_MAP, False, False),
]


@pytest.mark.parametrize("name,model,dim,npar,mtype,is_qr,is_var", CASES)
def test_cuda_lyapunov_matches_oracle(gpu_ctx, name, model, dim, npar, mtype, is_qr, is_var):
    print(f"case cuda_lyapunov {name}", flush=True)
    rng = np.random.default_rng(zlib.crc32(name.encode()))
    n = 96
    if model == L.MODEL_SHIMIZU:
        params = np.stack([rng.uniform(0.2, 2.0, n), rng.uniform(0.0, 1.2, n)], 1)
        inits = np.tile([1e-4, 0.0, 0.0], (n, 1)) + rng.uniform(-1e-5, 1e-5, (n, 3))
        h, t_skip, t_clp, t_calc, add_skip = 0.01, 20.0, 10.0, 50.0, 1.0
    else:
        params = np.stack([rng.uniform(1.0, 1.42, n), rng.uniform(0.05, 0.35, n)], 1)
        inits = np.tile([0.1, 0.1], (n, 1)) + rng.uniform(-0.05, 0.05, (n, 2))
        h, t_skip, t_clp, t_calc, add_skip = 0.01, 500.0, 100.0, 5000.0, 1.0
    eps = 1e-7
    dev = torch.device("cuda", 0)
    d_params = torch.from_numpy(np.ascontiguousarray(params.reshape(-1))).to(dev)
    d_inits = torch.from_numpy(np.ascontiguousarray(inits.reshape(-1))).to(dev)
    d_out = torch.full((n * dim,), -7.0, dtype=torch.float64, device=dev)
    gpu_ctx.cuda_lyapunov(n, is_qr, is_var, d_inits, dim, model, d_params, npar, eps, h, d_out, dim,
                          t_skip, t_clp, t_calc, add_skip, model_type=mtype)
    torch.cuda.synchronize()
    lam = d_out.cpu().numpy().reshape(n, dim)
    method = L.METHOD_QR_REF if is_qr else (L.METHOD_VAR_REF if is_var else L.METHOD_FD_REF)
    desc = ref_desc(model, dim, npar, method, n, h, eps, t_skip, t_clp, t_calc, add_skip, dim)
    lam_o, _, _, _ = O.sweep(O.make_desc(desc.to_dict()), first=0, count=n, params_in=params.reshape(-1),
                             inits_in=inits.reshape(-1), n_threads=8)
    np.testing.assert_array_equal(lam, lam_o)
    assert np.isfinite(lam).sum() > n  # a real spectrum, not an untouched buffer


def test_cuda_lyapunov_refusals(gpu_ctx):
    """QR on a map is refused as the reference refuses it
    (cuda/lyapunov.dp.cpp:302-305); a model type that is not the model's, and
    more than INT32_MAX tasks, are argument errors."""
    dev = torch.device("cuda", 0)
    b = torch.zeros(64, dtype=torch.float64, device=dev)
    with pytest.raises(L.LyapError) as e:
        gpu_ctx.cuda_lyapunov(4, True, False, b, 2, L.MODEL_HENON, b, 2, 1e-7, 1.0, b, 2, 10, 10, 10, 1.0,
                              model_type=L.TYPE_MAP)
    assert e.value.code == L.ERR_UNSUPPORTED
    with pytest.raises(L.LyapError) as e:
        gpu_ctx.cuda_lyapunov(4, False, True, b, 2, L.MODEL_HENON, b, 2, 1e-7, 1.0, b, 2, 10, 10, 10, 1.0,
                              model_type=L.TYPE_FLOW)
    assert e.value.code == L.ERR_INVALID
    import ctypes
    rc = L.load().lyap_cuda_lyapunov(gpu_ctx._h, 32, 32, (1 << 31) + 5, 0, 1, L.TYPE_FLOW,
                                     ctypes.c_void_p(b.data_ptr()), 3, L.MODEL_SHIMIZU,
                                     ctypes.c_void_p(b.data_ptr()), 2, 1e-7, 0.01,
                                     ctypes.c_void_p(b.data_ptr()), 3, 1.0, 1.0, 1.0, 1.0, None)
    assert rc == L.ERR_INVALID and b"task_num" in L.load().lyap_last_error()


SHIMIZU_CONFIG = {
    "hardware": {"num_GPUs": 1},
    "tool": {"lyapunov": {"samples_num": 1, "using_qr": False, "using_var": False}},
    "systems": [{
        "model_name": "shimizu", "lyapunov_exponent_num": 3, "epsilon": 1e-7,
        "integration_step": 1e-2, "skip_time_attractor": 20.0, "skip_time_slave_trajectory": 10.0,
        "skip_time_normalization": 1e-2, "total_time": 50.0, "initial_state": [1e-4, 0, 0],
        "initial_model_params": [0, 0], "first_sweep_param": [0, 0.2, 2, 12],
        "second_sweep_param": [1, 0, 1.2, 10]}],
}


def test_chaos_lyap_driver_on_shimizu_fd(tmp_path):
    """The whole lifecycle through the driver: the .npy rows are the oracle's
    values bit for bit (p1 outer), and the .json _Data rows are those values
    printed as std::to_string does (fixed, six decimals)."""
    print("case chaos_lyap shimizu FD 12x10", flush=True)
    (tmp_path / "config.json").write_text(json.dumps(SHIMIZU_CONFIG))
    r = subprocess.run([CLI, "shimizu"], cwd=tmp_path, capture_output=True, text=True, timeout=120)
    print(r.stdout[-2000:], r.stderr[-2000:], flush=True)
    assert r.returncode == 0
    rows = np.load(tmp_path / "Data" / "Lyapunov shimizu.npy")
    assert rows.shape == (120, 5)
    s = SHIMIZU_CONFIG["systems"][0]
    f, g = s["first_sweep_param"], s["second_sweep_param"]
    spec = dict(model=L.MODEL_SHIMIZU, dim=3, n_exps=3, n_params=2, method=L.METHOD_FD_REF,
                integrator=L.INT_DVERK, adaptive=0, precision=L.PREC_FP64, h=s["integration_step"],
                eps=s["epsilon"], t_transient=s["skip_time_attractor"],
                t_slave=s["skip_time_slave_trajectory"], t_average=s["total_time"],
                t_renorm=s["skip_time_normalization"], base_params=[0.0, 0.0],
                init_state=s["initial_state"], init_kind=L.INIT_NONE, init_param_offset=-1,
                sweep_idx1=0, start1=f[1], delta1=abs(f[2] - f[1]) / (f[3] - 1), n1=f[3],
                sweep_idx2=1, start2=g[1], delta2=abs(g[2] - g[1]) / (g[3] - 1), n2=g[3])
    desc = L.finalize(L.make_desc(spec))
    lam_o, _, _, _ = O.sweep(O.make_desc(desc.to_dict()), n_threads=8)
    i1, i2 = np.divmod(np.arange(120), g[3])
    np.testing.assert_array_equal(rows[:, 0], f[1] + i1 * desc.delta1)
    np.testing.assert_array_equal(rows[:, 1], g[1] + i2 * desc.delta2)
    np.testing.assert_array_equal(rows[:, 2:], lam_o)
    out = json.loads((tmp_path / "Data" / "Lyapunov shimizu.json").read_text())
    assert out["ID_Tool"] == "Lyapunov" and out["Model"]["Name"] == "shimizu"
    lines = out["_Data"].split("\n")
    assert lines[-1] == "" and len(lines) == 121
    for row, line in zip(rows, lines):
        assert line == " ".join(f"{v:f}" for v in row)


ROSSLER_ORDERED_CONFIG = {
    "hardware": {"num_GPUs": 1},
    "tool": {"lyapunov": {"samples_num": 1, "using_qr": False, "using_var": True}},
    "systems": [{
        "model_name": "rossler", "lyapunov_exponent_num": 3, "epsilon": 1e-7,
        "integration_step": 0.01, "skip_time_attractor": 5, "skip_time_slave_trajectory": 0,
        "skip_time_normalization": 0.5, "total_time": 20, "initial_state": [0, 0, 0],
        "initial_model_params": [0, 0, 0.2], "first_sweep_param": [0, 0.05, 0.4, 12],
        "second_sweep_param": [1, 2, 20, 10],
        "method": "var_trace", "integrator": "dverk", "adaptive": True, "rtol": 1e-9, "atol": 1e-9,
        "precision": "mixed", "seed": 11, "init_perturbation": {"kind": "uniform", "scale": 1.0},
        "order_probe": 50}],
}


def test_chaos_lyap_driver_heavy_first_rossler(tmp_path):
    """The driver with this build's optional keys (the sweep's method, adaptive
    DVERK, mixed form, trace closure) and the heavy-first order: the rows are the
    oracle's values bit for bit, in the reference's p1-outer order."""
    print("case chaos_lyap rossler var_trace order_probe 12x10", flush=True)
    (tmp_path / "config.json").write_text(json.dumps(ROSSLER_ORDERED_CONFIG))
    r = subprocess.run([CLI, "rossler"], cwd=tmp_path, capture_output=True, text=True, timeout=120)
    print(r.stdout[-2000:], r.stderr[-2000:], flush=True)
    assert r.returncode == 0
    rows = np.load(tmp_path / "Data" / "Lyapunov rossler.npy")
    assert rows.shape == (120, 5)
    s = ROSSLER_ORDERED_CONFIG["systems"][0]
    f, g = s["first_sweep_param"], s["second_sweep_param"]
    spec = dict(model=L.MODEL_ROSSLER, dim=3, n_exps=3, n_params=3, method=L.METHOD_VAR_TRACE,
                integrator=L.INT_DVERK, adaptive=1, precision=L.PREC_MIXED, h=s["integration_step"],
                rtol=s["rtol"], atol=s["atol"], eps=s["epsilon"], t_transient=s["skip_time_attractor"],
                t_slave=0.0, t_average=s["total_time"], t_renorm=s["skip_time_normalization"],
                base_params=s["initial_model_params"], init_state=s["initial_state"],
                init_kind=L.INIT_UNIFORM, init_scale=1.0, init_param_offset=-1, seed=s["seed"],
                sweep_idx1=0, start1=f[1], delta1=abs(f[2] - f[1]) / (f[3] - 1), n1=f[3],
                sweep_idx2=1, start2=g[1], delta2=abs(g[2] - g[1]) / (g[3] - 1), n2=g[3])
    desc = L.finalize(L.make_desc(spec))
    lam_o, st_o, _, _ = O.sweep(O.make_desc(desc.to_dict()), n_threads=8)
    ok = st_o == 0
    assert ok.sum() > 60
    np.testing.assert_array_equal(rows[ok, 2:], lam_o[ok])
    assert np.all(np.isnan(rows[~ok, 2:]))
