"""Heavy-first runs (option "order_probe"): a probe run counts every point's
steps, the full run takes the points most steps first and scatters its outputs
back to point order.  Every point's results are the same bits as the run in
the kernels' own order (and so as the oracle's), over the whole grid and over
a sub-range, for the one-lane, four-lane and Lorenz-96 kernels, whatever the
probe's length and the grid neighbourhood its keys take the maximum over."""
import ctypes

import numpy as np
import pytest
import torch

from chaos_oneapi_amd import lyap as L
from chaos_oneapi_amd import workloads as W
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def host(t):
    return tuple(x.cpu().numpy() for x in t)


def run_opts(ctx, desc, opts, **kw):
    for k, v in opts.items():
        ctx.set_option(k, v)
    try:
        out = host(ctx.run(desc, **kw))
        torch.cuda.synchronize()
    finally:
        for k in opts:
            ctx.set_option(k, 0)
    return out


def assert_same(a, b):
    for x, y in zip(a, b):
        np.testing.assert_array_equal(x, y)


CASES = {
    # (spec, kernel options): grids of one or two blocks, so that points start in later rounds
    "rossler-trace-lane": (dict(W.rossler_dverk(24, precision=L.PREC_MIXED, t_transient=5.0, t_average=20.0),
                                method=L.METHOD_VAR_TRACE, rtol=1e-9, atol=1e-9), {"lane_max_blocks": 1}),
    "lorenz-fp64-quad": (dict(W.lorenz_fixed(16, t_transient=2.0, t_average=8.0), adaptive=1, rtol=1e-9,
                              atol=1e-9), {"lane_max_blocks": 1, "lane_quad": 1}),
    "l96-8-mixed": (W.lorenz96_dopri(n=12, dim=8, precision=L.PREC_MIXED, t_transient=2.0, t_average=6.0),
                    {"l96_max_blocks": 2}),
}


@pytest.mark.parametrize("case", list(CASES))
@pytest.mark.parametrize("permille,radius", [(10, 2), (200, 0), (30, 8)])
def test_heavy_first_matches_spread_order(gpu_ctx, case, permille, radius):
    spec, opts = CASES[case]
    desc = L.finalize(L.make_desc(spec))
    n = desc.n1 * desc.n2
    print(f"case order {case} permille {permille} radius {radius} n={n}", flush=True)
    base = run_opts(gpu_ctx, desc, opts)
    gpu_ctx.set_option("order_radius", radius)
    try:
        got = run_opts(gpu_ctx, desc, dict(opts, order_probe=permille))
        # a sub-range of the grid (neighbourhoods cut at its ends)
        first, count = n // 3, n // 2
        sub = run_opts(gpu_ctx, desc, dict(opts, order_probe=permille), first=first, count=count)
    finally:
        gpu_ctx.set_option("order_radius", 2)
    assert_same(got, base)
    assert_same(sub, tuple(x[first:first + count] for x in base))


def test_refine_heavy_first_is_a_sorted_permutation(gpu_ctx):
    """The refinement's own ordering: a permutation of the given points whose
    neighbourhood-maximum probe counts do not increase along the list."""
    import torch.nn.functional as Fn

    from chaos_oneapi_amd import refine as R

    spec = dict(W.rossler_dverk(40, precision=L.PREC_MIXED, t_transient=5.0, t_average=20.0),
                method=L.METHOD_VAR_TRACE)
    fd = L.finalize(L.make_desc(spec))
    n = fd.n1 * fd.n2
    idx = torch.arange(0, n, 3, device="cuda", dtype=torch.int64)
    got = R.heavy_first(gpu_ctx, fd, idx, 100, 1)
    torch.cuda.synchronize()
    assert torch.equal(got.sort().values, idx)
    pd = L.SweepDesc.from_buffer_copy(fd)
    pd.t_transient *= 0.1
    pd.t_average = max(fd.t_average * 0.1, fd.t_renorm)
    _, _, na, nr = gpu_ctx.run(L.finalize(pd))
    keys = (na + nr).to(torch.float32).reshape(1, 1, fd.n1, fd.n2)
    keys = Fn.max_pool2d(keys, 3, stride=1, padding=1).reshape(-1)
    k = keys[got].cpu().numpy()
    assert np.all(np.diff(k) <= 0) and k[0] > k[-1]


def test_heavy_first_matches_oracle(gpu_ctx):
    spec = dict(W.rossler_dverk(10, precision=L.PREC_MIXED, t_transient=5.0, t_average=20.0),
                method=L.METHOD_VAR_TRACE, rtol=1e-9, atol=1e-9)
    desc = L.finalize(L.make_desc(spec))
    print("case order vs oracle rossler 10x10", flush=True)
    lam, st, na, nr = run_opts(gpu_ctx, desc, {"order_probe": 50, "lane_max_blocks": 1})
    o_lam, o_st, o_na, o_nr = O.sweep(O.make_desc(desc.to_dict()), first=0, count=desc.n1 * desc.n2)
    np.testing.assert_array_equal(st.astype(np.uint32), o_st)
    np.testing.assert_array_equal(na.astype(np.uint32), o_na)
    np.testing.assert_array_equal(nr.astype(np.uint32), o_nr)
    np.testing.assert_array_equal(lam, o_lam)


def test_heavy_first_with_lambda_only(gpu_ctx):
    """Null status and step-count outputs through the C-ABI: the scatter skips them."""
    spec, opts = CASES["rossler-trace-lane"]
    desc = L.finalize(L.make_desc(spec))
    n = desc.n1 * desc.n2
    base = run_opts(gpu_ctx, desc, opts)
    lam = torch.empty((n, desc.n_exps), dtype=torch.float64, device="cuda")
    lib = L.load()
    for k, v in dict(opts, order_probe=30).items():
        gpu_ctx.set_option(k, v)
    try:
        L.check(lib.lyap_sweep_run(gpu_ctx._h, ctypes.byref(desc), 0, n, None, None, None,
                                   ctypes.c_void_p(lam.data_ptr()), None, None, None,
                                   ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)))
        torch.cuda.synchronize()
    finally:
        for k in dict(opts, order_probe=0):
            gpu_ctx.set_option(k, 0)
    np.testing.assert_array_equal(lam.cpu().numpy(), base[0])


def test_order_probe_range(gpu_ctx):
    lib = L.load()
    for bad in (-1, 501):
        assert lib.lyap_ctx_set_option(gpu_ctx._h, b"order_probe", bad) == L.ERR_INVALID
    assert lib.lyap_ctx_set_option(gpu_ctx._h, b"order_probe", 0) == L.SUCCESS
    for bad in (-1, 9):
        assert lib.lyap_ctx_set_option(gpu_ctx._h, b"order_radius", bad) == L.ERR_INVALID
    assert lib.lyap_ctx_set_option(gpu_ctx._h, b"order_radius", 2) == L.SUCCESS


def test_heavy_first_with_parameter_rows(gpu_ctx):
    """Caller-supplied parameter rows (no grid): the points are ordered by their
    own probe counts, and every output is the unordered run's bit for bit."""
    spec, opts = CASES["rossler-trace-lane"]
    desc = L.finalize(L.make_desc(spec))
    n = 400
    g = torch.Generator().manual_seed(3)
    a = 0.05 + 0.35 * torch.rand(n, generator=g, dtype=torch.float64)
    c = 2.0 + 18.0 * torch.rand(n, generator=g, dtype=torch.float64)
    params = torch.stack([a, c, torch.full((n,), 0.2, dtype=torch.float64)], dim=1).contiguous().cuda()
    print("case order with 400 parameter rows", flush=True)
    base = run_opts(gpu_ctx, desc, opts, count=n, params=params)
    got = run_opts(gpu_ctx, desc, dict(opts, order_probe=30), count=n, params=params)
    assert_same(got, base)


def test_lane_grid_less_free_blocks(gpu_ctx):
    """lane_max_blocks < 0 leaves that many blocks of the occupancy-sized grid
    free (at least one block runs): the same bits as the full grid."""
    lib = L.load()
    assert lib.lyap_ctx_set_option(gpu_ctx._h, b"lane_max_blocks", -65537) == L.ERR_INVALID
    spec, _ = CASES["rossler-trace-lane"]
    desc = L.finalize(L.make_desc(spec))
    base = run_opts(gpu_ctx, desc, {})
    for free in (-32, -65536):
        assert_same(run_opts(gpu_ctx, desc, {"lane_max_blocks": free}), base)
