"""GPU sweep vs the CPU FP64 oracle on the same seeded grids.

Bar: |dlambda_i| <= 1e-4 for the FP64 path (the FP64 path is also expected to
match bit for bit: same operations in the same order, no contraction); the
mixed form is compared bit for bit with the oracle's mixed form and its error
against the FP64 oracle is reported.  Status words and step counters must
match exactly.
"""
import numpy as np
import pytest
import torch

from chaos_oneapi_amd import lyap as L
from chaos_oneapi_amd import workloads as W
from oracle import oracle as O

pytestmark = pytest.mark.gpu

TOL_FP64 = 1e-4


def run_both(ctx, spec, count=None, first=0):
    desc = L.finalize(L.make_desc(spec))
    n = desc.n1 * desc.n2 if count is None else count
    print(f"case model={desc.model} method={desc.method} integ={desc.integrator} "
          f"adapt={desc.adaptive} prec={desc.precision} n={n}", flush=True)
    lam, st, na, nr = ctx.run(desc, first=first, count=n)
    torch.cuda.synchronize()
    g = (lam.cpu().numpy(), st.cpu().numpy().astype(np.uint32), na.cpu().numpy().astype(np.uint32),
         nr.cpu().numpy().astype(np.uint32))
    od = O.make_desc(desc.to_dict())
    o = O.sweep(od, first=first, count=n)
    return g, o


def assert_match(g, o, tol, exact):
    lam_g, st_g, na_g, nr_g = g
    lam_o, st_o, na_o, nr_o = o
    np.testing.assert_array_equal(st_g, st_o)
    np.testing.assert_array_equal(na_g, na_o)
    np.testing.assert_array_equal(nr_g, nr_o)
    ok = st_o == 0
    diff = np.abs(lam_g[ok] - lam_o[ok])
    print(f"max |dlambda| = {diff.max() if diff.size else 0.0:.3e}, points ok {ok.sum()}/{len(ok)}",
          flush=True)
    assert np.all(np.isnan(lam_g[~ok]))
    assert diff.size == 0 or diff.max() <= tol
    if exact:
        np.testing.assert_array_equal(lam_g[ok], lam_o[ok])


@pytest.mark.parametrize("method", [L.METHOD_VAR, L.METHOD_VAR_TRACE])
@pytest.mark.parametrize("integ", [L.INT_DVERK, L.INT_DOPRI5])
@pytest.mark.parametrize("adaptive", [0, 1])
@pytest.mark.parametrize("prec", [L.PREC_FP64, L.PREC_MIXED])
@pytest.mark.parametrize("model", ["lorenz", "rossler", "shimizu"])
def test_var_lane_matches_oracle(gpu_ctx, model, integ, adaptive, prec, method):
    """The sweep's method, and its trace closure (two deviation vectors, the
    last exponent from the integrated Jacobian trace)."""
    if model == "lorenz":
        spec = W.lorenz_fixed(n=6, precision=prec, t_transient=5.0, t_average=20.0)
    elif model == "rossler":
        spec = W.rossler_dverk(n=6, precision=prec, t_transient=10.0, t_average=40.0)
    else:
        spec = W.lorenz_fixed(n=6, precision=prec, t_transient=5.0, t_average=20.0)
        spec.update(model=L.MODEL_SHIMIZU, n_params=2, base_params=[0.0, 0.0],
                    init_state=[0.1, 0.0, 0.0], init_scale=1e-3)
        spec.update(W.grid_fields(0, 0.2, 2.0, 6, 1, 0.0, 1.2, 6))
    spec.update(integrator=integ, adaptive=adaptive, rtol=1e-9, atol=1e-9, method=method)
    g, o = run_both(gpu_ctx, spec)
    assert_match(g, o, TOL_FP64, exact=True)


@pytest.mark.parametrize("method", [L.METHOD_VAR, L.METHOD_VAR_TRACE])
@pytest.mark.parametrize("integ", [L.INT_DVERK, L.INT_DOPRI5])
@pytest.mark.parametrize("adaptive", [0, 1])
@pytest.mark.parametrize("prec", [L.PREC_FP64, L.PREC_MIXED])
@pytest.mark.parametrize("model", ["lorenz", "rossler", "shimizu"])
def test_var_quad_matches_oracle(gpu_ctx, model, integ, adaptive, prec, method):
    """lane_quad = 1: four lanes per point (one deviation vector per lane, the
    base step and control repeated), on a grid of one block so that quads
    refill: the same bits as the oracle (and so as the one-lane kernel)."""
    if model == "rossler":
        spec = W.rossler_dverk(n=12, precision=prec, t_transient=10.0, t_average=40.0)
    else:
        spec = W.lorenz_fixed(n=12, precision=prec, t_transient=5.0, t_average=20.0)
        if model == "shimizu":
            spec.update(model=L.MODEL_SHIMIZU, n_params=2, base_params=[0.0, 0.0],
                        init_state=[0.1, 0.0, 0.0], init_scale=1e-3)
            spec.update(W.grid_fields(0, 0.2, 2.0, 12, 1, 0.0, 1.2, 12))
    spec.update(integrator=integ, adaptive=adaptive, rtol=1e-9, atol=1e-9, method=method)
    gpu_ctx.set_option("lane_quad", 1)
    gpu_ctx.set_option("lane_max_blocks", 1)
    try:
        g, o = run_both(gpu_ctx, spec)
    finally:
        gpu_ctx.set_option("lane_quad", 0)
        gpu_ctx.set_option("lane_max_blocks", 0)
    assert_match(g, o, TOL_FP64, exact=True)


@pytest.mark.parametrize("gather", [2, 16, 64])
@pytest.mark.parametrize("integ,adaptive,prec,method", [(L.INT_DVERK, 1, L.PREC_MIXED, L.METHOD_VAR),
                                                        (L.INT_DOPRI5, 1, L.PREC_FP64, L.METHOD_VAR),
                                                        (L.INT_DOPRI5, 0, L.PREC_MIXED, L.METHOD_VAR),
                                                        (L.INT_DVERK, 1, L.PREC_MIXED,
                                                         L.METHOD_VAR_TRACE)])
def test_var_lane_gathered_renorm_matches_oracle(gpu_ctx, gather, integ, adaptive, prec, method):
    """Lanes at a renormalisation wait for `gather` lanes of the wave, on a grid
    of 2 blocks so that lanes also refill: every point's bits are unchanged."""
    spec = W.rossler_dverk(n=24, precision=prec, t_transient=5.0, t_average=20.0)
    spec.update(integrator=integ, adaptive=adaptive, rtol=1e-9, atol=1e-9, method=method)
    gpu_ctx.set_option("lane_gather", gather)
    gpu_ctx.set_option("lane_max_blocks", 2)
    try:
        g, o = run_both(gpu_ctx, spec)
    finally:
        gpu_ctx.set_option("lane_gather", 0)
        gpu_ctx.set_option("lane_max_blocks", 0)
    assert_match(g, o, TOL_FP64, exact=True)


@pytest.mark.parametrize("model", ["rossler", "l96-8"])
def test_in_order_items_over_an_index_list_match_oracle(gpu_ctx, model):
    """item_order = 1 (the kernels take an index list in list order, as the
    longest-first rerun does) over a shuffled list with repeated points, on a
    grid small enough that lanes and slots refill: every point's bits match
    the oracle at that point."""
    if model == "rossler":
        spec = W.rossler_dverk(n=24, precision=L.PREC_MIXED, t_transient=5.0, t_average=20.0)
        spec.update(rtol=1e-9, atol=1e-9)
        knob = ("lane_max_blocks", 1)
    else:
        spec = W.lorenz96_dopri(n=5, dim=8, precision=L.PREC_MIXED, t_transient=2.0, t_average=6.0)
        gpu_ctx.set_option("l96_ppw_8", 4)
        knob = ("l96_max_blocks", 1)
    desc = L.finalize(L.make_desc(spec))
    n = desc.n1 * desc.n2
    rng = np.random.default_rng(3)
    idx = np.concatenate([rng.permutation(n), rng.integers(0, n, 7)]).astype(np.int64)
    print(f"case model={model} item_order=1 n={len(idx)}", flush=True)
    gpu_ctx.set_option("item_order", 1)
    gpu_ctx.set_option(*knob)
    try:
        lam, st, na, nr = gpu_ctx.run(desc, idx_list=torch.as_tensor(idx, device="cuda"))
        torch.cuda.synchronize()
    finally:
        gpu_ctx.set_option("item_order", 0)
        gpu_ctx.set_option(knob[0], 0)
        gpu_ctx.set_option("l96_ppw_8", 0)
    g = (lam.cpu().numpy(), st.cpu().numpy().astype(np.uint32), na.cpu().numpy().astype(np.uint32),
         nr.cpu().numpy().astype(np.uint32))
    o = O.sweep(O.make_desc(desc.to_dict()), first=0, count=n)
    o = tuple(a[idx] for a in o)
    assert_match(g, o, TOL_FP64, exact=True)


@pytest.mark.parametrize("method", [L.METHOD_VAR_REF, L.METHOD_QR_REF, L.METHOD_FD_REF])
@pytest.mark.parametrize("model", ["lorenz", "shimizu", "rossler"])
def test_reference_modes_match_oracle(gpu_ctx, model, method):
    """The reference's VAR, QR and FD modes on 81 points, renormalising every
    5 steps."""
    spec = W.lorenz_fixed(n=9, t_transient=2.0, t_average=5.0)
    spec.update(method=method, integrator=L.INT_DVERK, eps=1e-7, t_slave=1.0, t_renorm=0.05)
    if model == "shimizu":
        spec.update(model=L.MODEL_SHIMIZU, n_params=2, base_params=[0.0, 0.0],
                    init_state=[1e-4, 0.0, 0.0], init_kind=L.INIT_NONE)
        spec.update(W.grid_fields(0, 0.2, 2.0, 9, 1, 0.0, 1.2, 9))
    elif model == "rossler":
        spec.update(model=L.MODEL_ROSSLER, base_params=[0.0, 0.0, 0.2], init_state=[1.0, 1.0, 1.0],
                    init_scale=0.1)
        spec.update(W.grid_fields(0, 0.1, 0.3, 9, 1, 4.0, 6.0, 9))
    g, o = run_both(gpu_ctx, spec)
    assert_match(g, o, TOL_FP64, exact=True)


L96_LAYOUTS = [(8, 1), (16, 1), (16, 2), (16, 4), (32, 2), (32, 4)]


@pytest.mark.parametrize("n,waves", L96_LAYOUTS)
@pytest.mark.parametrize("integ,adaptive,prec", [(L.INT_DOPRI5, 1, L.PREC_FP64),
                                                 (L.INT_DOPRI5, 1, L.PREC_MIXED),
                                                 (L.INT_DVERK, 1, L.PREC_FP64),
                                                 (L.INT_DOPRI5, 0, L.PREC_FP64),
                                                 (L.INT_DVERK, 0, L.PREC_MIXED)])
def test_l96_matches_oracle(gpu_ctx, n, waves, integ, adaptive, prec):
    spec = W.lorenz96_dopri(n=3, dim=n, precision=prec, t_transient=2.0, t_average=6.0)
    spec.update(integrator=integ, adaptive=adaptive)
    gpu_ctx.set_option(f"l96_waves_{n}", waves)
    try:
        g, o = run_both(gpu_ctx, spec)
    finally:
        gpu_ctx.set_option(f"l96_waves_{n}", 0)
    assert_match(g, o, TOL_FP64, exact=True)


L96_WAVE_LAYOUTS = [(8, 1), (8, 2), (8, 4), (8, 8), (16, 1), (16, 2), (16, 4), (32, 1), (32, 2)]


@pytest.mark.parametrize("n,ppw", L96_WAVE_LAYOUTS)
@pytest.mark.parametrize("integ,adaptive,prec", [(L.INT_DOPRI5, 1, L.PREC_FP64),
                                                 (L.INT_DOPRI5, 1, L.PREC_MIXED),
                                                 (L.INT_DVERK, 1, L.PREC_MIXED),
                                                 (L.INT_DOPRI5, 0, L.PREC_FP64)])
def test_l96_wave_drain_matches_oracle(gpu_ctx, n, ppw, integ, adaptive, prec):
    """Wave-level drain (ppw points per wave64, slots refilling on their own):
    a grid of 2 workgroups over 25 points makes every slot take several points."""
    spec = W.lorenz96_dopri(n=5, dim=n, precision=prec, t_transient=2.0, t_average=6.0)
    spec.update(integrator=integ, adaptive=adaptive)
    gpu_ctx.set_option(f"l96_ppw_{n}", ppw)
    gpu_ctx.set_option("l96_max_blocks", 2)
    try:
        g, o = run_both(gpu_ctx, spec)
    finally:
        gpu_ctx.set_option(f"l96_ppw_{n}", 0)
        gpu_ctx.set_option("l96_max_blocks", 0)
    assert_match(g, o, TOL_FP64, exact=True)


@pytest.mark.parametrize("model", ["henon", "logistic"])
@pytest.mark.parametrize("method,prec", [(L.METHOD_VAR, L.PREC_FP64), (L.METHOD_VAR, L.PREC_MIXED),
                                         (L.METHOD_VAR_REF, L.PREC_FP64),
                                         (L.METHOD_FD_REF, L.PREC_FP64)])
def test_maps_match_oracle(gpu_ctx, model, method, prec):
    """Map-type models: the sweep's VAR method (FP64 and binary32 tangent
    products) and the reference's VAR/FD map branches; escaping Henon orbits
    included (a up to 1.42, b up to 0.35)."""
    if model == "henon":
        spec = W.henon(n=12, precision=prec, t_transient=100.0, t_average=2000.0)
    else:
        spec = W.logistic(n=96, precision=prec, t_transient=100.0, t_average=2000.0)
    spec.update(method=method, eps=1e-8)
    g, o = run_both(gpu_ctx, spec)
    assert_match(g, o, TOL_FP64, exact=True)
